"""The Fortran interface blocks INTEGRATION.md shows a maintainer of the
reference (sections 1-3 and 6) compile with flang and link against
libx2gpu.so: every bind(c) name they declare resolves to an exported symbol
of the library, with the calls written as the document writes them.  No
compute call is made (this runs without a GPU)."""
import os
import re
import subprocess

import pytest

from conftest import ROOT

FLANG = "/opt/rocm/lib/llvm/bin/flang"
DOC = os.path.join(ROOT, "INTEGRATION.md")
LIBDIR = os.path.join(ROOT, "x2dhf_amd", "_lib")


def fortran_blocks():
    txt = open(DOC).read()
    return re.findall(r"```fortran\n(.*?)```", txt, flags=re.S)


def test_document_has_the_binding_blocks():
    blocks = fortran_blocks()
    joined = "\n".join(blocks)
    for name in ("x2g_ctx_create", "x2g_grid_create", "x2g_mcsor_host", "x2g_relax_batch", "x2g_guess_orbital"):
        assert name in joined, name


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_interface_blocks_compile_and_link(tmp_path):
    blocks = fortran_blocks()
    bind_mod = next(b for b in blocks if b.lstrip().startswith("module x2gpu_bind"))
    batch = next(b for b in blocks if "x2g_relax_batch" in b)
    orb_call = next(b for b in blocks if b.lstrip().startswith("call putin"))
    guess = next(b for b in blocks if "x2g_guess_orbital(g" in b)
    # section 3: the item type and its interface, as a module
    batch_mod = "module x2gpu_batch\n  use, intrinsic :: iso_c_binding\n  implicit none\n" + batch + "end module\n"
    # section 6: the lcao type; its calls' interfaces are written here as a
    # maintainer would add them next to section 1's
    lcao_type = guess[:guess.index("end type") + len("end type")] + "\n"
    guess_calls = "\n".join(l for l in guess.splitlines() if l.strip().startswith("ierr"))
    guess_mod = ("module x2gpu_guess\n  use, intrinsic :: iso_c_binding\n  implicit none\n" + lcao_type +
                 "  interface\n"
                 "    integer(c_int) function x2g_guess_orbital(g, orb, lc) bind(c, name='x2g_guess_orbital')\n"
                 "      import :: c_int, c_ptr, x2g_lcao\n"
                 "      type(c_ptr), value :: g, orb\n      type(x2g_lcao), intent(in) :: lc\n    end function\n"
                 "    integer(c_int) function x2g_guess_coulomb(g, pot, c1, c2, two) bind(c, name='x2g_guess_coulomb')\n"
                 "      import :: c_int, c_ptr, c_double\n      type(c_ptr), value :: g, pot\n"
                 "      real(c_double), value :: c1, c2\n      integer(c_int), value :: two\n    end function\n"
                 "    integer(c_int) function x2g_guess_exchange(g, pot, c1, c2) bind(c, name='x2g_guess_exchange')\n"
                 "      import :: c_int, c_ptr, c_double\n      type(c_ptr), value :: g, pot\n"
                 "      real(c_double), value :: c1, c2\n    end function\n"
                 "  end interface\nend module\n")
    # section 2: the orbital relaxation call between the reference's own putin
    # and putout (left out here: they are the reference's routines)
    orb_stmts = "\n".join(l for l in orb_call.splitlines() if not l.strip().startswith("call put"))
    driver = f"""
subroutine orb_relax(g, wk2, lhs, fock2, ovforb, maxsor2, iorb, isym)
  use, intrinsic :: iso_c_binding
  use x2gpu_bind
  implicit none
  type(c_ptr), value :: g
  real(c_double) :: wk2(*), lhs(*), fock2(*), ovforb(*)
  integer :: maxsor2, iorb, ierr
  integer(c_int) :: isym
{orb_stmts}
end subroutine

subroutine starts(g, orb_d, coul_d, exch_d, lcao, c1, c2, two_centres)
  use, intrinsic :: iso_c_binding
  use x2gpu_guess
  implicit none
  type(c_ptr), value :: g, orb_d, coul_d, exch_d
  type(x2g_lcao) :: lcao
  real(c_double) :: c1, c2
  integer(c_int) :: two_centres
  integer :: ierr
{guess_calls}
end subroutine

subroutine batch(ctx, grids, items, n)
  use, intrinsic :: iso_c_binding
  use x2gpu_batch
  implicit none
  type(c_ptr), value :: ctx
  type(c_ptr) :: grids(*)
  type(x2g_relax_item) :: items(*)
  integer(c_int) :: n, ierr
  ierr = x2g_relax_batch(ctx, grids, items, n, 1_c_int, 10_c_int)
  if (ierr /= 0) stop 'x2g_relax_batch'
end subroutine

program integration_doc
  use, intrinsic :: iso_c_binding
  use x2gpu_bind
  implicit none
  type(c_ptr) :: ctx, g
  integer(c_int) :: rc
  real(c_double) :: coef(25), suppl(14), vec(4)
  ! linked, never called here (no device in this test)
  if (command_argument_count() > 99) then
     rc = x2g_ctx_create(0_c_int, ctx)
     rc = x2g_grid_create(ctx, 1_c_int, 1_c_int, coef, suppl, vec, g)
  end if
  print '(a)', 'linked'
end program
"""
    src = tmp_path / "doc.f90"
    src.write_text(bind_mod + "\n" + batch_mod + "\n" + guess_mod + "\n" + driver)
    exe = tmp_path / "doc"
    r = subprocess.run([FLANG, "-O0", "-J", str(tmp_path), "-o", str(exe), str(src), "-L" + LIBDIR, "-lx2gpu",
                        "-Wl,--no-as-needed", "-Wl,-rpath," + LIBDIR], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], capture_output=True, text=True).stdout
    exports = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIBDIR, "libx2gpu.so")],
                             capture_output=True, text=True).stdout
    used = sorted(set(re.findall(r"\b(x2g_[a-z0-9_]+)\b", nm)))
    assert {"x2g_mcsor_host", "x2g_relax_batch", "x2g_grid_create", "x2g_guess_orbital"} <= set(used), used
    for name in used:
        assert re.search(r"\bT " + name + r"\b", exports), name
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and "linked" in out.stdout, out.stdout + out.stderr
