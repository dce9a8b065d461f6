"""x2g_mcsor_noalloc (the mcsor entry point that allocates no device memory,
its workspace in a static device array) against x2g_mcsor, through the C
ABI: bit for bit on two calls in a row, the second on a smaller grid (the
array keeps the first call's contents), and a grid too large for the array
is refused."""
import numpy as np
import pytest

from conftest import grid_coef, load_golden

pytestmark = pytest.mark.gpu


def _synthetic(ctx, nni, nmu, seed):
    from x2dhf_amd import gpu, scf
    coef, suppl, vxi, veta = scf.grid_build(nni, nmu, 7.0, 7.0, 2.068, 40.0)
    g = gpu.Grid(ctx, nni, nmu, coef, suppl, vxi, veta)
    rng = np.random.default_rng(seed)
    work = rng.standard_normal((nni + 8) * (nmu + 8))
    lhs = suppl[7] + 0.1 * rng.standard_normal(nni * nmu)
    rhs = rng.standard_normal(nni * nmu)
    return g, work, lhs, rhs


def _both(ctx, g, isym, bkind, work, lhs, rhs, omega, nsw):
    from x2dhf_amd import gpu
    lhs_d, rhs_d = ctx.upload(lhs), ctx.upload(rhs)
    out = []
    for name in ("x2g_mcsor", "x2g_mcsor_noalloc"):
        w = ctx.upload(work)
        gpu.check(getattr(ctx.lib, name)(g.h, isym, bkind, w.ptr, lhs_d.ptr, rhs_d.ptr, omega, nsw))
        out.append(w.get())
        w.free()
    lhs_d.free()
    rhs_d.free()
    return out


def test_noalloc_equals_mcsor_on_two_calls_second_smaller():
    from x2dhf_amd import gpu
    with gpu.Context(0) as ctx:
        print("noalloc: 331x601 synthetic, then the 61x91 golden", flush=True)
        g, work, lhs, rhs = _synthetic(ctx, 331, 601, 1)
        a, b = _both(ctx, g, 1, 1, work, lhs, rhs, 1.9, 10)
        g.destroy()
        assert np.array_equal(a, b)
        z = load_golden("n2_61x91")
        nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
        g = gpu.Grid(ctx, nni, nmu, grid_coef(z), z["suppl"], z["vxi"], z["veta"])
        a, b = _both(ctx, g, int(z["k1_isym"][0]), 1, z["k1_win"], z["k1_lhs"], z["k1_rhs"],
                     float(z["k1_scal"][0]), int(z["k1_scal"][1]))
        g.destroy()
        assert np.array_equal(a, b)
        ref = z["k1_wout_mcsor"]
        assert np.abs(b - ref).max() < 1e-12 * np.abs(ref).max()


def test_noalloc_refuses_a_grid_beyond_its_scratch():
    from x2dhf_amd import gpu, scf
    with gpu.Context(0) as ctx:
        nni, nmu = 1501, 2401          # 29 MB padded: beyond the 24 MiB array for every sweep kind
        coef, suppl, vxi, veta = scf.grid_build(nni, nmu, 1.0, 1.0, 1.4, 40.0)
        g = gpu.Grid(ctx, nni, nmu, coef, suppl, vxi, veta)
        w = ctx.alloc((nni + 8) * (nmu + 8))
        l, r = ctx.upload(suppl[7]), ctx.alloc(nni * nmu)
        rc = ctx.lib.x2g_mcsor_noalloc(g.h, 1, 1, w.ptr, l.ptr, r.ptr, 1.9, 1)
        assert rc == -1
        assert b"scratch" in ctx.lib.x2g_last_error()
        for a in (w, l, r):
            a.free()
        g.destroy()
