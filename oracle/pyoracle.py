"""ctypes loader of the CPU oracle (oracle/x2oracle.c) -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this module, and only to check or time the product; the product path
(x2dhf_amd) never does.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libx2oracle.so")
_vp, _int, _dbl = C.c_void_p, C.c_int, C.c_double
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"oracle library {LIB_PATH} missing; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.x2o_putin.argtypes = [_int, _int, _int, _vp, _vp]
        L.x2o_putout.argtypes = [_int, _int, _vp, _vp]
        L.x2o_boundary.argtypes = [_int, _int, _int, _vp, _vp]
        L.x2o_mcsor.argtypes = [_int, _int, _int] + [_vp] * 7 + [_dbl, _int, _int]
        L.x2o_mcsor_f32.argtypes = [_int, _int, _int] + [_vp] * 7 + [_dbl, _int]
        L.x2o_residual.argtypes = [_int, _int] + [_vp] * 7
        L.x2o_orb_asympt_get.argtypes = [_int, _int, _vp, _vp, _vp, _vp]
        L.x2o_orb_asympt_set.argtypes = [_int, _int, _vp, _vp]
        L.x2o_mesh.argtypes = [_int, _int, _int, _vp]
        L.x2o_mesh.restype = C.c_long
        L.x2o_relax_ordered.argtypes = [_int, _int, _int] + [_vp] * 7 + [_dbl, _int, _int, _int, _vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def putin(nni, nmu, isym, fun, work=None):
    if work is None:
        work = np.zeros((nni + 8) * (nmu + 8))
    lib().x2o_putin(nni, nmu, isym, _p(np.ascontiguousarray(fun, np.float64)), _p(work))
    return work


def putout(nni, nmu, work):
    fun = np.empty(nni * nmu)
    lib().x2o_putout(nni, nmu, _p(fun), _p(np.ascontiguousarray(work)))
    return fun


def mcsor(nni, nmu, isym, work, lhs, rhs, b2d, d2d, coef16, exeven, omega, nsweeps, mode=0):
    """In-place relaxation of the padded `work` (mode 0: mcsor, 1: sor)."""
    lib().x2o_mcsor(nni, nmu, isym, _p(work), _p(np.ascontiguousarray(lhs, np.float64)),
                    _p(np.ascontiguousarray(rhs, np.float64)), _p(np.ascontiguousarray(b2d, np.float64)),
                    _p(np.ascontiguousarray(d2d, np.float64)), _p(np.ascontiguousarray(coef16, np.float64)),
                    _p(np.ascontiguousarray(exeven, np.float64)), float(omega), int(nsweeps), int(mode))
    return work


def mcsor_f32(nni, nmu, isym, work, lhs, rhs, b2d, d2d, coef16, exeven, omega, nsweeps):
    lib().x2o_mcsor_f32(nni, nmu, isym, _p(work), _p(np.ascontiguousarray(lhs, np.float32)),
                        _p(np.ascontiguousarray(rhs, np.float32)), _p(np.ascontiguousarray(b2d, np.float32)),
                        _p(np.ascontiguousarray(d2d, np.float32)), _p(np.ascontiguousarray(coef16, np.float64)),
                        _p(np.ascontiguousarray(exeven, np.float64)), float(omega), int(nsweeps))
    return work


def residual(nni, nmu, work, lhs, rhs, b2d, d2d, coef16):
    res = np.empty(nni * nmu)
    lib().x2o_residual(nni, nmu, _p(np.ascontiguousarray(work)), _p(np.ascontiguousarray(lhs)),
                       _p(np.ascontiguousarray(rhs)), _p(np.ascontiguousarray(b2d)), _p(np.ascontiguousarray(d2d)),
                       _p(np.ascontiguousarray(coef16)), _p(res))
    return res


def orb_tail(nni, nmu, vxisq, vetasq, fa, psi):
    """orbAsymptGet + orbAsymptSet on a copy of psi."""
    ed = np.empty(4 * nni)
    lib().x2o_orb_asympt_get(nni, nmu, _p(np.ascontiguousarray(vxisq)), _p(np.ascontiguousarray(vetasq)),
                             _p(np.ascontiguousarray(fa)), _p(ed))
    out = np.array(psi, dtype=np.float64, copy=True)
    lib().x2o_orb_asympt_set(nni, nmu, _p(out), _p(ed))
    return out


# the reference's mesh orderings (initCBAllocArrays.F90:2458-2810) by the
# 8-character name of its `order` card
ORDERINGS = {"col-wise": 0, "middle": 1, "middlemt": 2, "row-wise": 3, "rrow-wis": 4, "test1": 5, "test2": 6,
             "test3": 7, "test4": 8, "middle2": 9}


def mesh(nni, nmu, ordering):
    """Relaxation sequence of mesh ordering (name or code) as unpadded indices."""
    code = ORDERINGS.get(ordering, ordering)
    seq = np.empty((nni - 2) * (nmu - 5), dtype=np.int32)
    lib().x2o_mesh(nni, nmu, int(code), _p(seq))
    return seq


def relax_ordered(nni, nmu, isym, work, lhs, rhs, b2d, d2d, coef16, exeven, omega, nsweeps, mode, ordering):
    """In-place sor (mode 1) / mcsor (mode 0) of `work` in mesh ordering
    `ordering`; mode 2 is sor walking the ordering backwards (altsweeps)."""
    code = ORDERINGS.get(ordering, ordering)
    seq = np.empty((nni - 2) * (nmu - 5), dtype=np.int32)
    lib().x2o_relax_ordered(nni, nmu, isym, _p(work), _p(np.ascontiguousarray(lhs, np.float64)),
                            _p(np.ascontiguousarray(rhs, np.float64)), _p(np.ascontiguousarray(b2d, np.float64)),
                            _p(np.ascontiguousarray(d2d, np.float64)), _p(np.ascontiguousarray(coef16, np.float64)),
                            _p(np.ascontiguousarray(exeven, np.float64)), float(omega), int(nsweeps), int(mode),
                            int(code), _p(seq))
    return work
