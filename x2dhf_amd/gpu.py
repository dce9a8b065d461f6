"""ctypes binding of the x2gpu C ABI (include/x2gpu.h).

The library is the product: there is no CPU fallback.  Loading fails with an
ImportError naming the missing file when libx2gpu.so has not been built.
Device memory is managed through the C ABI (x2g_malloc/x2g_h2d/x2g_d2h), so
nothing here depends on torch.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libx2gpu.so")

X2G_OK = 0

_vp = C.c_void_p
_int = C.c_int
_dbl = C.c_double
_szt = C.c_size_t


class RelaxItem(C.Structure):
    """x2g_relax_item"""
    _fields_ = [("fun_d", _vp), ("lhs_d", _vp), ("rhs_d", _vp), ("tail_fock1_d", _vp),
                ("omega", _dbl), ("isym", _int), ("bkind", _int), ("prec", _int),
                ("grid_index", _int), ("rhs_f1_d", _vp), ("rhs_f2_d", _vp), ("rhs_f3_d", _vp),
                ("method", _int), ("ordering", _int)]


class Term(C.Structure):
    """x2g_term"""
    _fields_ = [("c", _dbl), ("a_d", _vp), ("b_d", _vp), ("cc_d", _vp)]


class Dot(C.Structure):
    """x2g_dot"""
    _fields_ = [("w_d", _vp), ("a_d", _vp), ("b_d", _vp), ("cc_d", _vp)]


MAX_ORTHO = 16


class LincombOp(C.Structure):
    """x2g_lincomb_op"""
    _fields_ = [("out_d", _vp), ("base_d", _vp), ("base_scale", _dbl), ("first_term", _int), ("nterms", _int),
                ("grid_index", _int)]


class NormOrthoItem(C.Structure):
    """x2g_normortho_item"""
    _fields_ = [("psi_d", _vp), ("others_d", _vp * MAX_ORTHO), ("nothers", _int), ("ihomo", _int),
                ("normalise", _int), ("grid_index", _int), ("norm_out_d", _vp)]


class EaHFItem(C.Structure):
    """x2g_eahf_item"""
    _fields_ = [("psi_d", _vp), ("cpot_d", _vp), ("xpot_d", _vp), ("isym", _int), ("m", _int),
                ("grid_index", _int), ("ee_out_d", _vp)]


class MomentItem(C.Structure):
    """x2g_moment_item"""
    _fields_ = [("psi1_d", _vp), ("psi2_d", _vp), ("dm", _int), ("grid_index", _int)]


class TailItem(C.Structure):
    """x2g_tail_item"""
    _fields_ = [("pot_d", _vp), ("kind", _int), ("dm", _int), ("grid_index", _int), ("mom", _dbl * 8)]


class Lcao(C.Structure):
    """x2g_lcao"""
    _fields_ = [("m", _int), ("n1", _int), ("l1", _int), ("n2", _int), ("l2", _int),
                ("c1", _dbl), ("za", _dbl), ("c2", _dbl), ("zb", _dbl)]


CTRL_MAXO = 16


class ScfCtrl(C.Structure):
    """x2g_scf_ctrl"""
    _fields_ = [("ethld", _dbl), ("nthld", _dbl), ("thren4init", _dbl), ("recalc_mm", _dbl), ("inc_thld", _dbl),
                ("ee", _dbl * CTRL_MAXO), ("eeold", _dbl * CTRL_MAXO), ("deltaee", _dbl * CTRL_MAXO),
                ("orbnorm", _dbl * CTRL_MAXO),
                ("eeconv", _dbl), ("eeconvprev", _dbl), ("dnmaxprev", _dbl), ("deltaeemax", _dbl),
                ("deltanormmax", _dbl),
                ("norb", _int), ("maxscf", _int), ("nscf2skip", _int), ("nenlast", _int), ("nnolast", _int),
                ("nscfextra", _int), ("maxresets", _int), ("polish", _int),
                ("ifixorb", _int * CTRL_MAXO), ("itouch", _int * CTRL_MAXO), ("mm_touch", _int * CTRL_MAXO),
                ("iscf", _int), ("nresets", _int), ("nresets4no", _int), ("norbenergythld", _int),
                ("norbnormthld", _int), ("noenergydec", _int), ("nonormdec", _int),
                ("lenergythld", _int), ("lnormthld", _int), ("lfixorb", _int), ("lfixorb4init", _int),
                ("mm_due", _int), ("prec", _int), ("polish_end", _int), ("stopcode", _int), ("stop", _int)]


# name -> (restype, argtypes)
_PROTOS = {
    "x2g_version": (_int, []),
    "x2g_last_error": (C.c_char_p, []),
    "x2g_ctx_create": (_int, [_int, C.POINTER(_vp)]),
    "x2g_ctx_destroy": (_int, [_vp]),
    "x2g_ctx_sync": (_int, [_vp]),
    "x2g_ctx_stream": (_vp, [_vp]),
    "x2g_mem_info": (_int, [_vp, C.POINTER(_szt), C.POINTER(_szt)]),
    "x2g_stats_enable": (_int, [_vp, _int]),
    "x2g_stats_get": (_int, [_vp, _vp]),
    "x2g_malloc": (_int, [_vp, _szt, C.POINTER(_vp)]),
    "x2g_free": (_int, [_vp, _vp]),
    "x2g_h2d": (_int, [_vp, _vp, _vp, _szt]),
    "x2g_d2h": (_int, [_vp, _vp, _vp, _szt]),
    "x2g_d2d": (_int, [_vp, _vp, _vp, _szt]),
    "x2g_fill": (_int, [_vp, _vp, _szt, _dbl]),
    "x2g_grid_create": (_int, [_vp, _int, _int, _vp, _vp, _vp, C.POINTER(_vp)]),
    "x2g_grid_destroy": (_int, [_vp]),
    "x2g_guess_orbital": (_int, [_vp, _vp, C.POINTER(Lcao)]),
    "x2g_guess_coulomb": (_int, [_vp, _vp, _dbl, _dbl, _int]),
    "x2g_guess_exchange": (_int, [_vp, _vp, _dbl, _dbl]),
    "x2g_grid_suppl": (_vp, [_vp, _int]),
    "x2g_set_sweep_engine": (_int, [_int, _int]),
    "x2g_set_levels_pipeline": (_int, [_int]),
    "x2g_putin": (_int, [_vp, _int, _vp, _vp]),
    "x2g_putout": (_int, [_vp, _vp, _vp]),
    "x2g_mcsor": (_int, [_vp, _int, _int, _vp, _vp, _vp, _dbl, _int]),
    "x2g_mcsor_host": (_int, [_vp, _int, _int, _vp, _vp, _vp, _dbl, _int]),
    "x2g_mcsor_noalloc": (_int, [_vp, _int, _int, _vp, _vp, _vp, _dbl, _int]),
    "x2g_set_autotune": (_int, [_int]),
    "x2g_set_sweep_plan": (_int, [_int, _int, _int]),
    "x2g_set_graphs": (_int, [_vp, _int]),
    "x2g_graph_counts": (_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "x2g_tune_report": (_int, [C.c_char_p, _int]),
    "x2g_relax_ordered": (_int, [_vp, _int, _int, _vp, _vp, _vp, _dbl, _int, _int, _int]),
    "x2g_relax_batch": (_int, [_vp, _vp, C.POINTER(RelaxItem), _int, _int, _int]),
    "x2g_residual_norms": (_int, [_vp, _vp, C.POINTER(RelaxItem), _int, _vp]),
    "x2g_relax_workspace_bytes": (C.c_longlong, [_int, _int, _int]),
    "x2g_lincomb": (_int, [_vp, _vp, _vp, _dbl, C.POINTER(Term), _int]),
    "x2g_wdot": (_int, [_vp, C.POINTER(Dot), _int, _vp]),
    "x2g_oneel": (_int, [_vp, _vp, _vp, _int, _int, _vp]),
    "x2g_moments": (_int, [_vp, _vp, _vp, _int, _int, _vp]),
    "x2g_pot_tail": (_int, [_vp, _vp, _int, _int, _int, _vp]),
    "x2g_setci": (_int, [_vp, _vp, _int]),
    "x2g_lincomb_batch": (_int, [_vp, _vp, C.POINTER(LincombOp), _int, C.POINTER(Term), _int]),
    "x2g_normortho_batch": (_int, [_vp, _vp, C.POINTER(NormOrthoItem), _int]),
    "x2g_eahf_batch": (_int, [_vp, _vp, C.POINTER(EaHFItem), _int]),
    "x2g_moments_batch": (_int, [_vp, _vp, C.POINTER(MomentItem), _int, _int, _vp]),
    "x2g_pot_tail_batch": (_int, [_vp, _vp, C.POINTER(TailItem), _int, _int]),
    "x2g_gather": (_int, [_vp, _vp, _vp, _int, _vp]),
    "x2g_scf_control": (_int, [_vp, _vp, _vp, _int]),
    "x2g_scf_ctrl_bytes": (_int, []),
}

_lib = None


def load():
    """Load libx2gpu.so (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"x2dhf_amd: HIP library {LIB_PATH} is missing; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class X2GError(RuntimeError):
    pass


def tune_report():
    """The sweep autotuner's tuned batch shapes (x2g_tune_report), one line each."""
    L = load()
    buf = C.create_string_buffer(1 << 16)
    L.x2g_tune_report(buf, len(buf))
    return [ln for ln in buf.value.decode().splitlines() if ln]


def check(rc):
    if rc != X2G_OK:
        raise X2GError(f"x2gpu error {rc}: {load().x2g_last_error().decode()}")


def _ptr(a):
    return a.ctypes.data_as(_vp)


class Context:
    """x2g_ctx: one device, one stream, workspaces."""

    def __init__(self, device=0):
        self.lib = load()
        h = _vp()
        check(self.lib.x2g_ctx_create(device, C.byref(h)))
        self.h = h
        self._bufs = []

    def close(self):
        if self.h:
            for b in self._bufs:
                if b.ptr:
                    self.lib.x2g_free(self.h, b.ptr)
                    b.ptr = None
            self._bufs = []
            self.lib.x2g_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        check(self.lib.x2g_ctx_sync(self.h))

    def alloc(self, n, dtype=np.float64):
        return DeviceArray(self, n, dtype)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        d = DeviceArray(self, arr.size, arr.dtype)
        d.set(arr)
        return d


class DeviceArray:
    """A device buffer of n elements allocated through x2g_malloc."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = _vp()
        check(ctx.lib.x2g_malloc(ctx.h, self.n * self.dtype.itemsize, C.byref(p)))
        self.ptr = p.value
        ctx._bufs.append(self)

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def set(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype).ravel()
        assert arr.size == self.n, (arr.size, self.n)
        check(self.ctx.lib.x2g_h2d(self.ctx.h, self.ptr, _ptr(arr), self.nbytes))

    def get(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(self.ctx.lib.x2g_d2h(self.ctx.h, _ptr(out), self.ptr, self.nbytes))
        return out

    def fill(self, v):
        assert self.dtype == np.float64
        check(self.ctx.lib.x2g_fill(self.ctx.h, self.ptr, self.n, float(v)))

    def free(self):
        if self.ptr:
            check(self.ctx.lib.x2g_free(self.ctx.h, self.ptr))
            self.ptr = None


class Grid:
    """x2g_grid: one (nni, nmu) grid with its coefficient arrays on the device.

    coef: 25 doubles {dmu1, dmu2, dni1, dni2, exeven, diag, hni, hmu, r};
    suppl: (14, nni*nmu) in the reference's i4b order; vec: vxi, vxisq, veta, vetasq.
    """

    def __init__(self, ctx, nni, nmu, coef, suppl, vxi, veta):
        self.ctx = ctx
        self.nni, self.nmu = int(nni), int(nmu)
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        assert coef.size == 25
        suppl = np.ascontiguousarray(suppl, dtype=np.float64).reshape(14, -1)
        assert suppl.shape[1] == nni * nmu
        vxi = np.asarray(vxi, dtype=np.float64)
        veta = np.asarray(veta, dtype=np.float64)
        vec = np.ascontiguousarray(np.concatenate([vxi, vxi * vxi, veta, veta * veta]))
        h = _vp()
        check(ctx.lib.x2g_grid_create(ctx.h, self.nni, self.nmu, _ptr(coef), _ptr(suppl), _ptr(vec), C.byref(h)))
        self.h = h
        self.coef = coef

    @property
    def n(self):
        return self.nni * self.nmu

    @property
    def npad(self):
        return (self.nni + 8) * (self.nmu + 8)

    def suppl_ptr(self, k):
        return self.ctx.lib.x2g_grid_suppl(self.h, k)

    def destroy(self):
        if self.h:
            self.ctx.lib.x2g_grid_destroy(self.h)
            self.h = None
