"""ctypes binding of the Fortran SCF driver (libx2scf.so) and molecule specs.

A job is one finite-difference Hartree-Fock calculation of a closed-shell
diatomic on one (nni x nmu) grid, with all arrays resident on the GPU.  The
driver keeps the reference's SCF control (doSCF.F90); see
x2dhf_amd/fortran/x2f_scf.f90.
"""
import ctypes as C
import os
import threading

import numpy as np

from . import gpu

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libx2scf.so")
MAXO = 16

_i, _d = C.c_int, C.c_double


class JobSpec(C.Structure):
    """x2f_spec (x2dhf_amd/fortran/x2f_api.f90)."""
    _fields_ = [("nni", _i), ("nmu", _i), ("norb", _i), ("maxscf", _i), ("ienterm", _i), ("inoterm", _i),
                ("nscfextra", _i), ("ienterm4init", _i), ("maxsororb", _i * 2), ("maxsorpot", _i * 2),
                ("mpole", _i), ("prec", _i), ("schedule", _i), ("device", _i),
                ("z1", _d), ("z2", _d), ("r", _d), ("rinf", _d), ("omega_orb", _d), ("omega_pot", _d),
                ("recalc_mm", _d),
                ("m", _i * MAXO), ("gu", _i * MAXO), ("n1", _i * MAXO), ("l1", _i * MAXO),
                ("n2", _i * MAXO), ("l2", _i * MAXO),
                ("occ", _d * MAXO), ("c1", _d * MAXO), ("c2", _d * MAXO), ("za", _d * MAXO), ("zb", _d * MAXO),
                ("nscf2skip", _i), ("nenlast", _i), ("nnolast", _i), ("polish", _i),
                ("orb_method", _i), ("pot_method", _i), ("ordering", _i), ("alt_sweeps", _i),
                ("host_control", _i)]


class JobResult(C.Structure):
    """x2f_result."""
    _fields_ = [("status", _i), ("iterations", _i), ("stopcode", _i), ("norb", _i),
                ("nsweep_orb", C.c_longlong), ("nsweep_pot", C.c_longlong), ("bytes", C.c_longlong),
                ("etot", _d), ("ekin", _d), ("enuc", _d), ("ecoul", _d), ("eexch", _d), ("virial", _d),
                ("delta_e", _d), ("delta_norm", _d),
                ("eps", _d * MAXO), ("norm", _d * MAXO), ("cmulti", (_d * 8) * MAXO)]


STOPCODES = {0: "running", 1: "orbital energy threshold", 2: "orbital norm threshold", 3: "iteration limit",
             4: "saturation", 5: "resets", 6: "diverged", 7: "not a number"}

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    gpu.load()  # libx2scf.so links libx2gpu.so
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"x2dhf_amd: Fortran driver {LIB_PATH} is missing; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.x2f_grid_build.argtypes = [_i, _i, _d, _d, _d, _d, vp, vp, vp]
    L.x2f_grid_build.restype = _i
    L.x2f_omega_defaults.argtypes = [_i, _i, _d, _d, C.POINTER(_d), C.POINTER(_d)]
    L.x2f_omega_defaults.restype = None
    L.x2f_job_create.argtypes = [C.POINTER(JobSpec), _i, C.POINTER(_i)]
    for n in ("x2f_job_prep", "x2f_job_finish", "x2f_job_run", "x2f_job_destroy"):
        getattr(L, n).argtypes = [_i]
        getattr(L, n).restype = _i
    L.x2f_job_iterate.argtypes = [_i, _i]
    L.x2f_job_iterate.restype = _i
    L.x2f_jobs_iterate_batched.argtypes = [_i, C.POINTER(_i), C.POINTER(_i)]
    L.x2f_jobs_iterate_batched.restype = _i
    L.x2f_job_result.argtypes = [_i, C.POINTER(JobResult)]
    L.x2f_job_result.restype = _i
    L.x2f_job_get.argtypes = [_i, _i, _i, vp]
    L.x2f_job_get.restype = _i
    L.x2f_job_put.argtypes = [_i, _i, _i, vp]
    L.x2f_job_put.restype = _i
    L.x2f_job_prep_restart.argtypes = [_i, vp, _i]
    L.x2f_job_prep_restart.restype = _i
    L.x2f_job_bytes.argtypes = [_i]
    L.x2f_job_bytes.restype = C.c_longlong
    L.x2f_job_ctx.argtypes = [_i]
    L.x2f_job_ctx.restype = vp
    L.x2f_job_create_ctx.argtypes = [C.POINTER(JobSpec), vp, C.POINTER(_i)]
    L.x2f_job_create_ctx.restype = _i
    L.x2f_spec_bytes.argtypes = [C.POINTER(JobSpec)]
    L.x2f_spec_bytes.restype = C.c_longlong
    L.x2f_spec_workspace_bytes.argtypes = [C.POINTER(JobSpec)]
    L.x2f_spec_workspace_bytes.restype = C.c_longlong
    L.x2f_control_probe.argtypes = [_i, _i, vp, vp, vp, _i, _i, _i, _i, C.POINTER(_i), C.POINTER(_i)]
    L.x2f_control_probe.restype = _i
    _lib = L
    return L


def grid_build(nni, nmu, z1, z2, r, rinf):
    """Host grid: (coef[25], suppl[14, nni*nmu], vxi, veta) from the Fortran builder."""
    L = load()
    coef = np.zeros(25)
    suppl = np.zeros((14, nni * nmu))
    vec = np.zeros(2 * nmu + 2 * nni)
    rc = L.x2f_grid_build(nni, nmu, z1, z2, r, rinf, coef.ctypes.data, suppl.ctypes.data, vec.ctypes.data)
    if rc != 0:
        raise ValueError(f"grid {nni}x{nmu}: sizes must be 1 mod 6 and >= 13")
    return coef, suppl, vec[:nmu].copy(), vec[2 * nmu:2 * nmu + nni].copy()


# ---------------------------------------------------------------------------
# molecules

class Orbital:
    """One closed-shell orbital: m (0 sigma, 1 pi, 2 delta), g/u label (+1, -1,
    0) and its LCAO guess c1*phi_A(n1,l1,m;za) + c2*phi_B(n2,l2,m;zb)."""

    def __init__(self, m, gu, c1, n1, l1, za, c2, n2, l2, zb):
        self.m, self.gu = m, gu
        self.c1, self.n1, self.l1, self.za = c1, n1, l1, za
        self.c2, self.n2, self.l2, self.zb = c2, n2, l2, zb

    @property
    def occ(self):
        return 2.0 if self.m == 0 else 4.0


def make_spec(z1, z2, r, nni, nmu, rinf, orbitals, maxscf=3000, ienterm=10, inoterm=8, omega=(-1.0, -1.0),
              prec=0, schedule=0, device=0, maxsor=(1, 10), mpole=4, nscfextra=3, ienterm4init=1,
              maxsorpot=None, recalc_mm=1.15, conv=(0, 0, 0), polish=0, orb_method=0, pot_method=0, ordering=0,
              alt_sweeps=0, host_control=0):
    """Job specification.  maxsor / maxsorpot: (macro-iterations, sweeps) of
    the orbital / potential relaxations (maxsorpot defaults to maxsor);
    conv: (nscf2skip, nenlast, nnolast) of the saturation test, 0 = default;
    polish: with prec = 1, the number of FP64-sweep iterations the job ends
    with (it then stops at the FP64 path's fixed point), 0 = none;
    orb_method / pot_method: 0 MCSOR, 1 SOR, and ordering: an X2G_ORDER_*
    code (0 col-wise, 1 middle, ... 9 middle2; deck._ORDERINGS), the
    relaxation scheme of the FP64 sweeps (0/0/0 is the
    pipelined col-wise colour sweep; the FP32 correction sweeps are always
    col-wise MCSOR); alt_sweeps: the altsweeps option, SOR sweeps backward
    for the odd orbitals (doSCF.F90:403-411; no effect on MCSOR);
    host_control: run the batched schedule's end-of-iteration control with
    the host rules instead of on the device (x2g_scf_control); the plain
    schedule always uses the host rules."""
    s = JobSpec()
    s.nni, s.nmu, s.norb = nni, nmu, len(orbitals)
    s.maxscf, s.ienterm, s.inoterm, s.nscfextra, s.ienterm4init = maxscf, ienterm, inoterm, nscfextra, ienterm4init
    s.maxsororb[0], s.maxsororb[1] = maxsor
    s.maxsorpot[0], s.maxsorpot[1] = maxsor if maxsorpot is None else maxsorpot
    s.mpole, s.prec, s.schedule, s.device = mpole, prec, schedule, device
    s.z1, s.z2, s.r, s.rinf = z1, z2, r, rinf
    s.omega_orb, s.omega_pot = omega
    s.recalc_mm = recalc_mm
    s.nscf2skip, s.nenlast, s.nnolast = conv
    s.polish = polish
    s.orb_method, s.pot_method, s.ordering = orb_method, pot_method, ordering
    s.alt_sweeps = 1 if alt_sweeps else 0
    s.host_control = 1 if host_control else 0
    for k, o in enumerate(orbitals):
        s.m[k], s.gu[k], s.occ[k] = o.m, o.gu, o.occ
        s.c1[k], s.n1[k], s.l1[k], s.za[k] = o.c1, o.n1, o.l1, o.za
        s.c2[k], s.n2[k], s.l2[k], s.zb[k] = o.c2, o.n2, o.l2, o.zb
    return s


def h2_spec(nni=91, nmu=121, rinf=40.0, r=1.4, **kw):
    """H2, one sigma-g orbital from hydrogenic 1s functions (configs 1)."""
    return make_spec(1.0, 1.0, r, nni, nmu, rinf, [Orbital(0, 1, 1.0, 1, 0, 1.0, 1.0, 1, 0, 1.0)], **kw)


def n2_spec(nni=331, nmu=601, rinf=40.0, r=2.068, **kw):
    """N2: 1 pi-u, 3 sigma-g, 2 sigma-u, 2 sigma-g, 1 sigma-u, 1 sigma-g with the
    hydrogenic LCAO guess of the reference's N2 deck (config 2)."""
    orbs = [Orbital(1, -1, 0.5, 2, 1, 5.0, 0.5, 2, 1, 5.0),
            Orbital(0, 1, 1.5, 2, 1, 5.0, 0.0, 2, 1, 5.0),
            Orbital(0, -1, 0.5, 2, 0, 7.0, -0.5, 2, 0, 7.0),
            Orbital(0, 1, 0.5, 2, 0, 7.0, 0.5, 2, 0, 7.0),
            Orbital(0, -1, 0.5, 1, 0, 7.0, -0.5, 1, 0, 7.0),
            Orbital(0, 1, 0.5, 1, 0, 7.0, 0.5, 1, 0, 7.0)]
    return make_spec(7.0, 7.0, r, nni, nmu, rinf, orbs, **kw)


def fh_spec(nni=481, nmu=901, rinf=50.0, r=1.733, **kw):
    """FH: 1 pi, 3 sigma, 2 sigma, 1 sigma from hydrogenic functions (config 3)."""
    orbs = [Orbital(1, 0, 1.0, 2, 1, 5.2, 0.0, 2, 1, 1.0),
            Orbital(0, 0, 0.7, 2, 1, 5.2, 0.5, 1, 0, 1.0),
            Orbital(0, 0, 1.0, 2, 0, 5.2, 0.0, 1, 0, 1.0),
            Orbital(0, 0, 1.0, 1, 0, 9.0, 0.0, 1, 0, 1.0)]
    return make_spec(9.0, 1.0, r, nni, nmu, rinf, orbs, **kw)


def lih_spec(nni=169, nmu=241, rinf=40.0, r=3.015, **kw):
    """LiH: 2 sigma, 1 sigma with the hydrogenic LCAO guess of the reference's
    LiH deck (test-sets/lih/set-01/input.data)."""
    orbs = [Orbital(0, 0, 1.0, 1, 0, 3.0, -1.0, 1, 0, 1.0),
            Orbital(0, 0, 1.0, 1, 0, 3.0, 1.0, 1, 0, 1.0)]
    return make_spec(3.0, 1.0, r, nni, nmu, rinf, orbs, **kw)


def bh_spec(nni=169, nmu=241, rinf=40.0, r=2.3289, **kw):
    """BH: 3 sigma, 2 sigma, 1 sigma with the hydrogenic LCAO guess of the
    reference's BH deck (test-sets/bh-elprop/set-01/input.data)."""
    orbs = [Orbital(0, 0, 1.0, 2, 1, 4.0, -1.0, 1, 0, 1.0),
            Orbital(0, 0, 1.0, 2, 0, 4.0, 0.0, 1, 0, 1.0),
            Orbital(0, 0, 1.0, 1, 0, 5.0, 0.0, 1, 0, 1.0)]
    return make_spec(5.0, 1.0, r, nni, nmu, rinf, orbs, **kw)


def spec_bytes(spec):
    """Device bytes a job of this specification holds (its own arrays), without
    creating it; raises ValueError for a specification the driver rejects."""
    b = int(load().x2f_spec_bytes(C.byref(spec)))
    if b < 0:
        raise ValueError("unsupported job specification")
    return b


def spec_workspace_bytes(spec):
    """Device workspace bytes a job of this specification adds to its context
    while it is relaxed (the library's relaxation and quadrature buffers)."""
    b = int(load().x2f_spec_workspace_bytes(C.byref(spec)))
    if b < 0:
        raise ValueError("unsupported job specification")
    return b


def control_probe(ee, norms, ee0, maxscf=1000, ienterm=10, inoterm=8, nscf2skip=0):
    """Run the SCF loop's host bookkeeping (energy_change and
    convergence_control of the Fortran driver) on given per-iteration orbital
    energies ee[nit, norb] and norms[nit, norb], starting from energies ee0;
    no device work.  Returns (stop reason, iteration it stopped at)."""
    ee = np.asfortranarray(np.asarray(ee, dtype=np.float64).T)
    nm = np.asfortranarray(np.asarray(norms, dtype=np.float64).T)
    e0 = np.ascontiguousarray(ee0, dtype=np.float64)
    norb, nit = ee.shape
    sc, it = _i(), _i()
    rc = load().x2f_control_probe(norb, nit, ee.ctypes.data, nm.ctypes.data, e0.ctypes.data, maxscf, ienterm,
                                  inoterm, nscf2skip, C.byref(sc), C.byref(it))
    if rc != 0:
        raise ValueError("x2f_control_probe: bad sizes")
    return STOPCODES.get(sc.value, str(sc.value)), it.value


# The Fortran job table is modified only by job creation and destruction;
# host threads that step different jobs serialise those two through this lock.
JOB_TABLE_LOCK = threading.Lock()


class Job:
    """One SCF job on the GPU.  share: another Job whose device context this
    job uses; ctx: an x2gpu context (gpu.Context) owned by the caller.  Jobs
    stepped together by iterate_batched must use one context."""

    def __init__(self, spec, share=None, ctx=None):
        self.L = load()
        self.spec = spec
        h = _i()
        with JOB_TABLE_LOCK:
            if ctx is not None:
                rc = self.L.x2f_job_create_ctx(C.byref(spec), ctx.h.value if hasattr(ctx.h, "value") else ctx.h,
                                               C.byref(h))
            else:
                rc = self.L.x2f_job_create(C.byref(spec), share.h if share is not None else 0, C.byref(h))
        if rc != 0:
            raise ValueError(f"x2f_job_create failed with code {rc}")
        self.h = h.value

    def prep(self):
        self._chk(self.L.x2f_job_prep(self.h))

    def prep_restart(self, ee):
        """Set-up pass after loading a checkpoint with put(): ee is the
        checkpoint's Lagrange-multiplier matrix (at least norb x norb)."""
        a = np.asfortranarray(np.asarray(ee, dtype=np.float64))
        self._chk(self.L.x2f_job_prep_restart(self.h, a.ctypes.data, a.shape[0]))

    def iterate(self, n=1):
        """Run up to n SCF iterations; True when the SCF stopped."""
        rc = self.L.x2f_job_iterate(self.h, n)
        if rc < 0:
            raise RuntimeError("x2f_job_iterate failed")
        return rc == 1

    def finish(self):
        self._chk(self.L.x2f_job_finish(self.h))

    def run(self):
        self._chk(self.L.x2f_job_run(self.h))
        return self.result()

    def result(self):
        r = JobResult()
        self._chk(self.L.x2f_job_result(self.h, C.byref(r)))
        n = r.norb
        return {
            "etot": r.etot, "ekin": r.ekin, "enuc": r.enuc, "ecoul": r.ecoul, "eexch": r.eexch,
            "virial": r.virial, "iterations": r.iterations, "stop": STOPCODES.get(r.stopcode, str(r.stopcode)),
            "eps": np.array(r.eps[:n]), "norm": np.array(r.norm[:n]),
            "cmulti": np.array([list(r.cmulti[i]) for i in range(n)]), "nsweep_orb": r.nsweep_orb,
            "nsweep_pot": r.nsweep_pot, "bytes": r.bytes, "delta_e": r.delta_e, "delta_norm": r.delta_norm,
            "done": bool(r.status),
        }

    def get(self, kind, i):
        """kind 0: orbital i, 1: Coulomb potential i, 2: exchange slot i (1-based)."""
        out = np.empty(self.spec.nni * self.spec.nmu)
        self._chk(self.L.x2f_job_get(self.h, kind, i, out.ctypes.data))
        return out

    def put(self, kind, i, values):
        """Load orbital i (kind 0), Coulomb potential i (1) or exchange slot i
        (2, 1-based) from a host array of nni*nmu values (before prep())."""
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if a.size != self.spec.nni * self.spec.nmu:
            raise ValueError(f"put: {a.size} values for a {self.spec.nni}x{self.spec.nmu} grid")
        self._chk(self.L.x2f_job_put(self.h, kind, i, a.ctypes.data))

    def bytes(self):
        return int(self.L.x2f_job_bytes(self.h))

    def ctx(self):
        """The job's x2g_ctx* (shared by every job created with share=)."""
        return self.L.x2f_job_ctx(self.h)

    def close(self):
        if self.h:
            with JOB_TABLE_LOCK:
                self.L.x2f_job_destroy(self.h)
            self.h = 0

    @staticmethod
    def _chk(rc):
        if rc != 0:
            raise RuntimeError(f"x2f call failed with code {rc}")


def iterate_batched(jobs, stopped=None):
    """One batched SCF iteration over jobs sharing one device context."""
    L = load()
    n = len(jobs)
    hs = (_i * n)(*[j.h for j in jobs])
    st = (_i * n)(*([0] * n if stopped is None else [int(s) for s in stopped]))
    rc = L.x2f_jobs_iterate_batched(n, hs, st)
    if rc != 0:
        raise RuntimeError("x2f_jobs_iterate_batched failed (jobs must share a context)")
    return [bool(v) for v in st]
