"""Parity of the HIP kernels (through the C ABI) with the reference's own
routines (golden vectors) and with the CPU oracle on seeded inputs.

Tolerances: FP64 kernels agree with the reference to 1e-12 relative to the
array's largest magnitude (the device forms fused multiply-adds, the
reference does not); FP32 sweeps agree with the FP32 oracle to 2e-5.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import grid_coef, load_golden
from oracle import pyoracle

pytestmark = pytest.mark.gpu

REL64 = 1e-12


@pytest.fixture(scope="module")
def ctx():
    from x2dhf_amd import gpu
    c = gpu.Context(0)
    yield c
    gpu.load().x2g_set_sweep_engine(0, -1)
    gpu.load().x2g_set_sweep_plan(0, 0, 0)
    c.close()


def make_grid(ctx, z):
    from x2dhf_amd import gpu
    nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
    return gpu.Grid(ctx, nni, nmu, grid_coef(z), z["suppl"], z["vxi"], z["veta"])


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def engines():
    # (engine, strip rows, block rows): the default sweep kernel with the
    # library's strip height and with a forced small strip (several strips per
    # problem), in blocks of 4 and of 2 rows, the first strip kernel likewise,
    # and the per-pass kernels
    return [(0, -1, 4), (0, 24, 4), (0, -1, 2), (0, 22, 2), (2, -1, 4), (2, 32, 4), (1, -1, 4)]


def set_engine(ctx, eng):
    from x2dhf_amd import gpu
    engine, rows, br = eng
    gpu.check(ctx.lib.x2g_set_sweep_engine(engine, rows))
    gpu.check(ctx.lib.x2g_set_sweep_plan(br if br != 4 else 0, 0, 0))


@pytest.mark.parametrize("case", ["h2_61x91", "n2_61x91"])
@pytest.mark.parametrize("pfx", ["k1_", "k2_"])
@pytest.mark.parametrize("eng", engines())
def test_mcsor_vs_reference(ctx, case, pfx, eng):
    from x2dhf_amd import gpu
    z = load_golden(case)
    if pfx + "win" not in z:
        pytest.skip("single potential")
    print(f"case={case} {pfx} engine={eng}", flush=True)
    set_engine(ctx, eng)
    g = make_grid(ctx, z)
    isym = int(z[pfx + "isym"][0])
    omega, nsw = float(z[pfx + "scal"][0]), int(z[pfx + "scal"][1])
    lhs, rhs = ctx.upload(z[pfx + "lhs"]), ctx.upload(z[pfx + "rhs"])
    for name, ns in (("wout_mcsor", nsw), ("wout_mcsor1", 1)):
        w = ctx.upload(z[pfx + "win"])
        gpu.check(ctx.lib.x2g_mcsor(g.h, isym, 1, w.ptr, lhs.ptr, rhs.ptr, omega, ns))
        e = relerr(w.get(), z[pfx + name])
        assert e < REL64, (name, e)
        w.free()
    g.destroy()


def test_putin_putout_vs_reference(ctx):
    from x2dhf_amd import gpu
    z = load_golden("n2_61x91")
    g = make_grid(ctx, z)
    for pfx, k in (("k1_", 0), ("k2_", 1)):
        isym = int(z[pfx + "isym"][0])
        f = ctx.upload(z["pots_asympt"][k])
        w = ctx.alloc(g.npad)
        gpu.check(ctx.lib.x2g_putin(g.h, isym, f.ptr, w.ptr))
        assert np.array_equal(w.get(), z[pfx + "win"])
        out = ctx.alloc(g.n)
        gpu.check(ctx.lib.x2g_putout(g.h, out.ptr, w.ptr))
        assert np.array_equal(out.get(), z["pots_asympt"][k])
    g.destroy()


def _setci_all(ctx, g, z, orbs):
    """Impose the g/u symmetry of every orbital (setCi) as the reference does
    before the SCF loop (doSCF.F90:334-336)."""
    from x2dhf_amd import gpu
    # N2 decks: orbitals 1 pi u, 1 sigma g, 1 sigma u, 1 sigma g, 1 sigma u, 1 sigma g;
    # ihomo is -1 for sigma-u and pi-g functions (initCBlocks, initCBAllocArrays.F90:981-987)
    ihomo = {1: [1], 6: [1, 1, -1, 1, -1, 1]}[len(orbs)]
    for k, o in enumerate(orbs):
        gpu.check(ctx.lib.x2g_setci(g.h, o.ptr, ihomo[k]))


@pytest.mark.parametrize("eng", engines())
def test_potential_batch_vs_reference(ctx, eng):
    """All Coulomb and exchange potentials of the N2 pi orbital (7 problems, both
    parities, Delta m = 0, 1, 2) relaxed in one batch = the reference's
    coulExchMCSOR (relaxPots.F90:196-322)."""
    from x2dhf_amd import gpu
    set_engine(ctx, eng)
    for case in ("h2_61x91", "n2_61x91"):
        z = load_golden(case)
        print(f"case={case} engine={eng}", flush=True)
        g = make_grid(ctx, z)
        norb = int(z["sizes"][2])
        orbs = [ctx.upload(z["orb_prep"][k]) for k in range(norb)]
        _setci_all(ctx, g, z, orbs)
        f3, e, gg = z["suppl"][7], z["suppl"][3], z["suppl"][9]
        omega = float(z["scal"][5])
        items, keep = [], []
        orb_h = [o.get() for o in orbs]
        for k, (in1, in2, dm, isym, _) in enumerate(z["potlist"]):
            rhs = orb_h[in1 - 1] * orb_h[in2 - 1] * gg
            lhs = f3 + float(dm * dm) * e if dm else f3.copy()
            fun, l, r = ctx.upload(z["pots_asympt"][k]), ctx.upload(lhs), ctx.upload(rhs)
            keep += [fun, l, r]
            items.append(gpu.RelaxItem(fun.ptr, l.ptr, r.ptr, None, omega, int(isym), 1, 0, 0))
        arr = (gpu.RelaxItem * len(items))(*items)
        grids = (gpu._vp * 1)(g.h.value)
        gpu.check(ctx.lib.x2g_relax_batch(ctx.h, grids, arr, len(items), 1, 10))
        for k in range(len(items)):
            got = keep[3 * k].get()
            err = relerr(got, z["pots_relaxed"][k])
            assert err < 1e-11, (case, k, err)
        g.destroy()


@pytest.mark.parametrize("eng", engines())
def test_orbital_relaxation_vs_reference(ctx, eng):
    """orbMCSOR (relaxOrbs.F90:150-241): lhs = fock1+diag, rhs = fock2, ten
    sweeps, then the exponential tail of the orbital."""
    from x2dhf_amd import gpu
    set_engine(ctx, eng)
    for case, isym in (("h2_61x91", 1), ("n2_61x91", -1)):
        z = load_golden(case)
        g = make_grid(ctx, z)
        norb, tr = int(z["sizes"][2]), int(z["sizes"][3])
        orbs = [ctx.upload(z["orb_prep"][k]) for k in range(norb)]
        _setci_all(ctx, g, z, orbs)
        diag = float(z["scal"][2])
        lhs, rhs, f1 = ctx.upload(z["fock1"] + diag), ctx.upload(z["fock2"]), ctx.upload(z["fock1"])
        psi = orbs[tr - 1]
        it = gpu.RelaxItem(psi.ptr, lhs.ptr, rhs.ptr, f1.ptr, float(z["scal"][3]), isym, 0, 0, 0)
        arr = (gpu.RelaxItem * 1)(it)
        grids = (gpu._vp * 1)(g.h.value)
        gpu.check(ctx.lib.x2g_relax_batch(ctx.h, grids, arr, 1, 1, 10))
        err = relerr(psi.get(), z["orb_relaxed"])
        assert err < 1e-11, (case, err)
        g.destroy()


def test_fock_lincomb_vs_reference(ctx):
    """fockHF (fock.F90:57-184) assembled with x2g_lincomb on the relaxed potentials."""
    from x2dhf_amd import gpu
    # H2: fock1 = f0 + eps f1 + f2 * (occ-1) J
    z = load_golden("h2_61x91")
    g = make_grid(ctx, z)
    S = [g.suppl_ptr(k) for k in range(14)]
    J = ctx.upload(z["pots_relaxed"][0])
    out = ctx.alloc(g.n)
    eps = float(z["ee_prep"][0, 0])
    terms = (gpu.Term * 2)(gpu.Term(eps, S[5], None, None), gpu.Term(1.0, S[6], J.ptr, None))
    gpu.check(ctx.lib.x2g_lincomb(g.h, out.ptr, S[4], 1.0, terms, 2))
    assert relerr(out.get(), z["fock1"]) < 1e-13
    g.destroy()
    # N2, pi orbital 1: Coulomb with occupations 3 (own shell) and 2; exchange
    # with every sigma (gec = 1) and the pi self-exchange (Delta m = 2, gec = 1)
    z = load_golden("n2_61x91")
    g = make_grid(ctx, z)
    S = [g.suppl_ptr(k) for k in range(14)]
    orbs = [ctx.upload(z["orb_prep"][k]) for k in range(6)]
    _setci_all(ctx, g, z, orbs)
    P = [ctx.upload(p) for p in z["pots_relaxed"]]   # (1,1),(1,2)..(1,6), Coulomb 1
    C = [P[6]] + [ctx.upload(z["coul_prep"][k]) for k in range(1, 6)]
    coul = ctx.alloc(g.n)
    t = [gpu.Term(3.0, C[0].ptr, None, None)] + [gpu.Term(2.0, C[k].ptr, None, None) for k in range(1, 6)]
    gpu.check(ctx.lib.x2g_lincomb(g.h, coul.ptr, None, 0.0, (gpu.Term * 6)(*t), 6))
    assert relerr(coul.get(), z["coulombptr"]) < 1e-13
    exch = ctx.alloc(g.n)
    t = [gpu.Term(1.0, P[k].ptr, orbs[k].ptr, None) for k in range(1, 6)] + [gpu.Term(1.0, P[0].ptr, orbs[0].ptr, None)]
    gpu.check(ctx.lib.x2g_lincomb(g.h, exch.ptr, None, 0.0, (gpu.Term * 6)(*t), 6))
    assert relerr(exch.get(), z["exchangeptr"]) < 1e-13
    f1 = ctx.alloc(g.n)
    eps = float(z["ee_prep"][0, 0])
    t = [gpu.Term(eps, S[5], None, None), gpu.Term(1.0, S[3], None, None), gpu.Term(1.0, S[6], coul.ptr, None)]
    gpu.check(ctx.lib.x2g_lincomb(g.h, f1.ptr, S[4], 1.0, (gpu.Term * 3)(*t), 3))
    assert relerr(f1.get(), z["fock1"]) < 1e-13
    f2 = ctx.alloc(g.n)
    t = [gpu.Term(1.0, S[6], exch.ptr, None)]
    gpu.check(ctx.lib.x2g_lincomb(g.h, f2.ptr, None, 0.0, (gpu.Term * 1)(*t), 1))
    assert relerr(f2.get(), z["fock2"]) < 1e-13
    g.destroy()


@pytest.mark.parametrize("case", ["h2_61x91", "n2_61x91"])
def test_quadratures_vs_reference(ctx, case):
    """norm (normOrtho.F90:41-50), oneelij (integrals.F90:406-480) and the
    multipole moments (scfUtils.F90:490-850) of the traced orbital."""
    from x2dhf_amd import gpu
    z = load_golden(case)
    g = make_grid(ctx, z)
    S = [g.suppl_ptr(k) for k in range(14)]
    norb, tr = int(z["sizes"][2]), int(z["sizes"][3])
    psi = ctx.upload(z["orb_relaxed"])
    out = np.zeros(3)
    d = (gpu.Dot * 1)(gpu.Dot(S[13], S[8], psi.ptr, psi.ptr))
    gpu.check(ctx.lib.x2g_wdot(g.h, d, 1, out.ctypes.data))
    assert abs(out[0] - z["norm_1"][tr - 1]) < 1e-13 * abs(z["norm_1"][tr - 1])
    # one-electron integral on the orthonormalised orbital
    isym, m = (1, 0) if case.startswith("h2") else (-1, 1)
    po = ctx.upload(z["orb_ortho"])
    gpu.check(ctx.lib.x2g_oneel(g.h, po.ptr, po.ptr, isym, m, out.ctypes.data))
    assert abs(out[0] - z["oneel_1"][0]) < 1e-11 * abs(z["oneel_1"][0]), (out[0], z["oneel_1"][0])
    # Coulomb moments of the traced orbital (even ones; odd vanish for homonuclear)
    mom = np.zeros(8)
    gpu.check(ctx.lib.x2g_moments(g.h, po.ptr, po.ptr, 0, 4, mom.ctypes.data))
    ref = z["cmulti_1"][:, tr - 1]
    for k in (1, 3):
        assert abs(mom[k] - ref[k]) < 1e-11 * max(abs(ref[k]), 1e-3), (k, mom[k], ref[k])
    if case.startswith("n2"):
        # exchange moments of the pairs (1, j): pair index k2(1,j) = 1 + j(j-1)/2
        orbs = [ctx.upload(z["orb_prep"][k]) for k in range(norb)]
        _setci_all(ctx, g, z, orbs)
        orbs[0] = po
        for j in range(1, 7):
            ipc = 1 + j * (j - 1) // 2
            dm = 2 if j == 1 else 1
            gpu.check(ctx.lib.x2g_moments(g.h, orbs[0].ptr, orbs[j - 1].ptr, dm, 4, mom.ctypes.data))
            for k, name in enumerate(("excdi", "excqu", "excoc", "exche")):
                r = z[name + "_1"][ipc - 1]
                assert abs(mom[k] - r) < 1e-10 * max(abs(r), 1e-4), (j, name, mom[k], r)
    g.destroy()


@pytest.mark.parametrize("case", ["h2_61x91", "n2_61x91"])
def test_potential_tails_vs_reference(ctx, case):
    """coulAsympt / exchAsympt (scfUtils.F90:275-329) from the moments of prepSCF."""
    from x2dhf_amd import gpu
    z = load_golden(case)
    g = make_grid(ctx, z)
    nni, nmu = g.nni, g.nmu
    for k, (in1, in2, dm, isym, _) in enumerate(z["potlist"]):
        pot = ctx.upload(z["pots_prep"][k])
        if in1 == in2 and dm == 0:   # Coulomb potential of orbital in1
            mom = np.ascontiguousarray(z["cmulti_prep"][:, in1 - 1])
            gpu.check(ctx.lib.x2g_pot_tail(g.h, pot.ptr, 0, 0, 4, mom.ctypes.data))
        else:
            ipc = in1 + in2 * (in2 - 1) // 2
            mom = np.array([z[n + "_prep"][ipc - 1] for n in ("excdi", "excqu", "excoc", "exche")])
            gpu.check(ctx.lib.x2g_pot_tail(g.h, pot.ptr, 1, int(dm), 4, mom.ctypes.data))
        got = pot.get().reshape(nmu, nni)[nmu - 4:]
        ref = z["pots_asympt"][k].reshape(nmu, nni)[nmu - 4:]
        assert relerr(got, ref) < 1e-12, (case, k)
    g.destroy()


def test_setci(ctx):
    from x2dhf_amd import gpu
    z = load_golden("h2_61x91")
    g = make_grid(ctx, z)
    rng = np.random.default_rng(5)
    a = rng.standard_normal(g.n)
    for ih in (1, -1):
        d = ctx.upload(a)
        gpu.check(ctx.lib.x2g_setci(g.h, d.ptr, ih))
        A = a.reshape(g.nmu, g.nni).copy()
        n2 = g.nni // 2 + 1
        for ni in range(1, n2):
            A[:, g.nni - ni] = ih * A[:, ni - 1]
        if ih == -1:
            A[:, n2 - 1] = 0
        assert np.array_equal(d.get(), A.ravel())
    g.destroy()


# ---------------------------------------------------------------------------
# Seeded synthetic problems at the configuration sizes (oracle comparison).

def synthetic_problem(nni, nmu, seed, rinf=40.0, r=2.0):
    """Grid-like coefficients (bpot(mu), d(nu), f3-type lhs) and a smooth rhs."""
    rng = np.random.default_rng(seed)
    hni = np.pi / (nni - 1)
    xmi0 = 2.0 * rinf / r
    hmu = np.log(xmi0 + np.sqrt(xmi0 * xmi0 - 1)) / (nmu - 1)
    aa1 = np.array([3.0, -32.0, 168.0, -672.0])
    aa2 = np.array([-9.0, 128.0, -1008.0, 8064.0])
    coef16 = np.concatenate([aa1 / 840 / hmu, aa2 / 5040 / hmu ** 2, aa1 / 840 / hni, aa2 / 5040 / hni ** 2])
    ex = np.array([210.0, -120.0, 45.0, -10.0, 1.0]) / 126.0
    diag = -14350.0 / 5040.0 * (1 / hmu ** 2 + 1 / hni ** 2)
    vmu = np.arange(nmu) * hmu
    vxi = np.cosh(vmu)
    vni = np.arange(nni) * hni
    veta = np.cos(vni)
    x1 = np.sqrt(np.maximum(vxi ** 2 - 1, 0))
    w5 = np.where(np.arange(nmu) == 0, 0.0, vxi / np.where(x1 == 0, 1, x1) - 2 * x1 / vxi)
    e1 = np.sqrt(1 - veta ** 2)
    d1 = np.where(e1 < 1e-15, 0.0, veta / np.where(e1 == 0, 1, e1))
    suppl = np.zeros((14, nni * nmu))
    suppl[1] = np.repeat(w5, nni)
    suppl[2] = np.tile(d1, nmu)
    suppl[0] = np.repeat(np.where(np.arange(nmu) == 0, 0.0, vxi / np.where(x1 == 0, 1, x1)), nni)
    lhs = (-2.0 / np.repeat(vxi ** 2, nni) + diag) * (1 + 0.05 * rng.random(nni * nmu))
    X, Y = np.meshgrid(vni, vmu)
    rhs = (np.exp(-((X - 1.0) ** 2 + (Y - 1.5) ** 2)) * np.sin(3 * X) + 0.01 * rng.standard_normal(X.shape)).ravel()
    fun = 0.1 * rng.standard_normal(nni * nmu)
    coef = np.concatenate([coef16, ex, [diag, hni, hmu, r]])
    return dict(nni=nni, nmu=nmu, coef=coef, suppl=suppl, vxi=vxi, veta=veta, lhs=lhs, rhs=rhs, fun=fun)


@pytest.mark.parametrize("shape", [(91, 121), (169, 241), (331, 601), (481, 901), (505, 901), (721, 241)])
@pytest.mark.parametrize("isym", [1, -1])
@pytest.mark.parametrize("engine", [(0, -1, 4), (0, -1, 2), (2, -1, 4)])
def test_relax_batch_large_vs_oracle(ctx, shape, isym, engine):
    """Batched relaxation (maxsor1=2 x maxsor2=10) at configuration sizes against the
    oracle's putin -> mcsor -> putout macro-iterations.  FP64 grids wider than the
    default kernel's LDS ring (481, 505) run the first strip kernel with one pass
    per barrier; 505 (nni = 0 mod 5) has another colour multiplier."""
    from x2dhf_amd import gpu
    nni, nmu = shape
    print(f"shape={shape} isym={isym} engine={engine}", flush=True)
    set_engine(ctx, engine)
    P = synthetic_problem(nni, nmu, seed=nni + nmu)
    g = gpu.Grid(ctx, nni, nmu, P["coef"], P["suppl"], P["vxi"], P["veta"])
    fun, l, r = ctx.upload(P["fun"]), ctx.upload(P["lhs"]), ctx.upload(P["rhs"])
    it = gpu.RelaxItem(fun.ptr, l.ptr, r.ptr, None, 1.9, isym, 1, 0, 0)
    gpu.check(ctx.lib.x2g_relax_batch(ctx.h, (gpu._vp * 1)(g.h.value), (gpu.RelaxItem * 1)(it), 1, 2, 10))
    got = fun.get()
    f = P["fun"].copy()
    for _ in range(2):
        w = pyoracle.putin(nni, nmu, isym, f)
        pyoracle.mcsor(nni, nmu, isym, w, P["lhs"], P["rhs"], P["suppl"][1], P["suppl"][2], P["coef"][:16],
                       P["coef"][16:21], 1.9, 10)
        f = pyoracle.putout(nni, nmu, w)
    assert relerr(got, f) < 1e-11
    g.destroy()


def test_uneven_batch_matches_single_problems(ctx):
    """Problems on four different grids relaxed in one batch give the same
    result as each relaxed alone."""
    from x2dhf_amd import gpu
    set_engine(ctx, (0, -1, 4))
    shapes = [(91, 121), (151, 241), (61, 91), (331, 601)]
    probs = [synthetic_problem(n, m, seed=7 + i) for i, (n, m) in enumerate(shapes)]
    grids = [gpu.Grid(ctx, p["nni"], p["nmu"], p["coef"], p["suppl"], p["vxi"], p["veta"]) for p in probs]
    gh = (gpu._vp * len(grids))(*[g.h.value for g in grids])
    bufs, items = [], []
    for i, p in enumerate(probs):
        f, l, r = ctx.upload(p["fun"]), ctx.upload(p["lhs"]), ctx.upload(p["rhs"])
        bufs.append((f, l, r))
        items.append(gpu.RelaxItem(f.ptr, l.ptr, r.ptr, None, 1.85, 1 if i % 2 == 0 else -1, 1, 0, i))
    gpu.check(ctx.lib.x2g_relax_batch(ctx.h, gh, (gpu.RelaxItem * 4)(*items), 4, 1, 10))
    together = [b[0].get() for b in bufs]
    for i, p in enumerate(probs):
        f = ctx.upload(p["fun"])
        it = gpu.RelaxItem(f.ptr, bufs[i][1].ptr, bufs[i][2].ptr, None, 1.85, items[i].isym, 1, 0, 0)
        gpu.check(ctx.lib.x2g_relax_batch(ctx.h, (gpu._vp * 1)(grids[i].h.value), (gpu.RelaxItem * 1)(it), 1, 1, 10))
        assert np.array_equal(f.get(), together[i]), i
    for g in grids:
        g.destroy()


@pytest.mark.parametrize("shape", [(91, 121), (331, 601), (505, 901), (523, 151), (721, 241)])
@pytest.mark.parametrize("engine", [(0, -1, 4), (0, -1, 2), (2, -1, 4)])
def test_fp32_correction_sweeps_vs_oracle(ctx, shape, engine):
    """With u = 0 one mixed-precision macro-iteration is exactly FP32 MCSOR from
    zero on rhs (the residual), so its interior must match the FP32 oracle
    (grids wider than 512 points: the widest ring class, two points per pass
    lane, and the per-point residual kernel)."""
    from x2dhf_amd import gpu
    nni, nmu = shape
    print(f"shape={shape} engine={engine}", flush=True)
    set_engine(ctx, engine)
    P = synthetic_problem(nni, nmu, seed=11)
    g = gpu.Grid(ctx, nni, nmu, P["coef"], P["suppl"], P["vxi"], P["veta"])
    f = ctx.upload(np.zeros(nni * nmu))
    l, r = ctx.upload(P["lhs"]), ctx.upload(P["rhs"])
    it = gpu.RelaxItem(f.ptr, l.ptr, r.ptr, None, 1.9, 1, 1, 1, 0)
    gpu.check(ctx.lib.x2g_relax_batch(ctx.h, (gpu._vp * 1)(g.h.value), (gpu.RelaxItem * 1)(it), 1, 1, 10))
    got = f.get().reshape(nmu, nni)
    w = np.zeros((nni + 8) * (nmu + 8), dtype=np.float32)
    pyoracle.mcsor_f32(nni, nmu, 1, w, P["lhs"], P["rhs"], P["suppl"][1], P["suppl"][2], P["coef"][:16],
                       P["coef"][16:21], 1.9, 10)
    ref = w.reshape(nmu + 8, nni + 8)[4:nmu + 4, 4:nni + 4].astype(np.float64)
    inter = (slice(1, nmu - 4), slice(1, nni - 1))
    assert relerr(got[inter], ref[inter]) < 2e-5
    g.destroy()


def test_mixed_precision_refines_to_fp64(ctx):
    """FP32 correction sweeps with FP64 residuals converge to the FP64 solution:
    after enough macro-iterations the FP64 residual is at FP64 level."""
    from x2dhf_amd import gpu
    set_engine(ctx, (0, -1, 4))
    nni, nmu = 169, 241
    P = synthetic_problem(nni, nmu, seed=13)
    g = gpu.Grid(ctx, nni, nmu, P["coef"], P["suppl"], P["vxi"], P["veta"])
    gh = (gpu._vp * 1)(g.h.value)
    l, r = ctx.upload(P["lhs"]), ctx.upload(P["rhs"])
    res = np.zeros(2)
    out = {}
    for prec in (0, 1):
        f = ctx.upload(P["fun"])
        it = (gpu.RelaxItem * 1)(gpu.RelaxItem(f.ptr, l.ptr, r.ptr, None, 1.9, 1, 1, prec, 0))
        gpu.check(ctx.lib.x2g_relax_batch(ctx.h, gh, it, 1, 200, 10))
        gpu.check(ctx.lib.x2g_residual_norms(ctx.h, gh, it, 1, res.ctypes.data))
        out[prec] = (f.get(), np.sqrt(res[0] / res[1]))
    print("relative residuals fp64 / mixed:", out[0][1], out[1][1])
    assert out[0][1] < 1e-10 and out[1][1] < 1e-10, (out[0][1], out[1][1])
    assert relerr(out[1][0], out[0][0]) < 1e-10
    g.destroy()


@pytest.mark.parametrize("nni,nmu,isym", [(91, 121, 1), (169, 241, -1), (331, 601, 1), (505, 901, -1), (523, 151, 1)])
def test_residual_norms_vs_oracle(ctx, nni, nmu, isym):
    """The FP64 residual rhs - (L + lhs) u of the correction equations (both
    residual kernels: the row kernel for nni <= 512, the per-point kernel for
    wider grids): its squared norm against the oracle's residual."""
    from x2dhf_amd import gpu
    from oracle import pyoracle
    P = synthetic_problem(nni, nmu, seed=nni + nmu)
    print(f"residual {nni}x{nmu} isym={isym}", flush=True)
    g = gpu.Grid(ctx, nni, nmu, P["coef"], P["suppl"], P["vxi"], P["veta"])
    gh = (gpu._vp * 1)(g.h.value)
    f, l, r = ctx.upload(P["fun"]), ctx.upload(P["lhs"]), ctx.upload(P["rhs"])
    it = (gpu.RelaxItem * 1)(gpu.RelaxItem(f.ptr, l.ptr, r.ptr, None, 1.9, isym, 1, 1, 0))
    out = np.zeros(2)
    gpu.check(ctx.lib.x2g_residual_norms(ctx.h, gh, it, 1, out.ctypes.data))
    w = pyoracle.putin(nni, nmu, isym, P["fun"])
    res = pyoracle.residual(nni, nmu, w, P["lhs"], P["rhs"], P["suppl"][1], P["suppl"][2], P["coef"][:16])
    rhs2 = P["rhs"].reshape(nmu, nni)[1:nmu - 4, 1:nni - 1]
    assert abs(out[0] - np.sum(res * res)) <= 1e-12 * np.sum(res * res)
    assert abs(out[1] - np.sum(rhs2 * rhs2)) <= 1e-12 * np.sum(rhs2 * rhs2)
    g.destroy()


@pytest.mark.parametrize("nni,nmu,nf", [(169, 241, 3), (331, 601, 2), (523, 151, 3)])
def test_rhs_as_product_of_factors(ctx, nni, nmu, nf):
    """An item may give its right-hand side as rhs_f1*rhs_f2*rhs_f3 (the
    potentials' psi_i psi_j g): FP32-correction relaxation (the residual forms
    the product; both residual kernels) and FP64 relaxation (the product is
    formed into rhs_d first) agree with the same item given rhs = the product
    computed by x2g_lincomb, bit for bit, and so do the residual norms."""
    from x2dhf_amd import gpu
    set_engine(ctx, (0, -1, 4))
    P = synthetic_problem(nni, nmu, seed=7 * nni + nf)
    rng = np.random.default_rng(nni)
    print(f"rhs factors {nni}x{nmu} nf={nf}", flush=True)
    g = gpu.Grid(ctx, nni, nmu, P["coef"], P["suppl"], P["vxi"], P["veta"])
    gh = (gpu._vp * 1)(g.h.value)
    fac = [ctx.upload(1.0 + 0.1 * rng.standard_normal(nni * nmu)) for _ in range(nf)]
    prod = ctx.upload(np.zeros(nni * nmu))
    t = (gpu.Term * 1)(gpu.Term(1.0, fac[0].ptr, fac[1].ptr, fac[2].ptr if nf == 3 else None))
    gpu.check(ctx.lib.x2g_lincomb(g.h, prod.ptr, None, C.c_double(0.0), t, 1))
    l = ctx.upload(P["lhs"])
    for prec in (1, 0):
        out = {}
        for mode in ("rhs", "factors"):
            f = ctx.upload(P["fun"])
            scratch = ctx.upload(np.zeros(nni * nmu))
            it = gpu.RelaxItem(f.ptr, l.ptr, prod.ptr if mode == "rhs" else scratch.ptr, None, 1.9, 1, 1, prec, 0)
            if mode == "factors":
                it.rhs_f1_d, it.rhs_f2_d = fac[0].ptr, fac[1].ptr
                it.rhs_f3_d = fac[2].ptr if nf == 3 else None
            its = (gpu.RelaxItem * 1)(it)
            norms = np.zeros(2)
            gpu.check(ctx.lib.x2g_residual_norms(ctx.h, gh, its, 1, norms.ctypes.data))
            gpu.check(ctx.lib.x2g_relax_batch(ctx.h, gh, its, 1, 2, 10))
            out[mode] = (f.get(), norms)
        print(f"prec {prec}: max |du| {np.max(np.abs(out['rhs'][0] - out['factors'][0])):.3e}")
        # the norms are summed with atomics (order varies between calls)
        assert np.allclose(out["rhs"][1], out["factors"][1], rtol=1e-13, atol=0.0)
        assert np.array_equal(out["rhs"][0], out["factors"][0])
    g.destroy()


def test_zero_sweeps_leave_the_function_unchanged(ctx):
    """maxsor2 = 0: no sweep runs, so the function comes back bit for bit,
    in FP64 and with the FP32 correction (whose start array is then never
    cleared or read)."""
    from x2dhf_amd import gpu
    set_engine(ctx, (0, -1, 4))
    nni, nmu = 169, 241
    P = synthetic_problem(nni, nmu, seed=3)
    g = gpu.Grid(ctx, nni, nmu, P["coef"], P["suppl"], P["vxi"], P["veta"])
    gh = (gpu._vp * 1)(g.h.value)
    l, r = ctx.upload(P["lhs"]), ctx.upload(P["rhs"])
    for prec in (0, 1):
        f = ctx.upload(P["fun"])
        it = (gpu.RelaxItem * 1)(gpu.RelaxItem(f.ptr, l.ptr, r.ptr, None, 1.9, 1, 1, prec, 0))
        gpu.check(ctx.lib.x2g_relax_batch(ctx.h, gh, it, 1, 2, 0))
        assert np.array_equal(f.get(), P["fun"]), prec
    g.destroy()


@pytest.mark.parametrize("case", ["h2_61x91", "n2_61x91"])
def test_mcsor_host_dropin_vs_reference(ctx, case):
    """x2g_mcsor_host -- the call the reference's orbMCSOR makes in place of
    mcsor (host arrays in and out, relaxOrbs.F90:393-395) -- on the reference's
    own padded input reproduces its mcsor output."""
    z = load_golden(case)
    g = make_grid(ctx, z)
    isym = int(z["k1_isym"][0])
    omega, nsw = float(z["k1_scal"][0]), int(z["k1_scal"][1])
    for name, ns in (("wout_mcsor", nsw), ("wout_mcsor1", 1)):
        work = np.ascontiguousarray(z["k1_win"], dtype=np.float64).copy()
        lhs = np.ascontiguousarray(z["k1_lhs"], dtype=np.float64)
        rhs = np.ascontiguousarray(z["k1_rhs"], dtype=np.float64)
        rc = ctx.lib.x2g_mcsor_host(g.h, isym, 1, work.ctypes.data, lhs.ctypes.data, rhs.ctypes.data,
                                    omega, ns)
        assert rc == 0, ctx.lib.x2g_last_error()
        e = relerr(work, z["k1_" + name])
        print(f"{case} x2g_mcsor_host {ns} sweeps: rel err {e:.2e}")
        assert e < REL64, (name, e)
    g.destroy()


@pytest.mark.parametrize("case", ["h2_61x91", "n2_61x91", "h2_61x91_middle", "n2_61x91_middle"])
@pytest.mark.parametrize("pfx", ["k1_", "k2_"])
def test_ordered_relaxation_bitexact_vs_reference(ctx, case, pfx):
    """x2g_relax_ordered: sor (10 sweeps, forward and backward) and one-thread
    mcsor (10 and 1 sweeps) in the col-wise and the reference's default middle ordering
    reproduce the reference's own routines bit for bit (the level schedule
    keeps its sequential order of dependencies; no fused multiply-adds), with
    the sweeps pipelined and one level of one sweep at a time."""
    from x2dhf_amd import gpu
    z = load_golden(case)
    if pfx + "win" not in z:
        pytest.skip("single potential")
    ordering = 1 if case.endswith("_middle") else 0
    print(f"case={case} {pfx} ordering={ordering}", flush=True)
    g = make_grid(ctx, z)
    isym = int(z[pfx + "isym"][0])
    omega, nsw = float(z[pfx + "scal"][0]), int(z[pfx + "scal"][1])
    lhs, rhs = ctx.upload(z[pfx + "lhs"]), ctx.upload(z[pfx + "rhs"])
    try:
        for pipe in (1, 0):
            gpu.check(ctx.lib.x2g_set_levels_pipeline(pipe))
            for method, name, ns in ((1, "wout_sor", nsw), (2, "wout_sor_bwd", nsw), (0, "wout_mcsor", nsw),
                                     (0, "wout_mcsor1", 1)):
                w = ctx.upload(z[pfx + "win"])
                gpu.check(ctx.lib.x2g_relax_ordered(g.h, isym, 1, w.ptr, lhs.ptr, rhs.ptr, omega, ns, method,
                                                    ordering))
                got = w.get()
                assert np.array_equal(got, z[pfx + name]), (pipe, name, relerr(got, z[pfx + name]))
                w.free()
    finally:
        gpu.check(ctx.lib.x2g_set_levels_pipeline(1))
    g.destroy()


ORDER_CODES = {"middlemt": 2, "row-wise": 3, "rrow-wis": 4, "test1": 5, "test2": 6, "test3": 7, "test4": 8,
               "middle2": 9}


@pytest.mark.parametrize("ordering", list(ORDER_CODES))
def test_other_orderings_bitexact_vs_reference(ctx, ordering):
    """x2g_relax_ordered in each of the reference's other mesh orderings:
    sor (10 sweeps, forward and backward) and one-thread mcsor (1 sweep) on
    the H2 61x91 deck with that `order` card reproduce the reference's own routines bit for bit
    (tests/golden/h2_61x91_orderings.npz)."""
    import os
    from conftest import GOLDEN
    from x2dhf_amd import gpu
    z = load_golden("h2_61x91_middle")
    v = np.load(os.path.join(GOLDEN, "h2_61x91_orderings.npz"))
    vec = lambda k: v[f"{ordering}/k1_{k}"]
    print(f"ordering={ordering}", flush=True)
    g = make_grid(ctx, z)
    isym = int(vec("isym")[0])
    omega, nsw = float(vec("scal")[0]), int(vec("scal")[1])
    lhs, rhs = ctx.upload(vec("lhs")), ctx.upload(vec("rhs"))
    try:
        for pipe in (1, 0):
            gpu.check(ctx.lib.x2g_set_levels_pipeline(pipe))
            for method, name, ns in ((1, "wout_sor", nsw), (2, "wout_sor_bwd", nsw), (0, "wout_mcsor1", 1)):
                w = ctx.upload(vec("win"))
                gpu.check(ctx.lib.x2g_relax_ordered(g.h, isym, 1, w.ptr, lhs.ptr, rhs.ptr, omega, ns, method,
                                                    ORDER_CODES[ordering]))
                got = w.get()
                assert np.array_equal(got, vec(name)), (pipe, name, relerr(got, vec(name)))
                w.free()
    finally:
        gpu.check(ctx.lib.x2g_set_levels_pipeline(1))
    g.destroy()


@pytest.mark.parametrize("shape", [(169, 241), (331, 601), (523, 151)])
def test_pipelined_ordered_sweeps_vs_oracle(ctx, shape):
    """The pipelined ordered sweeps at configuration sizes: sor forward and
    backward (10 sweeps) and one-thread mcsor (3 sweeps), even and odd
    symmetry, in the default middle, the col-wise and the test2 orderings,
    equal the oracle's sequential routine bit for bit (and so does one level
    of one sweep at a time, on the middle ordering)."""
    from x2dhf_amd import gpu
    nni, nmu = shape
    P = synthetic_problem(nni, nmu, seed=3 * nni + nmu)
    g = gpu.Grid(ctx, nni, nmu, P["coef"], P["suppl"], P["vxi"], P["veta"])
    lhs, rhs = ctx.upload(P["lhs"]), ctx.upload(P["rhs"])
    args = (P["lhs"], P["rhs"], P["suppl"][1], P["suppl"][2], P["coef"][:16], P["coef"][16:21], 1.8)
    try:
        for pipe, orderings in ((1, ("middle", "col-wise", "test2")), (0, ("middle",))):
            gpu.check(ctx.lib.x2g_set_levels_pipeline(pipe))
            for ordering in orderings:
                for isym in (1, -1):
                    win = pyoracle.putin(nni, nmu, isym, P["fun"])
                    for method, ns in ((1, 10), (2, 10), (0, 3)):
                        print(f"shape={shape} pipe={pipe} ordering={ordering} isym={isym} method={method}",
                              flush=True)
                        want = pyoracle.relax_ordered(nni, nmu, isym, win.copy(), *args, ns, method, ordering)
                        w = ctx.upload(win)
                        gpu.check(ctx.lib.x2g_relax_ordered(g.h, isym, 1, w.ptr, lhs.ptr, rhs.ptr, 1.8, ns, method,
                                                            pyoracle.ORDERINGS[ordering]))
                        got = w.get()
                        w.free()
                        assert np.array_equal(got, want), (ordering, isym, method, relerr(got, want))
    finally:
        gpu.check(ctx.lib.x2g_set_levels_pipeline(1))
    g.destroy()


def test_pipelined_ordered_sweeps_longer_than_one_pipeline(ctx):
    """More sweeps than one pipeline holds (64): the kernel runs 64, drains,
    then the remaining 6 -- still the oracle's sequential sor bit for bit
    (middle ordering, forward and backward, 91x121)."""
    from x2dhf_amd import gpu
    nni, nmu = 91, 121
    P = synthetic_problem(nni, nmu, seed=11)
    g = gpu.Grid(ctx, nni, nmu, P["coef"], P["suppl"], P["vxi"], P["veta"])
    lhs, rhs = ctx.upload(P["lhs"]), ctx.upload(P["rhs"])
    args = (P["lhs"], P["rhs"], P["suppl"][1], P["suppl"][2], P["coef"][:16], P["coef"][16:21], 1.7)
    gpu.check(ctx.lib.x2g_set_levels_pipeline(1))
    win = pyoracle.putin(nni, nmu, 1, P["fun"])
    for method in (1, 2):
        want = pyoracle.relax_ordered(nni, nmu, 1, win.copy(), *args, 70, method, "middle")
        w = ctx.upload(win)
        gpu.check(ctx.lib.x2g_relax_ordered(g.h, 1, 1, w.ptr, lhs.ptr, rhs.ptr, 1.7, 70, method, 1))
        got = w.get()
        w.free()
        assert np.array_equal(got, want), (method, relerr(got, want))
    g.destroy()


def test_ordered_potential_batch_vs_oracle(ctx):
    """All potentials of the N2 pi orbital in one x2g_relax_batch call with
    SOR in the middle ordering (each a workgroup of the level kernel; every
    other potential with the backward sweeps of the altsweeps option), two
    macro-iterations of putin, 10 sweeps, putout: equal to the oracle's
    sequential sor on each, bit for bit."""
    from x2dhf_amd import gpu
    z = load_golden("n2_61x91")
    nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
    g = make_grid(ctx, z)
    f3, e, gg = z["suppl"][7], z["suppl"][3], z["suppl"][9]
    omega = float(z["scal"][5])
    orb = z["orb_prep"]
    items, keep, want = [], [], []
    for in1, in2, dm, isym, _ in z["potlist"]:
        rhs = orb[in1 - 1] * orb[in2 - 1] * gg
        lhs = f3 + float(dm * dm) * e if dm else f3.copy()
        k = len(want)
        method = 1 if k % 2 == 0 else 2
        fun0 = z["pots_asympt"][k]
        w = fun0.copy()
        for _ in range(2):
            wp = pyoracle.putin(nni, nmu, int(isym), w)
            pyoracle.relax_ordered(nni, nmu, int(isym), wp, lhs, rhs, z["suppl"][1], z["suppl"][2], z["coef"][:16],
                                   z["coef"][16:21], omega, 10, method, "middle")
            w = pyoracle.putout(nni, nmu, wp)
        want.append(w)
        fun, l, r = ctx.upload(fun0), ctx.upload(lhs), ctx.upload(rhs)
        keep += [fun, l, r]
        it = gpu.RelaxItem(fun.ptr, l.ptr, r.ptr, None, omega, int(isym), 1, 0, 0)
        it.method, it.ordering = method, 1
        items.append(it)
    arr = (gpu.RelaxItem * len(items))(*items)
    gh = (C.c_void_p * 1)(g.h.value)
    gpu.check(ctx.lib.x2g_relax_batch(ctx.h, gh, arr, len(items), 2, 10))
    for k, w in enumerate(want):
        got = keep[3 * k].get()
        assert np.array_equal(got, w), (k, relerr(got, w))
    g.destroy()
