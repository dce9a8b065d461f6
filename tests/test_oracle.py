"""Pin the CPU oracle (oracle/x2oracle.c) before trusting it.

* Against the reference's own routines: vectors written by the reference
  mcsor/sor/putin (sormcsor.F90, inout.F90) through oracle/ref_harness.f90.
* Against what the mathematics fixes: the 9-point stencil weights, the
  boundary extrapolation weights and the 7-point Newton-Cotes weights on
  polynomials, and the fixed point of the sweep (zero residual).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden
from oracle import pyoracle


def _ok(a, b):
    return np.array_equal(a, b)


@pytest.mark.parametrize("case", ["h2_61x91", "n2_61x91"])
@pytest.mark.parametrize("pfx", ["k1_", "k2_"])
def test_oracle_sweeps_bitexact_vs_reference(case, pfx):
    z = load_golden(case)
    if pfx + "win" not in z:
        pytest.skip("case has a single potential")
    nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
    isym = int(z[pfx + "isym"][0])
    omega, nsw = float(z[pfx + "scal"][0]), int(z[pfx + "scal"][1])
    args = (z[pfx + "lhs"], z[pfx + "rhs"], z["suppl"][1], z["suppl"][2], z["coef"][:16], z["coef"][16:21], omega)
    for mode, name, ns in [(0, "wout_mcsor", nsw), (0, "wout_mcsor1", 1), (1, "wout_sor", nsw)]:
        w = pyoracle.mcsor(nni, nmu, isym, z[pfx + "win"].copy(), *args, ns, mode)
        assert _ok(w, z[pfx + name]), (case, pfx, name, np.abs(w - z[pfx + name]).max())
    # sor with the altsweeps backward sweep (isstep = -1, doSCF.F90:403-411); the
    # reference's backward result is not its forward one, so the direction is pinned
    w = pyoracle.relax_ordered(nni, nmu, isym, z[pfx + "win"].copy(), *args, nsw, 2, "col-wise")
    assert _ok(w, z[pfx + "wout_sor_bwd"]), (case, pfx, np.abs(w - z[pfx + "wout_sor_bwd"]).max())
    assert np.abs(z[pfx + "wout_sor_bwd"] - z[pfx + "wout_sor"]).max() > 1e-3


@pytest.mark.parametrize("case", ["h2_61x91", "n2_61x91"])
def test_oracle_putin_vs_reference(case):
    z = load_golden(case)
    nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
    for pfx, k in (("k1_", 0), ("k2_", 1)):
        if pfx + "win" not in z:
            continue
        isym = int(z[pfx + "isym"][0])
        w = pyoracle.putin(nni, nmu, isym, z["pots_asympt"][k])
        assert _ok(w, z[pfx + "win"])
        assert _ok(pyoracle.putout(nni, nmu, w), z["pots_asympt"][k])


@pytest.mark.parametrize("case", ["h2_61x91", "n2_61x91"])
def test_oracle_orbital_tail_vs_reference(case):
    """The relaxed orbital's tail rows mu = nmu-3..nmu are the decay extrapolation
    of its row nmu-4 with factors from fock1/f1 (relaxOrbs.F90:104-124)."""
    z = load_golden(case)
    nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
    f1 = z["suppl"][5]
    fa = np.where(np.abs(f1) <= np.finfo(float).eps, 0.0, z["fock1"] / np.where(f1 == 0, 1, f1))
    vxisq, vetasq = z["vxi"] ** 2, z["veta"] ** 2
    psi = z["orb_relaxed"]
    got = pyoracle.orb_tail(nni, nmu, vxisq, vetasq, fa, psi)
    assert _ok(got, psi)


def test_stencil_weights_exact_on_polynomials():
    """9-point first/second derivative weights (initCBAllocArrays.F90:1670-1673)
    differentiate polynomials of degree <= 8 exactly; the boundary weights
    exeven (2037-2041) reproduce even polynomials of degree <= 8 at the axis."""
    z = load_golden("h2_61x91")
    h = float(z["scal"][0])
    dni1, dni2 = z["coef"][8:12], z["coef"][12:16]
    ex = z["coef"][16:21]
    x0 = 0.37
    for deg in range(0, 9):
        f = lambda x: (x - 0.1) ** deg
        d1 = deg * (x0 - 0.1) ** (deg - 1) if deg >= 1 else 0.0
        d2 = deg * (deg - 1) * (x0 - 0.1) ** (deg - 2) if deg >= 2 else 0.0
        ks = np.array([4, 3, 2, 1])
        g1 = np.sum(dni1 * (f(x0 - ks * h) - f(x0 + ks * h)))
        g2 = np.sum(dni2 * (f(x0 - ks * h) + f(x0 + ks * h))) + (-14350.0 / 5040.0 / h ** 2) * f(x0)
        assert abs(g1 - d1) < 1e-8 * max(1, abs(d1)), (deg, g1, d1)
        assert abs(g2 - d2) < 1e-6 * max(1, abs(d2)), (deg, g2, d2)
    for p in range(0, 5):
        vals = np.array([(k * 1.0) ** (2 * p) for k in range(1, 6)])
        expect = 1.0 if p == 0 else 0.0
        assert abs(np.dot(ex, vals) - expect) < 1e-11


def test_newton_cotes_weights_exact():
    """The mu and nu quadrature weights built into wgt1/wgt2 are the composite
    7-point Newton-Cotes rule (41, 216, 27, 272, 27, 216, 41)*h/140, exact for
    polynomials of degree <= 7 (initCBAllocArrays.F90:1917-1952)."""
    n = 61
    h = 1.0 / (n - 1)
    w = np.zeros(n)
    w[0] = 41
    for i in range(1, n, 6):
        w[i:i + 5] = [216, 27, 272, 27, 216]
        w[i + 5] = 41 if i + 5 == n - 1 else 82
    w *= h / 140
    x = np.linspace(0, 1, n)
    for deg in range(8):
        assert abs(np.dot(w, x ** deg) - 1.0 / (deg + 1)) < 1e-13


def test_oracle_fixed_point_has_zero_residual():
    """After many sweeps of a well-posed problem the MCSOR iterate satisfies
    rhs - (L + lhs) u = 0: the residual restatement and the update formula
    describe the same operator."""
    z = load_golden("h2_61x91")
    nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
    w = z["k1_win"].copy()
    args = (z["k1_lhs"], z["k1_rhs"], z["suppl"][1], z["suppl"][2], z["coef"][:16], z["coef"][16:21], 1.95)
    r0 = pyoracle.residual(nni, nmu, w, z["k1_lhs"], z["k1_rhs"], z["suppl"][1], z["suppl"][2], z["coef"][:16])
    pyoracle.mcsor(nni, nmu, 1, w, *args, 2000, 0)
    r1 = pyoracle.residual(nni, nmu, w, z["k1_lhs"], z["k1_rhs"], z["suppl"][1], z["suppl"][2], z["coef"][:16])
    assert np.linalg.norm(r1) < 1e-9 * np.linalg.norm(r0)


def test_oracle_boundary_lines_even_odd():
    nni, nmu = 31, 41
    rng = np.random.default_rng(3)
    f = rng.standard_normal((nni + 8) * (nmu + 8))
    ex = np.array([210, -120, 45, -10, 1]) / 126.0
    g = f.copy()
    pyoracle.lib().x2o_boundary(nni, nmu, 1, pyoracle._p(g), pyoracle._p(ex))
    W = g.reshape(nmu + 8, nni + 8)
    # mu=1 row from mu=2..6 at the same nu
    for nu in (2, 10, nni - 1):
        assert abs(W[4, nu + 3] - np.dot(ex, W[5:10, nu + 3])) < 1e-12
    # nu=1 at mu=1 uses the new mu=1 row
    assert abs(W[4, 4] - np.dot(ex, W[4, 5:10])) < 1e-12
    h = f.copy()
    pyoracle.lib().x2o_boundary(nni, nmu, -1, pyoracle._p(h), pyoracle._p(ex))
    V = h.reshape(nmu + 8, nni + 8)
    assert np.all(V[4, 5:nni + 3] == 0) and np.all(V[4:nmu, 4] == 0) and np.all(V[4:nmu, nni + 3] == 0)


def _f32_vs_f64(z, pfx, nsweeps, bsign=1.0):
    """Relative max difference between the FP32 twin and the FP64 oracle (itself
    bit-exact against the reference's mcsor) after nsweeps sweeps from the same
    start; bsign = -1 plants a sign error in the FP32 twin's b(mu) d/dmu term."""
    nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
    isym = int(z[pfx + "isym"][0])
    omega = float(z[pfx + "scal"][0])
    b, d, coef, ex = z["suppl"][1], z["suppl"][2], z["coef"][:16], z["coef"][16:21]
    w64 = pyoracle.mcsor(nni, nmu, isym, z[pfx + "win"].copy(), z[pfx + "lhs"], z[pfx + "rhs"], b, d, coef, ex,
                         omega, nsweeps, 0)
    w32 = pyoracle.mcsor_f32(nni, nmu, isym, z[pfx + "win"].astype(np.float32), z[pfx + "lhs"], z[pfx + "rhs"],
                             bsign * b, d, coef, ex, omega, nsweeps)
    return np.abs(w32.astype(np.float64) - w64).max() / np.abs(w64).max()


@pytest.mark.parametrize("case", ["h2_61x91", "n2_61x91"])
@pytest.mark.parametrize("pfx", ["k1_", "k2_"])
def test_oracle_f32_twin_tracks_pinned_f64_sweep(case, pfx):
    """x2o_mcsor_f32 (the checker of the FP32 correction sweeps) follows the
    FP64 oracle to FP32 rounding after 1 and 10 sweeps on the reference's own
    inputs; the same comparison with a sign error planted in its b(mu) term
    (the oracle on the CPU, never the product) fails by orders of magnitude."""
    z = load_golden(case)
    if pfx + "win" not in z:
        pytest.skip("case has a single potential")
    e1, e10 = _f32_vs_f64(z, pfx, 1), _f32_vs_f64(z, pfx, 10)
    # one sweep: a few FP32 roundings of O(1) terms; ten sweeps at omega = 1.95
    # amplify them (measured up to 4.6e-5 on the H2 Coulomb problem)
    assert e1 < 1e-5, e1
    assert e10 < 2e-4, e10
    assert _f32_vs_f64(z, pfx, 1, -1.0) > 1e-2
    assert _f32_vs_f64(z, pfx, 10, -1.0) > 1e-2


def test_oracle_f32_twin_fixed_point():
    """Many FP32 twin sweeps reach the FP64 oracle's fixed point to within the
    FP32 conditioning of the operator (2.5e-4 relative measured on the H2
    Coulomb problem, whose FP64 residual there is 4e-12 relative), and the
    pinned FP64 residual of the FP32 iterate is at FP32 level; with a sign
    error planted in the twin's b(mu) term the iteration does not get there."""
    z = load_golden("h2_61x91")
    nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
    b, d, coef, ex = z["suppl"][1], z["suppl"][2], z["coef"][:16], z["coef"][16:21]
    lhs, rhs = z["k1_lhs"], z["k1_rhs"]
    w64 = pyoracle.mcsor(nni, nmu, 1, z["k1_win"].copy(), lhs, rhs, b, d, coef, ex, 1.95, 1500, 0)
    r64 = pyoracle.residual(nni, nmu, w64, lhs, rhs, b, d, coef)
    assert np.linalg.norm(r64) < 1e-10 * np.linalg.norm(rhs)

    def run(bsign):
        with np.errstate(all="ignore"):
            w = pyoracle.mcsor_f32(nni, nmu, 1, z["k1_win"].astype(np.float32), lhs, rhs, bsign * b, d, coef,
                                   ex, 1.95, 1500).astype(np.float64)
            r = pyoracle.residual(nni, nmu, w, lhs, rhs, b, d, coef)
            return (np.abs(w - w64).max() / np.abs(w64).max(), np.linalg.norm(r) / np.linalg.norm(rhs))

    dw, res = run(1.0)
    assert dw < 1e-3 and res < 1e-2, (dw, res)
    dw_bad, res_bad = run(-1.0)
    assert not (dw_bad < 1e-1), dw_bad
    assert not (res_bad < 1e-1), res_bad


@pytest.mark.parametrize("case", ["h2_61x91_middle", "n2_61x91_middle"])
@pytest.mark.parametrize("pfx", ["k1_", "k2_"])
def test_oracle_middle_ordering_bitexact_vs_reference(case, pfx):
    """sor (forward and backward) and mcsor in the reference's default `middle`
    mesh ordering (initCBAllocArrays.F90:2499-2540), 10 and 1 sweeps, against the reference's
    own routines run on the same deck (mcsor with one thread: with several,
    the reference's OpenMP passes race on the neighbouring points that this
    ordering puts into one pass)."""
    z = load_golden(case)
    if pfx + "win" not in z:
        pytest.skip("case has a single potential")
    nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
    isym = int(z[pfx + "isym"][0])
    omega, nsw = float(z[pfx + "scal"][0]), int(z[pfx + "scal"][1])
    args = (z[pfx + "lhs"], z[pfx + "rhs"], z["suppl"][1], z["suppl"][2], z["coef"][:16], z["coef"][16:21], omega)
    for mode, name, ns in [(0, "wout_mcsor", nsw), (0, "wout_mcsor1", 1), (1, "wout_sor", nsw), (2, "wout_sor_bwd", nsw)]:
        w = pyoracle.relax_ordered(nni, nmu, isym, z[pfx + "win"].copy(), *args, ns, mode, "middle")
        assert _ok(w, z[pfx + name]), (case, pfx, name, np.abs(w - z[pfx + name]).max())


OTHER_ORDERINGS = ("middlemt", "row-wise", "rrow-wis", "test1", "test2", "test3", "test4", "middle2")


@pytest.mark.parametrize("ordering", OTHER_ORDERINGS)
def test_oracle_other_orderings_bitexact_vs_reference(ordering):
    """sor (10 sweeps, forward and backward) and one-thread mcsor (1 sweep) in each of the
    reference's other mesh orderings (initCBAllocArrays.F90:2543-2810) against
    its own routines on the H2 61x91 deck with that `order` card
    (tests/golden/h2_61x91_orderings.npz; grid arrays from the same deck)."""
    z = load_golden("h2_61x91_middle")
    v = np.load(os.path.join(GOLDEN, "h2_61x91_orderings.npz"))
    g = lambda k: v[f"{ordering}/k1_{k}"]
    nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
    isym = int(g("isym")[0])
    omega, nsw = float(g("scal")[0]), int(g("scal")[1])
    args = (g("lhs"), g("rhs"), z["suppl"][1], z["suppl"][2], z["coef"][:16], z["coef"][16:21], omega)
    for mode, name, ns in [(0, "wout_mcsor1", 1), (1, "wout_sor", nsw), (2, "wout_sor_bwd", nsw)]:
        w = pyoracle.relax_ordered(nni, nmu, isym, g("win").copy(), *args, ns, mode, ordering)
        assert _ok(w, g(name)), (ordering, name, np.abs(w - g(name)).max())
    # the MCSOR vector pins the order itself: the middle ordering's passes
    # give another result.  (SOR results of orderings that keep the same
    # precedence between stencil neighbours coincide, e.g. col-wise and
    # row-wise: the cross stencil has no diagonal couplings.)
    w = pyoracle.relax_ordered(nni, nmu, isym, g("win").copy(), *args, 1, 0, "middle")
    assert not np.array_equal(w, g("wout_mcsor1"))


def test_mesh_orderings_are_permutations():
    """Every ordering visits every relaxed point once; col-wise is mu-major
    with nu fastest, middle starts at the middle point and ends at the corner
    (nu, mu) = (nni-1, nmu-4); test1, test3 and middle2 are one sequence
    (the reference writes it three times); no two other orderings agree."""
    for nni, nmu in ((61, 91), (91, 121), (151, 211)):
        n = (nni - 2) * (nmu - 5)
        seqs = {}
        for o in pyoracle.ORDERINGS:
            s = pyoracle.mesh(nni, nmu, o)
            seqs[o] = tuple(s.tolist())
            assert len(set(s.tolist())) == n, o
        assert seqs["test1"] == seqs["test3"] == seqs["middle2"]
        distinct = {seqs[o] for o in pyoracle.ORDERINGS if o not in ("test3", "middle2")}
        assert len(distinct) == len(pyoracle.ORDERINGS) - 2
        r = pyoracle.mesh(nni, nmu, "row-wise")
        assert np.array_equal(pyoracle.mesh(nni, nmu, "rrow-wis"), r[::-1])
        for o in ("col-wise", "middle"):
            s = pyoracle.mesh(nni, nmu, o)
            mu, nu = s // nni + 1, s % nni + 1
            assert len(set(s.tolist())) == n and mu.min() == 2 and mu.max() == nmu - 4
            assert nu.min() == 2 and nu.max() == nni - 1
        s = pyoracle.mesh(nni, nmu, "middle")
        assert (s[-1] // nni + 1, s[-1] % nni + 1) == (nmu - 4, nni - 1)
        assert s[0] // nni + 1 == (nmu + 7) // 2 - 4 and s[0] % nni + 1 == (nni + 7) // 2 - 4
        c = pyoracle.mesh(nni, nmu, "col-wise")
        assert np.array_equal(c, np.array([(mu - 1) * nni + nu - 1 for mu in range(2, nmu - 3)
                                           for nu in range(2, nni)]))
