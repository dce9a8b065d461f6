"""Batched synthetic Poisson problems (x2dhf_amd.poisson) against closed forms:
the potential of a unit Gaussian charge is erf(d / (sigma sqrt 2)) / d, and
x2dhf stores the Coulomb potential as (R/2) xi V (the form of its multipole
tail, scfUtils.F90:1051-1066).  The 8th-order stencil and the Dirichlet tail
from the multipole moments reproduce it to discretisation accuracy; the FP32
correction sweeps with FP64 residuals reach the same solution as FP64 sweeps."""
import numpy as np
import pytest
from scipy.special import erf

pytestmark = pytest.mark.gpu


def _gauss_problem(nni, nmu, r, sigma):
    return dict(nni=nni, nmu=nmu, r=r, rinf=40.0, gauss=[(0.0, sigma, 1.0)], electrons=1.0)


@pytest.mark.parametrize("prec", [0, 1])
def test_gaussian_charge_potential_closed_form(prec):
    from x2dhf_amd import poisson, scf
    p = _gauss_problem(169, 241, 2.0, 1.0)
    print(f"gaussian prec={prec}", flush=True)
    b = poisson.Batch([p], prec=prec)
    try:
        calls, res = b.solve(tol=1e-12, stall_checks=10)
        u = b.solution(0).reshape(241, 169)
    finally:
        b.close()
    _, _, vxi, veta = scf.grid_build(169, 241, 1.0, 1.0, 2.0, 40.0)
    xi, eta = np.meshgrid(vxi, veta, indexing="ij")
    r2 = 1.0
    dist = r2 * np.sqrt(np.maximum(xi * xi + eta * eta - 1.0, 0.0))
    exact = np.where(dist > 1e-12, erf(dist / np.sqrt(2.0)) / np.where(dist > 1e-12, dist, 1.0), np.sqrt(2 / np.pi))
    got = u / (r2 * xi)
    sel = (dist < 15.0)
    err = np.abs(got[sel] - exact[sel]).max()
    print(f"calls {calls[0]} residual {res[0]:.1e} max |V - erf(d/sqrt2)/d| = {err:.2e}")
    assert res[0] <= 1e-10
    assert err < 1e-9


def test_mixed_precision_batch_reaches_fp64_solution():
    from x2dhf_amd import poisson
    probs = poisson.make_problems(6, seed=3)
    sols = {}
    for prec in (0, 1):
        b = poisson.Batch(probs, prec=prec)
        try:
            calls, res = b.solve(tol=1e-12, stall_checks=10)
            print(f"prec={prec} residuals {np.array2string(res, precision=1)}", flush=True)
            assert res.max() <= 1e-9
            sols[prec] = [b.solution(i) for i in range(len(probs))]
        finally:
            b.close()
    for a, m in zip(sols[0], sols[1]):
        assert np.abs(a - m).max() < 1e-8 * np.abs(a).max()


def test_config4_256_problems_reach_tolerance_or_fp64_floor():
    """Configuration 4 at its stated size: the 256 seeded problems on the four
    grids, FP32 sweeps with FP64 refinement, relaxed together.  Every problem
    reaches max(1e-12, FLOOR_FACTOR x its FP64 floor) -- the residual that the
    rounding of its own FP64 solution causes (x2dhf_amd.poisson; the floor
    exceeds 1e-12 for most of these problems, see
    tests/test_poisson_floor.py).  A problem that stalls above that does not
    count as solved, and fails this test."""
    from x2dhf_amd import poisson
    print("256 problems, prec=1", flush=True)
    s = poisson.run(256, seed=0, prec=1, tol=1e-12)
    res, fl = np.array(s["residuals"]), np.array(s["floors"])
    above = res > 1e-12
    ratio = (res[above] / fl[above]).max() if above.any() else 0.0
    print(f"solve {s['solve_s']:.2f} s: solved {s['solved']}/256, at 1e-12 {s['at_tol']}, max residual "
          f"{res.max():.2e}, worst residual/floor {ratio:.2f}, floors {fl[above].min() if above.any() else 0:.1e}.."
          f"{fl.max():.1e}")
    assert s["solved"] == 256
