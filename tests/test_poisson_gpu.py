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
