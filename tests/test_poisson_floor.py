"""The FP64 floor of the configuration-4 problems' relative residual, pinned
on the CPU with the oracle's FP64 residual (oracle/x2oracle.c, x2o_residual):
for the exact solution u of a Gaussian charge distribution (the closed-form
potential erf(d / (sigma sqrt 2)) / d, stored as (R/2) xi V), any FP64 copy of
u carries a relative rounding delta, and rhs - A fl(u) = A (u * delta)
(A linear).  ||A (u * delta)|| / ||rhs|| is therefore the smallest residual an
FP64-stored solution reaches; for most of the seeded problems it exceeds the
configuration's 1e-12.  The GPU floor certificate (poisson.Batch.floors)
measures the same quantity with the device residual."""
import numpy as np
import pytest
from scipy.special import erf

from oracle import pyoracle


def closed_form_u(p, vxi, veta):
    r2 = 0.5 * p["r"]
    xi, eta = np.meshgrid(vxi, veta, indexing="ij")
    z = r2 * xi * eta
    rho2 = (r2 * r2) * np.maximum(xi * xi - 1, 0) * np.maximum(1 - eta * eta, 0)
    v, qs = np.zeros_like(z), 0.0
    for zk, sk, wk in p["gauss"]:
        dist = np.sqrt(rho2 + (z - zk) ** 2)
        q = wk * (2 * np.pi * sk * sk) ** 1.5
        qs += q
        v += q * np.where(dist > 1e-12, erf(dist / (sk * np.sqrt(2))) / np.where(dist > 1e-12, dist, 1),
                          np.sqrt(2 / np.pi) / sk)
    return (r2 * xi * v / qs).ravel()


def floor_and_residual(p, seed):
    from x2dhf_amd import poisson, scf
    nni, nmu = p["nni"], p["nmu"]
    coef, suppl, vxi, veta = scf.grid_build(nni, nmu, 1.0, 1.0, p["r"], p["rinf"])
    d = poisson.density(p, vxi, veta)
    d /= float(np.sum(suppl[13] * suppl[8] * d))
    rhs = d * suppl[9]
    u = closed_form_u(p, vxi, veta)
    rng = np.random.default_rng(seed)
    du = u * rng.uniform(-0.5, 0.5, u.size) * np.finfo(float).eps
    zero = np.zeros(nni * nmu)
    r_du = pyoracle.residual(nni, nmu, pyoracle.putin(nni, nmu, 1, du), suppl[7], zero, suppl[1], suppl[2],
                             coef[:16])
    r_u = pyoracle.residual(nni, nmu, pyoracle.putin(nni, nmu, 1, u), suppl[7], rhs, suppl[1], suppl[2], coef[:16])
    nb = np.linalg.norm(rhs)
    return np.linalg.norm(r_du) / nb, np.linalg.norm(r_u) / nb


def test_floor_exceeds_the_configured_tolerance_for_most_problems():
    from x2dhf_amd import poisson
    probs = poisson.make_problems(256, seed=0)
    sample = probs[::8]                               # 32 of the 256, all four grids
    fl = np.array([floor_and_residual(p, 7)[0] for p in sample])
    grids = np.array([p["nni"] for p in sample])
    print("floor by grid:", {int(n): f"{fl[grids == n].min():.1e}..{fl[grids == n].max():.1e}"
                             for n in np.unique(grids)})
    # the finer grids' h^-2 weights put their floor well above 1e-12
    assert (fl > 1e-12).sum() >= len(sample) // 2
    assert fl[grids >= 337].min() > 1e-12
    # and it is an FP64 effect: far below anything FP32 storage would allow
    assert fl.max() < 1e-9


@pytest.mark.parametrize("k", [0, 13])
def test_floor_scales_with_the_rounding_and_not_with_the_sample(k):
    """The floor is a property of the problem: two independent rounding
    samples agree to within 10%, and halving delta halves it (A is linear)."""
    from x2dhf_amd import poisson
    p = poisson.make_problems(256, seed=0)[k]
    a, _ = floor_and_residual(p, 1)
    b, _ = floor_and_residual(p, 2)
    assert abs(a - b) < 0.1 * max(a, b)
