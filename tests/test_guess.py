"""Starting functions of a job: the device kernels (x2g_guess_orbital /
_coulomb / _exchange, x2dhf_amd/csrc/x2gpu_guess.hip) against the CPU
restatement (oracle/guess_oracle.py), which is pinned here by closed forms:
the hydrogen 1s, 2s, 2p0 and 2p1 functions, their unit norm under the grid's
own quadrature (wgt2 * f4, the orbital norm of the SCF), and the 1/r tail of
the Thomas-Fermi potential."""
import numpy as np
import pytest

NNI, NMU, R, RINF = 91, 121, 1.4, 40.0


def _grid():
    from x2dhf_amd import scf
    return scf.grid_build(NNI, NMU, 1.0, 1.0, R, RINF)


def _polar(vxi, veta, side):
    """Distance and cos(theta) from centre A (z = -R/2, side 0) or B (side 1)."""
    xi, eta = np.meshgrid(vxi, veta, indexing="ij")
    z = (R / 2) * xi * eta
    rho2 = (R / 2) ** 2 * np.maximum(xi * xi - 1, 0) * np.maximum(1 - eta * eta, 0)
    zc = z + R / 2 if side == 0 else z - R / 2
    r = np.sqrt(rho2 + zc * zc)
    c = np.where(r > 1e-12, zc / np.where(r > 1e-12, r, 1.0), -1.0)
    if side == 1:
        c = -c  # centre B measures theta from the other end of the axis
    return r.ravel(), c.ravel(), np.sqrt(np.maximum(1 - c * c, 0)).ravel()


@pytest.mark.parametrize("side", [0, 1])
def test_oracle_hydrogen_closed_forms(side):
    from oracle import guess_oracle as go
    _, _, vxi, veta = _grid()
    r, c, s = _polar(vxi, veta, side)
    z = 1.7
    cases = {
        (1, 0, 0): z ** 1.5 * np.exp(-z * r) / np.sqrt(np.pi),
        (2, 0, 0): z ** 1.5 / (4 * np.sqrt(2 * np.pi)) * (2 - z * r) * np.exp(-z * r / 2),
        (2, 1, 0): z ** 2.5 / (4 * np.sqrt(2 * np.pi)) * r * np.exp(-z * r / 2) * c,
        (2, 1, 1): z ** 2.5 / (8 * np.sqrt(np.pi)) * r * np.exp(-z * r / 2) * s,
    }
    for (n, l, m), exact in cases.items():
        got = go.hydrogenic(vxi, veta, R, side, n, l, m, z)
        # m = 1 on the axis: sin(theta) is the square root of a rounding residue
        # of 1 - cos^2 (about 1e-16), so exact zeros come out near 1e-8
        tol = 1e-12 if m == 0 else 5e-8
        assert np.abs(got - exact).max() < tol * np.abs(exact).max(), (n, l, m)


def test_oracle_orbitals_unit_norm_on_the_grid():
    from oracle import guess_oracle as go
    _, suppl, vxi, veta = _grid()
    w = suppl[13] * suppl[8]                      # wgt2 * f4: the SCF's orbital norm
    # unit norm to the grid's discretisation error (larger for the diffuse 3d functions)
    for n, l, m, z, tol in [(1, 0, 0, 1.0, 1e-9), (2, 0, 0, 1.0, 1e-9), (2, 1, 0, 1.0, 1e-9),
                            (2, 1, 1, 1.0, 1e-9), (3, 2, 1, 1.0, 1e-5), (3, 2, 2, 1.0, 1e-5)]:
        for side in (0, 1):
            phi = go.hydrogenic(vxi, veta, R, side, n, l, m, z)
            norm = float(np.sum(w * phi * phi))
            assert abs(norm - 1.0) < tol, (n, l, m, side, norm)


def test_oracle_thomas_fermi_tail():
    from oracle import guess_oracle as go
    r = np.array([50.0, 200.0, 1000.0])
    xi = 2 * r / R                                 # eta = 0 puts the point at distance R xi / 2 from A
    v = go.pot_tf(R, np.zeros(3), xi)
    assert np.all(np.abs(v * r - 1.0) < 0.05 * np.array([1, 0.3, 0.1])) and np.all(np.diff(v * r) > 0)


def _device_grid():
    from x2dhf_amd import gpu
    coef, suppl, vxi, veta = _grid()
    ctx = gpu.Context(0)
    return ctx, gpu.Grid(ctx, NNI, NMU, coef, suppl, vxi, veta), suppl, vxi, veta


@pytest.mark.gpu
def test_device_orbital_guess_matches_oracle():
    from x2dhf_amd import gpu
    from oracle import guess_oracle as go
    ctx, g, suppl, vxi, veta = _device_grid()
    try:
        out = ctx.alloc(g.n)
        for lc in [(0, 1, 0, 1, 0, 1.0, 1.0, 1.0, 1.0), (0, 2, 1, 1, 0, 0.7, 5.2, 0.5, 1.0),
                   (1, 2, 1, 2, 1, 0.5, 5.0, -0.5, 5.0), (0, 2, 0, 1, 0, 1.0, 4.0, 0.0, 1.0),
                   (1, 3, 2, 2, 1, 1.0, 2.5, 0.3, 1.5)]:
            m, n1, l1, n2, l2, c1, za, c2, zb = lc
            print(f"guess orbital {lc}", flush=True)
            gpu.check(ctx.lib.x2g_guess_orbital(g.h, out.ptr, C_lcao(m, n1, l1, n2, l2, c1, za, c2, zb)))
            got = out.get()
            want = go.lcao(vxi, veta, R, m, c1, n1, l1, za, c2, n2, l2, zb)
            assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max(), lc
    finally:
        g.destroy()
        ctx.close()


@pytest.mark.gpu
def test_device_potential_guesses_match_oracle():
    from x2dhf_amd import gpu
    from oracle import guess_oracle as go
    ctx, g, suppl, vxi, veta = _device_grid()
    try:
        out = ctx.alloc(g.n)
        for c1, c2, two in [(0.5, 0.5, 1), (1.0, 0.0, 0), (0.7, -0.3, 1)]:
            print(f"guess potentials {c1} {c2} {two}", flush=True)
            gpu.check(ctx.lib.x2g_guess_coulomb(g.h, out.ptr, c1, c2, two))
            got, want = out.get(), go.coulomb(vxi, veta, R, suppl[8], c1, c2, two)
            # 1 - (t/w)^2 cancels near a nucleus, where the device's fused
            # multiply-adds and numpy's separate roundings part at 1e-13
            assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
            gpu.check(ctx.lib.x2g_guess_exchange(g.h, out.ptr, c1, c2))
            got, want = out.get(), go.exchange(vxi, veta, R, suppl[8], c1, c2)
            assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    finally:
        g.destroy()
        ctx.close()


def C_lcao(*a):
    import ctypes as C
    from x2dhf_amd import gpu
    return C.byref(gpu.Lcao(*a))
