"""The batched C ABI (several problems on several grids per launch, scalars
kept on the device) against the single-problem entry points, which the other
GPU tests pin to the reference's routines.  The batched reductions keep the
single ones' slicing and summation order, so quadratures, moments, tails and
linear combinations agree bit for bit; the folded norm+ortho update agrees
to rounding (1e-14 relative)."""
import ctypes as C

import numpy as np
import pytest

from conftest import grid_coef, load_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    from x2dhf_amd import gpu
    ctx = gpu.Context(0)
    gs = []
    for case in ("h2_61x91", "n2_61x91"):
        z = load_golden(case)
        nni, nmu = int(z["sizes"][0]), int(z["sizes"][1])
        gs.append((z, gpu.Grid(ctx, nni, nmu, grid_coef(z), z["suppl"], z["vxi"], z["veta"])))
    yield ctx, gs
    for _, g in gs:
        g.destroy()
    ctx.close()


def grids_arg(gs):
    from x2dhf_amd import gpu
    return (gpu._vp * len(gs))(*[g.h.value for _, g in gs])


def test_lincomb_batch_equals_single(env):
    from x2dhf_amd import gpu
    ctx, gs = env
    rng = np.random.default_rng(1)
    ops, terms, outs, expect = [], [], [], []
    for gi, (z, g) in enumerate(gs):
        a, b, base = (ctx.upload(rng.standard_normal(g.n)) for _ in range(3))
        outs += [a, b, base]
        t = [gpu.Term(0.7, a.ptr, b.ptr, None), gpu.Term(-1.3, b.ptr, None, None)]
        ref = ctx.alloc(g.n)
        gpu.check(ctx.lib.x2g_lincomb(g.h, ref.ptr, base.ptr, 0.25, (gpu.Term * 2)(*t), 2))
        out = ctx.alloc(g.n)
        outs += [ref, out]
        ops.append(gpu.LincombOp(out.ptr, base.ptr, 0.25, len(terms), 2, gi))
        terms += t
        expect.append((out, ref))
    gpu.check(ctx.lib.x2g_lincomb_batch(ctx.h, grids_arg(gs), (gpu.LincombOp * len(ops))(*ops), len(ops),
                                        (gpu.Term * len(terms))(*terms), len(terms)))
    for out, ref in expect:
        assert np.array_equal(out.get(), ref.get())


def test_normortho_batch_matches_reference_sequence(env):
    """Against norm (x2g_wdot + scaling), setci and ortho done one call at a time
    in the reference's order (normOrtho.F90:14-59, 103-190)."""
    from x2dhf_amd import gpu
    ctx, gs = env
    z, g = gs[1]                               # N2: sigma-g orbitals 2, 4, 6
    S = [g.suppl_ptr(k) for k in range(14)]
    orbs = [ctx.upload(z["orb_prep"][k]) for k in range(6)]
    ihomo = [1, 1, -1, 1, -1, 1]
    # orbital 2 against the final orbitals 4 and 6 (same symmetry, higher index)
    psi0 = z["orb_prep"][1] * 1.01 + 1e-3 * z["orb_prep"][3]
    # sequential reference
    p = ctx.upload(psi0)
    x = np.zeros(1)
    gpu.check(ctx.lib.x2g_wdot(g.h, (gpu.Dot * 1)(gpu.Dot(S[13], p.ptr, p.ptr, S[8])), 1, x.ctypes.data))
    gpu.check(ctx.lib.x2g_lincomb(g.h, p.ptr, p.ptr, 1.0 / np.sqrt(x[0]), None, 0))
    gpu.check(ctx.lib.x2g_setci(g.h, p.ptr, ihomo[1]))
    ov = np.zeros(2)
    d = (gpu.Dot * 2)(gpu.Dot(S[13], p.ptr, orbs[3].ptr, S[8]), gpu.Dot(S[13], p.ptr, orbs[5].ptr, S[8]))
    gpu.check(ctx.lib.x2g_wdot(g.h, d, 2, ov.ctypes.data))
    t = (gpu.Term * 2)(gpu.Term(-ov[0], orbs[3].ptr, None, None), gpu.Term(-ov[1], orbs[5].ptr, None, None))
    gpu.check(ctx.lib.x2g_lincomb(g.h, p.ptr, p.ptr, 1.0, t, 2))
    a = ov[0] ** 2 + ov[1] ** 2
    gpu.check(ctx.lib.x2g_lincomb(g.h, p.ptr, p.ptr, np.sqrt(1 + a) / (1 - a * a), None, 0))
    ref = p.get()
    # batched: the same orbital twice (two items) plus an H2 item
    q1, q2 = ctx.upload(psi0), ctx.upload(psi0)
    zh, gh = gs[0]
    h = ctx.upload(zh["orb_prep"][0] * 0.9)
    nrm = ctx.alloc(3)
    it = []
    for k, q in enumerate((q1, q2)):
        o = gpu.NormOrthoItem()
        o.psi_d = q.ptr
        o.others_d[0], o.others_d[1] = orbs[3].ptr, orbs[5].ptr
        o.nothers, o.ihomo, o.normalise, o.grid_index = 2, ihomo[1], 1, 1
        o.norm_out_d = nrm.ptr + 8 * k
        it.append(o)
    o = gpu.NormOrthoItem()
    o.psi_d, o.nothers, o.ihomo, o.normalise, o.grid_index, o.norm_out_d = h.ptr, 0, 1, 1, 0, nrm.ptr + 16
    it.append(o)
    gpu.check(ctx.lib.x2g_normortho_batch(ctx.h, grids_arg(gs), (gpu.NormOrthoItem * 3)(*it), 3))
    n = nrm.get()
    assert n[0] == x[0] and n[1] == x[0]
    for q in (q1, q2):
        got = q.get()
        assert np.abs(got - ref).max() < 1e-14 * np.abs(ref).max()
    # H2: plain normalisation (no others) then the g symmetry
    hv = h.get()
    w = np.zeros(1)
    hh = ctx.upload(hv)
    gpu.check(ctx.lib.x2g_wdot(gh.h, (gpu.Dot * 1)(gpu.Dot(gh.suppl_ptr(13), hh.ptr, hh.ptr, gh.suppl_ptr(8))), 1,
                               w.ctypes.data))
    assert abs(w[0] - 1.0) < 1e-13


def test_eahf_batch_equals_single_quadratures(env):
    from x2dhf_amd import gpu
    ctx, gs = env
    items, keep, expect = [], [], []
    out = ctx.alloc(4)
    rng = np.random.default_rng(3)
    for gi, (z, g) in enumerate(gs):
        S = [g.suppl_ptr(k) for k in range(14)]
        for k, (isym, m) in enumerate(((1, 0), (-1, 1)) if gi else ((1, 0), (1, 2))):
            psi = ctx.upload(z["orb_prep"][k % int(z["sizes"][2])])
            cp, xp = ctx.upload(rng.random(g.n)), ctx.upload(rng.standard_normal(g.n) * 1e-2)
            keep += [psi, cp, xp]
            one = np.zeros(3)
            gpu.check(ctx.lib.x2g_oneel(g.h, psi.ptr, psi.ptr, isym, m, one.ctypes.data))
            two = np.zeros(2)
            d = (gpu.Dot * 2)(gpu.Dot(S[13], psi.ptr, cp.ptr, psi.ptr), gpu.Dot(S[13], psi.ptr, xp.ptr, None))
            gpu.check(ctx.lib.x2g_wdot(g.h, d, 2, two.ctypes.data))
            expect.append(one[0] + (two[0] - two[1]))
            items.append(gpu.EaHFItem(psi.ptr, cp.ptr, xp.ptr, isym, m, gi, out.ptr + 8 * len(items)))
    gpu.check(ctx.lib.x2g_eahf_batch(ctx.h, grids_arg(gs), (gpu.EaHFItem * 4)(*items), 4))
    assert np.array_equal(out.get(), np.array(expect))


def test_moments_and_tails_batch_equal_single(env):
    from x2dhf_amd import gpu
    ctx, gs = env
    z, g = gs[1]
    orbs = [ctx.upload(z["orb_prep"][k]) for k in range(6)]
    pairs = [(0, 0, 0), (1, 1, 0), (0, 1, 1), (0, 0, 2), (2, 5, 0)]
    its = (gpu.MomentItem * len(pairs))(*[gpu.MomentItem(orbs[a].ptr, orbs[b].ptr, dm, 1) for a, b, dm in pairs])
    out = np.zeros(8 * len(pairs))
    gpu.check(ctx.lib.x2g_moments_batch(ctx.h, grids_arg(gs), its, len(pairs), 4, out.ctypes.data))
    for k, (a, b, dm) in enumerate(pairs):
        mom = np.zeros(8)
        gpu.check(ctx.lib.x2g_moments(g.h, orbs[a].ptr, orbs[b].ptr, dm, 4, mom.ctypes.data))
        assert np.array_equal(out[8 * k:8 * k + 4], mom[:4]), k
    # tails of a Coulomb and two exchange potentials
    rng = np.random.default_rng(4)
    specs = [(0, 0), (1, 1), (1, 2)]
    pots_b = [ctx.upload(rng.standard_normal(g.n)) for _ in specs]
    pots_s = [ctx.upload(p.get()) for p in pots_b]
    tails = []
    for (kind, dm), p in zip(specs, pots_b):
        t = gpu.TailItem(p.ptr, kind, dm, 1)
        for q in range(8):
            t.mom[q] = out[q] if q < 4 else 0.0
        tails.append(t)
    gpu.check(ctx.lib.x2g_pot_tail_batch(ctx.h, grids_arg(gs), (gpu.TailItem * 3)(*tails), 3, 4))
    for (kind, dm), pb, ps in zip(specs, pots_b, pots_s):
        gpu.check(ctx.lib.x2g_pot_tail(g.h, ps.ptr, kind, dm, 4, out[:8].ctypes.data))
        assert np.array_equal(pb.get(), ps.get())


def test_gather(env):
    from x2dhf_amd import gpu
    ctx, _ = env
    a, b = ctx.upload(np.arange(5.0)), ctx.upload(np.arange(10.0, 13.0))
    out = np.zeros(8)
    src = (gpu._vp * 2)(a.ptr, b.ptr)
    cnt = (C.c_int * 2)(5, 3)
    gpu.check(ctx.lib.x2g_gather(ctx.h, src, cnt, 2, out.ctypes.data))
    assert np.array_equal(out, [0, 1, 2, 3, 4, 10, 11, 12])


def test_empty_batches_are_no_ops(env):
    """Every batched entry point accepts an empty batch (n = 0: a server step
    with no job in some phase) and returns success without touching the
    device or its arguments."""
    from x2dhf_amd import gpu
    ctx, gs = env
    g = grids_arg(gs)
    out = np.full(4, 7.0)
    assert ctx.lib.x2g_relax_batch(ctx.h, g, None, 0, 1, 10) == 0
    assert ctx.lib.x2g_residual_norms(ctx.h, g, None, 0, out.ctypes.data) == 0
    assert ctx.lib.x2g_lincomb_batch(ctx.h, g, None, 0, None, 0) == 0
    assert ctx.lib.x2g_normortho_batch(ctx.h, g, None, 0) == 0
    assert ctx.lib.x2g_eahf_batch(ctx.h, g, None, 0) == 0
    assert ctx.lib.x2g_moments_batch(ctx.h, g, None, 0, 4, out.ctypes.data) == 0
    assert ctx.lib.x2g_pot_tail_batch(ctx.h, g, None, 0, 4) == 0
    assert ctx.lib.x2g_gather(ctx.h, None, None, 0, out.ctypes.data) == 0
    assert ctx.lib.x2g_scf_control(ctx.h, None, None, 0) == 0
    ctx.sync()
    assert np.array_equal(out, np.full(4, 7.0))
