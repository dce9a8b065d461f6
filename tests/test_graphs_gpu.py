"""Graph replay of the batched relaxation (x2g_set_graphs): a call signature
seen twice is captured into a HIP graph and later calls launch the graph.
Every replay must give the eager call's result bit for bit, with the
program's own eager work and synchronisations between the replays, after a
change of the sweep-plan settings (a new signature, captured anew), and over
whole served SCF requests."""
import ctypes as C

import numpy as np
import pytest

from test_kernels_gpu import synthetic_problem

pytestmark = pytest.mark.gpu

SHAPES = [(91, 121), (169, 241), (331, 601)]
CALLS = 6


def _counts(ctx):
    from x2dhf_amd import gpu
    c, r = C.c_longlong(), C.c_longlong()
    gpu.check(ctx.lib.x2g_graph_counts(ctx.h, C.byref(c), C.byref(r)))
    return c.value, r.value


def _sequence(graphs, replan_at=None):
    """CALLS relaxations of one mixed batch (FP32 correction items and FP64
    items on three grids), each followed by eager work on another buffer and
    a synchronisation; the functions after every call."""
    from x2dhf_amd import gpu
    ctx = gpu.Context(0)
    gpu.check(ctx.lib.x2g_set_sweep_plan(0, 0, 0))
    gpu.check(ctx.lib.x2g_set_graphs(ctx.h, 1 if graphs else 0))
    probs = [synthetic_problem(n, m, seed=31 + k) for k, (n, m) in enumerate(SHAPES)]
    grids = [gpu.Grid(ctx, p["nni"], p["nmu"], p["coef"], p["suppl"], p["vxi"], p["veta"]) for p in probs]
    gh = (gpu._vp * len(grids))(*[g.h.value for g in grids])
    keep, items = [], []
    for k, p in enumerate(probs):
        for prec in (1, 0):
            f, lo, r = ctx.upload(p["fun"]), ctx.upload(p["lhs"]), ctx.upload(p["rhs"])
            keep.append((f, lo, r))
            items.append(gpu.RelaxItem(f.ptr, lo.ptr, r.ptr, None, 1.9, 1 if k % 2 == 0 else -1, 1, prec, k))
    arr = (gpu.RelaxItem * len(items))(*items)
    other = ctx.alloc(1 << 20)
    out = []
    try:
        for call in range(CALLS):
            if call == replan_at:
                gpu.check(ctx.lib.x2g_set_sweep_plan(2, 0, 0))
            gpu.check(ctx.lib.x2g_relax_batch(ctx.h, gh, arr, len(items), 1, 10))
            other.fill(float(call))
            ctx.sync()
            out.append([f.get() for f, _, _ in keep])
        return out, _counts(ctx)
    finally:
        gpu.check(ctx.lib.x2g_set_sweep_plan(0, 0, 0))
        for g in grids:
            g.destroy()
        ctx.close()


def test_graph_replays_equal_eager_calls():
    print("eager sequence", flush=True)
    eager, c0 = _sequence(False)
    print("graph sequence", flush=True)
    graph, c1 = _sequence(True)
    print(f"captures/replays: eager {c0}, graphs {c1}")
    assert c0 == (0, 0)
    assert c1[0] == 1 and c1[1] >= CALLS - 3
    for call, (a, b) in enumerate(zip(eager, graph)):
        for k, (x, y) in enumerate(zip(a, b)):
            assert np.array_equal(x, y), (call, k)


def test_graph_replays_of_ordered_sor_equal_eager_calls():
    """Items relaxed in the reference's sequential point order (k_levels:
    SOR forward and backward, one-thread MCSOR, in the middle ordering) are
    replayed from the captured graph with the eager results, bit for bit."""
    from x2dhf_amd import gpu
    res = {}
    for graphs in (False, True):
        ctx = gpu.Context(0)
        try:
            gpu.check(ctx.lib.x2g_set_graphs(ctx.h, 1 if graphs else 0))
            probs = [synthetic_problem(n, m, seed=61 + k) for k, (n, m) in enumerate(SHAPES[:2])]
            grids = [gpu.Grid(ctx, p["nni"], p["nmu"], p["coef"], p["suppl"], p["vxi"], p["veta"]) for p in probs]
            gh = (gpu._vp * len(grids))(*[g.h.value for g in grids])
            keep, items = [], []
            for k, p in enumerate(probs):
                for method in (1, 2, 0):
                    f, lo, r = ctx.upload(p["fun"]), ctx.upload(p["lhs"]), ctx.upload(p["rhs"])
                    keep.append((f, lo, r))
                    it = gpu.RelaxItem(f.ptr, lo.ptr, r.ptr, None, 1.9, 1 if method != 2 else -1, 1, 0, k)
                    it.method, it.ordering = method, 1
                    items.append(it)
            arr = (gpu.RelaxItem * len(items))(*items)
            out = []
            print(f"ordered items, graphs={graphs}", flush=True)
            for call in range(4):
                gpu.check(ctx.lib.x2g_relax_batch(ctx.h, gh, arr, len(items), 1, 3))
                ctx.sync()
                out.append([f.get() for f, _, _ in keep])
            res[graphs] = (out, _counts(ctx))
            for g in grids:
                g.destroy()
        finally:
            ctx.close()
    print(f"captures/replays {res[True][1]}")
    assert res[False][1] == (0, 0) and res[True][1][0] == 1 and res[True][1][1] >= 2
    for call, (a, b) in enumerate(zip(res[False][0], res[True][0])):
        for k, (x, y) in enumerate(zip(a, b)):
            assert np.array_equal(x, y), (call, k)


def test_new_sweep_plan_is_captured_anew():
    """Changing the sweep plan mid-sequence makes a new signature: the graph of
    the old plan is not replayed, and the results still equal the eager ones."""
    eager, _ = _sequence(False, replan_at=3)
    graph, c = _sequence(True, replan_at=3)
    print(f"captures/replays {c}")
    assert c[0] >= 1
    for call, (a, b) in enumerate(zip(eager, graph)):
        for k, (x, y) in enumerate(zip(a, b)):
            assert np.array_equal(x, y), (call, k)


def test_served_requests_with_graphs_equal_eager():
    """Three requests served together with graph replay give the energies and
    iteration counts of the eager server, and the relaxations were replayed."""
    from x2dhf_amd import scf, serve
    specs = [scf.lih_spec(61, 91, 40.0, schedule=1, prec=1, ienterm=11, inoterm=11, maxscf=2000),
             scf.h2_spec(61, 91, 40.0, schedule=1, prec=1, ienterm=11, inoterm=11, maxscf=2000),
             scf.bh_spec(91, 121, 40.0, schedule=1, prec=1, ienterm=11, inoterm=11, maxscf=2000)]
    res = {}
    for graphs in (False, True):
        print(f"served, graphs={graphs}", flush=True)
        srv = serve.Server(max_active=8, graphs=graphs)
        try:
            reqs = [srv.submit(s, arrival=0.0) for s in specs]
            srv.run()
            res[graphs] = [r.result for r in reqs]
            counts = srv.graph_counts()
        finally:
            srv.close()
        print(f"captures/replays {counts}")
        if graphs:
            assert counts[0] >= 1 and counts[1] > 100
        else:
            assert counts == (0, 0)
    for a, b in zip(res[False], res[True]):
        assert a["iterations"] == b["iterations"]
        assert a["etot"] == b["etot"]
        assert np.array_equal(a["eps"], b["eps"])
