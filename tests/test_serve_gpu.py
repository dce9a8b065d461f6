"""Continuous batching (x2dhf_amd.serve): requests that arrive while others are
on the GPU join the running batch as they arrive, their iterations overlap
with the others', and each result is bit for bit the one the request gets
when it is served alone.  A clock advanced by one unit per server step makes
the arrivals deterministic."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class StepClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


def _specs():
    from x2dhf_amd import scf
    return [scf.n2_spec(91, 121, 40.0, schedule=1, prec=1, ienterm=11, inoterm=11, maxscf=2000),
            scf.lih_spec(61, 91, 40.0, schedule=1, prec=1, ienterm=11, inoterm=11, maxscf=2000),
            scf.h2_spec(61, 91, 40.0, schedule=1, prec=0, ienterm=11, inoterm=11, maxscf=2000),
            scf.bh_spec(91, 121, 40.0, schedule=1, prec=1, ienterm=11, inoterm=11, maxscf=2000)]


def _serve(specs, arrivals):
    from x2dhf_amd import serve
    clk = StepClock()
    srv = serve.Server(max_active=8, clock=clk)
    try:
        reqs = [srv.submit(s, arrival=a) for s, a in zip(specs, arrivals)]
        steps = 0
        while srv.pending or srv.active:
            srv.step()
            clk.t += 1.0
            steps += 1
            assert steps < 20000
        return reqs
    finally:
        srv.close()


def test_arrivals_join_the_running_batch_and_results_do_not_depend_on_the_mix():
    specs = _specs()
    print("served alone", flush=True)
    alone = [_serve([s], [0.0])[0].result for s in specs]
    print("served with staggered arrivals", flush=True)
    reqs = _serve(specs, [0.0, 15.0, 40.0, 41.0])
    # the later requests were admitted while the first was still iterating
    assert reqs[1].admitted < reqs[0].finished and reqs[2].admitted < reqs[0].finished
    assert all(r.admitted >= r.arrival for r in reqs)
    for a, r in zip(alone, reqs):
        m = r.result
        print(f"alone E={a['etot']:.15f} it={a['iterations']}   served E={m['etot']:.15f} it={m['iterations']} "
              f"(arrival {r.arrival:.0f}, admitted {r.admitted:.0f}, finished {r.finished:.0f})")
        assert m["iterations"] == a["iterations"]
        assert m["etot"] == a["etot"]
        assert np.array_equal(m["eps"], a["eps"])


def test_admission_respects_the_memory_budget():
    """A budget that holds the N2 job alone (but the two small jobs together):
    the small requests wait for the N2 job, in arrival order, and the
    admitted footprint never exceeds the budget."""
    from x2dhf_amd import scf, serve
    specs = _specs()[:3]                       # N2 91x121, LiH 61x91, H2 61x91
    fp = [serve.footprint(s) for s in specs]
    assert fp[0] + fp[1] > fp[0] and fp[1] + fp[2] <= fp[0]
    clk = StepClock()
    srv = serve.Server(max_active=8, hbm_budget=fp[0], clock=clk)
    try:
        reqs = [srv.submit(s, arrival=0.0) for s in specs]
        while srv.pending or srv.active:
            srv.step()
            assert srv.held <= srv.budget
            clk.t += 1.0
    finally:
        srv.close()
    assert reqs[1].admitted >= reqs[0].finished and reqs[2].admitted >= reqs[0].finished
    assert reqs[1].admitted <= reqs[2].admitted


def test_config5_widened_stream_each_result_as_alone_and_within_1e8_of_fp64():
    """Configuration 5's request mix: a seeded Poisson-arrival stream of
    closed-shell diatomics (Z up to 9 here, R in [1, 6], 1-5 orbitals, grids
    from 169x241 to 505x901), served with continuous batching in mixed
    precision.  Each served result is bit for bit the one the request gets
    served alone, and its total energy is within the stated 1e-8 Eh of the
    same request run alone with FP64 sweeps."""
    from x2dhf_amd import scf, serve
    stream = serve.poisson_stream(8, 20.0, seed=2, prec=1)
    names = [n for _, n, _ in stream]
    grids = {(s.nni, s.nmu) for _, _, s in stream}
    assert (169, 241) in grids and (505, 901) in grids
    assert {s.norb for _, _, s in stream} >= {1, 2, 3, 4, 5}
    print("served stream:", names, flush=True)
    # arrival seconds -> server steps (one step per clock unit)
    reqs = _serve([s for _, _, s in stream], [100.0 * a for a, _, _ in stream])
    assert any(r.admitted < max(q.finished for q in reqs) and r.admitted > 0 for r in reqs)
    for (a, name, spec), r in zip(stream, reqs):
        alone = _serve([spec], [0.0])[0].result
        s64 = type(spec).from_buffer_copy(spec)
        s64.prec = 0
        job = scf.Job(s64)
        try:
            job.prep()
            while not scf.iterate_batched([job])[0]:
                pass
            job.finish()
            fp64 = job.result()
        finally:
            job.close()
        m = r.result
        print(f"{name}: served E={m['etot']:.12f} it={m['iterations']} ({m['stop']}), alone it={alone['iterations']}, "
              f"FP64 E={fp64['etot']:.12f} dE={m['etot'] - fp64['etot']:.1e}", flush=True)
        assert m["iterations"] == alone["iterations"] and m["etot"] == alone["etot"]
        assert np.array_equal(m["eps"], alone["eps"])
        assert m["stop"] not in ("diverged", "not a number")
        assert abs(m["etot"] - fp64["etot"]) < 1e-8


def test_group_server_gives_the_single_server_results():
    """Two server groups on their own streams and host threads, fed from one
    queue, give every request the result it gets from a single server (and
    so, served alone), and both groups take requests."""
    from x2dhf_amd import serve
    specs = _specs()
    print("single server", flush=True)
    single = _serve(specs, [0.0] * len(specs))
    print("two groups", flush=True)
    grp = serve.GroupServer(groups=2, max_active=4)
    try:
        reqs = [grp.submit(s, arrival=0.0) for s in specs]
        grp.run()
        per_group = [len(srv.done) for srv in grp.servers]
    finally:
        grp.close()
    print(f"requests per group {per_group}")
    assert sorted(per_group) == [2, 2]
    for a, b in zip(single, reqs):
        assert b.result["iterations"] == a.result["iterations"]
        assert b.result["etot"] == a.result["etot"]
        assert np.array_equal(b.result["eps"], a.result["eps"])
