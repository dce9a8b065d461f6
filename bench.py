#!/usr/bin/env python3
"""Throughput benchmark of the MCSOR relaxation path on one MI355X.

Default workload (--workload stream): the served stream of closed-shell
diatomic HF requests -- Poisson arrivals at 20/s (seeded) of neutral
closed-shell diatomics with Z1, Z2 in [1, 10], R in [1, 6] bohr, 1-5 orbitals,
grids from 169x241 to 505x901 (x2dhf_amd.serve.poisson_stream), each
converged to orbital-energy changes below 1e-11 by the
continuous-batching server with FP32 MCSOR correction sweeps and FP64
residuals, norms, orthogonalisation, Lagrange multipliers and energies.  One
step = one stream of --count requests.  value = converged solves per hour.
Reported beside it: p50/p99 arrival-to-converged-energy latency, the same
requests served with FP64 sweeps (batched), the plain FP64 GPU path
(reference orbital order, one request at a time) on a sample of them, the
largest |dE| between the precisions, the roofline of the dominant sweep kernel
from HIP events around its launches, and the reference program (its
pthreads thread-pool build, compiled from its own sources) timed to
convergence on a bounded sample of the requests on this host's cores.

--workload n2: closed-shell N2 (Z1 = Z2 = 7, R = 2.068 bohr, 6 orbitals: 1 pi-u,
5 sigma), prolate spheroidal grid 331 x 601, R_inf = 40 bohr, hydrogenic LCAO
start, SCF to orbital-energy changes below 1e-12 (the reference's own N2
deck, tests/golden/decks/config2_n2_331x601.data; 337 x 601 rounds to 331 in
the reference).  Every SCF iteration relaxes 22 Coulomb/exchange potentials
and 6 orbitals with 10 MCSOR sweeps each.

One step = one wave of --batch requests of that workload served to
convergence by the continuous-batching server (x2dhf_amd.serve): FP32 MCSOR
correction sweeps with FP64 residuals, normalisation, orthogonalisation,
Lagrange multipliers and energies, all jobs of the wave stepped together.
value = converged solves per hour over the timed steps.  Reported beside it:
the plain FP64 GPU path (reference orbital order, one job, FP64 sweeps) on the
same device, p50/p99 time-to-converged-energy, |E - E_ref| against the
reference's FP64 CPU energy, the roofline of the dominant sweep kernel from
HIP events around its launches, and the reference's own CPU build timed on
this host's cores.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DECK = os.path.join(HERE, "tests", "golden", "decks", "config2_n2_331x601.data")
GOLDEN = os.path.join(HERE, "tests", "golden", "scf_config2_n2_331x601.json")
REF_EXE = os.path.join(HERE, "oracle", "_ref", "x2dhf_omp")
TRAFFIC = os.path.join(HERE, "profiles", "traffic_sweep_n2_331x601.json")
# the stream's dominant sweep launches: 2-row blocks of ring stride 528
TRAFFIC_STREAM = os.path.join(HERE, "profiles", "traffic_sweep_stream_mix.json")
# jobs the stream's server steps together (32 measured 3% faster than 16 on
# the default stream; 64 the same as 32)
STREAM_SLOTS = 32
# problems of the stream traffic probe: (nni, nmu, count), a mix of the
# stream's grid classes whose relaxed points (24.6 M, 394 MB of algorithmic
# bytes per FP32 sweep) match the stream's average sweep launch
STREAM_PROBE = ((505, 901, 28), (409, 721, 20), (361, 601, 20), (217, 361, 16), (169, 241, 16))
HBM_PEAK_GBS = 8000.0
METRIC = "converged SCF solves/hr + p50/p99 latency vs plain FP64 GPU at ΔE<1e-8 Eh; GB/s"


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed waves")
    ap.add_argument("--warmup", type=int, default=1, help="untimed waves")
    ap.add_argument("--batch", type=int, default=16, help="requests per wave (all admitted together)")
    ap.add_argument("--prec", type=int, default=1, choices=(0, 1), help="1: FP32 correction sweeps, 0: FP64")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 GPU solve")
    ap.add_argument("--no-cpu", action="store_true", help="skip the reference CPU baseline")
    ap.add_argument("--cpu-iters", default="61,81", help="SCF iterations of the two timed reference runs")
    ap.add_argument("--workload", choices=("stream", "n2", "poisson", "ordered"), default="stream",
                    help="stream: Poisson-arrival SCF requests of closed-shell diatomics (the headline line); "
                         "n2: waves of the N2 331x601 deck; poisson: one batch of seeded synthetic Poisson "
                         "problems on mixed grids; ordered: the reference's default relaxation (10 SOR "
                         "sweeps, middle ordering) of one 331x601 potential, pipelined against one level "
                         "per barrier")
    ap.add_argument("--count", type=int, default=None,
                    help="problems (poisson, default 256) or requests per stream (stream, default 64)")
    ap.add_argument("--rate", type=float, default=20.0, help="stream: mean arrivals per second")
    ap.add_argument("--plain-sample", type=int, default=6,
                    help="stream: requests also run one at a time on the plain FP64 GPU path")
    ap.add_argument("--no-graphs", action="store_true",
                    help="stream: skip serving the requests again with graph replay of the relaxations")
    ap.add_argument("--no-fp64", action="store_true", help="stream: skip serving the requests with FP64 sweeps")
    ap.add_argument("--cpu-seconds", type=float, default=25.0,
                    help="stream: CPU-baseline sample budget (whole converged reference runs, smallest first)")
    ap.add_argument("--seed", type=int, default=0, help="poisson / stream: seed of the synthetic workload")
    ap.add_argument("--traffic-probe", nargs="?", const="n2", choices=["n2", "stream"],
                    help="only run a copy kernel of known size and a few sweeps (for rocprofv3 --pmc): n2 = "
                         "--batch x 22 N2 331x601 potentials (FP32 and FP64), stream = the stream's widest "
                         "launch class (12 problems on 505x901 and 12 on 361x601, FP32, 2-row blocks)")
    ap.add_argument("--traffic-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="turn the counter_collection.csv files of the FETCH_SIZE and WRITE_SIZE runs of "
                         "--traffic-probe into the per-launch HBM traffic file read by the benchmark")
    ap.add_argument("--probe", choices=["n2", "stream"], default="n2",
                    help="which --traffic-probe the --traffic-summary files come from")
    return ap.parse_args()


def n2_spec(prec, schedule, maxscf=8000):
    from x2dhf_amd import scf
    return scf.n2_spec(331, 601, 40.0, ienterm=12, inoterm=12, maxscf=maxscf, prec=prec, schedule=schedule)


def plain_fp64():
    """One solve of the plain FP64 GPU path (reference order, FP64 sweeps)."""
    from x2dhf_amd import scf
    job = scf.Job(n2_spec(0, 0))
    try:
        t0 = time.perf_counter()
        res = job.run()
        return time.perf_counter() - t0, res
    finally:
        job.close()


def cpu_baseline(total_iters, iters):
    """The reference program (built from its sources by oracle/ref.mk) on the
    same deck: two runs of k1 < k2 SCF iterations (both past the frozen-orbital
    start), per-iteration time from their difference, extrapolated to the
    reference's own iteration count of the converged run."""
    k1, k2 = iters
    threads = min(16, os.cpu_count() or 1)
    sample = (f"reference x2dhf (OpenMP build of its own sources) on the N2 331x601 deck, runs of {k1} and {k2} "
              f"SCF iterations; per-iteration time x {total_iters} iterations of its converged run + set-up")
    if not os.path.exists(REF_EXE):
        return {"value": None, "unit": "solves/hr", "cores": threads, "kind": "reference",
                "sample": "reference binary not built here (oracle/_ref/x2dhf_omp)"}
    deck = open(DECK).read()
    times = []
    with tempfile.TemporaryDirectory() as d:
        for k in (k1, k2):
            path = os.path.join(d, f"deck{k}")
            with open(path, "w") as f:
                f.write(deck.replace("scf 8000 ", f"scf {k} "))
            t0 = time.perf_counter()
            with open(path) as fin, open(os.path.join(d, f"out{k}.lst"), "w") as fout:
                subprocess.run([REF_EXE], stdin=fin, stdout=fout, stderr=subprocess.STDOUT, cwd=d, check=True,
                               env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
            times.append(time.perf_counter() - t0)
            print(f"# cpu baseline: {k} iterations in {times[-1]:.2f} s ({threads} threads)", file=sys.stderr,
                  flush=True)
    per_it = (times[1] - times[0]) / (k2 - k1)
    setup = max(times[0] - k1 * per_it, 0.0)
    solve = setup + total_iters * per_it
    return {"value": 3600.0 / solve, "unit": "solves/hr", "cores": threads, "kind": "reference",
            "sample": sample, "seconds_per_iteration": per_it, "estimated_solve_s": solve}


def traffic_probe(batch):
    """For the rocprofv3 counter runs: a copy kernel of known size (x2g_lincomb
    out = 1.0*a on a 1201 x 1201 grid: 11.5 MB read + 11.5 MB written per
    launch, three launches), then one relaxation call (5 sweeps) in FP32 and
    one in FP64 of --batch x 22 potentials on N2 331x601 grids: the sweep
    launches of --batch jobs' potential relaxation in the benchmark."""
    import numpy as np
    from x2dhf_amd import gpu, scf
    ctx = gpu.Context(0)
    gpu.check(ctx.lib.x2g_set_autotune(0))  # the default plan (4-row blocks), no tuning launches
    rng = np.random.default_rng(0)
    cb = scf.grid_build(1201, 1201, 7.0, 7.0, 2.068, 40.0)
    big = gpu.Grid(ctx, 1201, 1201, *cb)
    a, out = ctx.upload(rng.standard_normal(big.n)), ctx.alloc(big.n)
    copy = (gpu.Term * 1)(gpu.Term(1.0, a.ptr, None, None))
    for _ in range(3):
        gpu.check(ctx.lib.x2g_lincomb(big.h, out.ptr, None, 0.0, copy, 1))
    gpu.check(ctx.lib.x2g_ctx_sync(ctx.h))
    coef, suppl, vxi, veta = scf.grid_build(331, 601, 7.0, 7.0, 2.068, 40.0)
    gs = [gpu.Grid(ctx, 331, 601, coef, suppl, vxi, veta) for _ in range(batch)]
    gh = (gpu._vp * batch)(*[g.h.value for g in gs])
    items, keep = [], []
    lhs = ctx.upload(suppl[7])
    for q in range(batch):
        for k in range(22):
            f, r = ctx.upload(0.01 * rng.standard_normal(gs[q].n)), ctx.upload(rng.standard_normal(gs[q].n))
            keep += [f, r]
            items.append(gpu.RelaxItem(f.ptr, lhs.ptr, r.ptr, None, 1.9, 1, 1, 1, q))
    n = len(items)
    gpu.check(ctx.lib.x2g_relax_batch(ctx.h, gh, (gpu.RelaxItem * n)(*items), n, 1, 5))
    for it in items:
        it.prec = 0
    gpu.check(ctx.lib.x2g_relax_batch(ctx.h, gh, (gpu.RelaxItem * n)(*items), n, 1, 5))
    gpu.check(ctx.lib.x2g_ctx_sync(ctx.h))
    for g in gs:
        g.destroy()
    big.destroy()
    ctx.close()
    print("traffic probe done", flush=True)


def traffic_probe_stream():
    """For the rocprofv3 counter runs: the copy kernel of traffic_probe, then
    one relaxation call (5 FP32 sweeps) of the problems of STREAM_PROBE in one
    batch, with the stream's usual plan for this class fixed (2-row blocks,
    as many workgroups per CU as the LDS holds; no autotuning launches)."""
    import numpy as np
    from x2dhf_amd import gpu, scf
    ctx = gpu.Context(0)
    lib = ctx.lib
    gpu.check(lib.x2g_set_autotune(0))
    gpu.check(lib.x2g_set_sweep_plan(2, 0, 0))
    rng = np.random.default_rng(0)
    cb = scf.grid_build(1201, 1201, 7.0, 7.0, 2.068, 40.0)
    big = gpu.Grid(ctx, 1201, 1201, *cb)
    a, out = ctx.upload(rng.standard_normal(big.n)), ctx.alloc(big.n)
    copy = (gpu.Term * 1)(gpu.Term(1.0, a.ptr, None, None))
    for _ in range(3):
        gpu.check(lib.x2g_lincomb(big.h, out.ptr, None, 0.0, copy, 1))
    gpu.check(lib.x2g_ctx_sync(ctx.h))
    gs, items, keep = [], [], []
    for nni, nmu, cnt in STREAM_PROBE:
        coef, suppl, vxi, veta = scf.grid_build(nni, nmu, 5.0, 3.0, 2.0, 40.0)
        g = gpu.Grid(ctx, nni, nmu, coef, suppl, vxi, veta)
        lhs = ctx.upload(suppl[7])
        keep.append(lhs)
        for _ in range(cnt):
            f, r = ctx.upload(0.01 * rng.standard_normal(g.n)), ctx.upload(rng.standard_normal(g.n))
            keep += [f, r]
            items.append(gpu.RelaxItem(f.ptr, lhs.ptr, r.ptr, None, 1.9, 1, 1, 1, len(gs)))
        gs.append(g)
    gh = (gpu._vp * len(gs))(*[g.h.value for g in gs])
    n = len(items)
    gpu.check(lib.x2g_relax_batch(ctx.h, gh, (gpu.RelaxItem * n)(*items), n, 1, 5))
    gpu.check(lib.x2g_ctx_sync(ctx.h))
    for g in gs:
        g.destroy()
    big.destroy()
    gpu.check(lib.x2g_set_sweep_plan(0, 0, 0))
    ctx.close()
    print("traffic probe done", flush=True)


def traffic_summary_stream(fetch_csv, write_csv):
    """HBM bytes per sweep launch of traffic_probe_stream, as traffic_summary:
    counters scaled by the copy kernel, averaged over the four FP32 sweeps
    that read their source (the first takes its zero start)."""
    import csv
    n_copy = 1201 * 1201 * 8

    def per_kernel(path):
        rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))

        def named(r, k):
            n = r["Kernel_Name"]
            return n.startswith(k) or f"::{k}(" in n or f"::{k}<" in n

        copy = [float(r["Counter_Value"]) * 1024 for r in rows if named(r, "k_lincomb")]
        sweep = [float(r["Counter_Value"]) * 1024 for r in rows if named(r, "k_sweep")]
        if len(copy) != 3 or len(sweep) != 5:
            raise SystemExit(f"{path}: expected 3 copy and 5 sweep dispatches, found {len(copy)} and {len(sweep)}")
        return sum(copy) / len(copy), sweep[1:]

    fc, fv = per_kernel(fetch_csv)
    wc, wv = per_kernel(write_csv)
    kf, kw = n_copy / fc, n_copy / wc
    algo = sum(16.0 * cnt * (nni - 2) * (nmu - 5) for nni, nmu, cnt in STREAM_PROBE)
    rd, wr = kf * sum(fv) / len(fv), kw * sum(wv) / len(wv)
    out = {"workload": " + ".join(f"{cnt} x {nni}x{nmu}" for nni, nmu, cnt in STREAM_PROBE) +
           " FP32 correction problems, one k_sweep<float,1,528,2> launch", "fetch_factor": kf, "write_factor": kw,
           "fp32": {"hbm_read_bytes": rd, "hbm_write_bytes": wr, "hbm_bytes_per_launch": rd + wr,
                    "algorithmic_bytes_per_launch": algo, "traffic_over_algorithmic": (rd + wr) / algo}}
    with open(TRAFFIC_STREAM, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def traffic_summary(fetch_csv, write_csv, batch):
    """HBM bytes per sweep launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs of --traffic-probe.  Each counter (KiB per dispatch) is scaled by
    what it reads for the probe's copy kernel (x2g_lincomb, bytes known), and
    averaged over the probe's FP32 sweep launches that read their source
    (launches 2-5: the first sweep of a correction takes its zero start
    without reading it) and its FP64 launches (the last five)."""
    import csv
    n_copy = 1201 * 1201 * 8  # bytes read and written by one copy launch

    def per_kernel(path):
        rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))
        # kernel names come demangled ("x2g::k_lincomb(double*, ...)") or bare ("k_lincomb")
        def named(r, k):
            n = r["Kernel_Name"]
            return n.startswith(k) or f"::{k}(" in n or f"::{k}<" in n

        copy = [float(r["Counter_Value"]) * 1024 for r in rows if named(r, "k_lincomb")]
        sweep = [float(r["Counter_Value"]) * 1024 for r in rows if named(r, "k_sweep")]
        if len(copy) != 3 or len(sweep) != 10:
            raise SystemExit(f"{path}: expected 3 copy and 10 sweep dispatches, found {len(copy)} and {len(sweep)}")
        return sum(copy) / len(copy), sweep[1:5], sweep[5:]

    fc, f32, f64 = per_kernel(fetch_csv)
    wc, w32, w64 = per_kernel(write_csv)
    kf, kw = n_copy / fc, n_copy / wc  # counter -> bytes factors from the copy kernel
    out = {"workload": f"{batch} x 22 potentials, N2 331x601, one sweep launch", "fetch_factor": kf,
           "write_factor": kw}
    algo = {4: 0.0, 8: 0.0}
    for t in (4, 8):
        algo[t] = batch * 22 * 4.0 * t * (331 - 2) * (601 - 5)
    for key, fv, wv, t in (("fp32", f32, w32, 4), ("fp64", f64, w64, 8)):
        rd = kf * sum(fv) / len(fv)
        wr = kw * sum(wv) / len(wv)
        out[key] = {"hbm_read_bytes": rd, "hbm_write_bytes": wr, "hbm_bytes_per_launch": rd + wr,
                    "algorithmic_bytes_per_launch": algo[t], "traffic_over_algorithmic": (rd + wr) / algo[t]}
    os.makedirs(os.path.dirname(TRAFFIC), exist_ok=True)
    with open(TRAFFIC, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def sweep_roofline(st, prec):
    """Roofline object of the sweep kernel of precision prec (1 FP32, 0 FP64)
    from the library's HIP-event statistics [launches, ms, bytes] x 2."""
    k = 1 if prec else 0
    launches, ms, nbytes = st[3 * k], st[3 * k + 1], st[3 * k + 2]
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms else None
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
            "kernel": "k_sweep<float>" if k else "k_sweep<double>",
            "algorithmic_bytes_per_launch": nbytes / launches if launches else None,
            "avg_launch_us": ms / launches * 1e3 if launches else None, "launches": launches}


def bench_ordered(args):
    """The reference's default relaxation on one problem: 10 SOR sweeps in the
    middle ordering (x2g_relax_ordered) of a Coulomb-type potential equation
    (lhs = f3 + diag) on the 331x601 grid of the He deck test-sets/he/set-03
    (Z = 2, R = 2 bohr, R_inf 200), with the sweeps pipelined (the default)
    and with one dependency level of one sweep per barrier; the two results
    must be identical.  Reported: sweeps per second of the pipelined schedule."""
    import numpy as np
    from x2dhf_amd import gpu, scf
    steps, warm = max(args.steps, 1) * 20, max(args.warmup, 1) * 3
    ctx = gpu.Context(0)
    coef, suppl, vxi, veta = scf.grid_build(331, 601, 2.0, 0.0, 2.0, 200.0)
    g = gpu.Grid(ctx, 331, 601, coef, suppl, vxi, veta)
    rng = np.random.default_rng(args.seed)
    X, Y = np.meshgrid(np.arange(331), np.arange(601))
    rhs = (np.exp(-((X - 160.0) ** 2 + (Y - 40.0) ** 2) / 800.0) + 1e-3 * rng.standard_normal(X.shape)).ravel()
    lhs_d, rhs_d = ctx.upload(suppl[7]), ctx.upload(rhs)
    work = ctx.alloc(g.npad)
    fun = ctx.upload(np.zeros(g.n))
    res, outs = {}, {}
    for pipe in (1, 0):
        gpu.check(ctx.lib.x2g_set_levels_pipeline(pipe))
        gpu.check(ctx.lib.x2g_putin(g.h, 1, fun.ptr, work.ptr))
        for _ in range(warm):
            gpu.check(ctx.lib.x2g_relax_ordered(g.h, 1, 1, work.ptr, lhs_d.ptr, rhs_d.ptr, 1.95, 10, 1, 1))
        gpu.check(ctx.lib.x2g_ctx_sync(ctx.h))
        t0 = time.perf_counter()
        for _ in range(steps):
            gpu.check(ctx.lib.x2g_relax_ordered(g.h, 1, 1, work.ptr, lhs_d.ptr, rhs_d.ptr, 1.95, 10, 1, 1))
        gpu.check(ctx.lib.x2g_ctx_sync(ctx.h))
        res[pipe] = (time.perf_counter() - t0) / steps
        outs[pipe] = work.get()
    gpu.check(ctx.lib.x2g_set_levels_pipeline(1))
    identical = bool(np.array_equal(outs[0], outs[1]))
    line = {
        "metric": "ordered SOR sweeps/s (one 331x601 problem, middle ordering, bit-exact sequential order)",
        "value": 10.0 / res[1], "unit": "sweeps/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": res[1] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": f"synthetic rhs (seed {args.seed}) on the he/set-03 grid",
        "config": {"workload": "10 SOR sweeps per call, middle ordering, 331x601, omega 1.95, even symmetry"},
        "per_level_barrier_ms": res[0] * 1e3, "speedup_vs_per_level": res[0] / res[1],
        "results_identical": identical,
    }
    g.destroy()
    ctx.close()
    print(json.dumps(line), flush=True)
    if not identical:
        sys.exit(1)


def bench_poisson(args):
    """Configuration 4: one batch of seeded synthetic Poisson problems (grids
    drawn from 91x121, 169x241, 337x601, 505x901; R in [1,6] bohr; densities
    of 1-8 Gaussians), FP32 correction sweeps with FP64 residuals, relaxed
    together until ||rhs - A u|| / ||rhs|| <= 1e-12 or no further progress;
    a problem is solved at 1e-12 or, where its FP64 floor lies higher, at
    FLOOR_FACTOR x that floor (x2dhf_amd.poisson)."""
    from x2dhf_amd import poisson
    count = args.count or 256
    out = poisson.run(count=count, seed=args.seed, prec=args.prec, tol=1e-12)
    st = out["stats"]
    sweeps = sum(out["calls"]) * 10
    line = {
        "metric": "converged synthetic Poisson solves/s (mixed grids, FP32 sweeps + FP64 residuals)",
        "value": out["solved"] / out["solve_s"], "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": out["solve_s"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.prec == 1 else "f64",
        "data": f"synthetic: {count} seeded problems (seed {args.seed}), densities of 1-8 Gaussians",
        "config": {"workload": "batched synthetic Poisson problems (Coulomb potential equation, Delta m = 0) "
                               "on grids 91x121 / 169x241 / 337x601 / 505x901, R in [1,6] bohr, Dirichlet "
                               "rows from the multipole moments, zero start, 10 MCSOR sweeps per relaxation call",
                   "problems": count, "tolerance": 1e-12},
        "max_relative_residual": out["max_residual"],
        "solved": out["solved"], "solved_at_1e-12": out["at_tol"],
        "max_residual_over_fp64_floor": max(r / f for r, f in zip(out["residuals"], out["floors"]) if f > 0)
        if any(f > 0 for f in out["floors"]) else None,
        "setup_s": out["setup_s"], "problem_sweeps": sweeps,
        "roofline": sweep_roofline(st, args.prec),
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def serve_stream(reqs, prec, stats=False, graphs=False, groups=1):
    """Serve (arrival_s, name, spec) requests with the continuous-batching
    server at the given precision (graphs: relaxations replayed from captured
    HIP graphs; groups > 1: that many server groups on their own streams and
    host threads, serve.GroupServer); returns (handles, elapsed_s, sweep
    stats summed over the groups' contexts)."""
    import ctypes as C
    from x2dhf_amd import gpu, scf, serve
    if groups > 1:
        server = serve.GroupServer(groups=groups, max_active=STREAM_SLOTS, graphs=graphs)
        ctxs = [srv.ctx for srv in server.servers]
    else:
        server = serve.Server(max_active=STREAM_SLOTS, graphs=graphs)
        ctxs = [server.ctx]
    lib = gpu.load()
    if stats:
        for c in ctxs:
            gpu.check(lib.x2g_stats_enable(c.h, 1))
    t0 = server.now()
    handles = []
    for t, name, spec in reqs:
        sp = scf.JobSpec.from_buffer_copy(spec)
        sp.prec = prec
        handles.append(server.submit(sp, arrival=t0 + t))
    server.run()
    for c in ctxs:
        gpu.check(lib.x2g_ctx_sync(c.h))
    elapsed = max(h.finished for h in handles) - t0
    tot = [0.0] * 6
    if stats:
        for c in ctxs:
            st = (C.c_double * 6)()
            gpu.check(lib.x2g_stats_get(c.h, st))
            tot = [a + b for a, b in zip(tot, st)]
    server.close()
    return handles, elapsed, tot


def cpu_baseline_stream(reqs, budget_s):
    """The reference program's OpenMP build (oracle/ref.mk omp, compiled from
    its own sources) on a bounded sample of the stream: whole runs of the
    requests in increasing grid size, orbitals and potentials by MCSOR on T
    threads (its mcsor / mcsor-ce cards, OMP_NUM_THREADS = T), one request at
    a time, until budget_s of wall time is spent; a run without a finite
    converged energy counts its time and no solve.  (The pthreads builds are
    not used: the PTHREAD build aborts and the thread-pool build's orbital
    sweeps diverge to NaN on these decks.)  The reference rounds each grid
    down to 1 mod 30 (its nnucalc / nmucalc), so its problems are up to 9%
    smaller than the GPU's."""
    from x2dhf_amd import deck
    cores = os.cpu_count() or 1
    threads = max(1, min(16, cores))
    base = {"unit": "solves/hr", "cores": threads, "host_cores": cores, "kind": "reference"}
    if not os.path.exists(REF_EXE):
        return dict(base, value=None, sample="reference binary not built here (oracle/_ref/x2dhf_omp)")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    order = sorted(range(len(reqs)), key=lambda i: reqs[i][2].nni * reqs[i][2].nmu)
    done, wall, names, failed = 0, 0.0, [], []
    with tempfile.TemporaryDirectory() as d:
        for i in order:
            if wall >= budget_s:
                break
            t, name, spec = reqs[i]
            path = os.path.join(d, f"deck{i}")
            with open(path, "w") as f:
                f.write(deck.to_deck(spec, title=name, mcsor_threads=threads))
            t0 = time.perf_counter()
            why = None
            try:
                with open(path) as fin, open(os.path.join(d, f"out{i}.lst"), "w") as fout:
                    subprocess.run([REF_EXE], stdin=fin, stdout=fout, stderr=subprocess.STDOUT, cwd=d,
                                   check=True, timeout=max(20.0, budget_s), env=env)
            except subprocess.TimeoutExpired:
                why = "no result within the time limit"
            except subprocess.CalledProcessError as e:
                why = f"exit status {e.returncode}"
            dt = time.perf_counter() - t0
            # the time of a run without a converged energy counts, with no solve
            wall += dt
            if why is None:
                txt = open(os.path.join(d, f"out{i}.lst")).read()
                etot = [ln for ln in txt.splitlines() if "total energy:" in ln]
                if not etot or "nan" in etot[-1].lower() or "iteration limit" in txt:
                    why = "no finite converged energy"
            if why:
                failed.append(f"{name} {dt:.1f}s ({why})")
                print(f"# cpu baseline: {name}: {why} after {dt:.1f} s", file=sys.stderr, flush=True)
                continue
            done += 1
            names.append(f"{name} {dt:.1f}s")
            print(f"# cpu baseline: {name} converged in {dt:.2f} s ({threads} threads)", file=sys.stderr,
                  flush=True)
    return dict(base, value=3600.0 * done / wall if wall else None, solves=done, wall_s=wall,
                runs_without_energy=failed,
                sample=f"reference x2dhf, OpenMP build of its own sources, MCSOR on {threads} threads for "
                       f"orbitals and potentials, col-wise order: the stream's requests "
                       f"in increasing grid size, one at a time, until {budget_s:.0f} s of wall time: {done} "
                       f"converged (" + "; ".join(names) + f"), {len(failed)} without a finite converged "
                       f"energy, whose time counts too")


def autotune_summary(report):
    """The sweep plans the autotuner chose: batches served by each choice of
    (block rows, workgroups per CU, split) over the tuned batch shapes."""
    import re
    chosen = {}
    for ln in report:
        m = re.search(r"br (\d+) wpc (\d+) split (\d+) \(batches (\d+)\)", ln)
        if m:
            k = f"br{m.group(1)} wpc{m.group(2)} split{m.group(3)}"
            chosen[k] = chosen.get(k, 0) + int(m.group(4))
    return {"shapes": len(report), "batches_by_plan": chosen}


def bench_stream(args):
    """Configuration 5, the headline: Poisson-arrival SCF requests (seeded,
    --rate per second) of closed-shell diatomics with Z1, Z2 in [1, 10], R in
    [1, 6] bohr, 1-5 orbitals, grids 169x241 .. 505x901, FP32 correction
    sweeps, served by the continuous-batching server; latency = arrival ->
    converged energy."""
    from x2dhf_amd import gpu, scf, serve
    count = args.count or 64
    for w in range(args.warmup):
        serve_stream(serve.poisson_stream(min(8, count), args.rate, seed=1000 + w, prec=args.prec), args.prec)
        print(f"# warmup stream {w + 1} done", file=sys.stderr, flush=True)
    timed, handles, elapsed, st = [], [], 0.0, [0.0] * 6
    for k in range(args.steps):
        reqs = serve.poisson_stream(count, args.rate, seed=args.seed + k, prec=args.prec)
        h, e, s = serve_stream(reqs, args.prec, stats=True)
        timed += reqs
        handles += h
        elapsed += e
        st = [a + b for a, b in zip(st, s)]
        print(f"# timed stream {k + 1}: {count} requests in {e:.2f} s", file=sys.stderr, flush=True)
    conv = [h.result["stop"] for h in handles]
    ok = [serve.converged(h.result) for h in handles]
    nconv = sum(ok)
    lat = [h.latency for h in handles]
    value = nconv / elapsed * 3600.0
    fp64 = None
    dE64 = []
    if not args.no_fp64 and args.prec == 1:
        h64_all, e64_all = [], 0.0
        for k in range(args.steps):
            reqs = serve.poisson_stream(count, args.rate, seed=args.seed + k, prec=0)
            h64, e64, _ = serve_stream(reqs, 0)
            h64_all += h64
            e64_all += e64
        dE64 = [abs(a.result["etot"] - b.result["etot"]) for a, b, c in zip(handles, h64_all, ok)
                if c and serve.converged(b.result)]
        n64 = sum(1 for h in h64_all if serve.converged(h.result))
        fp64 = {"solves_per_hr": n64 / e64_all * 3600.0, "elapsed_s": e64_all,
                "p50_latency_s": serve.percentile([h.latency for h in h64_all], 50),
                "p99_latency_s": serve.percentile([h.latency for h in h64_all], 99),
                "max_abs_dE_mixed_vs_fp64": max(dE64) if dE64 else None,
                "compared": len(dE64)}
        print(f"# FP64 served: {e64_all:.2f} s, max |dE| mixed vs FP64 {fp64['max_abs_dE_mixed_vs_fp64']} over "
              f"{len(dE64)} requests converged in both", file=sys.stderr, flush=True)
    graphs = None
    if not args.no_graphs:
        # the same requests again, relaxations replayed from captured HIP
        # graphs (sweep statistics off: they time eager launches only)
        hg_all, eg_all, caps, reps = [], 0.0, 0, 0
        for k in range(args.steps):
            reqs = serve.poisson_stream(count, args.rate, seed=args.seed + k, prec=args.prec)
            hg, eg, _ = serve_stream(reqs, args.prec, graphs=True)
            hg_all += hg
            eg_all += eg
        ng = sum(1 for h in hg_all if serve.converged(h.result))
        same = all(a.result["etot"] == b.result["etot"] for a, b in zip(handles, hg_all))
        graphs = {"solves_per_hr": ng / eg_all * 3600.0, "elapsed_s": eg_all,
                  "p50_latency_s": serve.percentile([h.latency for h in hg_all], 50),
                  "energies_identical_to_eager": same}
        print(f"# graph replay: {eg_all:.2f} s, {graphs['solves_per_hr']:.0f} solves/hr, energies identical: {same}",
              file=sys.stderr, flush=True)
    groups2 = None
    if not args.no_graphs:
        # the same requests again by two server groups on their own streams and
        # host threads (serve.GroupServer), relaxations replayed from graphs:
        # one group's host turn overlaps the other's kernels (per-launch sweep
        # times are not comparable with concurrent streams, so the roofline
        # comes from the single-server timed stream above)
        h2_all, e2_all = [], 0.0
        for k in range(args.steps):
            reqs = serve.poisson_stream(count, args.rate, seed=args.seed + k, prec=args.prec)
            h2, e2, _ = serve_stream(reqs, args.prec, graphs=True, groups=2)
            h2_all += h2
            e2_all += e2
        n2 = sum(1 for h in h2_all if serve.converged(h.result))
        same2 = all(a.result["etot"] == b.result["etot"] for a, b in zip(handles, h2_all))
        groups2 = {"groups": 2, "slots": STREAM_SLOTS, "solves_per_hr": n2 / e2_all * 3600.0, "elapsed_s": e2_all,
                   "p50_latency_s": serve.percentile([h.latency for h in h2_all], 50),
                   "energies_identical_to_eager": same2}
        print(f"# two groups + graph replay: {e2_all:.2f} s, {groups2['solves_per_hr']:.0f} solves/hr, energies "
              f"identical: {same2}", file=sys.stderr, flush=True)
    plain = None
    if not args.no_plain and args.plain_sample > 0:
        plain_s, dEp, its = 0.0, [], 0
        for (t, name, spec), h in list(zip(timed, handles))[:args.plain_sample]:
            ps = scf.JobSpec.from_buffer_copy(spec)
            ps.prec, ps.schedule = 0, 0
            job = scf.Job(ps)
            try:
                t1 = time.perf_counter()
                res = job.run()
                plain_s += time.perf_counter() - t1
            finally:
                job.close()
            if serve.converged(res) and serve.converged(h.result):
                dEp.append(abs(res["etot"] - h.result["etot"]))
            its += res["iterations"]
        n = min(args.plain_sample, len(timed))
        plain = {"sample": f"the first {n} requests of the timed stream, one at a time", "solve_s_total": plain_s,
                 "seconds_per_solve": plain_s / n, "solves_per_hr": n / plain_s * 3600.0,
                 "max_abs_dE_vs_served": max(dEp) if dEp else None}
        print(f"# plain FP64 GPU, {n} requests one at a time: {plain_s:.2f} s, max |dE| {plain['max_abs_dE_vs_served']}",
              file=sys.stderr, flush=True)
    roof = sweep_roofline(st, args.prec)
    tfile = TRAFFIC_STREAM if args.prec else TRAFFIC
    if os.path.exists(tfile) and roof.get("algorithmic_bytes_per_launch"):
        ratio = json.load(open(tfile)).get("fp32" if args.prec else "fp64", {}).get("traffic_over_algorithmic")
        if ratio:
            roof["traffic"] = ratio * roof["algorithmic_bytes_per_launch"]
            roof["traffic_source"] = ("measured/algorithmic HBM bytes of the probe launch in " +
                                      os.path.relpath(tfile, HERE) + " (rocprofv3 FETCH_SIZE / WRITE_SIZE runs of "
                                      "bench.py --traffic-probe) times this run's average algorithmic bytes per "
                                      "launch")
    cpu = None if args.no_cpu else cpu_baseline_stream(timed, args.cpu_seconds)
    by_mol = {}
    for (t, name, spec), h in zip(timed, handles):
        by_mol.setdefault(name.split(" ")[0], []).append(h.latency)
    line = {
        "metric": METRIC,
        "value": value, "unit": "solves/hr", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if args.prec == 1 else "f64",
        "data": f"synthetic: {count} seeded requests per stream (seed {args.seed}+step), Poisson arrivals at "
                f"{args.rate}/s, hydrogenic LCAO starts",
        "config": {"workload": "served stream (configuration 5): neutral closed-shell diatomics, Z1 and Z2 in "
                               "[1,10], R in [1,6] bohr, 1-5 orbitals, grids 169x241..505x901, R_inf 40, HF SCF "
                               "to orbital-energy changes < 1e-11 (10 MCSOR sweeps per relaxation), "
                               f"continuous batching (at most {STREAM_SLOTS} jobs stepped together, admission by HBM "
                               "footprint)", "requests_per_step": count, "rate_per_s": args.rate,
                   "precision": "FP32 MCSOR correction sweeps, FP64 residual/norm/orthogonalisation/"
                                "Lagrange multipliers/energy" if args.prec else "FP64"},
        "converged": nconv, "requests": len(handles), "stops": {c: conv.count(c) for c in sorted(set(conv))},
        "p50_latency_s": serve.percentile(lat, 50),
        "p99_latency_s": serve.percentile(lat, 99),
        "latency_p50_by_molecule_s": {k: serve.percentile(v, 50) for k, v in sorted(by_mol.items())},
        "fp64_served": fp64,
        "speedup_vs_fp64_served": value / fp64["solves_per_hr"] if fp64 else None,
        "graph_replay": graphs,
        "two_groups_graph_replay": groups2,
        "plain_fp64_gpu": plain,
        "speedup_vs_plain_fp64_gpu": value / plain["solves_per_hr"] if plain else None,
        "max_abs_dE_vs_fp64": max(dE64 + ([plain["max_abs_dE_vs_served"]] if plain and plain["max_abs_dE_vs_served"]
                                          is not None else [])) if dE64 else None,
        "roofline": roof,
        "autotune": autotune_summary(gpu.tune_report()),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.traffic_summary:
        if args.probe == "stream":
            traffic_summary_stream(*args.traffic_summary)
        else:
            traffic_summary(*args.traffic_summary, args.batch)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe == "stream":
        traffic_probe_stream()
        return
    if args.traffic_probe:
        traffic_probe(args.batch)
        return
    if args.workload == "poisson":
        bench_poisson(args)
        return
    if args.workload == "stream":
        bench_stream(args)
        return
    if args.workload == "ordered":
        bench_ordered(args)
        return
    import ctypes as C
    from x2dhf_amd import gpu, serve
    E_ref = json.load(open(GOLDEN))["etot"]
    ref_iters = len(json.load(open(GOLDEN))["iterations"])

    plain = None
    if not args.no_plain:
        t_plain, res = plain_fp64()
        plain = {"solve_s": t_plain, "solves_per_hr": 3600.0 / t_plain, "iterations": res["iterations"],
                 "abs_dE": abs(res["etot"] - E_ref)}
        print(f"# plain FP64 GPU: {t_plain:.2f} s, {res['iterations']} iterations, |dE| {plain['abs_dE']:.2e}",
              file=sys.stderr, flush=True)

    server = serve.Server(max_active=args.batch)
    lib = server.lib

    def wave():
        t = server.now()
        reqs = [server.submit(n2_spec(args.prec, 1), arrival=t) for _ in range(args.batch)]
        server.run()
        return reqs

    for w in range(args.warmup):
        wave()
        print(f"# warmup wave {w + 1} done", file=sys.stderr, flush=True)
    gpu.check(lib.x2g_ctx_sync(server.ctx.h))
    gpu.check(lib.x2g_stats_enable(server.ctx.h, 1))
    t0 = time.perf_counter()
    timed = []
    for s in range(args.steps):
        timed += wave()
        print(f"# timed wave {s + 1} done", file=sys.stderr, flush=True)
    gpu.check(lib.x2g_ctx_sync(server.ctx.h))
    elapsed = time.perf_counter() - t0
    st = (C.c_double * 6)()
    gpu.check(lib.x2g_stats_get(server.ctx.h, st))
    server.close()

    lat = [r.latency for r in timed]
    dE = max(abs(r.result["etot"] - E_ref) for r in timed)
    value = len(timed) / elapsed * 3600.0
    k = 1 if st[4] >= st[1] else 0            # dominant sweep precision by time
    launches, ms, nbytes = st[3 * k], st[3 * k + 1], st[3 * k + 2]
    avg_us = ms / launches * 1e3 if launches else None
    per_launch = nbytes / launches if launches else None
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms else None
    # HBM traffic per launch: the ratio of measured to algorithmic bytes of the
    # 352-potential launch (rocprofv3 counters on --traffic-probe, stored by
    # --traffic-summary) applied to this run's average launch
    traffic, ratio = None, None
    if os.path.exists(TRAFFIC) and per_launch:
        ratio = json.load(open(TRAFFIC)).get("fp32" if k == 1 else "fp64", {}).get("traffic_over_algorithmic")
        traffic = ratio * per_launch if ratio else None
    cpu = None if args.no_cpu else cpu_baseline(ref_iters, [int(x) for x in args.cpu_iters.split(",")])
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/hr",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.prec == 1 else "f64",
        "data": "synthetic: the reference's N2 deck (hydrogenic LCAO start), identical requests",
        "config": {"workload": "N2 R=2.068 bohr, 6 closed-shell orbitals, grid 331x601, R_inf=40, "
                               "HF SCF to orbital-energy change < 1e-12 (10 MCSOR sweeps per relaxation)",
                   "requests_per_step": args.batch, "served": "continuous batching, all requests of a wave "
                   "arrive together", "precision": "FP32 MCSOR correction sweeps, FP64 residual/norm/"
                   "orthogonalisation/Lagrange multipliers/energy" if args.prec else "FP64"},
        "p50_latency_s": serve.percentile(lat, 50),
        "p99_latency_s": serve.percentile(lat, 99),
        "max_abs_dE_vs_fp64_oracle": dE,
        "plain_fp64_gpu": plain,
        "speedup_vs_plain_fp64_gpu": value / plain["solves_per_hr"] if plain else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "kernel": "k_sweep<float>" if k == 1 else "k_sweep<double>",
                     "algorithmic_bytes_per_launch": per_launch, "avg_launch_us": avg_us,
                     "launches": launches, "traffic_over_algorithmic": ratio},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
