"""Continuous-batching SCF service on one GPU.

Requests (closed-shell diatomic HF jobs of any Z1, Z2, R, grid size and
precision) arrive over time.  The server keeps every admitted job's orbitals
and potentials resident in HBM and advances all admitted jobs together, one
batched SCF iteration per step (x2f_jobs_iterate_batched: every potential of
every job in one fused launch sequence, then every orbital, then the
normalisation / orthogonalisation / orbital energies with one device-to-host
transfer).  A job that converges is finished and freed at once, and queued
requests are admitted in arrival order as soon as a slot and enough device
memory are free, so new work joins the running batch instead of waiting for
it to drain.

Admission by footprint: a job is admitted when
    held + footprint(job) <= budget,
    footprint = spec_bytes + WORKSPACE_GROWTH * workspace_bytes,
where spec_bytes counts the job's own arrays (x2f_spec_bytes) and
workspace_bytes its share of the library's relaxation and quadrature
workspace (x2f_spec_workspace_bytes, from x2g_relax_workspace_bytes), with
room for the workspace's geometric growth; budget defaults to a fraction of
the device memory free when the server starts.  A request whose set-up runs
out of device memory all the same goes back to the head of the queue and
waits for a running job to retire.

Every job's result is independent of which other jobs share its batch: the
sweeps, quadratures and linear combinations of a job read and write only that
job's arrays, with a partition of the work that does not depend on the batch.
"""
import time
from dataclasses import dataclass, field

from . import gpu, scf

# the workspace grows by 1.5x steps, so a batch may hold up to 1.5x its sum
WORKSPACE_GROWTH = 1.5


@dataclass
class Request:
    rid: int
    spec: object
    arrival: float                 # seconds on the server clock
    footprint: int = 0             # bytes reserved while admitted
    admitted: float = None
    finished: float = None
    result: dict = None
    job: object = field(default=None, repr=False)

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class Server:
    """Continuous-batching server.  max_active caps the jobs stepped together
    (slots); hbm_budget (bytes) caps the footprint of the admitted jobs."""

    def __init__(self, device=0, max_active=16, hbm_budget=None, hbm_fraction=0.8, clock=time.perf_counter):
        self.ctx = gpu.Context(device)
        self.lib = gpu.load()
        if hbm_budget is None:
            free, total = gpu._szt(), gpu._szt()
            gpu.check(self.lib.x2g_mem_info(self.ctx.h, free, total))
            hbm_budget = int(hbm_fraction * free.value)
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.clock = clock
        self.t0 = clock()
        self.pending = []      # arrival order
        self.active = []
        self.done = []
        self.held = 0
        self.iterations = 0
        self.deferred = 0      # admissions put back for lack of device memory
        self._next_id = 0

    def now(self):
        return self.clock() - self.t0

    def submit(self, spec, arrival=None):
        """Queue a request arriving at `arrival` seconds on the server clock
        (default: now).  Raises ValueError for an unsupported specification or
        one that could never fit the budget."""
        fp = footprint(spec)
        if fp > self.budget:
            raise ValueError(f"request needs {fp} bytes, more than the budget {self.budget}")
        r = Request(self._next_id, spec, self.now() if arrival is None else float(arrival), fp)
        self._next_id += 1
        self.pending.append(r)
        self.pending.sort(key=lambda q: (q.arrival, q.rid))
        return r

    def _admit(self):
        t = self.now()
        while self.pending and self.pending[0].arrival <= t and len(self.active) < self.max_active:
            r = self.pending[0]
            if self.held + r.footprint > self.budget:
                break
            self.pending.pop(0)
            try:
                r.job = scf.Job(r.spec, ctx=self.ctx)
                r.job.prep()
            except (RuntimeError, ValueError) as e:
                # out of device memory after all: back to the head of the
                # queue until a running job retires (with nothing running,
                # the request can never fit and fails)
                if r.job is not None:
                    r.job.close()
                    r.job = None
                if not self.active:
                    raise RuntimeError(f"request {r.rid} cannot be set up on an idle device: {e}") from e
                self.pending.insert(0, r)
                self.deferred += 1
                break
            r.admitted = self.now()
            self.held += r.footprint
            self.active.append(r)

    def _retire(self, r):
        r.job.finish()
        r.result = r.job.result()
        r.job.close()
        r.job = None
        r.finished = self.now()
        self.held -= r.footprint
        self.done.append(r)

    def step(self):
        """Admit what can be admitted, then one batched iteration of every
        active job; returns the requests finished in this step."""
        self._admit()
        if not self.active:
            return []
        stopped = scf.iterate_batched([r.job for r in self.active])
        self.iterations += 1
        fin = [r for r, s in zip(self.active, stopped) if s]
        self.active = [r for r, s in zip(self.active, stopped) if not s]
        for r in fin:
            self._retire(r)
        return fin

    def run(self):
        """Serve until every submitted request is finished; idles (sleeps)
        until the next arrival when nothing is admitted."""
        while self.pending or self.active:
            if not self.active and self.pending and self.pending[0].arrival > self.now():
                time.sleep(min(0.05, self.pending[0].arrival - self.now()))
                continue
            self.step()
        return self.done

    def close(self):
        for r in self.active:
            if r.job is not None:
                r.job.close()
        self.active = []
        self.ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def footprint(spec):
    """Device bytes reserved for a job while it is admitted: its own arrays
    and its share of the library workspace (with room for growth)."""
    return int(scf.spec_bytes(spec) + WORKSPACE_GROWTH * scf.spec_workspace_bytes(spec))


# closed-shell diatomics of a served stream: (name, spec builder, equilibrium R)
STREAM_MOLECULES = (("H2", "h2_spec", 1.4), ("LiH", "lih_spec", 3.015), ("BH", "bh_spec", 2.3289),
                    ("FH", "fh_spec", 1.733))
STREAM_GRIDS = ((169, 241), (337, 601), (505, 901))


def poisson_stream(count, rate, seed=0, prec=1, rinf=40.0):
    """A seeded stream of requests: exponential inter-arrival times of mean
    1/rate seconds; each request a molecule of STREAM_MOLECULES with R drawn
    uniformly within 10% of its equilibrium value and a grid of STREAM_GRIDS,
    converged to orbital-energy and norm changes below 1e-12.
    Returns [(arrival_s, name, spec)]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for _ in range(count):
        t += float(rng.exponential(1.0 / rate))
        name, builder, req = STREAM_MOLECULES[int(rng.integers(len(STREAM_MOLECULES)))]
        nni, nmu = STREAM_GRIDS[int(rng.integers(len(STREAM_GRIDS)))]
        r = req * float(rng.uniform(0.9, 1.1))
        spec = getattr(scf, builder)(nni=nni, nmu=nmu, rinf=rinf, r=r, prec=prec, schedule=1, ienterm=12,
                                     inoterm=12, maxscf=8000)
        out.append((t, name, spec))
    return out


def percentile(values, q):
    """q-th percentile (0..100) by linear interpolation between order statistics."""
    v = sorted(values)
    if not v:
        return None
    k = (len(v) - 1) * q / 100.0
    lo = int(k)
    hi = min(lo + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)
