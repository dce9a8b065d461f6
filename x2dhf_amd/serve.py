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
import contextlib
import threading
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
    (slots); hbm_budget (bytes) caps the footprint of the admitted jobs;
    graphs replays the batched relaxations from captured HIP graphs."""

    def __init__(self, device=0, max_active=16, hbm_budget=None, hbm_fraction=0.8, clock=time.perf_counter,
                 graphs=False):
        self.ctx = gpu.Context(device)
        self.lib = gpu.load()
        # graphs: replay each batch composition's relaxation launches from a
        # captured HIP graph (x2g_set_graphs)
        gpu.check(self.lib.x2g_set_graphs(self.ctx.h, 1 if graphs else 0))
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
        # guards `pending` when it is a queue shared with other servers
        # (GroupServer); a server of its own needs none
        self._queue_lock = contextlib.nullcontext()

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
        admitted = 0
        while len(self.active) < self.max_active:
            with self._queue_lock:
                if not self.pending or self.pending[0].arrival > t:
                    break
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
                with self._queue_lock:
                    self.pending.insert(0, r)
                self.deferred += 1
                break
            r.admitted = self.now()
            self.held += r.footprint
            self.active.append(r)
            admitted += 1
        return admitted

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
        while True:
            with self._queue_lock:
                head = self.pending[0].arrival if self.pending else None
            if head is None and not self.active:
                break
            if not self.active and head is not None and head > self.now():
                time.sleep(min(0.05, max(0.0, head - self.now())))
                continue
            self.step()
            if not self.active:
                time.sleep(0.001)  # a shared queue's head went to another server, or does not fit
        return self.done

    def graph_counts(self):
        """(graphs captured, graph launches) of the server's context."""
        import ctypes as C
        c, r = C.c_longlong(), C.c_longlong()
        gpu.check(self.lib.x2g_graph_counts(self.ctx.h, C.byref(c), C.byref(r)))
        return c.value, r.value

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


# Served stream: closed-shell diatomics with 1 <= Z1, Z2 <= 10, neutral, whose
# closed-shell configuration has at most 5 orbitals (2, 4, 6, 8, 10 or 12
# electrons), R in [1, 6] bohr, on grids from 169x241 to 505x901.
STREAM_MAX_ORBITALS = 5
STREAM_ZMAX = 10
STREAM_R = (1.0, 6.0)
# admissible grid sizes (n - 1 divisible by 6) between the two ends, in
# steps that keep the nu:mu aspect of the end points
STREAM_NNI = tuple(range(169, 506, 48))      # 169, 217, ..., 505
STREAM_NMU = tuple(range(241, 902, 60))      # 241, 301, ..., 901


def _shells(z):
    """Atomic shells of a neutral atom of charge z up to n = 2 as
    (energy, kind, n, l, exponent): hydrogenic energies with Slater
    screening (the 1s exponent is z itself, as the reference's decks take
    it); kind "s" holds 2 electrons, "p" 6 (p0 sigma + p1 pi)."""
    out = [(-z * z / 2.0, "s", 1, 0, float(z))]
    if z > 2:
        zeff = z - 0.35 * (z - 3) - 0.85 * 2
        e2 = -zeff * zeff / 8.0
        out += [(e2 - 1e-6, "s", 2, 0, zeff), (e2, "p", 2, 1, zeff)]
    return out


def _p_shell(left):
    """Orbitals taken from a p shell with `left` electrons still to place:
    6 or more -> sigma (p0) and pi (p1); 4 -> pi; 2 -> sigma."""
    return ["p0", "p1"] if left >= 6 else (["p1"] if left == 4 else (["p0"] if left == 2 else []))


def closed_shell_orbitals(z1, z2):
    """Closed-shell configuration and hydrogenic LCAO start of the neutral
    diatomic (z1, z2), or None when it needs more than STREAM_MAX_ORBITALS
    orbitals or has an odd electron count.  Orbitals come highest first (the
    order of the reference's config card).  A homonuclear molecule gets g/u
    pairs of the atomic shells (sigma-g then sigma-u from an s shell;
    sigma-g from p0, pi-u from p1); a heteronuclear one takes the shells of
    both centres in energy order, each on its own centre."""
    nel = z1 + z2
    if nel % 2:
        return None
    orbs, left = [], nel
    if z1 == z2:
        for e, kind, n, l, ex in _shells(z1):
            if kind == "s":
                for gu, c2 in ((1, 1.0), (-1, -1.0)):
                    if left >= 2:
                        orbs.append(scf.Orbital(0, gu, 1.0, n, l, ex, c2, n, l, ex))
                        left -= 2
            else:
                for part in _p_shell(left):
                    if part == "p0":
                        orbs.append(scf.Orbital(0, 1, 1.0, n, l, ex, 1.0, n, l, ex))
                        left -= 2
                    else:
                        orbs.append(scf.Orbital(1, -1, 1.0, n, l, ex, 1.0, n, l, ex))
                        left -= 4
    else:
        shells = sorted([(e, 0, kind, n, l, ex) for e, kind, n, l, ex in _shells(z1)] +
                        [(e, 1, kind, n, l, ex) for e, kind, n, l, ex in _shells(z2)])
        for e, centre, kind, n, l, ex in shells:
            if left <= 0:
                break
            c1, c2 = (1.0, 0.0) if centre == 0 else (0.0, 1.0)
            za = ex if centre == 0 else 1.0
            zb = ex if centre == 1 else 1.0
            if kind == "s":
                orbs.append(scf.Orbital(0, 0, c1, n, l, za, c2, n, l, zb))
                left -= 2
            else:
                for part in _p_shell(left):
                    orbs.append(scf.Orbital(0 if part == "p0" else 1, 0, c1, n, l, za, c2, n, l, zb))
                    left -= 2 if part == "p0" else 4
    if left != 0 or not 1 <= len(orbs) <= STREAM_MAX_ORBITALS:
        return None
    return orbs[::-1]


STREAM_MOLECULES = tuple((z1, z2) for z1 in range(1, STREAM_ZMAX + 1) for z2 in range(1, z1 + 1)
                         if closed_shell_orbitals(z1, z2) is not None)


# Stopping rule of a served request: orbital-energy changes below 1e-11 for
# nscfextra + 1 iterations.  1e-11 is reachable in FP64 for every molecule of
# the stream (the orbital energies of the heavier ones carry 1e-11 of
# rounding noise, so with 1e-12 their runs end on the saturation test at a
# random point); the norm test is off (inoterm 16: a run may meet a norm
# threshold while its energy still moves by 1e-8), and the saturation test
# looks back 50 iterations instead of 20.
STREAM_STOP = dict(ienterm=11, inoterm=16, conv=(600, 50, 50), maxscf=8000)


def stream_request(rng, prec=1, rinf=40.0, **stop):
    """One request of the served stream: a molecule of STREAM_MOLECULES (Z1 >=
    Z2), R uniform in STREAM_R, a grid size class drawn uniformly from the
    sizes between 169x241 and 505x901, STREAM_STOP unless stop overrides it.
    Returns (name, spec)."""
    z1, z2 = STREAM_MOLECULES[int(rng.integers(len(STREAM_MOLECULES)))]
    r = float(rng.uniform(*STREAM_R))
    k = int(rng.integers(len(STREAM_NNI)))
    nni = STREAM_NNI[k]
    nmu = STREAM_NMU[round(k * (len(STREAM_NMU) - 1) / (len(STREAM_NNI) - 1))]
    orbs = closed_shell_orbitals(z1, z2)
    spec = scf.make_spec(float(z1), float(z2), r, nni, nmu, rinf, orbs, prec=prec, schedule=1,
                         **dict(STREAM_STOP, **stop))
    return f"Z{z1}-Z{z2} R={r:.3f} {nni}x{nmu} {len(orbs)}orb", spec


def converged(result):
    """Whether a finished request reached a converged energy: it stopped on
    the orbital-energy or norm threshold, or on the resets rule (its energy
    changes hover at the threshold: the rounding floor).  The saturation test
    certifies nothing (a slowly converging run meets it while its energy
    still drifts by 1e-7), nor does the iteration limit."""
    return result["stop"] in ("orbital energy threshold", "orbital norm threshold", "resets")


def poisson_stream(count, rate, seed=0, prec=1, rinf=40.0):
    """A seeded stream of requests: exponential inter-arrival times of mean
    1/rate seconds, each request from stream_request.  Returns
    [(arrival_s, name, spec)]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for _ in range(count):
        t += float(rng.exponential(1.0 / rate))
        name, spec = stream_request(rng, prec=prec, rinf=rinf)
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


class GroupServer:
    """`groups` servers on one GPU, each with its own device context (its own
    stream) and its own host thread, taking requests from one queue in
    arrival order: whichever group has a free slot and room in its share of
    the memory budget admits the head request.  While one group reads back
    its iteration's results and runs the host bookkeeping, and while it
    issues its many small launches, the other groups' kernels keep the GPU
    busy.  max_active and the budget are split evenly between the groups;
    results are those of a single server (no job depends on its batch)."""

    def __init__(self, groups=2, device=0, max_active=32, hbm_budget=None, hbm_fraction=0.8,
                 clock=time.perf_counter, graphs=False):
        self.clock = clock
        slots = max(1, max_active // groups)
        first = Server(device, slots, hbm_budget=hbm_budget, hbm_fraction=hbm_fraction, clock=clock, graphs=graphs)
        share = first.budget // groups
        first.budget = share
        self.servers = [first] + [Server(device, slots, hbm_budget=share, clock=clock, graphs=graphs)
                                  for _ in range(groups - 1)]
        self.budget = share
        self.t0 = clock()
        self.pending = []
        self._lock = threading.Lock()
        for srv in self.servers:
            srv.t0 = self.t0
            srv.pending = self.pending
            srv._queue_lock = self._lock
        self._next_id = 0

    def now(self):
        return self.clock() - self.t0

    def submit(self, spec, arrival=None):
        """Queue a request (as Server.submit; it must fit one group's budget)."""
        fp = footprint(spec)
        if fp > self.budget:
            raise ValueError(f"request needs {fp} bytes, more than a group's budget {self.budget}")
        r = Request(self._next_id, spec, self.now() if arrival is None else float(arrival), fp)
        self._next_id += 1
        with self._lock:
            self.pending.append(r)
            self.pending.sort(key=lambda q: (q.arrival, q.rid))
        return r

    def run(self):
        """Serve until every request is finished, one host thread per group."""
        errors = []

        def work(srv):
            try:
                srv.run()
            except BaseException as e:  # re-raised by run() after the other groups finish
                errors.append(e)
                with self._lock:
                    self.pending.clear()

        threads = [threading.Thread(target=work, args=(srv,)) for srv in self.servers]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return self.done

    @property
    def done(self):
        return [r for srv in self.servers for r in srv.done]

    @property
    def iterations(self):
        return sum(srv.iterations for srv in self.servers)

    def graph_counts(self):
        c = [srv.graph_counts() for srv in self.servers]
        return sum(a for a, _ in c), sum(b for _, b in c)

    def close(self):
        for srv in self.servers:
            srv.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
