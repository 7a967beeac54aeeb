"""serve -- a continuous-batching request loop on one GPU.

Two request kinds, over cells of any shape (up to the engine's limits):

  verify   a code is checked exhaustively: the uint8 distance transform gives
           the uncovered count, the covering radius and the distance
           histogram; the first uncovered words in ascending order are the
           witnesses (the answer of cov/verify_cov.py:109-230)
  search   a burst of tabu moves from a seeded start (random, greedy or a
           given seed code), the reference search of
           cov/engine/covsearch2e.c:1304-1580; the answer is the best code
           found, its uncovered count and the run's counters

Requests are admitted as they arrive while the HBM budget (the sum of
cov_mem_bytes of resident cells) and the cell slots allow; a request that
does not fit waits, one larger than the whole budget is refused.  Admitted
searches stay resident in one of four lanes by the size of their balls; the
searches of a lane advance together: every tick is one batched run
(coldcase_amd.engine.Batch, one evaluator, selection and commit launch per step for
all of them) of as many steps as the closest-to-done search still needs, up
to `tick_steps`; searches that finish leave and newly admitted ones join at
the lane's next tick.  The lanes tick independently on their own streams, and
verifies run on theirs while ticks are on the device.  A request's result does not depend on what else is resident: each
cell keeps its own counters, rng stream and tabu table.
"""
import collections
import hashlib
import math
import time
from dataclasses import dataclass, field

import numpy as np

from .engine import MODE_GRAPH, Batch, Cell, CovError, mem_bytes


@dataclass(eq=False)
class Request:
    rid: int
    kind: str                      # 'verify' or 'search'
    q: int
    n: int
    R: int
    M: int
    arrival: float = 0.0           # seconds after the start of the run
    code: object = None            # verify: the code; search: an optional seed code
    seed: int = 1
    moves: int = 100000
    greedy: bool = False
    tenure: int = 10
    stall: int = 5000
    kick: int = 2


@dataclass(eq=False)
class Result:
    rid: int
    kind: str
    arrival: float
    admitted: float = 0.0
    finished: float = 0.0
    error: str = ''
    # verify
    uncovered: int = -1
    witnesses: list = field(default_factory=list)
    radius: int = -1
    hist: list = field(default_factory=list)
    # search
    best_uncovered: int = -1
    iters: int = 0
    kicks: int = 0
    best_sha256: str = ''
    best_code: object = None       # search: the best code found (M x n digits)

    @property
    def latency(self):
        return self.finished - self.arrival

    def answer(self):
        """What the request computed, without timings."""
        if self.kind == 'verify':
            return (self.error, self.uncovered, tuple(map(tuple, self.witnesses)), self.radius, tuple(self.hist))
        return (self.error, self.best_uncovered, self.uncovered, self.iters, self.kicks, self.best_sha256)


def request_bytes(r):
    return mem_bytes(r.q, r.n, r.R, r.M)


def _digits(w, q, n):
    out = []
    for _ in range(n):
        out.append(w % q)
        w //= q
    return tuple(reversed(out))


def run_verify(r, witness_cap=5):
    """One verify request on a fresh cell (code upload, transform, witnesses)."""
    res = Result(r.rid, r.kind, r.arrival)
    code = np.ascontiguousarray(r.code, dtype=np.uint8).reshape(-1, r.n)
    with Cell(r.q, r.n, r.R, code.shape[0]) as c:
        c.set_code(code, recount=False)
        _, hist, rad = c.distance_transform()
        res.hist = [int(x) for x in hist]
        res.radius = int(rad)
        res.uncovered = int(r.q ** r.n - hist[:r.R + 1].sum())
        if res.uncovered and witness_cap:
            far, _ = c.gather_far(r.R, witness_cap)
            res.witnesses = [_digits(int(x), r.q, r.n) for x in far]
    return res


def start_search(r):
    c = Cell(r.q, r.n, r.R, r.M)
    c.search_params(tenure=r.tenure, stall=r.stall, kick=r.kick)
    seed_code = None if r.code is None else np.ascontiguousarray(r.code, dtype=np.uint8).reshape(-1, r.n)
    c.search_init(r.seed, seed_code=seed_code, greedy=r.greedy)
    return c


def finish_search(r, c):
    res = Result(r.rid, r.kind, r.arrival)
    st = c.state()
    best, bu = c.best()
    res.best_uncovered = int(bu)
    res.uncovered = int(st['uncovered'])
    res.iters = int(st['it'])
    res.kicks = int(st['kicks'])
    res.best_sha256 = hashlib.sha256(best.tobytes()).hexdigest()
    res.best_code = best
    return res


def cost_class(r):
    """The lane of a search request: 0..3 by its ball size, which sets the
    cost of its steps (a few hundred words to a few hundred thousand)."""
    v = ball_volume(r.q, r.n, r.R)
    return 0 if v < 1000 else (1 if v < 30000 else (2 if v < 120000 else 3))


NLANES = 4

# A cell's default evaluator grid (up to 1024 workgroups) is sized for a cell
# alone on the device, where its units have every CU to themselves.  In a
# lane, many cells' evaluators share the device and each of its workgroups
# pays the control block's copy and its slice of the kept-count changes, so
# resident searches take fewer: at most 256 (served throughput 1.86 -> 2.03
# req/s on the serving trace; 128 gives the same, 512 less).
LANE_EVAL_GROUPS = 256


def lane_launch(c):
    """The launch shape of a resident search: its evaluator capped at
    LANE_EVAL_GROUPS workgroups, its commit as alone."""
    e0, c0 = c.get_launch()
    if e0 > LANE_EVAL_GROUPS:
        c.set_launch(LANE_EVAL_GROUPS, c0)


class _Lane:
    """The resident searches of one cost class: they advance together, one
    batched tick at a time, on the lane's own stream."""

    def __init__(self):
        self.active = []                # [request, cell, steps left, admitted time]
        self.joining = []               # admitted while a tick runs: join at the next
        self.batch = None
        self.members = ()
        self.busy = False
        self.k = 0


class Server:
    """Continuous batching with HBM-budget admission (see the module text).

    hbm_budget  bytes of device memory the resident cells may use
    max_cells   resident searches at once
    tick_steps  steps per batched tick at most
    launch      optional callable(Cell) that sets the cell's launch shape
                (the autotuner's choice for batched runs); by default a
                resident search's evaluator takes at most LANE_EVAL_GROUPS
                workgroups (lane_launch)
    lane_cap    searches a lane advances at once (None: all it holds); the
                others wait, admitted, in arrival order until a member leaves

    Resident searches are split by cost class into lanes (cost_class); each
    lane is a batch on its own stream and ticks at its own pace, so a short
    search does not wait for the ticks of a long one.  A lane cap trades
    nothing in results (each search's trajectory is its own) for earlier
    completions of the first arrivals once a lane's batch fills the device.
    """

    def __init__(self, hbm_budget=64 << 30, max_cells=64, tick_steps=256, mode=MODE_GRAPH, launch=None,
                 lane_cap=None):
        self.budget = int(hbm_budget)
        self.max_cells = int(max_cells)
        self.tick_steps = int(tick_steps)
        self.mode = mode
        self.launch = launch if launch is not None else lane_launch
        self.lane_cap = None if lane_cap is None else max(1, int(lane_cap))
        self.peak_bytes = 0
        self.peak_cells = 0
        self.ticks = 0

    def run(self, requests, realtime=True):
        """Serve `requests` (arrival times in seconds from now); returns the
        results in request order."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        self._results = {}
        self._waiting = collections.deque()
        self._joining = []              # admitted searches not yet in a lane
        self._used = 0
        self._next = 0
        lanes = [_Lane() for _ in range(NLANES)]
        self._t0 = time.perf_counter()

        def resident():
            return sum(len(ln.active) + len(ln.joining) for ln in lanes) + len(self._joining)

        try:
            while self._next < len(reqs) or self._waiting or self._joining or \
                    any(ln.active or ln.joining or ln.busy for ln in lanes):
                did = self._admit(reqs, realtime, resident())
                pending = {}
                for a in self._joining:
                    pending.setdefault(cost_class(a[0]), []).append(a)
                self._joining = []
                for ci, ln in enumerate(lanes):
                    if ln.busy and ln.batch.query():
                        try:
                            ln.batch.sync()
                        except CovError:
                            self._fail_faulted(ln)
                        ln.busy = False
                        self._retire(ln)
                        did = True
                    if ci in pending:
                        ln.joining.extend(pending.pop(ci))
                    if not ln.busy and ln.joining:
                        room = len(ln.joining) if self.lane_cap is None else self.lane_cap - len(ln.active)
                        if room > 0:
                            ln.active.extend(ln.joining[:room])
                            ln.joining = ln.joining[room:]
                    if not ln.busy and ln.active:
                        cells = tuple(a[1] for a in ln.active)
                        if cells != ln.members:
                            self._drop_batch(ln)
                            ln.batch, ln.members = Batch(list(cells)), cells
                        ln.k = min(self.tick_steps, min(a[2] for a in ln.active))
                        ln.batch.launch(ln.k, mode=self.mode)
                        ln.busy = True
                        self.ticks += 1
                        did = True
                self.peak_cells = max(self.peak_cells, resident())
                if not did:
                    if not any(ln.busy for ln in lanes) and self._next < len(reqs) and realtime:
                        time.sleep(max(0.0, min(reqs[self._next].arrival - self._now(), 0.01)))
                    else:
                        time.sleep(0.0005)
        finally:
            for ln in lanes:
                if ln.busy:
                    try:
                        ln.batch.sync()
                    except CovError:
                        pass
                self._drop_batch(ln)
                for a in ln.active + ln.joining:
                    a[1].close()
            for a in self._joining:
                a[1].close()
        return [self._results[r.rid] for r in sorted(requests, key=lambda r: r.rid)]

    def _fail_faulted(self, ln):
        """A tick ended with a device-detected error in some of the lane's
        cells: those requests fail alone; the others' cells are unaffected
        (each has its own control block) and stay in the lane."""
        keep = []
        gone = []
        for a in ln.active:
            st = a[1].state()
            if st['fault']:
                try:
                    a[1].check_fault(st)
                except CovError as e:
                    err = str(e)
                self._results[a[0].rid] = Result(a[0].rid, a[0].kind, a[0].arrival, admitted=a[3],
                                                 finished=self._now(), error=err)
                self._used -= request_bytes(a[0])
                gone.append(a[1])
            else:
                keep.append(a)
        self._drop_batch(ln)
        for c in gone:
            c.close()
        ln.active = keep

    @staticmethod
    def _drop_batch(ln):
        # a batch holds its member cells: it goes before any of them is closed
        if ln.batch is not None:
            ln.batch.close()
        ln.batch, ln.members = None, ()

    def _retire(self, ln):
        """After a tick: searches that used their moves or covered leave."""
        left, gone = [], []
        for a in ln.active:
            a[2] -= ln.k
            if a[2] <= 0 or a[1].state()['done']:
                res = finish_search(a[0], a[1])
                res.admitted, res.finished = a[3], self._now()
                self._results[a[0].rid] = res
                self._used -= request_bytes(a[0])
                gone.append(a[1])
            else:
                left.append(a)
        if gone:
            self._drop_batch(ln)
            for c in gone:
                c.close()
        ln.active = left

    def _now(self):
        return time.perf_counter() - self._t0

    def _admit(self, reqs, realtime, resident):
        """Move arrivals into the queue, admit what fits (first fit in arrival
        order), run admitted verifies, start admitted searches; returns
        whether it did anything."""
        did = False
        t = self._now()
        if not realtime and not self._waiting and resident == 0 and self._next < len(reqs):
            t = max(t, reqs[self._next].arrival)
        while self._next < len(reqs) and reqs[self._next].arrival <= t:
            self._waiting.append(reqs[self._next])
            self._next += 1
            did = True
        for r in list(self._waiting):
            try:
                need, err = request_bytes(r), ''
            except CovError as e:
                need, err = None, str(e)
            if need is None or need > self.budget:
                self._waiting.remove(r)
                self._results[r.rid] = Result(r.rid, r.kind, r.arrival, admitted=self._now(),
                                              finished=self._now(), error=err or 'larger than the HBM budget')
                did = True
                continue
            if self._used + need > self.budget:
                continue
            if r.kind == 'search' and resident >= self.max_cells:
                continue
            self._waiting.remove(r)
            self._used += need
            self.peak_bytes = max(self.peak_bytes, self._used)
            did = True
            ta = self._now()
            if r.kind == 'verify':
                try:
                    res = run_verify(r)
                except CovError as e:
                    res = Result(r.rid, r.kind, r.arrival, error=str(e))
                res.admitted, res.finished = ta, self._now()
                self._results[r.rid] = res
                self._used -= need
            else:
                try:
                    c = start_search(r)
                    if self.launch:
                        self.launch(c)
                    self._joining.append([r, c, r.moves, ta])
                    resident += 1
                except CovError as e:
                    self._used -= need
                    self._results[r.rid] = Result(r.rid, r.kind, r.arrival, admitted=ta, finished=self._now(),
                                                  error=str(e))
        return did


def summarize(results):
    """p50 / p99 latency and throughput of a served run."""
    ok = [r for r in results if not r.error]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    span = (max(r.finished for r in ok) - min(r.arrival for r in ok)) if ok else 0.0
    out = {'requests': len(results), 'completed': len(ok), 'errors': len(results) - len(ok),
           'p50_s': float(np.percentile(lat, 50)), 'p99_s': float(np.percentile(lat, 99)),
           'max_s': float(lat.max()), 'throughput_rps': len(ok) / span if span > 0 else 0.0}
    for kind in ('verify', 'search'):
        lk = np.array([r.latency for r in ok if r.kind == kind])
        if lk.size:
            out[kind] = {'n': int(lk.size), 'p50_s': float(np.percentile(lk, 50)),
                         'p99_s': float(np.percentile(lk, 99))}
    return out


def poisson_workload(seed, n_req, rate, shapes, verify_frac=0.3, moves=100000, codes=None):
    """A seeded request trace: exponential inter-arrival times of mean
    1/rate, each request a verify (probability verify_frac) or a search
    burst of `moves` steps, on a shape drawn uniformly from `shapes`
    ((q, n, R, M) tuples).  Verify codes come from `codes[(q, n, R, M)]` when
    given (for instance the record codes), else uniform random digits;
    search seeds follow seed * 1000003 + rid * 7919."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n_req):
        t += float(rng.exponential(1.0 / rate))
        q, n, R, M = shapes[int(rng.integers(len(shapes)))]
        if rng.random() < verify_frac:
            code = (codes or {}).get((q, n, R, M))
            if code is None:
                code = rng.integers(0, q, size=(M, n)).astype(np.uint8)
            out.append(Request(rid, 'verify', q, n, R, M, arrival=t, code=code))
        else:
            out.append(Request(rid, 'search', q, n, R, M, arrival=t, seed=seed * 1000003 + rid * 7919,
                               moves=moves))
    return out


def ball_volume(q, n, R):
    return sum(math.comb(n, i) * (q - 1) ** i for i in range(R + 1))
