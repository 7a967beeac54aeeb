"""The serving loop's host logic on the CPU, with stand-in cells and batches
that record what the loop does to them: no batch is built from or launched
with a closed cell, no cell is closed or read while a batch holding it is in
flight, no cell is closed while a batch that holds it is open, a lane never
advances more searches than its cap, a cell that faults fails its own
request only, every request gets exactly one answer, and every admitted cell
is closed at the end."""
import random
import time

import numpy as np
import pytest

import coldcase_amd.serve as S


class _Cell:
    live = []

    def __init__(self, q, n, R, M):
        self.q, self.n, self.R, self.M = q, n, R, M
        self.closed = False
        self.inflight = 0
        self.it = 0
        self.done = False
        self.fault = 0
        self.holders = 0
        self.launch = (1024, 64)
        _Cell.live.append(self)

    def get_launch(self):
        return self.launch

    def set_launch(self, eval_groups, commit_groups):
        self.launch = (eval_groups, commit_groups)

    def search_params(self, **k):
        pass

    def search_init(self, seed, seed_code=None, greedy=False):
        self.it = 0

    def set_code(self, code, recount=True):
        pass

    def distance_transform(self):
        h = np.zeros(self.n + 1, np.int64)
        h[0] = self.q ** self.n
        return None, h, 0

    def gather_far(self, r, cap):
        return [], 0

    def state(self):
        assert not self.closed and self.inflight == 0, 'state read while in flight'
        return {'done': self.done, 'uncovered': 5, 'it': self.it, 'kicks': 0, 'fault': self.fault}

    def check_fault(self, st=None):
        st = self.state() if st is None else st
        if st['fault']:
            raise S.CovError('fault %d' % st['fault'])
        return st

    def best(self):
        assert not self.closed and self.inflight == 0, 'best read while in flight'
        return np.zeros((self.M, self.n), np.uint8), 1

    def close(self):
        assert self.inflight == 0, 'closing a cell in flight'
        self.closed = True

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class _Batch:
    cap = None

    def __init__(self, cells):
        assert cells and all(not c.closed for c in cells), 'batch with a closed cell'
        assert len(set(map(id, cells))) == len(cells)
        if _Batch.cap is not None:
            assert len(cells) <= _Batch.cap, 'lane above its cap'
        self.cells = list(cells)
        for c in self.cells:
            c.holders += 1
        self.busy_until = None
        self.closed = False
        self.k = 0

    def launch(self, k, mode=0):
        assert not self.closed and self.busy_until is None
        for c in self.cells:
            assert not c.closed, 'launch with a closed cell'
            c.inflight += 1
        self.k = k
        self.busy_until = time.perf_counter() + random.uniform(0, 0.001)

    def query(self):
        return self.busy_until is None or time.perf_counter() >= self.busy_until

    fault_rate = 0.0

    def sync(self):
        if self.busy_until is not None:
            faulted = False
            for c in self.cells:
                c.inflight -= 1
                if c.done:
                    continue
                if random.random() < _Batch.fault_rate:
                    c.fault, c.done, faulted = 2, True, True
                    continue
                c.it += self.k
                if random.random() < 0.01:
                    c.done = True
            self.busy_until = None
            if faulted:
                raise S.CovError('cov_batch_sync: a cell faulted')

    def close(self):
        assert self.busy_until is None, 'closing a busy batch'
        if not self.closed:
            for c in self.cells:
                c.holders -= 1
        self.closed = True


SHAPES = [(5, 7, 2, 400), (6, 6, 3, 41), (6, 8, 4, 169), (6, 9, 4, 660), (7, 9, 4, 1480), (6, 10, 4, 2784)]


@pytest.mark.parametrize('seed', range(12))
def test_lane_logic(monkeypatch, seed):
    monkeypatch.setattr(S, 'Cell', _Cell)
    monkeypatch.setattr(S, 'Batch', _Batch)
    monkeypatch.setattr(S, 'mem_bytes', lambda q, n, R, M: 1000 * M)
    rnd = random.Random(seed)
    random.seed(seed)
    moves = rnd.choice([300, 1000, 3000])
    reqs = S.poisson_workload(seed, 40, 2000.0, SHAPES, verify_frac=0.3, moves=moves)
    cap = rnd.choice([None, 1, 3])
    _Batch.cap = cap
    _Batch.fault_rate = 0.02 if seed % 3 == 2 else 0.0
    _Cell.live = []
    srv = S.Server(hbm_budget=rnd.choice([40 * 1000 * 2784, 10 ** 12]), max_cells=rnd.choice([8, 64]),
                   tick_steps=rnd.choice([64, 256]), lane_cap=cap)
    res = srv.run(reqs, realtime=rnd.random() < 0.5)
    assert [r.rid for r in res] == sorted(r.rid for r in reqs)
    failed = 0
    for q, r in zip(sorted(reqs, key=lambda x: x.rid), res):
        if r.error:
            assert _Batch.fault_rate > 0 and q.kind == 'search' and 'fault' in r.error
            failed += 1
        elif q.kind == 'search':
            assert 0 < r.iters <= q.moves       # its moves, or fewer once covered
    assert failed == sum(1 for c in _Cell.live if c.fault)
    assert all(c.closed for c in _Cell.live), 'leaked cell'
