#!/usr/bin/env python3
"""lns -- ruin-and-recreate covering-code search over exact gain and loss maps.

The engine of cov/gemm2/gpuchain.py on libcovhip.so: the same operations,
command line and RESULT line, with every count exact and computed on the
GPU.

  state      a set of codewords (word indices, MSB-first) and their coverage
             counts on the device; the count is kept up to date by ball
             updates and re-derived from scratch by a recount
  gain map   one count transform of the uncovered set gives, for every word
             x, how many uncovered words a codeword at x would cover
  loss       one count transform of the singly covered set gives, at every
             codeword, how many words only it covers
  recreate   lazy greedy: the best stale gains from the map are re-evaluated
             exactly (ball gathers), and a word is placed only when its exact
             gain is at least every other bound, so each placement is a true
             argmax of the gain (coverage gain is submodular, so stale
             gains only overestimate); the map is refreshed when the bounds
             run out or fall below the extraction threshold
  ruin       remove k codewords: lowest exact loss (with jitter), around a
             high-gain word ("hole"), around a random codeword ("cluster"),
             or at random (gpuchain.py:276-314)
  peel       drop codewords that cover no word alone, several at once when
             they lie more than 2R apart (gpuchain.py:317-358)
  lns        rounds of ruin + recreate at fixed M; a round that ends worse
             than it started is undone (back to the best code) except with
             probability 0.1 (gpuchain.py:361-409)

The extraction order here is ascending word index (the reference's is the
order of its atomics), and the random choices come from numpy's generator
seeded by -s, so a run is reproducible on this engine.

  python -m coldcase_amd.lns -q Q -n N -R R -M M [-t SECONDS] [-s SEED]
         --out FILE [--in SEEDFILE] [--kmin K] [--kmax K] [--quiet]
"""
import argparse
import heapq
import os
import sys
import time

import numpy as np

from .covsearch import read_code, write_code
from .engine import CT_PRIVATE, CT_UNCOVERED, Cell


class Engine:
    """One (q, n, R) cell searched by ruin and recreate on the GPU."""

    def __init__(self, q, n, R, seed=1, out=None, log=print, deadline=None):
        self.q, self.n, self.R = q, n, R
        self.pw = np.array([q ** (n - 1 - p) for p in range(n)], dtype=np.int64)
        self.rng = np.random.default_rng(seed)
        self.out, self.log, self.deadline = out, log, deadline
        self.cell = Cell(q, n, R, 1)          # the code lives here as a word list, not as cell rows
        self.code = np.zeros(0, np.int64)
        self.uncov = q ** n
        self.best_code, self.best_uncov = None, None
        self._pub = 0.0
        self.placed = []                      # (word, exact gain) of every placement, in order
        self.stats = dict(transforms=0, placements=0, removals=0, rounds=0, recounts=0)

    # ------------------------------------------------------------ plumbing
    def close(self):
        if self.cell is not None:
            self.cell.close()
            self.cell = None

    def time_left(self):
        return np.inf if self.deadline is None else self.deadline - time.time()

    def words(self, idx):
        """Word indices -> (k, n) digits."""
        idx = np.asarray(idx, dtype=np.int64).ravel()
        return ((idx[:, None] // self.pw[None, :]) % self.q).astype(np.uint8)

    def index(self, digits):
        return np.asarray(digits, dtype=np.int64).reshape(-1, self.n) @ self.pw

    # the names the reference's callers use (cov/gemm2/siege.py)
    word_index = index
    index_word = words

    def load(self, code):
        """Word indices (1-D) or digits (M x n); duplicates are dropped."""
        a = np.asarray(code)
        idx = self.index(a) if a.ndim == 2 else a.astype(np.int64)
        self.code = np.unique(idx)
        self.recount()
        self._keep_best()

    def recount(self):
        """Counts from scratch for the current code; returns the uncovered count."""
        self.cell.load_words(self.code)
        self.stats['recounts'] += 1
        self.uncov = int(self.cell.uncovered())
        return self.uncov

    verify_solved_on_device = recount

    def _keep_best(self):
        better = self.best_uncov is None or self.uncov < self.best_uncov or (
            self.uncov == self.best_uncov and len(self.code) < len(self.best_code))
        if better:
            self.best_uncov, self.best_code = self.uncov, self.code.copy()
            self.publish()

    def publish(self, force=False):
        """The best code to --out, atomically, at most once a second."""
        if self.out is None or self.best_code is None or not len(self.best_code):
            return
        if not force and time.time() - self._pub < 1.0:
            return
        self._pub = time.time()
        write_code(self.out, self.words(self.best_code), self.q)

    def _place(self, w, gain):
        self.cell.ball_update([w], +1)
        self.code = np.append(self.code, np.int64(w))
        self.uncov -= gain
        self.placed.append((int(w), int(gain)))
        self.stats['placements'] += 1

    def _remove(self, sel):
        """Remove the codewords at positions `sel` of the code."""
        sel = np.asarray(sel, dtype=np.int64)
        if not sel.size:
            return
        self.cell.ball_update(self.code[sel], -1)
        keep = np.ones(len(self.code), bool)
        keep[sel] = False
        self.code = self.code[keep]
        self.stats['removals'] += int(sel.size)

    # ------------------------------------------------------------ recreate
    def _bounds(self, cap):
        """A fresh gain map; (heap of (-gain, word) over words not in the code
        whose gain is at least the threshold, the threshold), the threshold
        chosen from the map's histogram so that about 0.9 cap words pass."""
        c = self.cell
        c.count_transform(CT_UNCOVERED)
        self.stats['transforms'] += 1
        top = c.map_max()
        if top <= 0:
            return [], 0
        thr = 1
        if top > 1:
            nb = 1024
            h = c.map_hist(nb, top)
            above = np.cumsum(h[::-1])[::-1]          # words in bins >= b
            ok = np.flatnonzero(above <= 0.9 * cap)
            b = int(ok[0]) if ok.size else nb - 1
            thr = max(1, b * top // nb + 1)
        words, vals, tot = c.map_extract(thr, cap)
        while tot > cap and thr < top:
            thr = min(top, thr + max(1, (top - thr) // 2))
            words, vals, tot = c.map_extract(thr, cap)
        inside = np.isin(words, self.code)
        heap = [(-int(v), int(w)) for w, v in zip(words[~inside], vals[~inside])]
        heapq.heapify(heap)
        return heap, thr

    def greedy_fill(self, target_m, batch=64, refresh_every=512, cap=1 << 18):
        """Place codewords of maximum exact gain until the code covers, has
        target_m words, or the time runs out; returns the uncovered count."""
        heap, thr, since = [], 1, refresh_every
        inside = set(self.code.tolist())
        while self.uncov > 0 and len(self.code) < target_m and self.time_left() > 5:
            if since >= refresh_every or not heap:
                heap, thr = self._bounds(cap)
                since = 0
                if not heap:
                    if thr > 1:                        # everything passing was placed: go lower
                        since = refresh_every
                        cap *= 2
                        continue
                    break
            popped = []
            while heap and len(popped) < batch:
                g, w = heapq.heappop(heap)
                if w not in inside:
                    popped.append(w)
            if not popped:
                since = refresh_every
                continue
            exact = self.cell.ball_gather(popped, 0)
            j = int(np.argmax(exact))
            best = int(exact[j])
            bound = -heap[0][0] if heap else 0
            if best >= 1 and best >= bound and best >= thr:
                self._place(popped[j], best)
                inside.add(popped[j])
                since += 1
                rest = [(-int(g), w) for k, (g, w) in enumerate(zip(exact, popped)) if k != j and g > 0]
            else:
                rest = [(-int(g), w) for g, w in zip(exact, popped) if g > 0]
                if best < max(thr, 1):
                    since = refresh_every                # the bounds below thr are unknown: refresh
            for item in rest:
                heapq.heappush(heap, item)
        return self.uncov

    # ---------------------------------------------------------------- ruin
    def losses(self):
        """Exact loss (words covered by it alone) of every codeword."""
        self.cell.count_transform(CT_PRIVATE)
        self.stats['transforms'] += 1
        return self.cell.map_read(self.code)

    def _nearest(self, centre, k):
        """Positions of the k codewords nearest to `centre` (digits), ties at random."""
        d = (self.words(self.code) != centre[None, :]).sum(axis=1)
        return np.argsort(d + self.rng.random(len(d)), kind='stable')[:k]

    def ruin(self, k, mode='low'):
        """Remove k codewords (at least one stays); returns their words."""
        m = len(self.code)
        k = min(int(k), m - 1)
        if k <= 0:
            return np.zeros(0, np.int64)
        if mode == 'random':
            sel = self.rng.choice(m, size=k, replace=False)
        elif mode == 'cluster':
            sel = self._nearest(self.words(self.code[self.rng.integers(m)])[0], k)
        elif mode == 'hole':
            self.cell.count_transform(CT_UNCOVERED)
            self.stats['transforms'] += 1
            top = self.cell.map_max()
            if top <= 0:
                return self.ruin(k, 'low')
            words, _, _ = self.cell.map_extract(max(1, top // 2), 4096)
            if not len(words):
                return self.ruin(k, 'low')
            sel = self._nearest(self.words(words[self.rng.integers(len(words))])[0], k)
        else:
            lo = self.losses().astype(np.float64)
            jitter = self.rng.exponential(max(1.0, 0.15 * lo.mean()), size=m)
            sel = np.argsort(lo + jitter, kind='stable')[:k]
        gone = self.code[sel].copy()
        self._remove(sel)
        self.uncov = int(self.cell.uncovered())
        return gone

    def peel(self):
        """Drop codewords that cover nothing alone until none is left;
        returns how many were dropped."""
        total = 0
        while self.time_left() > 5:
            lo = self.losses()
            zero = np.flatnonzero(lo == 0)
            if zero.size:
                zero = zero[self.cell.ball_gather(self.code[zero], 1) == 0]
            if not zero.size:
                break
            self.rng.shuffle(zero)
            if 2 * self.R < self.n:
                # words more than 2R apart share no ball word: removing one
                # cannot make another's coverage private
                dig = self.words(self.code[zero]).astype(np.int16)
                take = [0]
                for i in range(1, len(zero)):
                    if ((dig[take] != dig[i][None, :]).sum(axis=1) > 2 * self.R).all():
                        take.append(i)
                zero = zero[take]
            else:
                zero = zero[:1]
            self._remove(zero)
            total += int(zero.size)
        if total:
            self.uncov = int(self.cell.uncovered())
        return total

    # ----------------------------------------------------------------- LNS
    def lns(self, M, round_budget_s, kmin=8, kmax=None, log_every=20, max_rounds=None):
        """Rounds of ruin + recreate at M codewords until the code covers, the
        budget is spent or max_rounds have run; returns the uncovered count."""
        t0 = time.time()
        kmax = kmax or max(kmin + 1, M // 20)
        self.best_code, self.best_uncov = None, None
        if len(self.code) < M:
            self.greedy_fill(M)
        self._keep_best()
        modes = ('low', 'hole', 'hole', 'cluster', 'random')
        r = 0
        while self.uncov > 0 and (max_rounds is None or r < max_rounds):
            if time.time() - t0 > round_budget_s or self.time_left() < 10:
                break
            start = self.uncov
            k = int(self.rng.integers(kmin, kmax + 1))
            self.ruin(k, modes[int(self.rng.integers(len(modes)))])
            self.greedy_fill(M)
            self.stats['rounds'] += 1
            r += 1
            if self.uncov > start and self.rng.random() > 0.1:
                self.code = self.best_code.copy()        # undo: back to the best code
                self.recount()
            else:
                self._keep_best()
            if r % log_every == 0:
                self.log('  lns M=%d round=%d uncovered=%d best=%d k=%d (%.0fs)'
                         % (M, r, self.uncov, self.best_uncov, k, time.time() - t0))
            if r % 50 == 0:
                kept = self.uncov
                if self.recount() != kept:
                    self.log('  warning: incremental count %d, recount %d; keeping the recount'
                             % (kept, self.uncov))
        if self.uncov and self.best_code is not None and self.best_uncov < self.uncov:
            self.code = self.best_code.copy()
            self.recount()
        self._keep_best()
        return self.uncov


def main(argv=None):
    ap = argparse.ArgumentParser(prog='lns', description=__doc__.split('\n')[0])
    ap.add_argument('-q', type=int, required=True)
    ap.add_argument('-n', type=int, required=True)
    ap.add_argument('-R', type=int, required=True)
    ap.add_argument('-M', type=int, required=True)
    ap.add_argument('-t', type=float, default=60.0)
    ap.add_argument('-s', type=int, default=1)
    ap.add_argument('--out', required=True)
    ap.add_argument('--in', dest='seed_code')
    ap.add_argument('--kmin', type=int, default=8)
    ap.add_argument('--kmax', type=int)
    ap.add_argument('--rounds', type=int, default=None, help='stop after this many rounds')
    ap.add_argument('--quiet', action='store_true')
    a = ap.parse_args(argv)
    t_start = time.time()
    log = (lambda *x: None) if a.quiet else print
    e = Engine(a.q, a.n, a.R, seed=a.s, out=a.out, log=log, deadline=t_start + a.t)
    try:
        if a.seed_code:
            e.load(read_code(a.seed_code, a.n))
            if len(e.code) > a.M:
                e.ruin(len(e.code) - a.M, 'low')
        else:
            e.load(np.zeros(0, np.int64))
        unc = e.lns(a.M, round_budget_s=t_start + a.t - time.time() - 5, kmin=a.kmin, kmax=a.kmax,
                    max_rounds=a.rounds)
        if unc == 0:
            unc = e.recount()                   # a cover is confirmed by a recount from scratch
        e.publish(force=True)
        dt = time.time() - t_start
        moves = e.stats['placements'] + e.stats['removals']
        print('RESULT q=%d n=%d R=%d M=%d uncovered=%d iters=%d rate=%.1f transforms=%d time=%.1f'
              % (a.q, a.n, a.R, len(e.code), unc, e.stats['rounds'], moves / max(dt, 1e-9),
                 e.stats['transforms'], dt))
        sys.stdout.flush()
    finally:
        e.close()
    return 0


if __name__ == '__main__':
    sys.exit(main())
