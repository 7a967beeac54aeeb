"""autotune -- launch shapes chosen by measurement, per cell shape.

What is tuned:
  wave schedule  the evaluator's and the commit kernel's workgroup counts
                 (cov_set_launch), for one cell alone: coordinate descent over
                 a ladder around the engine's default, each candidate timed
                 on hipGraph-replayed steps of a real search of that shape,
                 in a palindromic order so that the search's drift over the
                 tuning run cancels to first order
  batch scale    for a batch, one factor applied to every member's tuned
                 counts (a batch shares the GPU, so each cell wants fewer
                 workgroups than alone), timed on the batch itself
  fleet groups   cells split by step cost into a few batches that run
                 concurrently on their own streams (engine.Fleet), each
                 with its own batch scale, chosen with the groups running
                 together
  tile           the distance transform's LDS bytes per tile
                 (cov_set_transform_tile), timed on a random code of the shape

Results are kept in a JSON table keyed by "q,n,R,M" (tuned.json beside this
file by default) and applied with `apply(cell)`.

    python -m coldcase_amd.autotune q n R M [--table PATH]
"""
import argparse
import json
import os
import time

import numpy as np

from .engine import GRAPH_STEPS, MODE_GRAPH, Batch, Cell, CovError

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tuned.json')


def _key(q, n, R, M):
    return '%d,%d,%d,%d' % (q, n, R, M)


def _ladder(d, lo, hi):
    vals = {max(lo, min(hi, int(round(d * f)))) for f in (0.125, 0.25, 0.5, 1.0, 2.0)}
    return sorted(vals)


def _time_steps(cell, steps):
    cell.sync()
    t = time.perf_counter()
    cell.run(steps, mode=MODE_GRAPH)
    return (time.perf_counter() - t) / steps


def _palindrome(cell, cands, setter, steps):
    """Time every candidate twice, in order then reversed; returns {cand: s/step}."""
    acc = {c: 0.0 for c in cands}
    for c in list(cands) + list(reversed(cands)):
        setter(c)
        cell.run(GRAPH_STEPS, mode=MODE_GRAPH)  # capture the graph outside the timing
        acc[c] += _time_steps(cell, steps)
    return {c: v / 2 for c, v in acc.items()}


def tune_search(q, n, R, M, seed=1, warmup=300, steps=256):
    """Best (eval_groups, commit_groups) for one cell of this shape alone."""
    with Cell(q, n, R, M) as cell:
        cell.search_params(tenure=10, stall=5000, kick=2)
        cell.search_init(seed)
        e0, c0 = cell.get_launch()
        cell.run(warmup, mode=MODE_GRAPH)
        if cell.state()['done']:
            return {'eval_groups': e0, 'commit_groups': c0, 'us_per_step': None, 'default_us': None}
        te = _palindrome(cell, _ladder(e0, 2, 4096), lambda e: cell.set_launch(e, c0), steps)
        be = min(te, key=te.get)
        tc = _palindrome(cell, _ladder(c0, 1, 4096), lambda c: cell.set_launch(be, c), steps)
        bc = min(tc, key=tc.get)
        return {'eval_groups': be, 'commit_groups': bc, 'us_per_step': tc[bc] * 1e6,
                'default_us': te.get(e0, float('nan')) * 1e6}


def tune_transform(q, n, R, M, seed=1, reps=5):
    """Best LDS bytes per transform tile for this shape."""
    rng = np.random.default_rng(seed)
    code = rng.integers(0, q, size=(M, n)).astype(np.uint8)
    best, times = None, {}
    with Cell(q, n, R, M) as cell:
        cell.set_code(code, recount=False)
        for lds in (16384, 24576, 32768, 40960, 49152, 63488):
            for nt in (256, 512, 1024):
                try:
                    cell.set_transform_tile(lds, nt)
                except CovError:
                    continue
                cell.distance_transform()
                t = time.perf_counter()
                for _ in range(reps):
                    cell.distance_transform()
                times[(lds, nt)] = (time.perf_counter() - t) / reps
        best = min(times, key=times.get)
    return {'lds': best[0], 'threads': best[1], 'transform_us': times[best] * 1e6}


def tune_batch(cells, factors=(1.0, 0.5, 0.25, 0.125, 0.0625), steps=64):
    """Scale every member's current launch shape by the factor that gives the
    fastest batched step; leaves the cells at that shape and returns it."""
    base = [c.get_launch() for c in cells]

    def setf(f):
        for c, (e, k) in zip(cells, base):
            c.set_launch(max(2, int(e * f)), max(1, int(k * f)))

    times = {}
    for f in list(factors) + list(reversed(factors)):
        setf(f)
        with Batch(cells) as b:
            b.run(GRAPH_STEPS)
            t = time.perf_counter()
            b.run(steps)
            times[f] = times.get(f, 0.0) + (time.perf_counter() - t) / steps
    best = min(times, key=times.get)
    setf(best)
    return {'factor': best, 'us_per_step': times[best] / 2 * 1e6}


def plan_groups(costs, ngroups=3):
    """Split cells into at most `ngroups` groups of about equal total cost,
    the most expensive cells together: indices sorted by decreasing cost,
    cut where the running sum passes each multiple of total / ngroups.
    `costs` are per-cell step times (or any proxy); returns lists of indices."""
    order = sorted(range(len(costs)), key=lambda i: (-costs[i], i))
    total = float(sum(costs)) or 1.0
    groups, cur, acc = [], [], 0.0
    for i in order:
        cur.append(i)
        acc += costs[i]
        if len(groups) < ngroups - 1 and acc >= total * (len(groups) + 1) / ngroups:
            groups.append(cur)
            cur = []
    if cur:
        groups.append(cur)
    return groups


def tune_fleet(cells, costs, ngroups=4, start=0.25, lo=0.0625, hi=1.0, passes=2, steps=96):
    """Group the cells by cost (plan_groups), then choose each group's batch
    scale with the groups running together: coordinate descent over the
    groups, each trying half and double its current scale, every candidate
    timed on `steps` steps of the whole fleet.  Leaves the cells at the chosen
    shapes; returns the groups (lists of cells), their scales and the index
    lists."""
    from .engine import Fleet
    idx = plan_groups(costs, ngroups)
    groups = [[cells[i] for i in g] for g in idx]
    base = [[c.get_launch() for c in g] for g in groups]
    fs = [start] * len(groups)

    def apply(fv):
        for g, b, f in zip(groups, base, fv):
            for c, (e, k) in zip(g, b):
                c.set_launch(max(2, int(e * f)), max(1, int(k * f)))

    def timed(fv):
        apply(fv)
        with Fleet(groups) as fl:
            fl.run(GRAPH_STEPS)
            t = time.perf_counter()
            fl.run(steps)
            return (time.perf_counter() - t) / steps

    best_t = timed(fs)
    for _ in range(passes):
        for gi in range(len(groups)):
            for cand in (fs[gi] / 2, fs[gi] * 2):
                if cand < lo or cand > hi:
                    continue
                trial = list(fs)
                trial[gi] = cand
                t = timed(trial)
                if t < best_t:
                    best_t, fs = t, trial
    apply(fs)
    return groups, fs, idx


class Table:
    """The tuned launch shapes, a JSON file keyed by "q,n,R,M"."""

    def __init__(self, path=TABLE):
        self.path = path
        self.rows = {}
        if os.path.exists(path):
            with open(path) as f:
                self.rows = json.load(f)

    def get(self, q, n, R, M):
        return self.rows.get(_key(q, n, R, M))

    def put(self, q, n, R, M, row):
        self.rows.setdefault(_key(q, n, R, M), {}).update(row)

    def save(self):
        tmp = self.path + '.part'
        with open(tmp, 'w') as f:
            json.dump(self.rows, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def apply(self, cell, scale=1.0):
        """Set the cell's tuned launch shape (times `scale`), if the table has
        its shape; returns whether it did."""
        row = self.get(cell.q, cell.n, cell.R, cell.M)
        if not row:
            return False
        if 'eval_groups' in row:
            cell.set_launch(max(2, int(row['eval_groups'] * scale)), max(1, int(row['commit_groups'] * scale)))
        if 'lds' in row:
            cell.set_transform_tile(row['lds'], row.get('threads', 0))
        return True


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('q', type=int)
    ap.add_argument('n', type=int)
    ap.add_argument('R', type=int)
    ap.add_argument('M', type=int)
    ap.add_argument('--table', default=TABLE)
    a = ap.parse_args(argv)
    t = Table(a.table)
    row = tune_search(a.q, a.n, a.R, a.M)
    row.update(tune_transform(a.q, a.n, a.R, a.M))
    t.put(a.q, a.n, a.R, a.M, row)
    t.save()
    print(json.dumps({_key(a.q, a.n, a.R, a.M): row}))
    return 0


if __name__ == '__main__':
    raise SystemExit(main())
