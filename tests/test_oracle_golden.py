"""The CPU oracle reproduces the reference solver's trajectories.

Every run in tests/golden/trajectories.json was produced by the reference
`covsearch2e` compiled from source (tests/golden/make_golden.py).  The oracle
must give the same best-uncovered count, iteration count, kick count and the
byte-identical published code file.
"""
import hashlib
import os

import pytest

from oracle.oracle import Oracle, format_code
from tests._util import DATA, golden_runs, prof_aggregates, prof_matches, read_code

RUNS = golden_runs()


def _id(r):
    return '%s-K%d(%d,%d)M%d-s%d-it%d' % (r['tag'], r['q'], r['n'], r['R'], r['M'], r['seed'], r['iters'])


def replay(r, threads=None):
    if threads is None:
        threads = 1 if r["q"] ** r["n"] < 3_000_000 else 4
    o = Oracle(r['q'], r['n'], r['R'], r['M'], r['seed'], threads=threads)
    f = r['flags']
    if '--wide' in f:
        o.set_wide()
    greedy = '--random-init' not in f
    tabu = int(f[f.index('--tabu') + 1]) if '--tabu' in f else 0
    stall = int(f[f.index('--stall') + 1]) if '--stall' in f else 5000
    cand = int(f[f.index('--cand') + 1]) if '--cand' in f else 100000
    seed_code = read_code(os.path.join(DATA, r['infile']), r['n']) if r['infile'] else None
    o.init(seed_code, greedy=greedy)
    if '--target' in f:
        o.set_target(int(f[f.index('--target') + 1]))
    o.run(r['iters'], tabu=tabu, stall=stall, kick=2, cand=cand)
    return o


@pytest.mark.parametrize('r', [r for r in RUNS if r['q'] ** r['n'] <= 3_000_000], ids=_id)
def test_oracle_matches_reference(r):
    o = replay(r)
    assert (o.best_uncovered, o.iters, o.kicks) == (r['uncovered'], r['iters_done'], r['kicks'])
    if 'ttt_it' in r:
        assert o.ttt_it == r['ttt_it']
    text = format_code(o.distinct_fill(o.best()), r['q'])
    assert hashlib.sha256(text.encode()).hexdigest() == r['out_sha256']


@pytest.mark.slow
@pytest.mark.parametrize('r', [r for r in RUNS if r['q'] ** r['n'] > 3_000_000
                               and r['q'] ** r['n'] * r['iters'] <= 5 * 10 ** 9], ids=_id)
def test_oracle_matches_reference_large(r):
    o = replay(r, threads=8)
    assert (o.best_uncovered, o.iters, o.kicks) == (r['uncovered'], r['iters_done'], r['kicks'])
    text = format_code(o.distinct_fill(o.best()), r['q'])
    assert hashlib.sha256(text.encode()).hexdigest() == r['out_sha256']


@pytest.mark.parametrize('r', [r for r in RUNS if 'prof' in r and r['q'] ** r['n'] <= 3_000_000], ids=_id)
def test_oracle_step_sums_match_reference_profile(r):
    """The per-step sums of the -DPROF build (mean candidates, sample, uncovered
    before each step, selected delta, distinct codewords) over the whole run."""
    import numpy as np
    o = Oracle(r['q'], r['n'], r['R'], r['M'], r['seed'])
    f = r['flags']
    o.init(None, greedy='--random-init' not in f)
    unc0 = o.uncovered
    tr = o.run(r['iters'], tabu=int(f[f.index('--tabu') + 1]),
               stall=int(f[f.index('--stall') + 1]) if '--stall' in f else 5000, trace=True)
    assert len(tr) == r['iters_done']
    assert prof_matches(prof_aggregates(tr, unc0), r['prof']) == {}
    # a planted mistake is caught: the uncovered count read after the step
    # instead of before it
    bad = prof_aggregates(np.vstack([tr[1:], tr[-1:]]), tr[0, 10])
    assert prof_matches(bad, r['prof']) != {}
