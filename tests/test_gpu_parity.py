"""GPU search vs the oracle, step by step, through the C ABI.

Every step's record (picked word, candidate count, chosen move, evaluated and
committed deltas, uncovered/best counts, kicks, rng draws consumed) must be
identical, and so must the final counters and best code.
"""
import numpy as np
import pytest

from oracle.oracle import TR_FIELDS, Oracle
from tests._util import golden_runs

pytestmark = pytest.mark.gpu


def _cell(q, n, R, M):
    from coldcase_amd import Cell
    return Cell(q, n, R, M)


def compare(q, n, R, M, seed, steps, tabu=10, stall=5000, cand=100000, greedy=False,
            seed_code=None, mode=0, check_cnt=True, threads=1, leave=0, wide=False):
    print('case K%d(%d,%d) M=%d seed=%d steps=%d stall=%d mode=%d leave=%d wide=%d'
          % (q, n, R, M, seed, steps, stall, mode, leave, wide), flush=True)
    cell = _cell(q, n, R, M)
    cell.search_params(tenure=tabu, stall=stall, kick=2, cand=cand)
    cell.set_leave_mode(leave)
    if wide:
        cell.search_wide(True)
    cell.search_init(seed, seed_code=seed_code, greedy=greedy)
    o = Oracle(q, n, R, M, seed, threads=threads)
    if wide:
        o.set_wide()
    o.init(seed_code, greedy=greedy)
    st0 = cell.state()
    assert st0['uncovered'] == o.uncovered, (st0, o.uncovered)
    assert st0['draws'] == o.draws
    assert (cell.code() == o.code()).all()
    gpu = cell.run(steps, mode=mode, trace=True)
    ref = o.run(steps, tabu=tabu, stall=stall, cand=cand, trace=True)
    assert gpu.shape == ref.shape
    diff = np.argwhere(gpu != ref)
    if diff.size:
        s, f = diff[0]
        raise AssertionError('step %d field %s: gpu %d oracle %d' % (s, TR_FIELDS[f], gpu[s, f], ref[s, f]))
    st = cell.state()
    assert (st['it'], st['best_uncovered'], st['kicks'], st['uncovered']) == \
        (o.iters, o.best_uncovered, o.kicks, o.uncovered)
    best, bu = cell.best()
    assert (best == o.best()).all()
    assert (cell.code() == o.code()).all()
    if check_cnt:
        assert (cell.cnt() == o.cnt()).all()
        assert cell.uncovered() == o.uncovered
    cell.close()
    return gpu


# the leaving side walked every step, or kept per codeword through the commits
# (the automatic choice keeps it on the K6(9,4)-sized cells below)
LEAVES = [1, 2]


@pytest.mark.parametrize('leave', LEAVES)
def test_k5_7_2_random_tabu(leave):
    compare(5, 7, 2, 400, 1, 2000, leave=leave)


@pytest.mark.parametrize('leave', LEAVES)
def test_k6_6_3_kicks(leave):
    tr = compare(6, 6, 3, 41, 2, 600, tabu=3, stall=20, leave=leave)
    assert tr[:, TR_FIELDS.index('kick')].sum() > 5


@pytest.mark.parametrize('leave', LEAVES)
def test_small_cells_greedy(leave):
    for (q, n, R, M, it) in [(3, 4, 1, 9, 200), (2, 7, 2, 8, 300), (7, 3, 1, 20, 300), (5, 4, 3, 3, 50)]:
        compare(q, n, R, M, 7, it, tabu=0, greedy=True, leave=leave)


@pytest.mark.parametrize('leave', LEAVES)
def test_cand_sample(leave):
    compare(6, 6, 3, 41, 5, 300, cand=3, leave=leave)


@pytest.mark.parametrize('leave', LEAVES)
def test_graph_mode_identical(leave):
    compare(5, 7, 2, 400, 3, 1000, mode=1, leave=leave)


@pytest.mark.parametrize('leave', [0, 1])
def test_k6_9_4_random_tabu(leave):
    compare(6, 9, 4, 660, 1, 200, threads=8, leave=leave)


def test_k6_8_4_kicks_kept():
    """Kicks on a cell where the counts are kept by default (average
    coverage above 4): the kicked codewords' balls move every count and XOR
    they touch, and their own rows are recounted."""
    tr = compare(6, 8, 4, 169, 4, 400, tabu=3, stall=15, threads=8)
    assert tr[:, TR_FIELDS.index('kick')].sum() > 3


def test_large_cells_random_init():
    """Random inits that end with no rng draws to spare (K6(10,4): 27840 of
    them) and shuffles longer than what is left after the pick: every later
    draw must come from a freshly generated ring slot."""
    compare(6, 10, 4, 2784, 1, 12, check_cnt=False, threads=8)
    compare(7, 9, 4, 1480, 1, 12, check_cnt=False, threads=8)
    compare(6, 9, 4, 1400, 1, 60, check_cnt=False, threads=8)


def test_selection_above_one_wave():
    """A dense code whose steps offer more than 256 candidate moves: the
    selection then runs over the whole workgroup in chunks rather than on one
    wave; both forms must equal the oracle, so the run must take both."""
    tr = compare(6, 9, 4, 1300, 3, 30, check_cnt=False, threads=8)
    nc = tr[:, TR_FIELDS.index('ncand')]
    assert (nc > 256).any() and (nc <= 256).any(), nc


def _prof_id(r):
    return 'K%d(%d,%d)M%d-s%d-it%d-%s' % (r['q'], r['n'], r['R'], r['M'], r['seed'], r['iters'],
                                         '_'.join(r['flags']))


@pytest.mark.parametrize('r', [r for r in golden_runs() if 'prof' in r], ids=_prof_id)
def test_step_sums_match_reference_profile(r):
    """Over whole runs of up to 10^4 moves (K6(9,4) seed 1 included), the
    engine's per-step records sum to the reference's -DPROF aggregates:
    mean candidates, sample size, uncovered count before each step, selected
    delta and distinct codewords (covsearch2e.c:1590-1618)."""
    from tests._util import prof_aggregates, prof_matches
    f = r['flags']
    tabu = int(f[f.index('--tabu') + 1])
    stall = int(f[f.index('--stall') + 1]) if '--stall' in f else 5000
    print('case prof %s' % _prof_id(r), flush=True)
    cell = _cell(r['q'], r['n'], r['R'], r['M'])
    cell.search_params(tenure=tabu, stall=stall, kick=2)
    cell.search_init(r['seed'], greedy='--random-init' not in f)
    unc0 = cell.state()['uncovered']
    tr = cell.run(r['iters'], mode=1, trace=True)
    st = cell.check_fault()
    _, bu = cell.best()
    cell.close()
    assert (bu, st['it'], st['kicks']) == (r['uncovered'], r['iters_done'], r['kicks'])
    assert prof_matches(prof_aggregates(tr, unc0), r['prof']) == {}


# --wide: every move of the candidate codewords while the list is short
def _seed41():
    import os
    from tests._util import DATA, read_code
    return read_code(os.path.join(DATA, 'seeds_K6_6_3_M41.txt'), 6)


@pytest.mark.parametrize('leave', LEAVES)
@pytest.mark.parametrize('mode', [0, 1])
def test_wide_k6_6_3_seed_file(leave, mode):
    tr = compare(6, 6, 3, 41, 1, 3000, seed_code=_seed41(), leave=leave, mode=mode, wide=True)
    assert (tr[:, 2] > 200).any()          # steps with the wide candidate set (n(q-1) = 30 per codeword)


def test_wide_kicks_and_sampling():
    compare(6, 6, 3, 41, 2, 2000, tabu=3, stall=20, seed_code=_seed41(), wide=True)
    compare(6, 6, 3, 41, 5, 1500, cand=50, seed_code=_seed41(), wide=True)


@pytest.mark.parametrize('leave', LEAVES)
def test_wide_k6_9_4(leave):
    tr = compare(6, 9, 4, 660, 1, 400, leave=leave, wide=True, threads=8)
    assert tr[-1, 2] > 300


def test_descent_many_removals():
    """Remove-and-repair over many removals (cov/engine/covsearch2e.c:1177-1189),
    run on the device with one wait at the end: K6(9,4)'s 661-word record code
    down to 650 words, then random K5(7,2) seeds 20-120 words over M = 400 with
    repeated words (ties at zero uniquely covered words, where the first goes),
    one cell taking a larger, a smaller and a larger seed in turn so that its
    descent buffer is reused and regrown."""
    import os
    from tests._util import DATA, read_code
    compare(6, 9, 4, 650, 3, 200, seed_code=read_code(os.path.join(DATA, 'K6_9_4_M661.txt'), 9), threads=8)
    rng = np.random.default_rng(5)
    cell = _cell(5, 7, 2, 400)
    cell.search_params(tenure=10, stall=5000, kick=2, cand=100000)
    for extra in (70, 20, 120):
        print('case descent K5(7,2) M=400 from %d words' % (400 + extra), flush=True)
        sc = rng.integers(0, 5, size=(400 + extra, 7), dtype=np.uint8)
        sc[7] = sc[2]
        sc[9] = sc[2]
        cell.search_init(3, seed_code=sc)
        o = Oracle(5, 7, 2, 400, 3).init(sc)
        assert cell.state()['uncovered'] == o.uncovered
        assert (cell.code() == o.code()).all()
        assert (cell.cnt() == o.cnt()).all()
        gpu = cell.run(300, trace=True)
        ref = o.run(300, tabu=10, stall=5000, cand=100000, trace=True)
        assert (gpu == ref).all()
        assert (cell.code() == o.code()).all()
    cell.close()
