"""Batched cells (one evaluator, selection and commit launch per step for all cells)
against the same cells run alone: every cell's counters, code, best code and
coverage counts must be identical, whatever the mix of shapes, launch modes,
run splits, kicks and early covers; one cell is also checked against the
oracle step for step."""
import hashlib
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from tests._util import cli_code_sha, golden_runs

pytestmark = pytest.mark.gpu

# (q, n, R, M, seed, tenure, stall)
MIX = [(5, 7, 2, 400, 1, 10, 5000), (6, 6, 3, 41, 2, 3, 20), (4, 8, 3, 60, 3, 10, 5000),
       (3, 8, 2, 120, 4, 10, 5000), (7, 5, 2, 60, 5, 10, 5000), (6, 9, 4, 660, 6, 10, 5000),
       (3, 4, 1, 9, 7, 10, 5000), (2, 10, 3, 70, 8, 10, 300)]


def _cell(spec, leave=0):
    from coldcase_amd import Cell
    q, n, R, M, seed, tenure, stall = spec
    c = Cell(q, n, R, M)
    c.search_params(tenure=tenure, stall=stall, kick=2)
    c.set_leave_mode(leave)
    c.search_init(seed)
    return c


def _snap(c):
    st = c.state()
    best, bu = c.best()
    return (st['it'], st['uncovered'], st['best_uncovered'], st['kicks'], st['draws'], st['done'],
            c.code().tobytes(), best.tobytes(), bu, c.cnt().tobytes())


def _solo(spec, steps, leave=0):
    c = _cell(spec, leave)
    c.run(steps, mode=1)
    s = _snap(c)
    c.close()
    return s


@pytest.mark.parametrize('mode', [0, 1])
def test_batch_equals_solo(mode):
    from coldcase_amd import Batch
    steps = [137, 250]
    print('case batch mix mode=%d steps=%s' % (mode, steps), flush=True)
    cells = [_cell(s) for s in MIX]
    with Batch(cells) as b:
        for k in steps:
            b.run(k, mode=mode)
    for spec, c in zip(MIX, cells):
        want = _solo(spec, sum(steps))
        got = _snap(c)
        assert got[:6] == want[:6], (spec, got[:6], want[:6])
        assert got[6:] == want[6:], spec
        c.close()


def test_batch_kept_equals_solo_walked():
    """Every cell of the mix keeps its leaving counts inside the batch
    (kicks included) and must end exactly where it ends alone with the
    leaving side walked."""
    from coldcase_amd import Batch
    from coldcase_amd.engine import LEAVE_KEPT, LEAVE_WALK
    print('case batch mix kept vs solo walked', flush=True)
    cells = [_cell(s, LEAVE_KEPT) for s in MIX]
    with Batch(cells) as b:
        b.run(200, mode=1)
        b.run(187, mode=0)
    assert cells[1].state()['kicks'] > 5
    for spec, c in zip(MIX, cells):
        assert _snap(c) == _solo(spec, 387, LEAVE_WALK), spec
        c.close()


def test_batch_kicks_and_covers():
    """A cell that covers early stops; the others go on; kicks happen inside
    the batch."""
    from coldcase_amd import Batch
    specs = [(3, 4, 1, 9, 7, 10, 5000), (6, 6, 3, 41, 2, 3, 20), (5, 7, 2, 400, 3, 10, 5000)]
    print('case batch kicks/covers', flush=True)
    cells = [_cell(s) for s in specs]
    with Batch(cells) as b:
        b.run(400, mode=1)
    st = [c.state() for c in cells]
    assert st[1]['kicks'] > 5
    for spec, c in zip(specs, cells):
        assert _snap(c) == _solo(spec, 400), spec
        c.close()


def test_batch_cell_matches_oracle():
    from coldcase_amd import Batch
    spec = (5, 7, 2, 400, 11, 10, 5000)
    print('case batch vs oracle', flush=True)
    cells = [_cell(spec), _cell((6, 6, 3, 41, 12, 3, 20))]
    with Batch(cells) as b:
        b.run(700, mode=1)
    o = Oracle(5, 7, 2, 400, 11).init()
    o.run(700, tabu=10, stall=5000)
    c = cells[0]
    st = c.state()
    assert (st['it'], st['uncovered'], st['best_uncovered'], st['kicks'], st['draws']) == \
        (o.iters, o.uncovered, o.best_uncovered, o.kicks, o.draws)
    assert (c.code() == o.code()).all() and (c.best()[0] == o.best()).all()
    assert (c.cnt() == o.cnt()).all()
    for c in cells:
        c.close()


def test_batch_rejects_bad_input():
    from coldcase_amd import Batch, CovError
    c = _cell(MIX[0])
    with pytest.raises(CovError):
        Batch([c, c])
    with pytest.raises(CovError):
        Batch([])
    c.close()


def test_fleet_equals_solo():
    """Cells split into concurrent batches by cost, run in chunks: every cell
    as alone."""
    from coldcase_amd import Fleet
    from coldcase_amd.autotune import plan_groups
    import math
    print('case fleet mix', flush=True)
    cells = [_cell(s) for s in MIX]
    costs = [math.comb(s[1], s[2]) * (s[0] - 1) ** s[2] * s[3] for s in MIX]
    groups = plan_groups(costs, 3)
    assert sorted(i for g in groups for i in g) == list(range(len(MIX)))
    with Fleet([[cells[i] for i in g] for g in groups]) as f:
        f.run(150, chunk=64)
        f.run(237, chunk=100)
    for spec, c in zip(MIX, cells):
        assert _snap(c) == _solo(spec, 387), spec
        c.close()


@pytest.mark.parametrize('spec', [(6, 6, 3, 41, 2, 3, 20), (5, 7, 2, 400, 3, 10, 5000)])
def test_best_code_after_long_stalls(spec):
    """Runs whose best improves again after more codeword changes than the
    device keeps a log of (about 4700 and 6300 here): the best code must still
    be the oracle's."""
    q, n, R, M, seed, tenure, stall = spec
    print('case best after stalls %s' % (spec,), flush=True)
    c = _cell(spec)
    c.run(12000, mode=1)
    o = Oracle(q, n, R, M, seed).init()
    o.run(12000, tabu=tenure, stall=stall)
    st = c.state()
    best, bu = c.best()
    assert (st['it'], st['uncovered'], st['best_uncovered'], st['kicks']) == \
        (o.iters, o.uncovered, o.best_uncovered, o.kicks)
    assert bu == o.best_uncovered
    assert (best == o.best()).all()
    assert (c.code() == o.code()).all()
    c.close()


def test_cells_created_while_the_device_is_busy():
    """Cells created and initialised while a long batch keeps another stream
    busy: every allocation is zeroed on the new cell's own stream, so its
    state is the one it has when created on an idle device (a zero fill left
    on another stream could land after the cell's initialisation)."""
    from coldcase_amd import Batch
    print('case create while busy', flush=True)
    busy = [_cell((6, 9, 4, 660, 11, 10, 5000)), _cell((7, 9, 4, 1480, 12, 10, 5000))]
    specs = [(5, 7, 2, 400, 21, 10, 5000), (6, 8, 4, 169, 22, 10, 5000), (6, 9, 4, 660, 23, 10, 5000)]
    with Batch(busy) as b:
        b.launch(20000)
        fresh = [_cell(s) for s in specs]
        assert not b.query(), 'the busy batch ended before the cells were created'
        for c in fresh:
            c.run(200, mode=1)
        b.sync()
    for spec, c in zip(specs, fresh):
        assert _snap(c) == _solo(spec, 200), spec
        c.close()
    for c in busy:
        c.close()


def _light(c):
    st = c.state()
    best, bu = c.best()
    return (st['it'], st['uncovered'], st['best_uncovered'], st['kicks'], st['draws'], st['done'], st['fault'],
            c.code().tobytes(), best.tobytes(), bu, hashlib.sha256(c.cnt().tobytes()).hexdigest())


def test_c4_grid_fleet():
    """The batched configuration itself: its 64 instances (the bench's own
    draw, coldcase_amd.workloads.c4_specs, seeds 1000003 + 7919 i), grouped
    by cost and scaled by the autotuner, advanced as a fleet of concurrent
    batches in chunks.  After 40 steps the instances with reference fixtures
    match them (best uncovered, iterations, kicks, published code); after 80
    every cell equals its solo run."""
    from coldcase_amd import Fleet
    from coldcase_amd.autotune import tune_fleet
    from coldcase_amd.workloads import c4_seed, c4_specs
    specs, redrawn = c4_specs()
    assert len(specs) == 64 and not redrawn
    print('case c4 fleet: 64 instances', flush=True)
    cells = [_cell((q, n, R, M, c4_seed(i), 10, 5000)) for i, (q, n, R, M) in enumerate(specs)]
    costs = [math.comb(n, R) * (q - 1) ** R + M for (q, n, R, M) in specs]
    groups, factors, idx = tune_fleet(cells, costs, ngroups=4, passes=1, steps=40)
    assert len(groups) == 4 and sorted(sum(idx, [])) == list(range(64))
    for i, c in enumerate(cells):
        c.search_init(c4_seed(i))
    gold = {(r['q'], r['n'], r['R'], r['M'], r['seed']): r for r in golden_runs() if r['tag'] == 'C4-exact'}
    assert len(gold) == 10
    with Fleet(groups) as f:
        f.run(40, chunk=20)
        seen = 0
        for i, (spec, c) in enumerate(zip(specs, cells)):
            r = gold.get(spec + (c4_seed(i),))
            if r is None:
                continue
            st = c.check_fault()
            best, bu = c.best()
            assert (bu, st['it'], st['kicks']) == (r['uncovered'], r['iters_done'], r['kicks']), spec
            assert cli_code_sha(best, spec[0], spec[3]) == r['out_sha256'], spec
            seen += 1
        assert seen == 10
        f.run(40, chunk=20)
    for i, (spec, c) in enumerate(zip(specs, cells)):
        solo = _cell(spec + (c4_seed(i), 10, 5000))
        solo.run(80, mode=1)
        want = _light(solo)
        solo.close()
        got = _light(c)
        assert got[:7] == want[:7], (i, spec, got[:7], want[:7])
        assert got[7:] == want[7:], (i, spec)
    for c in cells:
        c.close()


def test_wide_cells_in_a_batch_equal_solo():
    """--wide cells (and a standard one) in one batch end where they end alone."""
    import os
    from coldcase_amd import Batch, Cell
    from tests._util import DATA, read_code
    print('case batch wide', flush=True)
    seed41 = read_code(os.path.join(DATA, 'seeds_K6_6_3_M41.txt'), 6)
    specs = [(6, 6, 3, 41, 1, True), (6, 6, 3, 41, 2, True), (5, 6, 2, 70, 3, True), (6, 6, 3, 41, 4, False)]

    def make(sp):
        q, n, R, M, seed, wide = sp
        c = Cell(q, n, R, M)
        c.search_params(tenure=10)
        if wide:
            c.search_wide(True)
        c.search_init(seed, seed_code=seed41 if M == 41 else None)
        return c

    cells = [make(sp) for sp in specs]
    with Batch(cells) as b:
        b.run(700, mode=1)
        b.run(300, mode=0)
    for sp, c in zip(specs, cells):
        solo = make(sp)
        solo.run(1000, mode=1)
        assert _light(c) == _light(solo), sp
        solo.close()
        c.close()
    late = make(specs[3])
    with Batch([late]) as b:
        late.search_wide(True)
        with pytest.raises(Exception):
            b.run(10)
    late.close()
