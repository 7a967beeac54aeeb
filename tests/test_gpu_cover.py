"""Kernels (b), (c), (d) on the GPU against the oracle, brute force, closed
forms and the record codes, through the C ABI."""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from oracle.verify_oracle import ball_volume, cnt_by_balls, dilate_counts, mindist
from tests._codes import direct_sum, full_space, hamming_code
from tests._util import DATA, read_code

pytestmark = pytest.mark.gpu

SHAPES = [(2, 6, 2, 7), (3, 5, 2, 9), (4, 5, 3, 5), (5, 4, 1, 40), (6, 4, 2, 12), (7, 3, 1, 20),
          (3, 8, 3, 30), (5, 7, 2, 400), (6, 6, 3, 41), (2, 12, 4, 60), (8, 5, 2, 100), (16, 3, 1, 50)]


def cell(q, n, R, M):
    from coldcase_amd import Cell
    return Cell(q, n, R, M)


@pytest.mark.parametrize('q,n,R,M', SHAPES)
def test_recount_and_reductions(q, n, R, M):
    print('case recount K%d(%d,%d) M=%d' % (q, n, R, M), flush=True)
    rng = np.random.default_rng(q * 100 + n * 10 + R)
    code = rng.integers(0, q, size=(M, n)).astype(np.uint8)
    code[1] = code[0]                                   # a repeated codeword counts twice
    want = cnt_by_balls([tuple(c) for c in code], n, q, R)
    with cell(q, n, R, M) as c:
        c.set_code(code)
        got = c.cnt().astype(np.int64)
        assert (got == want).all()
        assert got.sum() == M * ball_volume(q, n, R)
        unc = int((want == 0).sum())
        assert c.uncovered() == unc
        assert c.state()['uncovered'] == unc
        zeros = np.flatnonzero(want == 0)
        idx, tot = c.gather_uncovered(7)
        assert tot == unc and (idx == zeros[:7]).all()
        for start in rng.integers(0, q ** n, size=5):
            after = zeros[zeros >= start]
            exp = int(after[0]) if after.size else (int(zeros[0]) if zeros.size else -1)
            assert c.first_uncovered_from(int(start)) == exp


@pytest.mark.parametrize('q,n,R,M', SHAPES[:8])
def test_distance_transform(q, n, R, M):
    print('case transform K%d(%d,%d) M=%d' % (q, n, R, M), flush=True)
    rng = np.random.default_rng(7 + q + n)
    code = rng.integers(0, q, size=(M, n)).astype(np.uint8)
    want = mindist(code, n, q)
    with cell(q, n, R, M) as c:
        c.set_code(code)
        dist, hist, rad = c.distance_transform(want_array=True)
        assert (dist == want).all()
        assert rad == int(want.max())
        assert (hist == np.bincount(want, minlength=n + 1)[:n + 1]).all()
        # #{dist <= r} is the dilation count after r rounds
        dil = dilate_counts(code, n, q, min(R + 1, n))
        for r in range(len(dil)):
            assert int(hist[:r + 1].sum()) == dil[r]
        # dist <= R  <=>  cnt > 0
        assert ((dist <= R) == (c.cnt() > 0)).all()


def test_ball_updates_single_and_multi():
    q, n, R, M = 5, 6, 2, 30
    rng = np.random.default_rng(3)
    code = rng.integers(0, q, size=(M, n)).astype(np.uint8)
    with cell(q, n, R, M) as c:
        c.set_code(code)
        base = c.cnt().astype(np.int64)
        w = int(rng.integers(q ** n))
        wd = tuple(int((w // q ** (n - 1 - p)) % q) for p in range(n))
        one = cnt_by_balls([wd], n, q, R)
        c.ball_update([w], +1)
        assert (c.cnt() == base + one).all()
        assert c.state()['uncovered'] == int(((base + one) == 0).sum())
        c.ball_update([w], -1)
        assert (c.cnt() == base).all()
        ws = rng.integers(0, q ** n, size=5)
        more = cnt_by_balls([tuple(int((x // q ** (n - 1 - p)) % q) for p in range(n)) for x in ws], n, q, R)
        c.ball_update(ws, +1)
        assert (c.cnt() == base + more).all()
        assert c.uncovered() == int(((base + more) == 0).sum())


@pytest.mark.parametrize('q,r', [(2, 3), (3, 2), (4, 2), (5, 2)])
def test_perfect_hamming(q, r):
    H = np.array(hamming_code(q, r), dtype=np.uint8)
    n = (q ** r - 1) // (q - 1)
    with cell(q, n, 1, len(H)) as c:
        c.set_code(H)
        assert (c.cnt() == 1).all()
        assert c.uncovered() == 0
        _, hist, rad = c.distance_transform()
        assert rad == 1
    with cell(q, n, 1, len(H) - 1) as c:
        c.set_code(H[:-1])
        assert c.uncovered() == 1 + n * (q - 1)


def test_direct_sum_and_single_word():
    D = np.array(direct_sum(hamming_code(3, 2), [(0, 0)]), dtype=np.uint8)
    with cell(3, 6, 3, len(D)) as c:
        c.set_code(D)
        assert c.uncovered() == 0
        assert c.distance_transform()[2] == 3
    one = np.zeros((1, 5), np.uint8)
    with cell(3, 5, 4, 1) as c:
        c.set_code(one)
        assert c.uncovered() == 2 ** 5

@pytest.mark.parametrize('q,n,R,M,lds', [(3, 8, 3, 30, 4096), (5, 7, 2, 400, 4096), (2, 12, 4, 60, 8192),
                                         (6, 6, 3, 41, 63488), (7, 5, 2, 9, 4096), (16, 3, 1, 50, 8192),
                                         (3, 11, 3, 25, 4096), (2, 16, 4, 30, 4096), (4, 9, 3, 20, 4096)])
def test_distance_transform_tile_plans(q, n, R, M, lds):
    """Small LDS budgets force many-pass plans with ragged row chunks; the
    codeword bitmap must be clean again for the next code."""
    print('case transform plans K%d(%d,%d) M=%d lds=%d' % (q, n, R, M, lds), flush=True)
    rng = np.random.default_rng(lds + q)
    with cell(q, n, R, M) as c:
        c.set_transform_tile(lds, 512 if lds < 32768 else 256)
        for rep in range(3):
            code = rng.integers(0, q, size=(M, n)).astype(np.uint8)
            want = mindist(code, n, q)
            c.set_code(code)
            dist, hist, rad = c.distance_transform(want_array=True)
            assert (dist == want).all(), rep
            assert rad == int(want.max())
            assert (hist == np.bincount(want, minlength=n + 1)[:n + 1]).all()


def axis_dp_dist(code, n, q):
    """dist[w] by the textbook axis DP over the whole array, one numpy pass
    per axis (min over the q words of a fiber, plus one); independent of the
    kernels' tiles, seed closed form and pass grouping."""
    N = q ** n
    d = np.full(N, 127, np.uint8)
    w = np.zeros(len(code), np.int64)
    for p in range(n):
        w = w * q + np.asarray(code)[:, p].astype(np.int64)
    d[w] = 0
    for j in range(n):
        v = d.reshape(q ** j, q, q ** (n - 1 - j))
        m = v.min(axis=1, keepdims=True)
        np.minimum(v, np.minimum(m, 126) + 1, out=v)
    return d


def test_axis_dp_reference_is_brute_force():
    rng = np.random.default_rng(5)
    for q, n, M in [(3, 6, 5), (4, 5, 9), (7, 4, 3), (2, 9, 4)]:
        code = rng.integers(0, q, size=(M, n)).astype(np.uint8)
        assert (axis_dp_dist(code, n, q) == mindist(code, n, q)).all()


@pytest.mark.parametrize('q,n,M,seed', [(7, 9, 1480, 1), (7, 9, 40, 2), (7, 9, 9000, 3), (6, 10, 2784, 4),
                                        (6, 9, 660, 5), (8, 8, 300, 6), (8, 9, 5000, 7), (5, 9, 3000, 8),
                                        (5, 10, 1000, 9), (3, 15, 2000, 10)])
def test_distance_transform_full_size(q, n, M, seed):
    """Whole distance arrays at the workload sizes (seed tiles in the closed
    form and by the row axes, two-stage and register column passes) against
    the numpy axis DP, with the histogram and the radius."""
    print('case transform full K%d(%d) M=%d' % (q, n, M), flush=True)
    rng = np.random.default_rng(seed)
    code = rng.integers(0, q, size=(M, n)).astype(np.uint8)
    want = axis_dp_dist(code, n, q)
    with cell(q, n, 1, M) as c:
        c.set_code(code, recount=False)
        dist, hist, rad = c.distance_transform(want_array=True)
        assert (dist == want).all()
        assert rad == int(want.max())
        assert (hist == np.bincount(want, minlength=n + 1)[:n + 1]).all()
        within, rad2 = c.distance_stats(3)
        assert within == int((want <= 3).sum()) and rad2 == rad


@pytest.mark.parametrize('fname,q,n,R', [('K6_9_4_M660.txt', 6, 9, 4), ('K7_9_4_M1480.txt', 7, 9, 4),
                                         ('K6_10_4_M2784.txt', 6, 10, 4), ('K6_8_4_M169.txt', 6, 8, 4)])
def test_record_codes_cover(fname, q, n, R):
    print('case record %s' % fname, flush=True)
    code = read_code(os.path.join(DATA, fname), n)
    with cell(q, n, R, len(code)) as c:
        c.set_code(code)
        assert c.uncovered() == 0
        assert c.state()['uncovered'] == 0
        _, hist, rad = c.distance_transform()
        assert rad <= R and int(hist.sum()) == q ** n
        assert int(c.cnt().astype(np.int64).sum()) == len(code) * ball_volume(q, n, R)


def test_recount_matches_oracle_k7_9_4():
    o = Oracle(7, 9, 4, 1480, 1).init(greedy=False)
    with cell(7, 9, 4, 1480) as c:
        c.set_code(o.code())
        assert (c.cnt() == o.cnt()).all()
        assert c.uncovered() == o.uncovered
        dist, hist, _ = c.distance_transform(want_array=True)
        assert int(hist[5:].sum()) == o.uncovered
        assert ((dist <= 4) == (o.cnt() > 0)).all()


@pytest.mark.parametrize('q,n,R,M', [(5, 7, 2, 400), (3, 8, 3, 30), (7, 5, 2, 9), (2, 12, 4, 60)])
def test_distance_stats(q, n, R, M):
    """The transform without its histogram: #{dist <= r} and the radius, and
    the far-word gather after it."""
    print('case distance stats K%d(%d,%d) M=%d' % (q, n, R, M), flush=True)
    rng = np.random.default_rng(100 + q)
    code = rng.integers(0, q, size=(M, n)).astype(np.uint8)
    want = mindist(code, n, q)
    with cell(q, n, R, M) as c:
        c.set_code(code, recount=False)
        for r in range(0, n + 1):
            within, rad = c.distance_stats(r)
            assert within == int((want <= r).sum()) and rad == int(want.max())
        far, tot = c.gather_far(R, 7)
        assert tot == int((want > R).sum())
        assert (far == np.flatnonzero(want > R)[:7]).all()
