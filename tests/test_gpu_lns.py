"""The ruin-and-recreate engine (coldcase_amd.lns, the engine of
cov/gemm2/gpuchain.py) on the GPU.  Its decisions rest on exact counts, so
they are checked exactly: every greedy placement has the largest gain of any
word (brute force over the whole space on the CPU), its incremental
uncovered count equals a recount and the oracle's count, a found cover is a
cover, peeling keeps a cover and leaves no codeword without private words."""
import os
import re

import numpy as np
import pytest

from oracle.verify_oracle import digits_table, mindist
from tests._util import DATA, read_code

pytestmark = pytest.mark.gpu


def _dist(q, n):
    d = digits_table(q, n).astype(np.int8)
    return (d[:, None, :] != d[None, :, :]).sum(axis=2).astype(np.int8)


def _count(code_idx, q, n, R):
    """The oracle's uncovered count of a code given by word indices."""
    d = digits_table(q, n)[np.asarray(code_idx, dtype=np.int64)]
    return int((mindist(d, n, q) > R).sum())


def _uncovered(code_idx, D, R):
    if not len(code_idx):
        return np.ones(D.shape[0], bool)
    return ~(D[:, np.asarray(code_idx)] <= R).any(axis=1)


def test_greedy_placements_are_true_argmax():
    from coldcase_amd.lns import Engine
    q, n, R = 4, 6, 2
    print('case lns greedy K4(6,2)', flush=True)
    D = _dist(q, n)
    e = Engine(q, n, R, seed=3, log=lambda *a: None)
    start = np.random.default_rng(1).choice(q ** n, size=5, replace=False)
    e.load(start)
    e.greedy_fill(30, batch=8, refresh_every=4, cap=64)
    code = list(np.unique(start))
    assert 0 < len(e.placed) <= 25 and (len(e.placed) == 25 or e.uncov == 0)
    for w, g in e.placed:
        unc = _uncovered(code, D, R)
        gains = (D <= R).astype(np.int32) @ unc.astype(np.int32)
        gains[code] = -1
        assert g == gains[w] == gains.max(), (w, g, gains[w], gains.max())
        code.append(w)
    assert e.uncov == int(_uncovered(code, D, R).sum()) == e.recount()
    e.close()


@pytest.mark.parametrize('q,n,R,M,seed', [(5, 4, 1, 64, 1), (3, 6, 1, 86, 2)])
def test_lns_takes_the_decisions_of_the_brute_force_cell(q, n, R, M, seed, tmp_path):
    """Every device answer is exact and ordered, so the engine on the GPU
    places, removes and restores exactly as on a brute-force stand-in cell,
    and finds the same cover; the published file is that cover, and peeling
    it with extra words keeps a cover in which every codeword covers some
    word alone."""
    import coldcase_amd.lns as L
    from tests._npcell import NumpyCell
    print('case lns K%d(%d,%d) M=%d seed %d vs stand-in' % (q, n, R, M, seed), flush=True)
    out = str(tmp_path / 'c.txt')
    e = L.Engine(q, n, R, seed=seed, out=out, log=lambda *a: None)
    e.load(np.zeros(0, np.int64))
    unc = e.lns(M, round_budget_s=120, kmin=2, kmax=6, max_rounds=300)
    real = L.Cell
    L.Cell = NumpyCell
    try:
        ref = L.Engine(q, n, R, seed=seed, log=lambda *a: None)
        ref.load(np.zeros(0, np.int64))
        ref.lns(M, round_budget_s=300, kmin=2, kmax=6, max_rounds=300)
    finally:
        L.Cell = real
    assert e.placed == ref.placed and e.stats == ref.stats
    assert unc == 0 == e.recount() and np.array_equal(np.sort(e.code), np.sort(ref.code))
    D = _dist(q, n)
    assert not _uncovered(e.code, D, R).any()
    e.publish(force=True)
    words = read_code(out, n)
    assert len(words) == M and not _uncovered(e.index(words), D, R).any()
    extra = np.setdiff1d(np.arange(0, q ** n, 7), e.code)[:10]
    e.load(np.concatenate([e.code, extra]))
    before = len(e.code)
    dropped = e.peel()
    assert dropped > 0 and len(e.code) == before - dropped
    assert e.uncov == 0 == e.recount() and not _uncovered(e.code, D, R).any()
    cnt = (D[:, e.code] <= R).sum(axis=1)
    private = [int(((D[:, c] <= R) & (cnt == 1)).sum()) for c in e.code]
    assert min(private) > 0 and np.array_equal(e.losses(), private)
    e.close()


def test_lns_rounds_keep_exact_counts_on_a_record_cell():
    """K6(6,3) from the reference benchmark's seed code: ruin of every kind,
    recreate, and the count after each round equals a recount."""
    from coldcase_amd.lns import Engine
    q, n, R, M = 6, 6, 3, 41
    print('case lns K6(6,3) seed file', flush=True)
    e = Engine(q, n, R, seed=5, log=lambda *a: None)
    e.load(read_code(os.path.join(DATA, 'seeds_K6_6_3_M41.txt'), n))
    assert len(e.code) == M
    for mode in ('low', 'hole', 'cluster', 'random'):
        e.ruin(4, mode)
        assert len(e.code) == M - 4
        assert e.uncov == e.recount() == _count(e.code, q, n, R)
        e.greedy_fill(M)
        assert len(e.code) == M and e.uncov == e.recount() == _count(e.code, q, n, R)
    e.lns(M, round_budget_s=60, max_rounds=30)
    assert e.uncov == e.recount() == _count(e.code, q, n, R)
    e.close()


def test_lns_cli(tmp_path, capsys):
    from coldcase_amd import lns
    print('case lns cli K3(5,1) M=27', flush=True)
    out = str(tmp_path / 'c.txt')
    rc = lns.main(['-q', '3', '-n', '5', '-R', '1', '-M', '27', '-t', '60', '-s', '2', '--out', out,
                   '--rounds', '200', '--kmin', '2', '--kmax', '5', '--quiet'])
    text = capsys.readouterr().out
    m = re.search(r'RESULT q=3 n=5 R=1 M=(\d+) uncovered=(\d+) iters=(\d+) rate=\S+ transforms=(\d+) time=', text)
    assert rc == 0 and m, text
    words = read_code(out, 5)
    idx = words.astype(np.int64) @ np.array([81, 27, 9, 3, 1])
    assert len(set(idx.tolist())) == len(idx) == int(m.group(1)) <= 27
    assert _count(idx, 3, 5, 1) == int(m.group(2))
