"""The ruin-and-recreate engine's host logic on the CPU, over a brute-force
stand-in for the device cell (tests/_npcell.py): on cells where greedy
placement alone needs more codewords, the rounds find a cover at fewer, the
oracle confirms it, and peeling keeps it."""
import numpy as np
import pytest

import coldcase_amd.lns as L
from oracle.verify_oracle import digits_table, mindist
from tests._npcell import NumpyCell

# (q, n, R, M, seed): greedy alone covers K5(4,1) at 69, K3(6,1) at 89, K4(4,1) at 33
CASES = [(5, 4, 1, 64, 1), (3, 6, 1, 86, 2), (4, 4, 1, 29, 1)]


@pytest.mark.parametrize('q,n,R,M,seed', CASES)
def test_rounds_cover_below_greedy(monkeypatch, q, n, R, M, seed):
    monkeypatch.setattr(L, 'Cell', NumpyCell)
    e = L.Engine(q, n, R, seed=seed, log=lambda *a: None)
    e.load(np.zeros(0, np.int64))
    g = L.Engine(q, n, R, seed=seed, log=lambda *a: None)
    g.load(np.zeros(0, np.int64))
    g.greedy_fill(10 ** 6)
    assert len(g.code) > M and g.uncov == 0
    assert e.lns(M, round_budget_s=300, kmin=2, kmax=6, max_rounds=300) == 0
    assert 0 < e.stats['rounds'] < 300 and len(e.code) == M == len(set(e.code.tolist()))
    dig = digits_table(q, n)[e.code]
    assert (mindist(dig, n, q) <= R).all()
    # peel on the cover plus some extra words: still a cover, smaller
    extra = np.setdiff1d(np.arange(0, q ** n, 7), e.code)[:10]
    e.load(np.concatenate([e.code, extra]))
    assert e.peel() > 0 and e.uncov == 0 and e.recount() == 0
    assert (mindist(digits_table(q, n)[e.code], n, q) <= R).all()
