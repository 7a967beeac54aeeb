"""Evaluator (kernel a) through the ABI: the shared-sphere delta of every
move equals the per-move walk of the oracle and the change of a full recount;
the commit applies exactly the evaluated delta (cov/search/covsearch.cu:572-576)."""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('q,n,R,M,seed', [(5, 7, 2, 400, 1), (6, 6, 3, 41, 2), (3, 8, 3, 30, 3),
                                          (4, 6, 0, 300, 4), (7, 4, 1, 60, 5), (6, 9, 4, 660, 6)])
def test_eval_all_moves_and_commit(q, n, R, M, seed):
    from coldcase_amd import Cell
    print('case eval K%d(%d,%d) M=%d' % (q, n, R, M), flush=True)
    rng = np.random.default_rng(seed)
    o = Oracle(q, n, R, M, seed).init(greedy=False)
    with Cell(q, n, R, M) as c:
        c.set_code(o.code())
        for trial in range(3):
            u = rng.integers(0, q, size=n).astype(np.uint8)
            cws = rng.choice(M, size=min(M, 12), replace=False).astype(np.int32)
            dv = c.eval_moves(u, cws)
            code = o.code()
            for t, i in enumerate(cws):
                for p in range(n):
                    if code[i, p] == u[p]:
                        continue
                    want = o.move_delta(int(i), p, int(u[p]))
                    assert dv[t, p] == want, (trial, int(i), p)
            # commit one evaluated move on both sides; deltas and counts agree
            t = int(rng.integers(len(cws)))
            i = int(cws[t])
            ps = [p for p in range(n) if code[i, p] != u[p]]
            if not ps:
                continue
            p = ps[0]
            d_gpu = c.commit(i, p, int(u[p]))
            d_ora = o.commit(i, p, int(u[p]))
            assert d_gpu == d_ora == dv[t, p]
            assert c.uncovered() == o.uncovered
            assert (c.code() == o.code()).all()
        assert (c.cnt() == o.cnt()).all()
        # the planes kept by the commit equal a fresh recount's
        c.recount()
        assert (c.cnt() == o.cnt()).all()
