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


@pytest.mark.parametrize('q,n,R,M,seed', [(5, 7, 2, 400, 1), (6, 6, 3, 41, 2), (3, 8, 3, 30, 3),
                                          (4, 6, 0, 300, 4), (7, 4, 1, 60, 5), (6, 9, 4, 660, 6)])
def test_eval_codewords_every_move(q, n, R, M, seed):
    """cov_eval_codewords: all n(q-1) moves of each codeword equal the
    oracle's wide evaluation (its exact uncovered set) and, on a sample, the
    per-move walk of the reference's base evaluator."""
    from coldcase_amd import Cell
    print('case eval_codewords K%d(%d,%d) M=%d' % (q, n, R, M), flush=True)
    rng = np.random.default_rng(seed + 50)
    o = Oracle(q, n, R, M, seed).init(greedy=False)
    with Cell(q, n, R, M) as c:
        c.set_code(o.code())
        cws = rng.choice(M, size=min(M, 10), replace=False).astype(np.int32)
        got = c.eval_codewords(cws)
    code = o.code()
    for t, i in enumerate(cws):
        assert np.array_equal(got[t], o.eval_codeword_wide(int(i))), int(i)
        for _ in range(4):
            p = int(rng.integers(n))
            v = int(rng.integers(q))
            if v != code[i, p]:
                assert got[t, p, v] == o.move_delta(int(i), p, v)


def test_every_move_delta_is_the_recount_change():
    """The defining identity, on every move of every codeword of a small
    cell: delta(i, p, v) = uncovered(code with c_i[p] := v) - uncovered(code),
    both counted from scratch on the CPU."""
    from coldcase_amd import Cell
    from oracle.verify_oracle import mindist
    q, n, R, M = 4, 5, 2, 12
    print('case recount identity K4(5,2) M=12', flush=True)
    code = np.random.default_rng(9).integers(0, q, size=(M, n)).astype(np.uint8)
    base = int((mindist(code, n, q) > R).sum())
    with Cell(q, n, R, M) as c:
        c.set_code(code)
        got = c.eval_codewords(np.arange(M, dtype=np.int32))
    for i in range(M):
        for p in range(n):
            for v in range(q):
                if v == code[i, p]:
                    continue
                moved = code.copy()
                moved[i, p] = v
                assert got[i, p, v] == int((mindist(moved, n, q) > R).sum()) - base, (i, p, v)
