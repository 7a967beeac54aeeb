"""The allocation-free gather (cov_gather_uncovered_noalloc) against
cov_gather_uncovered through the C ABI: the same total and the same
ascending indices on two calls in a row, the second on a smaller input, and
the stated limits refused with COV_ELIMIT."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cell(q, n, R, M, seed):
    from coldcase_amd import Cell
    c = Cell(q, n, R, M)
    c.set_code(np.random.default_rng(seed).integers(0, q, size=(M, n)).astype(np.uint8))
    return c


def test_noalloc_gather_equals_gather_on_two_calls():
    from coldcase_amd.engine import CovError
    print('case noalloc gather: K6(9,4) M=400 then K5(7,2) M=300', flush=True)
    big = _cell(6, 9, 4, 400, 1)          # tens of thousands uncovered
    small = _cell(5, 7, 2, 300, 2)        # fewer words, fewer uncovered
    want_b, tot_b = big.gather_uncovered(5000)
    got_b, gtot_b = big.gather_uncovered_noalloc(5000)
    assert gtot_b == tot_b == big.uncovered() and tot_b > 5000
    assert np.array_equal(got_b, want_b)
    want_s, tot_s = small.gather_uncovered(100)
    got_s, gtot_s = small.gather_uncovered_noalloc(100)
    assert gtot_s == tot_s == small.uncovered() and 100 < tot_s < tot_b
    assert np.array_equal(got_s, want_s)
    # the whole uncovered set of the small cell, and a cap of zero (total only)
    all_s, _ = small.gather_uncovered_noalloc(65536)
    assert len(all_s) == tot_s and np.array_equal(all_s, small.gather_uncovered(tot_s)[0])
    assert small.gather_uncovered_noalloc(0)[1] == tot_s
    cnt = small.cnt()
    assert np.array_equal(all_s, np.flatnonzero(cnt == 0))
    # the stated limits
    with pytest.raises(CovError):
        small.gather_uncovered_noalloc(65537)
    big.close()
    small.close()
    from coldcase_amd import Cell
    with Cell(7, 10, 1, 20) as huge:          # 7^10 words: above 2^26
        huge.set_code(np.zeros((20, 10), np.uint8))
        with pytest.raises(CovError):
            huge.gather_uncovered_noalloc(10)
