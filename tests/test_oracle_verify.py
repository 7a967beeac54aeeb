"""The numpy verifier restatements agree with each other, with brute force
and with the closed forms the reference's own verifier tests check
(cov/test_verify.py:60-123)."""
import random

import numpy as np
import pytest

from oracle.verify_oracle import (ball_volume, brute_uncovered, cnt_by_balls, dilate_counts,
                                  mindist)
from tests._codes import direct_sum, full_space, hamming_code


def unc_all(code, n, q, R):
    a = int((cnt_by_balls(code, n, q, R) == 0).sum())
    b = int((mindist(code, n, q) > R).sum())
    c = q ** n - dilate_counts(code, n, q, R)[R]
    return a, b, c


def test_random_codes_three_ways_and_brute():
    random.seed(20260818)
    for (q, n) in [(2, 6), (3, 5), (3, 6), (4, 5), (5, 4), (6, 4), (7, 3), (2, 9)]:
        for R in range(0, min(n, 4) + 1):
            M = random.randint(1, 12)
            code = list(dict.fromkeys(tuple(random.randrange(q) for _ in range(n)) for _ in range(M)))
            b = brute_uncovered(code, n, q, R)
            assert unc_all(code, n, q, R) == (b, b, b), (q, n, R, M)


def test_cnt_sums_to_M_times_V():
    rng = np.random.default_rng(5)
    for (q, n, R) in [(3, 5, 2), (5, 4, 1), (6, 4, 2)]:
        code = [tuple(r) for r in rng.integers(0, q, size=(9, n))]
        assert cnt_by_balls(code, n, q, R).sum() == len(code) * ball_volume(q, n, R)


def test_known_answers():
    full = full_space(2, 5)
    assert unc_all(full, 5, 2, 0) == (0, 0, 0)
    assert unc_all(full[:-1], 5, 2, 0) == (1, 1, 1)
    one = [tuple([0] * 5)]
    assert unc_all(one, 5, 3, 5)[0] == 0
    assert unc_all(one, 5, 3, 4) == (2 ** 5,) * 3          # (q-1)^n words at distance n
    # a single codeword's distance histogram is C(n,d)(q-1)^d
    d = mindist(one, 5, 3)
    from math import comb
    assert [int((d == k).sum()) for k in range(6)] == [comb(5, k) * 2 ** k for k in range(6)]


@pytest.mark.parametrize('q,r', [(2, 3), (3, 2), (4, 2), (5, 2)])
def test_perfect_hamming_codes(q, r):
    H = hamming_code(q, r)
    n = (q ** r - 1) // (q - 1)
    assert len(H) == q ** (n - r)
    cnt = cnt_by_balls(H, n, q, 1)
    assert (cnt == 1).all()                         # perfect: every word once
    assert int(mindist(H, n, q).max()) == 1
    # minus one word: exactly its ball, 1 + n(q-1) words, is uncovered
    assert unc_all(H[:-1], n, q, 1) == (1 + n * (q - 1),) * 3


def test_direct_sum_radius_adds():
    A = hamming_code(3, 2)                          # K_3(4,1) <= 9
    D = direct_sum(A, full_space(3, 2))             # radius 1 + 0
    assert int(mindist(D, 6, 3).max()) == 1
    D2 = direct_sum(A, [(0, 0)])                    # radius 1 + 2
    assert int(mindist(D2, 6, 3).max()) == 3
