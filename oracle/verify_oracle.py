"""numpy restatements of the reference verifiers (TEST INFRASTRUCTURE ONLY).

The checker for the GPU coverage, reduction and distance-transform kernels;
never imported by the product.

  brute_uncovered   cov/test_verify.py:33-48: scan every word against the code
  cnt_by_balls      coverage multiplicity by explicit ball enumeration, the
                    counting form of verify_pure (cov/verify_cov.py:109-152)
  mindist           exact min distance per word (cov/verify_independent.py:144-185)
  dilate_counts     covered count after r rounds of axis dilation
                    (cov/verify_independent.py:128-141)
  ball_volume       cov/verify_cov.py:101-102
  count_layers      N_d(x) = #{u in S : d(x,u) = d}, d = 0..R, by comparing
                    every word with every member of S (the brute force of
                    cov/gemm2/test_dct.py:26-39, MSB-first indices)
"""
import itertools
from math import comb

import numpy as np


def ball_volume(q, n, R):
    return sum(comb(n, i) * (q - 1) ** i for i in range(min(R, n) + 1))


def digits_table(q, n):
    """(q^n, n) uint8: row w holds the MSB-first digits of w."""
    N = q ** n
    idx = np.arange(N, dtype=np.int64)
    out = np.empty((N, n), dtype=np.uint8)
    for p in range(n):
        out[:, p] = (idx // q ** (n - 1 - p)) % q
    return out


def brute_uncovered(code, n, q, R):
    """Words of Z_q^n farther than R from every codeword, by exhaustive scan."""
    unc = 0
    for w in itertools.product(range(q), repeat=n):
        for c in code:
            d = 0
            for a, b in zip(c, w):
                if a != b:
                    d += 1
                    if d > R:
                        break
            if d <= R:
                break
        else:
            unc += 1
    return unc


def cnt_by_balls(code, n, q, R):
    """cnt[w] = #{codewords within R of w}, by enumerating each ball."""
    N = q ** n
    pw = [q ** (n - 1 - p) for p in range(n)]
    cnt = np.zeros(N, dtype=np.int64)
    for c in code:
        base = sum(int(c[p]) * pw[p] for p in range(n))
        cnt[base] += 1
        for r in range(1, min(R, n) + 1):
            for pos in itertools.combinations(range(n), r):
                choices = [[v for v in range(q) if v != c[p]] for p in pos]
                for vals in itertools.product(*choices):
                    w = base
                    for p, v in zip(pos, vals):
                        w += (v - int(c[p])) * pw[p]
                    cnt[w] += 1
    return cnt


def mindist(code, n, q, chunk=1 << 18):
    """dist[w] = min Hamming distance from w to the code (uint8)."""
    N = q ** n
    C = np.asarray(code, dtype=np.uint8).reshape(-1, n)
    out = np.empty(N, dtype=np.uint8)
    weights = np.array([q ** (n - 1 - p) for p in range(n)], dtype=np.int64)
    for s in range(0, N, chunk):
        m = min(chunk, N - s)
        idx = np.arange(s, s + m, dtype=np.int64)
        dig = np.empty((m, n), dtype=np.uint8)
        for p in range(n):
            dig[:, p] = (idx // weights[p]) % q
        best = np.full(m, n + 1, dtype=np.int16)
        for c in C:
            d = (dig != c[None, :]).sum(axis=1).astype(np.int16)
            np.minimum(best, d, out=best)
        out[s:s + m] = best
    return out


def dilate_counts(code, n, q, R):
    """covered[r] = #words within r of the code, for r = 0..R, by axis dilation."""
    N = q ** n
    w = np.array([q ** (n - 1 - p) for p in range(n)], dtype=np.int64)
    cov = np.zeros(N, dtype=bool)
    C = np.asarray(code, dtype=np.int64).reshape(-1, n)
    cov[(C * w).sum(axis=1)] = True
    out = [int(cov.sum())]
    for _ in range(R):
        nxt = np.zeros(N, dtype=bool)
        for p in range(n):
            v = cov.reshape(q ** p, q, q ** (n - 1 - p))
            nxt.reshape(q ** p, q, q ** (n - 1 - p))[:] |= v.any(axis=1, keepdims=True)
        cov = nxt
        out.append(int(cov.sum()))
    return out


def count_layers(S_idx, n, q, R):
    """(R+1, q^n) int32: N_d(x) for the multiset S of word indices."""
    N = q ** n
    digs = digits_table(q, n)
    out = np.zeros((R + 1, N), dtype=np.int32)
    for u in np.asarray(S_idx, dtype=np.int64):
        d = (digs != digs[u][None, :]).sum(axis=1)
        for r in range(R + 1):
            out[r] += (d == r)
    return out
