"""Codes with known covering properties, for known-answer tests.

Perfect q-ary Hamming codes of length (q^r-1)/(q-1) have covering radius 1
and every word is covered exactly once; a direct sum of codes of radii R1 and
R2 covers at radius R1+R2 (the facts cov/test_verify.py:90-123 checks).
"""
from itertools import product

# GF(q) for q in {2, 3, 4, 5, 7}: q = 4 is GF(2)[x]/(x^2+x+1), elements as
# bit patterns of their polynomial
_GF4_MUL = [[0, 0, 0, 0], [0, 1, 2, 3], [0, 2, 3, 1], [0, 3, 1, 2]]


def gf_add(q, a, b):
    return a ^ b if q == 4 else (a + b) % q


def gf_mul(q, a, b):
    return _GF4_MUL[a][b] if q == 4 else (a * b) % q


def gf_neg(q, a):
    return a if q == 4 else (-a) % q


def hamming_code(q, r):
    """Null space of the r x n matrix whose columns are the projective points."""
    cols = [v for v in product(range(q), repeat=r)
            if any(v) and v[next(i for i, x in enumerate(v) if x)] == 1]
    n = len(cols)
    unit = [cols.index(tuple(1 if j == i else 0 for j in range(r))) for i in range(r)]
    rest = [j for j in range(n) if j not in unit]
    code = []
    for info in product(range(q), repeat=len(rest)):
        c = [0] * n
        acc = [0] * r
        for j, val in zip(rest, info):
            c[j] = val
            for i in range(r):
                acc[i] = gf_add(q, acc[i], gf_mul(q, val, cols[j][i]))
        for i in range(r):
            c[unit[i]] = gf_neg(q, acc[i])
        code.append(tuple(c))
    return code


def full_space(q, n):
    return [tuple(w) for w in product(range(q), repeat=n)]


def direct_sum(A, B):
    return [a + b for a in A for b in B]
