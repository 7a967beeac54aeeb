#!/usr/bin/env python3
"""verify_cov -- exact coverage check for q-ary covering codes, on the GPU.

Drop-in for cov/verify_cov.py: the same functions, arguments, return values,
parse rules, CLI and exit codes, with the two methods running on MI355X
kernels instead of a bytearray and numpy bitsets:

  pure   ball marking: cnt[w] = #{codewords within R} by the LDS-tiled
         recount kernel, then the uncovered count by a streaming reduction
         and the witnesses by an ordered gather (cov/verify_cov.py:109-152)
  numpy  the uint8 axis-DP distance transform: dist[w] = min distance to the
         code; uncovered = #{dist > R} from its histogram.  Shares no kernel
         with `pure`, so `--method both` is a real cross-check
         (cov/verify_cov.py:169-230 plays that role in the reference)

Witnesses are the first `max_witness` uncovered words in ascending MSB-first
index order, as both reference methods return them.  `MemoryError` still
means "this method cannot run this instance" (here: q > 16, n > 16, q^n above
4e10, or more than 65535 distinct codewords for `pure`).

Usage (as the reference):
  python -m coldcase_amd.verify_cov CODEFILE -q Q -n N -R R [--method auto|pure|numpy|both]
  python -m coldcase_amd.verify_cov CODEFILE --json META.json
  python -m coldcase_amd.verify_cov CODEFILE -q 3 -n 11 -R 4 --radius
Exit status is 0 iff the code is a verified covering code of radius <= R.
"""
import argparse
import json
import os
import sys
import time
from math import comb

import numpy as np

DIGITS = '0123456789abcdefghijklmnopqrstuvwxyz'
MAX_Q, MAX_N, MAX_M, MAX_SPACE = 16, 16, 65535, 4 * 10 ** 10


# --------------------------------------------------------------------------
# I/O (cov/verify_cov.py:60-102)
# --------------------------------------------------------------------------

def parse_code(path, q, n):
    """Read a code file.  Returns (codewords, n_duplicates)."""
    words = []
    with open(path) as f:
        for lineno, raw in enumerate(f, 1):
            line = raw.split('#')[0].strip()
            if not line:
                continue
            if any(ch in line for ch in ' ,\t'):
                toks = [t for t in line.replace(',', ' ').split() if t]
                try:
                    w = [int(t) for t in toks]
                except ValueError:
                    raise SystemExit('%s:%d: cannot parse %r' % (path, lineno, line))
            else:
                w = []
                for ch in line:
                    k = DIGITS.find(ch.lower())
                    if k < 0:
                        raise SystemExit('%s:%d: bad digit %r' % (path, lineno, ch))
                    w.append(k)
            if len(w) != n:
                raise SystemExit('%s:%d: length %d, expected n=%d' % (path, lineno, len(w), n))
            for v in w:
                if not 0 <= v < q:
                    raise SystemExit('%s:%d: digit %d out of range for q=%d' % (path, lineno, v, q))
            words.append(tuple(w))
    if not words:
        raise SystemExit('%s: no codewords' % path)
    uniq = list(dict.fromkeys(words))
    return uniq, len(words) - len(uniq)


def fmt_word(w):
    if max(w) < 36:
        return ''.join(DIGITS[v] for v in w)
    return ' '.join(str(v) for v in w)


def ball_volume(q, n, R):
    return sum(comb(n, i) * (q - 1) ** i for i in range(min(R, n) + 1))


def _digits_of(idx, q, n):
    out = []
    for p in range(n):
        out.append((idx // q ** (n - 1 - p)) % q)
    return tuple(int(v) for v in out)


def _check_shape(code, n, q, need_m_limit):
    if q > MAX_Q or n > MAX_N:
        raise MemoryError('the GPU methods take q <= %d and n <= %d' % (MAX_Q, MAX_N))
    if q ** n > MAX_SPACE:
        raise MemoryError('q^n = %d above %d' % (q ** n, MAX_SPACE))
    if need_m_limit and len(code) > MAX_M:
        raise MemoryError('more than %d codewords for uint16 counters' % MAX_M)


def _as_array(code, n):
    a = np.array([tuple(c) for c in code], dtype=np.uint8)
    return a.reshape(-1, n)


# --------------------------------------------------------------------------
# method 1: ball marking (recount + ordered uncovered gather)
# --------------------------------------------------------------------------

def verify_pure(code, n, q, R, want_witness=True, max_witness=5):
    """Exhaustive check by coverage counts.  Returns (n_uncovered, witnesses)."""
    from .engine import Cell
    _check_shape(code, n, q, True)
    if R >= n:                     # every ball is the whole space
        return 0, []
    arr = _as_array(code, n)
    with Cell(q, n, R, arr.shape[0]) as cell:
        cell.set_code(arr)
        n_unc = int(cell.uncovered())
        wit = []
        if n_unc and want_witness:
            idx, _ = cell.gather_uncovered(max_witness)
            wit = [_digits_of(int(x), q, n) for x in idx]
    return n_unc, wit


# --------------------------------------------------------------------------
# method 2: distance transform
# --------------------------------------------------------------------------

def verify_numpy(code, n, q, R, want_witness=True, max_witness=5, n1=None):
    """Exhaustive check by the min-distance transform.  Returns
    (n_uncovered, witnesses).  `n1` is accepted for call compatibility; the
    transform has no split."""
    from .engine import Cell
    _check_shape(code, n, q, True)
    if R >= n:
        return 0, []
    arr = _as_array(code, n)
    with Cell(q, n, R, arr.shape[0]) as cell:
        cell.set_code(arr, recount=False)
        within, _ = cell.distance_stats(R)
        n_unc = int(q ** n - within)
        wit = []
        if n_unc and want_witness:
            far, _ = cell.gather_far(R, max_witness)
            wit = [_digits_of(int(x), q, n) for x in far]
    return n_unc, wit


def distance_histogram(code, n, q):
    """hist[d] = #{w : min distance to the code == d}, d = 0..n, and the
    covering radius (cov/verify_independent.py:144-185)."""
    from .engine import Cell
    _check_shape(code, n, q, True)
    arr = _as_array(code, n)
    with Cell(q, n, min(1, n - 1), arr.shape[0]) as cell:
        cell.set_code(arr)
        _, hist, rad = cell.distance_transform()
    return hist, rad


# --------------------------------------------------------------------------
# covering radius (cov/verify_cov.py:237-248)
# --------------------------------------------------------------------------

def covering_radius(code, n, q, method, rmax=None):
    """Smallest r such that the code covers Z_q^n.  Monotone -> binary search."""
    lo, hi = 0, rmax if rmax is not None else n
    while lo < hi:
        mid = (lo + hi) // 2
        unc, _ = method(code, n, q, mid, want_witness=False)
        if unc == 0:
            hi = mid
        else:
            lo = mid + 1
    return lo


# --------------------------------------------------------------------------
# driver (cov/verify_cov.py:255-348)
# --------------------------------------------------------------------------

def choose_method(n, q, R, M, forced):
    if forced != 'auto':
        return forced
    total = q ** n
    if total <= 3 * 10 ** 8 and M * ball_volume(q, n, R) <= 5 * 10 ** 7:
        return 'pure'
    return 'numpy'


def run(method, code, n, q, R, **kw):
    if method == 'pure':
        return verify_pure(code, n, q, R, **kw)
    return verify_numpy(code, n, q, R, **kw)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('codefile')
    ap.add_argument('-q', type=int, help='alphabet size')
    ap.add_argument('-n', type=int, help='code length')
    ap.add_argument('-R', type=int, help='claimed covering radius')
    ap.add_argument('--json', help='sidecar JSON with q, n, R (and optionally M)')
    ap.add_argument('--method', default='auto', choices=['auto', 'pure', 'numpy', 'both'])
    ap.add_argument('--radius', action='store_true', help='also compute the exact covering radius')
    ap.add_argument('--quiet', action='store_true')
    a = ap.parse_args(argv)

    q, n, R, claimed_M = a.q, a.n, a.R, None
    if a.json:
        meta = json.load(open(a.json))
        q = q if q is not None else meta['q']
        n = n if n is not None else meta['n']
        R = R if R is not None else meta['R']
        claimed_M = meta.get('M')
    if None in (q, n, R):
        ap.error('need q, n and R (directly or via --json)')
    if q < 2 or n < 1 or R < 0:
        ap.error('bad parameters')

    code, ndup = parse_code(a.codefile, q, n)
    M = len(code)
    V = ball_volume(q, n, R)

    p = (lambda *s: None) if a.quiet else (lambda *s: print(*s))
    p('code      : %s' % os.path.abspath(a.codefile))
    p('parameters: q=%d n=%d R=%d   q^n=%d   ball volume=%d' % (q, n, R, q ** n, V))
    p('codewords : M=%d%s' % (M, '  (%d duplicate lines removed)' % ndup if ndup else ''))
    if claimed_M is not None and claimed_M != M:
        p('WARNING   : sidecar claims M=%d but file has %d distinct codewords' % (claimed_M, M))
    if M * V < q ** n:
        p('IMPOSSIBLE: M*V = %d < q^n = %d (sphere-covering bound)' % (M * V, q ** n))

    methods = ['pure', 'numpy'] if a.method == 'both' else [choose_method(n, q, R, M, a.method)]
    results = {}
    for m in methods:
        t0 = time.time()
        try:
            unc, wit = run(m, code, n, q, R)
        except MemoryError as e:
            p('method %-5s: SKIPPED (%s)' % (m, e))
            continue
        results[m] = unc
        p('method %-5s: uncovered=%d   (%.2fs)' % (m, unc, time.time() - t0))
        if wit:
            p('  first uncovered words: %s' % ', '.join(fmt_word(w) for w in wit))
    if not results:
        p('RESULT    : NO METHOD COULD RUN')
        return 2
    vals = set(results.values())
    if len(vals) > 1:
        p('RESULT    : METHODS DISAGREE %s  -- DO NOT TRUST' % results)
        return 3
    unc = vals.pop()

    if a.radius:
        meth = verify_pure if 'pure' in results else verify_numpy
        cr = covering_radius(code, n, q, meth, rmax=n)
        p('covering radius (exact): %d' % cr)

    if unc == 0:
        p('RESULT    : VERIFIED  -- K_%d(%d,%d) <= %d' % (q, n, R, M))
        return 0
    p('RESULT    : NOT A COVERING CODE (%d of %d words uncovered)' % (unc, q ** n))
    return 1


if __name__ == '__main__':
    sys.exit(main())
