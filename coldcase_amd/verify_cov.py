#!/usr/bin/env python3
"""verify_cov -- exhaustive coverage check of a q-ary code on the GPU.

Answers the question of cov/verify_cov.py -- does every word of Z_q^n lie
within Hamming distance R of the code? -- by two methods that share no
kernel, so that running both is a real cross-check:

  pure   coverage counts: every codeword's radius-R ball is added into
         cnt[w] by the LDS-tiled recount kernel; the uncovered words are
         cnt == 0, counted by a streaming reduction and listed in ascending
         order by a two-pass gather
  numpy  distances: the uint8 axis-DP transform gives dist[w], the distance
         from w to the nearest codeword; the uncovered words are dist > R

(The names are those of the reference's methods, which callers select by.)

What callers rely on, and what is kept:
  parse_code(path, q, n) -> (codewords, duplicates): codewords as tuples in
      first-occurrence order; '#' starts a comment; a line is either base-36
      characters, one per coordinate, or integers separated by blanks or
      commas; a bad line ends the program with a message naming it
  verify_pure / verify_numpy(code, n, q, R, want_witness=True,
      max_witness=5) -> (uncovered, witnesses), the witnesses being the
      first uncovered words in ascending MSB-first index order; MemoryError
      means the method cannot take this instance (q or n above 16, q^n above
      4e10, more than 65535 distinct codewords)
  covering_radius(code, n, q, method), ball_volume(q, n, R)
  the command line: CODEFILE [-q Q -n N -R R | --json SIDECAR]
      [--method auto|pure|numpy|both] [--radius] [--quiet]; one line
      "method <name> : uncovered=<count> ..." per method run (the shape that
      cov/engine/selftest.sh:61-62 and cov/gemm/run_bench.py:56 parse); exit
      status 0 covered, 1 not covered, 2 no method could run, 3 the methods
      disagree (cov/verify_cov.py:274-348)
"""
import argparse
import json
import os
import sys
import time
from math import comb

import numpy as np

DIGITS = '0123456789abcdefghijklmnopqrstuvwxyz'
LIMIT_Q = LIMIT_N = 16
LIMIT_M = 65535
LIMIT_WORDS = 4 * 10 ** 10
# instances the ball-marking method takes by default: a bounded number of
# words and of ball updates (the reference's thresholds for `auto`)
PURE_WORDS, PURE_MARKS = 3 * 10 ** 8, 5 * 10 ** 7


def ball_volume(q, n, R):
    """|B_R| = sum_{i <= R} C(n, i) (q - 1)^i."""
    return sum(comb(n, i) * (q - 1) ** i for i in range(0, min(R, n) + 1))


# ------------------------------------------------------------------ reading
def _digits_of_line(text, q, n):
    """The coordinates on one (comment-free, non-empty) line, or ValueError
    with the reason."""
    separated = any(sep in text for sep in (' ', '\t', ','))
    if separated:
        fields = text.replace(',', ' ').split()
        try:
            vals = [int(f) for f in fields]
        except ValueError:
            raise ValueError('cannot parse %r' % text) from None
    else:
        vals = [DIGITS.find(ch) for ch in text.lower()]
        if -1 in vals:
            raise ValueError('bad digit %r' % text[vals.index(-1)])
    if len(vals) != n:
        raise ValueError('length %d, expected n=%d' % (len(vals), n))
    bad = [v for v in vals if v < 0 or v >= q]
    if bad:
        raise ValueError('digit %d out of range for q=%d' % (bad[0], q))
    return tuple(vals)


def parse_code(path, q, n):
    """Read a code file: (distinct codewords in first-occurrence order, the
    number of repeated lines dropped)."""
    seen = {}
    lines = 0
    with open(path) as fh:
        for number, raw in enumerate(fh, start=1):
            text = raw.partition('#')[0].strip()
            if not text:
                continue
            try:
                word = _digits_of_line(text, q, n)
            except ValueError as why:
                raise SystemExit('%s:%d: %s' % (path, number, why))
            lines += 1
            seen.setdefault(word, None)
    if not seen:
        raise SystemExit('%s: the file holds no codeword' % path)
    return list(seen), lines - len(seen)


def fmt_word(w):
    """A word as the code files write it."""
    return ''.join(DIGITS[v] for v in w) if all(v < 36 for v in w) else ' '.join(map(str, w))


# ----------------------------------------------------------------- methods
def _admissible(code, n, q):
    """The GPU methods' limits; MemoryError names the one exceeded."""
    if q > LIMIT_Q or n > LIMIT_N:
        raise MemoryError('q=%d n=%d: the device methods take q, n <= 16' % (q, n))
    if q ** n > LIMIT_WORDS:
        raise MemoryError('%d words: above the 4e10 the device methods take' % q ** n)
    if len(code) > LIMIT_M:
        raise MemoryError('%d codewords: above the 65535 of 16-bit counts' % len(code))


def _word_digits(index, q, n):
    out = [0] * n
    for p in range(n - 1, -1, -1):
        index, out[p] = divmod(index, q)
    return tuple(out)


def _code_array(code, n):
    return np.asarray([tuple(w) for w in code], dtype=np.uint8).reshape(-1, n)


def verify_pure(code, n, q, R, want_witness=True, max_witness=5):
    """Coverage counts: (uncovered, first uncovered words ascending)."""
    from .engine import Cell
    _admissible(code, n, q)
    if R >= n:
        return 0, []
    words = _code_array(code, n)
    with Cell(q, n, R, len(words)) as cell:
        cell.set_code(words)
        missing = int(cell.uncovered())
        first = []
        if missing and want_witness and max_witness > 0:
            first = [_word_digits(int(i), q, n) for i in cell.gather_uncovered(max_witness)[0]]
    return missing, first


def verify_numpy(code, n, q, R, want_witness=True, max_witness=5, n1=None):
    """Distances: (uncovered, first uncovered words ascending).  `n1`, the
    reference method's split point, is accepted and has no meaning here."""
    from .engine import Cell
    _admissible(code, n, q)
    if R >= n:
        return 0, []
    words = _code_array(code, n)
    with Cell(q, n, R, len(words)) as cell:
        cell.set_code(words, recount=False)
        near, _ = cell.distance_stats(R)
        missing = q ** n - int(near)
        first = []
        if missing and want_witness and max_witness > 0:
            first = [_word_digits(int(i), q, n) for i in cell.gather_far(R, max_witness)[0]]
    return missing, first


def distance_histogram(code, n, q):
    """(hist, radius): hist[d] = #{w : distance to the code = d} for d = 0..n
    and the covering radius, from one transform (the answer of
    cov/verify_independent.py:144-185)."""
    from .engine import Cell
    _admissible(code, n, q)
    words = _code_array(code, n)
    with Cell(q, n, min(1, n - 1), len(words)) as cell:
        cell.set_code(words, recount=False)
        _, hist, radius = cell.distance_transform()
    return hist, radius


def covering_radius(code, n, q, method, rmax=None):
    """The least r at which `method` finds nothing uncovered, by bisection
    over 0..rmax (coverage only grows with r)."""
    lo = 0
    hi = n if rmax is None else rmax
    while lo < hi:
        r = lo + (hi - lo) // 2
        if method(code, n, q, r, want_witness=False)[0]:
            lo = r + 1
        else:
            hi = r
    return lo


METHODS = {'pure': verify_pure, 'numpy': verify_numpy}


def choose_method(n, q, R, M, forced):
    """`forced` unless it is 'auto': then ball marking for instances of
    bounded size, else the distance transform."""
    if forced != 'auto':
        return forced
    small = q ** n <= PURE_WORDS and M * ball_volume(q, n, R) <= PURE_MARKS
    return 'pure' if small else 'numpy'


def run(method, code, n, q, R, **kw):
    return METHODS[method](code, n, q, R, **kw)


# -------------------------------------------------------------------- CLI
def _shape(args, parser):
    """(q, n, R, sidecar M or None): the flags win over the sidecar."""
    side = {}
    if args.json:
        with open(args.json) as fh:
            side = json.load(fh)
    q, n, R = (v if v is not None else side.get(k) for k, v in (('q', args.q), ('n', args.n), ('R', args.R)))
    if q is None or n is None or R is None:
        parser.error('q, n and R are needed, as flags or in the --json sidecar')
    if q < 2 or n < 1 or R < 0:
        parser.error('need q >= 2, n >= 1, R >= 0')
    return q, n, R, side.get('M')


def main(argv=None):
    parser = argparse.ArgumentParser(prog='verify_cov', description=__doc__.split('\n')[0],
                                     formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument('codefile')
    parser.add_argument('-q', type=int, help='alphabet size')
    parser.add_argument('-n', type=int, help='length')
    parser.add_argument('-R', type=int, help='covering radius to check')
    parser.add_argument('--json', help='sidecar JSON giving q, n, R (and M)')
    parser.add_argument('--method', default='auto', choices=('auto', 'pure', 'numpy', 'both'))
    parser.add_argument('--radius', action='store_true', help='report the exact covering radius too')
    parser.add_argument('--quiet', action='store_true')
    args = parser.parse_args(argv)
    q, n, R, side_m = _shape(args, parser)
    code, repeated = parse_code(args.codefile, q, n)
    M, V = len(code), ball_volume(q, n, R)
    say = (lambda text: None) if args.quiet else print

    say('file    %s' % os.path.abspath(args.codefile))
    say('shape   q=%d n=%d R=%d, %d words, balls of %d' % (q, n, R, q ** n, V))
    say('code    %d distinct codewords%s' % (M, ', %d repeated lines ignored' % repeated if repeated else ''))
    if side_m is not None and side_m != M:
        say('note    the sidecar gives M=%d, the file %d distinct codewords' % (side_m, M))
    if M * V < q ** n:
        say('note    M*|ball| = %d is below q^n = %d: no covering is possible' % (M * V, q ** n))

    chosen = ('pure', 'numpy') if args.method == 'both' else (choose_method(n, q, R, M, args.method),)
    counts = {}
    for name in chosen:
        start = time.perf_counter()
        try:
            missing, first = run(name, code, n, q, R)
        except MemoryError as why:
            say('method %s : not run (%s)' % (name, why))
            continue
        counts[name] = missing
        say('method %s : uncovered=%d  in %.3f s' % (name, missing, time.perf_counter() - start))
        if first:
            say('          lowest uncovered words: %s' % ' '.join(fmt_word(w) for w in first))

    if not counts:
        say('RESULT  undecided: no method could run on this instance')
        return 2
    if len(set(counts.values())) != 1:
        say('RESULT  the methods disagree (%s): the answer cannot be trusted'
            % ', '.join('%s %d' % kv for kv in counts.items()))
        return 3
    missing = next(iter(counts.values()))
    if args.radius:
        # one transform gives the exact radius: the largest distance to the code
        _, radius = distance_histogram(code, n, q)
        say('radius  %d (exact covering radius)' % radius)
    if missing:
        say('RESULT  not covering: %d of %d words lie farther than %d from the code' % (missing, q ** n, R))
        return 1
    say('RESULT  covering: K_%d(%d,%d) <= %d verified' % (q, n, R, M))
    return 0


if __name__ == '__main__':
    sys.exit(main())
