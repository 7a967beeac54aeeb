#!/usr/bin/env python3
"""covsearch -- the covsearch command-line contract, searched on the GPU.

A drop-in for the solver binary that cov/campaign.py runs through
COVSEARCH_BIN (cov/campaign.py:38,54-85) and that cov/engine/covengine
launches: the argv of cov/engine/covsearch2e.c:1042-1095, one RESULT line
(:1585-1589), the best code published atomically to --out (.part + rename,
at most once a second and at the end, always exactly M distinct codewords,
:791-868), exit status 0 iff the code covers (:1620).

The search itself runs on the device (coldcase_amd.engine.Cell): the same
trajectory as the reference for the same arguments, step for step.
Flags that only select an implementation in the reference (--preset, --st,
--eval, --walk, --hoist, --pf, --huge, --threads) are accepted and ignored,
since every one of them runs the identical search.  Flags that change the
search (--wide, --early, --fix0, --ucache, --upick, --sa, --noise) are not
part of this engine and are refused with exit status 2.
"""
import os
import sys
import time

import numpy as np

from .verify_cov import DIGITS

IGNORED_VALUE = {'--preset', '--st', '--eval', '--walk', '--threads', '--t0', '--cool', '--tmin', '--cpu',
                 '--target'}
IGNORED_FLAG = {'--hoist', '--pf', '--huge', '--quiet'}
REFUSED = {'--wide', '--early', '--fix0', '--ucache', '--upick', '--sa'}


def read_code(path, n):
    """The solver's own reader (cov/engine/covsearch2e.c:870-903): '#'
    comments, base-36 digits or separated integers, blank lines skipped."""
    out = []
    for line in open(path):
        line = line.split('#')[0]
        if any(ch in line for ch in ' ,\t'):
            toks = [t for t in line.replace(',', ' ').replace('\t', ' ').split() if t]
            w = [int(t) for t in toks][:64]
        else:
            w = [DIGITS.index(ch) for ch in line if ch in DIGITS][:64]
        if not w:
            continue
        if len(w) != n:
            sys.stderr.write('%s: bad codeword length %d\n' % (path, len(w)))
            sys.exit(1)
        out.append(w)
    return np.array(out, dtype=np.uint8).reshape(-1, n)


def distinct_fill(code, q, n, M):
    """Exactly M distinct codewords: first occurrences in order, then the
    smallest unused word indices (cov/engine/covsearch2e.c:791-832)."""
    pw = np.array([q ** (n - 1 - p) for p in range(n)], dtype=np.int64)
    idx = code.astype(np.int64) @ pw
    seen = set()
    rows = []
    for k, x in enumerate(idx[:M].tolist()):
        if x in seen:
            continue
        seen.add(x)
        rows.append(code[k])
    x = 0
    N = q ** n
    while len(rows) < M and x < N:
        if x not in seen:
            seen.add(x)
            rows.append(np.array([(x // int(pw[p])) % q for p in range(n)], dtype=np.uint8))
        x += 1
    return np.array(rows, dtype=np.uint8).reshape(-1, n)


def write_code(path, code, q):
    """Atomic publication: write <path>.part, then rename over path."""
    tmp = path + '.part'
    with open(tmp, 'w') as f:
        for w in code:
            if q <= 36:
                f.write(''.join(DIGITS[int(v)] for v in w) + '\n')
            else:
                f.write(' '.join(str(int(v)) for v in w) + '\n')
    os.replace(tmp, path)


def parse_args(argv):
    o = dict(q=0, n=0, R=-1, M=0, t=60.0, s=12345, inp=None, out=None, cand=100000, tabu=0, kick=2,
             iters=-1, stall=5000, greedy=True, quiet=False, noise=0)
    i = 0
    while i < len(argv):
        a = argv[i]
        has = i + 1 < len(argv)
        val = argv[i + 1] if has else None
        if a in ('-q', '-n', '-R', '-M') and has:
            o[a[1]] = int(val); i += 2
        elif a == '-t' and has:
            o['t'] = float(val); i += 2
        elif a == '-s' and has:
            o['s'] = int(val); i += 2
        elif a == '--in' and has:
            o['inp'] = val; i += 2
        elif a == '--out' and has:
            o['out'] = val; i += 2
        elif a in ('--cand', '--tabu', '--kick', '--stall', '--iters', '--noise') and has:
            o[a[2:]] = int(val); i += 2
        elif a == '--random-init':
            o['greedy'] = False; i += 1
        elif a == '--quiet':
            o['quiet'] = True; i += 1
        elif a in IGNORED_VALUE and has:
            i += 2
        elif a in IGNORED_FLAG:
            i += 1
        elif a in REFUSED:
            sys.stderr.write('option %s changes the search and is not supported by this engine\n' % a)
            sys.exit(2)
        else:
            sys.stderr.write('unknown/incomplete option %s\n' % a)
            sys.exit(2)
    if o['noise']:
        sys.stderr.write('option --noise changes the search and is not supported by this engine\n')
        sys.exit(2)
    return o


def main(argv=None):
    t_start = time.monotonic()
    argv = sys.argv[1:] if argv is None else argv
    o = parse_args(argv)
    q, n, R, M = o['q'], o['n'], o['R'], o['M']
    if q < 2 or n < 1 or R < 0 or M < 1:
        sys.stderr.write('usage: covsearch -q Q -n N -R R -M M [-t sec] [-s seed] [--in f] [--out f]\n'
                         '                 [--cand K] [--tabu T] [--kick K] [--iters I] [--stall S]\n'
                         '                 [--random-init] [--quiet]\n')
        return 2
    if q ** n > 4 * 10 ** 10:
        sys.stderr.write('q^n too large\n')
        return 2
    if M > 65535:
        sys.stderr.write('M > 65535 not supported (uint16 counters)\n')
        return 2
    verbose = not o['quiet']
    V = sum(__import__('math').comb(n, i) * (q - 1) ** i for i in range(min(R, n) + 1))
    if verbose:
        S = __import__('math').comb(n - 1, R) * (q - 1) ** R
        print('# covsearch q=%d n=%d R=%d M=%d  q^n=%d  ball=%d  sphere-move=%d  device=MI355X'
              % (q, n, R, M, q ** n, V, S))
        print('# sphere-covering bound: M >= %d' % ((q ** n + V - 1) // V))
        sys.stdout.flush()
    seed_code = read_code(o['inp'], n) if o['inp'] else None
    if seed_code is not None and (seed_code >= q).any():
        sys.stderr.write('digit out of range\n')
        return 1

    if R >= n:
        # every ball is the whole space: covered at once, nothing to search
        code = seed_code[:M] if seed_code is not None and len(seed_code) else np.zeros((0, n), np.uint8)
        if o['out']:
            write_code(o['out'], distinct_fill(code, q, n, M), q)
        el = time.monotonic() - t_start
        print('RESULT q=%d n=%d R=%d M=%d uncovered=0 iters=0 kicks=0 time=%.2f rate=0 cpu=%.3f '
              'cpurate=0 target=-1 ttt_cpu=-1.0000 ttt_wall=-1.0000 ttt_it=-1'
              % (q, n, R, M, el, time.process_time()))
        return 0

    from .engine import MODE_GRAPH, Cell
    cell = Cell(q, n, R, M)
    cell.search_params(tenure=o['tabu'], stall=o['stall'], kick=o['kick'], cand=o['cand'])
    cell.search_init(o['s'], seed_code=seed_code, greedy=o['greedy'])
    best, bu = cell.best()
    last_pub = -1e9
    if o['out']:
        write_code(o['out'], distinct_fill(best, q, n, M), q)
        last_pub = time.monotonic()
    if verbose:
        print('# init uncovered=%d' % cell.state()['uncovered'])
        sys.stdout.flush()
    t0 = time.monotonic()
    c0 = time.process_time()
    last_report = 0.0
    chunk = 64
    st = cell.state()
    while not st['done']:
        if time.monotonic() - t_start > o['t']:
            break
        if o['iters'] >= 0 and st['it'] >= o['iters']:
            break
        k = chunk if o['iters'] < 0 else min(chunk, o['iters'] - st['it'])
        prev_best = st['best_uncovered']
        cell.run(k, mode=MODE_GRAPH)
        st = cell.state()
        el = time.monotonic() - t0
        if chunk < 8192 and el < 0.5:
            chunk *= 2
        if st['best_uncovered'] < prev_best and o['out'] and (time.monotonic() - last_pub >= 1.0
                                                              or st['best_uncovered'] == 0):
            best, _ = cell.best()
            write_code(o['out'], distinct_fill(best, q, n, M), q)
            last_pub = time.monotonic()
        if verbose and el - last_report > 10.0:
            last_report = el
            print('# t=%6.1fs it=%-10d unc=%-10d best=%d (%.0f it/s)'
                  % (el, st['it'], st['uncovered'], st['best_uncovered'], st['it'] / max(el, 1e-9)))
            sys.stdout.flush()
    el = time.monotonic() - t0
    cpu = time.process_time() - c0
    it = st['it']
    print('RESULT q=%d n=%d R=%d M=%d uncovered=%d iters=%d kicks=%d time=%.2f rate=%.0f cpu=%.3f '
          'cpurate=%.0f target=-1 ttt_cpu=-1.0000 ttt_wall=-1.0000 ttt_it=-1'
          % (q, n, R, M, st['best_uncovered'], it, st['kicks'], el, it / (el if el > 0 else 1), cpu,
             it / (cpu if cpu > 0 else 1)))
    sys.stdout.flush()
    if o['out']:
        best, _ = cell.best()
        write_code(o['out'], distinct_fill(best, q, n, M), q)
    cell.close()
    return 0 if st['best_uncovered'] == 0 else 1


if __name__ == '__main__':
    sys.exit(main())
