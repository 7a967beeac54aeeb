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
since every one of them runs the identical search.  --wide (:1353-1376) is
supported, as the same search.  Flags that change the search otherwise
(--early, --fix0, --ucache, --upick, --sa, --noise) are not part of this
engine and are refused with exit status 2.

Stopping (:1305-1324): the device checks, before every step, the deadline
(-t seconds of wall clock from process start, or --cpu seconds of this
process's CPU time when that is given), the iteration budget (--iters) and
the target (--target T: stop once the best uncovered count is <= T).  The
time-to-target fields of the RESULT line (:1536-1542) give the iteration
whose move first reached the target, the wall seconds from the start of the
search to it (the device clock), and this process's CPU seconds to it.
While a --cpu budget runs, the host waits for the device by polling, so the
process's CPU time advances with the wall clock and the budget ends in about
as many wall seconds.
"""
import os
import sys
import time

import numpy as np

from .verify_cov import DIGITS

IGNORED_VALUE = {'--preset', '--st', '--eval', '--walk', '--threads', '--t0', '--cool', '--tmin'}
IGNORED_FLAG = {'--hoist', '--pf', '--huge', '--quiet'}
REFUSED = {'--early', '--fix0', '--ucache', '--upick', '--sa'}


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
             iters=-1, stall=5000, greedy=True, quiet=False, noise=0, target=-1, cpu=0.0, wide=False)
    i = 0
    while i < len(argv):
        a = argv[i]
        has = i + 1 < len(argv)
        val = argv[i + 1] if has else None
        if a in ('-q', '-n', '-R', '-M') and has:
            o[a[1]] = int(val); i += 2
        elif a == '-t' and has:
            o['t'] = float(val); i += 2
        elif a == '--target' and has:
            o['target'] = int(val); i += 2
        elif a == '--cpu' and has:
            o['cpu'] = float(val); i += 2
        elif a == '-s' and has:
            o['s'] = int(val); i += 2
        elif a == '--in' and has:
            o['inp'] = val; i += 2
        elif a == '--out' and has:
            o['out'] = val; i += 2
        elif a in ('--cand', '--tabu', '--kick', '--stall', '--iters', '--noise') and has:
            o[a[2:]] = int(val); i += 2
        elif a == '--wide':
            o['wide'] = True; i += 1
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
        hit = o['target'] >= 0
        print('RESULT q=%d n=%d R=%d M=%d uncovered=0 iters=0 kicks=0 time=%.2f rate=0 cpu=%.3f '
              'cpurate=0 target=%d ttt_cpu=%.4f ttt_wall=%.4f ttt_it=%d'
              % (q, n, R, M, el, time.process_time(), o['target'], 0.0 if hit else -1.0,
                 0.0 if hit else -1.0, 0 if hit else -1))
        return 0

    from .engine import GRAPH_STEPS, MODE_GRAPH, Cell
    cell = Cell(q, n, R, M)
    cell.search_params(tenure=o['tabu'], stall=o['stall'], kick=o['kick'], cand=o['cand'])
    if o['wide']:
        cell.search_wide(True)
    cell.search_init(o['s'], seed_code=seed_code, greedy=o['greedy'])
    best, bu = cell.best()
    last_pub = -1e9
    if o['out']:
        write_code(o['out'], distinct_fill(best, q, n, M), q)
        last_pub = time.monotonic()
    if verbose:
        print('# init uncovered=%d' % cell.state()['uncovered'])
        sys.stdout.flush()
    target, cpulimit = o['target'], o['cpu']
    t0 = time.monotonic()
    c0 = time.process_time()

    def budget_left():
        """Seconds left of the budget: CPU (--cpu) or wall from process start (-t)."""
        if cpulimit > 0.0:
            return cpulimit - (time.process_time() - c0)
        return o['t'] - (time.monotonic() - t_start)

    ttt_it, ttt_wall, ttt_cpu = -1, -1.0, -1.0
    left = budget_left()
    cell.limits(target, left if left > 0 else -1.0)
    if target >= 0 and bu <= target:
        ttt_it, ttt_wall, ttt_cpu = 0, 0.0, 0.0
    last_report = 0.0
    chunk = 64
    step_s = None               # seconds per step, from the last chunk
    st = cell.state()
    while not st['done']:
        left = budget_left()
        if left <= 0:
            break
        if o['iters'] >= 0 and st['it'] >= o['iters']:
            break
        k = chunk if o['iters'] < 0 else min(chunk, o['iters'] - st['it'])
        if step_s:
            # no more steps than the budget can hold, plus one graph: the
            # device stops at the deadline, and launches past it cost time too
            k = min(k, int(left / step_s) + GRAPH_STEPS)
        prev_best, prev_it = st['best_uncovered'], st['it']
        w0, p0 = time.monotonic(), time.process_time()
        if cpulimit > 0.0:
            # the device stops at the CPU budget's worth of wall seconds; the
            # host polls meanwhile, so its CPU time follows the wall clock
            cell.limits(target, left)
            cell.launch(k, mode=MODE_GRAPH)
            while not cell.query():
                pass
        else:
            cell.launch(k, mode=MODE_GRAPH)
            cell.sync()
        st = cell.check_fault()
        w1, p1 = time.monotonic(), time.process_time()
        if st['it'] - prev_it >= GRAPH_STEPS:
            step_s = (w1 - w0) / (st['it'] - prev_it)
        if target >= 0 and ttt_it < 0:
            it_hit, age = cell.ttt()
            if it_hit >= 0:
                ttt_it = it_hit
                ttt_wall = max(0.0, (time.monotonic() - t0) - age)
                # this process's CPU time at that moment, interpolated over the chunk
                f = (ttt_wall - (w0 - t0)) / (w1 - w0) if w1 > w0 else 1.0
                ttt_cpu = (p0 - c0) + min(max(f, 0.0), 1.0) * (p1 - p0)
        el = time.monotonic() - t0
        if w1 - w0 < 0.25 and chunk < 8192 and st['it'] - prev_it == k:
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
          'cpurate=%.0f target=%d ttt_cpu=%.4f ttt_wall=%.4f ttt_it=%d'
          % (q, n, R, M, st['best_uncovered'], it, st['kicks'], el, it / (el if el > 0 else 1), cpu,
             it / (cpu if cpu > 0 else 1), target, ttt_cpu, ttt_wall, ttt_it))
    sys.stdout.flush()
    if o['out']:
        best, _ = cell.best()
        write_code(o['out'], distinct_fill(best, q, n, M), q)
    cell.close()
    return 0 if st['best_uncovered'] == 0 else 1


if __name__ == '__main__':
    sys.exit(main())
