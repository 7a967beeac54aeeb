#!/usr/bin/env python3
"""Generate the golden trajectory fixtures from the reference solver.

Runs the reference `covsearch2e` (compiled unmodified from
cov/engine/covsearch2e.c by oracle/Makefile into oracle/_ref/) on a fixed list
of runs and records, per run, the RESULT fingerprint (best uncovered, iters,
kicks) and the sha256 of the code file it publishes (plus the text when it is
small).  A run is a pure function of its arguments, so these pin the whole
trajectory: rng consumption, candidate order, tie-breaks, tabu jitter, kicks.

Runs tagged `target` stop at --target T and also record the iteration that
first reached it (the RESULT line's ttt_it).  Runs listed in prof_runs() are
made with the -DPROF build (oracle/_ref/covsearch2e_prof) and also record its
per-step aggregates (mean candidates, sample size, uncovered count before
the step, selected delta and distinct codewords; cov/engine/covsearch2e.c:
1590-1618), which pin sums over every step of a long run.

Usage:  make -C oracle ref && python3 tests/golden/make_golden.py
Writes: tests/golden/trajectories.json
"""
import hashlib
import json
import math
import os
import re
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, 'oracle', '_ref', 'covsearch2e')
REF_PROF = os.path.join(ROOT, 'oracle', '_ref', 'covsearch2e_prof')
DATA = os.path.join(HERE, 'data')
RESULT = re.compile(r'RESULT .*?uncovered=(-?\d+) iters=(\d+) kicks=(\d+) .*?target=(-?\d+) .*?ttt_it=(-?\d+)')
PROF_KEYS = ('mean_ncand', 'mean_take', 'mean_uncovered', 'mean_bestd', 'mean_ncw')


def ball(q, n, R):
    return sum(math.comb(n, i) * (q - 1) ** i for i in range(min(R, n) + 1))


def runs():
    out = []

    def add(q, n, R, M, seed, iters, flags=(), infile=None, threads=1, tag=''):
        out.append(dict(q=q, n=n, R=R, M=M, seed=seed, iters=iters,
                        flags=list(flags), infile=infile, threads=threads, tag=tag))

    # the reference engine selftest cells (greedy init, tabu 0)
    for q, n, R, M, it in [(2, 6, 2, 12, 300), (3, 6, 2, 12, 300), (5, 5, 2, 30, 200),
                           (6, 6, 3, 41, 400), (6, 6, 3, 45, 400), (7, 4, 1, 60, 300),
                           (3, 11, 4, 81, 120), (6, 8, 4, 169, 40), (4, 5, 0, 1024, 5),
                           (5, 4, 3, 3, 50)]:
        for s in (1, 2):
            add(q, n, R, M, s, it, tag='engine-selftest')
    # solver test cases (seed 7) and the impossible size
    for q, n, R, M, it in [(3, 4, 1, 5, 200), (3, 4, 1, 9, 5000), (3, 5, 2, 6, 2000),
                           (4, 4, 2, 5, 2000), (5, 4, 1, 40, 3000), (6, 4, 2, 12, 3000),
                           (2, 7, 2, 8, 3000), (3, 3, 0, 20, 500), (3, 4, 3, 2, 200),
                           (7, 3, 1, 20, 2000), (6, 5, 2, 40, 4000), (3, 4, 1, 8, 200000)]:
        add(q, n, R, M, 7, it, tag='solver-test')
    # the tabu configuration of the benchmarks: random init, tenure 10
    T = ['--random-init', '--tabu', '10']
    for s in (1, 2, 3, 4, 5):
        for it in (1, 10, 100, 1000, 10000):
            add(5, 7, 2, 400, s, it, T, tag='C1')
    for s in (1, 2, 3):
        for it in (1, 3, 10, 30, 100, 300, 1000, 3000, 10000):
            add(6, 9, 4, 660, s, it, T, threads=8, tag='C2')
    for it in (1, 3, 10, 30, 100, 300):
        add(7, 9, 4, 1480, 1, it, T, threads=8, tag='C3')
    for it in (1, 3, 10, 30):
        add(6, 10, 4, 2784, 1, it, T, threads=8, tag='C5')
    # batched-grid shapes at M = ceil(1.3 q^n / V)
    for q, n, R in [(5, 7, 3), (6, 7, 2), (7, 7, 4), (8, 7, 3), (5, 8, 2), (6, 8, 3),
                    (7, 8, 2), (5, 9, 4), (6, 9, 3), (8, 8, 4)]:
        M = math.ceil(1.3 * q ** n / ball(q, n, R))
        it = 200 if q ** n < 2_000_000 else 20
        add(q, n, R, M, 11, it, T, threads=8, tag='C4-grid')
    # kick-heavy runs: pins the draw order of the kick perturbation
    for q, n, R, M, it, st in [(6, 6, 3, 41, 400, 20), (5, 7, 2, 400, 2000, 10),
                               (7, 4, 1, 60, 300, 7), (4, 6, 2, 20, 1500, 15),
                               (3, 7, 1, 40, 3000, 25), (6, 6, 3, 41, 3000, 50)]:
        for s in (1, 2, 3):
            add(q, n, R, M, s, it, T[:1] + ['--tabu', '3', '--stall', str(st)], tag='kick')
    # small candidate samples: partial Fisher-Yates
    for cand in (1, 3, 17):
        add(6, 6, 3, 41, 5, 500, T + ['--cand', str(cand)], tag='cand')
    # fixed starts from the reference benchmark seeds, and record codes
    add(6, 6, 3, 41, 1, 2000, ['--tabu', '10'], infile='seeds_K6_6_3_M41.txt', tag='seedfile')
    add(6, 8, 4, 169, 1, 60, ['--tabu', '10'], infile='seeds_K6_8_4_M169.txt', threads=8, tag='seedfile')
    add(6, 9, 4, 660, 1, 0, [], infile='K6_9_4_M660.txt', tag='record')
    # remove-and-repair descent 661 -> 660, then a few moves
    add(6, 9, 4, 660, 3, 5, ['--tabu', '10'], infile='K6_9_4_M661.txt', threads=8, tag='descent')
    # greedy init on a benchmark shape
    add(6, 9, 4, 660, 2, 10, ['--tabu', '10'], threads=8, tag='greedy')
    add(5, 7, 2, 400, 2, 1000, ['--tabu', '10'], tag='greedy')
    # --target: stop once the best uncovered count reaches T (the time-to-target
    # harnesses cov/opt/ttt.py:62 and cov/gemm/run_bench.py:68 run this way)
    for s, tg in ((1, 5000), (2, 4500), (3, 4200)):
        add(5, 7, 2, 400, s, 10000, T + ['--target', str(tg)], tag='target')
    add(5, 7, 2, 400, 4, 300, T + ['--target', '10'], tag='target')          # never reached
    for s in (1, 2, 3):
        add(6, 6, 3, 41, s, 20000, ['--tabu', '10', '--target', '30'], infile='seeds_K6_6_3_M41.txt',
            tag='target')
    add(6, 6, 3, 41, 4, 20000, ['--tabu', '3', '--stall', '50', '--target', '25'], infile='seeds_K6_6_3_M41.txt',
        tag='target')
    add(6, 8, 4, 169, 1, 3000, ['--tabu', '10', '--target', '200'], infile='seeds_K6_8_4_M169.txt', threads=8,
        tag='target')
    add(6, 9, 4, 660, 1, 10, ['--target', '0'], infile='K6_9_4_M660.txt', tag='target')   # met at the start
    # --wide: every move of a candidate codeword while the uncovered list is
    # valid and short (:1353-1376; cov/opt/ttt.py's p5b+wide variant)
    W = ['--wide', '--tabu', '10']
    for sd in (1, 2, 3):
        for it in (200, 2000):
            add(6, 6, 3, 41, sd, it, W, infile='seeds_K6_6_3_M41.txt', tag='wide')
        add(6, 6, 3, 41, sd, 20000, W + ['--target', '30'], infile='seeds_K6_6_3_M41.txt', tag='wide')
    for sd in (1, 2):
        add(6, 8, 4, 169, sd, 2000, W + ['--target', '20'], infile='seeds_K6_8_4_M169.txt', threads=8, tag='wide')
        add(6, 6, 3, 41, sd, 3000, ['--wide', '--tabu', '3', '--stall', '20'], infile='seeds_K6_6_3_M41.txt',
            tag='wide')
        add(5, 6, 2, 70, sd, 2000, ['--random-init'] + W, tag='wide')
    for it in (100, 1000, 3000):
        add(6, 9, 4, 660, 1, it, ['--random-init'] + W, threads=8, tag='wide')
    add(6, 6, 3, 41, 5, 2000, W + ['--cand', '50'], infile='seeds_K6_6_3_M41.txt', tag='wide')
    add(5, 7, 2, 400, 1, 3000, ['--random-init'] + W, tag='wide')        # the list never short enough
    # exact instances of the batched configuration (coldcase_amd/workloads.py:
    # c4_specs() index, its seed 1000003 + 7919 i), spanning the grid: the
    # largest space, the largest M, every radius
    sys.path.insert(0, ROOT)
    from coldcase_amd.workloads import c4_seed, c4_specs
    specs, _ = c4_specs()
    for i in (4, 6, 7, 10, 15, 20, 24, 36, 39, 59):
        q, n, R, M = specs[i]
        add(q, n, R, M, c4_seed(i), 40, T, threads=8, tag='C4-exact')
    return out


def prof_runs():
    """Runs also made with the -DPROF build, keyed like runs()."""
    T = ['--random-init', '--tabu', '10']
    out = []
    for s in (1, 2, 3, 4, 5):
        out.append(dict(q=5, n=7, R=2, M=400, seed=s, iters=10000, flags=T, infile=None, threads=1))
    out.append(dict(q=6, n=9, R=4, M=660, seed=1, iters=10000, flags=T, infile=None, threads=8))
    out.append(dict(q=5, n=7, R=2, M=400, seed=1, iters=2000, flags=T[:1] + ['--tabu', '3', '--stall', '10'],
                    infile=None, threads=1))
    return out


def prof_key(r):
    return (r['q'], r['n'], r['R'], r['M'], r['seed'], r['iters'], tuple(r['flags']), r['infile'])


def run_one(r, exe=REF):
    with tempfile.TemporaryDirectory() as td:
        outp = os.path.join(td, 'out.txt')
        cmd = [exe, '-q', str(r['q']), '-n', str(r['n']), '-R', str(r['R']), '-M', str(r['M']),
               '-s', str(r['seed']), '--iters', str(r['iters']), '--preset', 'p5b',
               '--threads', str(r['threads']), '-t', '1e9', '--quiet', '--out', outp] + r['flags']
        if r['infile']:
            cmd += ['--in', os.path.join(DATA, r['infile'])]
        p = subprocess.run(cmd, capture_output=True, text=True)
        m = RESULT.search(p.stdout)
        if not m:
            raise RuntimeError('no RESULT: %s\n%s' % (' '.join(cmd), p.stderr))
        text = open(outp).read()
    rec = dict(r)
    del rec['threads']
    rec['uncovered'] = int(m.group(1))
    rec['iters_done'] = int(m.group(2))
    rec['kicks'] = int(m.group(3))
    if int(m.group(4)) >= 0:
        rec['ttt_it'] = int(m.group(5))
    if exe == REF_PROF:
        prof = {}
        for line in p.stdout.splitlines():
            if line.startswith('# PROF'):
                for tok in line.split():
                    if '=' in tok:
                        k, v = tok.split('=', 1)
                        if k in PROF_KEYS:
                            prof[k] = float(v)
        rec['prof'] = prof
    rec['out_sha256'] = hashlib.sha256(text.encode()).hexdigest()
    if len(text) <= 2048:
        rec['out_text'] = text
    return rec


def main():
    if not os.path.exists(REF):
        sys.exit('build the reference first: make -C oracle ref')
    # --only TAG[,TAG]: regenerate those tags, keep every other record
    only = set(sys.argv[sys.argv.index('--only') + 1].split(',')) if '--only' in sys.argv else None
    todo = runs()
    kept = []
    if only is not None:
        todo = [r for r in todo if r['tag'] in only]
        kept = [r for r in json.load(open(os.path.join(HERE, 'trajectories.json')))['runs']
                if r['tag'] not in only]
    heavy = [r for r in todo if r['threads'] > 1]
    light = [r for r in todo if r['threads'] == 1]
    with ThreadPoolExecutor(max_workers=8) as ex:
        res_light = list(ex.map(run_one, light))
    res_heavy = [run_one(r) for r in heavy]
    res = res_light + res_heavy
    # the -DPROF aggregates, attached to the matching run (added if absent)
    byk = {prof_key(r): r for r in res}
    for pr in (prof_runs() if only is None else []):
        rec = run_one(pr, exe=REF_PROF)
        k = prof_key(pr)
        if k in byk:
            for f in ('uncovered', 'iters_done', 'kicks', 'out_sha256'):
                assert byk[k][f] == rec[f], ('the -DPROF build diverged', k, f)
            byk[k]['prof'] = rec['prof']
        else:
            rec['tag'] = 'prof'
            res.append(rec)
    res += kept
    res.sort(key=lambda r: (r['tag'], r['q'], r['n'], r['R'], r['M'], r['seed'], r['iters']))
    json.dump({'generator': 'tests/golden/make_golden.py',
               'solver': 'cov/engine/covsearch2e.c (reference, gcc -O3, --preset p5b)',
               'runs': res},
              open(os.path.join(HERE, 'trajectories.json'), 'w'), indent=1)
    print('%d runs written' % len(res))


if __name__ == '__main__':
    main()
