#!/usr/bin/env python3
"""Benchmark: tabu moves/s on K6(9,4) on one MI355X.

Workload (a "step" is one tabu move of the device-resident search):
  q=6, n=9, R=4 (10,077,696 words, ball 90,196), M=660 codewords, uniform
  random initial code from seed 1, tabu tenure 10, stall 5000, kick 2, all
  candidates evaluated; steps replayed from a captured hipGraph.
  W warmup steps, then K timed steps; inputs and state resident in HBM.

Printed: one JSON line with the metric, the evaluator kernel's roofline
(algorithmic bytes per launch / average launch time from HIP events on the
cell's stream, against 8 TB/s HBM) and the CPU baseline (the reference
solver compiled from source, else the oracle, on the host's cores over a
bounded prefix of the same search).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode graph|plain]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
Q, N_, R_, M_, SEED, TENURE = 6, 9, 4, 660, 1, 10


def cpu_baseline(cores, sample_iters):
    """Time the reference solver (or the oracle) on the host: K6(9,4) M=660,
    the same seed and settings, first `sample_iters` moves."""
    ref = os.path.join(ROOT, 'oracle', '_ref', 'covsearch2e')
    ora = os.path.join(ROOT, 'oracle', 'cso')
    if os.path.exists(ref):
        exe, kind = ref, 'reference'
    elif os.path.exists(ora):
        exe, kind = ora, 'port'
    else:
        return None
    cmd = [exe, '-q', str(Q), '-n', str(N_), '-R', str(R_), '-M', str(M_), '-s', str(SEED),
           '--random-init', '--tabu', str(TENURE), '--iters', str(sample_iters), '--preset', 'p5b',
           '--threads', str(cores), '-t', '1e9', '--quiet']
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    t0 = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True, env=env)
    wall = time.time() - t0
    m = re.search(r'iters=(\d+).*?time=([\d.]+) rate=(\d+)', p.stdout)
    if not m:
        return None
    it, tsec = int(m.group(1)), float(m.group(2))
    return {'value': round(it / tsec, 1) if tsec > 0 else None, 'unit': 'moves/s', 'cores': cores,
            'kind': kind,
            'sample': 'K6(9,4) M=660 seed 1 random init, tenure 10: the first %d moves (%.1f s search, '
                      '%.1f s with init), --preset p5b --threads %d' % (it, tsec, wall, cores)}


def traffic_from_profile():
    """HBM bytes per evaluator launch measured by rocprofv3 PMC counters,
    committed under profiles/ (None when absent)."""
    p = os.path.join(ROOT, 'profiles', 'eval_traffic_K6_9_4.json')
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get('bytes_per_launch')
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20000)
    ap.add_argument('--warmup', type=int, default=1000)
    ap.add_argument('--mode', choices=['graph', 'plain'], default='graph')
    ap.add_argument('--timing-steps', type=int, default=300)
    ap.add_argument('--cpu-iters', type=int, default=400)
    ap.add_argument('--no-cpu', action='store_true')
    a = ap.parse_args()
    if a.gpus != 1:
        print('bench.py measures one GPU: every box here has one GPU, and the search '
              'path shards by independent cells, not within one (--gpus %d refused)' % a.gpus,
              file=sys.stderr)
        return 2

    import torch
    from coldcase_amd import Cell
    from coldcase_amd.engine import MODE_GRAPH, MODE_PLAIN

    torch.cuda.set_device(0)
    mode = MODE_GRAPH if a.mode == 'graph' else MODE_PLAIN
    cell = Cell(Q, N_, R_, M_)
    cell.search_params(tenure=TENURE, stall=5000, kick=2)
    cell.search_init(SEED)
    cell.run(a.warmup, mode=mode)
    st0 = cell.state()
    torch.cuda.synchronize()
    cell.sync()
    t0 = time.perf_counter()
    cell.launch(a.steps, mode=mode)
    cell.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st1 = cell.state()
    moves = st1['it'] - st0['it']
    el = t1 - t0
    value = moves / el

    # roofline of the evaluator: per-launch HIP events on the cell's stream
    e0, _ = cell.search_bytes()
    cell.timing(True)
    cell.run(a.timing_steps, mode=MODE_PLAIN)
    ms, launches = cell.timing_read()
    cell.timing(False)
    e1, _ = cell.search_bytes()
    n_eval = int(launches[0])
    avg_s = ms[0] / 1e3 / max(n_eval, 1)
    bytes_per_launch = (e1 - e0) / max(n_eval, 1)
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    roof = {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
            'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic_from_profile(),
            'kernel': 'k_eval', 'avg_launch_us': round(avg_s * 1e6, 2),
            'alg_bytes_per_launch': int(bytes_per_launch),
            'commit_avg_us': round(ms[1] / 1e3 / max(int(launches[1]), 1) * 1e6, 2)}

    cpu = None
    if not a.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(cores, a.cpu_iters)

    out = {
        'metric': 'tabu moves/s on K6(9,4)',
        'value': round(value, 1),
        'unit': 'moves/s',
        'n_gpus': 1,
        'steps': moves,
        'warmup': a.warmup,
        'ms_per_step': round(el * 1e3 / max(moves, 1), 5),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'u16',
        'data': 'synthetic',
        'config': {'workload': 'K6(9,4) tabu search: q=6 n=9 R=4, M=660, random init seed %d, '
                               'tenure 10, %s steps' % (SEED, a.mode),
                   'words': Q ** N_, 'ball': 90196, 'M': M_, 'mode': a.mode,
                   'uncovered_at_window': [st0['uncovered'], st1['uncovered']],
                   'best_uncovered': st1['best_uncovered']},
        'roofline': roof,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out))
    cell.close()
    return 0


if __name__ == '__main__':
    sys.exit(main())
