#!/usr/bin/env python3
"""Benchmark: tabu moves/s on K6(9,4) on one MI355X.

Workload (a "step" is one tabu move of the device-resident search):
  q=6, n=9, R=4 (10,077,696 words, ball 90,196), M=660 codewords, uniform
  random initial code from seed 1, tabu tenure 10, stall 5000, kick 2, all
  candidates evaluated; steps replayed from a captured hipGraph.
  W warmup steps, then K timed steps (10^6 by default, the cell's stated
  run length); inputs and state resident in HBM.  The same cell then runs
  a shorter window with one launch sequence per step (plain mode).

Printed: one JSON line with the metric, the evaluator kernel's roofline
(algorithmic bytes per launch / average launch time from HIP events on the
cell's stream, against 8 TB/s HBM) and the CPU baseline (the reference
solver compiled from source, else the oracle, on the host's cores over a
bounded prefix of the same search).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode graph|plain]

Secondary workloads (one JSON line each, not the headline):
  --workload transform   K7(9,4) M=1480 random code: the uint8 axis-DP
                         distance transform (2*n*q^n algorithmic bytes) and
                         the uint16 coverage recount (2*q^n + 4*M*V bytes,
                         M*V atomics), K timed repetitions of each
  --workload batched     64 instances, (q,n,R) uniform over {5,6,7,8} x
                         {7,8,9} x {2,3,4} (numpy seed 2026, M > 65535
                         redrawn), M = ceil(1.3 q^n / |ball|), seeds
                         1000003 + 7919 i: batched (one launch pair per step
                         for all) vs plain (one cell at a time) over the same
                         prefix of steps, then the batched run to 10^5 moves
  --workload serve       Poisson arrivals (seed 7, 20 req/s) of verify and
                         10^5-move search requests over cells up to K6(10,4),
                         continuous batching with HBM-budget admission
  --workload copy        a plain 256 MiB device-to-device copy: the reference
                         kernel whose rocprofv3 FETCH_SIZE / WRITE_SIZE
                         readings scale those counters to bytes

The roofline's `traffic` (HBM bytes per launch of the dominant kernel) is
read from the counter summary under profiles/ when one is present for the
workload (a line "traffic_bytes: <bytes>"), else null.
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


def measured_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 counter summary, or None."""
    p = os.path.join(ROOT, 'profiles', name)
    if not os.path.exists(p):
        return None
    m = re.search(r'traffic_bytes:\s*(\d+)', open(p).read())
    return int(m.group(1)) if m else None


def bench_copy(a):
    """One device-to-device copy of 256 MiB, four times (torch copy_)."""
    import torch
    torch.cuda.set_device(0)
    nbytes = 256 << 20
    src = torch.ones(nbytes, dtype=torch.uint8, device='cuda')
    dst = torch.empty(nbytes, dtype=torch.uint8, device='cuda')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / 4
    print(json.dumps({'metric': 'device copy (counter calibration)', 'value': round(2 * nbytes / el / 1e9, 1),
                      'unit': 'GB/s', 'n_gpus': 1, 'steps': 4, 'warmup': 0, 'ms_per_step': round(el * 1e3, 4),
                      'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'u8',
                      'data': 'synthetic', 'config': {'workload': 'copy of %d bytes' % nbytes,
                                                      'read_bytes': nbytes, 'write_bytes': nbytes}}))
    return 0


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


def bench_transform(a):
    """K7(9,4), M=1480 uniform random codewords (numpy seed 3)."""
    import numpy as np
    import torch
    from coldcase_amd import Cell
    torch.cuda.set_device(0)
    q, n, R, M, V = 7, 9, 4, 1480, 182791
    N = q ** n
    code = np.random.default_rng(3).integers(0, q, size=(M, n)).astype(np.uint8)
    cell = Cell(q, n, R, M)
    if a.lds or a.threads:
        cell.set_transform_tile(a.lds, a.threads)
    cell.set_code(code)
    for _ in range(max(a.warmup, 1)):
        cell.distance_transform()
        cell.distance_stats(R)
        cell.recount()
    cell.sync()
    steps = a.steps
    t0 = time.perf_counter()
    for _ in range(steps):
        _, hist, rad = cell.distance_transform()
    cell.sync()
    el_t = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        cell.recount()
    cell.sync()
    el_r = (time.perf_counter() - t0) / steps
    # the transform as verify uses it (covered count and radius), then with the
    # full histogram, then the recount; HIP events around every launch
    cell.timing(True)
    for _ in range(steps):
        within, rad_s = cell.distance_stats(R)
    ms, launches = cell.timing_read()
    cell.timing(False)
    dt_us = ms[2] * 1e3 / steps            # all passes of one transform
    npass = int(launches[2]) // steps
    cell.timing(True)
    for _ in range(steps):
        cell.distance_transform()
        cell.recount()
    ms, launches = cell.timing_read()
    cell.timing(False)
    dt_hist_us = ms[2] * 1e3 / steps
    rc_us = ms[3] * 1e3 / max(int(launches[3]), 1)
    alg_t = 2 * n * N
    alg_r = 2 * N + 4 * M * V
    ach = alg_t / (dt_us * 1e-6) / 1e9
    out = {
        'metric': 'axis-DP distance transform on K7(9,4)', 'value': round(ach, 1), 'unit': 'GB/s (algorithmic)',
        'n_gpus': 1, 'steps': steps, 'warmup': a.warmup, 'ms_per_step': round(el_t * 1e3, 4),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'u8', 'data': 'synthetic',
        'config': {'workload': 'K7(9,4) M=1480 random code (numpy seed 3): distance transform + recount',
                   'words': N, 'passes_per_transform': npass,
                   'hist': [int(x) for x in hist], 'radius': rad, 'covered': within,
                   'uncovered': N - within},
        'roofline': {'bound': 'hbm', 'achieved': round(ach, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(ach / HBM_PEAK_GBS, 4),
                     'traffic': measured_traffic('pmc_k_dt_seed_k_dt_col_K7_9_4.md'), 'kernel': 'k_dt_seed+k_dt_col',
                     'avg_transform_us': round(dt_us, 2), 'alg_bytes_per_transform': alg_t,
                     'with_histogram_us': round(dt_hist_us, 2)},
        'recount': {'avg_us': round(rc_us, 2), 'wall_ms': round(el_r * 1e3, 4), 'alg_bytes': alg_r,
                    'achieved_GBs': round(alg_r / (rc_us * 1e-6) / 1e9, 1),
                    'atomics_per_s': round(M * V / (rc_us * 1e-6), 1)},
    }
    print(json.dumps(out))
    cell.close()
    return 0


def c4_specs(seed=2026, count=64):
    """The batched config's instances and how many draws were redrawn."""
    import math
    import numpy as np

    def ball(q, n, R):
        return sum(math.comb(n, i) * (q - 1) ** i for i in range(R + 1))

    rng = np.random.default_rng(seed)
    specs, redrawn = [], []
    while len(specs) < count:
        q = int(rng.choice([5, 6, 7, 8]))
        n = int(rng.choice([7, 8, 9]))
        R = int(rng.choice([2, 3, 4]))
        M = math.ceil(1.3 * q ** n / ball(q, n, R))
        if M > 65535:
            redrawn.append((q, n, R, M))
            continue
        specs.append((q, n, R, M))
    return specs, redrawn


def bench_batched(a):
    import torch
    from coldcase_amd import Batch, Cell, Fleet
    from coldcase_amd.autotune import tune_fleet
    from coldcase_amd.engine import MODE_GRAPH
    torch.cuda.set_device(0)
    specs, redrawn = c4_specs()
    moves, prefix = a.steps, a.prefix
    seeds = [1 * 1000003 + i * 7919 for i in range(len(specs))]

    def make(i, s):
        q, n, R, M = s
        c = Cell(q, n, R, M)
        c.search_params(tenure=10, stall=5000, kick=2)
        c.search_init(seeds[i])
        return c

    # plain: one cell at a time, its own launches (hipGraph steps), default shape
    plain_s = 0.0
    costs = []
    for i, s in enumerate(specs):
        with make(i, s) as c:
            c.run(32, mode=MODE_GRAPH)
            c.sync()
            t0 = time.perf_counter()
            c.run(prefix, mode=MODE_GRAPH)
            dt = time.perf_counter() - t0
            plain_s += dt
            costs.append(dt)
        print('plain %d/%d' % (i + 1, len(specs)), file=sys.stderr, flush=True)
    # batched: the cells grouped by their plain step time into concurrent
    # batches (one evaluator and one commit launch per step and group), each
    # group's batch scale autotuned
    cells = [make(i, s) for i, s in enumerate(specs)]
    groups, factors, gidx = tune_fleet(cells, costs, ngroups=a.groups)

    def restart():
        for i, c in enumerate(cells):
            c.search_init(seeds[i])

    # for reference: the same launch shapes as one batch of all cells
    restart()
    with Batch(cells) as b:
        b.run(32)
        t0 = time.perf_counter()
        b.run(prefix)
        one_batch_s = time.perf_counter() - t0
    restart()
    with Fleet(groups) as f:
        f.run(32)
        t0 = time.perf_counter()
        f.run(prefix)
        batched_prefix_s = time.perf_counter() - t0
        print('batched prefix done', file=sys.stderr, flush=True)
        done = 32 + prefix
        t0 = time.perf_counter()
        while done < moves:
            k = min(5000, moves - done)
            f.run(k)
            done += k
            print('batched %d/%d' % (done, moves), file=sys.stderr, flush=True)
        rest_s = time.perf_counter() - t0
    total_batched = batched_prefix_s + rest_s
    st = [c.state() for c in cells]
    # the timed region holds moves - 32 steps of every cell (32 warmed up the graphs)
    solves_batched = len(cells) / (total_batched * moves / (moves - 32)) if moves > 32 else 0.0
    plain_rate = len(cells) / (plain_s / prefix * moves)             # projected from the prefix
    batched_prefix_rate = len(cells) / (batched_prefix_s / prefix * moves)
    out = {
        'metric': 'batched vs plain solves/s (64 instances, 10^5 moves each)', 'value': round(solves_batched, 4),
        'unit': 'solves/s', 'n_gpus': 1, 'steps': moves, 'warmup': 32, 'ms_per_step': round(total_batched * 1e3 / max(moves - 32, 1), 4),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'u16', 'data': 'synthetic',
        'config': {'workload': '64 instances (q,n,R) from {5,6,7,8}x{7,8,9}x{2,3,4}, numpy seed 2026, '
                               'M = ceil(1.3 q^n/|ball|), seeds 1000003 + 7919 i, tenure 10',
                   'redrawn': redrawn, 'prefix_steps': prefix,
                   'groups': [len(g) for g in gidx], 'group_batch_scale': factors},
        'batched_full_solves_per_s': round(solves_batched, 4),
        'plain_solves_per_s_projected': round(plain_rate, 5),
        'ratio_full_vs_plain': round(solves_batched / plain_rate, 3),
        'prefix': {'plain_s': round(plain_s, 3), 'batched_s': round(batched_prefix_s, 3),
                   'one_batch_s': round(one_batch_s, 3),
                   'batched_solves_per_s_projected': round(batched_prefix_rate, 5),
                   'ratio': round(plain_s / batched_prefix_s, 3),
                   'ratio_one_batch': round(plain_s / one_batch_s, 3)},
        'covered': sum(1 for x in st if x['done']),
    }
    print(json.dumps(out))
    for c in cells:
        c.close()
    return 0


def bench_serve(a):
    import numpy as np
    import torch
    from coldcase_amd.serve import Server, poisson_workload, summarize
    torch.cuda.set_device(0)
    shapes = [(5, 7, 2, 400), (6, 6, 3, 41), (6, 8, 4, 169), (6, 9, 4, 660), (7, 9, 4, 1480), (6, 10, 4, 2784)]
    codes = {}
    data = os.path.join(ROOT, 'tests', 'golden', 'data')
    for fn, key in [('K6_10_4_M2784.txt', (6, 10, 4, 2784)), ('K7_9_4_M1480.txt', (7, 9, 4, 1480)),
                    ('K6_9_4_M660.txt', (6, 9, 4, 660)), ('K6_8_4_M169.txt', (6, 8, 4, 169)),
                    ('K6_6_3_M41.txt', (6, 6, 3, 41))]:
        p = os.path.join(data, fn)
        if os.path.exists(p):
            from coldcase_amd.verify_cov import parse_code
            code, _ = parse_code(p, key[0], key[1])
            codes[key] = np.array(code, dtype=np.uint8)
    reqs = poisson_workload(7, a.requests, 20.0, shapes, verify_frac=0.3, moves=a.steps, codes=codes)
    srv = Server(hbm_budget=int(a.budget_gb * (1 << 30)), max_cells=64, tick_steps=a.tick,
                 lane_cap=a.lane_cap or None)
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=True)
    el = time.perf_counter() - t0
    s = summarize(res)
    # plain: the same requests one at a time, each alone on the GPU (graph
    # steps, default launch shapes); a search's service time is its moves
    # times the shape's step time over a 256-step prefix, a verify's is timed
    from coldcase_amd import Cell
    from coldcase_amd.engine import MODE_GRAPH
    from coldcase_amd.serve import run_verify
    step_s, verify_s = {}, {}
    plain_total = 0.0
    for r in reqs:
        key = (r.q, r.n, r.R, r.M)
        if r.kind == 'search':
            if key not in step_s:
                with Cell(*key) as c:
                    c.search_params(tenure=r.tenure, stall=r.stall, kick=r.kick)
                    c.search_init(r.seed)
                    c.run(32, mode=MODE_GRAPH)
                    c.sync()
                    t1 = time.perf_counter()
                    c.run(256, mode=MODE_GRAPH)
                    step_s[key] = (time.perf_counter() - t1) / 256
            plain_total += r.moves * step_s[key]
        else:
            if key not in verify_s:
                run_verify(r)
                t1 = time.perf_counter()
                run_verify(r)
                verify_s[key] = time.perf_counter() - t1
            plain_total += verify_s[key]
    span = max(x.finished for x in res) - min(x.arrival for x in res)
    out = {'metric': 'p50/p99 solve latency (served)', 'value': round(s['p50_s'], 4), 'unit': 's',
           'n_gpus': 1, 'steps': a.steps, 'warmup': 0, 'ms_per_step': None, 'higher_is_better': False,
           'scaling': 'weak', 'vs_baseline': None, 'dtype': 'u16', 'data': 'synthetic',
           'config': {'workload': 'Poisson arrivals 20 req/s (seed 7), %d requests, 30%% verify / 70%% search '
                                  'bursts of %d moves over K5(7,2)..K6(10,4); record codes as verify inputs'
                                  % (a.requests, a.steps),
                      'hbm_budget_gb': a.budget_gb, 'tick_steps': a.tick, 'lane_cap': a.lane_cap or None},
           'latency': s, 'makespan_s': round(el, 3),
           'served_rps': round(len(res) / span, 4), 'plain_rps_projected': round(len(res) / plain_total, 4),
           'ratio_served_vs_plain': round(plain_total / span, 3),
           'plain_step_us': {'%d,%d,%d,%d' % k: round(v * 1e6, 1) for k, v in step_s.items()},
           'peak_cells': srv.peak_cells,
           'peak_gb': round(srv.peak_bytes / (1 << 30), 3), 'ticks': srv.ticks,
           'verify_record_uncovered': sorted({r.uncovered for q_, r in zip(reqs, res)
                                              if q_.kind == 'verify' and (q_.q, q_.n, q_.R, q_.M) in codes})}
    print(json.dumps(out))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=1000000)
    ap.add_argument('--warmup', type=int, default=1000)
    ap.add_argument('--mode', choices=['graph', 'plain'], default='graph')
    ap.add_argument('--timing-steps', type=int, default=300)
    ap.add_argument('--plain-steps', type=int, default=20000)
    ap.add_argument('--cpu-iters', type=int, default=4000)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--workload', choices=['search', 'transform', 'batched', 'serve', 'copy'], default='search')
    ap.add_argument('--lds', type=int, default=0, help='transform tile LDS bytes (0 = default)')
    ap.add_argument('--threads', type=int, default=0, help='transform workgroup threads (0 = default)')
    ap.add_argument('--prefix', type=int, default=1000, help='batched: steps timed in both modes')
    ap.add_argument('--groups', type=int, default=4, help='batched: concurrent batches, cells grouped by step cost')
    ap.add_argument('--requests', type=int, default=60, help='serve: requests in the trace')
    ap.add_argument('--tick', type=int, default=256, help='serve: steps per batched tick at most')
    ap.add_argument('--budget-gb', type=float, default=64.0, help='serve: HBM budget')
    ap.add_argument('--lane-cap', type=int, default=8, help='serve: searches a lane advances at once (0: all)')
    a = ap.parse_args()
    if a.gpus != 1:
        print('bench.py measures one GPU: every box here has one GPU, and the search '
              'path shards by independent cells, not within one (--gpus %d refused)' % a.gpus,
              file=sys.stderr)
        return 2
    if a.workload == 'copy':
        return bench_copy(a)
    if a.workload == 'transform':
        if a.steps == ap.get_default('steps'):
            a.steps = 20
        if a.warmup == ap.get_default('warmup'):
            a.warmup = 2
        return bench_transform(a)
    if a.workload == 'batched':
        if a.steps == ap.get_default('steps'):
            a.steps = 100000
        return bench_batched(a)
    if a.workload == 'serve':
        if a.steps == ap.get_default('steps'):
            a.steps = 100000
        return bench_serve(a)

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

    # the same cell, one launch sequence per step
    plain = None
    if a.plain_steps > 0 and mode == MODE_GRAPH and not st1['done']:
        p0 = cell.state()['it']
        t0 = time.perf_counter()
        cell.launch(a.plain_steps, mode=MODE_PLAIN)
        cell.sync()
        el_p = time.perf_counter() - t0
        pm = cell.state()['it'] - p0
        plain = {'value': round(pm / el_p, 1), 'steps': pm, 'ms_per_step': round(el_p * 1e3 / max(pm, 1), 5)}

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
            'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': measured_traffic('pmc_k_eval_K6_9_4.md'),
            'kernel': 'k_eval', 'avg_launch_us': round(avg_s * 1e6, 2),
            'alg_bytes_per_launch': int(bytes_per_launch),
            'commit_avg_us': round(ms[1] / 1e3 / max(int(launches[1]), 1) * 1e6, 2)}

    cpu = None
    if not a.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(cores, a.cpu_iters)

    out = {
        'metric': 'tabu moves/s on K6(9,4); batched vs plain solves/s; p50/p99 solve latency',
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
        'plain_mode': plain,
        'roofline': roof,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out))
    cell.close()
    return 0


if __name__ == '__main__':
    sys.exit(main())
