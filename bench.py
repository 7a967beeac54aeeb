#!/usr/bin/env python3
"""Benchmark: tabu moves/s on K6(9,4); batched vs plain solves/s; p50/p99
solve latency -- on one MI355X.

Headline workload (a "step" is one block of --moves-per-step tabu moves,
50,000 by default, so the default 20 steps are the cell's stated 10^6
moves):
  q=6, n=9, R=4 (10,077,696 words, ball 90,196), M=660 codewords, uniform
  random initial code from seed 1, tabu tenure 10, stall 5000, kick 2, all
  candidates evaluated; moves replayed from captured hipGraphs of 40 steps
  (50,000 moves = 1,250 graph launches, no plain remainder).  W warmup
  steps, then K timed steps; inputs and state resident in HBM.  The same
  cell then runs one step of moves with one launch sequence per move (plain
  mode), and 300 plain moves with HIP events around every launch on the
  cell's stream for the per-kernel durations.

One JSON line carries:
  value         moves/s over the K timed steps
  roofline      the kernel with the largest share of the step (event-timed):
                the reference's algorithmic bytes per move, B_move =
                M n + 2 sum_cw (S_f + E_cw) + 8 S (every candidate codeword's
                distance-R sphere walked in full; counted per move by the
                device from its candidate codewords and entering path) over
                that kernel's average duration, against 8 TB/s; `per_step`
                gives B_move over the whole step; `traffic` is the HBM bytes
                per launch of that kernel from the committed counter summary
  cpu_baseline  the reference solver compiled from source
                (oracle/_ref/covsearch2e, else the oracle's port) on the
                host's CPU share (16 threads), over the first moves of the
                same seeded search; `single_thread` the same at one thread;
                each with the GPU's rate over the identical moves
  batched       the 64-instance configuration: tuned concurrent batches vs
                one cell per launch over a prefix, the batched run to 10^5
                moves, and a CPU portfolio projection
  serve         the Poisson serving configuration: p50/p99 latency, served
                vs one-request-at-a-time throughput, and a serial CPU
                projection
  transform     K7(9,4) M=1480: the uint8 distance transform against 8 TB/s
                on its algorithmic bytes (2 n q^n) and on its measured
                traffic, and the coverage recount

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--parts search,cpu,batched,serve,transform]

Secondary entry points (one JSON line each): --workload transform | batched |
serve | copy (a plain 256 MiB device copy: the reference kernel whose
rocprofv3 FETCH_SIZE / WRITE_SIZE readings scale those counters to bytes).
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 1
KERNELS = {0: 'k_eval', 1: 'k_commit', 5: 'k_post'}   # k_commit selects the move first


def measured_traffic(name, kernel=None):
    """HBM bytes per launch from a committed rocprofv3 counter summary, or None:
    a line `traffic_bytes[<kernel>]: <bytes>`, or `traffic_bytes: <bytes>`."""
    p = os.path.join(ROOT, 'profiles', name)
    if not os.path.exists(p):
        return None
    text = open(p).read()
    m = None
    if kernel:
        m = re.search(r'traffic_bytes\[%s\]:\s*(\d+)' % re.escape(kernel), text)
    if m is None:
        m = re.search(r'traffic_bytes:\s*(\d+)', text)
    return int(m.group(1)) if m else None


def cpu_share():
    """Host threads of one GPU's share of the box (16), or fewer if the host has fewer."""
    return min(16, os.cpu_count() or 1)


def ref_exe():
    ref = os.path.join(ROOT, 'oracle', '_ref', 'covsearch2e')
    ora = os.path.join(ROOT, 'oracle', 'cso')
    if os.path.exists(ref):
        return ref, 'reference'
    if os.path.exists(ora):
        return ora, 'port'
    return None, None


def run_ref(q, n, R, M, seed, iters, threads, extra=()):
    """One run of the CPU solver; returns (iters, wall s, cpu s) from its RESULT line."""
    exe, _ = ref_exe()
    cmd = [exe, '-q', str(q), '-n', str(n), '-R', str(R), '-M', str(M), '-s', str(seed), '--random-init',
           '--tabu', '10', '--iters', str(iters), '--preset', 'p5b', '--threads', str(threads), '-t', '1e9',
           '--quiet'] + list(extra)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    p = subprocess.run(cmd, capture_output=True, text=True, env=env)
    m = re.search(r'iters=(\d+) kicks=\d+ time=([\d.]+) rate=\d+ cpu=([\d.]+)', p.stdout)
    if not m:
        raise RuntimeError('CPU solver gave no RESULT: %s\n%s' % (' '.join(cmd), p.stderr[-400:]))
    return int(m.group(1)), float(m.group(2)), float(m.group(3))


# ------------------------------------------------------------------ search
def bench_search(a):
    import torch
    from coldcase_amd import Cell
    from coldcase_amd.engine import GRAPH_STEPS, MODE_GRAPH, MODE_PLAIN
    from coldcase_amd.workloads import C2, TENURE
    q, n, R, M = C2
    torch.cuda.set_device(0)
    mode = MODE_GRAPH if a.mode == 'graph' else MODE_PLAIN
    mps = a.moves_per_step
    cell = Cell(q, n, R, M)
    cell.search_params(tenure=TENURE, stall=5000, kick=2)
    cell.search_init(SEED)

    def go(moves, md):
        # one launch of all the moves, or pieces of --chunk moves with a wait
        # between (counter collection cannot hold tens of thousands of
        # dispatches in flight)
        k = a.chunk if a.chunk > 0 else moves
        while moves > 0:
            cell.launch(min(k, moves), mode=md)
            cell.sync()
            moves -= k

    go(a.warmup * mps, mode)
    st0 = cell.state()
    b0 = cell.search_bytes()
    torch.cuda.synchronize()
    cell.sync()
    t0 = time.perf_counter()
    go(a.steps * mps, mode)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st1 = cell.check_fault()
    b1 = cell.search_bytes()
    moves = st1['it'] - st0['it']
    el = t1 - t0
    value = moves / el
    move_bytes = (b1[2] - b0[2]) / max(b1[3] - b0[3], 1)

    plain = None
    if a.plain_steps > 0 and mode == MODE_GRAPH and not st1['done']:
        p0 = cell.state()['it']
        t0 = time.perf_counter()
        go(a.plain_steps * mps, MODE_PLAIN)
        el_p = time.perf_counter() - t0
        pm = cell.check_fault()['it'] - p0
        plain = {'value': round(pm / el_p, 1), 'moves': pm, 'us_per_move': round(el_p * 1e6 / max(pm, 1), 3)}

    # per-kernel durations: HIP events around every launch on the cell's stream
    e0 = cell.search_bytes()
    cell.timing(True)
    cell.run(a.timing_steps, mode=MODE_PLAIN)
    ms, launches = cell.timing_read()
    cell.timing(False)
    e1 = cell.search_bytes()
    avg = {KERNELS[k]: ms[k] * 1e3 / max(int(launches[k]), 1) for k in KERNELS}
    step_us = sum(avg.values())
    bm = (e1[2] - e0[2]) / max(e1[3] - e0[3], 1)
    own_eval = (e1[0] - e0[0]) / max(int(launches[0]), 1)
    own_commit = (e1[1] - e0[1]) / max(int(launches[1]), 1)
    dom = max(avg, key=avg.get)
    ach = bm / (avg[dom] * 1e-6) / 1e9
    ach_step = bm / (step_us * 1e-6) / 1e9
    roof = {'bound': 'hbm', 'achieved': round(ach, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
            'frac': round(ach / HBM_PEAK_GBS, 5),
            'traffic': measured_traffic('step_kernels_K6_9_4.md', dom),
            'kernel': dom, 'avg_launch_us': round(avg[dom], 3),
            'alg_bytes_per_launch': int(bm),
            'alg_bytes_note': 'B_move = M n + 2 sum_cw (S_f + E_cw) + 8 S per move, the reference move with every '
                              'candidate sphere walked; one launch of each step kernel serves one move',
            'per_step': {'us': round(step_us, 3), 'achieved': round(ach_step, 1),
                         'frac': round(ach_step / HBM_PEAK_GBS, 5)},
            'kernels_us': {k: round(v, 3) for k, v in avg.items()},
            'kernels_share': {k: round(v / step_us, 4) for k, v in avg.items()},
            'own_bytes_per_launch': {'k_eval': int(own_eval), 'k_commit': int(own_commit)},
            'timed_moves': a.timing_steps}
    cell.close()
    out = {
        'metric': 'tabu moves/s on K6(9,4); batched vs plain solves/s; p50/p99 solve latency',
        'value': round(value, 1),
        'unit': 'moves/s',
        'n_gpus': 1,
        'steps': a.steps,
        'warmup': a.warmup,
        'ms_per_step': round(el * 1e3 / max(a.steps, 1), 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'u16',
        'data': 'synthetic',
        'config': {'workload': 'K6(9,4) tabu search: q=6 n=9 R=4, M=660, random init seed %d, tenure 10, '
                               '%d steps of %d moves (%s)' % (SEED, a.steps, mps, a.mode),
                   'words': q ** n, 'ball': 90196, 'M': M, 'mode': a.mode, 'moves_per_step': mps,
                   'moves': moves, 'us_per_move': round(el * 1e6 / max(moves, 1), 3),
                   'graph_launches': (a.steps * mps) // GRAPH_STEPS if mode == MODE_GRAPH else 0,
                   'plain_moves': (a.steps * mps) % GRAPH_STEPS if mode == MODE_GRAPH else a.steps * mps,
                   'window_moves': [st0['it'], st1['it']],
                   'uncovered_at_window': [st0['uncovered'], st1['uncovered']],
                   'best_uncovered': st1['best_uncovered'], 'kicks_in_window': st1['kicks'] - st0['kicks'],
                   'move_bytes_in_window': int(move_bytes)},
        'plain_mode': plain,
        'roofline': roof,
    }
    return out


# -------------------------------------------------------------- CPU baseline
def gpu_window_rate(moves):
    """GPU moves/s over moves 1..`moves` of the headline search (graphs
    captured beforehand on the same cell, so the window is the same moves)."""
    from coldcase_amd import Cell
    from coldcase_amd.engine import GRAPH_STEPS, MODE_GRAPH
    from coldcase_amd.workloads import C2, TENURE
    with Cell(*C2) as c:
        c.search_params(tenure=TENURE, stall=5000, kick=2)
        c.search_init(SEED)
        c.run(GRAPH_STEPS, mode=MODE_GRAPH)
        c.search_init(SEED)
        c.sync()
        t0 = time.perf_counter()
        c.run(moves, mode=MODE_GRAPH)
        el = time.perf_counter() - t0
        return moves / el, c.state()['best_uncovered']


def bench_cpu(a):
    exe, kind = ref_exe()
    if exe is None:
        return None
    from coldcase_amd.workloads import C2
    q, n, R, M = C2
    cores = a.cpu_threads or cpu_share()
    it_m, wall_m, cpu_m = run_ref(q, n, R, M, SEED, a.cpu_all_moves, cores)
    it_1, wall_1, cpu_1 = run_ref(q, n, R, M, SEED, a.cpu_1t_moves, 1)
    g_m, _ = gpu_window_rate(it_m)
    g_1, _ = gpu_window_rate(it_1)
    what = 'K6(9,4) M=660 seed 1 random init, tenure 10, moves 1..%d (%.2f s wall, %.2f s CPU), ' \
           '--preset p5b --threads %d'
    return {'value': round(it_m / wall_m, 1), 'unit': 'moves/s', 'cores': cores, 'kind': kind,
            'sample': what % (it_m, wall_m, cpu_m, cores),
            'gpu_same_moves': round(g_m, 1), 'gpu_over_cpu': round(g_m / (it_m / wall_m), 2),
            'single_thread': {'value': round(it_1 / wall_1, 1), 'unit': 'moves/s', 'cores': 1, 'kind': kind,
                              'sample': what % (it_1, wall_1, cpu_1, 1), 'gpu_same_moves': round(g_1, 1),
                              'gpu_over_cpu': round(g_1 / (it_1 / wall_1), 2)}}


# --------------------------------------------------------------- transform
def bench_transform(a, reps=20, warmup=2):
    import numpy as np
    import torch
    from coldcase_amd import Cell
    from coldcase_amd.workloads import C3, ball
    torch.cuda.set_device(0)
    q, n, R, M = C3
    V = ball(q, n, R)
    N = q ** n
    code = np.random.default_rng(3).integers(0, q, size=(M, n)).astype(np.uint8)
    cell = Cell(q, n, R, M)
    if a.lds or a.threads:
        cell.set_transform_tile(a.lds, a.threads)
    cell.set_code(code)
    for _ in range(warmup):
        cell.distance_transform()
        cell.distance_stats(R)
        cell.recount()
    cell.sync()
    cell.timing(True)
    for _ in range(reps):
        within, rad_s = cell.distance_stats(R)
    ms, launches = cell.timing_read()
    cell.timing(False)
    dt_us = ms[2] * 1e3 / reps
    npass = int(launches[2]) // reps
    cell.timing(True)
    for _ in range(reps):
        _, hist, rad = cell.distance_transform()
        cell.recount()
    ms, launches = cell.timing_read()
    cell.timing(False)
    dt_hist_us = ms[2] * 1e3 / reps
    rc_us = ms[3] * 1e3 / max(int(launches[3]), 1)
    # the count transform of the uncovered set (the gain map), kernel time of
    # all its passes per transform
    from coldcase_amd.engine import CT_UNCOVERED
    cell.count_transform(CT_UNCOVERED)
    cell.timing(True)
    for _ in range(max(reps // 4, 2)):
        cell.count_transform(CT_UNCOVERED)
    ms, launches = cell.timing_read()
    cell.timing(False)
    nct = max(reps // 4, 2)
    ct_us = ms[6] * 1e3 / nct
    ct_passes = int(launches[6]) // nct
    cell.close()
    alg_t = 2 * n * N
    alg_r = 2 * N + 4 * M * V
    ach = alg_t / (dt_us * 1e-6) / 1e9
    traffic = measured_traffic('transforms_K7_9_4.md')
    ct_traffic = measured_traffic('transforms_K7_9_4.md', 'k_ct')
    out = {'workload': 'K7(9,4) M=1480 random code (numpy seed 3): distance transform + recount',
           'words': N, 'passes_per_transform': npass, 'avg_transform_us': round(dt_us, 2),
           'with_histogram_us': round(dt_hist_us, 2), 'radius': rad, 'covered': int(within),
           'hist': [int(x) for x in hist],
           'roofline': {'bound': 'hbm', 'achieved': round(ach, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                        'frac': round(ach / HBM_PEAK_GBS, 4), 'alg_bytes_per_transform': alg_t,
                        'traffic': traffic, 'kernel': 'k_dt_seed+k_dt_tcol',
                        # what any transform moves at least: the distances written once and the
                        # codeword bitmap read once
                        'min_bytes_per_transform': N + N // 8},
           'recount': {'avg_us': round(rc_us, 2), 'alg_bytes': alg_r,
                       'achieved_GBs': round(alg_r / (rc_us * 1e-6) / 1e9, 1),
                       'atomics_per_s': round(M * V / (rc_us * 1e-6), 1)}}
    # count transform: the reference's formulation moves (R+1) int32 layers
    # in and out once per axis, 8 n (R+1) bytes per word; the passes here
    # read 2 B of counts, write and re-read the layers once per pass and
    # write the 4 B map
    alg_c = 8 * n * (R + 1) * N
    lay = 4 * (R + 1)
    own_c = N * (2 + lay + 2 * lay * (ct_passes - 2) + lay + 4) if ct_passes >= 2 else N * (2 + 4)
    out['count_transform'] = {'source': 'uncovered words (the gain map)', 'avg_us': round(ct_us, 2),
                              'passes': ct_passes, 'alg_bytes': alg_c,
                              'achieved_GBs': round(alg_c / (ct_us * 1e-6) / 1e9, 1),
                              'pass_bytes': own_c,
                              'pass_bytes_GBs': round(own_c / (ct_us * 1e-6) / 1e9, 1),
                              'pass_bytes_frac': round(own_c / (ct_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                              'traffic': ct_traffic}
    if ct_traffic:
        out['count_transform']['traffic_frac'] = round(ct_traffic / (ct_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
    if traffic:
        tr = traffic / (dt_us * 1e-6) / 1e9
        out['roofline']['traffic_achieved'] = round(tr, 1)
        out['roofline']['traffic_frac'] = round(tr / HBM_PEAK_GBS, 4)
    mb = (N + N // 8) / (dt_us * 1e-6) / 1e9
    out['roofline']['min_bytes_frac'] = round(mb / HBM_PEAK_GBS, 4)
    return out


# ----------------------------------------------------------------- batched
def cpu_portfolio(specs, seeds, moves, procs, total_moves):
    """The batched configuration on the CPU: every instance a single-thread
    reference process, `procs` at once (the portfolio discipline), timed over
    its first `moves` moves; each instance's 10^5-move time projected from
    that rate, then packed longest-first onto `procs` processes."""
    def one(i):
        q, n, R, M = specs[i]
        m = moves
        while True:
            it, wall, cpu = run_ref(q, n, R, M, seeds[i], m, 1)
            if wall >= 0.2 or m >= 100 * moves:
                return it / max(wall, 0.01)
            m *= 10

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=procs) as ex:
        rates = list(ex.map(one, range(len(specs))))
    el = time.perf_counter() - t0
    jobs = sorted((total_moves / r for r in rates), reverse=True)
    load = [0.0] * procs
    for j in jobs:
        k = load.index(min(load))
        load[k] += j
    makespan = max(load)
    return {'value': round(len(specs) / makespan, 6), 'unit': 'solves/s', 'cores': procs,
            'kind': ref_exe()[1], 'projected': True,
            'sample': 'each instance %d moves or more (at least 0.2 s) single-threaded, %d processes at once '
                      '(%.1f s); 10^5-move time '
                      'per instance from its rate, packed longest first' % (moves, procs, el),
            'makespan_s': round(makespan, 1), 'slowest_instance_moves_per_s': round(min(rates), 2)}


def bench_batched(a, with_cpu=True):
    import torch
    from coldcase_amd import Batch, Cell, Fleet
    from coldcase_amd.autotune import tune_fleet
    from coldcase_amd.engine import GRAPH_STEPS, MODE_GRAPH
    from coldcase_amd.workloads import TENURE, c4_seed, c4_specs
    torch.cuda.set_device(0)
    specs, redrawn = c4_specs()
    moves, prefix = a.batched_moves, a.prefix
    seeds = [c4_seed(i) for i in range(len(specs))]

    def make(i, s):
        q, n, R, M = s
        c = Cell(q, n, R, M)
        c.search_params(tenure=TENURE, stall=5000, kick=2)
        c.search_init(seeds[i])
        return c

    # plain: one cell at a time, its own launches (hipGraph steps), default shape
    plain_s = 0.0
    costs = []
    for i, s in enumerate(specs):
        with make(i, s) as c:
            c.run(GRAPH_STEPS, mode=MODE_GRAPH)
            c.sync()
            t0 = time.perf_counter()
            c.run(prefix, mode=MODE_GRAPH)
            dt = time.perf_counter() - t0
            plain_s += dt
            costs.append(dt)
    print('batched: plain prefix done', file=sys.stderr, flush=True)
    cells = [make(i, s) for i, s in enumerate(specs)]
    groups, factors, gidx = tune_fleet(cells, costs, ngroups=a.groups)
    print('batched: fleet tuned', file=sys.stderr, flush=True)

    def restart():
        for i, c in enumerate(cells):
            c.search_init(seeds[i])

    restart()
    with Batch(cells) as b:
        b.run(GRAPH_STEPS)
        t0 = time.perf_counter()
        b.run(prefix)
        one_batch_s = time.perf_counter() - t0
    restart()
    with Fleet(groups) as f:
        f.run(GRAPH_STEPS)
        t0 = time.perf_counter()
        f.run(prefix)
        batched_prefix_s = time.perf_counter() - t0
        done = GRAPH_STEPS + prefix
        t0 = time.perf_counter()
        while done < moves:
            k = min(20000, moves - done)
            f.run(k)
            done += k
            print('batched %d/%d' % (done, moves), file=sys.stderr, flush=True)
        rest_s = time.perf_counter() - t0
    total = batched_prefix_s + rest_s
    st = [c.check_fault() for c in cells]
    for c in cells:
        c.close()
    timed = moves - GRAPH_STEPS
    solves = len(cells) / (total * moves / timed)      # the warm-up steps charged at the measured rate
    plain_rate = len(cells) / (plain_s / prefix * moves)
    out = {'metric': 'solves/s, 64 instances x 10^5 moves', 'value': round(solves, 4), 'unit': 'solves/s',
           'workload': '64 instances (q,n,R) from {5,6,7,8}x{7,8,9}x{2,3,4}, numpy seed 2026, '
                       'M = ceil(1.3 q^n/|ball|), seeds 1000003 + 7919 i, tenure 10, 10^5 moves each',
           'redrawn': redrawn, 'groups': [len(g) for g in gidx], 'group_batch_scale': factors,
           'plain_solves_per_s_projected': round(plain_rate, 5),
           'ratio_vs_plain': round(solves / plain_rate, 3),
           'prefix': {'moves': prefix, 'plain_s': round(plain_s, 3), 'fleet_s': round(batched_prefix_s, 3),
                      'one_batch_s': round(one_batch_s, 3), 'ratio': round(plain_s / batched_prefix_s, 3),
                      'ratio_one_batch': round(plain_s / one_batch_s, 3)},
           'moves_per_instance': [x['it'] for x in st][:1] + [min(x['it'] for x in st), max(x['it'] for x in st)],
           'covered': sum(1 for x in st if x['done'])}
    if with_cpu and ref_exe()[0]:
        cpu = cpu_portfolio(specs, seeds, a.cpu_c4_moves, a.cpu_threads or cpu_share(), moves)
        cpu['gpu_over_cpu'] = round(solves / cpu['value'], 1)
        out['cpu_baseline'] = cpu
    return out


# ------------------------------------------------------------------- serve
def record_codes():
    import numpy as np
    from coldcase_amd.verify_cov import parse_code
    codes = {}
    data = os.path.join(ROOT, 'tests', 'golden', 'data')
    for fn, key in [('K6_10_4_M2784.txt', (6, 10, 4, 2784)), ('K7_9_4_M1480.txt', (7, 9, 4, 1480)),
                    ('K6_9_4_M660.txt', (6, 9, 4, 660)), ('K6_8_4_M169.txt', (6, 8, 4, 169)),
                    ('K6_6_3_M41.txt', (6, 6, 3, 41))]:
        p = os.path.join(data, fn)
        if os.path.exists(p):
            code, _ = parse_code(p, key[0], key[1])
            codes[key] = np.array(code, dtype=np.uint8)
    return codes


def cpu_serial(reqs, codes, threads, sample_moves):
    """The serving trace on the CPU, one request at a time in arrival order:
    a search's service time is its moves at the CPU solver's rate for that
    shape (measured over its first `sample_moves` moves at `threads`
    threads), a verify's is the numpy axis dilation of the oracle
    (cov/verify_independent.py:128-141) on one core, timed.  Latencies follow
    from a FIFO queue over the arrival times."""
    import numpy as np
    from oracle.verify_oracle import dilate_counts
    rate, vt = {}, {}
    for r in reqs:
        key = (r.q, r.n, r.R, r.M)
        if r.kind == 'search' and key not in rate:
            # grow the sample of a fast shape until it runs a measurable time
            moves = sample_moves
            while True:
                it, wall, _ = run_ref(*key, r.seed, moves, threads)
                if wall >= 0.2 or moves >= 100 * sample_moves:
                    break
                moves *= 10
            rate[key] = it / max(wall, 0.01)
        if r.kind == 'verify' and key not in vt:
            t0 = time.perf_counter()
            dilate_counts(np.asarray(r.code), r.n, r.q, r.R)
            vt[key] = time.perf_counter() - t0
    free = 0.0
    lat = []
    last = 0.0
    for r in sorted(reqs, key=lambda x: x.arrival):
        key = (r.q, r.n, r.R, r.M)
        svc = r.moves / rate[key] if r.kind == 'search' else vt[key]
        start = max(r.arrival, free)
        free = start + svc
        lat.append(free - r.arrival)
        last = free
    span = last - min(r.arrival for r in reqs)
    return {'p50_s': round(float(np.percentile(lat, 50)), 2), 'p99_s': round(float(np.percentile(lat, 99)), 2),
            'rps': round(len(reqs) / span, 6), 'cores': threads, 'kind': ref_exe()[1], 'projected': True,
            'search_moves_per_s': {'%d,%d,%d,%d' % k: round(v, 1) for k, v in rate.items()},
            'verify_s': {'%d,%d,%d,%d' % k: round(v, 3) for k, v in vt.items()},
            'sample': 'searches at the CPU solver rate over each shape\'s first %d moves or more (at least 0.2 s; '
                      '%d threads); verifies '
                      'by numpy axis dilation on one core; FIFO over the same arrivals' % (sample_moves, threads)}


def bench_serve(a, with_cpu=True):
    import torch
    from coldcase_amd import Cell
    from coldcase_amd.engine import GRAPH_STEPS, MODE_GRAPH
    from coldcase_amd.serve import Server, run_verify, summarize
    from coldcase_amd.workloads import c5_trace
    torch.cuda.set_device(0)
    codes = record_codes()
    reqs = c5_trace(a.requests, a.serve_moves, codes=codes)
    srv = Server(hbm_budget=int(a.budget_gb * (1 << 30)), max_cells=64, tick_steps=a.tick,
                 lane_cap=a.lane_cap or None)
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=True)
    el = time.perf_counter() - t0
    s = summarize(res)
    print('serve: trace done', file=sys.stderr, flush=True)
    # plain: the same requests one at a time, each alone on the GPU (graph
    # steps, default launch shapes); a search's service time is its moves
    # times the shape's step time over a 256-step prefix, a verify's is timed
    step_s, verify_s = {}, {}
    plain_total = 0.0
    for r in reqs:
        key = (r.q, r.n, r.R, r.M)
        if r.kind == 'search':
            if key not in step_s:
                with Cell(*key) as c:
                    c.search_params(tenure=r.tenure, stall=r.stall, kick=r.kick)
                    c.search_init(r.seed)
                    c.run(GRAPH_STEPS, mode=MODE_GRAPH)
                    c.sync()
                    t1 = time.perf_counter()
                    c.run(320, mode=MODE_GRAPH)
                    step_s[key] = (time.perf_counter() - t1) / 320
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
           'workload': 'Poisson arrivals 20 req/s (seed 7), %d requests, 30%% verify / 70%% search bursts of %d '
                       'moves over K5(7,2)..K6(10,4); record codes as verify inputs' % (a.requests, a.serve_moves),
           'hbm_budget_gb': a.budget_gb, 'tick_steps': a.tick, 'lane_cap': a.lane_cap or None,
           'latency': {k: (round(v, 4) if isinstance(v, float) else v) for k, v in s.items()},
           'makespan_s': round(el, 3), 'served_rps': round(len(res) / span, 4),
           'plain_rps_projected': round(len(res) / plain_total, 4),
           'ratio_served_vs_plain': round(plain_total / span, 3),
           'plain_us_per_move': {'%d,%d,%d,%d' % k: round(v * 1e6, 1) for k, v in step_s.items()},
           'peak_cells': srv.peak_cells, 'peak_gb': round(srv.peak_bytes / (1 << 30), 3), 'ticks': srv.ticks,
           'errors': s['errors'],
           'verify_record_uncovered': sorted({r.uncovered for q_, r in zip(sorted(reqs, key=lambda x: x.rid), res)
                                              if q_.kind == 'verify' and (q_.q, q_.n, q_.R, q_.M) in codes})}
    if with_cpu and ref_exe()[0]:
        cpu = cpu_serial(reqs, codes, a.cpu_threads or cpu_share(), a.cpu_c5_moves)
        cpu['served_over_cpu_rps'] = round(out['served_rps'] / cpu['rps'], 1)
        out['cpu_baseline'] = cpu
    return out


# --------------------------------------------------------------------- copy
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--moves-per-step', type=int, default=50000)
    ap.add_argument('--mode', choices=['graph', 'plain'], default='graph')
    ap.add_argument('--chunk', type=int, default=0, help='launch search windows in pieces of this many moves')
    ap.add_argument('--timing-steps', type=int, default=300, help='plain moves timed with per-launch events')
    ap.add_argument('--plain-steps', type=int, default=1, help='steps of plain-mode moves after the window')
    ap.add_argument('--parts', default='search,cpu,batched,serve,transform')
    ap.add_argument('--workload', choices=['search', 'transform', 'batched', 'serve', 'copy'], default='search')
    ap.add_argument('--cpu-threads', type=int, default=0, help='CPU baseline threads (0: the 16-thread share)')
    ap.add_argument('--cpu-all-moves', type=int, default=6000)
    ap.add_argument('--cpu-1t-moves', type=int, default=400)
    ap.add_argument('--cpu-c4-moves', type=int, default=100)
    ap.add_argument('--cpu-c5-moves', type=int, default=200)
    ap.add_argument('--lds', type=int, default=0, help='transform tile LDS bytes (0 = default)')
    ap.add_argument('--threads', type=int, default=0, help='transform workgroup threads (0 = default)')
    ap.add_argument('--prefix', type=int, default=1000, help='batched: moves timed in both modes')
    ap.add_argument('--batched-moves', type=int, default=100000)
    ap.add_argument('--groups', type=int, default=4, help='batched: concurrent batches, cells grouped by step cost')
    ap.add_argument('--requests', type=int, default=60, help='serve: requests in the trace')
    ap.add_argument('--serve-moves', type=int, default=100000)
    ap.add_argument('--tick', type=int, default=480, help='serve: steps per batched tick at most')
    ap.add_argument('--budget-gb', type=float, default=64.0, help='serve: HBM budget')
    ap.add_argument('--lane-cap', type=int, default=16, help='serve: searches a lane advances at once (0: all)')
    a = ap.parse_args()
    if a.gpus != 1:
        print('bench.py measures one GPU: every box here has one GPU, and the search path shards by '
              'independent cells, not within one (--gpus %d refused)' % a.gpus, file=sys.stderr)
        return 2
    if a.workload == 'copy':
        return bench_copy(a)
    if a.workload == 'transform':
        print(json.dumps(bench_transform(a)))
        return 0
    if a.workload == 'batched':
        print(json.dumps(bench_batched(a, with_cpu='cpu' in a.parts)))
        return 0
    if a.workload == 'serve':
        print(json.dumps(bench_serve(a, with_cpu='cpu' in a.parts)))
        return 0
    parts = set(a.parts.split(','))
    out = bench_search(a)
    print('search: %.1f moves/s' % out['value'], file=sys.stderr, flush=True)
    out['cpu_baseline'] = bench_cpu(a) if 'cpu' in parts else None
    if 'batched' in parts:
        out['batched'] = bench_batched(a, with_cpu='cpu' in parts)
        print('batched: %.4f solves/s' % out['batched']['value'], file=sys.stderr, flush=True)
    if 'serve' in parts:
        out['serve'] = bench_serve(a, with_cpu='cpu' in parts)
        print('serve: p50 %.2f s' % out['serve']['value'], file=sys.stderr, flush=True)
    if 'transform' in parts:
        out['transform'] = bench_transform(a)
    print(json.dumps(out))
    return 0


if __name__ == '__main__':
    sys.exit(main())
