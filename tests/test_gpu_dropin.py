"""The drop-ins on the GPU.

covsearch: every golden run of the reference solver (118 runs, compiled from
source, tests/golden/make_golden.py) replayed through the covsearch CLI
contract must print the same RESULT fields and publish a byte-identical --out
file.  verify_cov: verify_pure and verify_numpy against brute force on random
codes (uncovered count and the five smallest witnesses), the record codes,
the covering radius, and the CLI's exit codes.
"""
import hashlib
import os
import re
import time

import numpy as np
import pytest

from oracle.verify_oracle import mindist
from tests._util import DATA, golden_runs

pytestmark = pytest.mark.gpu

RUNS = golden_runs()


def _id(r):
    return '%s-K%d(%d,%d)M%d-s%d-it%d' % (r['tag'], r['q'], r['n'], r['R'], r['M'], r['seed'], r['iters'])


@pytest.mark.parametrize('r', RUNS, ids=_id)
def test_covsearch_cli_matches_reference(r, tmp_path, capsys):
    from coldcase_amd import covsearch
    print('case cli %s' % _id(r), flush=True)
    out = str(tmp_path / 'code.txt')
    argv = ['-q', str(r['q']), '-n', str(r['n']), '-R', str(r['R']), '-M', str(r['M']), '-s', str(r['seed']),
            '--iters', str(r['iters']), '-t', '1e9', '--quiet', '--out', out] + r['flags']
    if r['infile']:
        argv += ['--in', os.path.join(DATA, r['infile'])]
    rc = covsearch.main(argv)
    text = capsys.readouterr().out
    m = re.search(r'RESULT q=\d+ n=\d+ R=\d+ M=\d+ uncovered=(\d+) iters=(\d+) kicks=(\d+) .*target=(-?\d+) '
                  r'ttt_cpu=(\S+) ttt_wall=(\S+) ttt_it=(-?\d+)', text)
    assert m, text
    assert (int(m.group(1)), int(m.group(2)), int(m.group(3))) == (r['uncovered'], r['iters_done'], r['kicks'])
    f = r['flags']
    target = int(f[f.index('--target') + 1]) if '--target' in f else -1
    assert int(m.group(4)) == target
    assert int(m.group(7)) == r.get('ttt_it', -1)
    if r.get('ttt_it', -1) >= 0:
        assert 0.0 <= float(m.group(6)) < 600 and 0.0 <= float(m.group(5)) < 600
    else:
        assert float(m.group(5)) == -1.0 and float(m.group(6)) == -1.0
    assert rc == (0 if r['uncovered'] == 0 else 1)
    assert hashlib.sha256(open(out, 'rb').read()).hexdigest() == r['out_sha256']


def _result(text):
    m = re.search(r'RESULT .*uncovered=(\d+) iters=(\d+) kicks=(\d+) time=(\S+) .*cpu=(\S+) cpurate', text)
    assert m, text
    return int(m.group(1)), int(m.group(2)), int(m.group(3)), float(m.group(4)), float(m.group(5))


@pytest.mark.parametrize('budget', [['-t', '1.5'], ['--cpu', '0.6', '-t', '999999']], ids=['wall', 'cpu'])
def test_covsearch_cli_stops_at_its_budget(budget, tmp_path, capsys):
    """The device reads the clock before every step (covsearch2e.c:1305-1315):
    a run far longer than its budget ends within it, publishes M distinct
    codewords, and is a clean prefix of the unbounded trajectory."""
    from coldcase_amd import covsearch
    print('case cli budget %s' % budget, flush=True)
    out = str(tmp_path / 'code.txt')
    base = ['-q', '6', '-n', '9', '-R', '4', '-M', '660', '-s', '5', '--random-init', '--tabu', '10', '--quiet']
    t0 = time.monotonic()
    rc = covsearch.main(base + ['--out', out] + budget)
    wall = time.monotonic() - t0
    unc, iters, kicks, el, cpu = _result(capsys.readouterr().out)
    assert rc == 1 and 100 < iters < 10 ** 6
    lim = float(budget[1])
    assert wall < lim + 2.0, wall
    if budget[0] == '--cpu':
        assert lim - 0.05 <= cpu < lim + 0.25, cpu
    lines = open(out).read().split()
    assert len(lines) == 660 and len(set(lines)) == 660
    # the same run with its iteration count as the budget publishes the same code
    out2 = str(tmp_path / 'code2.txt')
    rc = covsearch.main(base + ['--out', out2, '--iters', str(iters), '-t', '1e9'])
    assert _result(capsys.readouterr().out)[:3] == (unc, iters, kicks)
    assert open(out2).read() == open(out).read()


def test_cell_target_matches_oracle():
    """--target on the device: the trace, the stop (done = 2) and the
    iteration that reached the target equal the oracle's."""
    from coldcase_amd import Cell
    from oracle.oracle import Oracle
    q, n, R, M, seed, target = 5, 7, 2, 400, 3, 4300
    print('case target K%d(%d,%d) M=%d seed %d target %d' % (q, n, R, M, seed, target), flush=True)
    with Cell(q, n, R, M) as c:
        c.search_params(tenure=10)
        c.search_init(seed)
        c.limits(target)
        tr = c.run(20000, trace=True)
        st = c.state()
        ttt_it, age = c.ttt()
    o = Oracle(q, n, R, M, seed).init(greedy=False)
    o.set_target(target)
    ref = o.run(20000, tabu=10, trace=True)
    assert tr.shape == ref.shape and (tr == ref).all()
    assert st['done'] == 2 and st['best_uncovered'] <= target
    assert ttt_it == o.ttt_it == st['it'] and age >= 0.0
    # lifting the target resumes the same trajectory
    with Cell(q, n, R, M) as c:
        c.search_params(tenure=10)
        c.search_init(seed)
        c.limits(target)
        c.run(20000)
        c.limits(-1)
        tr2 = c.run(50, trace=True)
    o.set_target(-1)
    ref2 = o.run(50, tabu=10, trace=True)
    assert (tr2 == ref2).all()


def test_cell_deadline_stops_before_a_step():
    from coldcase_amd import Cell
    print('case deadline K6(9,4)', flush=True)
    with Cell(6, 9, 4, 660) as c:
        c.search_params(tenure=10)
        c.search_init(1)
        c.limits(-1, -1.0)               # already past: no step begins
        assert c.run(100) == 0 and c.state()['it'] == 0
        # at about 40 us a step, 0.2 s of deadline is about 5,000 of the
        # 40,000 steps launched; the launches after it begin no step
        c.limits(-1, 0.2)
        ran = c.run(40000, mode=1)
        assert 1000 < ran < 20000, ran
        c.limits(-1, 0.0)                # no deadline
        assert c.run(10) == 10


def _word(w, q, n):
    return tuple(int(x) for x in np.base_repr(int(w), q).zfill(n))


@pytest.mark.parametrize('q,n,R,M', [(2, 8, 1, 20), (3, 6, 1, 40), (4, 5, 2, 9), (5, 6, 2, 60), (6, 5, 2, 30),
                                     (7, 4, 1, 50), (3, 9, 2, 60)])
def test_verify_methods_against_brute_force(q, n, R, M):
    from coldcase_amd.verify_cov import covering_radius, verify_numpy, verify_pure
    print('case verify K%d(%d,%d) M=%d' % (q, n, R, M), flush=True)
    rng = np.random.default_rng(q * 31 + n)
    code = [tuple(int(x) for x in w) for w in rng.integers(0, q, size=(M, n))]
    d = mindist(np.array(code, dtype=np.uint8), n, q)
    far = np.flatnonzero(d > R)
    want = (int(far.size), [_word(w, q, n) for w in far[:5]])
    assert verify_pure(code, n, q, R) == want
    assert verify_numpy(code, n, q, R) == want
    assert covering_radius(code, n, q, verify_numpy) == int(d.max())
    assert covering_radius(code, n, q, verify_pure) == int(d.max())


@pytest.mark.parametrize('fname,q,n,R', [('K6_9_4_M660.txt', 6, 9, 4), ('K7_9_4_M1480.txt', 7, 9, 4),
                                         ('K6_10_4_M2784.txt', 6, 10, 4), ('K6_6_3_M41.txt', 6, 6, 3)])
def test_verify_cli_record_codes(fname, q, n, R, capsys):
    from coldcase_amd import verify_cov
    print('case verify cli %s' % fname, flush=True)
    rc = verify_cov.main([os.path.join(DATA, fname), '-q', str(q), '-n', str(n), '-R', str(R), '--method', 'both'])
    text = capsys.readouterr().out
    assert rc == 0, text
    assert 'uncovered=0' in text
    # one radius short: not a covering, exit 1, both methods agree
    rc = verify_cov.main([os.path.join(DATA, fname), '-q', str(q), '-n', str(n), '-R', str(R - 1),
                          '--method', 'both', '--quiet'])
    assert rc == 1


def test_verify_cli_output_as_the_reference_scripts_read_it(tmp_path, capsys):
    """The CLI's per-method lines read with the reference scripts' own
    patterns: the regex of cov/gemm/run_bench.py:56 over `--method both`, and
    the awk of cov/engine/selftest.sh:61-62 over `--method pure`."""
    from coldcase_amd import verify_cov
    print('case verify cli patterns', flush=True)
    rng = np.random.default_rng(4)
    code = rng.integers(0, 5, size=(300, 7))
    p = tmp_path / 'c.txt'
    p.write_text(''.join(''.join(str(int(v)) for v in w) + '\n' for w in code))
    d = mindist(np.unique(code.astype(np.uint8), axis=0), 7, 5)
    want = int((d > 2).sum())
    rc = verify_cov.main([str(p), '-q', '5', '-n', '7', '-R', '2', '--method', 'both'])
    out = capsys.readouterr().out
    vals = re.findall(r"^method [a-z]+ *: uncovered=(\d+)", out, re.M)
    assert rc == 1 and len(vals) == 2 and {int(v) for v in vals} == {want}
    rc = verify_cov.main([str(p), '-q', '5', '-n', '7', '-R', '2', '--method', 'pure', '--radius'])
    out = capsys.readouterr().out
    # awk -F'uncovered=' '/method pure/{print $2+0}'
    got = [line.split('uncovered=')[1] for line in out.splitlines() if 'method pure' in line]
    assert len(got) == 1 and int(re.match(r'\s*(\d+)', got[0]).group(1)) == want
    assert 'radius  %d' % int(d.max()) in out
