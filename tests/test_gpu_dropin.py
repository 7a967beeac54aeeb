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
    m = re.search(r'RESULT q=\d+ n=\d+ R=\d+ M=\d+ uncovered=(\d+) iters=(\d+) kicks=(\d+)', text)
    assert m, text
    assert (int(m.group(1)), int(m.group(2)), int(m.group(3))) == (r['uncovered'], r['iters_done'], r['kicks'])
    assert rc == (0 if r['uncovered'] == 0 else 1)
    assert hashlib.sha256(open(out, 'rb').read()).hexdigest() == r['out_sha256']


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
