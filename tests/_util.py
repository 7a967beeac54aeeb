"""Shared helpers for the tests (code-file parsing of fixtures, paths)."""
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, 'tests', 'golden')
DATA = os.path.join(GOLDEN, 'data')
DIGITS = '0123456789abcdefghijklmnopqrstuvwxyz'


def read_code(path, n):
    out = []
    for raw in open(path):
        line = raw.split('#')[0].strip()
        if not line:
            continue
        if any(ch in line for ch in ' ,\t'):
            w = [int(t) for t in line.replace(',', ' ').split()]
        else:
            w = [DIGITS.index(ch.lower()) for ch in line]
        assert len(w) == n
        out.append(w)
    return np.array(out, dtype=np.uint8)


def golden_runs():
    return json.load(open(os.path.join(GOLDEN, 'trajectories.json')))['runs']


DSEL_RANDOM = 1 << 60
# printed precision of the reference's -DPROF aggregates (covsearch2e.c:1602-1612)
PROF_TOL = {'mean_ncand': 0.05, 'mean_take': 0.05, 'mean_uncovered': 0.05, 'mean_bestd': 0.005, 'mean_ncw': 0.05}


def prof_aggregates(trace, unc0, cand=100000):
    """The reference's -DPROF per-step means from a step trace (the layout
    shared by the oracle and the engine): candidates, sample size, the
    uncovered count at the top of each step (before its pick), the selected
    delta over steps that selected by delta, and distinct codewords."""
    import numpy as np
    tr = np.asarray(trace)
    it = len(tr)
    unc_top = np.concatenate([[unc0], tr[:-1, 10]])
    dsel = tr[:, 8]
    return {'mean_ncand': tr[:, 2].sum() / it, 'mean_take': np.minimum(tr[:, 2], cand).sum() / it,
            'mean_uncovered': unc_top.sum() / it, 'mean_bestd': dsel[dsel != DSEL_RANDOM].sum() / it,
            'mean_ncw': tr[:, 3].sum() / it}


def prof_matches(agg, ref):
    """Each aggregate equals the reference's printed value to its precision."""
    return {k: (agg[k], v) for k, v in ref.items() if abs(agg[k] - v) > PROF_TOL[k] + 1e-9}


def cli_code_sha(code, q, M):
    """sha256 of the code file the solver publishes for this best code:
    exactly M distinct words, one per line (covsearch2e.c:791-846)."""
    import hashlib
    from coldcase_amd.covsearch import distinct_fill
    code = np.asarray(code, dtype=np.uint8)
    text = ''.join(''.join(DIGITS[int(v)] for v in w) + '\n' for w in distinct_fill(code, q, code.shape[1], M))
    return hashlib.sha256(text.encode()).hexdigest()
