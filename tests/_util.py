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
