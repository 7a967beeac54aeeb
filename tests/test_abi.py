"""The C ABI library loads and exports every entry point include/covhip.h
declares; shape limits and memory sizing behave (no GPU needed)."""
import os
import re

import pytest

from tests._util import ROOT


def declared():
    src = open(os.path.join(ROOT, 'include', 'covhip.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(cov_[a-z_]+)\s*\(', src)))


def test_library_exports_every_declared_symbol():
    from coldcase_amd import engine
    L = engine.lib()
    names = declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
    assert sorted(engine.EXPORTS) == names


def test_version_and_errors():
    from coldcase_amd import engine
    L = engine.lib()
    assert L.cov_version() == 1
    assert L.cov_strerror(-1) == b'bad parameters'
    assert L.cov_strerror(-2) == b'q^n too large'


def test_mem_bytes_and_limits():
    from coldcase_amd.engine import CovError, mem_bytes
    b = mem_bytes(6, 9, 4, 660)
    # cnt and the coverer XOR (2 B/word each) + three bit planes + list +
    # distance array and bitmap
    assert 5.5 * 6 ** 9 < b < 8.0 * 6 ** 9
    assert mem_bytes(6, 10, 4, 2784) > 121e6
    for bad in [(1, 5, 1, 3), (17, 3, 1, 3), (3, 5, 5, 3), (3, 5, 1, 70000), (3, 17, 1, 3)]:
        with pytest.raises(CovError):
            mem_bytes(*bad)
    with pytest.raises(CovError):
        mem_bytes(9, 12, 2, 100)      # q^n above 4e10


def test_create_without_gpu_fails_loudly():
    if os.path.exists('/dev/kfd'):
        pytest.skip('GPU present')
    from coldcase_amd import Cell, CovError
    with pytest.raises(CovError):
        Cell(5, 7, 2, 400)


def test_graph_steps_constant_matches_the_header():
    from coldcase_amd import engine
    src = open(os.path.join(ROOT, 'include', 'covhip.h')).read()
    assert int(re.search(r'#define COV_GRAPH_STEPS (\d+)', src).group(1)) == engine.GRAPH_STEPS
