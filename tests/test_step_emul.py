"""Host emulation of the step kernels' word-index arithmetic (no GPU): the
cell plan that cov_create builds (coldcase_amd/csrc/cov_plan.h), its subset
tables in exact-size heap arrays under AddressSanitizer, and the index
functions the evaluator, commit, ball and kept-count code use (pattern_word,
the sphere-walk tables and walk_word, ball_word_at, pack_index, the division
magics, commit_offset and commit_word, the commit's subset rows derived from
the row avoiding the last position, slot_diff) against brute-force digit arithmetic,
for the workload cells and shapes at the edges of the plan: q = 2 (one
pattern per subset), q = 16, R = 0, R = 8, walks without tables ((q-1)^ceil(R/2)
above 128), and q^n above 2^32 (64-bit digit extraction)."""
import os
import shutil
import subprocess

import pytest

from tests._util import ROOT

SHAPES = [  # q n R centres
    (5, 7, 2, 7), (6, 9, 4, 5), (7, 9, 4, 4), (6, 10, 4, 4),
    (2, 16, 8, 5), (2, 12, 1, 7), (16, 3, 2, 7), (16, 8, 5, 3), (16, 6, 5, 3),
    (3, 12, 8, 4), (9, 11, 2, 6), (4, 8, 0, 7), (13, 4, 3, 7), (8, 9, 3, 4),
]


@pytest.fixture(scope='module')
def emul(tmp_path_factory):
    hipcc = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(hipcc):
        pytest.skip('hipcc not available')
    out = str(tmp_path_factory.mktemp('step') / 'step_emul')
    cmd = [hipcc, '--offload-arch=gfx950', '-x', 'hip', '-O1', '-g', '-std=c++17',
           '-Xarch_host', '-fsanitize=address', '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-I', os.path.join(ROOT, 'include'), '-I', os.path.join(ROOT, 'coldcase_amd', 'csrc'),
           os.path.join(ROOT, 'tests', 'host', 'step_emul.cpp'), '-o', out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


@pytest.mark.parametrize('shape', SHAPES, ids=lambda s: 'K%d(%d,%d)' % s[:3])
def test_step_index_math_matches_brute_force(emul, shape):
    q, n, R, centres = shape
    r = subprocess.run([emul, str(q), str(n), str(R), str(1000 + q * 31 + n), str(centres), '400000'],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith('ok'), r.stdout + r.stderr[-3000:]
