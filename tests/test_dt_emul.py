"""Host emulation of the distance-transform kernels (no GPU): the pass plan
and every per-thread phase of coldcase_amd/csrc/cov_dist.h run thread by
thread over exact-size heap buffers under AddressSanitizer, against a
brute-force minimum distance and its histogram, for shapes that exercise
one-pass, two-pass and many-pass plans, column passes and LDS-tile outer
passes, ragged row chunks, odd strides and small LDS budgets; seed tiles
in the closed form (few codewords) and by the row axes (many); two-stage
column passes (rows through LDS) and register column passes."""
import os
import shutil
import subprocess

import pytest

from tests._util import ROOT

SHAPES = [  # q n M seed lds
    (5, 7, 40, 1, 32768), (6, 6, 20, 2, 32768), (7, 5, 10, 3, 4096), (2, 12, 30, 4, 4096),
    (16, 3, 10, 5, 32768), (3, 10, 25, 6, 8192), (4, 8, 12, 7, 32768), (6, 7, 30, 8, 4096),
    (13, 4, 9, 9, 4096), (2, 16, 3, 10, 8192), (9, 5, 7, 11, 16384), (5, 8, 1, 12, 4096),
    (11, 4, 50, 13, 4096), (3, 11, 2, 14, 63488),
    # the cells of the workloads: K5(7,2), K6(9,4), K7(9,4), K6(10,4) (checked
    # against a full-array axis DP on the large ones)
    (5, 7, 400, 21, 32768), (6, 9, 660, 22, 32768), (7, 9, 1480, 23, 32768), (6, 10, 2784, 24, 32768),
    # wider workgroups
    (7, 5, 10, 3, 4096, 512), (3, 10, 25, 6, 8192, 1024), (2, 16, 3, 10, 8192, 512), (6, 9, 660, 25, 40960, 1024),
    # the light statistics (#{dist <= r} and the maximum) instead of the histogram
    (5, 7, 40, 31, 32768, 256, 2), (7, 9, 1480, 32, 40960, 512, 4), (2, 16, 3, 33, 8192, 256, 9), (3, 11, 25, 34, 4096, 256, 0),
    # seed groups of 64 (q = 4), padded groups of 27, 9 and 11 words, dense codes,
    # and a seed pass that is also the last pass (light statistics)
    (4, 9, 50, 51, 32768), (3, 12, 100, 52, 16384), (9, 5, 400, 53, 32768, 512), (11, 4, 300, 54, 8192),
    (8, 6, 2000, 55, 32768), (5, 9, 3000, 56, 49152, 256, 3), (16, 3, 10, 57, 32768, 256, 1),
    # seed tiles in the closed form and by the row axes in one pass, with the
    # full histogram and the light statistics; closed-form tiles of two blocks
    (5, 8, 1500, 67, 8192), (5, 8, 1400, 68, 8192, 512, 3), (7, 7, 1900, 69, 16384), (4, 10, 300, 70, 4096),
    # register column passes only (no two-stage pass); two-stage passes whose
    # workgroups loop over tiles (3 and 5 compute units)
    (7, 9, 1480, 23, 32768, 256, -1, 1, 0), (7, 9, 1480, 26, 49152, 512, 3, 1, 1, 3),
    (3, 12, 100, 52, 16384, 256, -1, 1, 1, 5),
    # outer passes in LDS tiles (the plan for spaces of 2^32 words and more)
    (7, 9, 1480, 41, 49152, 512, -1, 0), (6, 10, 2784, 42, 49152, 512, 4, 0), (2, 16, 3, 43, 8192, 256, -1, 0),
    (3, 11, 25, 44, 4096, 1024, 2, 0), (13, 4, 9, 45, 4096, 256, -1, 0),
]


@pytest.fixture(scope='module')
def emul(tmp_path_factory):
    hipcc = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(hipcc):
        pytest.skip('hipcc not available')
    out = str(tmp_path_factory.mktemp('dt') / 'dt_emul')
    cmd = [hipcc, '--offload-arch=gfx950', '-x', 'hip', '-O1', '-g', '-std=c++17',
           '-Xarch_host', '-fsanitize=address', '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-I', os.path.join(ROOT, 'include'), '-I', os.path.join(ROOT, 'coldcase_amd', 'csrc'),
           os.path.join(ROOT, 'tests', 'host', 'dt_emul.cpp'), '-o', out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def _id(s):
    light = '_r%d' % s[6] if len(s) > 6 and s[6] >= 0 else ''
    lds = '_ldsouter' if len(s) > 7 and not s[7] else ''
    tc = '_notcol' if len(s) > 8 and not s[8] else ''
    cu = '_cu%d' % s[9] if len(s) > 9 else ''
    return 'K%d(%d)M%d_lds%d_t%d%s%s%s%s' % (s[0], s[1], s[2], s[4], s[5] if len(s) > 5 else 256, light, lds, tc, cu)


@pytest.mark.parametrize('shape', SHAPES, ids=_id)
def test_transform_phases_match_brute_force(emul, shape):
    r = subprocess.run([emul] + [str(x) for x in shape], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith('ok'), r.stdout + r.stderr[-3000:]
