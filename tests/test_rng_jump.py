"""The rng jump matrices of the four-wave ring fill, against sequential
xoshiro256++ stepping (host build of tests/host/rng_jump.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_jump_matrices_match_sequential_steps(tmp_path):
    hipcc = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(hipcc):
        pytest.skip('hipcc not available')
    out = str(tmp_path / 'rng_jump')
    cmd = [hipcc, '--offload-arch=gfx950', '-x', 'hip', '-O2', '-std=c++17',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-I', os.path.join(ROOT, 'include'), '-I', os.path.join(ROOT, 'coldcase_amd', 'csrc'),
           os.path.join(ROOT, 'tests', 'host', 'rng_jump.cpp'), '-o', out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith('ok'), r.stdout + r.stderr[-2000:]
