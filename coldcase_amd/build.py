"""Build libcovhip.so in-tree with hipcc for gfx950.

    python -m coldcase_amd.build

Every HIP source is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'lib', 'libcovhip.so')
ARCH = os.environ.get('COV_OFFLOAD_ARCH', 'gfx950')


def hipcc():
    for c in ('/opt/rocm/bin/hipcc', 'hipcc'):
        if os.path.isabs(c) and os.path.exists(c):
            return c
    return 'hipcc'


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + '.tmp'
    cmd = [hipcc(), '--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC', '-shared',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-I', os.path.join(ROOT, 'include'), '-I', CSRC,
           os.path.join(CSRC, 'covhip.hip'), '-o', tmp]
    if verbose:
        print(' '.join(cmd))
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError('hipcc failed (%d)' % p.returncode)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose=True))
