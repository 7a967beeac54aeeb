"""Build libcovhip.so in-tree with hipcc for gfx950.

    python -m coldcase_amd.build

Every HIP source is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).

Kernels take the cell (a struct of some 660 bytes) by value.  The compiler
reads such an argument in place only while its private copy has at most
`instcombine-max-copied-from-constant-users` users (300 by default); past
that, every lane copies the whole struct into scratch at launch, which
costs the step kernels several microseconds.  The limit is raised, and the
build fails if any kernel still uses private memory: the per-kernel
register and LDS figures the compiler reports are written beside the
library (lib/resources.txt).
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'lib', 'libcovhip.so')
ARCH = os.environ.get('COV_OFFLOAD_ARCH', 'gfx950')
RESOURCES = os.path.join(HERE, 'lib', 'resources.txt')


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
           '-mllvm', '-instcombine-max-copied-from-constant-users=4096',
           '-Rpass-analysis=kernel-resource-usage',
           '-I', os.path.join(ROOT, 'include'), '-I', CSRC,
           os.path.join(CSRC, 'covhip.hip'), '-o', tmp]
    if verbose:
        print(' '.join(cmd))
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError('hipcc failed (%d)' % p.returncode)
    usage = kernel_resources(p.stderr)
    if not usage:
        raise RuntimeError('hipcc reported no kernel resource usage')
    with open(RESOURCES, 'w') as fh:
        fh.write('%-56s %6s %6s %6s %8s %8s %6s\n' % ('kernel', 'VGPRs', 'SGPRs', 'sspill', 'scratch', 'LDS', 'occ'))
        for name, u in sorted(usage.items()):
            fh.write('%-56s %6d %6d %6d %8d %8d %6d\n' % (name[:56], u.get('VGPRs', 0), u.get('TotalSGPRs', 0),
                                                          u.get('SGPRsSpill', 0), u.get('ScratchSize', 0),
                                                          u.get('LDS', 0), u.get('Occupancy', 0)))
    private = sorted(k for k, u in usage.items() if u.get('ScratchSize', 0) or u.get('VGPRsSpill', 0))
    if private:
        os.unlink(tmp)
        raise RuntimeError('kernels using private memory (scratch or spills): ' + ', '.join(private))
    os.replace(tmp, OUT)
    return OUT


_FIELDS = {'TotalSGPRs': 'TotalSGPRs', 'VGPRs': 'VGPRs', 'ScratchSize [bytes/lane]': 'ScratchSize',
           'Occupancy [waves/SIMD]': 'Occupancy', 'VGPRs Spill': 'VGPRsSpill', 'SGPRs Spill': 'SGPRsSpill',
           'LDS Size [bytes/block]': 'LDS'}


def kernel_resources(text):
    """{kernel: {field: value}} from hipcc's kernel-resource-usage remarks
    (device functions that are not kernels are reported too and kept)."""
    out, name = {}, None
    for line in text.splitlines():
        m = re.search(r'remark: Function Name: (\S+)', line)
        if m:
            name = m.group(1)
            out[name] = {}
            continue
        m = re.search(r'remark:\s+([^:]+?): (\d+)', line)
        if m and name and m.group(1) in _FIELDS:
            out[name][_FIELDS[m.group(1)]] = int(m.group(2))
    return out


if __name__ == '__main__':
    print(build(verbose=True))
