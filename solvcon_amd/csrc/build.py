"""Build libsolvcon_amd.so in-tree with hipcc for gfx950.

All HIP sources are compiled with -ffp-contract=off and without fast-math,
so the FP64 arithmetic of every kernel follows source order exactly; packed
FP32 instructions are disabled.

The device code of the multidimensional march (euler.hip) is also compiled
with -fno-signed-zeros.  It only lets the compiler drop additions of +0
(accumulators that start at zero): every value stays equal to the
reference's, infinities and NaNs included; only the sign of an exactly zero
result may differ.  Finite-math assumptions are not used: a state that
overflows propagates its infinities and NaNs exactly as the reference's
IEEE arithmetic does.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, 'libsolvcon_amd.so')
OBJDIR = os.path.join(HERE, '_build')

SOURCES = ['capi.hip', 'onedim.hip', 'mesh.hip', 'euler.hip', 'order.hip', 'ge.hip', 'spacetime.hip']

FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
    '-ffp-contract=off', '-fno-fast-math',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function',
]


def hipcc():
    for c in ('/opt/rocm/bin/hipcc', 'hipcc'):
        if os.path.exists(c) or c == 'hipcc':
            return c


PER_SOURCE = {
    'euler.hip': ['-Xarch_device', '-fno-signed-zeros'],
}


def _compile(src, extra=()):
    obj = os.path.join(OBJDIR, src.replace('.hip', '.o'))
    cmd = [hipcc(), *FLAGS, *PER_SOURCE.get(src, []), *extra, '-c',
           os.path.join(HERE, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = '\n'.join(l for l in r.stderr.splitlines()
                    if 'packed-fp32-ops' not in l)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (src, err))
    if err.strip():
        sys.stderr.write(err + '\n')
    return obj


def build(verbose=False, extra=()):
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, extra), SOURCES))
    tmp = OUT + '.tmp'
    cmd = [hipcc(), '--offload-arch=gfx950', '-shared', '-fPIC', *objs,
           '-o', tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n' + r.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print('built', OUT)
    return OUT


if __name__ == '__main__':
    build(verbose=True)
