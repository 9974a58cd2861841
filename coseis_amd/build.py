"""
Build the HIP solver library for gfx950, in-tree: coseis_amd/libsord_amd.so.

    python -m coseis_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libsord_amd.so')
SOURCES = ['sord_kernels.hip', 'sord_fused.hip', 'sord_adjoint.hip', 'sord_init.hip', 'sord_host.cpp']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = [
    '-std=c++17', '-O3', '--offload-arch=gfx950', '-fPIC',
    # packed FP32 instructions are disabled for all device code
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-I' + os.path.join(ROOT, 'include'), '-I' + SRC,
]


def build(verbose=False):
    """Compile the sources (in parallel) and link the library."""
    objs, procs = [], []
    for s in SOURCES:
        src = os.path.join(SRC, s)
        if not os.path.exists(src):
            continue
        obj = os.path.join(SRC, '_build', s + '.o')
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = [HIPCC] + FLAGS + ['-c', src, '-o', obj]
        if verbose:
            print(' '.join(cmd))
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for cmd, p in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, cmd)
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '-shared', '--offload-arch=gfx950', '-fPIC', '-o', tmp] + objs
    if verbose:
        print(' '.join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    build(verbose='-v' in sys.argv)
    print(LIB)
