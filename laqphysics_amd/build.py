"""Build the HIP engine library in-tree for gfx950.

    python -m laqphysics_amd.build        # -> laqphysics_amd/build/libhexhip.so

The library is a plain C-ABI shared object compiled by hipcc; Python loads it
with ctypes (laqphysics_amd/native.py), as the reference loads its native kernel
from build/ (python/native.py:7-8).
"""
import os
import shlex
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
SOURCES = ['csrc/hex_engine.hip', 'csrc/hex_krylov.hip', 'csrc/hex_separable.hip', 'csrc/hex_surface.hip',
           'csrc/hex_batch.hip', 'csrc/hex_apply2.hip', 'csrc/hex_ionic.hip', 'csrc/hex_yield.hip']
OUTPUT = PKG / 'build' / 'libhexhip.so'
ARCH = os.environ.get('HEX_OFFLOAD_ARCH', 'gfx950')
FLAGS = ['-O3', '-std=c++17', '-fPIC', '-shared', f'--offload-arch={ARCH}',
         # packed-FP32 VALU forms are kept out of every kernel
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-command-line-argument']


def build(verbose=False):
    OUTPUT.parent.mkdir(exist_ok=True)
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    tmp = OUTPUT.with_suffix('.so.tmp')
    cmd = [hipcc, *FLAGS, '-o', str(tmp), *[str(PKG / s) for s in SOURCES], '-L/opt/rocm/lib', '-lrocblas', '-lrocsolver',
           '-Wl,-rpath,/opt/rocm/lib']
    if verbose:
        print(' '.join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
