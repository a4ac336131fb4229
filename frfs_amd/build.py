"""Build the HIP shared library ``libfrfs_fq.so`` in-tree for gfx950.

Usage: ``python -m frfs_amd.build`` (or ``frfs_amd.build.build()``).
The library is written next to this file and is loaded by ``frfs_amd._lib``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libfrfs_fq.so')

SOURCES = ['fq_lib.hip', 'fq_aux.hip', 'sur_lib.hip']
HEADERS = ['fft.hpp', 'plan.hpp']

# Packed FP32 VALU forms are disabled for every HIP translation unit.
HIP_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found')


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    """Compile the library (always recompiles).  Returns the library path."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + '.tmp'
    cmd = [hipcc(), *HIP_FLAGS, *(extra or []), '-I', os.path.join(HERE, '..', 'include'),
           *srcs, '-lhipfft', '-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError('hipcc failed building libfrfs_fq.so')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
