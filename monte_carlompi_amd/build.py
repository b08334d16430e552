"""Build the HIP library in-tree: monte_carlompi_amd/libmcsnow.so (gfx950).

    python -m monte_carlompi_amd.build          # or __graft_entry__.build()

hipcc cross-compiles for gfx950 without a GPU.  Packed-FP32 instruction
selection is disabled (-packed-fp32-ops): hipcc would otherwise fuse
independent float ops into v_pk_* forms.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
INCLUDE = os.path.join(ROOT, 'include')
LIB = os.path.join(HERE, 'libmcsnow.so')
SOURCES = ['walk.hip', 'mlp.hip', 'records_text.hip']
ARCH = os.environ.get('MCS_OFFLOAD_ARCH', 'gfx950')

HIPCC_FLAGS = ['--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC', '-shared',
               '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
               '-munsafe-fp-atomics', '-Wall', '-Wno-unused-function']


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError('hipcc not found')


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-I', INCLUDE, '-I', CSRC] + srcs + ['-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError('hipcc failed building libmcsnow.so')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
