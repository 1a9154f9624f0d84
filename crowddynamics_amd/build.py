"""Build recipe of the HIP library (gfx950 only).

``python -m crowddynamics_amd.build`` compiles ``csrc/*.hip`` into
``crowddynamics_amd/libcdamd.so`` with hipcc.  Flags:

* ``--offload-arch=gfx950``: MI355X code object only.
* ``-ffp-contract=off``: no a*b+c fusion, the IEEE semantics of the
  reference's numba build (fp64 parity at the 1e-12 level).
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed-fp32 VALU
  instructions (the v_pk_* fp32 forms are avoided on these machines).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libcdamd.so')
SOURCES = ['cd_kernels.hip', 'cd_adjoint.hip', 'cd_field.hip', 'cd_state.hip']

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
    '-ffp-contract=off',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
]


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found')


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-I', os.path.join(HERE, '..', 'include'),
                                     '-o', tmp] + srcs
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    build(verbose='-v' in sys.argv)
