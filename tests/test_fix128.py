"""The exact pair-sum accumulator (Fix128, crowddynamics_amd/csrc/cd_device.h)
checked on the host: tests/native/fix128_check.hip compiled with hipcc."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _hipcc():
    for c in ('/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if c and os.path.exists(c):
            return c
    return None


@pytest.mark.skipif(_hipcc() is None, reason='hipcc not available')
def test_fix128_host(tmp_path):
    exe = str(tmp_path / 'fix128_check')
    src = os.path.join(HERE, 'native', 'fix128_check.hip')
    subprocess.run([_hipcc(), '--offload-arch=gfx950', '-O2', '-std=c++17', '-ffp-contract=off',
                    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-o', exe, src],
                   check=True, capture_output=True, timeout=300)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip().endswith('ok'), out.stdout
