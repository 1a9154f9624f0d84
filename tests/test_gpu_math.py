"""Device arithmetic that must equal the reference's bit for bit.

The reference takes lengths with np.hypot (core/vector2D.py:113 and every
distance of core/distance.py), which reaches the C library's hypot.  That
hypot is faithful but not correctly rounded, and distances that tie exactly
(equally oriented three-circle agents) are resolved by its last bit, so the
device's lengths must be the same bits, not merely close."""
import numpy as np
import pytest
import torch

from crowddynamics_amd import _lib
from crowddynamics_amd.device import stream_handle

pytestmark = pytest.mark.gpu


def _device_hypot(a, b):
    import ctypes as C
    xy = torch.as_tensor(np.stack([a, b], 1), dtype=torch.float64, device='cuda').contiguous()
    out = torch.empty(len(a), dtype=torch.float64, device='cuda')
    _lib.check(_lib.lib().cd_hypot(C.c_void_p(xy.data_ptr()), C.c_void_p(out.data_ptr()),
                                   len(a), stream_handle()), 'hypot')
    return out.cpu().numpy()


def test_hypot_equals_host_library_bits():
    rng = np.random.default_rng(11)
    n = 1 << 21
    parts = []
    # the crowd's scales: coordinates differences of a few metres
    parts.append((rng.uniform(-5, 5, n), rng.uniform(-5, 5, n)))
    # magnitudes spread over many binades, both orders
    e1 = rng.integers(-60, 60, n)
    e2 = e1 + rng.integers(-40, 40, n)
    parts.append((rng.uniform(-1, 1, n) * 2.0 ** e1, rng.uniform(-1, 1, n) * 2.0 ** e2))
    # the scaled branches: huge and tiny inputs, subnormals
    e3 = rng.integers(-1074, 1023, n)
    e4 = np.clip(e3 + rng.integers(-60, 60, n), -1074, 1023)
    parts.append((rng.uniform(0.5, 1, n) * 2.0 ** e3, rng.uniform(0.5, 1, n) * 2.0 ** e4))
    # special values
    sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, 1.7976931348623157e308, 1.0, 3.0])
    A, B = np.meshgrid(sp, sp)
    parts.append((A.ravel(), B.ravel()))
    for a, b in parts:
        with np.errstate(all='ignore'):
            want = np.hypot(a, b)
        got = _device_hypot(a, b)
        same = (got.view(np.uint64) == want.view(np.uint64)) | (np.isnan(got) & np.isnan(want))
        bad = np.flatnonzero(~same)
        assert len(bad) == 0, ('%d of %d differ, e.g. hypot(%r, %r) = %r on the device, %r on '
                               'the host' % (len(bad), len(a), a[bad[0]], b[bad[0]],
                                             got[bad[0]], want[bad[0]]))
