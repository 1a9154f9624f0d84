"""Shared helpers of the parity tests."""
import numpy as np


def assert_vec_close(got, ref, scale, rtol, what=''):
    """|got_i - ref_i| <= rtol * (scale_i + |ref_i|) per agent.

    `scale_i` is the sum of the magnitudes of the terms summed into ref_i: two
    summation orders of the same terms differ by at most a few ulps of that
    sum, so this is the tolerance that a correct reordering meets."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.asarray(scale, dtype=np.float64)
    if got.ndim == 2:
        err = np.hypot(*(got - ref).T)
        mag = np.hypot(*ref.T)
    else:
        err = np.abs(got - ref)
        mag = np.abs(ref)
    bound = rtol * (scale + mag) + 1e-300
    bad = ~(err <= bound)
    if bad.any():
        k = int(np.argmax(np.where(bad, err / bound, -1)))
        raise AssertionError('%s: %d/%d agents out of tolerance; worst agent %d err %.3e '
                             'bound %.3e got %s ref %s' % (what, bad.sum(), len(err), k,
                                                           err[k], bound[k], got[k], ref[k]))


def max_rel(got, ref, scale):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if got.ndim == 2:
        err = np.hypot(*(got - ref).T)
        mag = np.hypot(*ref.T)
    else:
        err = np.abs(got - ref)
        mag = np.abs(ref)
    return float(np.max(err / (np.asarray(scale) + mag + 1e-300))) if len(err) else 0.0
