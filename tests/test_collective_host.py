"""Pins of the oracle's collective-motion restatement (no device)."""
import numpy as np

from oracle import oracle as O


def test_line_intersect_known_answers():
    a, b = np.array([0.0, 0.0]), np.array([2.0, 2.0])
    assert O.line_intersect(a, b, np.array([0.0, 2.0]), np.array([2.0, 0.0]))
    assert not O.line_intersect(a, b, np.array([3.0, 0.0]), np.array([3.0, 5.0]))
    assert not O.line_intersect(a, b, np.array([1.0, 0.0]), np.array([3.0, 2.0]))  # parallel
    assert O.line_intersect(a, b, np.array([2.0, 2.0]), np.array([5.0, 0.0]))      # touching end


def test_nearest_neighbors_small_cases():
    p = np.array([(0.0, 0.0), (1.0, 0.0), (0.0, 2.0), (5.0, 0.0), (-1.5, 0.0)])
    nb = O.nearest_neighbors(p, sight=3.0, k=2, segments=[])
    assert list(nb[0]) == [1, 4]          # 1.0 and 1.5 m
    assert list(nb[3]) == [-1, -1]        # nobody within 3 m
    wall = [(0.5, -1.0, 0.5, 1.0)]        # blocks 0 <-> 1
    nb = O.nearest_neighbors(p, sight=3.0, k=2, segments=wall)
    assert list(nb[0]) == [4, 2]


def test_herding_single_neighbour_closed_form():
    """One neighbour ahead walking away: the direction is the normalised
    weighted average of the unit vector toward it and its unit velocity."""
    p = np.array([(0.0, 0.0), (2.0, 1.0)])
    v = np.array([(1.0, 0.0), (0.0, 3.0)])
    nb = np.array([[1], [-1]])
    w = 0.15
    d, has = O.herding_interaction(np.array([True, False]), p, v, nb, w, 0.45 * np.pi)
    e0 = np.array([2.0, 1.0]) / np.hypot(2.0, 1.0)
    want = w * e0 + (1 - w) * np.array([0.0, 1.0])
    want /= np.hypot(*want)
    np.testing.assert_allclose(d[0], want, rtol=1e-15)
    assert has[0] and not has[1]
    # a neighbour coming head-on (inside the cone) is not followed
    v2 = np.array([(1.0, 0.0), (-2.0, -1.0)])
    d, has = O.herding_interaction(np.array([True, False]), p, v2, nb, w, 0.45 * np.pi)
    assert not has[0]
