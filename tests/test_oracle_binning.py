"""Pins the oracle's block-list binning (the contract documented in
oracle/cd_oracle.c) against numpy and brute force."""
import numpy as np
import pytest

from crowddynamics_amd import examples as E
from oracle import oracle as O


def _numpy_bins(pos, c):
    o = pos.min(0)
    l = ((pos - o) / c).astype(np.int64)
    shape = l.max(0) + 1
    flat = l[:, 0] * shape[1] + l[:, 1]
    return o, l, tuple(shape), flat


@pytest.mark.parametrize('seed,n,c', [(0, 1, 0.6), (1, 50, 0.6), (2, 500, 0.6), (3, 500, 3.6),
                                      (4, 2000, 1.0)])
def test_counting_sort_equals_numpy_stable_argsort(seed, n, c):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-10, 10, (n, 2)) * rng.uniform(0.1, 1, 2)
    pidx, cnt, off, shape, origin, cell_of = O.add_to_cells(pos, c)
    o, l, shp, flat = _numpy_bins(pos, c)
    assert shape == shp
    assert np.array_equal(origin, o)
    assert np.array_equal(cell_of, flat)
    assert np.array_equal(pidx, np.argsort(flat, kind='stable'))
    assert np.array_equal(cnt, np.bincount(flat, minlength=shp[0] * shp[1]))
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(cnt)[:-1]]))


@pytest.mark.parametrize('seed,n,c', [(5, 300, 0.6), (6, 300, 3.6), (7, 800, 0.8)])
def test_pair_set_equals_brute_force(seed, n, c):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 12, (n, 2))
    pos[:10] = pos[10]          # coincident points
    pairs = O.iter_pairs(pos, c)
    _, l, _, _ = _numpy_bins(pos, c)
    i, j = np.triu_indices(n, 1)
    adj = (np.abs(l[i, 0] - l[j, 0]) <= 1) & (np.abs(l[i, 1] - l[j, 1]) <= 1)
    expected = set(zip(i[adj].tolist(), j[adj].tolist()))
    got = [tuple(p) for p in pairs.tolist()]
    assert len(got) == len(set(got))            # each pair exactly once
    assert set(got) == expected
    assert all(a < b for a, b in got)           # (min, max) argument order


def test_hallway_pair_set_and_forces_consistent():
    """Sum over the pair list of the per-pair contributions equals the
    block-list pass (same iteration order: bitwise)."""
    scn = E.hallway()
    arr = scn.agents.copy()
    arr['force'] = 0.0
    ref = arr.copy()
    O.agent_agent_block_list(ref, scn.cell_size)
    acc = np.zeros((len(arr), 2))
    for a, b in O.iter_pairs(arr['position'], scn.cell_size):
        out = O.pair(arr, a, b)
        if out is not None:
            acc[a] += out[0]
            acc[b] += out[1]
    assert np.array_equal(acc, ref['force'])


def test_bench_pair_tally_counts_the_oracle_pairs():
    """bench.py's pair tally (k-d tree under the binning's cell rule) counts
    exactly the oracle's candidate pairs: the 20k room and a dense packing."""
    import bench
    for arr, cell in ((E.room().agents, 0.6), (E.rollout(n=3000, density=2.0).agents, 3.6)):
        tally = bench.pair_tally(arr, cell, np.zeros(0, dtype=[('p0', 'f8', 2), ('p1', 'f8', 2)]))
        assert tally['candidate_pairs'] == len(O.iter_pairs(arr['position'], cell))
