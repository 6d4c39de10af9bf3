"""GPU parity of the supervoxel nearest neighbours (tgmm_knnCUDA_, the
reference's knnCUDA_, temporalLogicalRules/knnCUDA/knnCuda.cu:142-232)
through the C ABI: slot indices and distances bit-identical to the oracle's
replay of the reference's scan, for random and raster-ordered centroids
(where the kernel skips most reference blocks), ties, anisotropy, self
search, NaN centroids, fewer references than slots, and the edge cases."""
import numpy as np
import pytest

import oracle_lib as O
from tgmm_amd import em, synth

pytestmark = pytest.mark.gpu

SCALE = (1.0, 1.0, 5.0)


def _check(query, ref, K, scale=SCALE):
    gi, gd = em.knnCUDA_(query, ref, K, scale)
    oi, od = O.sv_knn(query, ref, K, scale)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gd, od, equal_nan=True)
    return gi, gd


@pytest.mark.parametrize("K", [1, 5, 11])
@pytest.mark.parametrize("integer", [True, False])
def test_random_centroids(K, integer):
    rng = np.random.default_rng(10 * K + integer)
    q = rng.uniform(0, 200, (3, 3000))
    r = rng.uniform(0, 200, (3, 5000))
    if integer:  # many exact ties
        q, r = np.round(q), np.round(r)
    print("svknn random K=%d integer=%d" % (K, integer), flush=True)
    _check(q.astype(np.float32).reshape(-1), r.astype(np.float32).reshape(-1), K)


def _raster(xyz):
    """Centroids sorted by (z, y, x), as a segmentation's raster-order labels."""
    order = np.lexsort((xyz[0], xyz[1], xyz[2]))
    return xyz[:, order]


def test_raster_ordered_self_and_forward_search():
    fr = synth.config1(seed=9, n_cells=150, target_voxels=30000)
    c = _raster(fr.centroids.reshape(3, fr.S))
    nxt = _raster(c + np.random.default_rng(1).normal(0, 1.5, c.shape))
    print("svknn raster S=%d" % fr.S, flush=True)
    a = c.astype(np.float32).reshape(-1)
    b = nxt.astype(np.float32).reshape(-1)
    gi, gd = _check(a, a, 11)
    assert np.all(gd.reshape(11, fr.S).min(0) == 0.0)  # each finds itself
    _check(a, b, 10)
    _check(b, a, 10)


def test_nan_centroids_follow_the_scan():
    rng = np.random.default_rng(3)
    q = rng.uniform(0, 50, (3, 500)).astype(np.float32)
    r = rng.uniform(0, 50, (3, 700)).astype(np.float32)
    q[1, 17] = np.nan
    r[2, 300] = np.nan
    _check(q.reshape(-1), r.reshape(-1), 5)


def test_fewer_references_than_slots_and_edges():
    rng = np.random.default_rng(4)
    q = rng.uniform(0, 9, (3, 70)).astype(np.float32).reshape(-1)
    r = rng.uniform(0, 9, (3, 3)).astype(np.float32).reshape(-1)
    gi, gd = _check(q, r, 8)
    assert np.all(gi.reshape(8, 70)[3:] == -1)
    gi, gd = em.knnCUDA_(q, np.zeros(0, np.float32), 4, SCALE)
    assert np.all(gd == np.float32(1e32)) and np.all(gi == -1)
    gi, gd = em.knnCUDA_(np.zeros(0, np.float32), r, 4, SCALE)
    assert gi.size == 0 and gd.size == 0
    gi, gd = em.knnCUDA_(q, r, 2, SCALE, with_dist=False)
    assert gd is None and np.array_equal(gi, O.sv_knn(q, r, 2, SCALE)[0])
    with pytest.raises(em.TgmmError):
        em.knnCUDA_(q, r, 12, SCALE)
