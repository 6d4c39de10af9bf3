"""CPU tests that pin the oracle's supervoxel nearest neighbours (orc_sv_knn,
the reference's knnCUDA_, temporalLogicalRules/knnCUDA/knnCuda.cu:43-232).

The reference is CUDA and cannot be built here. Pins:
  * the slot contents: every slot's distance is the numpy float32 distance of
    its index ((q - r) * s squared, summed x, y, z), and the slots hold the K
    smallest distances as a multiset (numpy brute force);
  * the slot layout, worked by hand from the replacement rule (a reference not
    above the threshold replaces the first slot holding the largest
    distance): increasing and decreasing distance sequences, ties (a later
    reference at the threshold evicts), fewer references than slots, no
    references (distances 1e32, indices untouched), anisotropic scale, and a
    query set searched against itself (each query finds itself at 0).
"""
import numpy as np
import pytest

import oracle_lib as O
from helpers import scan_replay


def np_dist(query, ref, scale):
    q = query.reshape(3, -1).astype(np.float32)
    r = ref.reshape(3, -1).astype(np.float32)
    s = np.asarray(scale, np.float32)
    d = np.zeros((q.shape[1], r.shape[1]), np.float32)
    for a in range(3):
        t = (q[a][:, None] - r[a][None, :]) * s[a]
        d = d + t * t
    return d


def line_refs(xs):
    """References on the x axis at xs (query at the origin)."""
    r = np.zeros((3, len(xs)), np.float32)
    r[0] = xs
    return r.reshape(-1)


ORIGIN = np.zeros(3, np.float32)


@pytest.mark.parametrize("K", [1, 5, 11])
def test_slots_hold_the_k_smallest_distances(K):
    rng = np.random.default_rng(K)
    Q, R = 300, 900
    query = (rng.uniform(0, 60, (3, Q))).astype(np.float32).reshape(-1)
    ref = np.round(rng.uniform(0, 60, (3, R))).astype(np.float32).reshape(-1)  # ties
    scale = (1.0, 1.0, 5.0)
    ind, dist = O.sv_knn(query, ref, K, scale)
    D = np_dist(query, ref, scale)
    ind, dist = ind.reshape(K, Q), dist.reshape(K, Q)
    for q in range(Q):
        assert np.array_equal(dist[:, q], D[q, ind[:, q]])
        assert np.array_equal(np.sort(dist[:, q]), np.sort(D[q])[:K])
        assert len(set(ind[:, q])) == K


def test_increasing_distances_fill_slots_in_order():
    ind, dist = O.sv_knn(ORIGIN, line_refs([1, 2, 3, 4, 5]), 3, (1, 1, 1))
    assert list(ind) == [0, 1, 2] and list(dist) == [1, 4, 9]


def test_decreasing_distances_replace_the_largest_slot():
    # slots (25,0) (16,1) (9,2); 4 evicts 25 at slot 0; 1 evicts 16 at slot 1
    ind, dist = O.sv_knn(ORIGIN, line_refs([5, 4, 3, 2, 1]), 3, (1, 1, 1))
    assert list(ind) == [3, 4, 2] and list(dist) == [4, 1, 9]


def test_a_later_reference_at_the_threshold_evicts_the_first_largest_slot():
    # all at distance 4: slots (0, 1); then 2 and 3 each replace slot 0
    ind, dist = O.sv_knn(ORIGIN, line_refs([2, -2, 2, -2]), 2, (1, 1, 1))
    assert list(ind) == [3, 1] and list(dist) == [4, 4]


def test_fewer_references_than_slots():
    ind, dist = O.sv_knn(ORIGIN, line_refs([3, 1]), 4, (1, 1, 1))
    assert list(ind) == [0, 1, -1, -1]
    assert list(dist) == [9, 1, np.float32(1e32), np.float32(1e32)]


def test_no_references_sets_distances_only():
    ind, dist = O.sv_knn(np.zeros(6, np.float32), np.zeros(0, np.float32), 3, (1, 1, 1))
    assert np.all(dist == np.float32(1e32))
    assert np.all(ind == -7)  # the wrapper's fill: left as given


def test_scale_weights_the_axes():
    ref = np.array([3, 0, 0, 0, 0, 1], np.float32)  # (3,0,0) and (0,0,1) in SoA
    assert O.sv_knn(ORIGIN, ref, 1, (1, 1, 1))[0][0] == 1
    assert O.sv_knn(ORIGIN, ref, 1, (1, 1, 5))[0][0] == 0  # z distance 25 > 9


def test_self_search_finds_itself():
    rng = np.random.default_rng(5)
    xyz = rng.uniform(0, 100, (3, 400)).astype(np.float32).reshape(-1)
    ind, dist = O.sv_knn(xyz, xyz, 6, (1, 1, 5))
    ind, dist = ind.reshape(6, 400), dist.reshape(6, 400)
    for q in range(400):
        assert q in ind[:, q]
        assert dist[list(ind[:, q]).index(q), q] == 0.0


# ------------------------------------------------------------------ NaN --
replay = scan_replay


def test_nan_partial_sums_leave_the_reference():
    # K = 1, query at the origin, threshold 1 after reference 0
    #   ref 1 (5, 0, NaN): x term 25 > 1, left before the NaN is reached
    #   ref 2 (0.5, 3, NaN): x + y partial sum 9.25 > 1, left
    #   ref 3 (0.5, 0.5, NaN): partial sums 0.25, 0.5; the NaN sum is taken
    ref = np.array([[1, 5, 0.5, 0.5], [0, 0, 3, 0.5], [0, np.nan, np.nan, np.nan]],
                   np.float32).reshape(-1)
    for n, want in ((2, 0), (3, 0), (4, 3)):
        r = ref.reshape(3, 4)[:, :n].reshape(-1)
        ind, dist = O.sv_knn(ORIGIN, r, 1, (1, 1, 1))
        assert ind[0] == want, (n, ind)
        assert np.array_equal(ind, replay(ORIGIN, r, 1, (1, 1, 1))[0])


def test_nan_in_slot_zero_takes_the_next_reference():
    # K = 2: ref 0 (1, 0, NaN) goes to slot 0 with a NaN distance, so the
    # threshold is NaN (findMaxPosition never moves off slot 0); ref 1, far
    # away, is then taken into slot 0; the threshold becomes 1e32 (slot 1) and
    # ref 2 fills slot 1
    ref = np.array([[1, 100, 2], [0, 0, 0], [np.nan, 0, 0]], np.float32).reshape(-1)
    ind, dist = O.sv_knn(ORIGIN, ref, 2, (1, 1, 1))
    assert list(ind) == [1, 2] and list(dist) == [10000, 4]
    ri, rd = replay(ORIGIN, ref, 2, (1, 1, 1))
    assert np.array_equal(ind, ri) and np.array_equal(dist, rd)


@pytest.mark.parametrize("K", [1, 3, 5])
def test_nan_centroids_match_the_plain_replay(K):
    rng = np.random.default_rng(100 + K)
    q = rng.uniform(0, 20, (3, 40)).astype(np.float32)
    r = rng.uniform(0, 20, (3, 150)).astype(np.float32)
    for a, i in ((0, 3), (1, 17), (2, 30), (2, 0)):
        r[a, i] = np.nan
    q[1, 5] = np.nan
    q[2, 9] = np.nan
    q[0, 11] = np.nan
    oi, od = O.sv_knn(q.reshape(-1), r.reshape(-1), K, (1, 1, 5))
    ri, rd = replay(q.reshape(-1), r.reshape(-1), K, (1, 1, 5))
    assert np.array_equal(oi, ri)
    assert np.array_equal(od, rd, equal_nan=True)


def test_plain_replay_matches_the_oracle_without_nan():
    rng = np.random.default_rng(7)
    q = np.round(rng.uniform(0, 12, (3, 30))).astype(np.float32).reshape(-1)
    r = np.round(rng.uniform(0, 12, (3, 120))).astype(np.float32).reshape(-1)  # ties
    for K in (1, 4, 11):
        oi, od = O.sv_knn(q, r, K, (1, 1, 5))
        ri, rd = replay(q, r, K, (1, 1, 5))
        assert np.array_equal(oi, ri) and np.array_equal(od, rd)
