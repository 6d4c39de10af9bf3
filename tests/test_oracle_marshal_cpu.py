"""CPU tests that pin the oracle's input marshalling (orc_marshal_frame):
raw uint16 supervoxel voxels to the EM's float inputs
(mainTrackingGaussianMixtureModel.cpp:598-669, 774-870).

Pins:
  * centroids: bit-equal to the reference's own
    supervoxel::weightedGaussianStatistics<uint16> (supervoxel.cpp:1410-1449),
    golden vectors in tests/golden/sv_centroid_vectors.npz (float sums beyond
    2^24, an all-zero supervoxel, single voxels, voxels in random order) and
    the live reference twin when it is built;
  * coordinates: numpy unravel_index in Fortran order (x fastest), as mylib's
    coord4idx (array.p:557-576);
  * intensities: Scale_Array_To_Range to [0, 1] over the whole volume (mylib
    array.p:2708-2820, not buildable here: its .p sources need mylib's code
    generator). The brightest voxel maps to exactly 1, the background and
    the volume minimum to exactly 0, every value lies within one float ulp of
    the exact rational (v - m) / (M - m), order is preserved, and a constant
    volume is left unscaled. The last bit (double product then float
    rounding) is restated, not pinned;
  * the local background subtraction is uint16 arithmetic (wraps below thr).
"""
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "sv_centroid_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    z = np.load(GOLD)
    return {k: z[k] for k in z.files}


def _marshal(g, thr=None, dims=None):
    return O.marshal_frame(g["idx"], g["raw"], g["ptr"], g["thr"] if thr is None else thr,
                           (g["dims"] if dims is None else dims).astype(np.int64))


def test_centroids_match_reference_golden(gold):
    q, im, c = _marshal(gold)
    S = gold["ptr"].size - 1
    got = c.reshape(3, S).T
    assert np.array_equal(got, gold["centroid"])
    assert np.all(gold["centroid"][gold["intensity"] == 0] == 0)
    assert gold["intensity"].max() > 2 ** 24  # rounding of the float sums is exercised


def test_centroids_match_live_reference_twin():
    T = O.reftwin()
    if T is None or not hasattr(T, "ref_sv_centroid"):
        pytest.skip("reference twin not built (needs the reference tree)")
    rng = np.random.default_rng(7)
    dims = np.array([40, 33, 11], np.uint64)
    vol = rng.integers(0, 65536, int(np.prod(dims))).astype(np.uint16)
    perm = rng.permutation(vol.size).astype(np.uint64)
    sizes = rng.integers(1, 300, 60)
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    idx = perm[: ptr[-1]]
    raw = vol[idx.astype(np.int64)]
    _, _, c = O.marshal_frame(idx, raw, ptr, np.zeros(60, np.uint16), dims.astype(np.int64))
    for s in range(60):
        want = np.zeros(3, np.float32)
        it = np.zeros(1, np.float32)
        seg = np.ascontiguousarray(idx[ptr[s]:ptr[s + 1]])
        T.ref_sv_centroid(vol.ctypes.data, dims.ctypes.data, seg.ctypes.data, seg.size,
                          O.p(want, O._fp), O.p(it, O._fp))
        assert np.array_equal(c[[s, s + 60, s + 120]], want), s


def test_coordinates_are_fortran_order_indices(gold):
    q, _, _ = _marshal(gold)
    N = gold["idx"].size
    x, y, z = np.unravel_index(gold["idx"].astype(np.int64), tuple(gold["dims"].astype(int)),
                               order="F")
    assert np.array_equal(q[:N], x.astype(np.float32))
    assert np.array_equal(q[N:2 * N], y.astype(np.float32))
    assert np.array_equal(q[2 * N:], z.astype(np.float32))


def _exact_scaled(v, m, M):
    return Fraction(int(v) - int(m), int(M) - int(m))


def _ulp_close(x, fr):
    x = np.float32(x)
    lo, hi = np.nextafter(x, np.float32(-1)), np.nextafter(x, np.float32(2))
    return Fraction(float(lo)) <= fr <= Fraction(float(hi))


def test_intensities_scaled_over_the_volume_with_background(gold):
    _, im, _ = _marshal(gold)
    sub = (gold["raw"].astype(np.int64) - np.repeat(gold["thr"].astype(np.int64),
                                                    np.diff(gold["ptr"]))) % 65536
    M = sub.max()
    # background voxels exist (N < volume size), so the minimum is 0
    assert gold["idx"].size < np.prod(gold["dims"])
    assert im[sub == M].min() == 1.0 and im[sub == M].max() == 1.0
    assert np.all(im[sub == 0] == 0.0)
    order = np.argsort(sub, kind="stable")
    assert np.all(np.diff(im[order]) >= 0)
    rng = np.random.default_rng(3)
    for i in rng.choice(sub.size, 2000, replace=False):
        assert _ulp_close(im[i], _exact_scaled(sub[i], 0, M)), (i, sub[i], im[i])


def test_full_coverage_uses_the_listed_minimum():
    dims = np.array([4, 3, 2], np.int64)
    n = int(np.prod(dims))
    idx = np.arange(n, dtype=np.uint64)[::-1].copy()
    raw = (np.arange(n) * 37 + 500).astype(np.uint16)
    ptr = np.array([0, 10, n], np.int64)
    thr = np.array([100, 50], np.uint16)
    _, im, _ = O.marshal_frame(idx, raw, ptr, thr, dims)
    sub = raw.astype(np.int64) - np.repeat(thr.astype(np.int64), [10, n - 10])
    m, M = sub.min(), sub.max()
    assert m > 0
    assert np.all(im[sub == m] == 0.0) and np.all(im[sub == M] == 1.0)
    for v, x in zip(sub, im):
        assert _ulp_close(x, _exact_scaled(v, m, M))


def test_constant_volume_is_not_scaled():
    dims = np.array([2, 2, 2], np.int64)
    idx = np.arange(8, dtype=np.uint64)
    _, im, _ = O.marshal_frame(idx, np.full(8, 900, np.uint16), np.array([0, 8], np.int64),
                               np.array([200], np.uint16), dims)
    assert np.all(im == 700.0)
    # an all-background frame with a zero-valued supervoxel: M == m == 0
    _, im, c = O.marshal_frame(idx[:3], np.zeros(3, np.uint16), np.array([0, 3], np.int64),
                               np.array([0], np.uint16), dims)
    assert np.all(im == 0.0) and np.all(c == 0.0)


def test_background_subtraction_is_uint16_arithmetic():
    dims = np.array([8, 1, 1], np.int64)
    idx = np.arange(3, dtype=np.uint64)
    _, im, _ = O.marshal_frame(idx, np.array([5, 7, 9], np.uint16), np.array([0, 3], np.int64),
                               np.array([7], np.uint16), dims)
    # 5 - 7 wraps to 65534 (the reference subtracts in unsigned short)
    assert im[0] == 1.0 and im[1] == 0.0
    assert _ulp_close(im[2], Fraction(2, 65534))


def test_index_outside_the_volume_is_an_error():
    with pytest.raises(ValueError):
        O.marshal_frame(np.array([24], np.uint64), np.ones(1, np.uint16),
                        np.array([0, 1], np.int64), np.zeros(1, np.uint16),
                        np.array([4, 3, 2], np.int64))
