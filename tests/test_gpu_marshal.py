"""GPU parity of the input marshalling (tgmm_marshal_frame,
tgmm_GMEMinitializeMemoryFromRaw) through the C ABI: bit-identical to the
oracle (coordinates, scaled intensities, centroids) on the reference's golden
supervoxels and on small edge cases, from host, device and whole-volume
inputs, and an EM fed from a raw frame equal to the same EM fed the
oracle-marshalled host arrays."""
import os

import numpy as np
import pytest

import oracle_lib as O
from helpers import raw_from_frame, records_equal
from tgmm_amd import em, synth

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "sv_centroid_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    z = np.load(GOLD)
    return {k: z[k] for k in z.files}


def _gpu(idx, raw, ptr, thr, dims, volume=None):
    q, im, c, p = em.marshal_frame(idx, raw, ptr, thr, dims, volume)
    N, S = int(ptr[-1]), len(ptr) - 1
    return (q.numpy(np.float32)[: 3 * N], im.numpy(np.float32)[:N],
            c.numpy(np.float32)[: 3 * S], p.numpy(np.int64)[: S + 1])


def _same(got, want):
    for g, w, name in zip(got, want, ("query", "img", "centroids")):
        assert np.array_equal(g, w), name


def test_golden_supervoxels(gold):
    args = (gold["idx"], gold["raw"], gold["ptr"], gold["thr"], gold["dims"].astype(np.int64))
    print("golden marshal N=%d S=%d" % (gold["idx"].size, gold["ptr"].size - 1), flush=True)
    got = _gpu(*args)
    _same(got, O.marshal_frame(*args))
    S = gold["ptr"].size - 1
    assert np.array_equal(got[2].reshape(3, S).T, gold["centroid"])
    assert np.array_equal(got[3], gold["ptr"])


def test_volume_and_device_inputs(gold):
    dims = gold["dims"].astype(np.int64)
    vol = np.zeros(int(np.prod(dims)), np.uint16)
    vol[gold["idx"].astype(np.int64)] = gold["raw"]
    ref = _gpu(gold["idx"], gold["raw"], gold["ptr"], gold["thr"], dims)
    dvol = em.DeviceArray(vol)
    _same(_gpu(gold["idx"], None, gold["ptr"], gold["thr"], dims, volume=dvol), ref)
    didx, draw, dthr = (em.DeviceArray(gold["idx"]), em.DeviceArray(gold["raw"]),
                        em.DeviceArray(gold["thr"]))
    _same(_gpu(didx, draw, gold["ptr"], dthr, dims), ref)


@pytest.mark.parametrize("case", ["full", "constant", "wrap", "empty_sv"])
def test_edge_cases(case):
    dims = np.array([4, 3, 2], np.int64)
    n = 24
    if case == "full":  # no background: the listed minimum is the volume's
        idx = np.arange(n, dtype=np.uint64)[::-1].copy()
        raw = (np.arange(n) * 37 + 500).astype(np.uint16)
        ptr = np.array([0, 10, n], np.int64)
        thr = np.array([100, 50], np.uint16)
    elif case == "constant":
        idx = np.arange(n, dtype=np.uint64)
        raw = np.full(n, 900, np.uint16)
        ptr = np.array([0, n], np.int64)
        thr = np.array([200], np.uint16)
    elif case == "wrap":
        idx = np.arange(3, dtype=np.uint64)
        raw = np.array([5, 7, 9], np.uint16)
        ptr = np.array([0, 3], np.int64)
        thr = np.array([7], np.uint16)
    else:  # empty and all-zero supervoxels
        idx = np.array([3, 4, 9], np.uint64)
        raw = np.array([0, 0, 40], np.uint16)
        ptr = np.array([0, 0, 2, 2, 3], np.int64)
        thr = np.zeros(4, np.uint16)
    print("marshal edge case %s" % case, flush=True)
    _same(_gpu(idx, raw, ptr, thr, dims), O.marshal_frame(idx, raw, ptr, thr, dims))


def test_index_outside_the_volume_is_an_error():
    with pytest.raises(em.TgmmError):
        _gpu(np.array([0, 24], np.uint64), np.ones(2, np.uint16), np.array([0, 2], np.int64),
             np.zeros(1, np.uint16), np.array([4, 3, 2], np.int64))
    # nothing is left behind: a good frame right after works
    idx = np.arange(4, dtype=np.uint64)
    args = (idx, np.arange(4).astype(np.uint16), np.array([0, 4], np.int64),
            np.zeros(1, np.uint16), np.array([4, 3, 2], np.int64))
    _same(_gpu(*args), O.marshal_frame(*args))


def test_em_from_raw_equals_em_from_marshalled_host_arrays():
    fr = synth.config1(seed=4321, n_cells=40, target_voxels=12000)
    raw_args = raw_from_frame(fr, (512, 512, 100), seed=3)
    q, im, c = O.marshal_frame(*raw_args)
    print("EM from raw: N=%d S=%d G=%d" % (fr.N, fr.S, fr.G), flush=True)
    em.regularizePrecisionMatrixConstants.setConstants(0.02, 0.2222, 9.0)
    scale = np.array([1, 1, 5], np.float32)
    em.copyScaleToDvice(scale)
    outs = []
    for source in ("raw", "host"):
        if source == "raw":
            d = em.GMEMinitializeMemoryFromRaw(raw_args, scale)
        else:
            d = em.GMEMinitializeMemoryWithSupervoxels(q, im, scale, fr.N, fr.S, c, fr.sv_ptr)
        gm = fr.gm.copy()
        em.GMEMvariationalInferenceCUDAWithSupervoxels(d, gm, fr.N, fr.G, fr.S, 5, -1.0)
        ind = em.copy_device_to_host(d.indCUDA, 5 * fr.S * 4, np.int32)
        img = em.copy_device_to_host(d.imgDataCUDA, fr.N * 4, np.float32)
        em.GMEMreleaseMemoryWithSupervoxels(d)
        outs.append((gm, ind, img))
    (g1, i1, m1), (g2, i2, m2) = outs
    assert np.array_equal(m1, im) and np.array_equal(m2, im)
    assert records_equal(g1, g2)
    assert np.array_equal(i1, i2)


def test_batch_from_raw_frames_equals_batch_from_host_arrays():
    # a batch whose frames arrive raw (marshalled on the device) against one
    # built from the oracle-marshalled float arrays
    fa = synth.config1(seed=61, n_cells=40, target_voxels=10000)
    fb = synth.config1(seed=62, n_cells=25, target_voxels=6000)
    raws = [raw_from_frame(f, (512, 512, 100), seed=i) for i, f in enumerate((fa, fb))]
    hosts = []
    for f, r in zip((fa, fb), raws):
        q, im, c = O.marshal_frame(*r)
        hosts.append(synth.Frame(q, im, c, f.sv_ptr, f.gm))
    opts = em.make_opts(K=5, max_iter=6, tol=-1.0)
    with em.Batch.from_raw(raws, [fa.gm, fb.gm], opts) as b:
        b.solve(em.SOLVE_GRAPH)
        got = b.fetch()
    with em.Batch(hosts, opts) as b:
        b.solve(em.SOLVE_GRAPH)
        want = b.fetch()
    for g, w in zip(got, want):
        assert records_equal(g["gm"], w["gm"])
        assert np.array_equal(g["rnk"], w["rnk"]) and np.array_equal(g["ind"], w["ind"])
        assert g["ll"] == w["ll"] and g["iterations"] == w["iterations"]
