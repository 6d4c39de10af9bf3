"""GPU parity of the local log-likelihood (calculateLocalLikelihood,
GMEMupdateCUDA.cu:2952-3050) through the C ABI: against the oracle on the
tracker's per-Gaussian supervoxel lists and on random lists with duplicates
and empty lists, against the EM's own log-likelihood for the list of all
supervoxels, and its argument errors."""
import numpy as np
import pytest

import oracle_lib as O
from helpers import rel_err
from tgmm_amd import em, synth

pytestmark = pytest.mark.gpu

K = 5


def _close(a, b):
    return abs(a - b) <= 1e-10 * max(1.0, abs(b))


@pytest.fixture(scope="module")
def solved():
    """One frame through the drop-in EM; the device arrays stay allocated."""
    fr = synth.config1(seed=777, n_cells=40, target_voxels=12000)
    print("local LL frame N=%d S=%d G=%d" % (fr.N, fr.S, fr.G), flush=True)
    em.regularizePrecisionMatrixConstants.setConstants(0.02, 0.2222, 9.0)
    scale = np.array([1, 1, 5], np.float32)
    d = em.GMEMinitializeMemoryWithSupervoxels(fr.query, fr.img, scale, fr.N, fr.S,
                                               fr.centroids, fr.sv_ptr)
    em.copyScaleToDvice(scale)
    gm = fr.gm.copy()
    em.GMEMvariationalInferenceCUDAWithSupervoxels(d, gm, fr.N, fr.G, fr.S, 6, -1.0)
    ind = em.copy_device_to_host(d.indCUDA, K * fr.S * 4, np.int32)
    yield fr, d, gm, ind
    em.GMEMreleaseMemoryWithSupervoxels(d)


def _gpu(solved, lists, prefill=None):
    fr, d, gm, _ = solved
    out = list(prefill) if prefill is not None else []
    return em.calculateLocalLikelihood(out, lists, d.queryCUDA, d.imgDataCUDA, d.rnkCUDA,
                                       d.indCUDA, gm, d.labelListPtrCUDA, fr.N, fr.G, fr.S)


def test_tracker_lists_match_oracle(solved):
    # each Gaussian's supervoxelIdx[0 .. supervoxelNum), as the division test
    # builds them (mainTrackingGaussianMixtureModel.cpp:1312-1317)
    fr, d, gm, ind = solved
    lists = [list(gm["supervoxelIdx"][k][: gm["supervoxelNum"][k]]) for k in range(fr.G)]
    assert sum(len(l) for l in lists) > fr.S // 2
    got = _gpu(solved, lists)
    want = O.local_loglik(fr, gm, ind, K, lists)
    assert len(got) == len(lists)
    bad = [(i, g, w) for i, (g, w) in enumerate(zip(got, want)) if not _close(g, w)]
    assert not bad, bad[:5]


def test_random_lists_duplicates_and_empty(solved):
    fr, d, gm, ind = solved
    rng = np.random.default_rng(5)
    lists = [list(rng.integers(0, fr.S, rng.integers(0, 40))) for _ in range(300)]
    lists[2] = []
    lists[7] = []
    lists[8] = [3, 3, 3]
    lists[9] = list(range(fr.S))  # one long list (one wave sums it)
    prefill = [-123.5] * 5  # shorter than the lists: kept where empty, extended with 0
    got = _gpu(solved, lists, prefill)
    want = O.local_loglik(fr, gm, ind, K, lists, out=prefill + [0.0] * (len(lists) - 5))
    for i, (g, w) in enumerate(zip(got, want)):
        if not lists[i]:
            assert g == w, i
        else:
            assert _close(g, w), (i, g, w)
    assert abs(got[8] - 3 * _gpu(solved, [[3]])[0]) <= 1e-12 * abs(got[8])


def test_deterministic_and_independent_of_the_other_lists(solved):
    fr = solved[0]
    rng = np.random.default_rng(6)
    lists = [list(rng.integers(0, fr.S, 12)) for _ in range(64)]
    a = _gpu(solved, lists)
    b = _gpu(solved, lists)
    assert a == b
    perm = rng.permutation(len(lists))
    c = _gpu(solved, [lists[i] for i in perm])
    assert [a[i] for i in perm] == c
    assert _gpu(solved, [lists[17]])[0] == a[17]


def test_all_supervoxels_equal_the_em_loglik():
    fr = synth.config1(seed=778, n_cells=30, target_voxels=8000)
    with em.Batch([fr], em.make_opts(K=K, max_iter=5)) as b:
        b.solve(em.SOLVE_GRAPH)
        res = b.fetch()[0]
    q, im = em.DeviceArray(fr.query), em.DeviceArray(fr.img)
    ind = em.DeviceArray(res["ind"].astype(np.int32))
    ptr = em.DeviceArray(fr.sv_ptr.astype(np.int64))
    got = em.calculateLocalLikelihood([], [list(range(fr.S))], q, im, None, ind, res["gm"], ptr,
                                      fr.N, fr.G, fr.S)
    # the EM's value comes from the FP32 voxel path, this one from FP64
    assert rel_err(got[0], res["ll"]) < 1e-6, (got[0], res["ll"])


def test_all_empty_and_no_lists(solved):
    assert _gpu(solved, [[], []], [1.0, 2.0]) == [1.0, 2.0]
    assert _gpu(solved, []) == []


def test_bad_index_is_rejected(solved):
    fr = solved[0]
    with pytest.raises(em.TgmmError):
        _gpu(solved, [[0, fr.S]])
    with pytest.raises(em.TgmmError):
        _gpu(solved, [[-1]])
