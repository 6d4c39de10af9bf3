"""GPU parity of the entry points beside the EM at the full size of the
single-frame workload (10,000 Gaussians, 4.86M voxels, 108,573 supervoxels):
the input marshalling bitwise against the oracle on the whole frame, the
local log-likelihood of every Gaussian's supervoxel list against the oracle,
and the supervoxel neighbour search bitwise against the oracle on sampled
queries (each query's scan is independent of the others)."""
import numpy as np
import pytest

import oracle_lib as O
from helpers import raw_from_frame
from tgmm_amd import em, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def frame():
    f = synth.config2()
    print("full-size frame N=%d S=%d G=%d" % (f.N, f.S, f.G), flush=True)
    return f


def test_marshal_whole_frame_bitwise(frame):
    args = raw_from_frame(frame, (2048, 2048, 400), seed=1)
    q, im, c, p = em.marshal_frame(*args)
    N, S = frame.N, frame.S
    want = O.marshal_frame(*args)
    assert np.array_equal(q.numpy(np.float32)[: 3 * N], want[0])
    assert np.array_equal(im.numpy(np.float32)[:N], want[1])
    assert np.array_equal(c.numpy(np.float32)[: 3 * S], want[2])
    # the coordinates are the frame's own
    assert np.array_equal(want[0], frame.query)


def test_local_loglik_all_gaussian_lists(frame):
    with em.Batch([frame], em.make_opts(K=5, max_iter=3, tol=-1.0)) as b:
        b.solve(em.SOLVE_GRAPH)
        res = b.fetch()[0]
    gm = np.ascontiguousarray(res["gm"])
    ind = res["ind"].astype(np.int32)
    lists = [list(gm["supervoxelIdx"][k][: gm["supervoxelNum"][k]]) for k in range(frame.G)]
    got = em.calculateLocalLikelihood(
        [], lists, em.DeviceArray(frame.query), em.DeviceArray(frame.img), None,
        em.DeviceArray(ind), gm, em.DeviceArray(frame.sv_ptr.astype(np.int64)), frame.N,
        frame.G, frame.S)
    want = O.local_loglik(frame, gm, ind, 5, lists)
    err = max(abs(a - b) / max(1.0, abs(b)) for a, b in zip(got, want))
    assert err <= 1e-10, err


def test_supervoxel_knn_sampled_queries(frame):
    c = frame.centroids.reshape(3, frame.S).astype(np.float32)
    c = np.ascontiguousarray(c[:, np.lexsort((c[0], c[1], c[2]))])
    a = c.reshape(-1)
    ind, dist = em.knnCUDA_(a, a, 11, (1.0, 1.0, 5.0))
    sample = np.sort(np.random.default_rng(8).choice(frame.S, 1500, replace=False))
    qs = np.ascontiguousarray(c[:, sample].reshape(-1))
    oi, od = O.sv_knn(qs, a, 11, (1.0, 1.0, 5.0))
    assert np.array_equal(ind.reshape(11, frame.S)[:, sample].reshape(-1), oi)
    assert np.array_equal(dist.reshape(11, frame.S)[:, sample].reshape(-1), od)
