"""CPU tests that pin the oracle's local log-likelihood (calculateLocalLikelihood,
GMEMupdateCUDA.cu:2952-3050, kernel :1989-2103), the input of the tracker's
cell-division likelihood-ratio test (mainTrackingGaussianMixtureModel.cpp:
1336-1375).

Pins:
  * a mixture of Gaussians evaluated by scipy (multivariate normal log-density
    and logsumexp): the list value is sum over listed supervoxels, voxels v,
    of img_v * log sum_j (alpha_j / sum_alpha) N(x_v; m_j, (nu_j W_j)^-1),
    with sum_alpha over every record, not only the candidates;
  * the list of all supervoxels reproduces orc_loglik (pinned by the same
    closed form and used by the EM's convergence test), and the per-Gaussian
    supervoxel lists the tracker builds partition it;
  * a supervoxel listed twice counts twice; an empty list leaves its output as
    given (the reference skips it after resizing the output vector);
  * a candidate with an indefinite precision makes the density NaN, and the
    voxel then adds nothing (the reference's `auxLikelihood > 0` test).
"""
import numpy as np
import pytest
from scipy.special import logsumexp
from scipy.stats import multivariate_normal

import oracle_lib as O
from tgmm_amd import synth
from tgmm_amd.records import GMM_DTYPE, w_full


def mixture_frame(rng, S=12, G=6, K=3):
    sizes = rng.integers(1, 9, S)
    ptr = np.zeros(S + 1, np.int64)
    ptr[1:] = np.cumsum(sizes)
    N = int(ptr[-1])
    query = (rng.normal(size=(3, N)) * 2 + 40).reshape(-1).astype(np.float32)
    img = rng.uniform(0.1, 1.0, N).astype(np.float32)
    gm = np.zeros(G, GMM_DTYPE)
    gm["m_k"] = 40 + rng.normal(size=(G, 3))
    for k in range(G):
        A = rng.normal(size=(3, 3))
        W = A @ A.T * 0.05 + np.eye(3) * 0.05
        gm["W_k"][k] = [W[0, 0], W[0, 1], W[0, 2], W[1, 1], W[1, 2], W[2, 2]]
    gm["nu_k"] = rng.uniform(4, 12, G)
    gm["alpha_k"] = rng.uniform(0.5, 5, G)
    gm["alpha_k"][G - 1] = 50.0  # never a candidate below, counts in sum_alpha
    ind = np.full((K, S), -1, np.int32)
    for s in range(S):
        n = rng.integers(1, K + 1)
        ind[:n, s] = rng.choice(G - 1, n, replace=False)
    frame = synth.Frame(query, img, np.zeros(3 * S, np.float32), ptr, gm)
    return frame, ind.reshape(-1), K


def scipy_sv_ll(frame, ind, K, s):
    N, S = frame.N, frame.S
    X = frame.query.reshape(3, N).T.astype(np.float64)[frame.sv_ptr[s]:frame.sv_ptr[s + 1]]
    gm = frame.gm
    tot = gm["alpha_k"].sum()
    terms = []
    for j in range(K):
        k = ind[s + j * S]
        if k < 0:
            continue
        prec = gm["nu_k"][k] * w_full(gm["W_k"][k])
        terms.append(np.log(gm["alpha_k"][k] / tot)
                     + multivariate_normal(mean=gm["m_k"][k], cov=np.linalg.inv(prec)).logpdf(X))
    w = frame.img[frame.sv_ptr[s]:frame.sv_ptr[s + 1]].astype(np.float64)
    return float(np.sum(w * logsumexp(np.stack(terms), axis=0)))


def test_local_loglik_mixture_closed_form():
    rng = np.random.default_rng(21)
    frame, ind, K = mixture_frame(rng)
    lists = [[0], [3, 4, 5], [11, 2, 2], [], list(range(frame.S))]
    out = O.local_loglik(frame, frame.gm, ind, K, lists, out=[0, 0, 0, 7.25, 0])
    for i, lst in enumerate(lists):
        if not lst:
            assert out[i] == 7.25  # left as given
            continue
        want = sum(scipy_sv_ll(frame, ind, K, s) for s in lst)
        assert abs(out[i] - want) <= 1e-12 * abs(want), (i, out[i], want)


def test_duplicates_count_twice_and_lists_are_additive():
    rng = np.random.default_rng(22)
    frame, ind, K = mixture_frame(rng)
    a, b = [1, 7, 9], [0, 4]
    o = O.local_loglik(frame, frame.gm, ind, K, [a, b, a + b, a + a])
    assert abs(o[2] - (o[0] + o[1])) <= 1e-13 * abs(o[2])
    assert abs(o[3] - 2 * o[0]) <= 1e-13 * abs(o[3])


def test_sum_alpha_is_over_all_records():
    rng = np.random.default_rng(23)
    frame, ind, K = mixture_frame(rng)
    lists = [[2, 3], [5]]
    base = O.local_loglik(frame, frame.gm, ind, K, lists)
    # scaling every alpha leaves the ratios alpha_k / sum_alpha unchanged
    g2 = frame.gm.copy()
    g2["alpha_k"] *= 3.0
    assert np.allclose(O.local_loglik(frame, g2, ind, K, lists), base, rtol=1e-13, atol=0)
    # raising only the non-candidate's alpha shifts each list by
    # (sum of its intensities) * log(old sum / new sum)
    g3 = frame.gm.copy()
    g3["alpha_k"][-1] += 10.0
    s0, s1 = frame.gm["alpha_k"].sum(), g3["alpha_k"].sum()
    shifted = O.local_loglik(frame, g3, ind, K, lists)
    for i, lst in enumerate(lists):
        w = sum(frame.img[frame.sv_ptr[s]:frame.sv_ptr[s + 1]].astype(np.float64).sum()
                for s in lst)
        assert abs(shifted[i] - (base[i] + w * np.log(s0 / s1))) <= 1e-12 * abs(base[i])


def test_indefinite_precision_drops_the_voxel():
    rng = np.random.default_rng(24)
    frame, ind, K = mixture_frame(rng)
    ind2 = ind.reshape(K, frame.S).copy()
    ind2[:, 0] = [0, 1, -1]
    g = frame.gm.copy()
    g["W_k"][1] = [0.1, 0.0, 0.0, -0.2, 0.0, 0.3]  # det < 0: sqrt gives NaN
    o = O.local_loglik(frame, g, ind2.reshape(-1), K, [[0], [0, 4]])
    assert o[0] == 0.0
    assert o[1] == O.local_loglik(frame, g, ind2.reshape(-1), K, [[4]])[0]


@pytest.fixture(scope="module")
def solved():
    fr = synth.config1(seed=31, n_cells=20, target_voxels=4000)
    opts = O.make_opts(K=5, max_iter=4)
    gm, rnk, ind, hist, it, ll = O.em_run(fr, opts)
    return fr, gm, ind


def test_all_supervoxels_equal_the_loglik(solved):
    fr, gm, ind = solved
    L = O.oracle()
    tot = L.orc_sum_alpha(gm.shape[0], gm.ctypes.data)
    ll = L.orc_loglik(fr.N, fr.S, O.p(fr.query, O._fp), O.p(fr.img, O._fp),
                      O.p(fr.sv_ptr, O._lp), gm.shape[0], gm.ctypes.data, 5,
                      O.p(ind, O._ip), tot)
    out = O.local_loglik(fr, gm, ind, 5, [list(range(fr.S))])
    assert abs(out[0] - ll) <= 1e-12 * abs(ll)


def test_per_gaussian_lists_partition_the_assigned_supervoxels(solved):
    # the tracker's lists: each Gaussian's supervoxelIdx[0 .. supervoxelNum)
    # (mainTrackingGaussianMixtureModel.cpp:1312-1317)
    fr, gm, ind = solved
    lists = [list(gm["supervoxelIdx"][k][: gm["supervoxelNum"][k]]) for k in range(gm.shape[0])]
    per = O.local_loglik(fr, gm, ind, 5, lists)
    flat = sorted(s for l in lists for s in l)
    assert len(flat) > fr.S // 2
    whole = O.local_loglik(fr, gm, ind, 5, [flat])[0]
    assert abs(per.sum() - whole) <= 1e-11 * abs(whole)
