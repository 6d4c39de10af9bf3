"""CPU tests that pin the FP64 oracle.

Pins, per oracle function:
  orc_digamma         GSL psi (the reference's CPU twin) golden vectors and live
                      twin; known values; recurrence
  orc_inv3/orc_det3   W W^-1 = I; rotated diagonals; singular sentinel
  orc_eig3_ordered    known eigenvalues in the reference's order; V D V^T = W
  orc_regularize_W    reference CPU twin (GaussianMixtureModel.h:283-368)
  orc_posterior       reference CPU twin (GaussianMixtureModel.h:252-272)
  orc_expectations    reference CPU twin (GaussianMixtureModel.h:229-243)
  orc_knn             a plain-Python replay of the reference's scan (ties,
                      dead Gaussians), the K smallest distances (numpy),
                      the supervoxel scan, hand-worked slot layouts
  orc_estep           Bishop PRML 10.46-10.49 closed form with distinct
                      Gaussians (scipy digamma), hand-worked two-Gaussian
                      cases, invariants
  orc_mstep           numpy weighted mean / scatter
  orc_loglik          scipy multivariate normal log-density
  orc_death, orc_converged, orc_assign_supervoxels  rule-by-rule cases
"""
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from helpers import rel_err, scan_replay
from tgmm_amd import synth
from tgmm_amd.records import GMM_DTYPE, w_full

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reftwin_vectors.npz"))
L = O.oracle()
dp, fp, ip = O._dp, O._fp, O._ip


def packed(M):
    return np.array([M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]])


def rot(rng):
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return Q


# --------------------------------------------------------------- digamma --
def test_digamma_vs_gsl_golden():
    x, ref = GOLD["psi_x"], GOLD["psi"]
    got = np.array([L.orc_digamma(float(v)) for v in x])
    pos = x > 0
    assert rel_err(got[pos], ref[pos]) < 1e-13
    # reflection for negative non-integers: cancellation near the poles
    assert rel_err(got[~pos], ref[~pos], floor=1.0) < 1e-9


def test_digamma_live_twin():
    T = O.reftwin()
    if T is None:
        pytest.skip("reference CPU twin not built here (no /root/reference)")
    rng = np.random.default_rng(5)
    for v in np.concatenate([rng.uniform(1e-6, 20, 200), rng.uniform(20, 1e5, 50)]):
        assert abs(L.orc_digamma(v) - T.ref_psi(v)) <= 1e-13 * max(1.0, abs(T.ref_psi(v)))


def test_digamma_known_values():
    g = 0.57721566490153286061
    assert abs(L.orc_digamma(1.0) + g) < 1e-15
    assert abs(L.orc_digamma(0.5) + g + 2 * math.log(2)) < 1e-15
    rng = np.random.default_rng(1)
    for v in rng.uniform(0.01, 50, 100):
        assert abs(L.orc_digamma(v + 1) - L.orc_digamma(v) - 1 / v) < 1e-12 * max(1, 1 / v)
    # the reference's value at non-positive integers (GMEM.cu:515-518)
    for v in (0.0, -1.0, -7.0):
        assert L.orc_digamma(v) == -1e300


# ----------------------------------------------------------- 3x3 algebra --
def test_det_inv():
    rng = np.random.default_rng(2)
    for _ in range(50):
        Q = rot(rng)
        d = rng.uniform(0.01, 3, 3)
        M = Q @ np.diag(d) @ Q.T
        w = packed(M)
        assert abs(L.orc_det3(O.p(w, dp)) - np.prod(d)) < 1e-12 * np.prod(d)
        inv = np.zeros(6)
        assert L.orc_inv3(O.p(w, dp), O.p(inv, dp)) == 0
        assert np.allclose(w_full(inv) @ M, np.eye(3), atol=1e-10)
    sing = packed(np.diag([1.0, 2.0, 0.0]))
    inv = np.zeros(6)
    assert L.orc_inv3(O.p(sing, dp), O.p(inv, dp)) == 1
    assert np.all(inv == 1e300)


def test_eig3_order_and_reconstruction():
    rng = np.random.default_rng(3)
    for _ in range(50):
        Q = rot(rng)
        ev = np.sort(rng.uniform(0.01, 2.0, 3))[::-1]  # a > b > c
        M = Q @ np.diag(ev) @ Q.T
        d, v = np.zeros(3), np.zeros(9)
        L.orc_eig3_ordered(O.p(packed(M), dp), O.p(d, dp), O.p(v, dp))
        # eig3 order: (largest, smallest, middle)
        assert np.allclose(d, [ev[0], ev[2], ev[1]], rtol=1e-12)
        V = v.reshape(3, 3).T
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-12)
        assert np.allclose(V @ np.diag(d) @ V.T, M, atol=1e-12)
    # singular branch: (0, larger, smaller)
    M = np.array([[1.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 3.0]])  # det exactly 0
    d, v = np.zeros(3), np.zeros(9)
    L.orc_eig3_ordered(O.p(packed(M), dp), O.p(d, dp), O.p(v, dp))
    assert d[0] == 0.0 and abs(d[1] - 3) < 1e-12 and abs(d[2] - 2) < 1e-12
    V = v.reshape(3, 3).T
    assert np.allclose(V @ np.diag(d) @ V.T, M, atol=1e-12)


def _regularize(W6, nu, w4, scale=(1, 1, 5)):
    w = np.ascontiguousarray(W6, dtype=np.float64).copy()
    sc = np.array(scale, np.float32)
    L.orc_regularize_W(O.p(w, dp), float(nu), int(w4), O.p(sc, fp), 0.02, 0.2222, 9.0)
    return w


@pytest.mark.parametrize("w4", [0, 1])
def test_regularize_vs_reference_twin_golden(w4):
    Wr, nur, ref = GOLD["Wr"], GOLD["nur"], GOLD["reg_w4%d" % w4]
    for i in range(Wr.shape[0]):
        got = _regularize(Wr[i], nur[i], w4)
        err = np.linalg.norm(got - ref[i]) / np.linalg.norm(ref[i])
        assert err < 1e-9, (i, got, ref[i])


def test_regularize_clamps_known_answer():
    # diagonal W' (scale (1,1,5)): eigenvalues clamped into [lmin, lmax]/nu,
    # then the eccentricity ratio 9 enforced pairwise in eig3 order
    nu = 10.0
    W = packed(np.diag([1.0, 1e-6, 0.005 * 25]))   # W' = diag(1, 1e-6, 0.005)
    got = _regularize(W, nu, 1)
    hi, lo = 0.2222 / nu, 0.02 / nu
    # d = (1 -> hi, 1e-6 -> lo, 0.005 stays); ratio hi/lo = 11.1 > 9 -> d0 = 9 lo
    expect = np.diag([9 * lo, lo, 0.005 * 25])
    assert np.allclose(w_full(got), expect, rtol=1e-12)


@pytest.mark.parametrize("variant", ["upd_reg0_w40", "upd_reg1_w40", "upd_reg0_w41",
                                     "upd_reg1_w41"])
def test_posterior_vs_reference_twin_golden(variant):
    reg = variant[7] == "1"
    w4 = int(variant[-1])
    prior, Nk, xk, Sk, ref = GOLD["prior"], GOLD["Nk"], GOLD["xk"], GOLD["Sk"], GOLD[variant]
    for i in range(prior.shape[0]):
        g = np.zeros(1, GMM_DTYPE)
        g["beta_o"], g["nu_o"], g["alpha_o"] = prior[i, 0], prior[i, 1], prior[i, 2]
        g["m_o"] = prior[i, 3:6]
        g["W_o"] = prior[i, 6:12]
        Ck = np.ascontiguousarray(Nk[i] * Sk[i])
        x = np.ascontiguousarray(xk[i])
        L.orc_posterior(g.ctypes.data, float(Nk[i]), O.p(x, dp), O.p(Ck, dp))
        if reg:
            sc = np.array([1, 1, 5], np.float32)
            W = np.ascontiguousarray(g["W_k"][0]).copy()
            L.orc_regularize_W(O.p(W, dp), float(g["nu_k"][0]), w4, O.p(sc, fp), 0.02,
                               0.2222, 9.0)
            g["W_k"] = W
        r = ref[i]
        assert rel_err(g["beta_k"], r[0]) < 1e-13
        assert rel_err(g["nu_k"], r[1]) < 1e-13
        assert rel_err(g["alpha_k"], r[2]) < 1e-13
        assert np.linalg.norm(g["m_k"][0] - r[3:6]) < 1e-12 * np.linalg.norm(r[3:6])
        assert np.linalg.norm(g["W_k"][0] - r[6:12]) < 1e-9 * np.linalg.norm(r[6:12])
        assert np.array_equal(g["W_o"][0], r[12:18])


def test_expectations_vs_reference_twin_golden():
    W, nu, eld = GOLD["W"], GOLD["nu"], GOLD["eld"]
    al, tot, elp = GOLD["alpha"], GOLD["total"], GOLD["elp"]
    for i in range(W.shape[0]):
        g = np.zeros(1, GMM_DTYPE)
        g["W_k"], g["nu_k"], g["alpha_k"] = W[i], nu[i], al[i]
        L.orc_expectations(1, g.ctypes.data, float(tot[i]))
        assert abs(g["expectedLogDetCovarianceCUDA"][0] - eld[i]) < 1e-12 * abs(eld[i])
        assert abs(g["expectedLogResponsivityCUDA"][0] - elp[i]) < 1e-12 * max(1, abs(elp[i]))
    dead = np.zeros(1, GMM_DTYPE)
    dead["m_o"] = -1e32
    L.orc_expectations(1, dead.ctypes.data, 10.0)
    assert dead["expectedLogDetCovarianceCUDA"][0] == -1e32
    assert dead["expectedLogResponsivityCUDA"][0] == -1e32


# -------------------------------------------------------------------- kNN --
def _knn(cent, S, gm, K, scale=(1, 1, 5)):
    ind = np.zeros(K * S, np.int32)
    sc = np.asarray(scale, np.float32)
    L.orc_knn(S, O.p(np.ascontiguousarray(cent, np.float32), fp), gm.shape[0], gm.ctypes.data,
              O.p(sc, fp), K, O.p(ind, ip))
    return ind


def _means_soa(gm):
    """The records' means as the reference's float copy (GMEM.cu:527-539)."""
    return np.ascontiguousarray(gm["m_k"].astype(np.float32).T.reshape(-1))


@pytest.mark.parametrize("K,G", [(5, 200), (8, 60), (5, 3), (16, 40)])
def test_knn_matches_the_plain_scan_replay_with_ties_and_dead(K, G):
    """The oracle's kNN equals a plain-Python replay of knnCuda.cu's scan,
    slot for slot, on integer coordinates (many exact ties at the threshold)
    with dead Gaussians (m = -1e32, an infinite distance) among them."""
    rng = np.random.default_rng(K * 100 + G)
    S = 120
    cent = rng.integers(0, 40, size=3 * S).astype(np.float32)
    gm = np.zeros(G, GMM_DTYPE)
    gm["m_k"] = rng.integers(0, 40, size=(G, 3)).astype(np.float64)
    gm["m_o"] = gm["m_k"]
    dead = rng.random(G) < 0.1
    gm["m_o"][dead] = -1e32
    gm["m_k"][dead] = -1e32
    ind = _knn(cent, S, gm, K)
    ref, _ = scan_replay(cent, _means_soa(gm), K, (1, 1, 5))
    assert np.array_equal(ind, ref)
    assert not np.isin(ind, np.flatnonzero(dead)).any()


@pytest.mark.parametrize("K", [1, 5, 8])
def test_knn_slots_hold_the_k_smallest_distances(K):
    """Whatever the slot layout, the filled slots hold the K smallest float
    distances (numpy brute force, the reference's operation order)."""
    rng = np.random.default_rng(40 + K)
    S, G = 200, 150
    cent = rng.uniform(0, 60, 3 * S).astype(np.float32)
    gm = np.zeros(G, GMM_DTYPE)
    gm["m_k"] = rng.uniform(0, 60, (G, 3))
    ind = _knn(cent, S, gm, K).reshape(K, S)
    c = cent.reshape(3, S).T
    m = gm["m_k"].astype(np.float32)
    sc = np.array([1, 1, 5], np.float32)
    d = np.zeros((S, G), np.float32)
    for a in range(3):
        t = (c[:, None, a] - m[None, :, a]) * sc[a]
        d = d + t * t
    for s in range(S):
        assert np.array_equal(np.sort(d[s, ind[:, s]]), np.sort(d[s])[:K])


def test_knn_equals_the_supervoxel_scan_on_the_same_points():
    """The EM's scan (knnCuda.cu:61-144) and knnCUDA_'s (the other knnCuda.cu)
    are the same rule: on living means they give the same slots."""
    rng = np.random.default_rng(44)
    S, G = 300, 500
    cent = np.round(rng.uniform(0, 30, 3 * S)).astype(np.float32)
    gm = np.zeros(G, GMM_DTYPE)
    gm["m_k"] = np.round(rng.uniform(0, 30, (G, 3)))
    for K in (1, 5, 11):
        a = _knn(cent, S, gm, K)
        b, _ = O.sv_knn(cent, _means_soa(gm), K, (1, 1, 5))
        assert np.array_equal(a, b)


def test_knn_slot_layout_by_hand():
    """Query at the origin, means on the x axis (scale 1):
      distances 25, 16, 9, 4, 1 with K = 3: slots (25,0) (16,1) (9,2); 4
      replaces the largest (slot 0), then 1 replaces 16 (slot 1): [3, 4, 2];
      distances 4, 4, 4, 4 with K = 2: (0, 1), then each later one replaces
      slot 0 (the first slot holding the threshold): [3, 1];
      a dead Gaussian between them is skipped; fewer living Gaussians than
      slots leave -1."""
    def line(xs, K, dead=()):
        gm = np.zeros(len(xs), GMM_DTYPE)
        gm["m_k"][:, 0] = xs
        for k in dead:
            gm["m_k"][k] = -1e32
            gm["m_o"][k] = -1e32
        return list(_knn(np.zeros(3, np.float32), 1, gm, K, (1, 1, 1)))
    assert line([5, 4, 3, 2, 1], 3) == [3, 4, 2]
    assert line([2, -2, 2, -2], 2) == [3, 1]
    assert line([5, 0, 4, 3], 2, dead=(1,)) == [3, 2]
    assert line([3, 1], 4) == [0, 1, -1, -1]
    assert line([3, 1, 7], 4, dead=(0, 2)) == [1, -1, -1, -1]


# ------------------------------------------------------------- E / M / LL --
def tiny_frame(rng, n_sv=20, per=7, G=4):
    S = n_sv
    ptr = np.arange(S + 1, dtype=np.int64) * per
    N = int(ptr[-1])
    q = rng.normal(size=(3, N)) * 3 + 50
    query = q.reshape(-1).astype(np.float32)
    img = rng.uniform(0.1, 1.0, N).astype(np.float32)
    cent = np.concatenate([query[a * N:(a + 1) * N].reshape(S, per).mean(1)
                           for a in range(3)]).astype(np.float32)
    gm = np.zeros(G, GMM_DTYPE)
    gm["m_k"] = 50 + rng.normal(size=(G, 3))
    gm["m_o"] = gm["m_k"]
    gm["W_k"] = [0.1, 0.01, 0, 0.12, 0, 0.3]
    gm["W_o"] = gm["W_k"]
    gm["nu_k"] = gm["nu_o"] = 6.0
    gm["beta_k"] = gm["beta_o"] = 2.0
    gm["alpha_k"] = gm["alpha_o"] = 3.0
    return N, S, query, img, cent, ptr, gm


def test_estep_invariants():
    rng = np.random.default_rng(11)
    N, S, query, img, cent, ptr, gm = tiny_frame(rng)
    K = 3
    ind = np.zeros(K * S, np.int32)
    sc = np.array([1, 1, 1], np.float32)
    L.orc_knn(S, O.p(cent, fp), gm.shape[0], gm.ctypes.data, O.p(sc, fp), K, O.p(ind, ip))
    L.orc_expectations(gm.shape[0], gm.ctypes.data, float(gm["alpha_k"].sum()))
    rnk = np.zeros(K * S, np.float32)
    L.orc_estep(N, S, O.p(query, fp), O.p(ptr, O._lp), gm.shape[0], gm.ctypes.data, K,
                O.p(ind, ip), O.p(rnk, fp))
    r = rnk.reshape(K, S)
    assert np.all(r >= 0) and np.all(r <= 1)
    assert np.allclose(r.sum(0), 1.0, atol=1e-6)
    # two identical Gaussians share every supervoxel equally
    g2 = np.concatenate([gm[:1], gm[:1]]).astype(GMM_DTYPE)
    ind2 = np.tile(np.array([0, 1], np.int32)[:, None], (1, S)).reshape(-1)
    L.orc_expectations(2, g2.ctypes.data, float(g2["alpha_k"].sum()))
    r2 = np.zeros(2 * S, np.float32)
    L.orc_estep(N, S, O.p(query, fp), O.p(ptr, O._lp), 2, g2.ctypes.data, 2, O.p(ind2, ip),
                O.p(r2, fp))
    assert np.allclose(r2, 0.5, atol=1e-7)
    # one candidate and padding: r = 1 and 0
    ind3 = np.concatenate([np.zeros(S, np.int32), -np.ones(S, np.int32)])
    r3 = np.zeros(2 * S, np.float32)
    L.orc_estep(N, S, O.p(query, fp), O.p(ptr, O._lp), 2, g2.ctypes.data, 2, O.p(ind3, ip),
                O.p(r3, fp))
    assert np.all(r3[:S] == 1.0) and np.all(r3[S:] == 0.0)


def test_mstep_matches_numpy_weighted_moments():
    rng = np.random.default_rng(12)
    N, S, query, img, cent, ptr, gm = tiny_frame(rng, G=2)
    K = 2
    ind = np.concatenate([np.zeros(S, np.int32), np.ones(S, np.int32)])
    r = rng.uniform(0.0, 1.0, S)
    r[:3] = 1e-4  # below the 1e-3 cut: excluded (GMEM.cu:1350)
    rnk = np.concatenate([r, 1 - r]).astype(np.float32)
    Nk, xk, Ck = np.zeros(2), np.zeros(6), np.zeros(12)
    L.orc_mstep(N, S, O.p(query, fp), O.p(img, fp), O.p(ptr, O._lp), 2, K, O.p(ind, ip),
                O.p(rnk, fp), O.p(Nk, dp), O.p(xk, dp), O.p(Ck, dp))
    X = query.reshape(3, N).T.astype(np.float64)
    svid = np.repeat(np.arange(S), 7)
    for k in range(2):
        rk = rnk[k * S:(k + 1) * S].astype(np.float64)
        rk = np.where(rk > np.float32(1e-3), rk, 0.0)
        w = rk[svid] * img.astype(np.float64)
        mean = np.average(X, axis=0, weights=w)
        cov = np.cov(X.T, aweights=w, bias=True) * w.sum()
        assert abs(Nk[k] - w.sum()) < 1e-12 * w.sum()
        assert np.allclose(xk[3 * k:3 * k + 3], mean, rtol=1e-13)
        assert np.allclose(w_full(Ck[6 * k:6 * k + 6]), cov, rtol=1e-10)


def test_loglik_single_gaussian_closed_form():
    from scipy.stats import multivariate_normal
    rng = np.random.default_rng(13)
    N, S, query, img, cent, ptr, gm = tiny_frame(rng, G=1)
    ind = np.zeros(S, np.int32)
    ll = L.orc_loglik(N, S, O.p(query, fp), O.p(img, fp), O.p(ptr, O._lp), 1, gm.ctypes.data, 1,
                      O.p(ind, ip), float(gm["alpha_k"][0]))
    X = query.reshape(3, N).T.astype(np.float64)
    prec = gm["nu_k"][0] * w_full(gm["W_k"][0])
    lp = multivariate_normal(mean=gm["m_k"][0], cov=np.linalg.inv(prec)).logpdf(X)
    assert abs(ll - np.sum(img * lp)) < 1e-10 * abs(ll)


def test_death_rule():
    g = np.zeros(4, GMM_DTYPE)
    g["alpha_k"] = [1.0, 1e-7, 1e-7, 5.0]
    g["W_k"] = [1, 0, 0, 1, 0, 1]
    g["fixed"][2] = 1
    g["W_k"][3, 1] = np.nan
    L.orc_death(4, g.ctypes.data, float(g["alpha_k"][:3].sum() + 5.0))
    assert g["m_o"][0, 0] == 0.0                # survives
    assert g["m_o"][1, 0] == -1e32 and g["alpha_k"][1] == 0  # pi < 1e-5
    assert g["m_o"][2, 0] == 0.0                # fixed survives the alpha rule
    assert g["m_o"][3, 0] == -1e32              # NaN W always dies
    assert np.array_equal(g["W_k"][3], [1, 0, 0, 1, 0, 1])


def test_convergence_rule():
    c = L.orc_converged
    assert c(-1.0, -10.0, -100.0, -1000.0, 1e-6) == 0
    assert c(-100.0, -100.0 * (1 + 1e-7), -90.0, -80.0, 1e-6) == 1
    assert c(-100.0, -90.0, -100.0, -80.0, 1e-6) == 1  # oscillation caught by ll-2
    assert c(-100.0, -90.0, -95.0, -80.0, -1.0) == 0   # tol < 0 never converges
    assert c(float("nan"), -1.0, -2.0, -3.0, -1.0) == 1  # NaN stops (comparison false)


def test_assign_supervoxels_rule():
    S, K, G = 40, 2, 2
    rnk = np.zeros(K * S, np.float32)
    ind = np.zeros(K * S, np.int32)
    rnk[:S] = 0.9          # slot 0: all supervoxels to Gaussian 0 (cap truncates)
    ind[S:] = 1
    rnk[S + 5] = 0.51      # slot 1: one supervoxel to Gaussian 1
    g = np.zeros(G, GMM_DTYPE)
    L.orc_assign_supervoxels(S, G, g.ctypes.data, K, O.p(rnk, fp), O.p(ind, ip), 16)
    assert g["supervoxelNum"][0] == 16 and list(g["supervoxelIdx"][0]) == list(range(16))
    assert g["supervoxelNum"][1] == 1 and g["supervoxelIdx"][1][0] == 5


def test_oracle_em_config1_recovers_cells():
    f = synth.config1()
    gm, rnk, ind, hist, it, ll = O.em_run(f, O.make_opts(K=5, max_iter=10))
    assert it == 10 and np.isfinite(ll)
    err = np.abs(gm["m_k"] - f.true_centres)
    assert np.median(err) < 0.3  # means move from ~1.6 voxel noise to the cells
    assert (gm["m_o"][:, 0] > -1e31).all()


def test_next_priors_helper_vs_reference_twin_golden():
    """The host restatement of updatePriors used to check the server's
    device-side prior propagation, against the reference's CPU code."""
    from helpers import next_priors
    st, pri, pct = GOLD["st"], GOLD["pri"], GOLD["pct"]
    for i in range(st.shape[0]):
        g = np.zeros(1, GMM_DTYPE)
        g["beta_k"], g["nu_k"], g["alpha_k"] = st[i, 0], st[i, 1], st[i, 2]
        g["m_k"] = st[i, 3:6]
        g["W_k"] = st[i, 6:12]
        g["nu_o"] = st[i, 1] - st[i, 12]  # so that nu_k - nu_o = N_k
        g["alpha_o"] = st[i, 13]
        out = next_priors(g, pct[0], pct[1], pct[2], 1000.0 + i)
        r = pri[i]
        assert abs(out["beta_o"][0] - r[0]) <= 1e-12 * abs(r[0])
        assert abs(out["nu_o"][0] - r[1]) <= 1e-12 * abs(r[1])
        assert abs(out["alpha_o"][0] - r[2]) <= 1e-12 * abs(r[2])
        assert np.array_equal(out["m_o"][0], r[3:6])
        assert np.allclose(out["W_o"][0], r[6:12], rtol=1e-12, atol=0)


# --------------------------------------------------- E-step closed forms --
def _spd(rng, lo, hi):
    Q = rot(rng)
    return Q @ np.diag(rng.uniform(lo, hi, 3)) @ Q.T


def _distinct_gaussians(rng, G, centre):
    """Gaussians whose beta, nu, alpha and W all differ, so that no
    per-Gaussian term of the E-step cancels between candidates."""
    gm = np.zeros(G, GMM_DTYPE)
    gm["m_k"] = centre + rng.normal(size=(G, 3)) * 2.0
    gm["m_o"] = gm["m_k"]
    for k in range(G):
        gm["W_k"][k] = packed(_spd(rng, 0.03, 0.6))
    gm["W_o"] = gm["W_k"]
    gm["beta_k"] = rng.uniform(0.4, 6.0, G)
    gm["nu_k"] = rng.uniform(3.2, 14.0, G)
    gm["alpha_k"] = rng.uniform(0.3, 25.0, G)
    return gm


def _bishop_expectations(gm):
    """PRML eqs. 10.65 and 10.66 (D = 3) with scipy's digamma:
    E[ln |Lambda_k|] = sum_i psi((nu_k + 1 - i)/2) + 3 ln 2 + ln |W_k|,
    E[ln pi_k] = psi(alpha_k) - psi(sum_j alpha_j).
    The reference sums i = 0, 1, 2 (GMEM.cu:569-572, and its CPU code
    GaussianMixtureModel.h:230-235), where Bishop's i runs 1..3: its digamma
    arguments are (nu + 1)/2, nu/2, (nu - 1)/2. That is the reference's
    definition, which the oracle follows."""
    from scipy.special import digamma
    eld = np.array([sum(digamma((g["nu_k"] + 1 - i) / 2.0) for i in (0, 1, 2)) +
                    3 * math.log(2.0) + math.log(np.linalg.det(w_full(g["W_k"])))
                    for g in gm])
    elp = digamma(gm["alpha_k"]) - digamma(gm["alpha_k"].sum())
    return eld, elp


def _bishop_log_rho(x, g, eld, elp):
    """ln rho of one voxel and one Gaussian, PRML eq. 10.46 with eq. 10.64:
    E[ln pi] + E[ln |Lambda|]/2 - (D/2) ln 2 pi - (D/beta + nu d^T W d)/2.
    The (D/2) ln 2 pi term is common to all candidates (the reference omits
    it, GMEM.cu:2281-2298); it cancels in the normalisation."""
    d = np.asarray(x, np.float64) - g["m_k"]
    maha = 3.0 / g["beta_k"] + g["nu_k"] * (d @ w_full(g["W_k"]) @ d)
    return elp + 0.5 * eld - 1.5 * math.log(2 * math.pi) - 0.5 * maha


def test_estep_matches_bishop_closed_form_with_distinct_gaussians():
    """Supervoxels of 1 to 12 voxels, candidates with different beta, nu,
    alpha and W: r_sk = exp(L_sk - logsumexp_j L_sj) with L_sk the sum over
    the supervoxel's voxels of ln rho (PRML 10.46-10.49), the expectations
    from scipy, independent of the oracle's own expectation code."""
    from scipy.special import logsumexp
    rng = np.random.default_rng(21)
    G, S, K = 7, 40, 4
    sizes = rng.integers(1, 13, S)
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    N = int(ptr[-1])
    X = 50.0 + rng.normal(size=(N, 3)) * 2.5
    query = np.ascontiguousarray(X.T.reshape(-1), np.float32)
    Xf = query.reshape(3, N).T.astype(np.float64)  # the voxels as the oracle sees them
    gm = _distinct_gaussians(rng, G, 50.0)
    eld, elp = _bishop_expectations(gm)
    gm["expectedLogDetCovarianceCUDA"] = eld
    gm["expectedLogResponsivityCUDA"] = elp
    ind = np.full((K, S), -1, np.int32)
    for s in range(S):
        n_c = K if s % 5 else 2  # some supervoxels with padded slots
        ind[:n_c, s] = rng.choice(G, n_c, replace=False)
    ind = ind.reshape(-1)
    rnk = np.zeros(K * S, np.float32)
    L.orc_estep(N, S, O.p(query, fp), O.p(ptr, O._lp), G, gm.ctypes.data, K, O.p(ind, ip),
                O.p(rnk, fp))
    rnk = rnk.reshape(K, S)
    indr = ind.reshape(K, S)
    for s in range(S):
        js = [j for j in range(K) if indr[j, s] >= 0]
        Ls = np.array([sum(_bishop_log_rho(Xf[v], gm[indr[j, s]], eld[indr[j, s]],
                                           elp[indr[j, s]])
                           for v in range(ptr[s], ptr[s + 1])) for j in js])
        want = np.exp(Ls - logsumexp(Ls))
        assert np.allclose(rnk[js, s], want, rtol=1e-6, atol=1e-7), (s, rnk[js, s], want)
        assert np.all(rnk[[j for j in range(K) if indr[j, s] < 0], s] == 0.0)
    # the candidates' terms matter: responsibilities are far from uniform
    assert np.mean(np.max(rnk, 0)) > 0.8


def test_oracle_expectations_match_bishop():
    rng = np.random.default_rng(22)
    gm = _distinct_gaussians(rng, 9, 0.0)
    eld, elp = _bishop_expectations(gm)
    L.orc_expectations(9, gm.ctypes.data, float(gm["alpha_k"].sum()))
    assert np.allclose(gm["expectedLogDetCovarianceCUDA"], eld, rtol=1e-13, atol=1e-13)
    assert np.allclose(gm["expectedLogResponsivityCUDA"], elp, rtol=1e-13, atol=1e-13)


def _hand_pair():
    """Two Gaussians with literal parameters and expectations:
      A: m = (1, 2, 4), W = 0.5 I, nu = 4, beta = 2, E[ln det] = 0.3,
         E[ln pi] = -0.2
      B: m = (2, 2, 3), W = diag(1, 0.5, 0.25), nu = 6, beta = 1,
         E[ln det] = -1.0, E[ln pi] = -1.5"""
    g = np.zeros(2, GMM_DTYPE)
    g["m_k"] = [[1, 2, 4], [2, 2, 3]]
    g["W_k"] = [[0.5, 0, 0, 0.5, 0, 0.5], [1, 0, 0, 0.5, 0, 0.25]]
    g["nu_k"] = [4, 6]
    g["beta_k"] = [2, 1]
    g["alpha_k"] = [1, 1]
    g["expectedLogDetCovarianceCUDA"] = [0.3, -1.0]
    g["expectedLogResponsivityCUDA"] = [-0.2, -1.5]
    return g


def test_estep_hand_worked_two_gaussians():
    """Supervoxel 0: one voxel x = (1, 2, 3).
      A: d = (0, 0, -1), d^T W d = 0.5, 3/beta + nu 0.5 = 1.5 + 2 = 3.5,
         L_A = -0.2 + 0.15 - 1.75 = -1.8
      B: d = (-1, 0, 0), d^T W d = 1, 3/beta + nu 1 = 3 + 6 = 9,
         L_B = -1.5 - 0.5 - 4.5 = -6.5
      r_A = 1 / (1 + exp(-4.7)), r_B = 1 - r_A.
    Supervoxel 1: voxels (1, 2, 3) and (1, 2, 5).
      A at (1, 2, 5): d = (0, 0, 1), the same term -1.8: L_A = -3.6
      B at (1, 2, 5): d = (-1, 0, 2), d^T W d = 1 + 0.25 * 4 = 2,
         3 + 12 = 15, -1.5 - 0.5 - 7.5 = -9.5: L_B = -16
      r_A = 1 / (1 + exp(-12.4)).
    Candidate order is B, A in slot order for supervoxel 1."""
    g = _hand_pair()
    query = np.array([1, 1, 1, 2, 2, 2, 3, 3, 5], np.float32)  # SoA, N = 3
    ptr = np.array([0, 1, 3], np.int64)
    ind = np.array([0, 1, 1, 0], np.int32)  # slot 0: A, B; slot 1: B, A
    rnk = np.zeros(4, np.float32)
    L.orc_estep(3, 2, O.p(query, fp), O.p(ptr, O._lp), 2, g.ctypes.data, 2, O.p(ind, ip),
                O.p(rnk, fp))
    rA0 = 1.0 / (1.0 + math.exp(-4.7))
    rA1 = 1.0 / (1.0 + math.exp(-12.4))
    assert rnk[0] == np.float32(rA0) and rnk[2] == np.float32(1.0 - rA0)
    assert rnk[3] == np.float32(rA1) and rnk[1] == np.float32(1.0 - rA1)


def test_estep_voxel_hand_worked_two_gaussians():
    """Voxel mode (GMEM.cu:2177-2242) on the voxels of the hand-worked case:
    the scores are the per-voxel terms above; the stored value is r times the
    intensity. Voxel (1, 2, 5): L_A = -1.8, L_B = -9.5. A third candidate far
    away (score below -100) is cut and takes no share."""
    g = np.concatenate([_hand_pair(), _hand_pair()[:1]]).astype(GMM_DTYPE)
    g["m_k"][2] = [40, 2, 4]  # d = (-39, 0, -1): 3/2 + 4 * 0.5 * 1522 > 200
    query = np.array([1, 1, 2, 2, 3, 5], np.float32)  # N = 2
    img = np.array([0.5, 2.0], np.float32)
    ind = np.array([0, 0, 1, 1, 2, 2], np.int32)
    rnk = np.zeros(6, np.float32)
    L.orc_estep_voxel(2, O.p(query, fp), O.p(img, fp), 3, g.ctypes.data, 3, O.p(ind, ip),
                      O.p(rnk, fp))
    r0 = 1.0 / (1.0 + math.exp(-4.7))
    r1 = 1.0 / (1.0 + math.exp(-7.7))
    assert rnk[0] == np.float32(0.5 * r0) and rnk[2] == np.float32(0.5 * (1 - r0))
    assert rnk[1] == np.float32(2.0 * r1) and rnk[3] == np.float32(2.0 * (1 - r1))
    assert rnk[4] == 0.0 and rnk[5] == 0.0


def test_mstep_pair_cut_compares_in_double():
    """The reference tests the float r against the double 1e-3 (GMEM.cu:1350,
    1195): r == 1e-3f (0.00100000005) enters the M-step, the float below it
    does not."""
    rng = np.random.default_rng(23)
    N, S, query, img, cent, ptr, gm = tiny_frame(rng, n_sv=2, G=1)
    ind = np.zeros(S, np.int32)
    for r, kept in ((np.float32(1e-3), True),
                    (np.nextafter(np.float32(1e-3), np.float32(0)), False)):
        rnk = np.array([r, 0.5], np.float32)
        Nk, xk, Ck = np.zeros(1), np.zeros(3), np.zeros(6)
        L.orc_mstep(N, S, O.p(query, fp), O.p(img, fp), O.p(ptr, O._lp), 1, 1, O.p(ind, ip),
                    O.p(rnk, fp), O.p(Nk, dp), O.p(xk, dp), O.p(Ck, dp))
        w1 = 0.5 * img[7:14].astype(np.float64).sum()
        w0 = float(r) * img[:7].astype(np.float64).sum()
        assert abs(Nk[0] - (w1 + (w0 if kept else 0.0))) < 1e-12 * w1, (r, Nk[0])
