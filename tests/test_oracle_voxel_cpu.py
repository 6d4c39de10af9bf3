"""CPU tests that pin the oracle's voxel mode (the per-voxel EM of
GMEMvariationalInferenceCUDA, GMEMupdateCUDA.cu:3420-3973).

No golden vectors or runnable reference exist for this path (the reference's
CPU voxel-mode code, variationalInference.cpp, needs mylib, a kd-tree and the
CUDA headers). Each test pins one rule through a property that a plausible
mistake breaks:
  * with unit intensities, voxel mode equals supervoxel mode on one-voxel
    supervoxels, whose oracle is pinned by the reference's own golden vectors
    (kNN query, E-step scores, M-step, log-likelihood, convergence);
  * stored responsibilities of a voxel sum to its intensity, and the M-step
    weight of a pair is the stored value (not value * intensity again);
  * a voxel whose every log score is <= -100 contributes nothing
    (GMEM.cu:2208), unlike in supervoxel mode;
  * a pair whose stored value (r * intensity) is <= 1e-3 is left out of the
    M-step (GMEM.cu:1199), unlike in supervoxel mode where r alone is tested.
"""
import numpy as np
import pytest

import oracle_lib as O
from tgmm_amd import synth

FIELDS = ["beta_k", "nu_k", "alpha_k", "m_k", "W_k", "expectedLogDetCovarianceCUDA",
          "expectedLogResponsivityCUDA", "beta_o", "nu_o", "alpha_o", "m_o", "W_o", "fixed"]


@pytest.fixture(scope="module")
def frame():
    return synth.config1(seed=77, n_cells=12, target_voxels=3000)


def single_voxel_supervoxels(f, img=None):
    """The same voxels as one-voxel supervoxels (centroid = the voxel)."""
    N = f.N
    return synth.Frame(f.query.copy(), (f.img if img is None else img).copy(), f.query.copy(),
                       np.arange(N + 1, dtype=np.int64), f.gm.copy())


def with_voxels(f, xyz, img):
    """f with extra voxels appended (coordinates xyz [n, 3], intensities img)."""
    N = f.N
    q = f.query.reshape(3, N)
    q2 = np.concatenate([q, np.asarray(xyz, np.float32).T], axis=1)
    im = np.concatenate([f.img, np.asarray(img, np.float32)])
    n2 = q2.shape[1]
    return synth.Frame(q2.reshape(-1).copy(), im, q2.reshape(-1).copy(),
                       np.arange(n2 + 1, dtype=np.int64), f.gm.copy())


def test_unit_intensity_voxel_mode_equals_one_voxel_supervoxels(frame):
    f1 = single_voxel_supervoxels(frame, img=np.ones(frame.N, np.float32))
    gv, rv, iv, hv, nv, llv = O.em_run(f1, O.make_opts(K=5, max_iter=6, voxel=True))
    gs, rs, is_, hs, ns, lls = O.em_run(f1, O.make_opts(K=5, max_iter=6))
    assert nv == ns == 6
    for k in FIELDS:
        assert np.array_equal(gv[k], gs[k]), k
    assert np.array_equal(hv, hs) and llv == lls
    assert np.array_equal(iv, is_)
    # candidates cut at -100 are exactly 0 in voxel mode, < e-90 otherwise
    assert np.allclose(rv, rs, rtol=0, atol=1e-30)
    # voxel mode leaves the supervoxel lists as given
    assert np.array_equal(gv["supervoxelNum"], frame.gm["supervoxelNum"])


def test_rows_sum_to_intensity_and_mstep_weight_is_stored_value(frame):
    rng = np.random.default_rng(3)
    img = rng.uniform(0.05, 1.0, frame.N).astype(np.float32)
    f1 = single_voxel_supervoxels(frame, img=img)
    N, G, K = f1.N, f1.G, 5
    _, rnk, ind, Nk, xk, Ck = O.em_partial(f1, O.make_opts(K=K, max_iter=1, voxel=True), 1)
    r = rnk.reshape(K, N).astype(np.float64)
    ii = ind.reshape(K, N)
    rows = r.sum(0)
    hit = rows > 0
    assert hit.mean() > 0.99
    assert np.allclose(rows[hit], img[hit], rtol=2e-6)
    # the M-step weight of pair (v, j) is the stored value, above 1e-3
    X = f1.query.reshape(3, N).T.astype(np.float64)
    for k in range(G):
        sel = (ii == k) & (rnk.reshape(K, N) > np.float32(1e-3))
        w = np.zeros(N)
        for j in range(K):
            w += np.where(sel[j], r[j], 0.0)
        if w.sum() == 0:
            assert Nk[k] == 0
            continue
        assert abs(Nk[k] - w.sum()) <= 1e-12 * w.sum()
        assert np.allclose(xk[k], np.average(X, axis=0, weights=w), rtol=1e-12)


def _stats(f, voxel):
    _, _, _, Nk, xk, Ck = O.em_partial(f, O.make_opts(K=5, max_iter=1, voxel=voxel), 1)
    return Nk, xk, Ck


def test_far_voxel_is_cut(frame):
    base = single_voxel_supervoxels(frame)
    c = frame.query.reshape(3, frame.N).max(1) + 400.0
    far = with_voxels(base, [c], [1.0])
    a, b = _stats(base, True), _stats(far, True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    # in supervoxel mode the far voxel joins its nearest Gaussian
    a, b = _stats(base, False), _stats(far, False)
    assert not np.array_equal(a[0], b[0])


def test_faint_pairs_are_left_out(frame):
    base = single_voxel_supervoxels(frame)
    live = frame.gm["m_k"][:4].astype(np.float32)
    faint = with_voxels(base, live, [1e-4] * 4)
    a, b = _stats(base, True), _stats(faint, True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    a, b = _stats(base, False), _stats(faint, False)
    assert not np.array_equal(a[0], b[0])
