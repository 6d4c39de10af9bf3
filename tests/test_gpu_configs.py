"""GPU parity at the benchmark workloads, against the FP64 oracle:

  * config 2 (10,000 Gaussians, 4.86M voxels, 20 iterations): records,
    expectations, log-likelihood, the last search's slots, the supervoxel
    lists (capped ones included);
  * config 5 (60,000 Gaussians, 30M voxels, K = 8, 25 iterations): the same;
  * config 3 (64 frames in one batch, 15 iterations): every 8th frame;
  * config 4 (served time-lapse streams at their full cell counts): the first
    frames of each stream equal the batch API given the same priors;
  * convergence at the tracker's tolerance (tol = 1e-6, three-history stop,
    GMEM.cu:4145-4148): frames that stop at different iterations stop where
    the oracle does, alone and in one batch;
  * the death rule (GMEM.cu:1813-1843) firing on the GPU as in the oracle.
The kNN slots are also checked exactly, free of the iterations' drift: the
last search equals the oracle's scan over the means it started from."""
import copy

import numpy as np
import pytest

import oracle_lib as O
from helpers import next_priors, device_alpha_sum, record_errors, rel_err
from tgmm_amd import em, synth
from tgmm_amd.records import GMM_DTYPE

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _solve(frames, K, iters, tol=-1.0):
    with em.Batch(frames, em.make_opts(K=K, max_iter=iters, tol=tol)) as b:
        b.solve(em.SOLVE_GRAPH)
        return b.fetch(), b.stats()


def _knn(f, gm, K):
    ind = np.zeros(K * f.S, np.int32)
    O.oracle().orc_knn(f.S, O.p(f.centroids, O._fp), gm.shape[0],
                       np.ascontiguousarray(gm).ctypes.data,
                       O.p(np.array([1, 1, 5], np.float32), O._fp), K, O.p(ind, O._ip))
    return ind


def _check_vs_oracle(f, o, K, iters, tol=-1.0):
    gm_o, rnk_o, ind_o, hist, it, ll = O.em_run(f, O.make_opts(K=K, max_iter=iters, tol=tol))
    e = record_errors(o["gm"], gm_o)
    e["ll"] = rel_err(o["ll"], ll)
    print("errors %s" % e, flush=True)
    assert o["iterations"] == it
    assert max(e.values()) < TOL, e
    same = o["ind"] == ind_o
    print("slots equal to the oracle's: %d of %d" % (same.sum(), same.size), flush=True)
    assert same.all()
    assert np.array_equal(o["gm"]["supervoxelNum"], gm_o["supervoxelNum"])
    assert np.array_equal(o["gm"]["supervoxelIdx"], gm_o["supervoxelIdx"])
    return gm_o, rnk_o


@pytest.fixture(scope="module")
def c2():
    f = synth.config2()
    print("config 2: N=%d S=%d G=%d" % (f.N, f.S, f.G), flush=True)
    return f


def test_config2_twenty_iterations_match_oracle(c2):
    out, st = _solve([c2], 5, 20)
    o = out[0]
    assert st["layout_mismatches"] == 0
    gm_o, rnk_o = _check_vs_oracle(c2, o, 5, 20)
    capped = gm_o["supervoxelNum"] == 16
    print("capped Gaussians: %d, deaths: %d" % (capped.sum(), (gm_o["m_o"][:, 0] < -1e31).sum()),
          flush=True)
    assert capped.sum() > 1000
    # responsibilities where they enter the M-step
    big = rnk_o > 1e-3
    assert rel_err(o["rnk"][big], rnk_o[big]) < 1e-4


def test_config2_last_search_is_the_reference_scan(c2):
    """The slots of a 20-iteration solve are the oracle's scan over the means
    iteration 20 started from (the final means of a 19-iteration solve)."""
    prev, _ = _solve([c2], 5, 19)
    cur, st = _solve([c2], 5, 20)
    assert st["layout_mismatches"] == 0
    assert np.array_equal(cur[0]["ind"], _knn(c2, prev[0]["gm"], 5))


@pytest.fixture(scope="module")
def c5():
    f = synth.config5()
    print("config 5: N=%d S=%d G=%d" % (f.N, f.S, f.G), flush=True)
    return f


def test_config5_full_solve_matches_oracle(c5):
    out, st = _solve([c5], 8, 25)
    o = out[0]
    assert st["layout_mismatches"] == 0
    gm_o, rnk_o = _check_vs_oracle(c5, o, 8, 25)
    # properties of the result: rows of r sum to 1, records finite, precisions
    # positive definite
    r = o["rnk"].reshape(8, c5.S)
    has = (o["ind"].reshape(8, c5.S) >= 0).any(0)
    assert np.allclose(r[:, has].sum(0), 1.0, atol=1e-5)
    live = o["gm"]["m_o"][:, 0] > -1e31
    g = o["gm"][live]
    assert np.isfinite(g["m_k"]).all() and np.isfinite(g["W_k"]).all()
    from tgmm_amd.records import w_full
    assert (np.linalg.eigvalsh(w_full(g["W_k"])) > 0).all()


def test_config5_last_search_is_the_reference_scan(c5):
    prev, _ = _solve([c5], 8, 2)
    cur, st = _solve([c5], 8, 3)
    assert st["layout_mismatches"] == 0
    assert np.array_equal(cur[0]["ind"], _knn(c5, prev[0]["gm"], 8))


def test_config3_batch_every_eighth_frame_matches_oracle():
    frames = synth.config3_frames()
    print("config 3: %d frames, N=%d" % (len(frames), sum(f.N for f in frames)), flush=True)
    out, st = _solve(frames, 5, 15)
    assert st["layout_mismatches"] == 0
    for i in range(0, len(frames), 8):
        _check_vs_oracle(frames[i], out[i], 5, 15)


def test_config4_streams_served_equal_batch():
    """Three frames of each of the four config-4 streams (2,000-5,500
    Gaussians, ~200 voxels per cell) through one server: the first frame of a
    stream equals the batch API on its records, the later ones equal the
    batch API on the host restatement of the device-side prior update."""
    streams = synth.config4_streams(n_frames=3)
    pct = (0.5, 0.5, 0.8)
    opts = em.make_opts(K=5, max_iter=10, tol=-1.0)
    with em.Server(opts, hbm_budget_bytes=8 << 30, max_frames_in_flight=8, beta_pct=pct[0],
                   nu_pct=pct[1], alpha_pct=pct[2]) as sv:
        ts = {}
        for t in range(3):
            for i, fr in enumerate(streams):
                ts[(i, t)] = sv.submit(fr[t], stream=i, with_gaussians=(t == 0))
        sv.wait(-1)
        got = {k: sv.fetch(v, streams[k[0]][0].G) for k, v in ts.items()}
    for i, fr in enumerate(streams):
        gm = fr[0].gm
        for t in range(3):
            f = synth.Frame(fr[t].query, fr[t].img, fr[t].centroids, fr[t].sv_ptr, gm)
            ref = _solve([f], 5, 10)[0][0]
            o = got[(i, t)]
            assert o["gm"].tobytes() == ref["gm"].tobytes(), (i, t)
            assert np.array_equal(o["rnk"], ref["rnk"]) and np.array_equal(o["ind"], ref["ind"])
            assert o["ll"] == ref["ll"]
            tot = device_alpha_sum(ref["gm"]["alpha_k"])
            gm = next_priors(ref["gm"], pct[0], pct[1], pct[2], tot)


def _conv_frames():
    """Frames of different sizes and cell densities: at tol = 1e-6 they stop
    at different iterations."""
    return [synth.config1(seed=s, n_cells=n, target_voxels=v)
            for s, n, v in ((31, 120, 25000), (32, 200, 40000), (33, 60, 9000),
                            (34, 300, 50000))]


def test_convergence_at_the_tracker_tolerance():
    """maxIterEM = 400, tolLikelihood = 1e-6 (parseConfigFile.h:152-153): each
    frame stops at the oracle's iteration, its results match, and each
    frame's result is bitwise the same alone and in one batch."""
    frames = _conv_frames()
    mix, _ = _solve(frames, 5, 400, tol=1e-6)
    iters = [o["iterations"] for o in mix]
    print("iterations at tol 1e-6: %s" % iters, flush=True)
    assert len(set(iters)) > 1  # they stop at different iterations
    for f, o in zip(frames, mix):
        _check_vs_oracle(f, o, 5, 400, tol=1e-6)
        alone = _solve([f], 5, 400, tol=1e-6)[0][0]
        assert alone["gm"].tobytes() == o["gm"].tobytes()
        assert alone["ll"] == o["ll"] and alone["iterations"] == o["iterations"]


def test_convergence_in_the_drop_in_entry_point():
    """The drop-in VI call at the tracker's settings stops where the oracle
    does (graph replays, convergence polled on the device flags)."""
    f = _conv_frames()[1]
    em.regularizePrecisionMatrixConstants.setConstants(0.02, 0.2222, 9.0)
    scale = np.array([1, 1, 5], np.float32)
    d = em.GMEMinitializeMemoryWithSupervoxels(f.query, f.img, scale, f.N, f.S, f.centroids,
                                               f.sv_ptr)
    em.copyScaleToDvice(scale)
    gm = f.gm.copy()
    em.GMEMvariationalInferenceCUDAWithSupervoxels(d, gm, f.N, f.G, f.S, 400, 1e-6)
    em.GMEMreleaseMemoryWithSupervoxels(d)
    gm_o = O.em_run(f, O.make_opts(K=5, max_iter=400, tol=1e-6))[0]
    e = record_errors(gm, gm_o)
    assert max(e.values()) < TOL, e
    assert np.array_equal(gm["supervoxelIdx"], gm_o["supervoxelIdx"])


def test_death_rule_fires_as_in_the_oracle():
    """Gaussians far from every supervoxel with a tiny Dirichlet prior die in
    the first iteration (alpha / sum alpha < 1e-5, GMEM.cu:1813-1843): the
    same ones as in the oracle, with the reference's dead state."""
    f = copy.copy(synth.config1(seed=41))
    g = f.gm.copy()
    far = g[:12].copy()
    far["m_k"] = far["m_k"] + np.array([0.0, 0.0, 400.0])  # far above the volume
    far["m_o"] = far["m_k"]
    far["alpha_o"] = far["alpha_k"] = 1e-3
    f.gm = np.concatenate([g, far]).astype(GMM_DTYPE)
    out, _ = _solve([f], 5, 6)
    o = out[0]
    gm_o = O.em_run(f, O.make_opts(K=5, max_iter=6))[0]
    dead = o["gm"]["m_o"][:, 0] < -1e31
    print("dead: %d" % dead.sum(), flush=True)
    assert dead[-12:].all() and dead.sum() >= 12
    assert np.array_equal(dead, gm_o["m_o"][:, 0] < -1e31)
    assert (o["gm"]["alpha_k"][dead] == 0).all()
    assert (o["gm"]["W_k"][dead] == [1, 0, 0, 1, 0, 1]).all()
    e = record_errors(o["gm"], gm_o)
    assert max(e.values()) < TOL, e
