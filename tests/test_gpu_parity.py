"""GPU parity: the HIP library against the FP64 oracle on seeded frames."""
import copy

import numpy as np
import pytest

import oracle_lib as O
from helpers import dup_gaussians, record_errors, rel_err
from tgmm_amd import em, synth
from tgmm_amd.records import GMM_DTYPE

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative, means / precisions / nu / beta / alpha / log-likelihood


@pytest.fixture(scope="module")
def c1():
    return synth.config1()


def _opts(K=5, iters=10, tol=-1.0):
    return em.make_opts(K=K, max_iter=iters, tol=tol)


def _oopts(K=5, iters=10, tol=-1.0):
    return O.make_opts(K=K, max_iter=iters, tol=tol)


def _run_batch(frames, opts, mode=em.SOLVE_GRAPH):
    with em.Batch(frames, opts) as b:
        b.solve(mode)
        return b.fetch()


def test_knn_bit_exact_first_iteration(c1):
    """After one iteration, ind holds the kNN of the initial means."""
    out = _run_batch([c1], _opts(iters=1))[0]
    ind_ref = np.zeros(5 * c1.S, np.int32)
    O.oracle().orc_knn(c1.S, O.p(c1.centroids, O._fp), c1.G, c1.gm.ctypes.data,
                       O.p(np.array([1, 1, 5], np.float32), O._fp), 5, O.p(ind_ref, O._ip))
    assert np.array_equal(out["ind"], ind_ref)


def _dense(c1):
    """4 copies of each mean with quarter-voxel offsets: ties and near ties."""
    f = copy.copy(c1)
    rng = np.random.default_rng(3)
    gms = [c1.gm]
    for r in range(3):
        g = c1.gm.copy()
        g["m_k"] += np.round(rng.normal(scale=2.0, size=(c1.G, 3)) * 4) / 4
        g["m_o"] = g["m_k"]
        gms.append(g)
    f.gm = np.concatenate(gms).astype(GMM_DTYPE)
    return f


def _brute_knn(f, gm, K):
    ind_ref = np.zeros(K * f.S, np.int32)
    O.oracle().orc_knn(f.S, O.p(f.centroids, O._fp), gm.shape[0], gm.ctypes.data,
                       O.p(np.array([1, 1, 5], np.float32), O._fp), K, O.p(ind_ref, O._ip))
    return ind_ref


@pytest.mark.parametrize("K", [5, 8])
def test_knn_bit_exact_dense(c1, K):
    """Many Gaussians per cell (ties and near ties): kNN must equal brute
    force exactly."""
    f = _dense(c1)
    out = _run_batch([f], _opts(K=K, iters=1))[0]
    assert np.array_equal(out["ind"], _brute_knn(f, f.gm, K))


@pytest.mark.parametrize("K", [5, 8])
def test_knn_bit_exact_with_previous_candidates(c1, K):
    """From the second iteration on, the previous candidates bound the search
    and prune grid cells: the candidates of iteration n equal brute force on
    the means iteration n starts from (the final means of an n-1 iteration
    run of the same frame)."""
    f = _dense(c1)
    for n in (2, 4):
        prev = _run_batch([f], _opts(K=K, iters=n - 1))[0]
        cur = _run_batch([f], _opts(K=K, iters=n))[0]
        assert np.array_equal(cur["ind"], _brute_knn(f, np.ascontiguousarray(prev["gm"]), K))


def test_first_iteration_matches_oracle(c1):
    out = _run_batch([c1], _opts(iters=1))[0]
    gm_o, rnk_o, ind_o, hist, it, ll = O.em_run(c1, _oopts(iters=1))
    assert it == out["iterations"] == 1
    assert np.array_equal(out["ind"], ind_o)
    assert np.max(np.abs(out["rnk"] - rnk_o)) < 1e-6
    e = record_errors(out["gm"], gm_o)
    assert max(e.values()) < TOL, e
    assert rel_err(out["ll"], ll) < TOL


def test_ten_iterations_match_oracle(c1):
    out = _run_batch([c1], _opts(iters=10))[0]
    gm_o, rnk_o, ind_o, hist, it, ll = O.em_run(c1, _oopts(iters=10))
    assert out["iterations"] == it == 10
    e = record_errors(out["gm"], gm_o)
    assert max(e.values()) < TOL, e
    assert rel_err(out["ll"], ll) < TOL
    assert np.array_equal(out["gm"]["supervoxelNum"], gm_o["supervoxelNum"])


def test_fractional_coordinates_match_oracle(c1):
    """Every other voxel half a voxel off the integer grid along x: offsets
    from the supervoxels' anchors that are not integers, so the
    log-likelihood reads the full coordinates instead of the packed offsets;
    the EM still matches the oracle."""
    f = copy.copy(c1)
    f.query = c1.query.copy()
    f.query[: c1.N] += (np.arange(c1.N) % 2).astype(np.float32) * np.float32(0.5)
    out = _run_batch([f], _opts(iters=5))[0]
    gm_o, rnk_o, ind_o, hist, it, ll = O.em_run(f, _oopts(iters=5))
    e = record_errors(out["gm"], gm_o)
    assert max(e.values()) < TOL, e
    assert rel_err(out["ll"], ll) < TOL


def test_far_voxel_offsets_match_oracle():
    """A supervoxel whose voxels lie more than 127 voxels from its anchor
    (its first voxel, moved 200 voxels along x): offsets beyond the packed
    form's range, so the log-likelihood reads the full coordinates."""
    cells = synth.sample_cells(41, 60, (512.0, 512.0, 40.0), 5.0, 3.0, 6.0)
    f = synth.rasterize(42, cells, (512.0, 512.0, 40.0), 5.0, 12000, max_sv_size=40)
    f.query = f.query.copy()
    f.query[0] += np.float32(200.0)
    out = _run_batch([f], _opts(iters=5))[0]
    gm_o, rnk_o, ind_o, hist, it, ll = O.em_run(f, _oopts(iters=5))
    e = record_errors(out["gm"], gm_o)
    assert max(e.values()) < TOL, e
    assert rel_err(out["ll"], ll) < TOL


def soft_frame(max_sv=2, offset=1.0):
    """Tiny supervoxels and two Gaussians per cell: thousands of (s, k) pairs
    with responsibilities strictly between 1e-3 and 0.999."""
    cells = synth.sample_cells(99, 120, (256.0, 256.0, 60.0), 5.0, 3.0, 6.0)
    f = synth.rasterize(100, cells, (256.0, 256.0, 60.0), 5.0, 20000, max_sv_size=max_sv)
    return dup_gaussians(f, offset=offset)


@pytest.mark.parametrize("max_sv", [2, 4])
def test_soft_assignments_match_oracle(max_sv):
    f = soft_frame(max_sv)
    out = _run_batch([f], _opts(iters=10))[0]
    gm_o, rnk_o, ind_o, hist, it, ll = O.em_run(f, _oopts(iters=10))
    soft = ((rnk_o > 1e-3) & (rnk_o < 0.999)).sum()
    assert soft > 1000  # the case really exercises soft responsibilities
    e = record_errors(out["gm"], gm_o)
    assert max(e.values()) < TOL, e
    assert rel_err(out["ll"], ll) < TOL
    # responsibilities after 10 iterations: the FP32 r of each side feeds its
    # M-step, so the Gaussians drift apart by ~1e-7 relative and a score of a
    # few hundred nats moves by ~1e-5: r agrees to 1e-4 relative where it
    # enters the M-step (r > 1e-3) and to 1e-6 absolute below that
    same = out["ind"] == ind_o
    big = same & (rnk_o > 1e-3)
    assert rel_err(out["rnk"][big], rnk_o[big]) < 1e-4
    small = same & ~big
    assert np.max(np.abs(out["rnk"][small] - rnk_o[small])) < 1e-6


def test_modes_bitwise_identical(c1):
    a = _run_batch([c1], _opts(iters=5), em.SOLVE_GRAPH)[0]
    b = _run_batch([c1], _opts(iters=5), em.SOLVE_EAGER)[0]
    c = _run_batch([c1], _opts(iters=5), em.SOLVE_HOSTSYNC)[0]
    for x in (b, c):
        assert a["gm"].tobytes() == x["gm"].tobytes()
        assert np.array_equal(a["rnk"], x["rnk"]) and np.array_equal(a["ind"], x["ind"])
        assert a["ll"] == x["ll"]


def test_alone_equals_in_the_mix(c1):
    """Each frame's result in a 3-frame batch is bitwise its result alone."""
    f2 = dup_gaussians(c1)
    f3 = synth.config1(seed=12345, n_cells=50, target_voxels=9000)
    mix = _run_batch([c1, f2, f3], _opts(iters=6))
    for f, m in zip([c1, f2, f3], mix):
        alone = _run_batch([f], _opts(iters=6))[0]
        assert alone["gm"].tobytes() == m["gm"].tobytes()
        assert np.array_equal(alone["rnk"], m["rnk"])
        assert alone["ll"] == m["ll"]


def test_legacy_entry_points(c1, tmp_path):
    em.regularizePrecisionMatrixConstants.setConstants(0.02, 0.2222, 9.0)
    scale = np.array([1, 1, 5], np.float32)
    d = em.GMEMinitializeMemoryWithSupervoxels(c1.query, c1.img, scale, c1.N, c1.S,
                                               c1.centroids, c1.sv_ptr)
    em.copyScaleToDvice(scale)
    gm = c1.gm.copy()
    dbg = str(tmp_path / "rnk.bin")
    em.GMEMvariationalInferenceCUDAWithSupervoxels(d, gm, c1.N, c1.G, c1.S, 10, -1.0,
                                                   debugPath=dbg)
    rnk = em.copy_device_to_host(d.rnkCUDA, 5 * c1.S * 4, np.float32)
    ind = em.copy_device_to_host(d.indCUDA, 5 * c1.S * 4, np.int32)
    em.GMEMreleaseMemoryWithSupervoxels(d)
    assert d.queryCUDA.value is None
    gm_o, rnk_o, ind_o, hist, it, ll = O.em_run(c1, _oopts(iters=10))
    e = record_errors(gm, gm_o)
    assert max(e.values()) < TOL, e
    assert np.array_equal(ind, ind_o)
    assert np.max(np.abs(rnk - rnk_o)) < 1e-5
    assert np.array_equal(gm["supervoxelIdx"], gm_o["supervoxelIdx"])
    raw = open(dbg, "rb").read()
    hdr = np.frombuffer(raw[:8], np.int32)
    assert hdr[0] == c1.S and hdr[1] == 5
    assert np.array_equal(np.frombuffer(raw[8 + 20 * c1.S:], np.int32), ind)


def test_legacy_requires_constants_and_limits(c1):
    L = em.load()
    # ref_nb above the reference limit -> TGMM_ERR_LIMIT, no exit()
    rc = L.tgmm_GMEMvariationalInferenceCUDAWithSupervoxels(
        None, None, None, None, None, None, None, None, None, 10, 70000, 0, 1, -1.0, 0, 0,
        1, None)
    assert rc == -2


def test_repeatable_and_order_independent(c1):
    """Three runs of a 3-frame batch are bitwise identical, and permuting the
    frames inside the batch changes nothing (per-frame reductions use every
    block of the frame's aligned range, in a fixed order)."""
    f3 = synth.config1(seed=777, n_cells=37, target_voxels=7000)
    f2 = dup_gaussians(f3)
    frames = [c1, f2, f3]
    runs = [_run_batch(frames, _opts(iters=4)) for _ in range(3)]
    perm = _run_batch([f3, c1, f2], _opts(iters=4))
    for i in range(3):
        for r in runs[1:]:
            assert runs[0][i]["gm"].tobytes() == r[i]["gm"].tobytes()
            assert runs[0][i]["ll"] == r[i]["ll"]
    for a, b in zip([runs[0][2], runs[0][0], runs[0][1]], perm):
        assert a["gm"].tobytes() == b["gm"].tobytes()
        assert a["ll"] == b["ll"]


# ------------------------------------------------ reference slot layout --
def _solve(frames, opts):
    with em.Batch(frames, opts) as b:
        b.solve(em.SOLVE_GRAPH)
        return b.fetch(), b.stats()


def _lattice(c1, step=3.0):
    """Integer centroids and Gaussian means on a lattice: many Gaussians at
    exactly the K-th distance of a supervoxel (ties the reference's slot rule
    resolves by position)."""
    f = copy.copy(c1)
    f.centroids = np.round(c1.centroids).astype(np.float32)
    g = c1.gm.copy()
    g["m_k"] = np.round(g["m_k"] / step) * step
    g["m_o"] = g["m_k"]
    f.gm = np.concatenate([g, g.copy()]).astype(GMM_DTYPE)  # every mean twice
    return f


@pytest.mark.parametrize("K", [5, 8])
def test_knn_ties_take_the_reference_scan_set(c1, K):
    """With exact ties at the K-th distance the candidate set is the one the
    reference's scan keeps (not the lower indices): the tied supervoxels are
    searched again in scan order, and the slots equal the oracle's replay."""
    f = _lattice(c1)
    out, st = _solve([f], _opts(K=K, iters=1))
    assert st["knn_ties_last_iteration"] > 0  # the tie path ran
    assert st["layout_mismatches"] == 0
    assert np.array_equal(out[0]["ind"], _brute_knn(f, f.gm, K))


@pytest.mark.parametrize("K", [5, 8])
def test_final_slot_layout_after_several_iterations(c1, K):
    """After n iterations the slots hold the last search's candidates in the
    reference's layout: the oracle's scan over the means iteration n started
    from (the final means of an (n-1)-iteration run)."""
    f = _dense(c1)
    for n in (2, 5):
        prev, _ = _solve([f], _opts(K=K, iters=n - 1))
        cur, st = _solve([f], _opts(K=K, iters=n))
        assert st["layout_mismatches"] == 0
        assert np.array_equal(cur[0]["ind"],
                              _brute_knn(f, np.ascontiguousarray(prev[0]["gm"]), K))


@pytest.mark.parametrize("max_sv", [2, 4])
def test_capped_supervoxel_lists_match_oracle(max_sv):
    """Gaussians that collect more than 16 supervoxels keep the first 16 of
    the reference's slot-major scan (GMEM.cu:4328-4350): the device's lists
    equal the oracle's rule applied to the device's own rnk/ind, and the full
    oracle EM's lists."""
    f = soft_frame(max_sv)
    out, st = _solve([f], _opts(iters=10))
    o = out[0]
    assert st["layout_mismatches"] == 0
    g = f.gm.copy()
    O.oracle().orc_assign_supervoxels(f.S, f.G, g.ctypes.data, 5, O.p(o["rnk"], O._fp),
                                      O.p(o["ind"], O._ip), 16)
    assert np.array_equal(o["gm"]["supervoxelNum"], g["supervoxelNum"])
    assert np.array_equal(o["gm"]["supervoxelIdx"], g["supervoxelIdx"])
    gm_o, rnk_o, ind_o, hist, it, ll = O.em_run(f, _oopts(iters=10))
    capped = gm_o["supervoxelNum"] == 16
    print("capped Gaussians: %d" % capped.sum(), flush=True)
    assert capped.sum() > 20
    assert np.array_equal(o["gm"]["supervoxelNum"], gm_o["supervoxelNum"])
    assert np.array_equal(o["gm"]["supervoxelIdx"], gm_o["supervoxelIdx"])
