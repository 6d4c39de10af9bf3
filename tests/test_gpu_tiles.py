"""GPU tests of the tunable launch shapes: every valid tile shape gives
bitwise the same results (the kernels fix their summation order
independently of the shape), and the autotuner returns a valid choice that
changes nothing but speed."""
import numpy as np
import pytest

from helpers import dup_gaussians
from tgmm_amd import em, synth

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(ll_group=16, ll_unroll=4, mstep_lanes=16, sv_block=256, knn_block=128, knn_chunk=4,
         knn_cell=2, knn_lanes=0),
    dict(ll_group=32, ll_unroll=2, mstep_lanes=8, sv_block=128, knn_block=64, knn_chunk=8,
         knn_cell=6, knn_lanes=2),
    dict(ll_group=64, ll_unroll=1, mstep_lanes=4, sv_block=64, knn_block=128, knn_chunk=4,
         knn_cell=4, knn_lanes=4),
    dict(ll_group=16, ll_unroll=8, mstep_lanes=16, sv_block=128, knn_block=256, knn_chunk=8,
         knn_cell=2, knn_lanes=2),
    dict(ll_group=64, ll_unroll=2, mstep_lanes=4, sv_block=256, knn_block=256, knn_chunk=4,
         knn_cell=6, knn_lanes=0),
]


@pytest.fixture(scope="module")
def frames():
    c1 = synth.config1()
    cells = synth.sample_cells(99, 120, (256.0, 256.0, 60.0), 5.0, 3.0, 6.0)
    soft = dup_gaussians(synth.rasterize(100, cells, (256.0, 256.0, 60.0), 5.0, 20000,
                                         max_sv_size=4))
    # supervoxels of up to 200 voxels: several passes of every unroll
    cells = synth.sample_cells(41, 60, (256.0, 256.0, 60.0), 5.0, 4.0, 7.0)
    big = synth.rasterize(42, cells, (256.0, 256.0, 60.0), 5.0, 30000, max_sv_size=200)
    assert np.diff(big.sv_ptr).max() > 64
    return [c1, soft, big]


def _run(frames, K, tiles, iters=4):
    with em.Batch(frames, em.make_opts(K=K, max_iter=iters, tol=-1.0)) as b:
        if tiles is not None:
            b.set_tiles(tiles)
            assert b.get_tiles() == tiles
        b.solve(em.SOLVE_GRAPH)
        return b.fetch()


def _same(a, b):
    for x, y in zip(a, b):
        assert x["gm"].tobytes() == y["gm"].tobytes()
        assert np.array_equal(x["rnk"], y["rnk"]) and np.array_equal(x["ind"], y["ind"])
        assert x["ll"] == y["ll"]


@pytest.mark.parametrize("K", [5, 8])
def test_every_tile_shape_gives_identical_results(frames, K):
    ref = _run(frames, K, None)
    for t in SHAPES:
        print("tiles", t, flush=True)
        _same(ref, _run(frames, K, t))


def test_invalid_tiles_are_rejected(frames):
    with em.Batch(frames[:1], em.make_opts(K=5, max_iter=2)) as b:
        bad = dict(em.DEFAULT_TILES, ll_group=8)
        with pytest.raises(em.TgmmError):
            b.set_tiles(bad)
        bad = dict(em.DEFAULT_TILES, ll_group=32, ll_unroll=8)
        with pytest.raises(em.TgmmError):
            b.set_tiles(bad)


def test_autotune_choice_is_valid_and_changes_nothing(frames):
    ref = _run(frames, 5, None, iters=5)
    with em.Batch(frames, em.make_opts(K=5, max_iter=5, tol=-1.0)) as b:
        tiles, report = b.autotune(reps=2, cache=False)
        print(tiles, "\n", report, flush=True)
        assert b.get_tiles() == tiles
        assert len(report.strip().splitlines()) == 34
        b.solve(em.SOLVE_GRAPH)
        out = b.fetch()
    _same(ref, out)


def _run_voxel(frames, K, tiles, iters=4):
    with em.Batch(frames, em.make_opts(K=K, max_iter=iters, tol=-1.0, voxel=True)) as b:
        if tiles is not None:
            b.set_tiles(tiles)
            assert b.get_tiles() == tiles
        else:
            t = b.get_tiles()
            assert (t["ll_group"], t["ll_unroll"]) == (1, 1)  # voxel mode's default
        b.solve(em.SOLVE_GRAPH)
        return b.fetch()


@pytest.mark.parametrize("K", [5, 8])
def test_voxel_mode_lane_per_voxel_equals_group_kernel(frames, K):
    # the lane-per-voxel log-likelihood (the default in voxel mode) against
    # the 16-, 32- and 64-lane group kernels on the same one-voxel supervoxels
    vf = [frames[0], synth.config1(seed=5, n_cells=60, target_voxels=9000)]
    ref = _run_voxel(vf, K, None)
    for gw, u in ((16, 4), (32, 2), (64, 1)):
        t = dict(em.DEFAULT_TILES, ll_group=gw, ll_unroll=u)
        print("voxel mode, ll tiles", gw, u, flush=True)
        _same(ref, _run_voxel(vf, K, t))


def test_lane_per_voxel_needs_voxel_mode(frames):
    with em.Batch(frames[:1], em.make_opts(K=5, max_iter=2)) as b:
        with pytest.raises(em.TgmmError):
            b.set_tiles(dict(em.DEFAULT_TILES, ll_group=1, ll_unroll=1))
    with em.Batch(frames[:1], em.make_opts(K=5, max_iter=2, voxel=True)) as b:
        with pytest.raises(em.TgmmError):
            b.set_tiles(dict(em.DEFAULT_TILES, ll_group=1, ll_unroll=2))


def test_voxel_mode_autotune_includes_lane_per_voxel(frames):
    vf = frames[:1]
    ref = _run_voxel(vf, 5, None, iters=3)
    with em.Batch(vf, em.make_opts(K=5, max_iter=3, tol=-1.0, voxel=True)) as b:
        tiles, report = b.autotune(reps=2, cache=False)
        print(tiles, "\n", report, flush=True)
        assert report.splitlines()[0].startswith("loglik 1x1")
        assert len(report.strip().splitlines()) == 35
        b.solve(em.SOLVE_GRAPH)
        out = b.fetch()
    _same(ref, out)
