"""GPU tests of degenerate inputs: fewer Gaussians than candidates, empty
supervoxels, fixed and dead input Gaussians (against the oracle), and frames
with no supervoxels or no Gaussians inside a batch (skipped: the reference
has no defined behaviour there, its kNN launch of an empty grid fails)."""
import numpy as np
import pytest

import oracle_lib as O
from helpers import record_errors, rel_err
from tgmm_amd import em, synth

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def base():
    return synth.config1(seed=515, n_cells=30, target_voxels=6000)


def _run(frames, K=5, iters=6):
    with em.Batch(frames, em.make_opts(K=K, max_iter=iters)) as b:
        b.solve(em.SOLVE_GRAPH)
        return b.fetch()


def _check(frame, K=5, iters=6):
    print("frame N=%d S=%d G=%d K=%d" % (frame.N, frame.S, frame.G, K), flush=True)
    out = _run([frame], K, iters)[0]
    gm_o, rnk_o, ind_o, hist, it, ll = O.em_run(frame, O.make_opts(K=K, max_iter=iters))
    assert out["iterations"] == it == iters
    e = record_errors(out["gm"], gm_o)
    assert max(e.values()) < TOL, e
    assert rel_err(out["ll"], ll) < TOL
    return out, ind_o


@pytest.mark.parametrize("G", [1, 3])
def test_fewer_gaussians_than_candidates(base, G):
    f = synth.Frame(base.query, base.img, base.centroids, base.sv_ptr, base.gm[:G].copy())
    out, ind_o = _check(f, K=5)
    # slots beyond the G candidates are padding (-1, r = 0)
    ind = out["ind"].reshape(5, f.S)
    assert np.all(ind[G:] == -1) and np.all(out["rnk"].reshape(5, f.S)[G:] == 0)
    assert np.array_equal(out["ind"][: G * f.S], ind_o[: G * f.S])


def test_empty_supervoxels(base):
    # every third supervoxel gets an empty twin (zero voxels)
    ptr = list(base.sv_ptr)
    new_ptr, cent = [ptr[0]], []
    c = base.centroids.reshape(3, base.S)
    for s in range(base.S):
        new_ptr.append(ptr[s + 1])
        cent.append(c[:, s])
        if s % 3 == 0:
            new_ptr.append(ptr[s + 1])
            cent.append(c[:, s])
    cent = np.stack(cent, axis=1).astype(np.float32).reshape(-1)
    f = synth.Frame(base.query, base.img, cent, np.array(new_ptr, np.int64), base.gm.copy())
    _check(f)


def test_fixed_and_dead_input_gaussians(base):
    gm = base.gm.copy()
    gm["fixed"][::7] = 1
    dead = np.arange(3, gm.shape[0], 11)
    gm["m_o"][dead] = -1e32
    gm["m_k"][dead] = -1e32
    gm["alpha_k"][dead] = 0.0
    f = synth.Frame(base.query, base.img, base.centroids, base.sv_ptr, gm)
    out, _ = _check(f)
    # the flags come back as given, and fixed Gaussians keep their posterior
    # (GMEM.cu:1610-1611 and 1702-1703: the update and the regularisation skip them)
    assert np.array_equal(out["gm"]["fixed"], gm["fixed"])
    fx = gm["fixed"] == 1
    live = fx & (gm["m_o"][:, 0] > -1e31)
    assert np.array_equal(out["gm"]["m_k"][live], gm["m_k"][live])
    assert np.array_equal(out["gm"]["W_k"][live], gm["W_k"][live])
    # dead ones stay dead and are never candidates
    assert np.all(out["gm"]["m_o"][dead, 0] < -1e31)
    assert not np.isin(out["ind"], dead).any()


def test_empty_and_gaussianless_frames_are_skipped(base):
    empty = synth.Frame(np.zeros(0, np.float32), np.zeros(0, np.float32),
                        np.zeros(0, np.float32), np.zeros(1, np.int64), base.gm[:2].copy())
    no_g = synth.Frame(base.query, base.img, base.centroids, base.sv_ptr, base.gm[:0].copy())
    outs = _run([base, empty, no_g])
    alone = _run([base])[0]
    assert outs[0]["gm"].tobytes() == alone["gm"].tobytes() and outs[0]["ll"] == alone["ll"]
    # no supervoxels: not iterated, records come back as given
    assert outs[1]["iterations"] == 0
    for k in ("m_k", "W_k", "nu_k", "beta_k", "alpha_k"):
        assert np.array_equal(outs[1]["gm"][k], empty.gm[k]), k
    # no Gaussians: not iterated, every slot padding
    assert outs[2]["iterations"] == 0
    assert np.all(outs[2]["ind"] == -1) and np.all(outs[2]["rnk"] == 0)
