"""GPU tests of the continuous-batching server: results equal the batch API
bitwise, device-side prior propagation equals the host restatement of
updatePriors, and every frame's result is the same alone and in the mix."""
import numpy as np
import pytest

from helpers import device_alpha_sum, next_priors
from tgmm_amd import em, synth

pytestmark = pytest.mark.gpu

PCT = (0.5, 0.5, 0.8)


def opts(iters=5):
    return em.make_opts(K=5, max_iter=iters, tol=-1.0)


def server(n_slots=4):
    return em.Server(opts(), hbm_budget_bytes=1 << 30, max_frames_in_flight=n_slots,
                     beta_pct=PCT[0], nu_pct=PCT[1], alpha_pct=PCT[2])


def same(a, b):
    assert a["gm"].tobytes() == b["gm"].tobytes()
    assert np.array_equal(a["rnk"], b["rnk"]) and np.array_equal(a["ind"], b["ind"])
    assert a["ll"] == b["ll"] and a["iterations"] == b["iterations"]


def batch_one(f, gm):
    f2 = synth.Frame(f.query, f.img, f.centroids, f.sv_ptr, gm)
    with em.Batch([f2], opts()) as b:
        b.solve(em.SOLVE_GRAPH)
        return b.fetch()[0]


def test_independent_frames_match_batch():
    frames = [synth.config1(seed=s, n_cells=40 + 10 * s, target_voxels=6000 + 1000 * s)
              for s in range(1, 4)]
    with server() as sv:
        tickets = [sv.submit(f) for f in frames]
        sv.wait(-1)
        outs = [sv.fetch(t, f.G) for t, f in zip(tickets, frames)]
    for f, o in zip(frames, outs):
        same(o, batch_one(f, f.gm))


def test_stream_priors_on_device_match_host_restatement():
    fr = synth.timelapse(5, 60, 3, (256.0, 256.0, 60.0), voxels_per_cell=150)
    with server() as sv:
        t0 = sv.submit(fr[0], stream=0)
        t1 = sv.submit(fr[1], stream=0, with_gaussians=False)
        t2 = sv.submit(fr[2], stream=0, with_gaussians=False)
        sv.wait(-1)
        o = [sv.fetch(t, 60) for t in (t0, t1, t2)]
    ref0 = batch_one(fr[0], fr[0].gm)
    same(o[0], ref0)
    prev = ref0
    for t in (1, 2):
        # the stream's alpha total is the device's frame sum (its own order)
        tot = device_alpha_sum(prev["gm"]["alpha_k"])
        assert abs(tot - prev["gm"]["alpha_k"].sum()) <= 1e-13 * tot
        g = next_priors(prev["gm"], PCT[0], PCT[1], PCT[2], tot)
        ref = batch_one(fr[t], g)
        same(o[t], ref)
        prev = ref


def test_each_result_same_alone_and_in_the_mix():
    streams = [synth.timelapse(11 + i, 40 + 15 * i, 3, (200.0, 200.0, 50.0),
                               voxels_per_cell=120) for i in range(3)]
    alone = []
    for i, fr in enumerate(streams):
        with server(n_slots=1) as sv:
            ts = [sv.submit(fr[0], stream=i)] + [sv.submit(f, stream=i, with_gaussians=False)
                                                 for f in fr[1:]]
            sv.wait(-1)
            alone.append([sv.fetch(t, fr[0].G) for t in ts])
    with server(n_slots=3) as sv:
        ts = {}
        for t in range(3):  # interleaved arrivals
            for i, fr in enumerate(streams):
                ts[(i, t)] = sv.submit(fr[t], stream=i, with_gaussians=(t == 0))
                sv.step()
        sv.wait(-1)
        mix = {k: sv.fetch(v, streams[k[0]][0].G) for k, v in ts.items()}
    for i in range(3):
        for t in range(3):
            same(mix[(i, t)], alone[i][t])
