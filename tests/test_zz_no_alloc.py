"""GPU test of the allocation-free marshalling entry point
(tgmm_marshal_frame_noalloc): through the C ABI, two calls in a row (the
second on a smaller frame) give bitwise the outputs of tgmm_marshal_frame on
the same device inputs; sizes beyond its scratch array are refused."""
import ctypes

import numpy as np
import pytest

from helpers import raw_from_frame
from tgmm_amd import em, synth

pytestmark = pytest.mark.gpu


def _marshal(fn, fr, dims, seed):
    idx, raw, ptr, thr, d = raw_from_frame(fr, dims, seed=seed)
    dev = (em.DeviceArray(idx), em.DeviceArray(raw), em.DeviceArray(thr))
    f, keep = em.raw_frame(dev[0], dev[1], ptr, dev[2], d)
    N, S = fr.N, fr.S
    outs = (em.DeviceArray(nbytes=12 * N), em.DeviceArray(nbytes=4 * N),
            em.DeviceArray(nbytes=12 * S), em.DeviceArray(nbytes=8 * (S + 1)))
    rc = fn(ctypes.byref(f), outs[0].ptr, outs[1].ptr, outs[2].ptr, outs[3].ptr)
    assert rc == 0, em.load().tgmm_last_error()
    return (outs[0].numpy(np.float32)[: 3 * N], outs[1].numpy(np.float32)[:N],
            outs[2].numpy(np.float32)[: 3 * S], outs[3].numpy(np.int64)[: S + 1])


def test_noalloc_equals_marshal_frame_twice_in_a_row():
    L = em.load()
    L.tgmm_marshal_frame_noalloc.argtypes = L.tgmm_marshal_frame.argtypes
    big = synth.config1(seed=71, n_cells=300, target_voxels=60000)
    small = synth.config1(seed=72, n_cells=40, target_voxels=5000)
    dims = (512, 512, 100)
    for fr, seed in ((big, 1), (small, 2)):
        print("noalloc marshal N=%d S=%d" % (fr.N, fr.S), flush=True)
        want = _marshal(L.tgmm_marshal_frame, fr, dims, seed)
        got = _marshal(L.tgmm_marshal_frame_noalloc, fr, dims, seed)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
        assert np.array_equal(got[0], fr.query)


def test_noalloc_refuses_frames_beyond_its_scratch():
    L = em.load()
    L.tgmm_marshal_frame_noalloc.argtypes = L.tgmm_marshal_frame.argtypes
    S = 2_100_000  # one voxel each; the supervoxel pointer alone needs 16.8 MB
    ptr = np.arange(S + 1, dtype=np.int64)
    idx = em.DeviceArray(np.arange(S, dtype=np.uint64))
    raw = em.DeviceArray(np.full(S, 500, np.uint16))
    thr = em.DeviceArray(np.zeros(S, np.uint16))
    f, keep = em.raw_frame(idx, raw, ptr, thr, (2048, 2048, 1))
    outs = (em.DeviceArray(nbytes=12 * S), em.DeviceArray(nbytes=4 * S),
            em.DeviceArray(nbytes=12 * S))
    rc = L.tgmm_marshal_frame_noalloc(ctypes.byref(f), outs[0].ptr, outs[1].ptr, outs[2].ptr,
                                      None)
    assert rc == -2  # TGMM_ERR_LIMIT
    assert b"scratch" in L.tgmm_last_error()
