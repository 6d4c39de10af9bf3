"""GPU test of the C++ drop-in: tests/native/compat_frame, a program that calls
the reference's own C++ names through include/tgmm_em_compat.hpp,
include/tgmm_utils_knn_compat.hpp and include/tgmm_knn_compat.hpp (device
selection, the EM's init, solve and release, the local log-likelihood and the
supervoxel neighbour search, in the tracker's order, then the voxel-mode EM),
runs as a child process on one frame; its records, local log-likelihoods and
neighbour slots equal, bitwise, those of the same calls made through the
Python mirror of the C ABI."""
import os
import subprocess

import numpy as np
import pytest

from helpers import records_equal
from tgmm_amd import em, synth
from tgmm_amd.records import GMM_DTYPE

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "compat_frame")
SCALE = np.array([1.0, 1.0, 5.0], np.float32)
LAM = (0.02, 0.2222, 9.0)
ITERS = 5


def test_cpp_compat_program_equals_python_path(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/native/compat_frame is not built (python -m tgmm_amd.build)")
    fr = synth.config1(seed=2024, n_cells=50, target_voxels=12000)
    N, S, G = fr.N, fr.S, fr.G
    rng = np.random.default_rng(1)
    lists = [list(rng.integers(0, S, rng.integers(0, 12))) for _ in range(40)]
    lptr = np.zeros(len(lists) + 1, np.int64)
    lptr[1:] = np.cumsum([len(l) for l in lists])
    lidx = np.concatenate([np.asarray(l, np.int32) for l in lists]).astype(np.int32)
    inp, out = tmp_path / "frame.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        fh.write(np.array([N, S, G, len(lists), lidx.size, ITERS], np.int64).tobytes())
        for a, t in ((fr.query, np.float32), (fr.img, np.float32), (fr.centroids, np.float32),
                     (SCALE, np.float32), (fr.sv_ptr, np.int64)):
            fh.write(np.ascontiguousarray(a, t).tobytes())
        fh.write(em.as_records(fr.gm).tobytes())
        fh.write(lptr.tobytes())
        fh.write(lidx.tobytes())
        fh.write(np.array(LAM, np.float64).tobytes())
    print("compat program: N=%d S=%d G=%d" % (N, S, G), flush=True)
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    raw = np.fromfile(out, np.uint8)
    o = 0
    gm_cpp = np.frombuffer(raw[o:o + G * GMM_DTYPE.itemsize].tobytes(), GMM_DTYPE)
    o += G * GMM_DTYPE.itemsize
    ll_cpp = np.frombuffer(raw[o:o + 8 * len(lists)].tobytes(), np.float64)
    o += 8 * len(lists)
    ind_cpp = np.frombuffer(raw[o:o + 44 * S].tobytes(), np.int32)
    o += 44 * S
    dist_cpp = np.frombuffer(raw[o:o + 44 * S].tobytes(), np.float32)
    o += 44 * S
    gv_cpp = np.frombuffer(raw[o:o + G * GMM_DTYPE.itemsize].tobytes(), GMM_DTYPE)
    o += G * GMM_DTYPE.itemsize
    dev_cpp = int(np.frombuffer(raw[o:o + 4].tobytes(), np.int32)[0])

    # the same calls through the Python mirror
    em.regularizePrecisionMatrixConstants.setConstants(*LAM)
    d = em.GMEMinitializeMemoryWithSupervoxels(fr.query, fr.img, SCALE, N, S, fr.centroids,
                                               fr.sv_ptr)
    em.copyScaleToDvice(SCALE)
    gm = em.as_records(fr.gm).copy()
    em.GMEMvariationalInferenceCUDAWithSupervoxels(d, gm, N, G, S, ITERS, -1.0)
    ll = em.calculateLocalLikelihood([], lists, d.queryCUDA, d.imgDataCUDA, d.rnkCUDA,
                                     d.indCUDA, gm, d.labelListPtrCUDA, N, G, S)
    em.GMEMreleaseMemoryWithSupervoxels(d)
    ind, dist = em.knnCUDA_(fr.centroids, fr.centroids, 11, SCALE)

    assert records_equal(gm_cpp, gm)
    assert list(ll_cpp) == ll
    assert np.array_equal(ind_cpp, ind) and np.array_equal(dist_cpp, dist)
    # the device the tracker selects: the visible one with the most memory
    assert dev_cpp == 0
    # voxel mode through the same names equals the batch API's voxel mode
    with em.Batch([fr], em.make_opts(K=5, max_iter=ITERS, tol=-1.0, voxel=True)) as b:
        b.solve(em.SOLVE_GRAPH)
        gv = b.fetch()[0]["gm"]
    assert records_equal(gv_cpp, gv)
