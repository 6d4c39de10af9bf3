"""C-ABI checks that need no GPU: the library loads, exports every function
include/tgmm_em.h declares, keeps the 288-byte record layout, and reports
argument errors as status codes (never exit())."""
import ctypes
import os
import re

import numpy as np
import pytest

from tgmm_amd import em, synth
from tgmm_amd.records import GMM_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tgmm_em.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(tgmm_\w+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    L = em.load()
    names = declared_functions()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_record_layout():
    offs = {"beta_k": 0, "nu_k": 8, "alpha_k": 16, "m_k": 24, "W_k": 48,
            "expectedLogDetCovarianceCUDA": 96, "expectedLogResponsivityCUDA": 104,
            "splitScore": 112, "beta_o": 120, "nu_o": 128, "alpha_o": 136, "m_o": 144,
            "W_o": 168, "fixed": 216, "supervoxelIdx": 220, "supervoxelNum": 284}
    for k, v in offs.items():
        assert GMM_DTYPE.fields[k][1] == v, k
    assert GMM_DTYPE.itemsize == 288
    assert synth._load().tgmm_synth_record_size() == 288


def test_device_count_without_gpu_is_not_an_error():
    assert em.device_count() >= 0


def test_bad_options_are_status_codes():
    L = em.load()
    f = synth.config1(n_cells=10, target_voxels=2000)
    for K, code in ((0, -1), (17, -1)):
        with pytest.raises(em.TgmmError) as e:
            em.Batch([f], em.make_opts(K=K))
        assert e.value.code == code
    with pytest.raises(em.TgmmError) as e:
        em.Batch([f], em.make_opts(lambda_min=-1.0))
    assert e.value.code == -5


def test_legacy_limits_and_missing_constants_are_status_codes():
    L = em.load()
    rc = L.tgmm_GMEMvariationalInferenceCUDAWithSupervoxels(
        None, None, None, None, None, None, None, None, None, 10, 70000, 0, 1, -1.0, 0, 0, 1,
        None)
    assert rc == -2  # ref_nb > 65535
    rc = L.tgmm_GMEMvariationalInferenceCUDAWithSupervoxels(
        None, None, None, None, None, None, None, None, None, 65535 * 1024 + 1, 10, 0, 1, -1.0,
        0, 0, 1, None)
    assert rc == -2  # query_nb above 1024 * 65535
    assert b"Too many" in L.tgmm_last_error()


def test_local_likelihood_argument_errors_are_status_codes():
    """Host-side checks of tgmm_calculateLocalLikelihood, before any device
    call: malformed list pointers and indices outside [0, numLabels)."""
    L = em.load()
    out = np.full(2, 5.0)
    ll = ctypes.c_longlong

    def call(ptr, idx, n, S=10, G=0):
        ptr = np.asarray(ptr, np.int64)
        idx = np.asarray(idx if len(idx) else [0], np.int32)
        return L.tgmm_calculateLocalLikelihood(out.ctypes.data, ptr.ctypes.data,
                                               idx.ctypes.data, n, None, None, None, None,
                                               None, None, 100, G, S)
    assert call([0, 1], [3], -1) == -1        # negative list count
    assert call([1, 1, 1], [3], 2) == -1      # list_ptr[0] != 0
    assert call([0, 2, 1], [3, 4], 2) == -1   # decreasing
    assert call([0, 1, 2], [3, 10], 2) == -1  # index == numLabels
    assert b"outside" in L.tgmm_last_error()
    assert call([0, 1, 2], [-1, 2], 2) == -1
    assert call([0, 1, 2], [1, 2], 2) == -1   # NULL device arrays
    assert call([0, 0, 0], [], 2) == 0        # every list empty: nothing to do
    assert call([0], [], 0) == 0
    assert np.all(out == 5.0)


def test_marshal_and_knn_argument_errors_are_status_codes():
    """Host-side checks of tgmm_marshal_frame, tgmm_GMEMinitializeMemoryFromRaw
    and tgmm_knnCUDA_ that return before any device call."""
    L = em.load()

    def frame(N, ptr, dims=(4, 3, 2)):
        idx = np.arange(max(N, 1), dtype=np.uint64)
        raw = np.ones(max(N, 1), np.uint16)
        thr = np.zeros(max(len(ptr) - 1, 1), np.uint16)
        f, keep = em.raw_frame(idx[:N] if N else idx, raw[:N] if N else raw, ptr, thr, dims)
        f.N = N
        return f, keep + [idx, raw, thr]

    out = np.zeros(64, np.float32)
    ptr = out.ctypes.data
    f, k = frame(3, [0, 2, 3])
    f.dims[0] = 0
    assert L.tgmm_marshal_frame(ctypes.byref(f), ptr, ptr, ptr, None) == -1  # dims
    f, k = frame(3, [0, 2, 4])
    assert L.tgmm_marshal_frame(ctypes.byref(f), ptr, ptr, ptr, None) == -1  # ptr[S] != N
    f, k = frame(3, [0, 2, 1, 3])
    assert L.tgmm_marshal_frame(ctypes.byref(f), ptr, ptr, ptr, None) == -1  # decreasing
    f, k = frame(30, [0, 30])
    assert L.tgmm_marshal_frame(ctypes.byref(f), ptr, ptr, ptr, None) == -1  # > 24 voxels
    assert b"disjoint" in L.tgmm_last_error()
    f, k = frame(3, [0, 3])
    f.dims[:] = [1 << 20, 1 << 20, 1 << 14]
    assert L.tgmm_marshal_frame(ctypes.byref(f), ptr, ptr, ptr, None) == -2  # 2^54 voxels
    # the drop-in initialiser counts supervoxels in an unsigned short
    big = np.zeros(70001, np.int64)
    f, k = frame(0, big)
    pp = [ctypes.c_void_p() for _ in range(7)]
    sc = np.ones(3, np.float32)
    rc = L.tgmm_GMEMinitializeMemoryFromRaw(*[ctypes.byref(x) for x in pp[:2]], sc.ctypes.data,
                                            *[ctypes.byref(x) for x in pp[2:]], ctypes.byref(f))
    assert rc == -2 and all(not x.value for x in pp)
    # supervoxel kNN: K beyond maxKNN; no references fills the distances only
    q = np.zeros(6, np.float32)
    ind = np.full(8, 5, np.int32)
    dist = np.zeros(8, np.float32)
    assert L.tgmm_knnCUDA_(ind.ctypes.data, dist.ctypes.data, q.ctypes.data, q.ctypes.data, 2,
                           2, 12, sc.ctypes.data, 0) == -2
    assert L.tgmm_knnCUDA_(ind.ctypes.data, dist.ctypes.data, q.ctypes.data, q.ctypes.data, 2,
                           0, 4, sc.ctypes.data, 0) == 0
    assert np.all(dist == np.float32(1e32)) and np.all(ind == 5)


def test_generator_is_deterministic():
    a = synth.config1()
    b = synth.config1()
    assert np.array_equal(a.query, b.query) and np.array_equal(a.img, b.img)
    assert a.gm.tobytes() == b.gm.tobytes()
    assert a.sv_ptr[0] == 0 and a.sv_ptr[-1] == a.N
    assert np.all(np.diff(a.sv_ptr) >= 1) and np.all(np.diff(a.sv_ptr) <= 64)
    assert 0.0 <= a.img.min() and a.img.max() == 1.0


def test_compat_header_compiles_and_links(tmp_path):
    """The tracker's drop-in links: tests/native/link_tracker.cpp calls every
    function the tracker takes from GMEMupdateCUDA.h, UtilsCUDA/knnCuda.h
    (selectCUDAdeviceMaxMemory, setDeviceCUDA, uploadScaleCUDA) and
    knnCUDA/knnCuda.h through include/tgmm_em_compat.hpp,
    include/tgmm_utils_knn_compat.hpp and include/tgmm_knn_compat.hpp, and is
    linked against libtgmm_em.so alone (an unresolved reference in the
    program fails the link). Run without arguments it calls nothing."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    lib = os.path.join(ROOT, "tgmm_amd")
    exe = str(tmp_path / "link_tracker")
    cmd = [gxx, "-std=c++11", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "native", "link_tracker.cpp"), "-L", lib, "-ltgmm_em",
           "-Wl,-rpath," + lib, "-Wl,--unresolved-symbols=ignore-in-shared-libs", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", "-C", "--undefined-only", exe], capture_output=True, text=True)
    for name in ("tgmm_selectCUDAdeviceMaxMemory", "tgmm_setDeviceCUDA", "tgmm_uploadScaleCUDA",
                 "tgmm_GMEMinitializeMemory", "tgmm_GMEMreleaseMemory",
                 "tgmm_GMEMvariationalInferenceCUDAWithSupervoxels", "tgmm_knnCUDA_"):
        assert name in nm.stdout, name
