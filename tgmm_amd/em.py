"""Python host mirror of the EM library's C-ABI (include/tgmm_em.h).

Two layers, as in the C header:

* the reference's host API, same names and argument meaning
  (nmeth.3036-S2/src/UtilsCUDA/GMEMupdateCUDA.h:82, 92-98, 132-144):
  ``GMEMinitializeMemoryWithSupervoxels``, ``copyScaleToDvice``,
  ``GMEMvariationalInferenceCUDAWithSupervoxels``,
  ``GMEMreleaseMemoryWithSupervoxels`` plus
  ``regularizePrecisionMatrixConstants.setConstants`` (constants.h:38-44).
  Errors raise ``TgmmError`` where the reference calls ``exit()``.
* ``Batch``: many frames in one device workspace, graph-replayed iterations.

The library is loaded from the package directory; importing this module on
a machine without the built library raises immediately (there is no CPU
fallback).
"""
import ctypes
import json
import os
import types

import numpy as np

from .records import GMM_DTYPE

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtgmm_em.so")

TGMM_OK = 0
ERR_NAMES = {-1: "TGMM_ERR_ARG", -2: "TGMM_ERR_LIMIT", -3: "TGMM_ERR_OOM",
             -4: "TGMM_ERR_DEVICE", -5: "TGMM_ERR_NOT_SET", -6: "TGMM_ERR_IO",
             -7: "TGMM_ERR_BUSY"}
SOLVE_EAGER, SOLVE_GRAPH, SOLVE_HOSTSYNC = 0, 1, 2
LEGACY_K = 5


class TgmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERR_NAMES.get(code, "?"), code, msg))
        self.code = code


class EmOpts(ctypes.Structure):
    _fields_ = [
        ("K", ctypes.c_int),
        ("max_iter", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("lambda_min", ctypes.c_double),
        ("lambda_max", ctypes.c_double),
        ("max_excentricity", ctypes.c_double),
        ("w4dof", ctypes.c_int),
        ("scale", ctypes.c_float * 3),
        ("max_sv_per_gaussian", ctypes.c_int),
        ("voxel_mode", ctypes.c_int),
    ]


def make_opts(K=5, max_iter=10, tol=-1.0, lambda_min=0.02, lambda_max=0.2222,
              max_excentricity=9.0, w4dof=True, scale=(1.0, 1.0, 5.0), cap=16, voxel=False):
    """Defaults: parseConfigFile.h:193-195 constants, W4DOF as the tracker
    (main.cpp:124), anisotropy (1, 1, 5). voxel=True selects the per-voxel
    EM of GMEMvariationalInferenceCUDA (GMEMupdateCUDA.cu:3420-3973)."""
    o = EmOpts()
    o.K, o.max_iter, o.tol = int(K), int(max_iter), float(tol)
    o.lambda_min, o.lambda_max = float(lambda_min), float(lambda_max)
    o.max_excentricity = float(max_excentricity)
    o.w4dof = 1 if w4dof else 0
    o.scale[:] = [float(x) for x in scale]
    o.max_sv_per_gaussian = int(cap)
    o.voxel_mode = 1 if voxel else 0
    return o


class FrameDesc(ctypes.Structure):
    _fields_ = [
        ("n_voxels", ctypes.c_longlong),
        ("n_supervoxels", ctypes.c_longlong),
        ("n_gaussians", ctypes.c_int),
        ("query", ctypes.c_void_p),
        ("img", ctypes.c_void_p),
        ("centroids", ctypes.c_void_p),
        ("sv_ptr", ctypes.c_void_p),
        ("gm", ctypes.c_void_p),
        ("rnk", ctypes.c_void_p),
        ("ind", ctypes.c_void_p),
        ("ll", ctypes.c_double),
        ("iterations", ctypes.c_int),
        ("status", ctypes.c_int),
    ]


class ServerOpts(ctypes.Structure):
    _fields_ = [
        ("em", EmOpts),
        ("hbm_budget_bytes", ctypes.c_longlong),
        ("max_frames_in_flight", ctypes.c_int),
        ("beta_percentage_of_nk", ctypes.c_double),
        ("nu_percentage_of_nk", ctypes.c_double),
        ("alpha_percentage", ctypes.c_double),
        ("borrow_inputs", ctypes.c_int),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("iterate_s", ctypes.c_double), ("retire_s", ctypes.c_double),
                ("steps", ctypes.c_longlong), ("frame_iterations", ctypes.c_longlong),
                ("in_flight", ctypes.c_longlong), ("queued", ctypes.c_longlong),
                ("iterate_gpu_s", ctypes.c_double)]


class BatchStats(ctypes.Structure):
    _fields_ = [("knn_ties_last_iteration", ctypes.c_int), ("layout_mismatches", ctypes.c_int),
                ("workspace_bytes", ctypes.c_longlong)]


class Tiles(ctypes.Structure):
    """Launch shapes of the tunable kernels (tgmm_tiles). Results are bitwise
    the same for every valid choice."""
    _fields_ = [("ll_group", ctypes.c_int), ("ll_unroll", ctypes.c_int),
                ("mstep_lanes", ctypes.c_int), ("sv_block", ctypes.c_int),
                ("knn_block", ctypes.c_int), ("knn_chunk", ctypes.c_int),
                ("knn_cell", ctypes.c_int), ("knn_lanes", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

    @classmethod
    def from_dict(cls, d):
        t = cls()
        for k, _ in cls._fields_:
            setattr(t, k, int(d[k]))
        return t


DEFAULT_TILES = dict(ll_group=16, ll_unroll=4, mstep_lanes=16, sv_block=256, knn_block=256,
                     knn_chunk=8, knn_cell=6, knn_lanes=2)



_lib = None
_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)


class RawFrame(ctypes.Structure):
    """tgmm_raw_frame: one frame as raw supervoxel voxels."""
    _fields_ = [
        ("N", ctypes.c_longlong),
        ("S", ctypes.c_longlong),
        ("dims", ctypes.c_longlong * 3),
        ("pixel_idx", ctypes.c_void_p),
        ("raw", ctypes.c_void_p),
        ("volume", ctypes.c_void_p),
        ("sv_ptr", ctypes.c_void_p),
        ("thr", ctypes.c_void_p),
        ("device_inputs", ctypes.c_int),
    ]


def raw_frame(pixel_idx, raw, sv_ptr, thr, dims, volume=None):
    """(RawFrame, keep-alive list) from numpy arrays (host) or DeviceArrays
    (device) for pixel_idx / raw / thr, all of one kind; volume: DeviceArray
    of the whole uint16 image, read when raw is None."""
    keep = []
    kinds = {isinstance(a, DeviceArray) for a in (pixel_idx, raw, thr) if a is not None}
    if len(kinds) > 1:
        raise TypeError("pixel_idx, raw and thr must all be host arrays or all DeviceArrays")

    def host_or_dev(a, dtype):
        if a is None:
            return None
        if isinstance(a, DeviceArray):
            keep.append(a)
            return a.ptr.value
        h = np.ascontiguousarray(a, dtype)
        keep.append(h)
        return h.ctypes.data
    ptr = np.ascontiguousarray(sv_ptr, np.int64)
    keep.append(ptr)
    f = RawFrame()
    f.N = int(ptr[-1])
    f.S = ptr.shape[0] - 1
    f.dims[:] = [int(d) for d in dims]
    f.pixel_idx = host_or_dev(pixel_idx, np.uint64)
    f.raw = host_or_dev(raw, np.uint16)
    f.volume = host_or_dev(volume, np.uint16)
    f.sv_ptr = ptr.ctypes.data
    f.thr = host_or_dev(thr, np.uint16)
    f.device_inputs = 1 if kinds == {True} else 0
    return f, keep


def marshal_frame(pixel_idx, raw, sv_ptr, thr, dims, volume=None):
    """EM inputs built on the device from a raw frame (tgmm_marshal_frame).
    Returns DeviceArrays (query [3N], img [N], centroids [3S], sv_ptr [S+1])."""
    f, keep = raw_frame(pixel_idx, raw, sv_ptr, thr, dims, volume)
    q = DeviceArray(nbytes=3 * f.N * 4)
    im = DeviceArray(nbytes=f.N * 4)
    c = DeviceArray(nbytes=3 * f.S * 4)
    p = DeviceArray(nbytes=(f.S + 1) * 8)
    _check(load().tgmm_marshal_frame(ctypes.byref(f), q.ptr, im.ptr, c.ptr, p.ptr))
    return q, im, c, p


def GMEMinitializeMemoryFromRaw(frame_arrays, scale):
    """tgmm_GMEMinitializeMemoryFromRaw: the seven device arrays of
    GMEMinitializeMemoryWithSupervoxels, filled from a raw frame given as
    (pixel_idx, raw, sv_ptr, thr, dims[, volume]). Returns a DevicePointers."""
    f, keep = raw_frame(*frame_arrays)
    sc = np.ascontiguousarray(scale, np.float32)
    d = DevicePointers()
    _check(load().tgmm_GMEMinitializeMemoryFromRaw(
        ctypes.byref(d.queryCUDA), ctypes.byref(d.imgDataCUDA), sc.ctypes.data,
        ctypes.byref(d.rnkCUDA), ctypes.byref(d.indCUDA), ctypes.byref(d.indCUDAtr),
        ctypes.byref(d.centroidLabelPositionCUDA), ctypes.byref(d.labelListPtrCUDA),
        ctypes.byref(f)))
    return d


def load():
    """Load libtgmm_em.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libtgmm_em.so is not built at %s; run "
                          "python -m tgmm_amd.build" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.tgmm_last_error.restype = ctypes.c_char_p
    L.tgmm_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    L.tgmm_GMEMinitializeMemoryWithSupervoxels.argtypes = [
        _pp, _vp, _pp, _vp, _vp, ctypes.c_longlong, _pp, _pp, _pp, ctypes.c_ushort,
        _pp, _vp, _pp, _vp]
    L.tgmm_copyScaleToDvice.argtypes = [_vp]
    L.tgmm_setRegularizePrecisionMatrixConstants.argtypes = [ctypes.c_double] * 3
    L.tgmm_GMEMvariationalInferenceCUDAWithSupervoxels.argtypes = [
        _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_longlong, ctypes.c_int,
        ctypes.c_ushort, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_char_p]
    L.tgmm_GMEMreleaseMemoryWithSupervoxels.argtypes = [_pp] * 7
    L.tgmm_GMEMvariationalInferenceCUDA.argtypes = [
        _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
        ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]
    L.tgmm_calculateLocalLikelihood.argtypes = [
        _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_longlong,
        ctypes.c_int, ctypes.c_int]
    L.tgmm_marshal_frame.argtypes = [ctypes.POINTER(RawFrame), _vp, _vp, _vp, _vp]
    L.tgmm_GMEMinitializeMemoryFromRaw.argtypes = [_pp, _pp, _vp, _pp, _pp, _pp, _pp, _pp,
                                                   ctypes.POINTER(RawFrame)]
    L.tgmm_knnCUDA_.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_longlong, ctypes.c_int,
                                ctypes.c_int, _vp, ctypes.c_int]
    L.tgmm_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(FrameDesc), ctypes.c_int,
                                    ctypes.POINTER(EmOpts), _pp]
    L.tgmm_batch_reset.argtypes = [_vp]
    L.tgmm_batch_solve.argtypes = [_vp, ctypes.c_int]
    L.tgmm_batch_fetch.argtypes = [_vp, ctypes.POINTER(FrameDesc), ctypes.c_int]
    L.tgmm_batch_time_kernel.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]
    L.tgmm_batch_destroy.argtypes = [_vp]
    L.tgmm_batch_stats.argtypes = [_vp, ctypes.POINTER(BatchStats)]
    L.tgmm_GMEMinitializeMemory.argtypes = [_pp, _vp, _pp, _vp, _vp, ctypes.c_longlong, _pp,
                                            _pp, _pp]
    L.tgmm_GMEMreleaseMemory.argtypes = [_pp] * 5
    L.tgmm_selectCUDAdeviceMaxMemory.argtypes = [ctypes.POINTER(ctypes.c_int)]
    L.tgmm_setDeviceCUDA.argtypes = [ctypes.c_int]
    L.tgmm_uploadScaleCUDA.argtypes = [_vp]
    L.tgmm_batch_set_tiles.argtypes = [_vp, ctypes.POINTER(Tiles)]
    L.tgmm_batch_get_tiles.argtypes = [_vp, ctypes.POINTER(Tiles)]
    L.tgmm_batch_autotune.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(Tiles),
                                      ctypes.c_char_p, ctypes.c_int]
    L.tgmm_server_set_tiles.argtypes = [_vp, ctypes.POINTER(Tiles)]
    L.tgmm_server_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
    L.tgmm_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(ServerOpts), _pp]
    L.tgmm_server_submit.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(FrameDesc),
                                     ctypes.POINTER(ctypes.c_longlong)]
    L.tgmm_server_wait.argtypes = [_vp, ctypes.c_longlong]
    L.tgmm_server_step.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
    L.tgmm_server_fetch.argtypes = [_vp, ctypes.c_longlong, ctypes.POINTER(FrameDesc),
                                    ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]
    L.tgmm_server_query.argtypes = [_vp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_int)]
    L.tgmm_server_destroy.argtypes = [_vp]
    _lib = L
    return L


_hip = None


def pinned(a):
    """A page-locked (hipHostMalloc) copy of numpy array `a`, or `a` itself if
    page-locked memory cannot be had. The memory is freed with the array."""
    import weakref
    a = np.ascontiguousarray(a)
    try:
        _hip_lib()
        p = ctypes.c_void_p()
        if _hip.hipHostMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 1)), 0) != 0:
            return a
    except OSError:
        return a
    buf = (ctypes.c_char * max(a.nbytes, 1)).from_address(p.value)
    out = np.frombuffer(buf, dtype=a.dtype, count=a.size).reshape(a.shape)
    out[...] = a
    # freed when the buffer object goes, i.e. after the last array viewing it
    weakref.finalize(buf, _hip.hipHostFree, ctypes.c_void_p(p.value))
    return out


def _tiles_cache_path():
    """Autotuner cache: $TGMM_TILES_CACHE, else ~/.cache/tgmm_amd/tiles.json."""
    return os.environ.get("TGMM_TILES_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "tgmm_amd", "tiles.json")


def _check(rc):
    if rc != TGMM_OK:
        raise TgmmError(rc, load().tgmm_last_error().decode())
    return rc


def device_count():
    n = ctypes.c_int(0)
    _check(load().tgmm_device_count(ctypes.byref(n)))
    return n.value


def as_records(a):
    """C-contiguous array with exactly the 288-byte record layout (numpy
    operations such as concatenate may repack a padded structured dtype)."""
    a = np.asarray(a)
    if a.dtype != GMM_DTYPE:
        a = a.astype(GMM_DTYPE)
    return np.ascontiguousarray(a)


def _ptr(a):
    return a.ctypes.data if a is not None else None


# ---------------------------------------------------------------- drop-in --
class DevicePointers:
    """The seven device arrays GMEMinitializeMemoryWithSupervoxels returns."""

    def __init__(self):
        self.queryCUDA = ctypes.c_void_p()
        self.imgDataCUDA = ctypes.c_void_p()
        self.rnkCUDA = ctypes.c_void_p()
        self.indCUDA = ctypes.c_void_p()
        self.indCUDAtr = ctypes.c_void_p()
        self.centroidLabelPositionCUDA = ctypes.c_void_p()
        self.labelListPtrCUDA = ctypes.c_void_p()


def GMEMinitializeMemoryWithSupervoxels(query, imgData, scale, query_nb, numLabels,
                                        centroidLabelPosition, labelListPtrHOST):
    """Allocate and upload one frame (GMEMupdateCUDA.h:132-137). Host arrays:
    query f32 SoA [3N], imgData f32 [N], scale f32 [3], centroids f32 SoA [3S],
    labelListPtr i64 [S+1]. Returns a DevicePointers."""
    L = load()
    q = np.ascontiguousarray(query, np.float32)
    im = np.ascontiguousarray(imgData, np.float32)
    sc = np.ascontiguousarray(scale, np.float32)
    ce = np.ascontiguousarray(centroidLabelPosition, np.float32)
    pt = np.ascontiguousarray(labelListPtrHOST, np.int64)
    d = DevicePointers()
    _check(L.tgmm_GMEMinitializeMemoryWithSupervoxels(
        ctypes.byref(d.queryCUDA), _ptr(q), ctypes.byref(d.imgDataCUDA), _ptr(im), _ptr(sc),
        int(query_nb), ctypes.byref(d.rnkCUDA), ctypes.byref(d.indCUDA),
        ctypes.byref(d.indCUDAtr), int(numLabels), ctypes.byref(d.centroidLabelPositionCUDA),
        _ptr(ce), ctypes.byref(d.labelListPtrCUDA), _ptr(pt)))
    return d


def copyScaleToDvice(scale):
    """Anisotropy for the regularisation (GMEMupdateCUDA.h:82)."""
    sc = np.ascontiguousarray(scale, np.float32)
    _check(load().tgmm_copyScaleToDvice(_ptr(sc)))


class regularizePrecisionMatrixConstants:
    """Mirror of the reference's host statics (constants.h:19-45)."""

    @staticmethod
    def setConstants(lambdaMin, lambdaMax, maxExcentricity):
        _check(load().tgmm_setRegularizePrecisionMatrixConstants(
            float(lambdaMin), float(lambdaMax), float(maxExcentricity)))


def GMEMvariationalInferenceCUDAWithSupervoxels(d, vecGMtemp, query_nb, ref_nb, numLabels,
                                                maxIterEM, tolLikelihood, devCUDA=0,
                                                frame=0, W4DOF=True, debugPath=""):
    """One frame's EM (GMEMupdateCUDA.h:92-98). vecGMtemp: GMM_DTYPE array,
    updated in place. Device rnk/ind in `d` hold the last iteration."""
    if vecGMtemp.dtype != GMM_DTYPE or not vecGMtemp.flags["C_CONTIGUOUS"]:
        raise TypeError("vecGMtemp must be a C-contiguous GMM_DTYPE array (288-byte records)")
    _check(load().tgmm_GMEMvariationalInferenceCUDAWithSupervoxels(
        d.queryCUDA, d.imgDataCUDA, d.rnkCUDA, None, d.indCUDA, d.indCUDAtr,
        d.centroidLabelPositionCUDA, d.labelListPtrCUDA, vecGMtemp.ctypes.data,
        int(query_nb), int(ref_nb), int(numLabels), int(maxIterEM), float(tolLikelihood),
        int(devCUDA), int(frame), 1 if W4DOF else 0,
        debugPath.encode() if debugPath else None))


def GMEMvariationalInferenceCUDA(queryCUDA, imgDataCUDA, rnkCUDA, indCUDA, indCUDAtr,
                                 vecGMtemp, query_nb, ref_nb, maxIterEM, tolLikelihood,
                                 devCUDA=0, frame=0, W4DOF=True, debugPath=""):
    """One frame's per-voxel EM (GMEMupdateCUDA.h:87-91). Device arrays (see
    DeviceArray): query f32 SoA [3N], img f32 [N], rnk f32 [5N], ind i32
    [5N], indTr i32 [5N] or None. vecGMtemp: GMM_DTYPE array, updated in
    place (supervoxelIdx/Num unchanged). rnk holds r * img of the last
    iteration, slot-major."""
    if vecGMtemp.dtype != GMM_DTYPE or not vecGMtemp.flags["C_CONTIGUOUS"]:
        raise TypeError("vecGMtemp must be a C-contiguous GMM_DTYPE array (288-byte records)")
    dp = lambda a: None if a is None else (a.ptr if isinstance(a, DeviceArray) else a)
    _check(load().tgmm_GMEMvariationalInferenceCUDA(
        dp(queryCUDA), dp(imgDataCUDA), dp(rnkCUDA), None, dp(indCUDA), dp(indCUDAtr),
        vecGMtemp.ctypes.data, int(query_nb), int(ref_nb), int(maxIterEM),
        float(tolLikelihood), int(devCUDA), int(frame), 1 if W4DOF else 0,
        debugPath.encode() if debugPath else None))


def calculateLocalLikelihood(localLikelihood, listSupervoxels, queryCUDA, imgDataCUDA,
                             rnkCUDA, indCUDA, vecGMtemp, labelListPtrCUDA, query_nb, ref_nb,
                             numLabels):
    """Local log-likelihood of each list of supervoxel indices
    (GMEMupdateCUDA.h:116-127). localLikelihood: a Python list, resized to
    len(listSupervoxels) (new entries 0.0) and written for every non-empty
    list, as the reference does with its vector; also returned. Device
    pointers (or DeviceArray) as after GMEMvariationalInferenceCUDAWithSupervoxels;
    vecGMtemp: GMM_DTYPE host records [ref_nb]."""
    if vecGMtemp.dtype != GMM_DTYPE or not vecGMtemp.flags["C_CONTIGUOUS"]:
        raise TypeError("vecGMtemp must be a C-contiguous GMM_DTYPE array (288-byte records)")
    n = len(listSupervoxels)
    del localLikelihood[n:]
    localLikelihood.extend([0.0] * (n - len(localLikelihood)))
    ptr = np.zeros(n + 1, np.int64)
    ptr[1:] = np.cumsum([len(x) for x in listSupervoxels])
    idx = (np.concatenate([np.asarray(x, np.int32) for x in listSupervoxels])
           if ptr[-1] else np.zeros(1, np.int32)).astype(np.int32)
    out = np.array(localLikelihood + [0.0], np.float64)
    dp = lambda a: None if a is None else (a.ptr if isinstance(a, DeviceArray) else a)
    _check(load().tgmm_calculateLocalLikelihood(
        out.ctypes.data, ptr.ctypes.data, idx.ctypes.data, n, dp(queryCUDA), dp(imgDataCUDA),
        dp(rnkCUDA), dp(indCUDA), vecGMtemp.ctypes.data, dp(labelListPtrCUDA), int(query_nb),
        int(ref_nb), int(numLabels)))
    localLikelihood[:] = [float(x) for x in out[:n]]
    return localLikelihood


def knnCUDA_(query, ref, KNN, scale, devCUDA=0, with_dist=True):
    """Supervoxel nearest neighbours (knnCuda.h:56-58): query [3Q], ref [3R]
    SoA centroids on the host. Returns (ind, dist) [KNN*Q] slot-major, dist
    None when with_dist is False. With no references ind is -1 throughout
    (the C entry point leaves it as given) and dist 1e32."""
    q = np.ascontiguousarray(query, np.float32)
    r = np.ascontiguousarray(ref, np.float32)
    sc = np.ascontiguousarray(scale, np.float32)
    Q, R = q.size // 3, r.size // 3
    ind = np.full(max(KNN * Q, 1), -1, np.int32)
    dist = np.zeros(max(KNN * Q, 1), np.float32) if with_dist else None
    _check(load().tgmm_knnCUDA_(ind.ctypes.data, dist.ctypes.data if with_dist else None,
                                q.ctypes.data, r.ctypes.data, Q, R, int(KNN), sc.ctypes.data,
                                int(devCUDA)))
    return ind[: KNN * Q], (dist[: KNN * Q] if with_dist else None)


def GMEMreleaseMemoryWithSupervoxels(d):
    """Free the seven arrays and null them (GMEMupdateCUDA.h:140-144)."""
    _check(load().tgmm_GMEMreleaseMemoryWithSupervoxels(
        ctypes.byref(d.queryCUDA), ctypes.byref(d.imgDataCUDA), ctypes.byref(d.rnkCUDA),
        ctypes.byref(d.indCUDA), ctypes.byref(d.indCUDAtr),
        ctypes.byref(d.centroidLabelPositionCUDA), ctypes.byref(d.labelListPtrCUDA)))


def _hip_lib():
    """libamdhip64 through ctypes, for host-side buffer management only."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
    _hip.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t,
                                   ctypes.c_uint]
    _hip.hipHostFree.argtypes = [ctypes.c_void_p]
    _hip.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
    _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    _hip.hipFree.argtypes = [_vp]
    return _hip


class DeviceArray:
    """A device buffer (hipMalloc) holding a copy of a numpy array, or
    `nbytes` of zeros; freed with the object. For callers of the drop-in
    entry points that take device pointers."""

    def __init__(self, host=None, nbytes=None):
        H = _hip_lib()
        self.ptr = ctypes.c_void_p()
        n = host.nbytes if host is not None else int(nbytes)
        if H.hipMalloc(ctypes.byref(self.ptr), max(n, 1)) != 0:
            raise TgmmError(-3, "hipMalloc of %d bytes failed" % n)
        self.nbytes = n
        if host is not None:
            h = np.ascontiguousarray(host)
            if H.hipMemcpy(self.ptr, h.ctypes.data, n, 1) != 0:  # HostToDevice
                raise TgmmError(-4, "hipMemcpy H2D failed")
        else:
            z = np.zeros(max(n, 1), np.uint8)
            if H.hipMemcpy(self.ptr, z.ctypes.data, max(n, 1), 1) != 0:
                raise TgmmError(-4, "hipMemcpy H2D failed")

    def numpy(self, dtype):
        return copy_device_to_host(self.ptr, self.nbytes, dtype)

    def __del__(self):
        try:
            if self.ptr:
                _hip_lib().hipFree(self.ptr)
                self.ptr = ctypes.c_void_p()
        except Exception:
            pass


def copy_device_to_host(dev_ptr, nbytes, dtype):
    """hipMemcpy of a device buffer (e.g. rnkCUDA) into a new numpy array."""
    _hip_lib()
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
    rc = _hip.hipMemcpy(out.ctypes.data, dev_ptr, nbytes, 2)  # DeviceToHost
    if rc != 0:
        raise TgmmError(-4, "hipMemcpy D2H failed (%d)" % rc)
    return out


# ------------------------------------------------------------------- batch --
class Batch:
    """Frames of uneven size in one device workspace.

    frames: objects with ``query, img, centroids, sv_ptr, gm`` numpy arrays
    (the boundary layout; see tgmm_amd.synth.Frame)."""

    def __init__(self, frames, opts, device=0):
        self.L = load()
        self.opts = opts
        self.frames = list(frames)
        self._keep = []
        n = len(self.frames)
        self.desc = (FrameDesc * n)()
        self.voxel = bool(opts.voxel_mode)
        for i, f in enumerate(self.frames):
            img = np.ascontiguousarray(f.img, np.float32)
            if self.voxel:  # one-voxel supervoxels: the supervoxel arrays are unused
                ce = np.zeros(3, np.float32)
                pt = np.arange(2, dtype=np.int64)
            else:
                ce = np.ascontiguousarray(f.centroids, np.float32)
                pt = np.ascontiguousarray(f.sv_ptr, np.int64)
            arrs = [np.ascontiguousarray(f.query, np.float32), img, ce, pt, as_records(f.gm)]
            self._keep.append(arrs)
            d = self.desc[i]
            d.n_voxels = arrs[1].shape[0]
            d.n_supervoxels = arrs[1].shape[0] if self.voxel else arrs[3].shape[0] - 1
            d.n_gaussians = arrs[4].shape[0]
            d.query, d.img, d.centroids, d.sv_ptr, d.gm = [a.ctypes.data for a in arrs]
        self.h = ctypes.c_void_p()
        _check(self.L.tgmm_batch_create(device, self.desc, n, ctypes.byref(opts),
                                        ctypes.byref(self.h)))
        self._keep = None  # host copies are no longer needed after create

    @classmethod
    def from_raw(cls, raw_frames, gms, opts, device=0):
        """A supervoxel-mode batch whose frames arrive as raw uint16 supervoxel
        voxels: each frame is built on the device (tgmm_marshal_frame) and
        the workspace is filled from those device arrays (tgmm_batch_create
        takes device pointers for query, img and centroids). raw_frames:
        (pixel_idx, raw, sv_ptr, thr, dims[, volume]) per frame; gms: the
        frames' Gaussian records."""
        if opts.voxel_mode:
            raise TgmmError(-1, "Batch.from_raw builds supervoxel-mode batches")
        self = cls.__new__(cls)
        self.L = load()
        self.opts = opts
        self.voxel = False
        self.h = ctypes.c_void_p()
        n = len(raw_frames)
        self.desc = (FrameDesc * n)()
        keep, frames = [], []
        for i, (rf, gm) in enumerate(zip(raw_frames, gms)):
            q, im, c, _ = marshal_frame(*rf)
            ptr = np.ascontiguousarray(rf[2], np.int64)
            g = as_records(gm)
            keep += [q, im, c, ptr, g]
            N, S = int(ptr[-1]), ptr.shape[0] - 1
            d = self.desc[i]
            d.n_voxels, d.n_supervoxels, d.n_gaussians = N, S, g.shape[0]
            d.query, d.img, d.centroids = q.ptr.value, im.ptr.value, c.ptr.value
            d.sv_ptr, d.gm = ptr.ctypes.data, g.ctypes.data
            frames.append(types.SimpleNamespace(sv_ptr=ptr, gm=g,
                                                img=np.broadcast_to(np.float32(0), (N,))))
        self.frames = frames
        _check(self.L.tgmm_batch_create(device, self.desc, n, ctypes.byref(opts),
                                        ctypes.byref(self.h)))
        self._keep = None  # the device inputs were copied into the workspace
        return self

    def reset(self):
        _check(self.L.tgmm_batch_reset(self.h))

    def solve(self, mode=SOLVE_GRAPH):
        _check(self.L.tgmm_batch_solve(self.h, mode))

    def fetch(self, records_only=False, into=None):
        """List of dicts: gm (records), rnk/ind (slot-major [K, S]; [K, N] in
        voxel mode; None with records_only), ll, iterations. `into`: a list of
        record arrays (one per frame) to write the records into."""
        K = self.opts.K
        outs = []
        for i, f in enumerate(self.frames):
            S = (int(np.asarray(f.img).shape[0]) if self.voxel
                 else int(np.asarray(f.sv_ptr).shape[0] - 1))
            gm = (into[i] if into is not None
                  else np.zeros(int(np.asarray(f.gm).shape[0]), GMM_DTYPE))
            d = self.desc[i]
            d.gm = gm.ctypes.data
            if records_only:
                rnk = ind = None
                d.rnk = d.ind = None
            else:
                rnk = np.zeros(K * S, np.float32)
                ind = np.zeros(K * S, np.int32)
                d.rnk, d.ind = rnk.ctypes.data, ind.ctypes.data
            outs.append(dict(gm=gm, rnk=rnk, ind=ind))
        _check(self.L.tgmm_batch_fetch(self.h, self.desc, len(self.frames)))
        for i, o in enumerate(outs):
            o["ll"] = self.desc[i].ll
            o["iterations"] = self.desc[i].iterations
        return outs

    def stats(self):
        """kNN ties in the last iteration, final-layout mismatches (0) and
        workspace bytes (tgmm_batch_stats)."""
        st = BatchStats()
        _check(self.L.tgmm_batch_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in BatchStats._fields_}

    def set_tiles(self, tiles):
        t = tiles if isinstance(tiles, Tiles) else Tiles.from_dict(tiles)
        _check(self.L.tgmm_batch_set_tiles(self.h, ctypes.byref(t)))

    def get_tiles(self):
        t = Tiles()
        _check(self.L.tgmm_batch_get_tiles(self.h, ctypes.byref(t)))
        return t.as_dict()

    def signature(self):
        """Workload key for the tile cache: the mode, K and log2 buckets of
        the total supervoxels, Gaussians and voxels per supervoxel."""
        S = sum(int(np.asarray(f.sv_ptr).shape[0] - 1) for f in self.frames)
        G = sum(int(np.asarray(f.gm).shape[0]) for f in self.frames)
        N = sum(int(np.asarray(f.img).shape[0]) for f in self.frames)
        b = lambda x: int(round(np.log2(max(x, 1))))
        return "%sK%d_S%d_G%d_v%d_f%d" % ("vox_" if self.opts.voxel_mode else "", self.opts.K,
                                          b(S), b(G), b(N / max(S, 1)), b(len(self.frames)))

    def autotune(self, reps=3, cache=True):
        """Pick the tile shapes by timing the candidates on this batch
        (tgmm_batch_autotune), or reuse the choice cached for the same
        workload signature. Leaves the batch in its initial state. Returns
        (tiles dict, report text or None when cached)."""
        key = self.signature()
        path = _tiles_cache_path()
        table = {}
        if cache and os.path.exists(path):
            try:
                with open(path) as fh:
                    table = json.load(fh)
            except (OSError, ValueError):
                table = {}
        if cache and key in table:
            self.set_tiles(table[key])
            self.reset()
            return table[key], None
        t = Tiles()
        buf = ctypes.create_string_buffer(4096)
        _check(self.L.tgmm_batch_autotune(self.h, int(reps), ctypes.byref(t), buf, 4096))
        d = t.as_dict()
        if cache:
            table[key] = d
            try:
                os.makedirs(os.path.dirname(path), exist_ok=True)
                tmp = path + ".%d" % os.getpid()
                with open(tmp, "w") as fh:
                    json.dump(table, fh, indent=1, sort_keys=True)
                os.replace(tmp, path)
            except OSError:
                pass
        return d, buf.value.decode()

    def time_kernel(self, name, reps):
        ms = ctypes.c_double()
        by = ctypes.c_double()
        _check(self.L.tgmm_batch_time_kernel(self.h, name.encode(), int(reps),
                                             ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    def close(self):
        if self.h:
            self.L.tgmm_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ server --
class Server:
    """Continuous-batching server for concurrent time-lapse streams
    (tgmm_server_* in include/tgmm_em.h)."""

    def __init__(self, opts, hbm_budget_bytes=8 << 30, max_frames_in_flight=16,
                 beta_pct=0.5, nu_pct=0.5, alpha_pct=0.8, device=0, borrow=False):
        """borrow=True: the server reads the frames' arrays in place (they are
        kept referenced here until fetched); page-locked arrays (pinned())
        then upload without blocking the host."""
        self.L = load()
        self.opts = opts
        self.borrow = bool(borrow)
        so = ServerOpts()
        so.em = opts
        so.hbm_budget_bytes = int(hbm_budget_bytes)
        so.max_frames_in_flight = int(max_frames_in_flight)
        so.beta_percentage_of_nk = beta_pct
        so.nu_percentage_of_nk = nu_pct
        so.alpha_percentage = alpha_pct
        so.borrow_inputs = 1 if self.borrow else 0
        self.h = ctypes.c_void_p()
        _check(self.L.tgmm_server_create(device, ctypes.byref(so), ctypes.byref(self.h)))
        self.shapes = {}
        self._held = {}

    def submit(self, frame, stream=-1, with_gaussians=True):
        """Queue a frame (host arrays are copied); returns a ticket. With
        with_gaussians=False the priors come from the stream's previous
        posterior on the device."""
        d = FrameDesc()
        q = np.ascontiguousarray(frame.query, np.float32)
        im = np.ascontiguousarray(frame.img, np.float32)
        ce = np.ascontiguousarray(frame.centroids, np.float32)
        pt = np.ascontiguousarray(frame.sv_ptr, np.int64)
        d.n_voxels = im.shape[0]
        d.n_supervoxels = pt.shape[0] - 1
        d.query, d.img, d.centroids, d.sv_ptr = q.ctypes.data, im.ctypes.data, ce.ctypes.data, \
            pt.ctypes.data
        gm = None
        if with_gaussians:
            gm = as_records(frame.gm)
            d.n_gaussians = gm.shape[0]
            d.gm = gm.ctypes.data
        t = ctypes.c_longlong()
        _check(self.L.tgmm_server_submit(self.h, int(stream), ctypes.byref(d), ctypes.byref(t)))
        self.shapes[t.value] = int(d.n_supervoxels)
        if self.borrow:
            self._held[t.value] = (q, im, ce, pt)
        return t.value

    def set_tiles(self, tiles):
        t = tiles if isinstance(tiles, Tiles) else Tiles.from_dict(tiles)
        _check(self.L.tgmm_server_set_tiles(self.h, ctypes.byref(t)))

    def stats(self):
        """Cumulative host seconds per phase and step counts (tgmm_server_stats)."""
        st = ServerStats()
        _check(self.L.tgmm_server_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def step(self):
        p = ctypes.c_int()
        _check(self.L.tgmm_server_step(self.h, ctypes.byref(p)))
        return p.value

    def wait(self, ticket=-1):
        _check(self.L.tgmm_server_wait(self.h, int(ticket)))

    def query(self, ticket):
        """0 queued, 1 in flight, 2 done (ready to fetch)."""
        st = ctypes.c_int()
        _check(self.L.tgmm_server_query(self.h, int(ticket), ctypes.byref(st)))
        return st.value

    def fetch(self, ticket, n_gaussians):
        S = self.shapes[ticket]
        K = self.opts.K
        gm = np.zeros(int(n_gaussians), GMM_DTYPE)
        rnk = np.zeros(K * S, np.float32)
        ind = np.zeros(K * S, np.int32)
        d = FrameDesc()
        d.gm, d.rnk, d.ind = gm.ctypes.data, rnk.ctypes.data, ind.ctypes.data
        ts, ta, td = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(self.L.tgmm_server_fetch(self.h, int(ticket), ctypes.byref(d), ctypes.byref(ts),
                                        ctypes.byref(ta), ctypes.byref(td)))
        del self.shapes[ticket]
        self._held.pop(ticket, None)
        return dict(gm=gm, rnk=rnk, ind=ind, ll=d.ll, iterations=d.iterations,
                    t_submit=ts.value, t_admit=ta.value, t_done=td.value)

    def close(self):
        if self.h:
            self.L.tgmm_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
