"""ctypes bindings to the oracle (oracle/liboracle.so) and the reference's
CPU twin (oracle/_ref/libtgmm_reftwin.so). Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REFTWIN_SO = os.path.join(ROOT, "oracle", "_ref", "libtgmm_reftwin.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_longlong)
_vp = ctypes.c_void_p
_ll = ctypes.c_longlong


class EmOpts(ctypes.Structure):
    """tgmm_em_opts from include/tgmm_em.h"""
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


def make_opts(K=5, max_iter=10, tol=-1.0, lmin=0.02, lmax=0.2222, emax=9.0,
              w4dof=1, scale=(1.0, 1.0, 5.0), cap=16, voxel=False):
    o = EmOpts()
    o.K, o.max_iter, o.tol = K, max_iter, tol
    o.lambda_min, o.lambda_max, o.max_excentricity = lmin, lmax, emax
    o.w4dof = w4dof
    o.scale[:] = list(scale)
    o.max_sv_per_gaussian = cap
    o.voxel_mode = 1 if voxel else 0
    return o


def p(a, t):
    return a.ctypes.data_as(t)


_orc = None


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so not built (make -C oracle)")
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_digamma.restype = ctypes.c_double
        L.orc_digamma.argtypes = [ctypes.c_double]
        L.orc_det3.restype = ctypes.c_double
        L.orc_det3.argtypes = [_dp]
        L.orc_inv3.argtypes = [_dp, _dp]
        L.orc_eig3_ordered.argtypes = [_dp, _dp, _dp]
        L.orc_regularize_W.argtypes = [_dp, ctypes.c_double, ctypes.c_int, _fp,
                                       ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double]
        L.orc_posterior.argtypes = [_vp, ctypes.c_double, _dp, _dp]
        L.orc_knn.argtypes = [_ll, _fp, ctypes.c_int, _vp, _fp, ctypes.c_int, _ip]
        L.orc_sum_alpha.restype = ctypes.c_double
        L.orc_sum_alpha.argtypes = [ctypes.c_int, _vp]
        L.orc_expectations.argtypes = [ctypes.c_int, _vp, ctypes.c_double]
        L.orc_estep.argtypes = [_ll, _ll, _fp, _lp, ctypes.c_int, _vp,
                                ctypes.c_int, _ip, _fp]
        L.orc_estep_voxel.argtypes = [_ll, _fp, _fp, ctypes.c_int, _vp, ctypes.c_int, _ip,
                                      _fp]
        L.orc_mstep.argtypes = [_ll, _ll, _fp, _fp, _lp, ctypes.c_int,
                                ctypes.c_int, _ip, _fp, _dp, _dp, _dp]
        L.orc_update.argtypes = [ctypes.c_int, _vp, _dp, _dp, _dp,
                                 ctypes.POINTER(EmOpts)]
        L.orc_death.argtypes = [ctypes.c_int, _vp, ctypes.c_double]
        L.orc_loglik.restype = ctypes.c_double
        L.orc_loglik.argtypes = [_ll, _ll, _fp, _fp, _lp, ctypes.c_int, _vp,
                                 ctypes.c_int, _ip, ctypes.c_double]
        L.orc_local_loglik.argtypes = [_ll, _ll, _fp, _fp, _lp, ctypes.c_int, _vp,
                                       ctypes.c_int, _ip, ctypes.c_int, _lp, _ip, _dp]
        L.orc_marshal_frame.argtypes = [_ll, _ll, _vp, _vp, _lp, _vp, _lp, _fp, _fp, _fp]
        L.orc_sv_knn.argtypes = [_ip, _fp, _fp, _fp, _ll, ctypes.c_int, ctypes.c_int, _fp]
        L.orc_converged.argtypes = [ctypes.c_double] * 5
        L.orc_assign_supervoxels.argtypes = [_ll, ctypes.c_int, _vp,
                                             ctypes.c_int, _fp, _ip, ctypes.c_int]
        L.orc_em_run.argtypes = [_ll, _ll, ctypes.c_int, _fp, _fp, _fp, _lp,
                                 _vp, ctypes.POINTER(EmOpts), _fp, _ip, _dp,
                                 _ip, _dp]
        L.orc_em_partial.argtypes = [_ll, _ll, ctypes.c_int, _fp, _fp, _fp,
                                     _lp, _vp, ctypes.POINTER(EmOpts),
                                     ctypes.c_int, _fp, _ip, _dp, _dp, _dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _orc = L
    return _orc


def em_run(frame, opts, gm=None):
    """Full oracle EM on a synth.Frame; returns (gm, rnk, ind, ll_hist, iters, ll)."""
    L = oracle()
    gm = (frame.gm if gm is None else gm).copy()
    S, K = (frame.N if opts.voxel_mode else frame.S), opts.K
    rnk = np.zeros(S * K, np.float32)
    ind = np.zeros(S * K, np.int32)
    hist = np.zeros(max(opts.max_iter, 1), np.float64)
    iters = ctypes.c_int(0)
    ll = ctypes.c_double(0)
    rc = L.orc_em_run(frame.N, S, frame.G, p(frame.query, _fp), p(frame.img, _fp),
                      p(frame.centroids, _fp), p(frame.sv_ptr, _lp), gm.ctypes.data,
                      ctypes.byref(opts), p(rnk, _fp), p(ind, _ip), p(hist, _dp),
                      ctypes.byref(iters), ctypes.byref(ll))
    assert rc == 0, rc
    return gm, rnk, ind, hist[: iters.value], iters.value, ll.value


def em_partial(frame, opts, iters, gm=None):
    """Oracle EM stopped after the M-step statistics of iteration `iters`."""
    L = oracle()
    gm = (frame.gm if gm is None else gm).copy()
    S, K, G = (frame.N if opts.voxel_mode else frame.S), opts.K, frame.G
    rnk = np.zeros(S * K, np.float32)
    ind = np.zeros(S * K, np.int32)
    Nk = np.zeros(G)
    xk = np.zeros(3 * G)
    Ck = np.zeros(6 * G)
    rc = L.orc_em_partial(frame.N, S, G, p(frame.query, _fp), p(frame.img, _fp),
                          p(frame.centroids, _fp), p(frame.sv_ptr, _lp),
                          gm.ctypes.data, ctypes.byref(opts), iters, p(rnk, _fp),
                          p(ind, _ip), p(Nk, _dp), p(xk, _dp), p(Ck, _dp))
    assert rc == 0
    return gm, rnk, ind, Nk, xk.reshape(G, 3), Ck.reshape(G, 6)


def local_loglik(frame, gm, ind, K, lists, out=None):
    """Oracle local log-likelihood of each list of supervoxel indices (lists:
    sequence of integer sequences); empty lists keep `out`'s value (0 by
    default)."""
    L = oracle()
    gm = np.ascontiguousarray(gm)
    ind = np.ascontiguousarray(ind, np.int32)
    ptr = np.zeros(len(lists) + 1, np.int64)
    ptr[1:] = np.cumsum([len(x) for x in lists])
    idx = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32) for x in lists])
                               if ptr[-1] else np.zeros(1, np.int32), np.int32)
    out = np.zeros(len(lists)) if out is None else np.array(out, np.float64)
    L.orc_local_loglik(frame.N, frame.S, p(frame.query, _fp), p(frame.img, _fp),
                       p(frame.sv_ptr, _lp), gm.shape[0], gm.ctypes.data, K, p(ind, _ip),
                       len(lists), p(ptr, _lp), p(idx, _ip), p(out, _dp))
    return out


def marshal_frame(pixel_idx, raw, sv_ptr, thr, dims):
    """Oracle input marshalling; returns (query [3N], img [N], centroids [3S])."""
    L = oracle()
    idx = np.ascontiguousarray(pixel_idx, np.uint64)
    rw = np.ascontiguousarray(raw, np.uint16)
    ptr = np.ascontiguousarray(sv_ptr, np.int64)
    th = np.ascontiguousarray(thr, np.uint16)
    dm = np.ascontiguousarray(dims, np.int64)
    N, S = idx.shape[0], th.shape[0]
    q = np.zeros(max(3 * N, 1), np.float32)
    im = np.zeros(max(N, 1), np.float32)
    c = np.zeros(max(3 * S, 1), np.float32)
    rc = L.orc_marshal_frame(N, S, idx.ctypes.data, rw.ctypes.data, p(ptr, _lp), th.ctypes.data,
                             p(dm, _lp), p(q, _fp), p(im, _fp), p(c, _fp))
    if rc:
        raise ValueError("index outside the volume")
    return q[: 3 * N], im[:N], c[: 3 * S]


def sv_knn(query, ref, K, scale, with_dist=True):
    """Oracle supervoxel kNN (knnCUDA_): (ind [K*Q], dist [K*Q]) slot-major."""
    L = oracle()
    q = np.ascontiguousarray(query, np.float32)
    r = np.ascontiguousarray(ref, np.float32)
    Q, R = q.size // 3, r.size // 3
    ind = np.full(max(K * Q, 1), -7, np.int32)
    dist = np.zeros(max(K * Q, 1), np.float32)
    sc = np.ascontiguousarray(scale, np.float32)
    L.orc_sv_knn(p(ind, _ip), p(dist, _fp) if with_dist else None, p(q, _fp), p(r, _fp), Q, R,
                 K, p(sc, _fp))
    return ind[: K * Q], dist[: K * Q]


_twin = None


def reftwin():
    """The reference's CPU twin (None when it was not built here)."""
    global _twin
    if _twin is None:
        if not os.path.exists(REFTWIN_SO):
            return None
        L = ctypes.CDLL(REFTWIN_SO)
        L.ref_psi.restype = ctypes.c_double
        L.ref_psi.argtypes = [ctypes.c_double]
        L.ref_expected_logdet.restype = ctypes.c_double
        L.ref_expected_logdet.argtypes = [_dp, ctypes.c_double]
        L.ref_expected_logresp.restype = ctypes.c_double
        L.ref_expected_logresp.argtypes = [ctypes.c_double, ctypes.c_double]
        L.ref_update_regularize.argtypes = [_dp, ctypes.c_double, _dp, _dp,
                                            ctypes.c_int, ctypes.c_int, _fp,
                                            ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, _dp]
        L.ref_regularize.argtypes = [_dp, ctypes.c_double, ctypes.c_int, _fp,
                                     ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double]
        L.ref_update_priors.argtypes = [_dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, _dp]
        if hasattr(L, "ref_sv_centroid"):
            L.ref_sv_centroid.argtypes = [_vp, _vp, _vp, _ll, _fp, _fp]
        _twin = L
    return _twin
