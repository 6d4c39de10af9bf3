"""Writes tests/golden/reftwin_vectors.npz: input/output vectors of the
reference's own CPU update equations (GaussianMixtureModel.h:229-368,
GaussianMixtureModel.cpp:892-910, GSL psi.c), and
tests/golden/sv_centroid_vectors.npz: supervoxels over a uint16 volume with
the centroid and intensity the reference's
supervoxel::weightedGaussianStatistics<uint16> (supervoxel.cpp:1410-1449)
gives them, through the wrapper built in oracle/_ref (make -C oracle ref).
Run in a tree where /root/reference exists:

    python tests/golden/make_golden.py             # both files
    python tests/golden/make_golden.py centroids   # the centroid file only
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


def rand_spd(rng, n, scale_diag=(0.002, 0.05), aniso=25.0, w4dof=False):
    out = np.zeros((n, 6))
    for i in range(n):
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        d = rng.uniform(*scale_diag, size=3)
        M = Q @ np.diag(d) @ Q.T
        M[2, 2] *= aniso
        M[0, 2] *= np.sqrt(aniso)
        M[2, 0] = M[0, 2]
        M[1, 2] *= np.sqrt(aniso)
        M[2, 1] = M[1, 2]
        if w4dof:
            M[0, 2] = M[2, 0] = M[1, 2] = M[2, 1] = 0
        out[i] = [M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]]
    return out


def main():
    T = O.reftwin()
    if T is None:
        raise SystemExit("oracle/_ref/libtgmm_reftwin.so not built (make -C oracle ref)")
    rng = np.random.default_rng(20261202)
    dp = O._dp
    fp = O._fp
    # digamma
    xs = np.concatenate([np.geomspace(1e-7, 1e4, 200), -np.linspace(0.05, 9.95, 60) + 1e-3,
                         np.array([0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 11.5, 12.0, 12.5])])
    psi = np.array([T.ref_psi(float(x)) for x in xs])
    # expectations
    W = rand_spd(rng, 64)
    nu = rng.uniform(3.5, 400, 64)
    eld = np.array([T.ref_expected_logdet(O.p(np.ascontiguousarray(W[i]), dp), float(nu[i]))
                    for i in range(64)])
    al = rng.uniform(0.0, 500, 64)
    tot = rng.uniform(600, 5e4, 64)
    elp = np.array([T.ref_expected_logresp(float(al[i]), float(tot[i])) for i in range(64)])
    # posterior + regularisation
    n = 256
    prior = np.zeros((n, 12))
    prior[:, 0] = rng.uniform(1e-3, 50, n)     # beta_o
    prior[:, 1] = rng.uniform(3, 200, n)       # nu_o
    prior[:, 2] = rng.uniform(0, 10, n)        # alpha_o
    prior[:, 3:6] = rng.uniform(0, 2000, (n, 3))
    prior[:, 6:12] = rand_spd(rng, n, (0.01, 0.3))
    Nk = rng.uniform(0, 400, n)
    Nk[:8] = 0.0
    xk = prior[:, 3:6] + rng.normal(scale=3.0, size=(n, 3))
    Sk = np.linalg.inv(rand_spd_full(rng, n)) if False else rand_cov(rng, n)
    scale = np.array([1, 1, 5], np.float32)
    outs = {}
    for reg in (0, 1):
        for w4 in (0, 1):
            res = np.zeros((n, 18))
            for i in range(n):
                T.ref_update_regularize(O.p(np.ascontiguousarray(prior[i]), dp), float(Nk[i]),
                                        O.p(np.ascontiguousarray(xk[i]), dp),
                                        O.p(np.ascontiguousarray(Sk[i]), dp), reg, w4,
                                        O.p(scale, fp), 0.02, 0.2222, 9.0,
                                        O.p(res[i], dp))
            outs["upd_reg%d_w4%d" % (reg, w4)] = res
    # regularisation only, with clamps active (wide eigenvalue spread)
    Wr = rand_spd(rng, 256, (1e-5, 2.0))
    nur = rng.uniform(3, 300, 256)
    regs = {}
    for w4 in (0, 1):
        R = Wr.copy()
        for i in range(256):
            T.ref_regularize(O.p(R[i], dp), float(nur[i]), w4, O.p(scale, fp), 0.02, 0.2222, 9.0)
        regs["reg_w4%d" % w4] = R
    # priors for the next frame
    st = np.zeros((64, 14))
    st[:, 0] = rng.uniform(1, 300, 64)
    st[:, 1] = rng.uniform(4, 300, 64)
    st[:, 2] = rng.uniform(0, 300, 64)
    st[:, 3:6] = rng.uniform(0, 2000, (64, 3))
    st[:, 6:12] = rand_spd(rng, 64, (0.001, 0.05))
    st[:, 12] = rng.uniform(0, 300, 64)
    st[:, 13] = rng.uniform(0, 5, 64)
    pri = np.zeros((64, 12))
    pct = (0.5, 0.5, 0.8)
    for i in range(64):
        T.ref_update_priors(O.p(st[i], dp), pct[0], pct[1], pct[2], 1000.0 + i, O.p(pri[i], dp))
    np.savez_compressed(os.path.join(HERE, "reftwin_vectors.npz"), psi_x=xs, psi=psi,
                        W=W, nu=nu, eld=eld, alpha=al, total=tot, elp=elp, prior=prior, Nk=Nk,
                        xk=xk, Sk=Sk, Wr=Wr, nur=nur, st=st, pri=pri, pct=np.array(pct),
                        **outs, **regs)
    print("wrote reftwin_vectors.npz")


def rand_cov(rng, n):
    out = np.zeros((n, 6))
    for i in range(n):
        A = rng.normal(size=(3, 3)) * np.array([3.0, 3.0, 0.8])
        M = A @ A.T / 3.0 + np.diag([0.2, 0.2, 0.05])
        out[i] = [M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]]
    return out


def centroid_vectors():
    """Disjoint supervoxels (random blobs, voxels listed in random order) over
    a 97 x 83 x 29 volume: intensities 100-4000, some near 65535 (sums beyond
    2^24, so the float accumulation order shows), one all-zero supervoxel (the
    N_k < 1e-5 branch) and single-voxel ones."""
    T = O.reftwin()
    if T is None or not hasattr(T, "ref_sv_centroid"):
        raise SystemExit("oracle/_ref/libtgmm_reftwin.so not built (make -C oracle ref)")
    rng = np.random.default_rng(20261203)
    dims = np.array([97, 83, 29], np.uint64)
    vol_size = int(np.prod(dims))
    used = np.zeros(vol_size, bool)
    vol = np.zeros(vol_size, np.uint16)
    idx_all, ptr, thr = [], [0], []
    for s in range(300):
        c = rng.integers([0, 0, 0], dims.astype(np.int64))
        ext = rng.integers(1, [9, 9, 4])
        lo = np.maximum(c - ext, 0)
        hi = np.minimum(c + ext + 1, dims.astype(np.int64))
        xs, ys, zs = np.meshgrid(*[np.arange(a, b) for a, b in zip(lo, hi)], indexing="ij")
        lin = (xs + dims[0] * (ys + dims[1] * zs)).reshape(-1).astype(np.int64)
        lin = lin[~used[lin]]
        n = int(rng.integers(1, 400)) if s % 7 else 1
        if lin.size == 0:
            continue
        lin = rng.permutation(lin)[:n]
        used[lin] = True
        kind = s % 11
        if kind == 3:
            v = np.zeros(lin.size, np.uint16)
        elif kind in (5, 9):
            v = rng.integers(60000, 65536, lin.size).astype(np.uint16)
        else:
            v = rng.integers(100, 4001, lin.size).astype(np.uint16)
        vol[lin] = v
        idx_all.append(lin.astype(np.uint64))
        ptr.append(ptr[-1] + lin.size)
        thr.append(int(v.min()) - int(rng.integers(0, 2)) if v.min() > 0 else 0)
    idx = np.concatenate(idx_all)
    ptr = np.array(ptr, np.int64)
    S = ptr.size - 1
    cent = np.zeros((S, 3), np.float32)
    inten = np.zeros(S, np.float32)
    for s in range(S):
        seg = np.ascontiguousarray(idx[ptr[s]:ptr[s + 1]])
        T.ref_sv_centroid(vol.ctypes.data, dims.ctypes.data, seg.ctypes.data, seg.size,
                          O.p(cent[s], O._fp), O.p(inten[s:s + 1], O._fp))
    np.savez_compressed(os.path.join(HERE, "sv_centroid_vectors.npz"), dims=dims, idx=idx,
                        raw=vol[idx.astype(np.int64)], ptr=ptr,
                        thr=np.array(thr, np.uint16), centroid=cent, intensity=inten)
    print("wrote sv_centroid_vectors.npz (%d supervoxels, %d voxels)" % (S, idx.size))


if __name__ == "__main__":
    if sys.argv[1:] != ["centroids"]:
        main()
    centroid_vectors()
