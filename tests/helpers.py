"""Shared helpers for the parity tests."""
import numpy as np

from tgmm_amd.records import GMM_DTYPE, is_dead


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(b), floor)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(den > 0, np.abs(a - b) / den, np.abs(a - b))
    return float(np.max(r)) if r.size else 0.0


def record_errors(gm, ref):
    """Max relative errors over the Gaussians alive on either side (a
    Gaussian one side killed and the other kept compares its -1e32 mean and
    fails): means (per component, relative to |m|), precision (Frobenius),
    nu, beta, alpha, and the expectations E[ln det Lambda] ("eld") and
    E[ln pi] ("elp"), relative with a floor of 1 (they are logarithms near 0),
    plus "dead", the number of Gaussians whose dead/alive state differs."""
    dead_a, dead_b = is_dead(gm), is_dead(ref)
    live = ~dead_a | ~dead_b
    out = {}
    if not live.any():
        return dict(m=0.0, W=0.0, nu=0.0, beta=0.0, alpha=0.0, eld=0.0, elp=0.0, dead=0)
    a, b = gm[live], ref[live]
    out["m"] = float(np.max(np.linalg.norm(a["m_k"] - b["m_k"], axis=1) /
                            np.linalg.norm(b["m_k"], axis=1)))
    dw = np.linalg.norm(a["W_k"] - b["W_k"], axis=1)
    out["W"] = float(np.max(dw / np.linalg.norm(b["W_k"], axis=1)))
    for f in ("nu_k", "beta_k", "alpha_k"):
        out[f[:-2]] = rel_err(a[f], b[f], floor=1e-300)
    out["eld"] = rel_err(a["expectedLogDetCovarianceCUDA"], b["expectedLogDetCovarianceCUDA"],
                         floor=1.0)
    out["elp"] = rel_err(a["expectedLogResponsivityCUDA"], b["expectedLogResponsivityCUDA"],
                         floor=1.0)
    out["dead"] = int(np.sum(dead_a != dead_b))
    return out


def dup_gaussians(frame, offset=1.5, seed=7):
    """Two Gaussians per cell (the second displaced by `offset` voxels), so
    that supervoxels are shared and responsibilities are soft."""
    rng = np.random.default_rng(seed)
    g = frame.gm
    g2 = g.copy()
    d = rng.normal(size=(g.shape[0], 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    g2["m_k"] += offset * d
    g2["m_o"] = g2["m_k"]
    out = np.concatenate([g, g2]).astype(GMM_DTYPE)
    import copy
    f2 = copy.copy(frame)
    f2.gm = out
    return f2


def next_priors(gm, beta_pct, nu_pct, alpha_pct, alpha_total):
    """A stream's next-frame Gaussians from the previous posterior:
    GaussianMixtureModel::updatePriors (GaussianMixtureModel.cpp:892-910) with
    N_k = nu_k - nu_o, then nu_k *= 2 and W_k /= 2 (main.cpp:961-969)."""
    g = gm.copy()
    nk = g["nu_k"] - g["nu_o"]
    g["m_o"] = g["m_k"]
    g["beta_o"] = np.maximum(beta_pct * nk, 1e-5)
    nu_o = np.maximum(nu_pct * nk, 3.0)
    g["W_o"] = g["W_k"] * g["nu_k"][:, None] / nu_o[:, None]
    g["nu_o"] = nu_o
    if alpha_total >= 0:
        g["alpha_o"] = alpha_total * 1e-5 * alpha_pct
    g["nu_k"] = g["nu_k"] * 2.0
    g["W_k"] = g["W_k"] / 2.0
    return g


def _butterfly64(v):
    """Sum over 64 lanes by xor-butterfly (strides 32 .. 1); lane 0's value."""
    a = list(v)
    for m in (32, 16, 8, 4, 2, 1):
        a = [a[i] + a[i ^ m] for i in range(64)]
    return a[0]


def device_alpha_sum(alpha):
    """The device's per-frame sum of alpha_k, in its summation order: blocks of
    64 Gaussians over the 256-aligned range, each summed by a 64-lane
    butterfly; the block partials summed by one 64-lane wave, lane l taking
    partials l, l+64, ... into four interleaved sums, then a butterfly."""
    a = [float(x) for x in np.asarray(alpha, np.float64)]
    nb = ((max(len(a), 1) + 255) // 256) * 4
    a += [0.0] * (nb * 64 - len(a))
    parts = [_butterfly64(a[64 * b:64 * (b + 1)]) for b in range(nb)]
    acc = []
    for t in range(64):
        a4 = [0.0, 0.0, 0.0, 0.0]
        i = t
        while i + 192 < nb:
            for q in range(4):
                a4[q] += parts[i + 64 * q]
            i += 256
        q = 0
        while i < nb:
            a4[q & 3] += parts[i]
            i += 64
            q += 1
        acc.append((a4[0] + a4[1]) + (a4[2] + a4[3]))
    return _butterfly64(acc)


def records_equal(a, b):
    """Bitwise equality of every named field of two record arrays (the three
    padding bytes after `fixed` are not compared: numpy's copy of a padded
    structured array need not copy them)."""
    if a.shape != b.shape:
        return False
    return all(np.ascontiguousarray(a[k]).tobytes() == np.ascontiguousarray(b[k]).tobytes()
               for k in a.dtype.names if k)


def raw_from_frame(frame, dims, seed=0, bright=3000):
    """A raw frame (pixel_idx, raw, sv_ptr, thr, dims) with the frame's
    integer voxel coordinates and supervoxels: raw = thr + round(img * bright)
    with thr drawn per supervoxel, voxels in their list order."""
    rng = np.random.default_rng(seed)
    N, S = frame.N, frame.S
    q = frame.query.reshape(3, N).astype(np.int64)
    d = np.asarray(dims, np.int64)
    idx = (q[0] + d[0] * (q[1] + d[1] * q[2])).astype(np.uint64)
    thr = rng.integers(50, 400, S).astype(np.uint16)
    per = np.repeat(thr.astype(np.int64), np.diff(frame.sv_ptr))
    raw = (per + np.round(frame.img.astype(np.float64) * bright)).astype(np.uint16)
    return idx, raw, frame.sv_ptr.astype(np.int64), thr, d


def scan_replay(query, ref, K, scale):
    """A plain-Python replay of the reference's nearest-neighbour scan (the
    EM's knnKernel, UtilsCUDA/knnCuda.cu:61-144, and knnCUDA_'s kernel,
    temporalLogicalRules/knnCUDA/knnCuda.cu:56-140; findMaxPosition is the
    same in both), float32 operation by operation, for small inputs. query and
    ref are SoA (x.., y.., z..). The references are scanned in index order;
    the x term, the x + y partial sum and the full sum are each tested against
    the threshold (`dist > minDistThr` leaves the reference); a reference that
    passes replaces the slot at minPos, then the first slot holding the
    largest distance becomes minPos (`minDist < distArray[ii]`, so a NaN in
    slot 0 keeps the threshold NaN). Returns slot-major (ind, dist) [K * Q];
    slots never filled keep (-1, 1e32)."""
    f32 = np.float32
    q = np.asarray(query, f32).reshape(3, -1)
    r = np.asarray(ref, f32).reshape(3, -1)
    s = [f32(x) for x in scale]
    Q, R = q.shape[1], r.shape[1]
    ind = np.full((K, Q), -1, np.int32)
    dist = np.full((K, Q), f32(1e32), f32)
    with np.errstate(over="ignore", invalid="ignore"):
        for i in range(Q):
            md = [f32(1e32)] * K
            mi = [-1] * K
            thr, pos = f32(1e32), 0
            for ii in range(R):
                a = (q[0, i] - r[0, ii]) * s[0]
                d = a * a
                if d > thr:
                    continue
                a = (q[1, i] - r[1, ii]) * s[1]
                d = d + a * a
                if d > thr:
                    continue
                a = (q[2, i] - r[2, ii]) * s[2]
                d = d + a * a
                if d > thr:
                    continue
                md[pos], mi[pos] = d, ii
                thr, pos = md[0], 0
                for j in range(1, K):
                    if thr < md[j]:
                        thr, pos = md[j], j
            ind[:, i] = mi
            dist[:, i] = md
    return ind.reshape(-1), dist.reshape(-1)
