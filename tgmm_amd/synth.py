"""Seeded synthetic frames for the five measured workloads.

Thin ctypes wrapper over ``csrc/synth.cpp`` (host C++). The generator builds
the tracker's boundary layout (SoA float voxel coordinates, float [0,1]
intensities, supervoxel pointer array, intensity-weighted centroids) and the
initial per-Gaussian records (priors alpha_o=1, beta_o=1, nu_o=4, W_o=I/25,
means = true centres + N(0, 2^2) voxels).
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from .records import GMM_DTYPE

_HERE = os.path.dirname(os.path.abspath(__file__))
SYNTH_SO = os.path.join(_HERE, "libtgmm_synth.so")


class _Cells(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("centre", ctypes.POINTER(ctypes.c_double)),
        ("radius", ctypes.POINTER(ctypes.c_double)),
        ("quat", ctypes.POINTER(ctypes.c_double)),
        ("bright", ctypes.POINTER(ctypes.c_double)),
    ]


class _Frame(ctypes.Structure):
    _fields_ = [
        ("n_voxels", ctypes.c_longlong),
        ("n_supervoxels", ctypes.c_longlong),
        ("n_cells", ctypes.c_int),
        ("query", ctypes.POINTER(ctypes.c_float)),
        ("img", ctypes.POINTER(ctypes.c_float)),
        ("centroids", ctypes.POINTER(ctypes.c_float)),
        ("sv_ptr", ctypes.POINTER(ctypes.c_longlong)),
        ("sv_cell", ctypes.POINTER(ctypes.c_int)),
        ("raw", ctypes.POINTER(ctypes.c_ushort)),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_SO):
            raise RuntimeError(
                "libtgmm_synth.so is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(SYNTH_SO)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.tgmm_synth_sample_cells.argtypes = [
            ctypes.c_uint64, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.POINTER(_Cells)]
        lib.tgmm_synth_rasterize.argtypes = [
            ctypes.c_uint64, ctypes.POINTER(_Cells), dp, ctypes.c_double,
            ctypes.c_double, ctypes.c_int, ctypes.POINTER(_Frame)]
        lib.tgmm_synth_init_gaussians.argtypes = [
            ctypes.c_uint64, ctypes.POINTER(_Cells), ctypes.c_double,
            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_void_p]
        lib.tgmm_synth_free.argtypes = [ctypes.POINTER(_Frame)]
        assert lib.tgmm_synth_record_size() == GMM_DTYPE.itemsize
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Cells:
    centre: np.ndarray   # (n, 3) voxel units
    radius: np.ndarray   # (n, 3) physical units
    quat: np.ndarray     # (n, 4)
    bright: np.ndarray   # (n,)

    @property
    def n(self):
        return self.centre.shape[0]

    def _c(self):
        self.centre = np.ascontiguousarray(self.centre, dtype=np.float64)
        self.radius = np.ascontiguousarray(self.radius, dtype=np.float64)
        self.quat = np.ascontiguousarray(self.quat, dtype=np.float64)
        self.bright = np.ascontiguousarray(self.bright, dtype=np.float64)
        return _Cells(self.n, _dp(self.centre), _dp(self.radius),
                      _dp(self.quat), _dp(self.bright))


@dataclass
class Frame:
    """One frame in the reference's boundary layout (host numpy arrays)."""
    query: np.ndarray       # (3N,) float32 SoA
    img: np.ndarray         # (N,) float32
    centroids: np.ndarray   # (3S,) float32 SoA
    sv_ptr: np.ndarray      # (S+1,) int64
    gm: np.ndarray          # (G,) GMM_DTYPE initial records
    sv_cell: np.ndarray = None
    true_centres: np.ndarray = None
    meta: dict = field(default_factory=dict)

    @property
    def N(self):
        return int(self.img.shape[0])

    @property
    def S(self):
        return int(self.sv_ptr.shape[0] - 1)

    @property
    def G(self):
        return int(self.gm.shape[0])


def sample_cells(seed, n, vol, aniso_z=5.0, rmin=3.0, rmax=6.0, pair_frac=0.0,
                 pd_min=6.0, pd_max=10.0, bsig=0.0):
    lib = _load()
    c = Cells(np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 4)), np.zeros(n))
    cc = c._c()
    v = np.ascontiguousarray(vol, dtype=np.float64)
    rc = lib.tgmm_synth_sample_cells(seed, n, _dp(v), aniso_z, rmin, rmax,
                                     pair_frac, pd_min, pd_max, bsig,
                                     ctypes.byref(cc))
    if rc != 0:
        raise RuntimeError("tgmm_synth_sample_cells failed: %d" % rc)
    return c


def rasterize(seed, cells, vol, aniso_z, target_voxels, max_sv_size=64,
              init_seed=None, init_sd=2.0, alpha_o=1.0, beta_o=1.0, nu_o=4.0,
              w_diag=1.0 / 25.0):
    lib = _load()
    cc = cells._c()
    v = np.ascontiguousarray(vol, dtype=np.float64)
    fr = _Frame()
    rc = lib.tgmm_synth_rasterize(seed, ctypes.byref(cc), _dp(v), aniso_z,
                                  float(target_voxels), max_sv_size,
                                  ctypes.byref(fr))
    if rc != 0:
        raise RuntimeError("tgmm_synth_rasterize failed: %d" % rc)
    try:
        N, S = fr.n_voxels, fr.n_supervoxels
        query = np.ctypeslib.as_array(fr.query, shape=(3 * max(N, 1),))[: 3 * N].copy()
        img = np.ctypeslib.as_array(fr.img, shape=(max(N, 1),))[:N].copy()
        cent = np.ctypeslib.as_array(fr.centroids, shape=(3 * max(S, 1),))[: 3 * S].copy()
        ptr = np.ctypeslib.as_array(fr.sv_ptr, shape=(S + 1,)).copy()
        svc = np.ctypeslib.as_array(fr.sv_cell, shape=(max(S, 1),))[:S].copy()
    finally:
        lib.tgmm_synth_free(ctypes.byref(fr))
    gm = np.zeros(cells.n, dtype=GMM_DTYPE)
    rc = lib.tgmm_synth_init_gaussians(
        seed ^ 0x5A5A5A5A if init_seed is None else init_seed, ctypes.byref(cc),
        init_sd, alpha_o, beta_o, nu_o, w_diag, gm.ctypes.data)
    if rc != 0:
        raise RuntimeError("tgmm_synth_init_gaussians failed: %d" % rc)
    return Frame(query, img, cent, ptr, gm, svc, cells.centre.copy())


# ---- the measured workloads -------------------------------------------------

CONFIG_SEEDS = {1: 0xC0FFEE01, 2: 0xC0FFEE02, 3: 0xC0FFEE03, 4: 0xC0FFEE04,
                5: 0xC0FFEE05}


def config1(seed=CONFIG_SEEDS[1], n_cells=200, target_voxels=40000):
    """200 cells in 512x512x100, aZ=5, radii U[3,6], ~40k voxels, ~800 sv."""
    vol = (512.0, 512.0, 100.0)
    cells = sample_cells(seed, n_cells, vol, 5.0, 3.0, 6.0)
    f = rasterize(seed + 1, cells, vol, 5.0, target_voxels)
    f.meta = dict(config=1, K=5, iters=10, vol=vol, aniso_z=5.0)
    return f


def config2(seed=CONFIG_SEEDS[2], n_cells=10000, target_voxels=5_000_000):
    """10k cells in 2048x2048x400, radii U[3,8], 15% close pairs, ~5M voxels."""
    vol = (2048.0, 2048.0, 400.0)
    cells = sample_cells(seed, n_cells, vol, 5.0, 3.0, 8.0, pair_frac=0.15)
    f = rasterize(seed + 1, cells, vol, 5.0, target_voxels)
    f.meta = dict(config=2, K=5, iters=20, vol=vol, aniso_z=5.0)
    return f


def config3_frames(seed=CONFIG_SEEDS[3], n_frames=64, total_voxels=25_000_000,
                   gmin=100, gmax=20000):
    """64 frames, G log-uniform [100, 20000], 30-600 voxels per Gaussian
    rescaled to ~25M voxels in total; volume per frame at config-1 density."""
    rng = np.random.default_rng(seed)
    G = np.exp(rng.uniform(np.log(gmin), np.log(gmax), n_frames)).astype(int)
    vpg = rng.uniform(30, 600, n_frames)
    vpg *= total_voxels / float(np.sum(G * vpg))
    dens = 200.0 / (512.0 * 512.0 * 100.0)
    frames = []
    for i in range(n_frames):
        vol_cells = G[i] / dens
        x = (vol_cells * 512.0 / 100.0) ** (1.0 / 3.0)
        vol = (x, x, x * 100.0 / 512.0)
        cells = sample_cells(seed + 7919 * (i + 1), int(G[i]), vol, 5.0, 3.0, 6.0)
        f = rasterize(seed + 7919 * (i + 1) + 1, cells, vol, 5.0, G[i] * vpg[i])
        f.meta = dict(config=3, K=5, iters=15, vol=vol, aniso_z=5.0)
        frames.append(f)
    return frames


def config4_streams(seed=CONFIG_SEEDS[4], n_frames=200, cells=(2000, 3000, 4000, 5500),
                    voxels_per_cell=200.0, workers=4):
    """Time-lapse streams (one per entry of `cells`, seed + stream): cells
    move by N(0, 1.5^2) voxels per frame and divide with p = 0.005, so a
    stream of 200 frames grows by about e (5,500 -> ~15,000 cells);
    ~200 voxels per cell, volume at config-1 density for the first frame.
    Streams are generated in parallel threads (the generator releases the
    GIL); returns a list of frame lists."""
    from concurrent.futures import ThreadPoolExecutor
    dens = 200.0 / (512.0 * 512.0 * 100.0)

    def one(i):
        x = (cells[i] / dens * 512.0 / 100.0) ** (1.0 / 3.0)
        vol = (x, x, x * 100.0 / 512.0)
        fr = timelapse(seed + i, cells[i], n_frames, vol, 5.0, 3.0, 6.0,
                       voxels_per_cell=voxels_per_cell)
        for f in fr:
            f.meta.update(config=4, stream=i, K=5, iters=10, vol=vol, aniso_z=5.0)
        return fr

    with ThreadPoolExecutor(max(1, min(workers, len(cells)))) as ex:
        return list(ex.map(one, range(len(cells))))


def config5(seed=CONFIG_SEEDS[5], n_cells=60000, target_voxels=30_000_000):
    """60k cells in 4096x4096x500, radii U[3,8], lognormal brightness 0.8,
    ~30M voxels; K=8, 25 iterations."""
    vol = (4096.0, 4096.0, 500.0)
    cells = sample_cells(seed, n_cells, vol, 5.0, 3.0, 8.0, bsig=0.8)
    f = rasterize(seed + 1, cells, vol, 5.0, target_voxels)
    f.meta = dict(config=5, K=8, iters=25, vol=vol, aniso_z=5.0)
    return f


def timelapse(seed, n_cells, n_frames, vol, aniso_z=5.0, rmin=3.0, rmax=6.0,
              voxels_per_cell=200.0, motion_sd=1.5, p_divide=0.005):
    """Frames of a synthetic time-lapse: cells move by N(0, motion_sd^2)
    voxels per frame and divide with probability p_divide (daughters 3
    voxels either side of the mother along a random direction). Frame 0
    carries initial Gaussians (one per cell); later frames carry the same
    Gaussian count (a served stream keeps its Gaussians)."""
    rng = np.random.default_rng(seed)
    cells = sample_cells(seed, n_cells, vol, aniso_z, rmin, rmax)
    frames = []
    for t in range(n_frames):
        f = rasterize(seed * 1000 + t, cells, vol, aniso_z, voxels_per_cell * cells.n,
                      init_seed=seed + 17)
        if t == 0:
            f.gm = f.gm[:n_cells].copy()
        else:
            f.gm = None
        f.meta = dict(stream_seed=seed, t=t, cells=cells.n)
        frames.append(f)
        # motion and divisions for the next frame
        c = cells.centre + rng.normal(scale=motion_sd, size=cells.centre.shape)
        c[:, 2] = cells.centre[:, 2] + rng.normal(scale=motion_sd / aniso_z, size=cells.n)
        div = rng.random(cells.n) < p_divide
        if div.any():
            d = rng.normal(size=(int(div.sum()), 3))
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            d *= 3.0
            d[:, 2] /= aniso_z
            kids = c[div] + d
            c[div] -= d
            c = np.concatenate([c, kids])
            cells = Cells(c, np.concatenate([cells.radius, cells.radius[div]]),
                          np.concatenate([cells.quat, cells.quat[div]]),
                          np.concatenate([cells.bright, cells.bright[div]]))
        else:
            cells = Cells(c, cells.radius, cells.quat, cells.bright)
        for a in range(3):
            np.clip(cells.centre[:, a], 0, vol[a] - 1, out=cells.centre[:, a])
    return frames
