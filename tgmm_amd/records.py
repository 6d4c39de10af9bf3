"""Host-side view of the 288-byte per-Gaussian record.

The layout is the reference's ``GaussianMixtureModelCUDA``
(nmeth.3036-S2/src/UtilsCUDA/GMEMupdateCUDA.h:22-59), which is also the C
struct in ``include/tgmm_em.h``: 27 doubles, the ``fixed`` byte, 16 supervoxel
indices and their count.
"""
import numpy as np

GMM_DTYPE = np.dtype(
    {
        "names": [
            "beta_k", "nu_k", "alpha_k", "m_k", "W_k",
            "expectedLogDetCovarianceCUDA", "expectedLogResponsivityCUDA",
            "splitScore", "beta_o", "nu_o", "alpha_o", "m_o", "W_o",
            "fixed", "supervoxelIdx", "supervoxelNum",
        ],
        "formats": [
            "<f8", "<f8", "<f8", ("<f8", 3), ("<f8", 6),
            "<f8", "<f8", "<f8", "<f8", "<f8", "<f8", ("<f8", 3), ("<f8", 6),
            "u1", ("<i4", 16), "<i4",
        ],
        "offsets": [0, 8, 16, 24, 48, 96, 104, 112, 120, 128, 136, 144, 168,
                    216, 220, 284],
        "itemsize": 288,
    }
)

assert GMM_DTYPE.itemsize == 288


def empty_records(n):
    """Zero-initialised record array of length n."""
    return np.zeros(int(n), dtype=GMM_DTYPE)


def is_dead(rec):
    """Dead-Gaussian test used throughout the reference (m_o[0] < -1e31)."""
    return rec["m_o"][..., 0] < -1e31


def w_full(w6):
    """Packed [xx, xy, xz, yy, yz, zz] -> full symmetric 3x3 (last axes)."""
    w6 = np.asarray(w6)
    out = np.empty(w6.shape[:-1] + (3, 3), dtype=w6.dtype)
    idx = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for c, (a, b) in enumerate(idx):
        out[..., a, b] = w6[..., c]
        out[..., b, a] = w6[..., c]
    return out
