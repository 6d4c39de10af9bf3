"""The kernels' per-Gaussian FP64 math (tgmm_amd/csrc/tgmm_device.h), built
for the host, against the reference's CPU twin (golden vectors), the oracle,
and closed forms. Runs without a GPU."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as O
from tgmm_amd.records import w_full

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "native", "libdevmath_host.so")
GOLD = np.load(os.path.join(HERE, "golden", "reftwin_vectors.npz"))

if not os.path.exists(SO):
    pytest.skip("tests/native/libdevmath_host.so not built", allow_module_level=True)
D = ctypes.CDLL(SO)
D.dm_digamma.restype = ctypes.c_double
D.dm_digamma.argtypes = [ctypes.c_double]
D.dm_regularize.argtypes = [O._dp, ctypes.c_double, ctypes.c_int, O._fp, ctypes.c_double,
                            ctypes.c_double, ctypes.c_double]
D.dm_eig3.argtypes = [O._dp, O._dp, O._dp]
SC = np.array([1, 1, 5], np.float32)


def test_digamma_matches_gsl():
    x, ref = GOLD["psi_x"], GOLD["psi"]
    got = np.array([D.dm_digamma(float(v)) for v in x])
    pos = x > 0
    assert np.max(np.abs(got[pos] - ref[pos]) / np.abs(ref[pos])) < 1e-13
    assert D.dm_digamma(0.0) == -1e300 and D.dm_digamma(-2.0) == -1e300


@pytest.mark.parametrize("w4", [0, 1])
def test_regularize_matches_reference_twin(w4):
    for i in range(GOLD["Wr"].shape[0]):
        W = np.ascontiguousarray(GOLD["Wr"][i]).copy()
        D.dm_regularize(O.p(W, O._dp), float(GOLD["nur"][i]), w4, O.p(SC, O._fp), 0.02, 0.2222,
                        9.0)
        r = GOLD["reg_w4%d" % w4][i]
        assert np.linalg.norm(W - r) < 1e-12 * np.linalg.norm(r)


def test_eig3_closed_form_order_and_vectors():
    rng = np.random.default_rng(9)
    for _ in range(200):
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        ev = rng.uniform(0.001, 1.0, 3)
        if rng.random() < 0.3:
            ev[1] = ev[2]  # repeated eigenvalue
        M = Q @ np.diag(ev) @ Q.T
        w = np.array([M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]])
        d, v = np.zeros(3), np.zeros(9)
        D.dm_eig3(O.p(w, O._dp), O.p(d, O._dp), O.p(v, O._dp))
        s = np.sort(ev)
        assert np.allclose(d, [s[2], s[0], s[1]], atol=1e-7 * s[2])
        V = v.reshape(3, 3).T
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-9)
        assert np.allclose(V @ np.diag(d) @ V.T, M, atol=1e-7 * s[2])


def test_pair_cut_compares_in_double():
    """The kernels' M-step pair test (pair_significant) is the reference's
    float-against-double comparison (GMEM.cu:1350, 1195)."""
    D.dm_pair_significant.argtypes = [ctypes.c_float]
    one = np.float32(1e-3)
    assert D.dm_pair_significant(float(one)) == 1
    assert D.dm_pair_significant(float(np.nextafter(one, np.float32(0)))) == 0
    assert D.dm_pair_significant(0.5) == 1 and D.dm_pair_significant(0.0) == 0
    assert D.dm_pair_significant(float("nan")) == 0
