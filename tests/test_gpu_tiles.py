"""Batched tile Cholesky kernels (MFMA) vs numpy on dense SPD matrices."""
import numpy as np
import pytest

from emagddm_amd import _lib

pytestmark = pytest.mark.gpu


def run(prec, A, p):
    import ctypes as C
    lib = _lib.load()
    m = A.shape[0]
    A = np.ascontiguousarray(A)
    out = np.zeros_like(A)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc = lib.emag_tile_cholesky_test(prec, m, p, dp(A), dp(out))
    return rc, out


def expected(A, p):
    L11 = np.linalg.cholesky(A[:p, :p])
    L21 = np.linalg.solve(L11, A[:p, p:]).T
    S = A[p:, p:] - L21 @ L21.T
    E = np.zeros_like(A)
    E[:p, :p] = L11
    E[p:, :p] = L21
    E[p:, p:] = np.tril(S)
    return E


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-5)])
@pytest.mark.parametrize("m,p", [(64, 64), (128, 64), (192, 128), (256, 128), (320, 192)])
def test_tile_partial_cholesky(gpu_available, prec, tol, m, p):
    rng = np.random.default_rng(m + p)
    G = rng.normal(size=(m, m))
    A = G @ G.T / m + np.eye(m)
    print(f"[tiles] prec={prec} m={m} p={p}", flush=True)
    rc, out = run(prec, A, p)
    assert rc == 0
    E = expected(A, p)
    assert np.abs(out - E).max() <= tol * np.abs(E).max()


def test_tile_detects_indefinite(gpu_available):
    A = np.eye(128)
    A[70, 70] = -1.0
    rc, _ = run(0, A, 128)
    assert rc == _lib.NUMERICAL_ERROR


@pytest.mark.parametrize("prec", [0, 1])
def test_mfma_accumulator_layout(gpu_available, prec):
    """The tile kernels assume D[row][l%16] for lane l, register r with
    row = l/16 + 4r (f64) or 4(l/16) + r (f32)."""
    import ctypes as C
    lib = _lib.load()
    out = np.zeros(512)
    assert lib.emag_mfma_layout_probe(prec, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    out = out.reshape(2, 64, 4)
    print("p-index per lane/reg:", out[0, ::16].tolist(), "q:", out[1, :4, 0].tolist())
    lane = np.arange(64)[:, None]
    reg = np.arange(4)[None, :]
    rows = (lane // 16 + 4 * reg) if prec == 0 else (4 * (lane // 16) + reg)
    assert np.array_equal(out[0], rows)
    assert np.array_equal(out[1], np.broadcast_to(lane % 16, (64, 4)))
