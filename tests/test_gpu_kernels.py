"""FP64 dense kernels of the stabilization against LAPACK semantics (GPU)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.fixture(scope="module")
def L():
    from z2_gauge_zfy_amd import _lib
    return _lib.lib()


@pytest.mark.parametrize("n", [16, 70, 144])
@pytest.mark.parametrize("ta,tb,triA,triB", [(0, 0, 0, 0), (1, 1, 0, 0), (1, 0, 0, 0), (0, 1, 0, 0),
                                             (0, 0, 1, 0), (0, 1, 0, 1)])
def test_gemm(L, n, ta, tb, triA, triB):
    from z2_gauge_zfy_amd._lib import check
    rng = np.random.default_rng(n + 10 * ta + 100 * tb + 1000 * triA)
    A = np.asfortranarray(rng.standard_normal((n, n)))
    B = np.asfortranarray(rng.standard_normal((n, n)))
    Cm = np.zeros((n, n), order="F")
    check(L.z2_test_gemm(n, _pd(A), ta, _pd(B), tb, triA, triB, _pd(Cm)))
    a = np.triu(A) if triA else A
    b = np.triu(B) if triB else B
    a = a.T if ta else a
    b = b.T if tb else b
    np.testing.assert_allclose(Cm, a @ b, rtol=1e-12, atol=1e-12 * n)


@pytest.mark.parametrize("n", [16, 100, 256, 600])
def test_qrp_matches_oracle(L, n):
    from z2_gauge_zfy_amd._lib import check
    rng = np.random.default_rng(n)
    A0 = rng.standard_normal((n, n)) @ np.diag(np.logspace(0, 10, n))[:, rng.permutation(n)]
    A = np.asfortranarray(A0.copy())
    D = np.zeros(n)
    ip = np.zeros(n, np.int32)
    tau = np.zeros(n)
    Q = np.zeros((n, n), order="F")
    check(L.z2_test_qrp(n, _pd(A), _pd(D), ip.ctypes.data_as(C.POINTER(C.c_int32)), _pd(tau), _pd(Q)))
    Mo, Do, ipo, tauo = O.qdrp(A0)
    assert np.array_equal(ip, ipo)  # same pivot sequence
    np.testing.assert_allclose(D, Do, rtol=1e-12)
    np.testing.assert_allclose(np.triu(A), np.triu(Mo), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(Q.T @ Q, np.eye(n), atol=1e-13)
    np.testing.assert_allclose(Q @ (np.triu(A) * D[:, None]), A0[:, ip - 1], rtol=1e-11,
                               atol=1e-11 * np.abs(A0).max())


@pytest.mark.parametrize("n", [16, 100, 144, 256, 400, 512])
def test_qr_blocked(L, n):
    """Blocked QR with columns pre-ordered by norm: A P = Q diag(D) R, Q orthogonal,
    R the Householder R of the ordered matrix (numpy's LAPACK QR, same sign rule)."""
    from z2_gauge_zfy_amd._lib import check
    rng = np.random.default_rng(100 + n)
    A0 = rng.standard_normal((n, n)) @ np.diag(np.logspace(0, 10, n))[:, rng.permutation(n)]
    A = np.asfortranarray(A0.copy())
    D = np.zeros(n)
    ip = np.zeros(n, np.int32)
    tau = np.zeros(n)
    Q = np.zeros((n, n), order="F")
    check(L.z2_test_qr_blocked(n, _pd(A), _pd(D), ip.ctypes.data_as(C.POINTER(C.c_int32)), _pd(tau),
                               _pd(Q)))
    norms = np.linalg.norm(A0, axis=0)
    assert sorted(ip - 1) == list(range(n))
    assert np.all(np.diff(norms[ip - 1]) <= 0)  # decreasing column norms
    Ap = A0[:, ip - 1]
    R = np.triu(A) * D[:, None]
    np.testing.assert_allclose(np.abs(np.diag(np.triu(A))), 1.0, rtol=1e-14)
    np.testing.assert_allclose(Q.T @ Q, np.eye(n), atol=1e-13)
    np.testing.assert_allclose(Q @ R, Ap, rtol=1e-11, atol=1e-11 * np.abs(A0).max())
    Rn = np.linalg.qr(Ap, mode="r")
    scale = np.abs(Rn).max(axis=1, keepdims=True)
    assert np.abs(np.abs(R) - np.abs(Rn)).max() < 1e-10 * scale.max()
    np.testing.assert_allclose(D, np.abs(np.diag(Rn)), rtol=1e-9)


@pytest.mark.parametrize("n", [16, 130])
@pytest.mark.parametrize("right", [0, 1])
def test_trsm(L, n, right):
    from z2_gauge_zfy_amd._lib import check
    rng = np.random.default_rng(n + right)
    R = np.asfortranarray(np.triu(rng.standard_normal((n, n))) * 0.1 + np.diag(rng.choice([-1.0, 1.0], n)))
    X0 = rng.standard_normal((n, n))
    X = np.asfortranarray(X0.copy())
    check(L.z2_test_trsm(n, _pd(R), _pd(X), right))
    want = np.linalg.solve(np.triu(R), X0) if right == 0 else np.linalg.solve(np.triu(R), X0.T).T
    np.testing.assert_allclose(X, want, rtol=1e-10, atol=1e-10)
