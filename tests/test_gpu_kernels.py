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


def _hip():
    """the HIP runtime the engine library itself links (one runtime per process)"""
    from z2_gauge_zfy_amd import _lib
    _lib.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def qdrp_on_device(fn, mats):
    """run a batched QDRP entry point on device copies of `mats` ([nmat, n, n])
    through the C ABI (null stream); returns A, D, ipvt, tau, Q on the host"""
    from z2_gauge_zfy_amd._lib import check
    h = _hip()
    nmat, n, _ = mats.shape
    host = {"A": np.ascontiguousarray(np.transpose(mats, (0, 2, 1))),  # column-major per matrix
            "D": np.zeros((nmat, n)), "ip": np.zeros((nmat, n), np.int32), "tau": np.zeros((nmat, n)),
            "Q": np.zeros((nmat, n, n))}
    dev = {}
    try:
        for k, a in host.items():
            p = C.c_void_p()
            assert h.hipMalloc(C.byref(p), a.nbytes) == 0
            dev[k] = p
            assert h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        check(fn(n, nmat, dev["A"], dev["D"], dev["ip"], dev["tau"], dev["Q"], None))
        assert h.hipDeviceSynchronize() == 0
        for k, a in host.items():
            assert h.hipMemcpy(a.ctypes.data, dev[k], a.nbytes, 2) == 0  # device to host
    finally:
        for p in dev.values():
            h.hipFree(p)
    t = lambda x: np.transpose(x, (0, 2, 1))  # back to [nmat][row][col]
    return t(host["A"]), host["D"], host["ip"], host["tau"], t(host["Q"])


@pytest.mark.parametrize("n, nmat", [(144, 3), (256, 2)])
def test_qdrp_batched_device_pointers(L, n, nmat):
    """z2_qdrp_batched on matrices in device memory: each matrix's factorization is
    the single-matrix blocked QR's, bit for bit, and A P = Q diag(D) R"""
    rng = np.random.default_rng(7 + n)
    mats = np.stack([rng.standard_normal((n, n)) @ np.diag(np.logspace(0, 8, n)) for _ in range(nmat)])
    A, D, ip, tau, Q = qdrp_on_device(L.z2_qdrp_batched, mats)
    for k in range(nmat):
        A1 = np.asfortranarray(mats[k].copy())
        D1, ip1, tau1 = np.zeros(n), np.zeros(n, np.int32), np.zeros(n)
        Q1 = np.zeros((n, n), order="F")
        from z2_gauge_zfy_amd._lib import check
        check(L.z2_test_qr_blocked(n, _pd(A1), _pd(D1), ip1.ctypes.data_as(C.POINTER(C.c_int32)), _pd(tau1),
                                   _pd(Q1)))
        assert np.array_equal(A[k], A1) and np.array_equal(D[k], D1)
        assert np.array_equal(ip[k], ip1) and np.array_equal(tau[k], tau1) and np.array_equal(Q[k], Q1)
        R = np.triu(A[k]) * D[k][:, None]
        np.testing.assert_allclose(Q[k] @ R, mats[k][:, ip[k] - 1], rtol=1e-11,
                                   atol=1e-11 * np.abs(mats[k]).max())
