"""z2_qdrp_batched_noalloc (workspace in the library's static device array)
against z2_qdrp_batched (workspace allocated per call): identical results, on two
calls in a row, the second with a smaller input (fewer and smaller matrices), so
that the shared static workspace is reused with stale contents from the first
call; and an input beyond the static workspace is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from z2_gauge_zfy_amd import _lib
    return _lib.lib()


def _mats(n, nmat, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.standard_normal((n, n)) @ np.diag(np.logspace(0, 6, n)) for _ in range(nmat)])


def test_noalloc_matches_allocating_entry_point(L):
    from test_gpu_kernels import qdrp_on_device
    for n, nmat, seed in ((256, 4, 11), (100, 2, 12)):  # second call smaller
        mats = _mats(n, nmat, seed)
        ref = qdrp_on_device(L.z2_qdrp_batched, mats)
        got = qdrp_on_device(L.z2_qdrp_batched_noalloc, mats)
        for a, b, name in zip(got, ref, ("A", "D", "ipvt", "tau", "Q")):
            assert np.array_equal(a, b), (n, nmat, name)
        print(f"n={n} nmat={nmat}: identical", flush=True)


def test_noalloc_rejects_inputs_beyond_the_workspace(L):
    import z2_gauge_zfy_amd as z
    from test_gpu_kernels import _hip
    from z2_gauge_zfy_amd._lib import check
    h = _hip()
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), 64) == 0  # never touched: the call fails first
    try:
        n, nmat = 512, 17  # 17 * 512^2 > 4194304 doubles
        with pytest.raises(z.Z2Error, match="exceeds"):
            check(L.z2_qdrp_batched_noalloc(n, nmat, p, p, p, p, p, None))
    finally:
        h.hipFree(p)
