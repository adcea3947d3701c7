"""The allocation-free tridiagonalization entry point (skh_sytrd_d_device_noalloc) against
skh_sytrd_d_device: the same kernels on the same input give the same bits (factors, tau, eta),
for two calls in a row, the second with a smaller (odd, padded leading dimension) matrix."""
import ctypes
import sys

import numpy as np
import pytest

from conftest import skew_gauss

pytestmark = pytest.mark.gpu

skl = pytest.importorskip("skewlinearalgebra")
from skewlinearalgebra import _lib  # noqa: E402


def _factor(torch, entry, A, ld, nb):
    n = A.shape[0]
    dev = torch.device("cuda:0")
    work = torch.zeros((n, ld), dtype=torch.float64, device=dev)   # column c at row c: ld-strided columns
    work[:, :n] = torch.tensor(A.T.copy(), device=dev)
    tau = torch.full((n,), 5.0, dtype=torch.float64, device=dev)
    eta = torch.full((n,), 5.0, dtype=torch.float64, device=dev)
    o = _lib.opts(nb, skl.ENGINE_LARGE)
    rc = getattr(_lib.lib(), entry)(n, work.data_ptr(), ld, tau.data_ptr(), eta.data_ptr(),
                                    torch.cuda.current_stream().cuda_stream, ctypes.byref(o))
    _lib.check(entry, rc)
    _lib.check("skh_device_status", _lib.lib().skh_device_status(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    F = work.cpu().numpy()[:, :n].T   # column-major factors as an n x n array
    return F, tau[: n - 1].cpu().numpy(), eta[: n - 1].cpu().numpy()


def test_noalloc_matches_device_entry():
    torch = pytest.importorskip("torch")
    for n, ld, nb in ((3000, 3000, 0), (999, 1000, 64)):
        print("noalloc n=%d ld=%d nb=%d" % (n, ld, nb), flush=True)
        sys.stdout.flush()
        A = skew_gauss(n, 7000 + n)
        Fz, tz, ez = _factor(torch, "skh_sytrd_d_device_noalloc", A, ld, nb)
        Fd, td, ed = _factor(torch, "skh_sytrd_d_device", A, ld, nb)
        assert np.array_equal(np.tril(Fz, -1), np.tril(Fd, -1))
        assert np.array_equal(tz, td)
        assert np.array_equal(ez, ed)
        iu = np.triu_indices(n)
        assert np.array_equal(Fz[iu], A[iu])   # diagonal and upper triangle left as given
        nA = np.linalg.norm(A)
        assert abs(np.sqrt(2.0 * np.sum(ez * ez)) - nA) <= 1e-13 * nA


def test_noalloc_rejects_sizes_beyond_its_scratch():
    L = _lib.lib()
    n = 8194
    rc = L.skh_sytrd_d_device_noalloc(n, None, n, None, None, None, None)
    assert rc == -1
    assert b"fixed scratch" in L.skh_last_error()
