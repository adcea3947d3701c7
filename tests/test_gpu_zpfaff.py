"""GPU parity of the complex skew-symmetric Pfaffian (Parlett-Reid with pivoting) against the oracle.

Reference behaviour: _pfaffian!(A::AbstractMatrix{<:Complex}) and _logabspfaffian!(...) in
src/pfaffian.jl:97-137 and :193-233, behind pfaffian!/logabspfaffian! (:155-158, :259-263) which
reject input that is not exactly skew-symmetric.  The kernels and the oracle do the same complex
arithmetic (separately rounded products and sums, Smith's division), so the Pfaffian itself is
bit-identical to the oracle's; log|Pf| and the phase go through log and atan2, whose device and host
versions may differ in the last bit: log|Pf| within 1e-12 * n and cis(phase) within 1e-10 * n / 64.
Also bit for bit: a matrix's result does not depend on the batch it is in, and scaling A by 2 scales
Pf by exactly 2^(n/2).
"""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
skl = pytest.importorskip("skewlinearalgebra")
from skewlinearalgebra import _lib  # noqa: E402


def say(msg):
    print("[case] " + msg, flush=True)


def zsym(n, seed):
    """A = R - R^T with Re and Im of R iid N(0, 1): complex skew-symmetric, not skew-Hermitian."""
    rng = np.random.default_rng(seed)
    R = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return np.asfortranarray(R - R.T)


def _fixture(c):
    A = np.array(c["re"]) + 1j * np.array(c["im"])
    ex = complex(float(Fraction(c["pf"][0])), float(Fraction(c["pf"][1])))
    return A, ex


def test_zpfaffian_golden(golden):
    """Every exact Gaussian-rational fixture (the reference's own test inputs, pivoting cases, a
    singular one), in one batch."""
    cases = [_fixture(c) for c in golden["complex_pf"]]
    say("zpfaff golden x%d" % len(cases))
    pf = skl.pfaffian_z_batched([A for A, _ in cases])
    lp, ph = skl.logabspfaffian_z_batched([A for A, _ in cases])
    for (A, ex), p, l, h in zip(cases, pf, lp, ph):
        assert abs(p - ex) <= 1e-12 * max(1.0, abs(ex)), (A.shape, p, ex)
        assert p == oracle.pfaffian(A)
        if ex == 0:
            assert l == -np.inf and h == 0
        else:
            assert abs(l - np.log(abs(ex))) <= 1e-12 * max(1.0, abs(np.log(abs(ex))))
            assert abs(h - ex / abs(ex)) <= 1e-12


@pytest.mark.parametrize("n", [2, 4, 6, 16, 31, 64, 130, 256])
def test_zpfaffian_vs_oracle(n):
    A = zsym(n, 900 + n)
    say("zpfaff n=%d" % n)
    pf = skl.pfaffian(A)
    ref = oracle.pfaffian(A)
    if n % 2:
        assert pf == 0 and ref == 0
    else:
        assert pf == ref, (pf, ref)
    lp, ph = skl.logabspfaffian(A)
    lr, hr = oracle.logabspfaffian(A)
    if n % 2:
        assert lp == -np.inf and ph == 0
    else:
        assert abs(lp - lr) <= 1e-12 * n * max(1.0, abs(lr))
        assert abs(ph - hr) <= 1e-10 * max(1.0, n / 64)


@pytest.mark.parametrize("n", [300, 512, 1024])
def test_zpfaffian_log_vs_oracle(n):
    """Sizes whose Pfaffian overflows Float64: compare log|Pf| and the phase."""
    A = zsym(n, 1900 + n)
    say("zpfaff log n=%d" % n)
    lp, ph = skl.logabspfaffian(A)
    lr, hr = oracle.logabspfaffian(A)
    assert abs(lp - lr) <= 1e-12 * n * abs(lr), (lp, lr)
    assert abs(ph - hr) <= 1e-10 * n / 64, (ph, hr)


@pytest.mark.parametrize("n", [2048, 4096])
def test_zpfaffian_large_det(n):
    """Large sizes (the whole-GPU kernel): Pf^2 = det, i.e. 2 log|Pf| = log|det| and
    cis(phase)^2 = det/|det|."""
    A = zsym(n, 77)
    say("zpfaff n=%d vs slogdet" % n)
    lp, ph = skl.logabspfaffian(A)
    sgn, ld = np.linalg.slogdet(A)
    assert abs(2 * lp - ld) <= 1e-11 * n * abs(ld), (2 * lp, ld)
    assert abs(ph * ph - sgn) <= 1e-8, (ph * ph, sgn)


def test_zpfaffian_mixed_large_and_small():
    """Even members of 512 and more take the whole-GPU kernel, the rest one workgroup each, in
    one call; odd sizes beyond 2048 are answered on the host (0, and (-Inf, 0))."""
    sizes = [600, 40, 513, 7, 512, 130, 2049]
    mats = [zsym(s, 700 + i) for i, s in enumerate(sizes)]
    say("zpfaff mixed %s" % sizes)
    lp, ph = skl.logabspfaffian_z_batched(mats)
    pf = skl.pfaffian_z_batched(mats[1:4] + [mats[5]])
    for i, A in enumerate(mats):
        if sizes[i] % 2:
            assert lp[i] == -np.inf and ph[i] == 0
            continue
        lr, hr = oracle.logabspfaffian(A)
        assert abs(lp[i] - lr) <= 1e-12 * sizes[i] * abs(lr), (sizes[i], lp[i], lr)
        assert abs(ph[i] - hr) <= 1e-10 * max(1.0, sizes[i] / 64), (sizes[i], ph[i], hr)
    for A, p in zip(mats[1:4] + [mats[5]], pf):
        assert p == oracle.pfaffian(A)


def test_zpfaffian_coop_zero_pivot_and_swaps():
    """Whole-GPU kernel: a matrix whose later column vanishes gives 0 / (-Inf, 0); a Pfaffian of
    moderate size that fits Float64 matches the oracle."""
    n = 512
    A = zsym(n, 5) * 0.05
    A[:, 300] = 0
    A[300, :] = 0
    say("zpfaff coop zero column")
    assert skl.pfaffian(A) == 0
    lp, ph = skl.logabspfaffian(A)
    assert lp == -np.inf and ph == 0
    B = zsym(n, 6) * 0.05
    ref = oracle.pfaffian(B)
    got = skl.pfaffian(B)
    assert got == ref, (got, ref)


def test_zpfaffian_batch_independent_and_variants():
    """A mixed batch (sizes 0..96 and one 300 that selects the large-LDS kernel variant) gives every
    matrix the same bits as its own single-matrix call and the small-only batch."""
    rng = np.random.default_rng(5)
    sizes = [int(s) for s in rng.integers(0, 97, size=200)] + [0, 1, 2]
    mats = [zsym(s, 3000 + i) for i, s in enumerate(sizes)]
    say("zpfaff batch of %d" % len(mats))
    small = skl.pfaffian_z_batched(mats)
    big = skl.pfaffian_z_batched(mats + [zsym(300, 1)])
    assert np.array_equal(small.view(np.float64), big[:-1].view(np.float64))
    for i in (0, 7, 100, len(mats) - 1):
        one = skl.pfaffian_z_batched([mats[i]])
        assert np.array_equal(one.view(np.float64), small[i:i + 1].view(np.float64))
    for i, (A, p) in enumerate(zip(mats, small)):
        ref = oracle.pfaffian(A)
        if A.shape[0] == 0:
            assert p == 1
        elif A.shape[0] % 2:
            assert p == 0
        else:
            assert p == ref, (i, A.shape, p, ref)
    lp, ph = skl.logabspfaffian_z_batched(mats)
    assert lp[sizes.index(0)] == 0 and ph[sizes.index(0)] == 1


def test_zpfaffian_scaling_exact():
    """Pf(2A) = 2^(n/2) Pf(A) bit for bit: every step scales exactly by a power of two."""
    A = zsym(32, 11)
    say("zpfaff scaling n=32")
    p1, p2 = skl.pfaffian_z_batched([A, 2 * A])
    assert p2 == p1 * 2.0 ** 16


def test_zpfaffian_zero_pivot_and_swaps():
    """A zero pivot in a later step gives 0 / (-Inf, 0); pivoting keeps the value (Pf is
    invariant under the elimination order) and matches the oracle, which pivots the same way."""
    n = 6
    A = np.zeros((n, n), complex)
    A[1, 0], A[0, 1] = 2 + 1j, -(2 + 1j)
    A[3, 2], A[2, 3] = 1j, -1j          # column 4 of the trailing block stays zero
    say("zpfaff zero pivot")
    pf = skl.pfaffian(A)
    lp, ph = skl.logabspfaffian(A)
    assert pf == 0 and lp == -np.inf and ph == 0
    # a pivot that must swap at every step: the largest entries sit on the farthest rows
    B = zsym(8, 4) * 1e-3
    for k in range(0, 8, 2):
        B[7, k], B[k, 7] = 10 + k, -(10 + k)
    ref = oracle.pfaffian(B)
    got = skl.pfaffian(B)
    assert got == ref
    assert abs(got * got - np.linalg.det(B)) <= 1e-10 * abs(np.linalg.det(B))


def test_zpfaffian_device_entry():
    """The device entry points on torch-allocated HBM give the host entry's bits."""
    torch = pytest.importorskip("torch")
    ns = np.array([64, 0, 33, 128], dtype=np.int64)
    mats = [zsym(int(s), 40 + i) for i, s in enumerate(ns)]
    offs = np.zeros(len(ns), dtype=np.int64)
    offs[1:] = np.cumsum(ns[:-1] ** 2)
    packed = np.concatenate([M.ravel(order="F") for M in mats])
    dA = torch.from_numpy(packed.view(np.float64).copy()).cuda()
    dpf = torch.zeros(2 * len(ns), dtype=torch.float64, device="cuda")
    dlp = torch.zeros(len(ns), dtype=torch.float64, device="cuda")
    dph = torch.zeros(2 * len(ns), dtype=torch.float64, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    say("zpfaff device entry")
    L = _lib.lib()
    nsp, offp = ctypes.c_void_p(ns.ctypes.data), ctypes.c_void_p(offs.ctypes.data)
    _lib.check("pf", L.skh_pfaffian_z_batched_device(len(ns), nsp, offp, ctypes.c_void_p(dA.data_ptr()),
                                                     ctypes.c_void_p(dpf.data_ptr()), st, None))
    _lib.check("lp", L.skh_logabspfaffian_z_batched_device(len(ns), nsp, offp, ctypes.c_void_p(dA.data_ptr()),
                                                           ctypes.c_void_p(dlp.data_ptr()),
                                                           ctypes.c_void_p(dph.data_ptr()), st, None))
    torch.cuda.synchronize()
    host_pf = skl.pfaffian_z_batched(mats)
    host_lp, host_ph = skl.logabspfaffian_z_batched(mats)
    assert np.array_equal(dpf.cpu().numpy(), host_pf.view(np.float64))
    assert np.array_equal(dlp.cpu().numpy(), host_lp)
    assert np.array_equal(dph.cpu().numpy(), host_ph.view(np.float64))


def test_zpfaffian_argument_errors():
    with pytest.raises(ValueError):
        skl.pfaffian_z_batched([np.zeros((8194, 8194), complex)])
    with pytest.raises(skl.ArgumentError):
        skl.pfaffian(np.array([[0, 1j], [1j, 0]]))     # skew-Hermitian, not skew-symmetric
