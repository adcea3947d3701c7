"""CPU-side checks of the C ABI: the library loads, exports every declared symbol, rejects bad
arguments before touching a device, and its host eigensolver (the part of the path that stays
on the CPU) agrees bit-for-bit with the oracle's restatement of skewtrieigvals!/skewtrieigen."""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import skew_gauss

skl = pytest.importorskip("skewlinearalgebra")
from skewlinearalgebra import _lib  # noqa: E402


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = _lib.declared_symbols()
    assert len(names) >= 12
    for name in names:
        assert hasattr(L, name), name
    assert L.skh_version().startswith(b"skewhip")


def _eig_host(ev):
    ev = np.array(ev, dtype=np.float64, copy=True)
    vals = np.zeros(ev.size + 1)
    rc = _lib.lib().skh_skewtrieigvals_d(ev.size + 1, ev.ctypes.data, vals.ctypes.data)
    assert rc == 0
    return vals


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 10, 11, 64, 65, 300, 1001])
def test_host_eigvals_bitwise_vs_oracle(n):
    A = skew_gauss(n, 900 + n)
    _, _, eta = oracle.skewblockedhess(A)
    np.testing.assert_array_equal(_eig_host(eta), oracle.skewtrieigvals(eta))


def test_host_eigvals_fixtures(golden):
    # some of these reach the symmetric-tridiagonal fallback, where the oracle calls LAPACK
    # dstev and the library its own QL solver: same eigenvalues, rounding may differ
    for ev in golden["issue116_ev"] + golden["issue118_ev"] + [[3.0, 0.0, 5.0], [0.0, 1.0, 1.0]]:
        np.testing.assert_allclose(_eig_host(ev), oracle.skewtrieigvals(ev), rtol=0, atol=1e-14)


def test_host_eigvals_fallback_path():
    """The stagnation detector (src/skeweigen.jl:132-137) hands over to a symmetric
    tridiagonal solver; the eigenvalues must still be right (sorted comparison)."""
    ev = np.array([2.0, 1e-11, 2.0, 1e-11, 2.0] * 3)
    got = np.sort(_eig_host(ev))
    T = np.diag(ev, -1) - np.diag(ev, 1)
    ref = np.sort(np.linalg.eigvals(T).imag)
    np.testing.assert_allclose(got, ref, atol=1e-12)


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("n", [2, 3, 4, 7, 8, 33, 64])
def test_host_eigen_bitwise_vs_oracle(n, variant):
    A = skew_gauss(n, 300 + n)
    _, _, eta = oracle.skewblockedhess(A)
    ev = eta.copy()
    N = n
    vals = np.zeros(N)
    vr = np.zeros((N, N), order="F")
    vi = np.zeros((N, N), order="F")
    rc = _lib.lib().skh_skewtrieigen_d(N, ev.ctypes.data, vals.ctypes.data, vr.ctypes.data, vi.ctypes.data,
                                       variant)
    assert rc == 0
    ov, oV = oracle.skewtrieigen(eta, "divided" if variant == 0 else "merged")
    np.testing.assert_array_equal(vals, ov)
    np.testing.assert_array_equal(vr + 1j * vi, oV)


def test_host_eigen_fixtures(golden):
    for ev in golden["issue116_ev"] + golden["issue118_ev"]:
        ev = np.array(ev, float)
        N = ev.size + 1
        vals = np.zeros(N)
        vr = np.zeros((N, N), order="F")
        vi = np.zeros((N, N), order="F")
        e2 = ev.copy()
        assert _lib.lib().skh_skewtrieigen_d(N, e2.ctypes.data, vals.ctypes.data, vr.ctypes.data,
                                             vi.ctypes.data, 1) == 0
        V = vr + 1j * vi
        T = np.diag(ev, -1) - np.diag(ev, 1)
        assert np.linalg.norm(V @ np.diag(1j * vals) @ V.conj().T - T) <= 1e-9 * max(1, np.linalg.norm(T))


def test_argument_errors_before_device():
    L = _lib.lib()
    o = _lib.opts()
    tau = np.zeros(4)
    assert L.skh_sytrd_d(-1, None, 1, None, None, None, ctypes.byref(o)) == -1
    A = np.zeros((4, 4), order="F")
    assert L.skh_sytrd_d(4, A.ctypes.data, 3, tau.ctypes.data, tau.ctypes.data, None, ctypes.byref(o)) == -3
    bad = _lib.opts(nb=1000)
    assert L.skh_sytrd_d(4, A.ctypes.data, 4, tau.ctypes.data, tau.ctypes.data, None, ctypes.byref(bad)) == -7
    assert b"nb" in L.skh_last_error()
    lp, sg = ctypes.c_double(), ctypes.c_double()
    # odd n answers without a device (src/pfaffian.jl:169)
    A3 = np.zeros((3, 3), order="F")
    assert L.skh_logpf_d(3, A3.ctypes.data, 3, ctypes.byref(lp), ctypes.byref(sg), None) == 0
    assert lp.value == -np.inf and sg.value == 0.0
    assert L.skh_skewtrieigen_d(4, A.ctypes.data, tau.ctypes.data, A.ctypes.data, A.ctypes.data, 7) == -6
    # complex Pfaffians: sizes beyond 8192 and negative offsets are rejected before any device work
    ns = np.array([4, 8194], dtype=np.int64)
    offs = np.array([0, 16], dtype=np.int64)
    out = np.zeros(8)
    assert L.skh_pfaffian_z_batched(2, ns.ctypes.data, offs.ctypes.data, out.ctypes.data, out.ctypes.data, None) == -2
    assert b"8192" in L.skh_last_error()
    offs[1] = -1
    ns[1] = 2
    assert L.skh_logabspfaffian_z_batched(2, ns.ctypes.data, offs.ctypes.data, out.ctypes.data, out.ctypes.data,
                                          out.ctypes.data, None) == -3
    assert L.skh_pfaffian_z_batched(-1, None, None, None, None, None) == -1
    assert L.skh_pfaffian_z_batched(0, None, None, None, None, None) == 0


def test_python_api_host_paths():
    """Pieces of the Python mirror that need no device."""
    with pytest.raises(skl.ArgumentError):
        skl.SkewHermitian(np.ones((3, 3)))
    with pytest.raises(skl.ArgumentError):
        skl.logabspfaffian(np.ones((4, 4)))
    assert skl.logabspfaffian(np.zeros((5, 5))) == (-np.inf, 0.0)
    # complex input takes the skew-symmetric check of pfaffian! (src/pfaffian.jl:155-158, :259-263)
    for f in (skl.pfaffian, skl.pfaffian_, skl.logabspfaffian, skl.logabspfaffian_):
        with pytest.raises(skl.ArgumentError, match="skew-symmetric"):
            f(np.array([[0, 1j], [1j, 0]]))
    assert len(skl.pfaffian_z_batched([])) == 0
    S = skl.skewhermitian(np.array([[1, 2], [3, 4]]))
    np.testing.assert_array_equal(S.data, [[0.0, -0.5], [0.5, 0.0]])   # issue #98
    T = skl.SkewHermTridiagonal(np.array([3.0, 4.0, 5.0]))
    np.testing.assert_array_equal(T.to_dense(), [[0, -3, 0, 0], [3, 0, -4, 0], [0, 4, 0, -5], [0, 0, 5, 0]])
    lp, sg = skl.logabspfaffian(T)
    assert abs(sg * np.exp(lp) - skl.pfaffian(T)) < 1e-12 and skl.pfaffian(T) == 15.0
    vals = skl.eigvals(T)
    np.testing.assert_allclose(np.sort(vals.imag), np.sort(np.linalg.eigvals(T.to_dense()).imag), atol=1e-12)


def test_skewtrieigvals_range_readme(golden):
    """eigvals(A, 0, 15) and eigvals(A, 1:3) of the README matrix (test/runtests.jl:24-25), on
    the tridiagonal part the library runs on the host (reference eta as input)."""
    from skewlinearalgebra import _lib
    L = _lib.lib()
    g = golden["readme"]
    eta = np.array(g["eta"])
    l1, l2 = g["eig_pos"]
    vals = np.zeros(4)
    m = ctypes.c_int64(0)
    assert L.skh_skewtrieigvals_range_d(4, eta.ctypes.data, 0, 0, 0.0, 15.0, 1, vals.ctypes.data, ctypes.byref(m)) == 0
    np.testing.assert_allclose(vals[:m.value], [l1, l2], rtol=1e-12)
    assert L.skh_skewtrieigvals_range_d(4, eta.ctypes.data, 1, 3, 0.0, 0.0, 0, vals.ctypes.data, ctypes.byref(m)) == 0
    np.testing.assert_allclose(vals[:m.value], [l1, l2, -l2], rtol=1e-12)
    # empty selections and argument checks
    assert L.skh_skewtrieigvals_range_d(4, eta.ctypes.data, 0, 0, 20.0, 30.0, 1, vals.ctypes.data, ctypes.byref(m)) == 0
    assert m.value == 0
    assert L.skh_skewtrieigvals_range_d(4, eta.ctypes.data, 3, 2, 0.0, 0.0, 0, vals.ctypes.data, ctypes.byref(m)) == 0
    assert m.value == 0
    assert L.skh_skewtrieigvals_range_d(4, eta.ctypes.data, 0, 3, 0.0, 0.0, 0, vals.ctypes.data, ctypes.byref(m)) == -3
    assert L.skh_skewtrieigvals_range_d(4, eta.ctypes.data, 2, 5, 0.0, 0.0, 0, vals.ctypes.data, ctypes.byref(m)) == -3
    # Python mirror on a SkewHermTridiagonal (host only)
    T = skl.SkewHermTridiagonal(eta)
    np.testing.assert_allclose(skl.eigvals(T, vl=0, vh=15).imag, [l1, l2], rtol=1e-12)
    np.testing.assert_allclose(skl.eigvals(T, irange=(1, 3)).imag, [l1, l2, -l2], rtol=1e-12)
    np.testing.assert_allclose(skl.eigvals(T, irange=range(2, 5)).imag, [l2, -l2, -l1], rtol=1e-12)


def test_skewtrieigvals_range_vs_dense():
    """Index and value ranges against the dense symmetric tridiagonal eigenvalues (numpy)."""
    from skewlinearalgebra import _lib
    L = _lib.lib()
    rng = np.random.default_rng(77)
    for n in (1, 2, 5, 40, 41):
        ev = rng.standard_normal(max(n - 1, 0))
        H = np.diag(ev, -1) + np.diag(ev, 1)
        lam = np.linalg.eigvalsh(H) if n > 0 else np.zeros(0)
        vals = np.zeros(n)
        m = ctypes.c_int64(0)
        il, iu = 1 + n // 4, max(1 + n // 4, (3 * n) // 4)
        assert L.skh_skewtrieigvals_range_d(n, ev.ctypes.data, il, iu, 0.0, 0.0, 0, vals.ctypes.data,
                                            ctypes.byref(m)) == 0
        np.testing.assert_allclose(vals[:m.value], -lam[il - 1:iu], atol=1e-13 * max(1.0, np.abs(lam).max()))
        vl, vh = -0.7, 1.3
        assert L.skh_skewtrieigvals_range_d(n, ev.ctypes.data, 0, 0, vl, vh, 1, vals.ctypes.data,
                                            ctypes.byref(m)) == 0
        sel = lam[(lam > -vh) & (lam <= -vl)]
        np.testing.assert_allclose(vals[:m.value], -sel, atol=1e-13 * max(1.0, np.abs(lam).max()))


def test_skewchol_and_zpfaffian_mirror_arguments():
    """Argument handling of the skew-Cholesky and complex-Pfaffian mirrors that needs no device:
    the skew-Hermitian check and its message (src/cholesky.jl:148), complex input to skewchol
    (defined for real matrices only, :28, :58), the pivot argument, and empty batches."""
    with pytest.raises(skl.ArgumentError, match="skew-Hermitian"):
        skl.skewchol(np.ones((3, 3)))
    with pytest.raises(TypeError):
        skl.skewchol(np.array([[0, 1j], [1j, 0]]))
    with pytest.raises(TypeError):
        skl.skewchol(np.zeros((2, 2)), pivot="rowmaximum")
    assert skl.skewchol_batched([]) == []
    lp, ph = skl.logabspfaffian_z_batched([])
    assert len(lp) == 0 and len(ph) == 0
    # NaN never passes the exact skew-symmetry checks (NaN != NaN), as in Julia
    A = np.array([[0, np.nan], [-np.nan, 0]], dtype=complex)
    with pytest.raises(skl.ArgumentError):
        skl.pfaffian(A)
    assert skl.skewchol.__defaults__[0] is skl.RowMaximum   # the reference's default pivoting


def test_skewfun_argument_errors_before_device():
    """Matrix-function entries reject bad arguments, and log of an odd-sized real skew matrix is a
    DomainError (src/exp.jl:52), before any device work."""
    L = _lib.lib()
    o = _lib.opts()
    A = np.zeros((3, 3), order="F")
    out = np.zeros((3, 3), order="F")
    p, q = A.ctypes.data, out.ctypes.data
    assert L.skh_skewfun_d(-1, None, 1, 0, None, None, ctypes.byref(o)) == -1
    assert L.skh_skewfun_d(3, None, 3, 0, q, None, ctypes.byref(o)) == -2
    assert L.skh_skewfun_d(3, p, 2, 0, q, None, ctypes.byref(o)) == -3
    assert L.skh_skewfun_d(3, p, 3, 12, q, None, ctypes.byref(o)) == -4
    assert L.skh_skewfun_z(3, p, 3, -1, q, None, ctypes.byref(o)) == -4
    assert L.skh_skewfun_d(3, p, 3, 0, None, None, ctypes.byref(o)) == -5
    assert L.skh_skewfun_d(3, p, 3, 5, q, None, ctypes.byref(o)) == -6
    assert L.skh_skewfun_d(3, p, 3, 1, q, None, ctypes.byref(o)) == _lib.ERR_DOMAIN
    assert L.skh_skewfun_tri_d(3, p, 1, q, None) == _lib.ERR_DOMAIN
    assert L.skh_skewfun_d(0, None, 1, 0, None, None, ctypes.byref(o)) == 0
    with pytest.raises(skl.DomainError):
        skl.log(skl.SkewHermitian(np.zeros((5, 5))))
    with pytest.raises(skl.ArgumentError):
        skl.exp(np.ones((3, 3)))
    assert skl.exp(np.zeros((0, 0))).shape == (0, 0)


def test_complex_tridiagonal_entries_before_device():
    """skh_skewtrieigvals_z runs on the host: argument errors, and the eigenvalues of a 2x2 complex
    SkewHermTridiagonal([a], [d1, d2]) (S = [[-d1, |a|], [|a|, -d2]]) against the closed form."""
    L = _lib.lib()
    mu = np.zeros(2)
    assert L.skh_skewtrieigvals_z(-1, None, None, mu.ctypes.data) == -1
    assert L.skh_skewtrieigvals_z(2, None, None, mu.ctypes.data) == -2
    assert L.skh_skewtrieigen_z(2, mu.ctypes.data, None, None, None) == -4
    ev = np.array([3.0 + 4.0j])
    dv = np.array([1.0, -2.0])
    assert L.skh_skewtrieigvals_z(2, ev.ctypes.data, dv.ctypes.data, mu.ctypes.data) == 0
    S = np.array([[-1.0, 5.0], [5.0, 2.0]])
    np.testing.assert_allclose(mu, np.linalg.eigvalsh(S), rtol=1e-15)
    vals = skl.eigvals(skl.SkewHermTridiagonal(ev, dv))
    np.testing.assert_allclose(np.imag(vals), -np.linalg.eigvalsh(S)[::-1], rtol=1e-15)


def test_tuning_table_roundtrip(tmp_path):
    """The plan tuner's JSON table: the batch key depends only on the size histogram, count and
    largest n (not on the order of the matrices), and entries survive a write and a read."""
    import numpy as np
    import skewlinearalgebra as skl
    a = np.array([32, 64, 64, 512, 96], dtype=np.int64)
    b = np.array([64, 512, 32, 96, 64], dtype=np.int64)
    c = np.array([64, 512, 32, 96, 96], dtype=np.int64)
    assert skl._batch_key(a) == skl._batch_key(b)
    assert skl._batch_key(a) != skl._batch_key(c)
    path = str(tmp_path / "tune.json")
    entry = {"config": {"engine": "blocked", "kernel_nmax": 512, "panel_width": 8, "workgroups": 512}, "ms": 1.5}
    skl._write_table(path, {skl._batch_key(a): entry})
    back = skl._read_table(path)
    assert back[skl._batch_key(b)] == entry
    assert skl._read_table(str(tmp_path / "missing.json")) == {}
