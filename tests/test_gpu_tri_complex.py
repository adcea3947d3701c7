"""GPU parity of eigvals / eigen / svd of a complex SkewHermTridiagonal(ev, dvim).

Reference behaviour (src/tridiag.jl): to_symtridiagonal (:558-583) turns iA into the real symmetric
tridiagonal S = SymTridiagonal(-dvim, |i ev|) by diagonal phases Q; eigvals! (:462-468) returns
-i mu for the eigenvalues mu of S, ascending mu reversed (sortby applied first); the (vl, vh) form
(:477-481) returns i mu for mu in (-vh, -vl]; eigen! (:508-515) returns Eigen(-i mu, Q Z).  The
reference's own test (test/runtests.jl, "tridiag.jl" testset, T = ComplexF32, n in {1, 2, 10, 11},
SkewHermTridiagonal(rand(T, n - 1), rand(real(T), n))) checks Q diag(values) Q^H = Matrix(A), sorted
imaginary parts against the dense eigen, svd reconstruction and svdvals against the dense ones; the
same checks run here (tolerance 1e-12 relative, with numpy's dense solvers as the independent
reference), plus the exact values of the documented 4x4 C = SkewHermTridiagonal([3, 4, 5], [6, 7, 8, 9]).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
skl = pytest.importorskip("skewlinearalgebra")


def say(msg):
    print("[case] " + msg, flush=True)


def rand_tri(n, seed, with_dvim=True):
    rng = np.random.default_rng(seed)
    ev = rng.random(n - 1) + 1j * rng.random(n - 1)
    dv = rng.random(n) if with_dvim else None
    return skl.SkewHermTridiagonal(ev, dv)


@pytest.mark.parametrize("n", [1, 2, 10, 11, 300])
@pytest.mark.parametrize("with_dvim", [True, False])
def test_complex_tridiagonal_eigen(n, with_dvim):
    A = rand_tri(n, 60 + n, with_dvim)
    B = A.to_dense()
    say("complex tridiagonal eigen n=%d dvim=%s" % (n, with_dvim))
    E = skl.eigen(A)
    V, lam = E.vectors, E.values
    scale = max(1.0, np.abs(B).max())
    assert np.abs(V @ np.diag(lam) @ V.conj().T - B).max() <= 1e-12 * n * scale
    assert np.abs(V.conj().T @ V - np.eye(n)).max() <= 1e-12 * n
    assert np.all(np.real(lam) == 0)
    mu = -np.imag(lam)
    assert np.all(np.diff(mu) >= 0)                      # eigen!: ascending mu, values -i mu
    dense = np.sort(np.imag(np.linalg.eigvals(B)))
    assert np.abs(np.sort(np.imag(lam)) - dense).max() <= 1e-12 * n * scale
    vals = skl.eigvals(A)                                # -i mu with mu reversed: eigen's order reversed
    assert np.all(np.real(vals) == 0)
    np.testing.assert_allclose(np.imag(vals), np.imag(lam)[::-1], rtol=1e-12, atol=1e-12 * scale)
    assert np.array_equal(skl.eigvecs(A), V)
    S = skl.svd(A)
    assert np.abs(S.U @ np.diag(S.S) @ S.Vt - B).max() <= 1e-12 * n * scale
    np.testing.assert_allclose(skl.svdvals(A), np.linalg.svd(B, compute_uv=False), rtol=1e-11, atol=1e-12)


def test_complex_tridiagonal_sortby_range_and_docs_case():
    A = rand_tri(12, 5)
    B = A.to_dense()
    say("complex tridiagonal sortby / (vl, vh) / 4x4")
    mu = np.sort(np.linalg.eigvalsh(1j * B))
    # sortby = abs: mu sorted by |mu|, reversed, negated
    got = skl.eigvals(A, sortby=abs)
    want = -np.array(sorted(mu, key=abs))[::-1]
    np.testing.assert_allclose(np.imag(got), want, rtol=1e-12, atol=1e-13)
    # (vl, vh): mu in (-vh, -vl], returned as i mu
    vl, vh = -0.5, 1.0
    got = skl.eigvals(A, vl=vl, vh=vh)
    keep = mu[(mu > -vh) & (mu <= -vl)]
    np.testing.assert_allclose(np.imag(got), keep, rtol=1e-12, atol=1e-13)
    with pytest.raises(NotImplementedError):
        skl.eigvals(A, irange=(1, 3))
    # test/runtests.jl: C = SkewHermTridiagonal(complex.([3,4,5]), [6,7,8,9]) == [6im -3 0 0; 3 7im -4 0; ...]
    C = skl.SkewHermTridiagonal(np.array([3, 4, 5], dtype=complex), np.array([6.0, 7, 8, 9]))
    D = np.array([[6j, -3, 0, 0], [3, 7j, -4, 0], [0, 4, 8j, -5], [0, 0, 5, 9j]])
    assert np.array_equal(C.to_dense(), D)
    E = skl.eigen(C)
    assert np.abs(E.vectors @ np.diag(E.values) @ E.vectors.conj().T - D).max() <= 1e-13 * 10
    np.testing.assert_allclose(np.sort(np.imag(E.values)), np.sort(np.imag(np.linalg.eigvals(D))), rtol=1e-13)
