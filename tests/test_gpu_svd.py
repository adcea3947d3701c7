"""svd / svdvals on the skew path (reference src/eigen.jl:110-144, src/tridiag.jl:526-555), checked as
the reference's tests do (test/runtests.jl:201-203, :369-371): U diag(S) Vt reproduces A, and the
singular values match a dense SVD.  Real input goes through the GPU eigvals / eigen."""
import numpy as np
import pytest

from conftest import skew_gauss

pytestmark = pytest.mark.gpu
skl = pytest.importorskip("skewlinearalgebra")


def say(msg):
    print("[case] " + msg, flush=True)


@pytest.mark.parametrize("n", [1, 2, 5, 10, 11, 64, 300])
def test_svd_real(n):
    A = skew_gauss(n, 70 + n)
    say("svd real n=%d" % n)
    S = skl.svd(skl.SkewHermitian(A))
    assert np.linalg.norm(S.U @ np.diag(S.S) @ S.Vt - A) <= 1e-12 * max(1, n) * max(1.0, np.linalg.norm(A))
    assert np.all(S.S >= 0) and np.all(np.diff(S.S) <= 0)
    sv = skl.svdvals(A)
    np.testing.assert_allclose(sv, np.linalg.svd(A, compute_uv=False), rtol=1e-10, atol=1e-12 * max(1.0, np.linalg.norm(A)))


@pytest.mark.parametrize("n", [4, 9, 33])
def test_svd_tridiagonal(n):
    rng = np.random.default_rng(n)
    T = skl.SkewHermTridiagonal(rng.standard_normal(n - 1))
    say("svd tridiagonal n=%d" % n)
    B = T.to_dense()
    S = skl.svd(T)
    assert np.linalg.norm(S.U @ np.diag(S.S) @ S.Vt - B) <= 1e-12 * n * np.linalg.norm(B)
    np.testing.assert_allclose(skl.svdvals(T), np.linalg.svd(B, compute_uv=False), rtol=1e-10, atol=1e-13)


def test_svd_complex_and_readme():
    rng = np.random.default_rng(3)
    M = rng.standard_normal((12, 12)) + 1j * rng.standard_normal((12, 12))
    A = (M - M.conj().T) / 2
    say("svd complex and README")
    S = skl.svd(skl.SkewHermitian(A))
    assert np.linalg.norm(S.U @ np.diag(S.S) @ S.Vt - A) <= 1e-12 * np.linalg.norm(A)
    np.testing.assert_allclose(S.S, np.linalg.svd(A, compute_uv=False), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(skl.svdvals(skl.SkewHermitian(A)), np.linalg.svd(A, compute_uv=False), rtol=1e-12)
    # README 4x4 (test/runtests.jl:27): svdvals = [i l1, i l1, i l2, i l2] in modulus
    R = np.array([[0, 2, -7, 4], [-2, 0, -8, 3], [7, 8, 0, 1], [-4, -3, -1, 0]], dtype=float)
    vals = np.sort(np.abs(np.imag(np.linalg.eigvals(R))))[::-1]
    np.testing.assert_allclose(skl.svdvals(R), vals, rtol=1e-12)


@pytest.mark.parametrize("n", [2, 11, 100])
def test_schur_and_eigvecs(n):
    """test/runtests.jl:135-136: A = Z T Z^H; eigvecs(T) = eigen(T).vectors (src/tridiag.jl:524)."""
    A = skew_gauss(n, 170 + n)
    say("schur n=%d" % n)
    F = skl.schur(skl.SkewHermitian(A))
    assert np.linalg.norm(F.vectors @ F.Schur @ F.vectors.conj().T - A) <= 1e-12 * n * max(1.0, np.linalg.norm(A))
    assert np.array_equal(F.Schur, np.diag(np.diag(F.Schur)))
    T = skl.SkewHermTridiagonal(np.random.default_rng(n).standard_normal(n - 1))
    np.testing.assert_array_equal(skl.eigvecs(T), skl.eigen(T).vectors)


@pytest.mark.parametrize("n", [5, 200])
def test_svd_complex_from_gpu_eigen(n):
    """svd(Hermitian(iA)) of the stdlib (eigen sorted by decreasing |mu|, negative-mu columns of V
    negated) with Vt scaled by -i (src/eigen.jl:138-142): A = U S Vt, U and V unitary, S sorted, and S
    equal to the dense singular values."""
    rng = np.random.default_rng(50 + n)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = (M - M.conj().T) / 2
    say("svd complex n=%d" % n)
    S = skl.svd(A)
    assert np.linalg.norm(S.U @ np.diag(S.S) @ S.Vt - A) <= 1e-12 * n * np.linalg.norm(A)
    assert np.abs(S.U.conj().T @ S.U - np.eye(n)).max() <= 1e-12 * n
    assert np.abs(S.Vt @ S.Vt.conj().T - np.eye(n)).max() <= 1e-12 * n
    assert np.all(np.diff(S.S) <= 0) and np.all(S.S >= 0)
    np.testing.assert_allclose(S.S, np.linalg.svd(A, compute_uv=False), rtol=1e-11, atol=1e-12)
