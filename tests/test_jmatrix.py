"""JMatrix, the J of SkewCholesky (reference src/jmatrix.jl), host-only: the reference's own
testset (test/runtests.jl:456-486) for element types Int32, Float32, Float64, n in 1, 2, 10, 11 and
SGN = +1, -1, plus the defining identity J * J = -I for even n."""
import numpy as np
import pytest

skl = pytest.importorskip("skewlinearalgebra")


@pytest.mark.parametrize("sgn", [1, -1])
@pytest.mark.parametrize("n", [1, 2, 10, 11])
@pytest.mark.parametrize("T", [np.int32, np.float32, np.float64])
def test_jmatrix_reference_testset(T, n, sgn):
    rng = np.random.default_rng(n + 100 * (sgn + 1))
    A = rng.random((n, n)).astype(T) if np.dtype(T).kind == "f" else rng.integers(0, 10, (n, n)).astype(T)
    J = skl.JMatrix(n, sgn, T)
    vec = np.zeros(n - 1, dtype=T)
    vec[0::2] = -sgn
    Jtest = skl.SkewHermTridiagonal(vec)
    assert J.shape == (n, n)
    Jd = np.asarray(J)
    assert J == Jd and J == Jtest.to_dense() and J == J.to_tridiagonal()
    assert np.array_equal(J.to_tridiagonal().ev, vec)
    np.testing.assert_allclose(A @ J, A @ Jtest.to_dense(), rtol=1e-6)
    np.testing.assert_allclose(J @ A, Jtest.to_dense() @ A, rtol=1e-6)
    assert -J == -Jd
    assert J.T == -Jd and J.adjoint() == -Jd
    if n % 2 == 0:
        assert J.inv() == -Jd
        np.testing.assert_allclose(skl.ldiv(J, A), np.linalg.solve(Jd.astype(np.float64), A), rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(skl.rdiv(A, J), A @ np.linalg.inv(Jd.astype(np.float64)), rtol=1e-6, atol=1e-6)
        np.testing.assert_array_equal(J @ (J @ np.eye(n)), -np.eye(n))
    else:
        with pytest.raises(np.linalg.LinAlgError):
            J.inv()
    for k in list(range(-4, 5)) + [n, n + 1]:
        np.testing.assert_array_equal(J.diag(k), np.diagonal(Jd, k) if abs(k) < n else np.zeros(0))
    assert J.tr() == 0
    assert bool(J.det()) == (n % 2 == 0)
    assert abs(float(J.det()) - np.linalg.det(Jd.astype(np.float64))) < 1e-12
    for i in range(n):
        for j in range(n):
            assert J[i, j] == Jd[i, j]


def test_jmatrix_errors():
    with pytest.raises(skl.ArgumentError):
        skl.JMatrix(-1)
    with pytest.raises(skl.ArgumentError):
        skl.JMatrix(3, 2)
    with pytest.raises(ValueError):
        skl.JMatrix(4) @ np.ones((3, 2))
    with pytest.raises(IndexError):
        skl.JMatrix(2)[2, 0]
