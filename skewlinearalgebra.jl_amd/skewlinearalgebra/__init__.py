"""Host-side mirror of SkewLinearAlgebra.jl's tridiagonalization API over libskewhip (gfx950).

Names follow the reference (Julia ``f!`` -> Python ``f_``):

  SkewHermitian, SkewHermTridiagonal            src/skewhermitian.jl:3-27, src/tridiag.jl:4-51
  isskewhermitian, isskewsymmetric             src/skewhermitian.jl:84-111
  skewhermitian, skewhermitian_                 src/skewhermitian.jl:35,119-132
  hessenberg, hessenberg_                       src/hessenberg.jl:5-15
  eigvals, eigvals_                             src/eigen.jl:2-16,43-45; src/tridiag.jl:446-460,484-504
  pfaffian, pfaffian_                           src/pfaffian.jl:74-94,139-158
  logabspfaffian, logabspfaffian_               src/pfaffian.jl:167-191,235-257
  logabspfaffian_batched (no reference counterpart: the reference user loops)

The tridiagonalization runs on the GPU (skh_sytrd_d); the O(n^2) implicit QR on the skew
tridiagonal runs on the host in the same library (skh_skewtrieigvals_d), as in the reference.
Errors: ArgumentError (a ValueError) for non-skew input, as the reference's ArgumentError.
"""
import ctypes
import hashlib
import json
import os

import numpy as np

from . import _lib
from ._lib import ENGINE_AUTO, ENGINE_BLOCKED, ENGINE_LARGE, ENGINE_TEAM1, SkhError

__all__ = [
    "SkewHermitian", "SkewHermTridiagonal", "Hessenberg", "HessenbergQ", "ArgumentError",
    "isskewhermitian", "isskewsymmetric", "skewhermitian", "skewhermitian_",
    "hessenberg", "hessenberg_", "eigvals", "eigvals_", "eigen", "eigen_", "Eigen", "pfaffian", "pfaffian_",
    "logabspfaffian", "logabspfaffian_", "logabspfaffian_batched", "logabspfaffian_grad_batched",
    "LogAbsPfaffian", "SkhError", "pfaffian_z_batched", "logabspfaffian_z_batched",
    "JMatrix", "SkewCholesky", "RowMaximum", "NoPivot", "skewchol", "skewchol_", "skewchol_batched", "ldiv", "rdiv",
    "svd", "svdvals", "SVD", "schur", "Schur", "eigvecs", "exp", "log", "cis", "cos", "sin", "sincos", "sinh", "cosh", "tan", "tanh", "cot",
    "coth", "DomainError", "SingularException",
    "ENGINE_AUTO", "ENGINE_LARGE", "ENGINE_TEAM1", "ENGINE_BLOCKED", "version", "LogPfPlan", "LogPfGradPlan",
]


class ArgumentError(ValueError):
    """Julia's ArgumentError: invalid (e.g. non skew-symmetric) input."""


def version():
    return _lib.lib().skh_version().decode()


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


# ----------------------------------------------------------------------------------------
# types
# ----------------------------------------------------------------------------------------
def isskewhermitian(A):
    """A == -A^H exactly (src/skewhermitian.jl:84-94)."""
    if isinstance(A, SkewHermitian):
        return True
    A = np.asarray(A)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ArgumentError("axes of A do not match: %s" % (A.shape,))
    return bool(np.array_equal(A, -A.conj().T))


def isskewsymmetric(A):
    """A == -A^T exactly (src/skewhermitian.jl:101-109)."""
    if isinstance(A, SkewHermitian) and not np.iscomplexobj(A.data):
        return True
    A = np.asarray(A)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ArgumentError("axes of A do not match: %s" % (A.shape,))
    return bool(np.array_equal(A, -A.T))


class SkewHermitian:
    """Zero-copy wrapper of a skew-Hermitian matrix (both triangles stored), src/skewhermitian.jl:3-27."""

    def __init__(self, A):
        if isinstance(A, SkewHermitian):
            self.data = A.data
            return
        A = np.asarray(A)
        if not isskewhermitian(A):
            raise ArgumentError("matrix `A` must be skew-Hermitian (equal `-A'`)")
        self.data = A

    @property
    def shape(self):
        return self.data.shape

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.data, dtype=dtype)

    def copy(self):
        S = SkewHermitian.__new__(SkewHermitian)
        S.data = self.data.copy()
        return S


def skewhermitian_(A):
    """In place (A - A^H)/2 (src/skewhermitian.jl:119-132); returns a SkewHermitian view."""
    n = A.shape[0]
    if A.shape != (n, n):
        raise ArgumentError("matrix is not square")
    for i in range(n):
        A[i, i] = 0 if not np.iscomplexobj(A) else 1j * A[i, i].imag
        a = (A[i, :i] - A[:i, i].conj()) / 2
        A[i, :i] = a
        A[:i, i] = -a.conj()
    return SkewHermitian(A)


def skewhermitian(A):
    """(A - A^H)/2 as a new SkewHermitian (src/skewhermitian.jl:35); a number -> imag(a)."""
    if np.isscalar(A):
        return np.imag(A)
    A = np.asarray(A)
    dt = np.result_type(A.dtype, np.float64) if A.dtype.kind in "iub" else np.result_type(A.dtype, np.float32)
    return skewhermitian_(np.array(A, dtype=dt, order="F", copy=True))


class SkewHermTridiagonal:
    """Skew-Hermitian tridiagonal: subdiagonal ev, optional imaginary diagonal dvim (src/tridiag.jl:4-51)."""

    def __init__(self, ev, dvim=None):
        self.ev = np.asarray(ev)
        self.dvim = None if dvim is None else np.asarray(dvim, dtype=np.real(self.ev).dtype)
        if self.dvim is not None and not np.iscomplexobj(self.ev):
            raise ArgumentError("dvim requires a complex subdiagonal")

    @property
    def shape(self):
        n = self.ev.size + 1
        return (n, n)

    def to_dense(self):
        """Matrix(::SkewHermTridiagonal) (src/tridiag.jl:131-151)."""
        n = self.ev.size + 1
        dt = np.complex128 if (np.iscomplexobj(self.ev) or self.dvim is not None) else np.float64
        M = np.zeros((n, n), dtype=dt)
        idx = np.arange(n - 1)
        M[idx + 1, idx] = self.ev
        M[idx, idx + 1] = -np.conj(self.ev)
        if self.dvim is not None:
            M[np.arange(n), np.arange(n)] = 1j * self.dvim
        return M

    def copy(self):
        return SkewHermTridiagonal(self.ev.copy(), None if self.dvim is None else self.dvim.copy())


class HessenbergQ:
    """Q = H_1 ... H_{n-1} from 'L'-layout factors and tau (LAPACK orgtr('L') convention)."""

    def __init__(self, factors, tau):
        self.factors = factors
        self.tau = tau

    def det(self):
        """det(Q) for real reflectors: (-1)^{count(tau != 0)}."""
        return -1.0 if int(np.count_nonzero(self.tau)) % 2 else 1.0


class Hessenberg:
    """Hessenberg(H::SkewHermTridiagonal, uplo='L', factors, tau) (src/hessenberg.jl:5-13)."""

    def __init__(self, H, factors, tau):
        self.H = H
        self.uplo = "L"
        self.factors = factors
        self.tau = tau

    @property
    def Q(self):
        return HessenbergQ(self.factors, self.tau)


# ----------------------------------------------------------------------------------------
# tridiagonalization
# ----------------------------------------------------------------------------------------
def _sytrd_inplace(F, nb=0, engine=ENGINE_AUTO):
    """F: float64 Fortran-ordered n x n, overwritten by the factors.  Returns (tau, eta, ntau)."""
    n = F.shape[0]
    tau = np.zeros(max(n - 1, 0))
    eta = np.zeros(max(n - 1, 0))
    nz = ctypes.c_int64(0)
    o = _lib.opts(nb, engine)
    rc = _lib.lib().skh_sytrd_d(n, _ptr(F), max(n, 1), _ptr(tau), _ptr(eta), ctypes.byref(nz), ctypes.byref(o))
    _lib.check("skh_sytrd_d", rc)
    return tau, eta, nz.value


def hessenberg_(A, nb=0, engine=ENGINE_AUTO):
    """hessenberg!(A::SkewHermitian{Float64}): A.data becomes the factors (src/hessenberg.jl:5-13)."""
    S = A if isinstance(A, SkewHermitian) else SkewHermitian(A)
    D = S.data
    if np.iscomplexobj(D):
        # complex branch (src/hessenberg.jl:7-8): T = SkewHermTridiagonal(complex(eta), imag(diag(A)))
        if D.dtype != np.complex128 or not D.flags.f_contiguous:
            raise ArgumentError("hessenberg_ needs a ComplexF64 column-major (Fortran-ordered) matrix; use hessenberg")
        n = D.shape[0]
        tau = np.zeros(max(n - 1, 0), dtype=np.complex128)
        eta = np.zeros(max(n - 1, 0))
        dvim = np.zeros(n)
        o = _lib.opts(nb, engine)
        _lib.check("skh_sytrd_z", _lib.lib().skh_sytrd_z(n, _ptr(D), max(n, 1), _ptr(tau), _ptr(eta), _ptr(dvim),
                                                          ctypes.byref(o)))
        return Hessenberg(SkewHermTridiagonal(eta.astype(np.complex128), dvim), D, tau)
    if D.dtype != np.float64 or not D.flags.f_contiguous:
        raise ArgumentError("hessenberg_ needs a Float64 column-major (Fortran-ordered) matrix; use hessenberg")
    tau, eta, _ = _sytrd_inplace(D, nb, engine)
    return Hessenberg(SkewHermTridiagonal(eta), D, tau)


def hessenberg(A, nb=0, engine=ENGINE_AUTO):
    """hessenberg(A::SkewHermitian) = hessenberg!(copyeigtype(A)) (src/hessenberg.jl:15)."""
    D = A.data if isinstance(A, SkewHermitian) else np.asarray(A)
    if not isinstance(A, SkewHermitian) and not isskewhermitian(D):
        raise ArgumentError("matrix `A` must be skew-Hermitian (equal `-A'`)")
    F = np.array(D, dtype=_eigtype(D), order="F", copy=True)
    S = SkewHermitian.__new__(SkewHermitian)
    S.data = F
    return hessenberg_(S, nb, engine)


def _eigtype(D):
    """LinearAlgebra.eigtype as this build computes it: Float64, or ComplexF64 for complex input."""
    return np.complex128 if np.iscomplexobj(D) else np.float64


# ----------------------------------------------------------------------------------------
# eigenvalues (host implicit QR on the tridiagonal, as the reference)
# ----------------------------------------------------------------------------------------
def _tri_eigvals_host(ev):
    ev = np.array(ev, dtype=np.float64, copy=True)
    n = ev.size + 1
    vals = np.zeros(n)
    rc = _lib.lib().skh_skewtrieigvals_d(n, _ptr(ev), _ptr(vals))
    if rc == 2:
        raise RuntimeError("Maximum number of iterations reached, the algorithm didn't converge")
    _lib.check("skh_skewtrieigvals_d", rc)
    return vals


def _tri_is_complex(A):
    return isinstance(A, SkewHermTridiagonal) and (np.iscomplexobj(A.ev) or A.dvim is not None)


def _tri_z(A, vectors):
    """mu (ascending eigenvalues of to_symtridiagonal(A)[1]) and, with vectors, V = Diagonal(Q) Z for a
    complex SkewHermTridiagonal (src/tridiag.jl:462-468, :508-515, :558-583)."""
    ev = np.ascontiguousarray(A.ev, dtype=np.complex128)
    n = ev.size + 1
    dv = None if A.dvim is None else np.ascontiguousarray(A.dvim, dtype=np.float64)
    mu = np.zeros(n)
    L = _lib.lib()
    if not vectors:
        rc = L.skh_skewtrieigvals_z(n, _ptr(ev), _ptr(dv) if dv is not None else None, _ptr(mu))
        V = None
    else:
        V = np.zeros((n, n), dtype=np.complex128, order="F")
        rc = L.skh_skewtrieigen_z(n, _ptr(ev), _ptr(dv) if dv is not None else None, _ptr(mu), _ptr(V))
    if rc == 2:
        raise RuntimeError("Maximum number of iterations reached, the algorithm didn't converge")
    _lib.check("skh_skewtrieigen_z" if vectors else "skh_skewtrieigvals_z", rc)
    return mu, V


def _range_args(irange, vl, vh):
    """Julia's eigvals!(A, irange::UnitRange) / eigvals!(A, vl, vh): irange = (il, iu), 1-based, inclusive."""
    if irange is not None and (vl is not None or vh is not None):
        raise ArgumentError("give either irange or (vl, vh)")
    if irange is not None:
        if isinstance(irange, range):
            il, iu = irange.start, irange.stop - 1
        else:
            il, iu = int(irange[0]), int(irange[-1])
        return il, iu, 0.0, 0.0, 0
    if vl is None or vh is None:
        raise ArgumentError("both vl and vh are needed")
    return 0, 0, float(vl), float(vh), 1


def eigvals_(A, sortby=None, nb=0, engine=ENGINE_AUTO, irange=None, vl=None, vh=None):
    """eigvals!(A::SkewHermitian | SkewHermTridiagonal{<:Real}, sortby) -> complex(0, vals);
    with irange = (il, iu) or (vl, vh): the range forms (src/eigen.jl:8-16,29-41,57-73,
    src/tridiag.jl:452-460,525-537)."""
    if irange is not None or vl is not None or vh is not None:
        il, iu, lo, hi, by_value = _range_args(irange, vl, vh)
        if _tri_is_complex(A):
            # src/tridiag.jl:477-481: the eigenvalues mu of to_symtridiagonal(A)[1] in (-vh, -vl], returned
            # as complex(0, mu) (not negated, as the reference does).  The irange form (:470-475) passes a
            # decreasing range to LAPACK and cannot run; it is not offered here.
            if not by_value:
                raise NotImplementedError("eigvals(::SkewHermTridiagonal{<:Complex}, irange) has no working "
                                          "reference form; use (vl, vh)")
            mu, _ = _tri_z(A, False)
            return mu[(mu > -hi) & (mu <= -lo)] * 1j
        if isinstance(A, SkewHermTridiagonal):
            ev = np.array(A.ev, dtype=np.float64, copy=True)
            n = ev.size + 1
            vals = np.zeros(n)
            m = ctypes.c_int64(0)
            rc = _lib.lib().skh_skewtrieigvals_range_d(n, _ptr(ev), il, iu, lo, hi, by_value, _ptr(vals), ctypes.byref(m))
            _lib.check("skh_skewtrieigvals_range_d", rc)
            return vals[:m.value] * 1j
        S = A if isinstance(A, SkewHermitian) else SkewHermitian(A)
        D = S.data
        n = D.shape[0]
        if np.iscomplexobj(D):
            # src/eigen.jl:36-41: eigenvalues lambda of Hermitian(iA) in (-vh, -vl], ascending, negated.
            # (The reference's complex irange form passes a negative-step range to LAPACK and cannot
            # run; it is not offered here.)
            if not by_value:
                raise NotImplementedError("eigvals(::SkewHermitian{<:Complex}, irange) has no working "
                                          "reference form; use (vl, vh)")
            allv = eigvals_(S, None, nb, engine).imag     # -lambda for ascending lambda
            lam = -allv
            keep = (lam > -hi) & (lam <= -lo)
            return (-lam[keep]) * 1j
        F = np.asfortranarray(D, dtype=np.float64)
        vals = np.zeros(max(n, 1))
        m = ctypes.c_int64(0)
        o = _lib.opts(nb, engine)
        rc = _lib.lib().skh_eigvals_range_d(n, _ptr(F), max(n, 1), il, iu, lo, hi, by_value, _ptr(vals),
                                            ctypes.byref(m), ctypes.byref(o))
        if rc == 2:
            raise RuntimeError("Maximum number of iterations reached, the algorithm didn't converge")
        _lib.check("skh_eigvals_range_d", rc)
        return vals[:m.value] * 1j
    if _tri_is_complex(A):
        # src/tridiag.jl:462-468: mu ascending, sorted by sortby, reversed, negated
        mu, _ = _tri_z(A, False)
        if sortby is not None:
            mu = np.array(sorted(mu, key=sortby))
        return -mu[::-1] * 1j
    if isinstance(A, SkewHermTridiagonal):
        vals = _tri_eigvals_host(A.ev)
    else:
        S = A if isinstance(A, SkewHermitian) else SkewHermitian(A)
        n = S.data.shape[0]
        if np.iscomplexobj(S.data):
            # src/eigen.jl:18-27: -eigvals(Hermitian(iA)), with sortby: sort, reverse, negate
            D = np.asfortranarray(S.data, dtype=np.complex128)
            vals = np.zeros(n)
            o = _lib.opts(nb, engine)
            rc = _lib.lib().skh_eigvals_z(n, _ptr(D), max(n, 1), _ptr(vals), ctypes.byref(o))
            if rc == 2:
                raise RuntimeError("Maximum number of iterations reached, the algorithm didn't converge")
            _lib.check("skh_eigvals_z", rc)
            if sortby is not None:
                lam = np.array(sorted(-vals, key=sortby))
                vals = -lam[::-1]
            return vals * 1j
        if n == 1:
            return np.array([S.data[0, 0]], dtype=np.complex128)   # src/eigen.jl:51
        H = hessenberg_(S, nb, engine)
        vals = _tri_eigvals_host(H.H.ev)
    if sortby is not None:
        vals = np.array(sorted(vals, key=sortby))
    return vals * 1j


def eigvals(A, sortby=None, nb=0, engine=ENGINE_AUTO, irange=None, vl=None, vh=None):
    """eigvals(A) = eigvals!(copyeigtype(A), sortby | irange | vl, vh) (src/eigen.jl:43-45,
    src/tridiag.jl:484-489)."""
    if isinstance(A, SkewHermTridiagonal):
        return eigvals_(A.copy(), sortby, irange=irange, vl=vl, vh=vh)
    D = A.data if isinstance(A, SkewHermitian) else np.asarray(A)
    S = SkewHermitian(np.array(D, dtype=_eigtype(D), order="F", copy=True))
    return eigvals_(S, sortby, nb, engine, irange=irange, vl=vl, vh=vh)


class Eigen:
    """LinearAlgebra.Eigen: values (complex, purely imaginary) and vectors (complex n x n)."""

    def __init__(self, values, vectors):
        self.values = values
        self.vectors = vectors

    def __iter__(self):
        return iter((self.values, self.vectors))


def _tri_eigen_host(ev, variant, gpu=None):
    """skewtrieigen (host QR); from n = 256 (or with gpu=True) its rotations are applied to the
    eigenvector halves on the GPU (same bits)."""
    ev = np.array(ev, dtype=np.float64, copy=True)
    n = ev.size + 1
    vals = np.zeros(n)
    vr = np.zeros((n, n), order="F")
    vi = np.zeros((n, n), order="F")
    name = "skh_skewtrieigen_d_gpu" if (gpu if gpu is not None else n >= 256) else "skh_skewtrieigen_d"
    rc = getattr(_lib.lib(), name)(n, _ptr(ev), _ptr(vals), _ptr(vr), _ptr(vi), variant)
    if rc == 2:
        raise RuntimeError("Maximum number of iterations reached, the algorithm didn't converge")
    _lib.check(name, rc)
    return vals, vr, vi


def _cplx(re, im):
    """re + i im as one complex array, filled in place (no complex temporaries)."""
    out = np.empty(re.shape, dtype=np.complex128, order="F")
    out.real = re
    out.imag = im
    return out


def eigen_(A, nb=0, engine=ENGINE_AUTO):
    """eigen!(A): SkewHermitian{<:Real} -> GPU tridiagonalization + Q + host skewtrieigen_divided!
    + GPU back-transform (src/eigen.jl:75-97); SkewHermTridiagonal{<:Real} -> skewtrieigen_merged!
    on the host (src/tridiag.jl:506)."""
    if _tri_is_complex(A):
        # src/tridiag.jl:508-515: Eigen(mu .* -im, Diagonal(Q) Z), the eigenpairs of the symmetric part on the GPU
        mu, V = _tri_z(A, True)
        return Eigen(-1j * mu, V)
    if isinstance(A, SkewHermTridiagonal):
        vals, vr, vi = _tri_eigen_host(A.ev, 1)
        return Eigen(1j * vals, _cplx(vr, vi))
    D = A.data if isinstance(A, SkewHermitian) else np.asarray(A)
    n = D.shape[0]
    if np.iscomplexobj(D):
        # src/eigen.jl:101-106: Eigen(complex(0, -lambda), eigenvectors of Hermitian(iA))
        F = np.asfortranarray(D, dtype=np.complex128)
        vals = np.zeros(n)
        V = np.zeros((n, n), dtype=np.complex128, order="F")
        o = _lib.opts(nb, engine)
        rc = _lib.lib().skh_eigen_z(n, _ptr(F), max(n, 1), _ptr(vals), _ptr(V), ctypes.byref(o))
        if rc == 2:
            raise RuntimeError("Maximum number of iterations reached, the algorithm didn't converge")
        _lib.check("skh_eigen_z", rc)
        return Eigen(1j * vals, V)
    F = np.asfortranarray(D, dtype=np.float64)
    vals = np.zeros(n)
    vr = np.zeros((n, n), order="F")
    vi = np.zeros((n, n), order="F")
    o = _lib.opts(nb, engine)
    rc = _lib.lib().skh_eigen_d(n, _ptr(F), max(n, 1), _ptr(vals), _ptr(vr), _ptr(vi), ctypes.byref(o))
    if rc == 2:
        raise RuntimeError("Maximum number of iterations reached, the algorithm didn't converge")
    _lib.check("skh_eigen_d", rc)
    return Eigen(1j * vals, _cplx(vr, vi))


def eigen(A, nb=0, engine=ENGINE_AUTO):
    """eigen(A) = eigen!(copyeigtype(A)) (src/eigen.jl:108, src/tridiag.jl:523)."""
    if isinstance(A, SkewHermTridiagonal):
        return eigen_(A.copy())
    D = A.data if isinstance(A, SkewHermitian) else np.asarray(A)
    if not isinstance(A, SkewHermitian) and not isskewhermitian(D):
        raise ArgumentError("matrix `A` must be skew-Hermitian (equal `-A'`)")
    return eigen_(SkewHermitian(np.array(D, dtype=_eigtype(D), order="F", copy=True)), nb, engine)


# ----------------------------------------------------------------------------------------
# singular values (src/eigen.jl:110-144, src/tridiag.jl:526-555): from eigvals / eigen
# ----------------------------------------------------------------------------------------
class SVD:
    """LinearAlgebra.SVD(U, S, Vt): A = U diag(S) Vt."""

    def __init__(self, U, S, Vt):
        self.U = U
        self.S = S
        self.Vt = Vt

    def __iter__(self):
        return iter((self.U, self.S, self.Vt))


def _svd_target(A):
    if isinstance(A, (SkewHermitian, SkewHermTridiagonal)):
        return A
    D = np.asarray(A)
    if not isskewhermitian(D):
        raise ArgumentError("matrix `A` must be skew-Hermitian (equal `-A'`)")
    return SkewHermitian(D)


def svdvals(A, nb=0, engine=ENGINE_AUTO):
    """svdvals(A): |imag(eigvals(A))| sorted in decreasing order (src/eigen.jl:110-114, real;
    src/tridiag.jl:526-530).  A complex SkewHermitian takes svdvals(Hermitian(iA)) (:116), i.e. the
    |eigenvalues| of iA, from the GPU eigenvalues of A."""
    A = _svd_target(A)
    vals = np.abs(np.imag(eigvals(A, nb=nb, engine=engine)))
    return np.sort(vals)[::-1].copy()


def svd(A, nb=0, engine=ENGINE_AUTO):
    """svd(A) from eigen(A) (src/eigen.jl:119-137, src/tridiag.jl:534-553): U = the eigenvectors
    ordered by decreasing |lambda| (a stable sort, as Julia's sortperm), S = |lambda|, V = -i U
    with the columns of negative lambda negated; returns SVD(U, S, V^H).  A complex SkewHermitian
    goes through svd(Hermitian(iA)) with Vt scaled by -i (:138-142); the stdlib's svd of a Hermitian
    matrix is its eigen-decomposition sorted by decreasing |mu| with negative-mu columns of V negated,
    and iA has the eigenvalues mu = -imag(lambda) and the eigenvectors of A (GPU eigen here)."""
    A = _svd_target(A)
    if isinstance(A, SkewHermitian) and np.iscomplexobj(A.data):
        E = eigen(A, nb=nb, engine=engine)
        mu = -np.imag(E.values)
        order = np.argsort(-np.abs(mu), kind="stable")
        mu = mu[order]
        U = E.vectors[:, order]
        V = U.copy()
        neg = mu < 0
        mu[neg] = -mu[neg]
        V[:, neg] = -V[:, neg]
        return SVD(U, mu, V.conj().T * (-1j))
    E = eigen(A, nb=nb, engine=engine)
    vals = np.imag(E.values).copy()
    order = np.argsort(-np.abs(vals), kind="stable")
    vals = vals[order]
    U = E.vectors[:, order]
    V = U * (-1j)
    neg = vals < 0
    vals[neg] = -vals[neg]
    V[:, neg] = -V[:, neg]
    return SVD(U, vals, V.conj().T)


class Schur:
    """LinearAlgebra.Schur(T, Z, values): A = Z T Z^H; for a skew matrix T is diagonal."""

    def __init__(self, T, Z, values):
        self.T = T
        self.Schur = T
        self.Z = Z
        self.vectors = Z
        self.values = values

    def __iter__(self):
        return iter((self.T, self.Z, self.values))


def schur(A, nb=0, engine=ENGINE_AUTO):
    """schur(A::SkewHermitian{<:Real}) (src/skewhermitian.jl:238-243): from eigen(A), T =
    Diagonal(values) as a dense matrix, Z = the eigenvectors."""
    E = eigen(A, nb=nb, engine=engine)
    return Schur(np.diag(E.values), E.vectors, E.values)


def eigvecs(A, nb=0, engine=ENGINE_AUTO):
    """eigvecs(A) = eigen(A).vectors (src/tridiag.jl:524)."""
    return eigen(A, nb=nb, engine=engine).vectors


# ----------------------------------------------------------------------------------------
# matrix functions (src/exp.jl:3-296): eigen-decomposition, products and LU solves on the GPU
# ----------------------------------------------------------------------------------------
class DomainError(ValueError):
    """Julia's DomainError: log of a singular skew matrix (src/exp.jl:52, :68)."""


class SingularException(ValueError):
    """LinearAlgebra.SingularException: the divisor in tan / cot / tanh / coth is singular."""


_FN = {"exp": 0, "log": 1, "cis": 2, "cos": 3, "sin": 4, "sincos": 5, "sinh": 6, "cosh": 7, "tan": 8,
       "tanh": 9, "cot": 10, "coth": 11}
_SKEW_OUT = ("sin", "sinh", "tan", "tanh", "cot", "coth")


def _skewfun(A, name, nb=0, engine=ENGINE_AUTO):
    """f(A) for f = name, A a SkewHermitian, a SkewHermTridiagonal or a skew-Hermitian matrix; the input
    is copied (copyeigtype, src/exp.jl:48).  Dense results: exp / log a plain matrix, cis / cos / cosh the
    Hermitian matrix, sin / sinh / tan / tanh / cot / coth a SkewHermitian, sincos the pair."""
    fn = _FN[name]
    L = _lib.lib()
    if isinstance(A, SkewHermTridiagonal) and not np.iscomplexobj(A.ev) and A.dvim is None:
        ev = np.array(A.ev, dtype=np.float64)
        n = ev.size + 1
        cplx = False

        def call(o1, o2):
            return L.skh_skewfun_tri_d(n, _ptr(ev), fn, o1, o2)
    else:
        if isinstance(A, SkewHermTridiagonal):
            D = A.to_dense()
        elif isinstance(A, SkewHermitian):
            D = A.data
        else:
            D = np.asarray(A)
            if not isskewhermitian(D):
                raise ArgumentError("matrix `A` must be skew-Hermitian (equal `-A'`)")
        cplx = np.iscomplexobj(D)
        F = np.array(D, dtype=np.complex128 if cplx else np.float64, order="F", copy=True)
        n = F.shape[0]
        o = _lib.opts(nb, engine)
        entry = L.skh_skewfun_z if cplx else L.skh_skewfun_d

        def call(o1, o2):
            return entry(n, _ptr(F), max(n, 1), fn, o1, o2, ctypes.byref(o))
    dt = np.complex128 if (cplx or name == "cis") else np.float64
    out1 = np.zeros((n, n), dtype=dt, order="F")
    out2 = np.zeros((n, n), dtype=dt, order="F") if name == "sincos" else None
    rc = call(_ptr(out1), _ptr(out2) if out2 is not None else None)
    if rc == _lib.ERR_DOMAIN:
        raise DomainError("Logarithm of a singular matrix doesn't exist")
    if rc == _lib.ERR_SINGULAR:
        raise SingularException(L.skh_last_error().decode(errors="replace"))
    if rc == _lib.ERR_NOCONV:
        raise RuntimeError("Maximum number of iterations reached, the algorithm didn't converge")
    _lib.check("skh_skewfun(%s)" % name, rc)
    if name == "sincos":
        return _skew_view(out1), out2
    return _skew_view(out1) if name in _SKEW_OUT else out1


def _skew_view(D):
    """SkewHermitian over D without the O(n^2) check, for results the library made exactly skew-Hermitian
    (skewhermitian! on the GPU: D[c, r] = -conj(D[r, c]) bit for bit)."""
    S = SkewHermitian.__new__(SkewHermitian)
    S.data = D
    return S


def exp(A, nb=0, engine=ENGINE_AUTO):
    """exp(A) (src/exp.jl:3-48): real Q1 = Qr C - Qim S, Q2 = Qr S + Qim C, exp(A) = Q1 Qr^T + Q2 Qim^T with
    C, S = cos, sin of the eigenvalues' imaginary parts; complex V exp(Lambda) V^H."""
    return _skewfun(A, "exp", nb, engine)


def log(A, nb=0, engine=ENGINE_AUTO):
    """log(A) (src/exp.jl:50-97): as exp with log|lambda| and sign(lambda) pi/2; a real A of odd size or
    with a zero eigenvalue raises DomainError."""
    return _skewfun(A, "log", nb, engine)


def cis(A, nb=0, engine=ENGINE_AUTO):
    """cis(A) = exp(iA) as a Hermitian matrix (src/exp.jl:99-121, :204)."""
    return _skewfun(A, "cis", nb, engine)


def cos(A, nb=0, engine=ENGINE_AUTO):
    """cos(A), Hermitian (src/exp.jl:123-149, :205)."""
    return _skewfun(A, "cos", nb, engine)


def sin(A, nb=0, engine=ENGINE_AUTO):
    """sin(A) = skewhermitian!(...) (src/exp.jl:151-202, :206)."""
    return _skewfun(A, "sin", nb, engine)


def sincos(A, nb=0, engine=ENGINE_AUTO):
    """(sin(A), cos(A)) (src/exp.jl:208-256)."""
    return _skewfun(A, "sincos", nb, engine)


def sinh(A, nb=0, engine=ENGINE_AUTO):
    """sinh(A) = skewhermitian!(exp(A)) (src/exp.jl:257)."""
    return _skewfun(A, "sinh", nb, engine)


def cosh(A, nb=0, engine=ENGINE_AUTO):
    """cosh(A) = hermitian!(exp(A)), conjugated for complex A (src/exp.jl:258-263)."""
    return _skewfun(A, "cosh", nb, engine)


def tan(A, nb=0, engine=ENGINE_AUTO):
    """tan(A) = skewhermitian!(cos(A) \\ sin(A)) (src/exp.jl:265-268)."""
    return _skewfun(A, "tan", nb, engine)


def tanh(A, nb=0, engine=ENGINE_AUTO):
    """tanh(A) = skewhermitian!((E + I) \\ (E - I)), E = exp(2A) (src/exp.jl:270-273)."""
    return _skewfun(A, "tanh", nb, engine)


def cot(A, nb=0, engine=ENGINE_AUTO):
    """cot(A) = skewhermitian!(sin(A) \\ cos(A)) (src/exp.jl:275-278)."""
    return _skewfun(A, "cot", nb, engine)


def coth(A, nb=0, engine=ENGINE_AUTO):
    """coth(A) = skewhermitian!((E - I) \\ (E + I)), E = exp(2A) (src/exp.jl:280-284)."""
    return _skewfun(A, "coth", nb, engine)


# ----------------------------------------------------------------------------------------
# Pfaffians
# ----------------------------------------------------------------------------------------
def _tri_logpf(ev):
    n = ev.size + 1
    if n % 2 == 1:
        return -np.inf, 0.0
    lp, sg = 0.0, 1.0
    for i in range(0, n - 1, 2):
        lp += np.log(abs(ev[i]))
        sg *= np.sign(-ev[i])
    return float(lp), float(sg)


def logabspfaffian_(A, nb=0, engine=ENGINE_AUTO):
    """(log|Pf(A)|, sign); odd n -> (-Inf, 0).  A (SkewHermitian) becomes the factors."""
    if isinstance(A, SkewHermTridiagonal):
        return _tri_logpf(np.asarray(A.ev, dtype=np.float64))   # src/pfaffian.jl:181-191
    if _is_complex(A):
        return logabspfaffian(A)
    S = A if isinstance(A, SkewHermitian) else None
    if S is None:
        D = np.asarray(A)
        if not isskewhermitian(D):
            raise ArgumentError("Pfaffian requires a skew-Hermitian matrix")
        S = SkewHermitian(D)
    n = S.data.shape[0]
    if n % 2 == 1:
        return -np.inf, 0.0
    if S.data.dtype != np.float64 or not S.data.flags.f_contiguous:
        return logabspfaffian(S, nb, engine)
    tau, eta, nz = _sytrd_inplace(S.data, nb, engine)
    lp, sg = _tri_logpf(eta)
    return lp, (-sg if nz % 2 else sg)


def logabspfaffian(A, nb=0, engine=ENGINE_AUTO):
    """logabspfaffian(A) = logabspfaffian!(copyeigtype(A)) (src/pfaffian.jl:236,246)."""
    if isinstance(A, SkewHermTridiagonal):
        return logabspfaffian_(A)
    if _is_complex(A):
        lp, ph = logabspfaffian_z_batched([_complex_skew(A)])
        return float(lp[0]), complex(ph[0])
    D = A.data if isinstance(A, SkewHermitian) else np.asarray(A)
    if not isinstance(A, SkewHermitian):
        if not isskewhermitian(D):
            raise ArgumentError("Pfaffian requires a skew-Hermitian matrix")
    F = np.array(D, dtype=np.float64, order="F", copy=True)
    n = F.shape[0]
    if n % 2 == 1:
        return -np.inf, 0.0
    lp, sg = ctypes.c_double(), ctypes.c_double()
    o = _lib.opts(nb, engine)
    rc = _lib.lib().skh_logpf_d(n, _ptr(F), max(n, 1), ctypes.byref(lp), ctypes.byref(sg), ctypes.byref(o))
    _lib.check("skh_logpf_d", rc)
    return lp.value, sg.value


def pfaffian_(A, nb=0, engine=ENGINE_AUTO):
    """pfaffian! for real input (src/pfaffian.jl:74-94,155-158): prod(-ev_odd) * sign(det Q)."""
    if isinstance(A, SkewHermTridiagonal):
        ev = np.asarray(A.ev, dtype=np.float64)
        n = ev.size + 1
        if n % 2 == 1:
            return 0.0
        return float(np.prod(-ev[0::2]))
    if _is_complex(A):
        return pfaffian(A)
    S = A if isinstance(A, SkewHermitian) else None
    if S is None:
        D = np.asarray(A)
        if not isskewhermitian(D):
            raise ArgumentError("Pfaffian requires a skew-symmetric matrix")
        S = SkewHermitian(np.array(D, dtype=np.float64, order="F", copy=True))
    n = S.data.shape[0]
    if n % 2 == 1:
        return 0.0
    if S.data.dtype != np.float64 or not S.data.flags.f_contiguous:
        S = SkewHermitian(np.array(S.data, dtype=np.float64, order="F", copy=True))
    tau, eta, nz = _sytrd_inplace(S.data, nb, engine)
    pf = 1.0
    for i in range(0, n - 1, 2):
        pf *= -eta[i]
    return -pf if nz % 2 else pf


def pfaffian(A, nb=0, engine=ENGINE_AUTO):
    """pfaffian(A) = pfaffian!(copyeigtype(A)) (src/pfaffian.jl:148)."""
    if isinstance(A, SkewHermTridiagonal):
        return pfaffian_(A)
    if _is_complex(A):
        return complex(pfaffian_z_batched([_complex_skew(A)])[0])
    D = A.data if isinstance(A, SkewHermitian) else np.asarray(A)
    if not isinstance(A, SkewHermitian) and not isskewhermitian(D):
        raise ArgumentError("Pfaffian requires a skew-symmetric matrix")
    return pfaffian_(SkewHermitian(np.array(D, dtype=np.float64, order="F", copy=True)), nb, engine)


def _is_complex(A):
    if isinstance(A, SkewHermTridiagonal):
        return False
    return np.iscomplexobj(A.data if isinstance(A, SkewHermitian) else A)


def _complex_skew(A):
    """The complex matrix of A after pfaffian!'s check (src/pfaffian.jl:155-158, :259-263):
    A == -transpose(A) exactly, else ArgumentError."""
    D = np.asarray(A.data if isinstance(A, SkewHermitian) else A)
    if not isskewsymmetric(D):
        raise ArgumentError("Pfaffian requires a skew-symmetric matrix")
    return D


def _pack_z(mats):
    mats = [np.asarray(M, dtype=np.complex128) for M in mats]
    for M in mats:
        if M.ndim != 2 or M.shape[0] != M.shape[1]:
            raise ArgumentError("each matrix must be square")
    ns = np.array([M.shape[0] for M in mats], dtype=np.int64)
    offs = np.zeros(len(mats), dtype=np.int64)
    if len(mats) > 1:
        offs[1:] = np.cumsum(ns[:-1] ** 2)
    packed = np.concatenate([M.ravel(order="F") for M in mats]) if mats else np.zeros(1, np.complex128)
    return ns, offs, np.ascontiguousarray(packed)


def pfaffian_z_batched(mats):
    """Pfaffians of complex skew-symmetric matrices (uneven sizes, n <= 8192) in one GPU call:
    Parlett-Reid with pivoting as _pfaffian!(A::AbstractMatrix{<:Complex}) (src/pfaffian.jl:97-137).
    Only the strictly lower triangle is read: the skew-symmetry check is the caller's (pfaffian
    makes it).  Odd n or a zero pivot give 0.  Returns a complex128 array."""
    ns, offs, packed = _pack_z(mats)
    pf = np.zeros(len(ns), dtype=np.complex128)
    if len(ns) == 0:
        return pf
    rc = _lib.lib().skh_pfaffian_z_batched(len(ns), _ptr(ns), _ptr(offs), _ptr(packed), _ptr(pf), None)
    _lib.check("skh_pfaffian_z_batched", rc)
    return pf


def logabspfaffian_z_batched(mats):
    """(log|Pf|, cis(phase)) of complex skew-symmetric matrices in one GPU launch, as
    _logabspfaffian!(A::AbstractMatrix{<:Complex}) (src/pfaffian.jl:193-233); odd n or a zero
    pivot give (-Inf, 0).  Returns (float64 array, complex128 array)."""
    ns, offs, packed = _pack_z(mats)
    lp = np.zeros(len(ns))
    ph = np.zeros(len(ns), dtype=np.complex128)
    if len(ns) == 0:
        return lp, ph
    rc = _lib.lib().skh_logabspfaffian_z_batched(len(ns), _ptr(ns), _ptr(offs), _ptr(packed), _ptr(lp), _ptr(ph),
                                                  None)
    _lib.check("skh_logabspfaffian_z_batched", rc)
    return lp, ph


# ----------------------------------------------------------------------------------------
# skew-Cholesky (src/cholesky.jl, src/jmatrix.jl)
# ----------------------------------------------------------------------------------------
class RowMaximum:
    """Complete pivoting (LinearAlgebra.RowMaximum), the default of skewchol."""


class NoPivot:
    """No pivoting (LinearAlgebra.NoPivot)."""


class JMatrix:
    """JMatrix{T, SGN}(n) (src/jmatrix.jl:9-121): the block-diagonal skew matrix with blocks
    [0 SGN; -SGN 0] (J[i, i+1] = SGN, J[i+1, i] = -SGN for even i, 0-based; a trailing zero row and
    column for odd n).  A structured matrix: products with J permute and negate rows or columns
    instead of multiplying; numpy.asarray gives the dense form."""

    def __init__(self, n, sgn=+1, dtype=np.float64):
        if n < 0:
            raise ArgumentError("size %d must be >= 0" % n)
        if sgn not in (+1, -1):
            raise ArgumentError("SGN parameter must be +-1")
        self.n = int(n)
        self.sgn = int(sgn)
        self.dtype = np.dtype(dtype)

    @property
    def shape(self):
        return (self.n, self.n)

    def __getitem__(self, ij):
        i, j = ij
        if not (0 <= i < self.n and 0 <= j < self.n):
            raise IndexError("index %s out of bounds for a %dx%d JMatrix" % (ij, self.n, self.n))
        if i == j + 1 and j % 2 == 0:
            return self.dtype.type(-self.sgn)
        if j == i + 1 and i % 2 == 0:
            return self.dtype.type(self.sgn)
        return self.dtype.type(0)

    def __array__(self, dtype=None, copy=None):
        M = np.zeros((self.n, self.n), dtype=dtype or self.dtype)
        i = np.arange(0, self.n - 1, 2)
        M[i, i + 1] = self.sgn
        M[i + 1, i] = -self.sgn
        return M

    def __eq__(self, other):
        if isinstance(other, JMatrix):
            return self.n == other.n and (self.sgn == other.sgn or self.n < 2)
        if isinstance(other, SkewHermTridiagonal):
            other = other.to_dense()
        other = np.asarray(other)
        return other.shape == self.shape and bool(np.array_equal(np.asarray(self), other))

    __hash__ = None

    def __neg__(self):
        return JMatrix(self.n, -self.sgn, self.dtype)

    @property
    def T(self):
        """transpose(J) = -J (:88)"""
        return -self

    def transpose(self):
        return -self

    def adjoint(self):
        """J' = -J (:89)"""
        return -self

    def __matmul__(self, A):
        """J * A (:51-67): B[i] = SGN A[i+1], B[i+1] = -SGN A[i] for even i; a zero last row for
        odd n."""
        A = np.asarray(A)
        if A.shape[0] != self.n:
            raise ValueError("DimensionMismatch: J has second dimension %d, A has first dimension %d"
                             % (self.n, A.shape[0]))
        B = np.zeros(A.shape, dtype=np.result_type(self.dtype, A.dtype))
        B[0:self.n - 1:2] = self.sgn * A[1:self.n:2]
        B[1:self.n:2] = -self.sgn * A[0:self.n - 1:2]
        return B

    def __rmatmul__(self, A):
        """A * J (:69-85): B[:, i] = -SGN A[:, i+1], B[:, i+1] = SGN A[:, i] for even i; a zero
        last column for odd n."""
        A = np.asarray(A)
        if A.shape[-1] != self.n:
            raise ValueError("DimensionMismatch: A has second dimension %d, J has first dimension %d"
                             % (A.shape[-1], self.n))
        B = np.zeros(A.shape, dtype=np.result_type(self.dtype, A.dtype))
        B[..., 0:self.n - 1:2] = -self.sgn * A[..., 1:self.n:2]
        B[..., 1:self.n:2] = self.sgn * A[..., 0:self.n - 1:2]
        return B

    def inv(self):
        """inv(J) = -J for even n; odd n is singular (:90-93)"""
        if self.n % 2:
            raise np.linalg.LinAlgError("SingularException(%d)" % self.n)
        return -self

    def tr(self):
        return self.dtype.type(0)

    def det(self):
        return self.dtype.type(self.n % 2 == 0)

    def diag(self, k=0):
        """diag(J, k) (:97-105)"""
        v = np.zeros(max(0, self.n - abs(k)), dtype=self.dtype)
        if k == 1:
            v[0::2] = self.sgn
        elif k == -1:
            v[0::2] = -self.sgn
        return v

    def to_tridiagonal(self):
        """SkewHermTridiagonal(J) (:34-38): ev[0::2] = -SGN"""
        ev = np.zeros(max(self.n - 1, 0), dtype=self.dtype)
        ev[0::2] = -self.sgn
        return SkewHermTridiagonal(ev)


def ldiv(J, A):
    """J \\ A = inv(J) * A (src/jmatrix.jl:86)"""
    return J.inv() @ A


def rdiv(A, J):
    """A / J = A * inv(J) (src/jmatrix.jl:87)"""
    return np.asarray(A) @ J.inv()


class SkewCholesky:
    """SkewCholesky(R, p) (src/cholesky.jl:3-26): R^T J R = A[p, p], J = JMatrix(n, +1);
    p is 0-based."""

    def __init__(self, R, p, steps=None):
        self.R = R
        self.J = JMatrix(R.shape[0], +1)
        self.p = p
        self.steps = steps

    def __iter__(self):
        return iter((self.R, self.J, self.p))


def _pivot_flag(pivot):
    if pivot is RowMaximum or isinstance(pivot, RowMaximum):
        return 1
    if pivot is NoPivot or isinstance(pivot, NoPivot):
        return 0
    raise TypeError("pivot must be RowMaximum() or NoPivot()")


def _real_skew(A, check):
    D = A.data if isinstance(A, SkewHermitian) else np.asarray(A)
    if np.iscomplexobj(D):
        raise TypeError("skewchol is defined for real skew-symmetric matrices (src/cholesky.jl:28, :58)")
    if check and not isinstance(A, SkewHermitian) and not isskewhermitian(D):
        raise ArgumentError("Pfaffian requires a skew-Hermitian matrix")
    return D


def skewchol_batched(mats, pivot=RowMaximum):
    """skewchol of each matrix (uneven sizes, n <= 4096) in one GPU launch; returns a list of
    SkewCholesky.  Each matrix must be exactly skew-symmetric (checked)."""
    flag = _pivot_flag(pivot)
    mats = [np.asarray(_real_skew(M, True), dtype=np.float64) for M in mats]
    ns = np.array([M.shape[0] for M in mats], dtype=np.int64)
    offs = np.zeros(len(mats), dtype=np.int64)
    if len(mats) > 1:
        offs[1:] = np.cumsum(ns[:-1] ** 2)
    packed = np.concatenate([M.ravel(order="F") for M in mats]) if mats else np.zeros(1)
    R = np.zeros_like(packed)
    p = np.zeros(max(1, int(ns.sum())), dtype=np.int64)
    steps = np.zeros(max(1, len(mats)), dtype=np.int64)
    if mats:
        rc = _lib.lib().skh_skewchol_d_batched(len(mats), _ptr(ns), _ptr(offs), _ptr(packed), _ptr(R), _ptr(p),
                                               _ptr(steps), flag, None)
        _lib.check("skh_skewchol_d_batched", rc)
    out, pacc = [], 0
    for k, n in enumerate(ns):
        n = int(n)
        Rk = R[offs[k]:offs[k] + n * n].reshape((n, n), order="F").copy()
        out.append(SkewCholesky(Rk, p[pacc:pacc + n].copy(), int(steps[k])))
        pacc += n
    return out


def skewchol(A, pivot=RowMaximum):
    """skewchol(A, pivot) (src/cholesky.jl:134-151): R^T J R = A[p, p] with RowMaximum (the
    default), R^T J R = A with NoPivot.  A (not modified) must be exactly skew-symmetric."""
    return skewchol_batched([_real_skew(A, True)], pivot)[0]


def skewchol_(A, pivot=RowMaximum):
    """skewchol!(A, pivot) (src/cholesky.jl:109-127): as skewchol, with A (a float64 array or a
    SkewHermitian's data) overwritten by R, which the result then shares."""
    D = _real_skew(A, True)
    C = skewchol(D, pivot)
    if isinstance(D, np.ndarray) and D.dtype == np.float64:
        D[...] = C.R
        C.R = D
    return C


def logabspfaffian_batched(mats, nb=0, engine=ENGINE_AUTO):
    """log|Pf| and sign of each matrix in `mats` (uneven sizes) in one persistent GPU launch.

    No reference counterpart (the reference user loops logabspfaffian, src/pfaffian.jl:236).
    Returns (logpf, sign) float64 arrays."""
    mats = [np.asarray(M, dtype=np.float64) for M in mats]
    for M in mats:
        if M.ndim != 2 or M.shape[0] != M.shape[1]:
            raise ArgumentError("each matrix must be square")
    ns = np.array([M.shape[0] for M in mats], dtype=np.int64)
    offs = np.zeros(len(mats), dtype=np.int64)
    if len(mats) > 1:
        offs[1:] = np.cumsum(ns[:-1] ** 2)
    packed = np.concatenate([M.ravel(order="F") for M in mats]) if mats else np.zeros(1)
    lp = np.zeros(len(mats))
    sg = np.zeros(len(mats))
    o = _lib.opts(nb, engine)
    rc = _lib.lib().skh_logpf_batched_d(len(mats), _ptr(ns), _ptr(offs), _ptr(packed), _ptr(lp), _ptr(sg),
                                        ctypes.byref(o))
    _lib.check("skh_logpf_batched_d", rc)
    return lp, sg


def _pack(mats):
    mats = [np.asarray(M, dtype=np.float64) for M in mats]
    for M in mats:
        if M.ndim != 2 or M.shape[0] != M.shape[1]:
            raise ArgumentError("each matrix must be square")
        if not isskewhermitian(M):
            raise ArgumentError("Pfaffian requires a skew-Hermitian matrix")
    ns = np.array([M.shape[0] for M in mats], dtype=np.int64)
    offs = np.zeros(len(mats), dtype=np.int64)
    if len(mats) > 1:
        offs[1:] = np.cumsum(ns[:-1] ** 2)
    packed = np.concatenate([M.ravel(order="F") for M in mats]) if mats else np.zeros(1)
    return mats, ns, offs, packed


def logabspfaffian_grad_batched(mats, nb=0, engine=ENGINE_AUTO):
    """(logpf, sign, grads) with grads[b] = d log|Pf(A_b)| / dA_b = -1/2 inv(A_b) (entrywise).

    The reverse rule a ChainRulesCore rrule of logabspfaffian would return for a unit seed; the
    reference has no AD code, its CPU counterpart is inv(A::SkewHermitian) (src/skewhermitian.jl:233)."""
    mats, ns, offs, packed = _pack(mats)
    lp = np.zeros(len(mats))
    sg = np.zeros(len(mats))
    g = np.zeros_like(packed)
    o = _lib.opts(nb, engine)
    rc = _lib.lib().skh_logpf_grad_batched_d(len(mats), _ptr(ns), _ptr(offs), _ptr(packed), _ptr(lp), _ptr(sg),
                                             _ptr(g), ctypes.byref(o))
    _lib.check("skh_logpf_grad_batched_d", rc)
    grads = [g[offs[k]:offs[k] + ns[k] ** 2].reshape(ns[k], ns[k], order="F") for k in range(len(mats))]
    return lp, sg, grads


def _device_batch_call(fn_name, A, want_grad):
    """A: torch float64 CUDA tensor (B, n, n) of skew matrices (row-major storage of A^T is
    column-major storage of A, and A^T = -A, so the grad/factor layouts are handled below)."""
    import torch
    if A.dim() != 3 or A.shape[1] != A.shape[2] or A.dtype != torch.float64 or not A.is_cuda:
        raise ArgumentError("expected a float64 CUDA tensor of shape (B, n, n)")
    B, n = A.shape[0], A.shape[1]
    # column-major view of each matrix = row-major storage of its transpose
    At = A.transpose(1, 2).contiguous()
    ns = np.full(B, n, dtype=np.int64)
    offs = (np.arange(B, dtype=np.int64) * n * n)
    if (n * n) % 2 == 1 and B > 1:
        raise ArgumentError("device batches need an even n*n stride; pad or call per matrix")
    lp = torch.empty(B, dtype=torch.float64, device=A.device)
    sg = torch.empty(B, dtype=torch.float64, device=A.device)
    st = torch.cuda.current_stream(A.device).cuda_stream
    o = _lib.opts()
    L = _lib.lib()
    if want_grad:
        gt = torch.empty_like(At)
        rc = L.skh_logpf_grad_batched_d_device(B, _ptr(ns), _ptr(offs), At.data_ptr(), lp.data_ptr(), sg.data_ptr(),
                                               gt.data_ptr(), st, ctypes.byref(o))
        _lib.check("skh_logpf_grad_batched_d_device", rc)
        return lp, sg, gt.transpose(1, 2)
    rc = L.skh_logpf_batched_d_device(B, _ptr(ns), _ptr(offs), At.data_ptr(), lp.data_ptr(), sg.data_ptr(), st,
                                      ctypes.byref(o))
    _lib.check("skh_logpf_batched_d_device", rc)
    return lp, sg, None


def _batch_key(ns):
    """Tuning-table key of a batch layout: its size histogram (hashed), count and largest n."""
    ns = np.asarray(ns, dtype=np.int64)
    vals, cnts = np.unique(ns, return_counts=True)
    h = hashlib.sha1(np.concatenate([vals, cnts]).astype(np.int64).tobytes()).hexdigest()[:16]
    return "%s/%d/%d" % (h, len(ns), int(ns.max()) if len(ns) else 0)


def _read_table(path):
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("plans", {}) if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _write_table(path, entries):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump({"format": "skewhip tuning table 1", "library": version(), "plans": entries}, f, indent=1,
                  sort_keys=True)
    os.replace(tmp, path)


class LogPfPlan:
    """A batched log|Pf| plan over device-resident matrices (C ABI skh_logpf_batched_plan_*).

    No reference counterpart (the reference user loops logabspfaffian, src/pfaffian.jl:236).
    `ns`, `offsets`: host int64 arrays (column-major matrices at even element offsets of one
    float64 device buffer).  `run(dA, dlogpf, dsgn, stream)` enqueues one evaluation without
    host synchronisation, allocation or host copies (capturable in a HIP graph, e.g.
    torch.cuda.graph); `tune(...)` times the kernel configurations and keeps the fastest;
    `status(stream)` waits and raises if the watchdog stopped a launch."""

    def __init__(self, ns, offsets):
        self._ns = np.ascontiguousarray(ns, dtype=np.int64)
        self._offs = np.ascontiguousarray(offsets, dtype=np.int64)
        if self._ns.shape != self._offs.shape or self._ns.ndim != 1:
            raise ArgumentError("ns and offsets must be 1-D arrays of the same length")
        h = ctypes.c_void_p()
        o = _lib.opts()
        _lib.check("skh_logpf_batched_plan_create",
                   _lib.lib().skh_logpf_batched_plan_create(len(self._ns), _ptr(self._ns), _ptr(self._offs),
                                                            ctypes.byref(o), ctypes.byref(h)))
        self._h = h

    def run(self, dA, dlogpf, dsgn, stream=None):
        _lib.check("skh_logpf_batched_plan_exec",
                   _lib.lib().skh_logpf_batched_plan_exec(self._h, dA, dlogpf, dsgn, stream))

    def tune(self, dA, dlogpf, dsgn, stream=None, table=None):
        """Time the launch configurations (engine, kernel NMAX, panel width, work-groups) on these
        inputs and keep the fastest; returns its time in ms.  With a tuning table (a JSON file: the
        `table` path, else $SKEWHIP_TUNE_TABLE), a batch layout already in the table takes its stored
        configuration without timing, and a new one is timed and added."""
        path = table if table is not None else os.environ.get("SKEWHIP_TUNE_TABLE")
        key = _batch_key(self._ns)
        entries = _read_table(path) if path else {}
        if key in entries:
            self.set_config(entries[key]["config"])
            return float(entries[key]["ms"])
        best = ctypes.c_double(0.0)
        _lib.check("skh_logpf_batched_plan_tune",
                   _lib.lib().skh_logpf_batched_plan_tune(self._h, dA, dlogpf, dsgn, stream, ctypes.byref(best)))
        if path:
            entries[key] = {"config": self.config(), "ms": best.value, "matrices": int(len(self._ns)),
                            "nmax": int(self._ns.max()) if len(self._ns) else 0}
            _write_table(path, entries)
        return best.value

    def config(self):
        """The launch configuration: engine ("blocked" or "unblocked"), kernel NMAX, panel width,
        work-groups."""
        out = np.zeros(4, dtype=np.int64)
        _lib.check("skh_plan_config", _lib.lib().skh_plan_config(self._h, _ptr(out)))
        return {"engine": "blocked" if int(out[0]) == ENGINE_BLOCKED else "unblocked", "kernel_nmax": int(out[1]),
                "panel_width": int(out[2]), "workgroups": int(out[3])}

    def set_config(self, cfg):
        v = np.array([ENGINE_BLOCKED if cfg["engine"] == "blocked" else ENGINE_TEAM1, cfg["kernel_nmax"],
                      cfg["panel_width"], cfg["workgroups"]], dtype=np.int64)
        _lib.check("skh_plan_set_config", _lib.lib().skh_plan_set_config(self._h, _ptr(v)))

    def status(self, stream=None):
        _lib.check("skh_plan_status", _lib.lib().skh_plan_status(self._h, stream))

    def info(self):
        out = np.zeros(4, dtype=np.int64)
        _lib.check("skh_plan_info", _lib.lib().skh_plan_info(self._h, _ptr(out)))
        return {"workgroups": int(out[0]), "kernel_nmax": int(out[1]), "tuned": bool(out[2]),
                "matrices": int(out[3])}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().skh_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LogPfGradPlan(LogPfPlan):
    """A batched log|Pf| + gradient plan over device-resident matrices (C ABI
    skh_logpf_grad_batched_plan_*): `run(dA, dlogpf, dsgn, dgrad, stream)` enqueues the forward
    factorization and the hand-written backward pass (-1/2 A^{-1} into dgrad, the input's packing)
    without host synchronisation, allocation, host copies or memset, so it can be captured in a HIP
    graph.  No reference counterpart (the reference has no AD code; its CPU route to the gradient is
    inv(A::SkewHermitian), src/skewhermitian.jl:233)."""

    def __init__(self, ns, offsets):
        self._ns = np.ascontiguousarray(ns, dtype=np.int64)
        self._offs = np.ascontiguousarray(offsets, dtype=np.int64)
        if self._ns.shape != self._offs.shape or self._ns.ndim != 1:
            raise ArgumentError("ns and offsets must be 1-D arrays of the same length")
        h = ctypes.c_void_p()
        o = _lib.opts()
        _lib.check("skh_logpf_grad_batched_plan_create",
                   _lib.lib().skh_logpf_grad_batched_plan_create(len(self._ns), _ptr(self._ns), _ptr(self._offs),
                                                                 ctypes.byref(o), ctypes.byref(h)))
        self._h = h

    def run(self, dA, dlogpf, dsgn, dgrad, stream=None):
        _lib.check("skh_logpf_grad_batched_plan_exec",
                   _lib.lib().skh_logpf_grad_batched_plan_exec(self._h, dA, dlogpf, dsgn, dgrad, stream))

    def tune(self, *a, **k):
        raise NotImplementedError("gradient plans are not tuned")


def _make_autograd():
    try:
        import torch
    except ImportError:   # pragma: no cover
        return None

    class LogAbsPfaffian(torch.autograd.Function):
        """torch.autograd counterpart of a ChainRulesCore rrule for logabspfaffian.

        forward(A) -> (logpf, sign) for a (B, n, n) float64 CUDA batch; backward returns
        dlogpf[:, None, None] * (-1/2 A^{-1}), computed on the GPU from the stored reflectors."""

        @staticmethod
        def forward(ctx, A):
            lp, sg, g = _device_batch_call("grad", A, True)
            ctx.save_for_backward(g)
            ctx.mark_non_differentiable(sg)
            return lp, sg

        @staticmethod
        def backward(ctx, dlp, dsg):
            (g,) = ctx.saved_tensors
            return dlp[:, None, None] * g

    return LogAbsPfaffian


LogAbsPfaffian = _make_autograd()
