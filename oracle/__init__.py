"""CPU oracle for the skew tridiagonalization path -- TEST INFRASTRUCTURE ONLY.

This package wraps ``liboracle.so`` (built from ``skew_oracle.c`` by the Makefile here),
a C restatement of SkewLinearAlgebra.jl's CPU algorithm:

* ``skewblockedhess``   -- src/hessenberg.jl:144-174 (with skewlatrd! :92-131, skewhess! :70-90,
                            householder! :26-44, setnb :133-142)
* ``skewtrieigvals``     -- src/skeweigen.jl:105-156 (+ reducetozero :11-47,
                            implicitstep_novec :73-103, backup :463-468)
* ``skewtrieigen``       -- src/skeweigen.jl:195-320 (merged) and :322-457 (divided)
* ``logabspfaffian`` / ``pfaffian`` -- src/pfaffian.jl:74-94, 167-191 (real, via the
                            tridiagonalization) and :95-137, :193-233 (complex skew-symmetric,
                            Parlett-Reid with pivoting)

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
use it, and only as the checker / the CPU baseline.  The shipped library never calls it.
"""
import ctypes
import glob
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64


def build():
    """Compile liboracle.so in place (gcc via the Makefile here)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def openblas_path():
    """numpy's bundled ILP64 OpenBLAS (symbols carry the scipy_ prefix), or None."""
    import numpy
    libdir = os.path.join(os.path.dirname(os.path.dirname(numpy.__file__)), "numpy.libs")
    hits = sorted(glob.glob(os.path.join(libdir, "libscipy_openblas64_*.so")))
    return hits[0] if hits else None


def lib(use_blas=True):
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_blas_load.argtypes = [ctypes.c_char_p]
        L.orc_blas_set_threads.argtypes = [ctypes.c_int]
        L.orc_setnb.argtypes = [_i64]
        L.orc_setnb.restype = _i64
        L.orc_skewblockedhess_d.argtypes = [_i64, _dp, _i64, _dp, _dp, _i64]
        L.orc_skewblockedhess_z.argtypes = [_i64, ctypes.c_void_p, _i64, ctypes.c_void_p, _dp, _i64]
        L.orc_skewtrieigvals_d.argtypes = [_i64, _dp, _dp]
        L.orc_skewtrieigen_d.argtypes = [_i64, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_logabspfaffian_d.argtypes = [_i64, _dp, _i64, _dp, _dp]
        L.orc_logabspfaffian_tri_d.argtypes = [_i64, _dp, _dp, _dp]
        L.orc_pfaffian_d.argtypes = [_i64, _dp, _i64]
        L.orc_pfaffian_d.restype = ctypes.c_double
        L.orc_pfaffian_z_out.argtypes = [_i64, _dp, _i64, _dp]
        L.orc_logabspfaffian_z.argtypes = [_i64, _dp, _i64, _dp, _dp]
        L.orc_skewchol_d.argtypes = [_i64, _dp, _i64, ctypes.c_int, _dp]
        L.orc_skewchol_d.restype = _i64
        _LIB = L
        if use_blas:
            p = openblas_path()
            if p is not None:
                L.orc_blas_load(p.encode())
    return _LIB


def set_threads(t):
    lib().orc_blas_set_threads(int(t))


def _ptr(a):
    return a.ctypes.data_as(_dp)


def setnb(n):
    return int(lib().orc_setnb(int(n)))


def skewblockedhess(A, nb=0):
    """Reference-style tridiagonalization.  Returns (factors, tau, eta); A is not modified."""
    A = np.asarray(A)
    n = A.shape[0]
    if np.iscomplexobj(A):
        F = np.array(A, dtype=np.complex128, order="F", copy=True)
        tau = np.zeros(max(n - 1, 0), dtype=np.complex128)
        eta = np.zeros(max(n - 1, 0), dtype=np.float64)
        lib().orc_skewblockedhess_z(n, F.ctypes.data_as(ctypes.c_void_p), n,
                                    tau.ctypes.data_as(ctypes.c_void_p), _ptr(eta), int(nb))
        return F, tau, eta
    F = np.array(A, dtype=np.float64, order="F", copy=True)
    tau = np.zeros(max(n - 1, 0))
    eta = np.zeros(max(n - 1, 0))
    lib().orc_skewblockedhess_d(n, _ptr(F), n, _ptr(tau), _ptr(eta), int(nb))
    return F, tau, eta


def skewtrieigvals(ev):
    """Imaginary parts of the eigenvalues of SkewHermTridiagonal(ev), reference order."""
    ev = np.array(ev, dtype=np.float64, copy=True)
    N = ev.size + 1
    vals = np.zeros(N)
    rc = lib().orc_skewtrieigvals_d(N, _ptr(ev), _ptr(vals))
    if rc != 0:
        raise RuntimeError("Maximum number of iterations reached, the algorithm didn't converge")
    return vals


def skewtrieigen(ev, variant="divided"):
    """(vals_imag, vectors) for SkewHermTridiagonal(ev); variant 'divided' or 'merged'."""
    ev = np.array(ev, dtype=np.float64, copy=True)
    N = ev.size + 1
    vals = np.zeros(N)
    vr = np.zeros((N, N), order="F")
    vi = np.zeros((N, N), order="F")
    rc = lib().orc_skewtrieigen_d(N, _ptr(ev), _ptr(vals), _ptr(vr), _ptr(vi),
                                  0 if variant == "divided" else 1)
    if rc != 0:
        raise RuntimeError("skewtrieigen failed (rc=%d)" % rc)
    return vals, vr + 1j * vi


def logabspfaffian(A):
    if np.iscomplexobj(A):
        return logabspfaffian_z(A)
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    n = A.shape[0]
    lp = ctypes.c_double()
    sg = ctypes.c_double()
    lib().orc_logabspfaffian_d(n, _ptr(A), n, ctypes.byref(lp), ctypes.byref(sg))
    return lp.value, sg.value


def logabspfaffian_tri(ev):
    ev = np.ascontiguousarray(ev, dtype=np.float64)
    lp = ctypes.c_double()
    sg = ctypes.c_double()
    lib().orc_logabspfaffian_tri_d(ev.size + 1, _ptr(ev), ctypes.byref(lp), ctypes.byref(sg))
    return lp.value, sg.value


def pfaffian(A):
    if np.iscomplexobj(A):
        return pfaffian_z(A)
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    return lib().orc_pfaffian_d(A.shape[0], _ptr(A), A.shape[0])


def pfaffian_z(A):
    """Complex skew-symmetric Pfaffian (Parlett-Reid, src/pfaffian.jl:97-137)."""
    A = np.array(A, dtype=np.complex128, order="F", copy=True)
    out = np.zeros(2)
    lib().orc_pfaffian_z_out(A.shape[0], _ptr(A), A.shape[0], _ptr(out))
    return complex(out[0], out[1])


def logabspfaffian_z(A):
    """(log|Pf|, cis(phase)) of a complex skew-symmetric matrix (src/pfaffian.jl:193-233)."""
    A = np.array(A, dtype=np.complex128, order="F", copy=True)
    lp = ctypes.c_double()
    ph = np.zeros(2)
    lib().orc_logabspfaffian_z(A.shape[0], _ptr(A), A.shape[0], ctypes.byref(lp), _ptr(ph))
    return lp.value, complex(ph[0], ph[1])


def skewchol(A, pivot=True):
    """Skew-Cholesky R^T J R = A[p, p] (src/cholesky.jl:28-107; RowMaximum if pivot, else
    NoPivot with p = 0..n-1).  Returns (R upper triangular, p 0-based, steps completed)."""
    B = np.array(A, dtype=np.float64, order="F", copy=True)
    n = B.shape[0]
    p = np.zeros(max(n, 1), dtype=np.int64)
    steps = lib().orc_skewchol_d(n, _ptr(B), max(n, 1), 1 if pivot else 0, _ptr(p))
    return np.triu(B), p[:n], int(steps)


def jmatrix(n):
    """JMatrix{+1}(n) dense (src/jmatrix.jl:23-30): J[i, i+1] = 1, J[i+1, i] = -1 for even i."""
    J = np.zeros((n, n))
    for i in range(0, n - 1, 2):
        J[i, i + 1] = 1.0
        J[i + 1, i] = -1.0
    return J


def form_q(F, tau):
    """Q = H_1 ... H_{n-1} from 'L'-layout factors (LAPACK orgtr('L') convention).

    Test helper: reflector i has v[i+1] = 1 (implicit; the stored entry is not read) and
    v[i+2:] = F[i+2:, i]; H_i = I - tau_i v v^H.
    """
    n = F.shape[0]
    dt = np.result_type(F.dtype, tau.dtype)
    Q = np.eye(n, dtype=dt)
    for i in range(n - 2, -1, -1):
        t = tau[i]
        if t == 0:
            continue
        v = np.zeros(n, dtype=dt)
        v[i + 1] = 1.0
        v[i + 2:] = F[i + 2:, i]
        # Q <- H_i Q  (apply from the left, building the product right to left)
        Q -= t * np.outer(v, v.conj() @ Q)
    return Q


# ----------------------------------------------------------------------------------------
# Matrix functions (src/exp.jl:3-296), restated in numpy on the oracle's own eigen-decomposition
# ----------------------------------------------------------------------------------------
def skeweigen(A):
    """skeweigen! (src/eigen.jl:75-97) of a real skew A: (vals_imag, Qr, Qim) with the eigenvectors
    Q Qr + i Q Qim from the oracle's tridiagonalization, its Q and skewtrieigen_divided!."""
    A = np.array(A, dtype=np.float64)
    n = A.shape[0]
    if n == 1:
        return np.zeros(1), np.ones((1, 1)), np.zeros((1, 1))
    F, tau, eta = skewblockedhess(A)
    Q = form_q(F, tau)
    vals, V = skewtrieigen(eta, "divided")
    return vals, Q @ V.real, Q @ V.imag


def _skewhermitian_(X):
    """skewhermitian!(X) (src/skewhermitian.jl:119-131): lower a = (X[i,j] - conj(X[j,i])) / 2, upper
    -conj(a), diagonal 0 (real) or i imag (complex)."""
    X = np.array(X)
    L = np.tril(X, -1)
    U = np.triu(X, -0)
    a = (L - np.conj(U).T) / 2
    a = np.tril(a, -1)
    out = a - np.conj(a).T
    d = np.diagonal(X)
    out[np.diag_indices_from(out)] = (1j * d.imag) if np.iscomplexobj(X) else 0.0
    return out


def _hermitian_view(X):
    """Matrix(Hermitian(X)) with uplo 'U': the upper triangle mirrored, the diagonal made real."""
    U = np.triu(X, 1)
    out = U + np.conj(U).T
    out = out.astype(X.dtype)
    out[np.diag_indices_from(out)] = np.diagonal(X).real
    return out


def _cosh_part(X):
    """cosh's hermitian!(exp(A)) (+ conj for complex) as a dense matrix (src/exp.jl:258-263, :286-296):
    upper a = (X[j,i] + conj(X[i,j])) / 2, lower conj(a), real diagonal."""
    U = np.triu(X, 1)
    L = np.tril(X, -1)
    a = (U + np.conj(L).T) / 2
    out = a + np.conj(a).T
    out = out.astype(X.dtype)
    out[np.diag_indices_from(out)] = np.diagonal(X).real
    return out


def _fun_from_eigs_real(vals, Qr, Qim, fn, n):
    if fn == "log":
        if n % 2 == 1 or np.any(vals == 0):
            raise ValueError("DomainError: Logarithm of a singular matrix doesn't exist")
    if fn in ("exp", "log"):
        if fn == "exp":
            a, b = np.cos(vals), np.sin(vals)
        else:
            a, b = np.log(np.abs(vals)), np.sign(vals) * np.pi / 2
        Q1 = Qr * a - Qim * b                      # src/exp.jl:26-28
        Q2 = Qr * b + Qim * a
        return Q1 @ Qr.T + Q2 @ Qim.T
    E = np.exp(-vals)                              # src/exp.jl:134-142, :177-185
    Q1, Q2 = Qr * E, Qim * E
    Cos = Q1 @ Qr.T + Q2 @ Qim.T
    Sin = Q2 @ Qr.T - Q1 @ Qim.T
    if fn == "cis":                                # src/exp.jl:99-109: Hermitian(Q E Q^H)
        return _hermitian_view(Cos + 1j * Sin)
    if fn == "cos":
        return _hermitian_view(Cos)
    return _skewhermitian_(Sin), _hermitian_view(Cos)


def _fun_from_eigs_complex(vals, V, fn):
    if fn in ("exp", "log"):
        d = np.exp(1j * vals) if fn == "exp" else np.log(1j * vals)
        return V @ (d[:, None] * V.conj().T)          # src/exp.jl:37-46, :86-95
    P1 = (V * np.exp(-vals)) @ V.conj().T
    if fn == "cis":
        return _hermitian_view(P1)
    P2 = (V * np.exp(vals)) @ V.conj().T
    if fn == "cos":                                   # src/exp.jl:137-149
        return _hermitian_view((P1 + P2) / 2)
    return _skewhermitian_((P1 - P2) * (-0.5j)), _hermitian_view((P1 + P2) / 2)


def _fun(eigs, fn, n, cplx):
    """fn of the matrix whose eigen-decomposition eigs(scale) gives (src/exp.jl:48-284)."""
    def base(name, scale=1.0):
        e = eigs(scale)
        if cplx:
            return _fun_from_eigs_complex(e[0], e[1], name)
        return _fun_from_eigs_real(e[0], e[1], e[2], name, n)
    if fn in ("exp", "log", "cis", "cos"):
        return base(fn)
    if fn in ("sin", "sincos", "tan", "cot"):
        S, C = base("sincos")
        if fn == "sin":
            return S
        if fn == "sincos":
            return S, C
        return _skewhermitian_(np.linalg.solve(C, S) if fn == "tan" else np.linalg.solve(S, C))
    if fn == "sinh":
        return _skewhermitian_(base("exp"))
    if fn == "cosh":
        return _cosh_part(base("exp"))
    E = base("exp", 2.0)                               # tanh, coth: exp(2A) (src/exp.jl:270-284)
    I = np.eye(n)
    return _skewhermitian_(np.linalg.solve(E + I, E - I) if fn == "tanh" else np.linalg.solve(E - I, E + I))


def skewfun(A, fn):
    """exp/log/cis/cos/sin/sincos/sinh/cosh/tan/tanh/cot/coth of a skew matrix (src/exp.jl), dense
    results; real A through skeweigen!, complex A through eigen of Hermitian(iA) (src/eigen.jl:101-106:
    values -lambda, the upper triangle of iA)."""
    A = np.asarray(A)
    n = A.shape[0]
    if np.iscomplexobj(A):
        def eigs(scale):
            lam, V = np.linalg.eigh(1j * A * scale, UPLO="U")
            return -lam, V
        return _fun(eigs, fn, n, True)
    return _fun(lambda scale: skeweigen(A * scale), fn, n, False)


def skewfun_tri(ev, fn):
    """The same for SkewHermTridiagonal{<:Real}(ev), through skewtrieigen_merged! (src/tridiag.jl:506)."""
    ev = np.asarray(ev, dtype=np.float64)
    n = ev.size + 1

    def eigs(scale):
        if n == 1:
            return np.zeros(1), np.ones((1, 1)), np.zeros((1, 1))
        vals, V = skewtrieigen(ev * scale, "merged")
        return vals, V.real, V.imag
    return _fun(eigs, fn, n, False)
