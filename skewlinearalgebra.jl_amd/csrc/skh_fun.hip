// skh_fun.hip -- the O(n^2) steps of the matrix functions of a skew matrix (exp, log, cis, cos, sin,
// sincos, sinh, cosh, tan, tanh, cot, coth; reference src/exp.jl:3-296).  The O(n^3) work (the
// eigen-decomposition, the products V D V^T / V D V^H, the LU solves) runs in the library's
// tridiagonalization and rocBLAS / rocSOLVER; these kernels do what sits between those calls:
// scaling the columns of the eigenvectors by a function of the eigenvalues, combining products, the
// symmetrizations skewhermitian! / hermitian! / Hermitian(upper) and the shifts E +- I.
//
// All matrices are n x n, column-major with ld n; complex ones are interleaved (re, im) pairs.
// Products and sums are rounded separately (contraction off), as in the reference's broadcasts.
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace skh {
namespace fun {

constexpr int TB = 256;

inline unsigned blocks(int64_t cnt) { return (unsigned)((cnt + TB - 1) / TB); }

// mode 0 (exp, log): T1 = X a - Y b, T2 = X b + Y a;  mode 1 (cis, cos, sin): T1 = X a, T2 = Y a,
// with a, b scaling the columns (T[i, j] uses a[j], b[j])
__global__ __launch_bounds__(TB) void k_colmix(int64_t n, const double* __restrict__ X, const double* __restrict__ Y,
                                               const double* __restrict__ a, const double* __restrict__ b, int mode,
                                               double* __restrict__ T1, double* __restrict__ T2)
{
    const int64_t idx = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (idx >= n * n) return;
    const int64_t j = idx / n;
    const double x = X[idx], y = Y[idx], aj = a[j];
    if (mode == 0) {
        const double bj = b[j];
        T1[idx] = x * aj - y * bj;
        T2[idx] = x * bj + y * aj;
    } else {
        T1[idx] = x * aj;
        T2[idx] = y * aj;
    }
}

// T = V diag(d), complex V and d
__global__ __launch_bounds__(TB) void k_zcolscale(int64_t n, const double2* __restrict__ V,
                                                  const double2* __restrict__ d, double2* __restrict__ T)
{
    const int64_t idx = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (idx >= n * n) return;
    const double2 v = V[idx], s = d[idx / n];
    T[idx] = make_double2(v.x * s.x - v.y * s.y, v.x * s.y + v.y * s.x);
}

// In place, from P1 = V e^{-mu} V^H and P2 = V e^{mu} V^H: mode 0: P2 <- C = (P1 + P2) / 2;
// mode 1: P1 <- S = (P1 - P2) (-i/2);  mode 2: both.  Each thread reads and writes its own element only.
__global__ __launch_bounds__(TB) void k_zsincos(int64_t n, double2* P1, double2* P2, int mode)
{
    const int64_t idx = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (idx >= n * n) return;
    const double2 p = P1[idx], q = P2[idx];
    if (mode != 1) P2[idx] = make_double2((p.x + q.x) * 0.5, (p.y + q.y) * 0.5);
    if (mode != 0) {
        const double dx = p.x - q.x, dy = p.y - q.y;   // (dx + i dy) (-i / 2) = dy / 2 - i dx / 2
        P1[idx] = make_double2(dy * 0.5, -(dx * 0.5));
    }
}

__device__ __forceinline__ bool pair_index(int64_t n, int64_t idx, int64_t& r, int64_t& c)
{
    // one thread per (r, c) with r >= c (lower triangle and diagonal), counted over the full square
    if (idx >= n * n) return false;
    c = idx / n;
    r = idx - c * n;
    return r >= c;
}

// In place, on real (cplx = 0) or interleaved complex matrices:
//   kind 0, skewhermitian!(A) (src/skewhermitian.jl:119-131): diagonal -> 0 (real) / i imag (complex);
//           a = (A[r, c] - conj(A[c, r])) / 2, A[r, c] = a, A[c, r] = -conj(a)
//   kind 1, Hermitian(A) with uplo 'U', made dense: A[r, c] = conj(A[c, r]) for r > c, real diagonal
//   kind 2, (A + A^H) / 2 with upper-triangle rounding: a = (A[c, r] + conj(A[r, c])) / 2 stored at (c, r),
//           conj(a) at (r, c), real diagonal.  This is cosh's hermitian! + Hermitian (+ conj for complex)
//           of src/exp.jl:258-263 and :286-296 as a dense matrix
__global__ __launch_bounds__(TB) void k_sympart(int64_t n, double* __restrict__ A, int cplx, int kind)
{
    int64_t r, c;
    if (!pair_index(n, (int64_t)blockIdx.x * TB + threadIdx.x, r, c)) return;
    if (!cplx) {
        double* lo = A + c * n + r;
        double* up = A + r * n + c;
        if (r == c) {
            if (kind == 0) *lo = 0.0;
            return;
        }
        if (kind == 0) {
            const double a = (*lo - *up) * 0.5;
            *lo = a;
            *up = -a;
        } else if (kind == 1) {
            *lo = *up;
        } else {
            const double a = (*up + *lo) * 0.5;
            *up = a;
            *lo = a;
        }
        return;
    }
    double2* Z = reinterpret_cast<double2*>(A);
    double2* lo = Z + c * n + r;
    double2* up = Z + r * n + c;
    if (r == c) {
        const double2 d = *lo;
        *lo = kind == 0 ? make_double2(0.0, d.y) : make_double2(d.x, 0.0);
        return;
    }
    const double2 l = *lo, u = *up;
    if (kind == 0) {
        const double2 a = make_double2((l.x - u.x) * 0.5, (l.y + u.y) * 0.5);   // (l - conj(u)) / 2
        *lo = a;
        *up = make_double2(-a.x, a.y);
    } else if (kind == 1) {
        *lo = make_double2(u.x, -u.y);
    } else {
        const double2 a = make_double2((u.x + l.x) * 0.5, (u.y - l.y) * 0.5);   // (u + conj(l)) / 2
        *up = a;
        *lo = make_double2(a.x, -a.y);
    }
}

// A[i, i] += v (real part)
__global__ __launch_bounds__(TB) void k_adddiag(int64_t n, double* __restrict__ A, int cplx, double v)
{
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int64_t at = (i * n + i) * (cplx ? 2 : 1);
    A[at] = A[at] + v;
}

// cis of a real skew matrix as a dense Hermitian (uplo 'U') matrix from its real and imaginary parts
__global__ __launch_bounds__(TB) void k_cis_out(int64_t n, const double* __restrict__ re,
                                                const double* __restrict__ im, double2* __restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (idx >= n * n) return;
    const int64_t c = idx / n, r = idx - c * n;
    if (r < c) out[idx] = make_double2(re[idx], im[idx]);
    else if (r == c) out[idx] = make_double2(re[idx], 0.0);
    else {
        const int64_t t = r * n + c;
        out[idx] = make_double2(re[t], -im[t]);
    }
}

__global__ __launch_bounds__(TB) void k_scale(int64_t cnt, double* __restrict__ p, double s)
{
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i < cnt) p[i] = p[i] * s;
}

}  // namespace fun
}  // namespace skh

using namespace skh;

extern "C" int skh_internal_fun_colmix(int64_t n, const double* X, const double* Y, const double* a, const double* b,
                                       int mode, double* T1, double* T2, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(fun::k_colmix, dim3(fun::blocks(n * n)), dim3(fun::TB), 0, st, n, X, Y, a, b, mode, T1, T2);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_fun_zcolscale(int64_t n, const double* V, const double* d, double* T, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(fun::k_zcolscale, dim3(fun::blocks(n * n)), dim3(fun::TB), 0, st, n, (const double2*)V,
                       (const double2*)d, (double2*)T);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_fun_zsincos(int64_t n, double* P1, double* P2, int mode, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(fun::k_zsincos, dim3(fun::blocks(n * n)), dim3(fun::TB), 0, st, n, (double2*)P1, (double2*)P2,
                       mode);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_fun_sympart(int64_t n, double* A, int cplx, int kind, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(fun::k_sympart, dim3(fun::blocks(n * n)), dim3(fun::TB), 0, st, n, A, cplx, kind);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_fun_adddiag(int64_t n, double* A, int cplx, double v, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(fun::k_adddiag, dim3(fun::blocks(n)), dim3(fun::TB), 0, st, n, A, cplx, v);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_fun_cis_out(int64_t n, const double* re, const double* im, double* out, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(fun::k_cis_out, dim3(fun::blocks(n * n)), dim3(fun::TB), 0, st, n, re, im, (double2*)out);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_fun_scale(int64_t cnt, double* p, double s, hipStream_t st)
{
    if (cnt <= 0) return 0;
    hipLaunchKernelGGL(fun::k_scale, dim3(fun::blocks(cnt)), dim3(fun::TB), 0, st, cnt, p, s);
    return (int)hipGetLastError();
}
