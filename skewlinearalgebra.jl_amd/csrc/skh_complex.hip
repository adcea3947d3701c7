// skh_complex.hip -- kernels of the complex (ComplexF64) skew-Hermitian path: the cooperative
// tridiagonalization (device code in skh_zengine.hpp), the imaginary diagonal of T, and the
// pieces of the explicit Q = H_0 ... H_{n-2} (compact WY with rocBLAS products on the host side).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "skh_internal.h"
#include "skh_sync.hpp"
#include "skh_zengine.hpp"

namespace skh {

// Columns k in [k_begin, k_end) (k_end = n - 1 for the last segment); the norm partials and head
// of column k_begin come from the previous launch (or zinit for k_begin = 0).
__global__ __launch_bounds__(WG, 1) void k_zfused(double2* A, int64_t lda, int64_t n, zc::ZWork w, double2* tau,
                                                  double* eta, unsigned* bar, int* err, int64_t k_begin,
                                                  int64_t k_end)
{
    using namespace zc;
    __shared__ ZShared sh;
    __shared__ int s_ok;
    const int G = gridDim.x, g = blockIdx.x;
    unsigned epoch = 0;
    if (k_begin == 0) {
        zinit(A, lda, n, w, g, G);
        if (!grid_sync_co(bar, ++epoch, err, &s_ok)) return;
    }
    for (int64_t k = k_begin; k < k_end; k++) {
        const ZScal s = zscalars(w, G, sh);
        if (g == 0 && threadIdx.x == 0) {
            st(tau + k, s.tau);
            eta[k] = s.beta;
        }
        if (k == n - 2) {
            // last (length-1) reflector: the factor holds its unit, and the 1x1 trailing block is
            // unchanged by it (|1 - tau| = 1) but still owes the previous reflector's update
            if (g == 0 && threadIdx.x == 0) zlast(A, lda, n, w, s);
            break;
        }
        zpass(A, lda, n, k, k >= 1, s, w, sh, g, G);
        if (!grid_sync_co(bar, ++epoch, err, &s_ok)) return;
        zrows(A, lda, n, k, s, w, sh, g, G, G);
        if (!grid_sync_co(bar, ++epoch, err, &s_ok)) return;
    }
}

// dvim[i] = Im A[i, i]
__global__ void k_zdiag(const double2* A, int64_t lda, int64_t n, double* dvim)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dvim[i] = A[i + i * lda].y;
}

// V (m x nbp, rows c0+1..n-1) of reflectors c0..c0+nbp-1 with explicit unit and zeros above it
__global__ void k_zform_v(const double2* F, int64_t ldf, int64_t n, int64_t c0, int nbp, double2* V, int64_t ldv)
{
    const int k = blockIdx.y;
    const int64_t m = n - c0 - 1;
    for (int64_t rr = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; rr < m; rr += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = c0 + 1 + rr;
        const int64_t head = c0 + k + 1;
        double2 v = make_double2(0.0, 0.0);
        if (r == head) v = make_double2(1.0, 0.0);
        else if (r > head) v = F[r + (c0 + k) * ldf];
        V[rr + k * ldv] = v;
    }
}

// T (nbp x nbp upper) from Gm = V^H V and tau: T[k,k] = tau_k, T[0:k,k] = -tau_k T[0:k,0:k] Gm[0:k,k]
// complex T of the compact WY form (as k_larft): thread i builds row i of T in LDS, no barrier; nbp <= 64
__global__ __launch_bounds__(64) void k_zlarft(const double2* Gm, int64_t ldg, const double2* tau, int nbp, double2* T,
                                               int64_t ldt)
{
    using namespace zc;
    __shared__ double2 row[64 * 64];
    const int i = threadIdx.x;
    if (i >= nbp) return;
    for (int k = 0; k < nbp; k++) {
        cd v = cd{0.0, 0.0};
        const cd tk = ld(tau + k);
        if (i < k) {
            cd acc = cd{0.0, 0.0};
            for (int l = i; l < k; l++) acc = acc + ld(row + l * 64 + i) * ld(Gm + l + k * ldg);
            v = cd{0.0, 0.0} - tk * acc;
        } else if (i == k) {
            v = tk;
        }
        st(row + k * 64 + i, v);
    }
    for (int k = 0; k < nbp; k++) T[i + k * ldt] = row[k * 64 + i];
}

__global__ void k_zeye(double2* Q, int64_t ldq, int64_t n)
{
    const int64_t c = blockIdx.y;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        Q[r + c * ldq] = make_double2(r == c ? 1.0 : 0.0, 0.0);
}

// Zc[r, j] = i^r C[r, j]: eigenvectors of D^H (iT) D mapped back to iT (D = diag(i^r))
__global__ void k_zphase(const double* C, int64_t n, double2* Zc)
{
    const int64_t j = blockIdx.y;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const double c = C[r + j * n];
        double2 z;
        switch (r & 3) {
            case 0: z = make_double2(c, 0.0); break;
            case 1: z = make_double2(0.0, c); break;
            case 2: z = make_double2(-c, 0.0); break;
            default: z = make_double2(0.0, -c); break;
        }
        Zc[r + j * n] = z;
    }
}

}  // namespace skh

using namespace skh;

extern "C" int skh_internal_zphase(const double* C, int64_t n, double2* Zc, hipStream_t st)
{
    hipLaunchKernelGGL(k_zphase, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 64)), (unsigned)n),
                       dim3(256), 0, st, C, n, Zc);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_zfused_occupancy(int* blocks_per_cu)
{
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)k_zfused, WG, 0);
}

extern "C" int skh_internal_launch_zfused(double2* A, int64_t lda, int64_t n, const zc::ZWork* w, double2* tau,
                                          double* eta, unsigned* bar, int* err, int grid, int64_t k_begin,
                                          int64_t k_end, hipStream_t st)
{
    zc::ZWork wc = *w;
    void* args[] = {(void*)&A,   (void*)&lda, (void*)&n,   (void*)&wc,      (void*)&tau,
                    (void*)&eta, (void*)&bar, (void*)&err, (void*)&k_begin, (void*)&k_end};
    return (int)hipLaunchCooperativeKernel((const void*)k_zfused, dim3(grid), dim3(WG), args, 0, st);
}

extern "C" int skh_internal_zdiag(const double2* A, int64_t lda, int64_t n, double* dvim, hipStream_t st)
{
    hipLaunchKernelGGL(k_zdiag, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1024))), dim3(256),
                       0, st, A, lda, n, dvim);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_zform_v(const double2* F, int64_t ldf, int64_t n, int64_t c0, int nbp, double2* V,
                                    int64_t ldv, hipStream_t st)
{
    const int64_t m = n - c0 - 1;
    hipLaunchKernelGGL(k_zform_v, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((m + 255) / 256, 64)), nbp),
                       dim3(256), 0, st, F, ldf, n, c0, nbp, V, ldv);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_zlarft(const double2* Gm, int64_t ldg, const double2* tau, int nbp, double2* T, int64_t ldt,
                                   hipStream_t st)
{
    if (nbp > 64) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zlarft, dim3(1), dim3(64), 0, st, Gm, ldg, tau, nbp, T, ldt);
    return (int)hipGetLastError();
}

extern "C" int skh_internal_zeye(double2* Q, int64_t ldq, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zeye, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 64)), (unsigned)n),
                       dim3(256), 0, st, Q, ldq, n);
    return (int)hipGetLastError();
}
