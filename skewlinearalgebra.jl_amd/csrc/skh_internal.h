// skh_internal.h -- structures shared between the kernels (skh_kernels.hip) and the C ABI
// (skh_api.cpp).  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "skh_device.hpp"

struct PanelLaunch {
    double* A;
    int64_t lda;
    int64_t n;
    int64_t c0;
    int nbp;
    skh::Work w;
    double* tau;
    double* eta;
    unsigned* bar;
    unsigned base;
    int* err;
    unsigned long long* ph;   // optional phase clock (4 counters) or null
    int grid;
    int nbmax;                // panel width bound selecting the kernel instance
};

struct BatchArgs {
    const double* A;       // packed input matrices (column-major, lda = n_b each)
    const int64_t* ns;     // sizes
    const int64_t* offs;   // element offsets into A
    const int64_t* order;  // processing order (largest first)
    int64_t batch;
    double* slots;         // one work slot per workgroup
    int64_t slot_stride;   // doubles per slot
    int nb;                // panel width
    unsigned* queue;       // work-queue counter (zeroed before each launch)
    double* logpf;
    double* sgn;
    double* fac;           // optional: factors out (strictly lower, at offs[b], ld n_b)
    double* taueta;        // optional: tau | eta out (2 n_b doubles at toffs[b])
    const int64_t* toffs;
};

extern "C" {
int skh_internal_launch_panel(const PanelLaunch* pl, hipStream_t st);
int skh_internal_panel_occupancy(int nbmax, int* blocks_per_cu);
int skh_internal_launch_rank2k(double* A, int64_t lda, const double* U, int64_t ldu, int64_t n, int64_t rlo,
                               int64_t clo, int64_t c0, int nbp, hipStream_t st);
int skh_internal_launch_finish(double* A, int64_t lda, int64_t n, double* tau, double* eta, hipStream_t st);
int skh_internal_zero_u32(unsigned* p, int count, hipStream_t st);
int skh_internal_launch_batched(const BatchArgs* a, int grid, hipStream_t st);
int skh_internal_batched_occupancy(int nb, int* blocks_per_cu);
int skh_internal_mfma_probe(double* dout, hipStream_t st);
int skh_internal_eye(double* Q, int64_t ldq, int64_t n, hipStream_t st);
int skh_internal_form_v(const double* F, int64_t ldf, int64_t n, int64_t c0, int nbp, double* V, int64_t ldv,
                        hipStream_t st);
int skh_internal_larft(const double* Gm, int64_t ldg, const double* tau, int nbp, double* T, int64_t ldt,
                       hipStream_t st);
// batched helpers of the gradient (matrix index = blockIdx.z / blockIdx.x as documented in the kernels)
int skh_internal_eye_batched(double* const* Q, int64_t n, int count, hipStream_t st);
int skh_internal_form_v_batched(const double* const* F, const int64_t* toffs_tau, int64_t n, int64_t c0, int nbp,
                                double* V, int64_t vstride, int count, hipStream_t st);
int skh_internal_larft_batched(const double* Gm, int64_t gstride, const double* taueta, const int64_t* toffs,
                               int64_t c0, int nbp, double* T, int64_t tstride, int count, hipStream_t st);
int skh_internal_bidiag_solve(double* const* Q, const double* taueta, const int64_t* toffs, int64_t n, double* Zt,
                              int64_t zstride, int count, hipStream_t st);
int skh_internal_skew_out(const double* K, int64_t kstride, double* const* G, int64_t n, double alpha, int count,
                          hipStream_t st);
}

// doubles needed by one batched work slot for matrices up to nmax with panel width nb
inline int64_t skh_slot_doubles(int64_t nmax, int nb)
{
    const int64_t ld = (nmax + 1) & ~int64_t(1);
    const int64_t nt = (nmax + skh::TB - 1) / skh::TB + 1;
    int64_t s = ld * nmax + ld * nb + 2 * ld + nt + nt * 2 * skh::NBMAX + 2 * skh::NBMAX;
    s += nt * ld + nt * ld;   // Prow, Pcol
    return (s + 31) & ~int64_t(31);   // keep every slot 256-byte aligned
}
