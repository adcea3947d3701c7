// skh_internal.h -- structures shared between the kernels (skh_kernels.hip, skh_compact.hip)
// and the C ABI (skh_api.cpp).  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "skh_device.hpp"
#include "skh_zengine.hpp"
#include "skh_zpanel.hpp"

struct PanelLaunch {
    double* A;
    int64_t lda;
    int64_t n;
    int64_t c0;
    int nbp;
    skh::Work w;
    double* tau;
    double* eta;
    int* err;
    unsigned long long* ph;   // optional phase clock (4 counters) or null
    int grid;
    int nbmax;                // panel width bound selecting the kernel instance
    int multi;                // 1 if a workgroup may own several 64-row blocks (grid < blocks)
};

// Compact engine (one workgroup per matrix, n <= nmax of the instantiation).
struct CompactArgs {
    const double* A;             // packed inputs, column-major, matrix b at offs[b] (even), ld n_b
    const int64_t* ns;
    const int64_t* offs;
    const int64_t* items;        // matrices of this launch in processing order (largest first)
    int64_t count;
    double* slots;               // one work slot per workgroup (when fac is null)
    int64_t slot_stride;         // doubles per slot: >= nmax^2 + 2*NMAX of the kernel, even
    unsigned* queue;             // work-queue counter (zeroed before each launch)
    int* err;                    // error flag: bit 1 = watchdog
    unsigned long long budget;   // watchdog, 100 MHz ticks (0 = none)
    double* logpf;
    double* sgn;
    double* fac;                 // optional: factorization in place here (input copied in first)
    double* taueta;              // optional: tau | eta out (2 n_b doubles at toffs[b]); required with fac
    const int64_t* toffs;
    double* wbuf;                // blocked engine: W panel workspace of each workgroup (NMAX x NB doubles)
    int64_t wstride;             // doubles between two workgroups' W workspaces
};

extern "C" {
int skh_internal_zfused_occupancy(int* blocks_per_cu);
int skh_internal_launch_zfused(double2* A, int64_t lda, int64_t n, const skh::zc::ZWork* w, double2* tau, double* eta,
                               unsigned* bar, int* err, int grid, int64_t k_begin, int64_t k_end, int fresh,
                               hipStream_t st);
// blocked complex path: panel kernel (one workgroup per CU), diagonal write-back, MFMA rank-2k update
int skh_internal_zpanel_occupancy(int* blocks_per_cu);
int skh_internal_launch_zpanel(double2* A, int64_t lda, int64_t n, int64_t c0, int nbp, const skh::zp::ZPWork* w,
                               double2* tau, double* eta, int* err, int grid, hipStream_t st);
int skh_internal_launch_zdfin(double2* A, int64_t lda, const double2* Dfin, int64_t c0, int nbp, hipStream_t st);
int skh_internal_zdiag(const double2* A, int64_t lda, int64_t n, double* dvim, hipStream_t st);
int skh_internal_zform_v(const double2* F, int64_t ldf, int64_t n, int64_t c0, int nbp, double2* V, int64_t ldv,
                         hipStream_t st);
int skh_internal_zlarft(const double2* Gm, int64_t ldg, const double2* tau, int nbp, double2* T, int64_t ldt,
                        hipStream_t st);
int skh_internal_zeye(double2* Q, int64_t ldq, int64_t n, hipStream_t st);
int skh_internal_zphase(const double* C, int64_t n, double2* Zc, hipStream_t st);
// matrix functions of a skew matrix (csrc/skh_fun.hip): column mixing of eigenvectors, combinations and
// symmetrizations (kind 0 skewhermitian!, 1 Hermitian upper made dense, 2 Hermitian part), diagonal shift
int skh_internal_fun_colmix(int64_t n, const double* X, const double* Y, const double* a, const double* b, int mode,
                            double* T1, double* T2, hipStream_t st);
int skh_internal_fun_zcolscale(int64_t n, const double* V, const double* d, double* T, hipStream_t st);
int skh_internal_fun_zsincos(int64_t n, double* P1, double* P2, int mode, hipStream_t st);
int skh_internal_fun_sympart(int64_t n, double* A, int cplx, int kind, hipStream_t st);
int skh_internal_fun_adddiag(int64_t n, double* A, int cplx, double v, hipStream_t st);
int skh_internal_fun_cis_out(int64_t n, const double* re, const double* im, double* out, hipStream_t st);
int skh_internal_fun_scale(int64_t cnt, double* p, double s, hipStream_t st);
// host (host_tridiag.cpp): eigenvalues of the symmetric tridiagonal (d, e), ascending into d; 0 or 2
int skh_internal_symtri_eigvals(int64_t n, double* d, const double* e);
int skh_internal_compact_occupancy(int nmax, int* blocks_per_cu);
// complex skew-symmetric Pfaffians (skh_zpfaff.hip): mode 0 writes Pf (2 doubles per matrix) to o1,
// mode 1 writes log|Pf| to o1 and cis(phase) (2 doubles per matrix) to o2
int skh_internal_zpfaff(const double2* A, const int64_t* ns, const int64_t* offs, const int64_t* order, int64_t batch,
                        int64_t nmax, double2* slots, int64_t slot, int grid, int mode, double* o1, double* o2,
                        hipStream_t st);
int skh_internal_zpfaff_occupancy(int64_t nmax, int* per_cu);
int skh_internal_zpfaff_nmax();
int skh_internal_zpfaff_coop(const double2* A, double2* L, int64_t n, int mode, double* o1, double* o2, double2* W,
                             double* part, unsigned* bar, int* err, int grid, hipStream_t st);
int skh_internal_zpfaff_coop_occupancy(int* per_cu);
// host skew-tridiagonal QR (host_tridiag.cpp) with its Qodd / Qeven rotations handed to `apply`
typedef int (*skh_rot_apply_fn)(void* ctx, int64_t nrows, const int32_t* ko, const double* cso, int64_t nro,
                                double* Qo, int64_t no, const int32_t* ke, const double* cse, int64_t nre, double* Qe,
                                int64_t ne);
// eigenpairs of SymTridiagonal(0, e) (n - 1 off-diagonal values): d ascending, Z n x n column-major;
// 0 on success, 2 if the solver did not converge, another status on failure
typedef int (*skh_symtri_fn)(void* ctx, int64_t n, const double* e, double* d, double* Z);
int skh_internal_skewtrieigen(int64_t N, double* evp, double* vals, double* vr, double* vi, int variant,
                              skh_rot_apply_fn apply, skh_symtri_fn symtri, void* ctx);
// to_symtridiagonal of a complex SkewHermTridiagonal and the eigenvalues of a symmetric tridiagonal (host)
void skh_internal_to_symtri_z(int64_t n, const double* ev, const double* dvim, double* d, double* e, double* q);
int skh_internal_symtri_eigvals(int64_t n, double* d, const double* e);
// rotation chains on the GPU (skh_rot.hip): seg* hold nseg* + 1 chain starts into the lists
int skh_internal_rot_chains(double* Qo, int64_t ldo, const int32_t* ko, const double* cso, const int64_t* sego,
                            int64_t nsego, double* Qe, int64_t lde, const int32_t* ke, const double* cse,
                            const int64_t* sege, int64_t nsege, int64_t nrows, hipStream_t st);
// skew-Cholesky (skh_chol.hip): matrix b at A + offs[b] and R + offs[b] (ld n[b]), p + poffs[b]
int skh_internal_skewchol(const double* A, double* R, const int64_t* ns, const int64_t* offs, const int64_t* poffs,
                          const int64_t* order, int64_t batch, int64_t nmax, int pivot, int64_t* p, int64_t* steps,
                          int grid, hipStream_t st);
int skh_internal_skewchol_occupancy(int64_t nmax, int* per_cu);
int skh_internal_skewchol_nmax();
// one large matrix over the whole GPU (cooperative): W 4n doubles, part 3 * grid doubles
int skh_internal_skewchol_coop(const double* A, double* U, int64_t n, int pivot, int64_t* p, int64_t* steps,
                               double* W, double* part, unsigned* bar, int* err, int grid, hipStream_t st);
int skh_internal_skewchol_coop_occupancy(int* per_cu);
int skh_internal_launch_compact(const CompactArgs* a, int nmax, int grid, hipStream_t st);
// blocked engine (skh_cblock.hpp): instances (nmax, nb) in {(256, 8), (256, 16), (512, 8)}; nonzero if absent
int skh_internal_launch_cblock(const CompactArgs* a, int nmax, int nb, int grid, hipStream_t st);
int skh_internal_cblock_occupancy(int nmax, int nb, int* blocks_per_cu);
int skh_internal_logpf_from(const double* eta, const double* tau, int64_t n, double* logpf, double* sgn, hipStream_t st);
int skh_internal_launch_panel(const PanelLaunch* pl, hipStream_t st);
// 1 if SKH_PANEL_COOPERATIVE asks for cooperative launches of the barrier-free panel kernels
int skh_internal_panel_cooperative();
// Segments of device memory to fill with the panel kernel's empty marker: entries [lo, hi) of
// `rows` rows, ld doubles apart, at p (one launch for all segments).  first / chunks: the block plan
// of a fill done by 64-thread workgroups of another launch (set by the launcher).
struct FillSegs {
    static constexpr int MAXSEG = 12;
    double* p[MAXSEG];
    int64_t rows[MAXSEG];
    int64_t ld[MAXSEG];
    int64_t lo[MAXSEG];
    int64_t hi[MAXSEG];
    int64_t first[MAXSEG + 1];   // first workgroup of each segment, [nseg] = total
    int64_t chunks[MAXSEG];      // workgroups per row of each segment
    int nseg;
};
int skh_internal_fill_empty(const FillSegs* fs, int nseg, hipStream_t st);
int skh_internal_panel_occupancy(int nbmax, int* blocks_per_cu);
// the complex rank-2k update after a panel; with fs (nseg > 0) the same launch also fills those segments
int skh_internal_launch_zrank2k(double2* A, int64_t lda, const double2* U, int64_t ldu, int64_t n, int64_t e,
                                int64_t c0, int nbp, const FillSegs* fs, int nseg, hipStream_t st);
// the rank-2k update after a panel; with fs (nseg > 0) the same launch also fills those segments
int skh_internal_launch_rank2k(double* A, int64_t lda, const double* U, int64_t ldu, int64_t n, int64_t e,
                               int64_t c0, int nbp, const FillSegs* fs, int nseg, hipStream_t st);
int skh_internal_launch_finish(double* A, int64_t lda, int64_t n, double* tau, double* eta, hipStream_t st);
int skh_internal_zero_u32(unsigned* p, int count, hipStream_t st);
// fixed device scratch of the allocation-free factorization (skh_noalloc.hip): setup looks up the
// array's address on `device` (which must be current); scratch returns the control block (2 KiB) and
// the workspace after it
int skh_internal_noalloc_setup(int device);
int skh_internal_noalloc_scratch(int device, unsigned** ctl, double** ws, size_t* ws_doubles);
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
// backward pass of the batched log|Pf| gradient (skh_grad.hip): batched FP64 MFMA GEMM
// C_z = alpha op(A_z) op(B_z) + beta C_z (X_z = X + z sX, ta / tb: transpose), and the helpers on
// members at base + goff[z] (factors in, gradient out) with strided workspaces
int skh_internal_bgemm(int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                       int64_t sA, const double* B, int64_t ldb, int64_t sB, double beta, double* C, int64_t ldc,
                       int64_t sC, int count, hipStream_t st);
int skh_internal_grad_eye(double* Q, int64_t qstride, int64_t n, int count, hipStream_t st);
int skh_internal_grad_form_v(const double* F, const int64_t* goff, int64_t n, int64_t c0, int nbp, double* V,
                             int64_t vstride, int count, hipStream_t st);
int skh_internal_grad_bidiag(const double* Q, int64_t qstride, const double* taueta, const int64_t* toff, int64_t n,
                             double* Zt, int64_t zstride, int count, hipStream_t st);
int skh_internal_grad_skew_out(const double* K, int64_t kstride, double* G, const int64_t* goff, int64_t n,
                               double alpha, int count, hipStream_t st);
int skh_internal_grad_zero(double* G, const int64_t* goff, const int64_t* gn, int count, hipStream_t st);
}

