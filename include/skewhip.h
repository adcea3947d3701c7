/*
 * skewhip.h -- C ABI of libskewhip.so, the MI355X (gfx950, FP64) implementation of
 * SkewLinearAlgebra.jl's skew tridiagonalization path.
 *
 * Conventions (what Julia's ccall would bind):
 *   - column-major matrices, 0-based indices, int64_t sizes and leading dimensions;
 *   - "host" entry points take caller-owned host memory (Julia arrays under GC.@preserve)
 *     and block until the result is back in that memory;
 *   - "_device" entry points take device (HBM) pointers and a hipStream_t (passed as void*; NULL
 *     is HIP's default stream), and are asynchronous with respect to the host unless stated
 *     otherwise;
 *   - return status: 0 success; -k invalid argument number k (1-based); > 0 runtime error
 *     (SKH_ERR_*).  skh_last_error() gives a message for the calling thread.
 * Calls are serialised by an internal mutex; the library owns cached device workspaces.
 * The device work of one call uses those workspaces and one set of grid-barrier control words,
 * so the work of two _device calls must not overlap on the GPU: issue them on one stream, or
 * synchronise between calls on different streams (plans own their buffers and are exempt).
 * The large engine's panel kernels need all their workgroups resident at once; their grid is the
 * occupancy bound and they use a plain launch, so work of other streams or processes running on
 * the GPU at the same time can keep workgroups out past the bounded waits (SKH_ERR_BARRIER).
 * SKH_PANEL_COOPERATIVE=1 in the environment selects cooperative launches, which the runtime
 * places co-resident beside other work, at about 22 us more per panel.
 */
#ifndef SKEWHIP_H
#define SKEWHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SKH_OK 0
#define SKH_ERR_HIP 1          /* HIP runtime error (message in skh_last_error) */
#define SKH_ERR_NOCONV 2       /* implicit QR hit its iteration limit */
#define SKH_ERR_BARRIER 3      /* a bounded wait of a persistent kernel timed out (a workgroup not resident or lost) */
#define SKH_ERR_NOT_SKEW 4     /* input is not skew-symmetric (logabspfaffian! of a Matrix) */
#define SKH_ERR_UNSUPPORTED 5  /* requested option is not available */
#define SKH_ERR_DOMAIN 6       /* log of a singular skew matrix (Julia's DomainError) */
#define SKH_ERR_SINGULAR 7     /* LU met an exactly zero pivot in tan/cot/tanh/coth (SingularException) */

/* engine selection in skh_opts.flags */
#define SKH_ENGINE_AUTO 0      /* single-matrix entries: one workgroup for even n <= 320, LARGE above */
#define SKH_ENGINE_LARGE 1     /* persistent multi-workgroup panel + MFMA rank-2k update */
#define SKH_ENGINE_TEAM1 2     /* one workgroup per matrix (the batched engine) */
#define SKH_ENGINE_BLOCKED 3   /* one workgroup per matrix, panels of opts.nb = 8 (n <= 512) or 16 (n <= 256)
                                  reflectors with an MFMA rank-2k update per panel (0 = 8) */
#define SKH_FLAG_TIMING 0x100  /* record per-kernel HIP events (read with skh_diag_timings) */
#define SKH_FLAG_PHASES 0x200  /* also run the panel kernel's phase clock (workgroup 0; slows it) */

typedef struct skh_opts {
    int64_t nb;     /* panel width; 0 = autotuned table; must be <= 128 */
    int64_t flags;  /* SKH_ENGINE_* */
} skh_opts;

/* Library lifetime.  skh_init selects the HIP device; the other calls init device 0 lazily. */
int skh_init(int device);
int skh_finalize(void);
const char *skh_version(void);
const char *skh_last_error(void);

/*
 * Blocked Householder skew tridiagonalization, real Float64.
 * Replaces skewblockedhess!(S::SkewHermitian{Float64}) -> (tau, eta)
 *   (reference src/hessenberg.jl:144-174; callers hessenberg! :5-13, skeweigvals! src/eigen.jl:49-73,
 *    skeweigen! src/eigen.jl:75-91, _pfaffian!/_logabspfaffian! src/pfaffian.jl:74-84,167-179).
 * On return the strictly lower triangle of A holds the 'L'-layout factors (reflector i in
 * A[i+1:n, i] with the unit A[i+1,i] = 1 stored; the last entry A[n-1,n-2] keeps eta[n-2]), the
 * diagonal and upper triangle are left as given.  tau, eta have n-1 entries.
 * *ntau_nonzero (may be NULL) = number of tau != 0, so det(Q) = (-1)^ntau_nonzero.
 */
int skh_sytrd_d(int64_t n, double *A, int64_t lda, double *tau, double *eta, int64_t *ntau_nonzero,
                const skh_opts *opts);
int skh_sytrd_d_device(int64_t n, double *dA, int64_t lda, double *dtau, double *deta, void *stream,
                       const skh_opts *opts);

/*
 * skh_sytrd_d_device without any allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, kernels and results as skh_sytrd_d_device (SKH_FLAG_TIMING and
 * SKH_FLAG_PHASES are ignored), for n <= SKH_NOALLOC_MAX_N; larger n is rejected (status -1).  The
 * workspace and control words live in a fixed device array of about 120 MiB (1.571e7 doubles) that
 * each GPU the library runs on holds until the process ends.  Every call on a GPU shares
 * that GPU's array and starts by clearing its control words: calls that overlap on one GPU race (issue
 * them on one stream, or synchronise between them).
 */
#define SKH_NOALLOC_MAX_N 8192
int skh_sytrd_d_device_noalloc(int64_t n, double *dA, int64_t lda, double *dtau, double *deta, void *stream,
                               const skh_opts *opts);

/*
 * Status of the device work of the last _device entry call (the allocation-free one included):
 * waits for `stream`, then returns SKH_ERR_BARRIER (details in skh_last_error) if one of its
 * persistent kernels abandoned a bounded wait (its results are then invalid), else 0.  The _device
 * entries of the large engine do not wait for their kernels, so this is where they report it; each
 * such call clears the flag when it starts.  No counterpart in the reference (its calls are synchronous).
 */
int skh_device_status(void *stream);

/*
 * log|Pf(A)| and sign for a real skew-symmetric A (A is not modified).
 * Replaces _logabspfaffian!(A::SkewHermitian{<:Real}) (reference src/pfaffian.jl:167-179):
 * odd n -> (-Inf, 0).
 */
int skh_logpf_d(int64_t n, const double *A, int64_t lda, double *logpf, double *sgn, const skh_opts *opts);
/* device variant (dA 16-byte aligned, *dlogpf and *dsgn in HBM), the same engine choice (SKH_ENGINE_AUTO:
 * one workgroup for even n <= 320, the cooperative engine above).  Enqueued on `stream`; the one-workgroup
 * engine returns once its launch is done. */
int skh_logpf_d_device(int64_t n, const double *dA, int64_t lda, double *dlogpf, double *dsgn, void *stream,
                       const skh_opts *opts);

/*
 * Batched log|Pf| + sign over `batch` real skew matrices of uneven sizes, packed column-major
 * (matrix b is n[b] x n[b] at element offset offset[b] of Apacked, lda = n[b]).
 * No reference counterpart: the reference user loops logabspfaffian (src/pfaffian.jl:236).
 * One persistent launch, one workgroup per matrix, largest matrices first.
 */
int skh_logpf_batched_d(int64_t batch, const int64_t *n, const int64_t *offset, const double *Apacked,
                        double *logpf, double *sgn, const skh_opts *opts);
/* device variant: dA, dlogpf, dsgn in HBM; n and offset are host arrays (used for scheduling). */
int skh_logpf_batched_d_device(int64_t batch, const int64_t *n, const int64_t *offset, const double *dA,
                               double *dlogpf, double *dsgn, void *stream, const skh_opts *opts);

/*
 * Batched log|Pf| + sign + gradient d log|Pf(A)| / dA = -1/2 A^{-1} (entrywise convention,
 * i.e. the cotangent a ChainRulesCore rrule of logabspfaffian returns for a unit seed).
 * No reference counterpart (the reference has no AD code; its CPU oracle for this quantity is
 * inv(A::SkewHermitian), src/skewhermitian.jl:233).  grad uses the same packing as Apacked.
 * A^{-1} = Q T^{-1} Q^T is formed from the stored reflectors with the closed-form inverse of the
 * skew tridiagonal T, every product in a hand-written batched FP64 MFMA GEMM.  Odd n: logpf = -Inf,
 * sign 0, grad = 0.  The device variant keeps the plan of its last batch layout (sizes and offsets)
 * and only enqueues when called again with the same layout; it blocks only while it builds that plan
 * and when members with n > 1024 are present (those are factored one at a time by the cooperative engine).
 */
int skh_logpf_grad_batched_d(int64_t batch, const int64_t *n, const int64_t *offset, const double *Apacked,
                             double *logpf, double *sgn, double *grad, const skh_opts *opts);
int skh_logpf_grad_batched_d_device(int64_t batch, const int64_t *n, const int64_t *offset, const double *dA,
                                    double *dlogpf, double *dsgn, double *dgrad, void *stream, const skh_opts *opts);

/*
 * Batched log|Pf| plans (no reference counterpart: the reference user loops logabspfaffian,
 * src/pfaffian.jl:236).  The sizes and offsets of the batch are fixed at creation (even offsets,
 * every even n <= 1024) and uploaded once; the plan owns its device work slots and control
 * words.  skh_logpf_batched_plan_exec enqueues one evaluation of the whole batch on `stream`
 * (a fill kernel and one persistent launch) without allocating, synchronising or copying from
 * the host, so it can be captured into a hipGraph and replayed.  An untuned plan takes the blocked
 * engine (panels of 8 reflectors, MFMA update) when every member has n <= 512, else the unblocked
 * one.  skh_logpf_batched_plan_tune times the kernel configurations (engine, kernel NMAX, panel
 * width 8 or 16, 1 to 4 work-groups per CU) on real data and keeps the fastest (*best_ms may be NULL).  skh_plan_status waits for `stream` and reports a
 * watchdog stop (SKH_ERR_BARRIER).  Overlapping executions of one plan race (they share its
 * slots and queue).  skh_plan_info: {work-groups, kernel NMAX, tuned, matrices}.
 */
typedef struct skh_plan_s *skh_plan;
int skh_logpf_batched_plan_create(int64_t batch, const int64_t *n, const int64_t *offset, const skh_opts *opts,
                                  skh_plan *plan);
int skh_logpf_batched_plan_exec(skh_plan plan, const double *dA, double *dlogpf, double *dsgn, void *stream);
int skh_logpf_batched_plan_tune(skh_plan plan, const double *dA, double *dlogpf, double *dsgn, void *stream,
                                double *best_ms);
/*
 * Gradient plans: the same for skh_logpf_grad_batched_d_device (log|Pf|, sign and -1/2 A^{-1} into dgrad,
 * which uses the input's packing; even n <= 1024).  Creation uploads the control arrays and allocates the
 * workspaces (tau | eta, the Q formation's buffers for the largest size group); exec enqueues the forward
 * launch and the backward kernels (hand-written: batched FP64 MFMA GEMMs, compact-WY helpers, the bidiagonal
 * solve and the skew output) without allocating, synchronising, copying from the host or memset, so it can be
 * captured into a hipGraph.  skh_plan_status, skh_plan_info and skh_plan_destroy take either kind of plan.
 */
int skh_logpf_grad_batched_plan_create(int64_t batch, const int64_t *n, const int64_t *offset, const skh_opts *opts,
                                       skh_plan *plan);
int skh_logpf_grad_batched_plan_exec(skh_plan plan, const double *dA, double *dlogpf, double *dsgn, double *dgrad,
                                     void *stream);
int skh_plan_status(skh_plan plan, void *stream);
/* The plan's launch configuration, out4 = {engine (SKH_ENGINE_TEAM1: unblocked, SKH_ENGINE_BLOCKED), kernel
 * NMAX, panel width (0 for the unblocked engine), work-groups}; skh_plan_set_config sets it (log|Pf| plans),
 * e.g. from a table of tuned configurations kept by the caller (the Python mirror keeps one as JSON). */
int skh_plan_config(skh_plan plan, int64_t *out4);
int skh_plan_set_config(skh_plan plan, const int64_t *in4);
int skh_plan_info(skh_plan plan, int64_t *out4);
int skh_plan_destroy(skh_plan plan);

/*
 * Explicit Q = H_0 ... H_{n-2} (n x n) from the factors and tau of skh_sytrd_d (device memory).
 * Replaces Matrix(F.Q) for F = hessenberg(A::SkewHermitian) (HessenbergQ ctor src/hessenberg.jl:3,
 * used by skeweigen! src/eigen.jl:85; LAPACK orgtr('L') convention: the unit of each reflector is
 * implicit).  Blocked compact-WY on the GPU.
 */
int skh_orgtr_d_device(int64_t n, const double *dF, int64_t ldf, const double *dtau, double *dQ, int64_t ldq,
                       void *stream);

/*
 * Complex skew-Hermitian tridiagonalization, ComplexF64 (interleaved re/im pairs, i.e. Julia's
 * Matrix{ComplexF64}; lda counts complex elements).
 * Replaces the complex branch of skewblockedhess!(S::SkewHermitian{ComplexF64}) -> (tau, eta)
 *   (reference src/hessenberg.jl:92-131 with :96-104, :123-126, :161, tail :70-90) and
 *   hessenberg!(A::SkewHermitian{<:Complex}) (:5-13).
 * On return the strictly lower triangle holds the factors (the unit of every reflector stored,
 * including the last length-1 reflector), the diagonal holds T's (imaginary) diagonal, the upper
 * triangle is left as given.  tau: n-1 complex; eta: n-1 real (the subdiagonal of T); dvim: n
 * (imaginary parts of T's diagonal: SkewHermTridiagonal(eta, dvim)).
 */
int skh_sytrd_z(int64_t n, double *A, int64_t lda, double *tau, double *eta, double *dvim, const skh_opts *opts);
/* device variant: all pointers in HBM (A 16-byte aligned); asynchronous on `stream`. */
int skh_sytrd_z_device(int64_t n, double *dA, int64_t lda, double *dtau, double *deta, double *ddvim, void *stream,
                       const skh_opts *opts);
/* Explicit unitary Q = H_0 ... H_{n-2} from complex factors and tau (device memory, LAPACK
 * zungtr('L') convention), replacing Matrix(F.Q) for the complex Hessenberg (src/hessenberg.jl:3). */
int skh_orgtr_z_device(int64_t n, const double *dF, int64_t ldf, const double *dtau, double *dQ, int64_t ldq,
                       void *stream);
/*
 * Eigenvalues / eigen-decomposition of a complex skew-Hermitian A (host memory, not modified).
 * Replace eigvals!/eigen!(A::SkewHermitian{<:Complex}) (reference src/eigen.jl:18-27, :101-106,
 * which call LAPACK on Hermitian(iA)): vals[k] = -lambda_k for the ascending eigenvalues lambda of
 * iA, so the eigenvalues are complex(0, vals[k]); V (eigen only): n x n complex eigenvectors
 * (interleaved, leading dimension n).  Here: GPU tridiagonalization A = Q T Q^H, the real
 * symmetric tridiagonal D^H (iT) D (D = diag(i^r)) solved on the host (eigvals) or by GPU divide
 * and conquer (eigen), vectors Q D Z on the GPU.
 */
int skh_eigvals_z(int64_t n, const double *A, int64_t lda, double *vals, const skh_opts *opts);
int skh_eigen_z(int64_t n, const double *A, int64_t lda, double *vals, double *V, const skh_opts *opts);

/*
 * Pfaffians of a batch of complex skew-SYMMETRIC matrices (A^T = -A, ComplexF64, interleaved
 * re/im; matrix b is n[b] x n[b] column-major at complex-element offset offset[b], lda = n[b]).
 * Replaces _pfaffian!(A::AbstractMatrix{<:Complex}) (reference src/pfaffian.jl:97-137, reached
 * from pfaffian!/pfaffian, :155-158 and :144) and _logabspfaffian!(A::AbstractMatrix{<:Complex})
 * (:193-233, from logabspfaffian!, :259-263): Parlett-Reid elimination with pivoting.
 * Only the strictly lower triangle is read (the reference's isskewsymmetric check is the caller's);
 * A is not modified.  n[b] <= 8192.  pf: 2 doubles (re, im) per matrix; odd n or a zero pivot
 * give 0.  logpf: log|Pf| per matrix, phase: cis(phase) as 2 doubles per matrix; odd n or a zero
 * pivot give (-Inf, 0).  No reference counterpart for the batching itself (the reference user
 * loops pfaffian).  Even n >= 512: one cooperative launch over the whole GPU per matrix; the
 * rest one workgroup per matrix, largest first.
 */
int skh_pfaffian_z_batched(int64_t batch, const int64_t *n, const int64_t *offset, const double *Apacked, double *pf,
                           const skh_opts *opts);
int skh_logabspfaffian_z_batched(int64_t batch, const int64_t *n, const int64_t *offset, const double *Apacked,
                                 double *logpf, double *phase, const skh_opts *opts);
/*
 * Skew-Cholesky factorization R^T J R = A[p, p] of a real skew-symmetric A (P. Benner et al., ETNA
 * 11, 2000), J = JMatrix{+1}: 2 x 2 blocks [0 1; -1 0] on the diagonal.
 * Replaces skewchol(A, pivot) / skewchol!(A, pivot) (reference src/cholesky.jl:113-151) and
 * _skewchol!(A::SkewHermitian{<:Real}, ::RowMaximum) (:58-107, pivot = 1, the reference default)
 * and (..., ::NoPivot) (:28-56, pivot = 0).  Only the strict upper triangle of A is read (A is taken
 * to be exactly skew-symmetric: the reference's isskewhermitian check is the caller's); A is not
 * modified.  R (n x n, ldr): upper triangular, strict lower triangle zero.  p: n int64, 0-based
 * (p = 0..n-1 without pivoting).  steps: 2 x 2 steps completed (n/2 unless the pivot search stops
 * below tol = 1e-15 ||A||_F, leaving the trailing square in R); may be NULL.  n <= 4096.
 * Matrices with n >= 512 are factored one at a time by a cooperative kernel over the whole GPU,
 * smaller ones one workgroup per matrix, largest first; both give the same bits.
 */
int skh_skewchol_d(int64_t n, const double *A, int64_t lda, double *R, int64_t ldr, int64_t *p, int pivot,
                   int64_t *steps, const skh_opts *opts);
/* Batched: matrix b at element offset offset[b] of Apacked and of Rpacked (ld n[b]); its p at
 * p + (n[0] + ... + n[b-1]); steps: batch entries (may be NULL).  No reference counterpart for the
 * batching itself (the reference user loops skewchol). */
int skh_skewchol_d_batched(int64_t batch, const int64_t *n, const int64_t *offset, const double *Apacked,
                           double *Rpacked, int64_t *p, int64_t *steps, int pivot, const skh_opts *opts);
/* device variant (dA, dR, dp, dsteps in HBM; n, offset host arrays): returns once the members with
 * n >= 512 are done and the control arrays of the rest are uploaded; those finish on `stream`. */
int skh_skewchol_d_batched_device(int64_t batch, const int64_t *n, const int64_t *offset, const double *dA,
                                  double *dR, int64_t *dp, int64_t *dsteps, int pivot, void *stream,
                                  const skh_opts *opts);

/* device variants: dA (16-byte aligned) and the outputs in HBM; n and offset are host arrays.
 * Enqueued on `stream`; the call returns once the members with even n >= 512 are done and the
 * control arrays of the rest are uploaded. */
int skh_pfaffian_z_batched_device(int64_t batch, const int64_t *n, const int64_t *offset, const double *dA,
                                  double *dpf, void *stream, const skh_opts *opts);
int skh_logabspfaffian_z_batched_device(int64_t batch, const int64_t *n, const int64_t *offset, const double *dA,
                                        double *dlogpf, double *dphase, void *stream, const skh_opts *opts);

/*
 * Eigen-decomposition of a real skew-symmetric A (host memory, A not modified).
 * Replaces skeweigen!/eigen!(A::SkewHermitian{<:Real}) (reference src/eigen.jl:75-97): GPU
 * tridiagonalization and Q, host skewtrieigen_divided!, GPU back-transform Q*Qr, Q*Qim.
 * vals: imaginary parts of the n eigenvalues (reference order); vr, vi: real and imaginary
 * parts of the eigenvector matrix (n x n, column-major, leading dimension n).
 */
int skh_eigen_d(int64_t n, const double *A, int64_t lda, double *vals, double *vr, double *vi, const skh_opts *opts);

/*
 * Host implicit-QR eigensolver of the real skew tridiagonal SkewHermTridiagonal(ev), n = size.
 * Replaces skewtrieigvals! (reference src/skeweigen.jl:105-156, backup :463-468): vals receives
 * the imaginary parts of the eigenvalues in the reference's order; ev (n-1) is overwritten.
 */
int skh_skewtrieigvals_d(int64_t n, double *ev, double *vals);
/*
 * Eigenvalues by index or value range.  by_value = 0: the eigenvalues of
 * H = SymTridiagonal(0, eta) with 1-based indices il..iu of its ascending spectrum; by_value = 1:
 * those in (-vh, -vl]; each negated, so the eigenvalues are complex(0, vals[k]), k < *m.
 * skh_skewtrieigvals_range_d replaces skewtrieigvals!(S::SkewHermTridiagonal{<:Real}, irange) /
 * (S, vl, vh) (reference src/tridiag.jl:525-537, via eigvals! :452-460); skh_eigvals_range_d
 * replaces eigvals!(A::SkewHermitian{<:Real}, irange) / (A, vl, vh) (src/eigen.jl:8-16, 57-73:
 * GPU tridiagonalization first; n == 1 returns the single entry unless (vl, vh) excludes it).
 */
int skh_skewtrieigvals_range_d(int64_t n, const double *ev, int64_t il, int64_t iu, double vl, double vh, int by_value,
                               double *vals, int64_t *m);
int skh_eigvals_range_d(int64_t n, const double *A, int64_t lda, int64_t il, int64_t iu, double vl, double vh,
                        int by_value, double *vals, int64_t *m, const skh_opts *opts);
/*
 * Eigenvectors as well: variant 0 = skewtrieigen_divided! (src/skeweigen.jl:322-457, used by
 * eigen(::SkewHermitian)), 1 = skewtrieigen_merged! (:195-320, used by eigen(::SkewHermTridiagonal)).
 * vr, vi: n x n real and imaginary parts of the eigenvector matrix (column-major).
 */
int skh_skewtrieigen_d(int64_t n, double *ev, double *vals, double *vr, double *vi, int variant);
/* The same with the sweeps' Givens rotations applied to the eigenvector halves Qodd / Qeven on the
 * GPU (one thread per row, the chain's carry in a register) instead of on the host: the O(n^3)
 * part of skewtrieigen_divided!/merged! (src/skeweigen.jl:158-193, :246-262, :376-392).  Host
 * arrays in and out; results bit-identical to skh_skewtrieigen_d.  skh_eigen_d takes this path
 * from n = 256. */
int skh_skewtrieigen_d_gpu(int64_t n, double *ev, double *vals, double *vr, double *vi, int variant);
/* Complex SkewHermTridiagonal(ev, dvim) (ev: n - 1 interleaved complex values, dvim: n reals or NULL):
 * to_symtridiagonal (src/tridiag.jl:558-583) makes i A = Q S Q^H with S real symmetric tridiagonal.
 * skh_skewtrieigvals_z: mu (n) = the eigenvalues of S ascending, by a host QL (eigvals!, src/tridiag.jl:
 * 462-468: the eigenvalues of A are -i mu).  skh_skewtrieigen_z: the same mu by rocSOLVER's divide and
 * conquer on the GPU, and V (n x n interleaved complex) = Diagonal(Q) Z, Z the eigenvectors of S
 * (eigen!, src/tridiag.jl:508-515: Eigen(-i mu, V)). */
int skh_skewtrieigvals_z(int64_t n, const double *ev, const double *dvim, double *mu);
int skh_skewtrieigen_z(int64_t n, const double *ev, const double *dvim, double *mu, double *V);

/* ---- matrix functions (src/exp.jl:3-296) -------------------------------------------------------
 * f(A) of a skew matrix from its eigen-decomposition, with the eigenvectors kept on the GPU: their
 * columns scaled by a function of the eigenvalues, the products V D V^T / V D V^H in rocBLAS, the
 * divisions of tan / cot / tanh / coth by LU in rocSOLVER, and the wrappers the reference applies
 * (skewhermitian!, Hermitian, hermitian!) made dense.  fn selects the function: */
#define SKH_FN_EXP 0     /* exp(A), a plain matrix (src/exp.jl:3-48) */
#define SKH_FN_LOG 1     /* log(A); real: even n and no zero eigenvalue, else SKH_ERR_DOMAIN (:50-97) */
#define SKH_FN_CIS 2     /* cis(A) = exp(iA), Hermitian, always complex output (:99-121, :204) */
#define SKH_FN_COS 3     /* cos(A), Hermitian (:123-149) */
#define SKH_FN_SIN 4     /* sin(A), skew-Hermitian (:151-202) */
#define SKH_FN_SINCOS 5  /* out1 = sin(A), out2 = cos(A) (:208-256) */
#define SKH_FN_SINH 6    /* skewhermitian!(exp(A)) (:257) */
#define SKH_FN_COSH 7    /* hermitian!(exp(A)); complex: its conjugate (:258-263) */
#define SKH_FN_TAN 8     /* skewhermitian!(cos(A) \ sin(A)) (:265-268) */
#define SKH_FN_TANH 9    /* skewhermitian!((E + I) \ (E - I)), E = exp(2A) (:270-273) */
#define SKH_FN_COT 10    /* skewhermitian!(sin(A) \ cos(A)) (:275-278) */
#define SKH_FN_COTH 11   /* skewhermitian!((E - I) \ (E + I)) (:280-284) */
/* Real skew A (n x n, lda; host).  out1 (and out2 for SKH_FN_SINCOS): n x n, column-major, ld n;
 * real except for SKH_FN_CIS, whose out1 holds n x n interleaved complex values (2 n^2 doubles).
 * Replaces exp/log/cis/cos/sin/sincos/sinh/cosh/tan/tanh/cot/coth(::SkewHermitian{<:Real}),
 * src/exp.jl:48, :97, :204-206, :256-284. */
int skh_skewfun_d(int64_t n, const double *A, int64_t lda, int fn, double *out1, double *out2,
                  const skh_opts *opts);
/* The same for SkewHermTridiagonal{<:Real}(ev) (ev: n - 1 values), through eigen! of the
 * tridiagonal (skewtrieigen_merged!, src/tridiag.jl:506). */
int skh_skewfun_tri_d(int64_t n, const double *ev, int fn, double *out1, double *out2);
/* Complex skew-Hermitian A (interleaved pairs, lda in complex elements); every output interleaved
 * complex n x n (src/exp.jl:37-46, :86-95, :112-121, :137-149, :181-202, :235-255). */
int skh_skewfun_z(int64_t n, const double *A, int64_t lda, int fn, double *out1, double *out2,
                  const skh_opts *opts);

/* Diagnostics: per-kernel times of the last skh_sytrd_d_device call made with SKH_FLAG_TIMING:
 * out10 = {panel ms, panel launches, rank-2k ms, rank-2k launches, panel algorithmic bytes,
 *          rank-2k algorithmic flops, then the first four phase-clock values of
 *          skh_diag_phase_clock (zero unless SKH_FLAG_PHASES was set)}.
 * Waits for the recorded events. */
int skh_diag_timings(double *out10);
/* Diagnostics: the panel kernel's phase clock of the last SKH_FLAG_PHASES factorization, in ms
 * (workgroup 0, summed over its launches): [0] each launch's first row pass, [1] barrier arrivals,
 * [2] phase B (with the barrier wait), [3] row passes (with the wait for the matvec partials), then
 * the row pass split into [4] loads and that wait, [5] correction of w, [6] column update and
 * partials, [7] hand-over, and phase B into [11] first tile, [12] barrier wait, [8] loads and
 * scalars, [9] t, [10] matvec.  count <= 13 values are written. */
int skh_diag_phase_clock(double *out, int count);
/* Diagnostics: v_mfma_f64_16x16x4f64 lane-layout probe (768 doubles, see the test suite). */
int skh_diag_mfma_probe(double *out768);

#ifdef __cplusplus
}
#endif
#endif
