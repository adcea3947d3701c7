/*
 * skew_oracle.c -- CPU restatement of SkewLinearAlgebra.jl's skew tridiagonalization path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the GPU product is checked
 * against (pytest under tests/, the smoke check and the cpu_baseline leg of bench.py).
 * Nothing in the shipped library links, loads or calls it.
 *
 * Every routine follows the reference algorithm step by step (same panel width policy,
 * same BLAS call shapes, same serial skew-combine loop, same deflation rules) so that
 * (a) its results can be pinned against the reference's own fixtures and LAPACK, and
 * (b) its run time is an honest CPU baseline for the reference's method.
 *
 * BLAS: when orc_blas_load() is given the path of an ILP64 OpenBLAS built with the
 * "scipy_" symbol prefix (numpy's bundled copy), gemv/gemm/nrm2/dstev go there (threaded,
 * as Julia's libblastrampoline -> OpenBLAS would).  Otherwise plain C loops are used.
 *
 * Layout: column-major, 0-based, leading dimension lda.  Complex numbers are C99
 * double _Complex (same memory layout as Julia's ComplexF64).
 */
#include <complex.h>
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex zc;
typedef int64_t i64;

/* ------------------------------------------------------------------------- */
/* BLAS dispatch                                                              */
/* ------------------------------------------------------------------------- */
typedef void (*dgemv_fn)(const char *, const i64 *, const i64 *, const double *, const double *,
                         const i64 *, const double *, const i64 *, const double *, double *,
                         const i64 *, size_t);
typedef void (*dgemm_fn)(const char *, const char *, const i64 *, const i64 *, const i64 *,
                         const double *, const double *, const i64 *, const double *, const i64 *,
                         const double *, double *, const i64 *, size_t, size_t);
typedef double (*dnrm2_fn)(const i64 *, const double *, const i64 *);
typedef void (*zgemv_fn)(const char *, const i64 *, const i64 *, const zc *, const zc *, const i64 *,
                         const zc *, const i64 *, const zc *, zc *, const i64 *, size_t);
typedef void (*zgemm_fn)(const char *, const char *, const i64 *, const i64 *, const i64 *,
                         const zc *, const zc *, const i64 *, const zc *, const i64 *, const zc *,
                         zc *, const i64 *, size_t, size_t);
typedef double (*dznrm2_fn)(const i64 *, const zc *, const i64 *);
typedef void (*dstev_fn)(const char *, const i64 *, double *, double *, double *, const i64 *,
                         double *, i64 *, size_t);
typedef void (*setthreads_fn)(int);
typedef int (*getthreads_fn)(void);

static struct {
    void *h;
    dgemv_fn dgemv;
    dgemm_fn dgemm;
    dnrm2_fn dnrm2;
    zgemv_fn zgemv;
    zgemm_fn zgemm;
    dznrm2_fn dznrm2;
    dstev_fn dstev;
    setthreads_fn setthreads;
    getthreads_fn getthreads;
} B;

int orc_blas_load(const char *path)
{
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    B.dgemv = (dgemv_fn)dlsym(h, "scipy_dgemv_64_");
    B.dgemm = (dgemm_fn)dlsym(h, "scipy_dgemm_64_");
    B.dnrm2 = (dnrm2_fn)dlsym(h, "scipy_dnrm2_64_");
    B.zgemv = (zgemv_fn)dlsym(h, "scipy_zgemv_64_");
    B.zgemm = (zgemm_fn)dlsym(h, "scipy_zgemm_64_");
    B.dznrm2 = (dznrm2_fn)dlsym(h, "scipy_dznrm2_64_");
    B.dstev = (dstev_fn)dlsym(h, "scipy_dstev_64_");
    B.setthreads = (setthreads_fn)dlsym(h, "scipy_openblas_set_num_threads64_");
    B.getthreads = (getthreads_fn)dlsym(h, "scipy_openblas_get_num_threads64_");
    if (!B.dgemv || !B.dgemm || !B.dnrm2 || !B.zgemv || !B.zgemm || !B.dznrm2) {
        memset(&B, 0, sizeof(B));
        dlclose(h);
        return -2;
    }
    B.h = h;
    return 0;
}

void orc_blas_unload(void) { memset(&B, 0, sizeof(B)); }
int orc_blas_loaded(void) { return B.h != NULL; }
void orc_blas_set_threads(int t) { if (B.setthreads) B.setthreads(t); }
int orc_blas_get_threads(void) { return B.getthreads ? B.getthreads() : 1; }

/* y = alpha*op(A)*x + beta*y, op = 'N' | 'T' */
static void dgemv(char tr, i64 m, i64 n, double alpha, const double *A, i64 lda, const double *x,
                  i64 incx, double beta, double *y, i64 incy)
{
    i64 ylen = (tr == 'N') ? m : n, xlen = (tr == 'N') ? n : m;
    if (ylen <= 0) return;
    if (B.dgemv && m > 0 && n > 0) {
        B.dgemv(&tr, &m, &n, &alpha, A, &lda, x, &incx, &beta, y, &incy, 1);
        return;
    }
    for (i64 i = 0; i < ylen; i++) y[i * incy] = (beta == 0.0) ? 0.0 : beta * y[i * incy];
    if (xlen <= 0) return;
    if (tr == 'N') {
        for (i64 j = 0; j < n; j++) {
            double t = alpha * x[j * incx];
            for (i64 i = 0; i < m; i++) y[i * incy] += t * A[i + j * lda];
        }
    } else {
        for (i64 j = 0; j < n; j++) {
            double s = 0.0;
            for (i64 i = 0; i < m; i++) s += A[i + j * lda] * x[i * incx];
            y[j * incy] += alpha * s;
        }
    }
}

/* C = alpha*A*B^T + beta*C (only the 'N','T' shape is needed) */
static void dgemm_nt(i64 m, i64 n, i64 k, double alpha, const double *A, i64 lda, const double *Bm,
                     i64 ldb, double beta, double *C, i64 ldc)
{
    if (m <= 0 || n <= 0) return;
    if (B.dgemm && k > 0) {
        char ta = 'N', tb = 'T';
        B.dgemm(&ta, &tb, &m, &n, &k, &alpha, A, &lda, Bm, &ldb, &beta, C, &ldc, 1, 1);
        return;
    }
    for (i64 j = 0; j < n; j++) {
        for (i64 i = 0; i < m; i++) C[i + j * ldc] = (beta == 0.0) ? 0.0 : beta * C[i + j * ldc];
        for (i64 l = 0; l < k; l++) {
            double t = alpha * Bm[j + l * ldb];
            for (i64 i = 0; i < m; i++) C[i + j * ldc] += t * A[i + l * lda];
        }
    }
}

static double dnrm2(i64 n, const double *x, i64 incx)
{
    if (n <= 0) return 0.0;
    if (B.dnrm2) return B.dnrm2(&n, x, &incx);
    /* scaled sum of squares (overflow-safe), as reference BLAS nrm2 */
    double scale = 0.0, ssq = 1.0;
    for (i64 i = 0; i < n; i++) {
        double a = fabs(x[i * incx]);
        if (a != 0.0) {
            if (scale < a) {
                ssq = 1.0 + ssq * (scale / a) * (scale / a);
                scale = a;
            } else {
                ssq += (a / scale) * (a / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

/* complex: op = 'N' | 'C' */
static void zgemv(char tr, i64 m, i64 n, zc alpha, const zc *A, i64 lda, const zc *x, i64 incx,
                  zc beta, zc *y, i64 incy)
{
    i64 ylen = (tr == 'N') ? m : n, xlen = (tr == 'N') ? n : m;
    if (ylen <= 0) return;
    if (B.zgemv && m > 0 && n > 0) {
        B.zgemv(&tr, &m, &n, &alpha, A, &lda, x, &incx, &beta, y, &incy, 1);
        return;
    }
    for (i64 i = 0; i < ylen; i++) y[i * incy] = (beta == 0.0) ? 0.0 : beta * y[i * incy];
    if (xlen <= 0) return;
    if (tr == 'N') {
        for (i64 j = 0; j < n; j++) {
            zc t = alpha * x[j * incx];
            for (i64 i = 0; i < m; i++) y[i * incy] += t * A[i + j * lda];
        }
    } else {
        for (i64 j = 0; j < n; j++) {
            zc s = 0.0;
            for (i64 i = 0; i < m; i++) s += conj(A[i + j * lda]) * x[i * incx];
            y[j * incy] += alpha * s;
        }
    }
}

/* C = alpha*A*B^H + beta*C */
static void zgemm_nc(i64 m, i64 n, i64 k, zc alpha, const zc *A, i64 lda, const zc *Bm, i64 ldb,
                     zc beta, zc *C, i64 ldc)
{
    if (m <= 0 || n <= 0) return;
    if (B.zgemm && k > 0) {
        char ta = 'N', tb = 'C';
        B.zgemm(&ta, &tb, &m, &n, &k, &alpha, A, &lda, Bm, &ldb, &beta, C, &ldc, 1, 1);
        return;
    }
    for (i64 j = 0; j < n; j++) {
        for (i64 i = 0; i < m; i++) C[i + j * ldc] = (beta == 0.0) ? 0.0 : beta * C[i + j * ldc];
        for (i64 l = 0; l < k; l++) {
            zc t = alpha * conj(Bm[j + l * ldb]);
            for (i64 i = 0; i < m; i++) C[i + j * ldc] += t * A[i + l * lda];
        }
    }
}

static double dznrm2(i64 n, const zc *x, i64 incx)
{
    if (n <= 0) return 0.0;
    if (B.dznrm2) return B.dznrm2(&n, x, &incx);
    double scale = 0.0, ssq = 1.0;
    for (i64 i = 0; i < n; i++) {
        double parts[2] = {creal(x[i * incx]), cimag(x[i * incx])};
        for (int p = 0; p < 2; p++) {
            double a = fabs(parts[p]);
            if (a != 0.0) {
                if (scale < a) {
                    ssq = 1.0 + ssq * (scale / a) * (scale / a);
                    scale = a;
                } else {
                    ssq += (a / scale) * (a / scale);
                }
            }
        }
    }
    return scale * sqrt(ssq);
}

/* ------------------------------------------------------------------------- */
/* setnb: src/hessenberg.jl:133-142                                            */
/* ------------------------------------------------------------------------- */
i64 orc_setnb(i64 n)
{
    if (n <= 12) return (n - 4 > 1) ? n - 4 : 1;
    if (n <= 100) return 10;
    return 60;
}

/* ------------------------------------------------------------------------- */
/* householder!: src/hessenberg.jl:26-44                                       */
/* x has length n (n >= 1); returns tau, beta; x becomes v with v[0] = 1.       */
/* ------------------------------------------------------------------------- */
static void householder_d(double *x, i64 n, double *tau, double *beta)
{
    if (n == 1) { /* real: no final 1x1 reflection (hessenberg.jl:27-29) */
        *tau = 0.0;
        *beta = x[0];
        return;
    }
    double xnorm = dnrm2(n - 1, x + 1, 1);
    double alpha = x[0];
    if (xnorm != 0.0 || alpha != 0.0) {
        double b = ((alpha > 0.0) ? -1.0 : 1.0) * hypot(alpha, xnorm);
        *tau = 1.0 - alpha / b;
        double s = 1.0 / (alpha - b);
        x[0] = 1.0;
        for (i64 k = 1; k < n; k++) x[k] *= s;
        *beta = b;
    } else {
        *tau = 0.0;
        for (i64 k = 0; k < n; k++) x[k] = 0.0;
        *beta = 0.0;
    }
}

static void householder_z(zc *x, i64 n, zc *tau, double *beta)
{
    double xnorm = dznrm2(n - 1, x + 1, 1);
    zc alpha = x[0];
    if (xnorm != 0.0 || alpha != 0.0) {
        /* hypot(alpha::Complex, xnorm) = sqrt(|alpha|^2 + xnorm^2), computed safely */
        double b = ((creal(alpha) > 0.0) ? -1.0 : 1.0) * hypot(cabs(alpha), xnorm);
        *tau = 1.0 - alpha / b;
        zc s = 1.0 / (alpha - b);
        x[0] = 1.0;
        for (i64 k = 1; k < n; k++) x[k] *= s;
        *beta = b;
    } else {
        *tau = 0.0;
        for (i64 k = 0; k < n; k++) x[k] = 0.0;
        *beta = 0.0;
    }
}

/* ------------------------------------------------------------------------- */
/* skewhess!: src/hessenberg.jl:70-90 (with lefthouseholder!/ger2! :46-68)      */
/* A is the nt x nt tail block (lda), tau/eta have nt-1 entries.                */
/* ------------------------------------------------------------------------- */
static void skewhess_d(i64 nt, double *A, i64 lda, double *tau, double *eta)
{
    double *atmp = (double *)malloc(sizeof(double) * (nt > 0 ? nt : 1));
    for (i64 i = 0; i < nt - 1; i++) {
        double ts, al;
        double *v = &A[(i + 1) + i * lda];
        i64 m = nt - i - 1;
        householder_d(v, m, &ts, &al);
        eta[i] = al;
        double *A22 = &A[(i + 1) + (i + 1) * lda];
        double *s = atmp + i + 1;
        /* lefthouseholder!: s = A22^T v ; A22 += (-tau) v s^T  (BLAS ger skips alpha == 0) */
        dgemv('T', m, m, 1.0, A22, lda, v, 1, 0.0, s, 1);
        if (ts != 0.0) {
            for (i64 c = 0; c < m; c++) {
                double t = -ts * s[c];
                for (i64 r = 0; r < m; r++) A22[r + c * lda] += v[r] * t;
            }
        }
        /* s = A22 v ; lower triangle -= tau s v^T, mirror to the upper triangle */
        dgemv('N', m, m, 1.0, A22, lda, v, 1, 0.0, s, 1);
        for (i64 j = 0; j < m; j++) {
            A22[j + j * lda] -= ts * s[j] * v[j];
            for (i64 k = j + 1; k < m; k++) {
                A22[k + j * lda] -= ts * s[k] * v[j];
                A22[j + k * lda] = -A22[k + j * lda];
            }
        }
        tau[i] = ts;
    }
    free(atmp);
}

static void skewhess_z(i64 nt, zc *A, i64 lda, zc *tau, double *eta)
{
    zc *atmp = (zc *)malloc(sizeof(zc) * (nt > 0 ? nt : 1));
    for (i64 i = 0; i < nt - 1; i++) {
        zc ts;
        double al;
        zc *v = &A[(i + 1) + i * lda];
        i64 m = nt - i - 1;
        householder_z(v, m, &ts, &al);
        eta[i] = al;
        zc *A22 = &A[(i + 1) + (i + 1) * lda];
        zc *s = atmp + i + 1;
        /* s = A22^H v ; A22 += (-conj(tau)) v s^H  (zgerc) */
        zgemv('C', m, m, 1.0, A22, lda, v, 1, 0.0, s, 1);
        zc t2 = -conj(ts);
        if (t2 != 0.0) {
            for (i64 c = 0; c < m; c++) {
                zc t = t2 * conj(s[c]);
                for (i64 r = 0; r < m; r++) A22[r + c * lda] += v[r] * t;
            }
        }
        zgemv('N', m, m, 1.0, A22, lda, v, 1, 0.0, s, 1);
        for (i64 j = 0; j < m; j++) {
            A22[j + j * lda] -= ts * s[j] * conj(v[j]);
            for (i64 k = j + 1; k < m; k++) {
                A22[k + j * lda] -= ts * s[k] * conj(v[j]);
                A22[j + k * lda] = -conj(A22[k + j * lda]);
            }
        }
        tau[i] = ts;
    }
    free(atmp);
}

/* ------------------------------------------------------------------------- */
/* skewlatrd!: src/hessenberg.jl:92-131 (real branch)                          */
/* At: the s x s panel view A[i0:n, i0:n]; W: s x nb (ldw)                      */
/* ------------------------------------------------------------------------- */
static void skewlatrd_d(double *At, i64 lda, double *eta, double *W, i64 ldw, double *tau, i64 s,
                        i64 nb)
{
    for (i64 i = 0; i < nb; i++) {
        if (i > 0) {
            dgemv('N', s - i, i, 1.0, &At[i], lda, &W[i], ldw, 1.0, &At[i + i * lda], 1);
            dgemv('N', s - i, i, -1.0, &W[i], ldw, &At[i], lda, 1.0, &At[i + i * lda], 1);
        }
        double ts, al;
        double *v = &At[(i + 1) + i * lda];
        i64 m = s - i - 1;
        householder_d(v, m, &ts, &al);
        eta[i] = al;
        double *w = &W[(i + 1) + i * ldw];
        dgemv('N', m, m, 1.0, &At[(i + 1) + (i + 1) * lda], lda, v, 1, 0.0, w, 1);
        if (i > 0) {
            double *t = &W[i * ldw];
            dgemv('T', m, i, 1.0, &W[i + 1], ldw, v, 1, 0.0, t, 1);
            dgemv('N', m, i, 1.0, &At[i + 1], lda, t, 1, 1.0, w, 1);
            dgemv('T', m, i, 1.0, &At[i + 1], lda, v, 1, 0.0, t, 1);
            dgemv('N', m, i, -1.0, &W[i + 1], ldw, t, 1, 1.0, w, 1);
        }
        for (i64 k = 0; k < m; k++) w[k] *= ts;
        tau[i] = ts;
    }
}

/* complex branch of skewlatrd!: src/hessenberg.jl:96-104,123-126 */
static void skewlatrd_z(zc *At, i64 lda, double *eta, zc *W, i64 ldw, zc *tau, zc *tempconj, i64 s,
                        i64 nb)
{
    for (i64 i = 0; i < nb; i++) {
        if (i > 0) {
            for (i64 j = 0; j < i; j++) tempconj[j] = conj(W[i + j * ldw]);
            zgemv('N', s - i, i, 1.0, &At[i], lda, tempconj, 1, 1.0, &At[i + i * lda], 1);
            for (i64 j = 0; j < i; j++) tempconj[j] = conj(At[i + j * lda]);
            zgemv('N', s - i, i, -1.0, &W[i], ldw, tempconj, 1, 1.0, &At[i + i * lda], 1);
        }
        zc ts;
        double al;
        zc *v = &At[(i + 1) + i * lda];
        i64 m = s - i - 1;
        householder_z(v, m, &ts, &al);
        eta[i] = al;
        zc *w = &W[(i + 1) + i * ldw];
        zgemv('N', m, m, 1.0, &At[(i + 1) + (i + 1) * lda], lda, v, 1, 0.0, w, 1);
        if (i > 0) {
            zc *t = &W[i * ldw];
            zgemv('C', m, i, 1.0, &W[i + 1], ldw, v, 1, 0.0, t, 1);
            zgemv('N', m, i, 1.0, &At[i + 1], lda, t, 1, 1.0, w, 1);
            zgemv('C', m, i, 1.0, &At[i + 1], lda, v, 1, 0.0, t, 1);
            zgemv('N', m, i, -1.0, &W[i + 1], ldw, t, 1, 1.0, w, 1);
        }
        for (i64 k = 0; k < m; k++) w[k] *= ts;
        /* alpha = tau * dot(w, v) / 2, dot(x,y) = x^H y */
        zc d = 0.0;
        for (i64 k = 0; k < m; k++) d += conj(w[k]) * v[k];
        zc a = ts * d / 2.0;
        for (i64 k = 0; k < m; k++) w[k] += a * v[k];
        tau[i] = ts;
    }
}

/* ------------------------------------------------------------------------- */
/* skewblockedhess!: src/hessenberg.jl:144-174                                 */
/* A (n x n, lda) is overwritten by the 'L'-layout factors; tau, eta: n-1.      */
/* nb_override > 0 replaces setnb (used only to exercise other blockings).      */
/* ------------------------------------------------------------------------- */
void orc_skewblockedhess_d(i64 n, double *A, i64 lda, double *tau, double *eta, i64 nb_override)
{
    if (n <= 1) return;
    i64 nb = nb_override > 0 ? nb_override : orc_setnb(n);
    double *W = (double *)calloc((size_t)n * nb, sizeof(double));
    i64 ldo = n - nb > 0 ? n - nb : 1;
    double *Om = (double *)malloc(sizeof(double) * (size_t)ldo * ldo);
    i64 oldi = -1; /* 0-based start of the last panel */
    for (i64 i = 0; i <= n - nb - 2; i += nb) {
        i64 s = n - i;
        double *At = &A[i + i * lda];
        skewlatrd_d(At, lda, eta + i, W, n, tau + i, s, nb);
        i64 mo = n - nb - i;
        dgemm_nt(mo, mo, nb, 1.0, &A[(i + nb) + i * lda], lda, &W[nb], n, 0.0, Om, ldo);
        i64 sg = i + nb; /* first row/col of the trailing block */
        for (i64 k = 0; k < mo; k++) {
            A[(sg + k) + (sg + k) * lda] += Om[k + k * ldo] - Om[k + k * ldo];
            for (i64 j = k + 1; j < mo; j++) {
                A[(sg + j) + (sg + k) * lda] += Om[j + k * ldo] - Om[k + j * ldo];
                A[(sg + k) + (sg + j) * lda] = -A[(sg + j) + (sg + k) * lda];
            }
        }
        oldi = i;
    }
    /* reference: oldi(1-based) += nb; the tail runs when oldi < n.  0-based: t0 = oldi-1. */
    i64 t0 = (oldi < 0) ? nb - 1 : oldi + nb;
    if (t0 + 1 < n) skewhess_d(n - t0, &A[t0 + t0 * lda], lda, tau + t0, eta + t0);
    free(W);
    free(Om);
}

void orc_skewblockedhess_z(i64 n, zc *A, i64 lda, zc *tau, double *eta, i64 nb_override)
{
    if (n <= 1) return;
    i64 nb = nb_override > 0 ? nb_override : orc_setnb(n);
    zc *W = (zc *)calloc((size_t)n * nb, sizeof(zc));
    zc *tempconj = (zc *)calloc((size_t)nb, sizeof(zc));
    i64 ldo = n - nb > 0 ? n - nb : 1;
    zc *Om = (zc *)malloc(sizeof(zc) * (size_t)ldo * ldo);
    i64 oldi = -1;
    for (i64 i = 0; i <= n - nb - 2; i += nb) {
        i64 s = n - i;
        zc *At = &A[i + i * lda];
        skewlatrd_z(At, lda, eta + i, W, n, tau + i, tempconj, s, nb);
        i64 mo = n - nb - i;
        zgemm_nc(mo, mo, nb, 1.0, &A[(i + nb) + i * lda], lda, &W[nb], n, 0.0, Om, ldo);
        i64 sg = i + nb;
        for (i64 k = 0; k < mo; k++) {
            A[(sg + k) + (sg + k) * lda] += Om[k + k * ldo] - conj(Om[k + k * ldo]);
            for (i64 j = k + 1; j < mo; j++) {
                A[(sg + j) + (sg + k) * lda] += Om[j + k * ldo] - conj(Om[k + j * ldo]);
                A[(sg + k) + (sg + j) * lda] = -conj(A[(sg + j) + (sg + k) * lda]);
            }
        }
        oldi = i;
    }
    i64 t0 = (oldi < 0) ? nb - 1 : oldi + nb;
    if (t0 + 1 < n) skewhess_z(n - t0, &A[t0 + t0 * lda], lda, tau + t0, eta + t0);
    free(W);
    free(tempconj);
    free(Om);
}

/* ------------------------------------------------------------------------- */
/* Host implicit QR on the real skew tridiagonal: src/skeweigen.jl             */
/* The helpers below use 1-based accessors so each line mirrors the reference. */
/* ------------------------------------------------------------------------- */
#define E1(k) ev[(k) - 1]
#define G1(k) G[(k) - 1]

/* getgivens: src/skeweigen.jl:3-7 */
static void getgivens(double a, double b, double *c, double *s)
{
    double nm = hypot(a, b);
    if (nm == 0.0) {
        *c = 1.0;
        *s = 0.0;
        return;
    }
    *c = a / nm;
    *s = b / nm;
}

/* reducetozero: src/skeweigen.jl:11-47 (n even, ev has n entries) */
static void reducetozero(double *ev, double *G, i64 n)
{
    if (n == 0) return;
    double bulge = 0.0, al, be, ga, c, s;
    if (n > 2) {
        al = E1(n - 2);
        be = E1(n - 1);
        ga = E1(n);
        getgivens(-be, ga, &c, &s);
        G1(n - 1) = c;
        G1(n) = -s;
        E1(n - 2) *= c;
        E1(n - 1) = c * be - s * ga;
        E1(n) = 0.0;
        bulge = -s * al;
        for (i64 i = n - 2; i >= 4; i -= 2) {
            al = E1(i - 2);
            be = E1(i - 1);
            getgivens(-be, bulge, &c, &s);
            G1(i - 1) = c;
            G1(i) = -s;
            E1(i - 2) *= c;
            E1(i - 1) = c * be - s * bulge;
            bulge = -s * al;
        }
    }
    be = (n == 2) ? E1(2) : bulge;
    al = E1(1);
    getgivens(-al, be, &c, &s);
    G1(1) = c;
    G1(2) = -s;
    E1(1) = c * al - s * be;
    if (n == 2) E1(2) = s * al + be * c;
}

/* implicitstep_novec: src/skeweigen.jl:73-103.  n = last active index of ev. */
static i64 implicitstep_novec(double *ev, i64 n, i64 start)
{
    double bulge = 0.0;
    double shift = E1(n) * E1(n); /* getshift(ev[n-1:n]) */
    double tol = 2.220446049250313e-16;
    for (i64 i = start; i <= n - 1; i++) {
        double al = (i > start) ? E1(i - 1) : 0.0;
        double be = E1(i), ga = E1(i + 1);
        double x1 = -al * al - be * be + shift;
        double x2 = -al * bulge + be * ga;
        double c, s;
        if (i > start) getgivens(al, bulge, &c, &s);
        else getgivens(x1, x2, &c, &s);
        if (i > 1) E1(i - 1) = c * al + s * bulge;
        E1(i) = c * be - s * ga;
        E1(i + 1) = s * be + c * ga;
        if (i < n - 1) {
            double ze = E1(i + 2);
            E1(i + 2) *= c;
            bulge = s * ze;
            if (fabs(bulge) < tol && fabs(E1(i)) < tol) return i + 1;
        }
    }
    return start;
}

/* Symmetric tridiagonal (zero diagonal, off-diagonal e) eigenvalues, ascending.
 * LAPACK dstev when BLAS is loaded (the reference calls LAPACK here); otherwise
 * Sturm-sequence bisection. */
static int symtri_eigvals(i64 n, const double *e, double *w)
{
    if (n <= 0) return 0;
    if (B.dstev) {
        double *d = (double *)calloc((size_t)n, sizeof(double));
        double *ee = (double *)malloc(sizeof(double) * (size_t)(n > 1 ? n - 1 : 1));
        for (i64 k = 0; k < n - 1; k++) ee[k] = e[k];
        char jobz = 'N';
        i64 ldz = 1, info = 0;
        double z = 0, work = 0;
        B.dstev(&jobz, &n, d, ee, &z, &ldz, &work, &info, 1);
        for (i64 k = 0; k < n; k++) w[k] = d[k];
        free(d);
        free(ee);
        return (int)info;
    }
    double bnd = 0.0;
    for (i64 k = 0; k < n - 1; k++) bnd = fmax(bnd, fabs(e[k]));
    bnd = 2.0 * bnd + 1e-300;
    for (i64 idx = 0; idx < n; idx++) {
        double lo = -bnd, hi = bnd;
        for (int it = 0; it < 200; it++) {
            double mid = 0.5 * (lo + hi);
            /* count eigenvalues < mid */
            i64 cnt = 0;
            double q = -mid;
            if (q < 0) cnt++;
            for (i64 k = 1; k < n; k++) {
                if (q == 0.0) q = 1e-300;
                q = -mid - e[k - 1] * e[k - 1] / q;
                if (q < 0) cnt++;
            }
            if (cnt > idx) hi = mid;
            else lo = mid;
            if (hi - lo <= 4.440892098500626e-16 * fmax(fabs(lo), fabs(hi))) break;
        }
        w[idx] = 0.5 * (lo + hi);
    }
    return 0;
}

/* skewtrieigvals!: src/skeweigen.jl:105-156 (+ backup :463-468).
 * N = matrix size; ev (N-1) is overwritten; vals receives the imaginary parts of the
 * eigenvalues in the reference's order.  Returns 0, or 1 when max_iter is hit. */
int orc_skewtrieigvals_d(i64 N, double *ev, double *vals)
{
    for (i64 k = 0; k < N; k++) vals[k] = 0.0;
    i64 n = N;
    double *G = NULL;
    if (n % 2 == 1) {
        n -= 1;
        G = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        reducetozero(ev, G, n);
    }
    double tol = 2.220446049250313e-16 * 10.0;
    i64 max_iter = 30 * n, iter = 0, count_static = 0, start = 1;
    double mem = 1.0;
    while (n > 2 && iter < max_iter) {
        start = implicitstep_novec(ev, n - 1, start);
        while (n > 2 && fabs(E1(n - 1)) <= tol) {
            vals[n - 1] = 0.0;
            n -= 1;
        }
        while (n > 2 && fabs(E1(n - 2)) <= tol * fabs(E1(n - 1))) {
            vals[n - 2] = E1(n - 1); /* eigofblock: (+i*k, -i*k) */
            vals[n - 1] = -E1(n - 1);
            n -= 2;
        }
        if (start > n - 2) start = 1;
        if (n > 1 && fabs(mem - E1(n - 1)) < 0.0001 * fabs(E1(n - 1))) {
            count_static += 1;
            if (count_static > 4) {
                double *w = (double *)malloc(sizeof(double) * (size_t)n);
                symtri_eigvals(n, ev, w);
                for (i64 k = 0; k < n; k++) vals[k] = -w[k];
                free(w);
                free(G);
                return 0;
            }
        } else {
            count_static = 0;
        }
        mem = (n > 1) ? E1(n - 1) : 0.0;
        iter += 1;
    }
    free(G);
    if (n == 2) {
        vals[0] = E1(1);
        vals[1] = -E1(1);
        return 0;
    } else if (n == 1) {
        vals[0] = 0.0;
        return 0;
    } else if (n == 0) {
        return 0;
    }
    return 1;
}

/* implicitstep_vec!: src/skeweigen.jl:158-193.  Qodd is no x no, Qeven ne x ne (col-major). */
static i64 implicitstep_vec(double *ev, double *Qeven, i64 lde, double *Qodd, i64 ldo, i64 n, i64 Nh,
                            i64 start)
{
    double bulge = 0.0;
    double shift = E1(n) * E1(n);
    double tol = 10.0 * 2.220446049250313e-16;
    for (i64 i = start; i <= n - 1; i++) {
        double al = (i > start) ? E1(i - 1) : 0.0;
        double be = E1(i), ga = E1(i + 1);
        double x1 = -al * al - be * be + shift;
        double x2 = -al * bulge + be * ga;
        double c, s;
        if (i > start) getgivens(al, bulge, &c, &s);
        else getgivens(x1, x2, &c, &s);
        if (i > 1) E1(i - 1) = c * al + s * bulge;
        E1(i) = c * be - s * ga;
        E1(i + 1) = s * be + c * ga;
        if (i < n - 1) {
            double ze = E1(i + 2);
            E1(i + 2) *= c;
            bulge = s * ze;
            if (fabs(bulge) < tol && fabs(E1(i)) < tol) start = i + 1;
        }
        double *Q = (i % 2 == 1) ? Qodd : Qeven;
        i64 ld = (i % 2 == 1) ? ldo : lde;
        i64 k = (i + 1) / 2; /* 1-based column */
        for (i64 j = 1; j <= Nh; j++) {
            double sg = Q[(j - 1) + (k - 1) * ld], om = Q[(j - 1) + k * ld];
            Q[(j - 1) + (k - 1) * ld] = c * sg + s * om;
            Q[(j - 1) + k * ld] = -s * sg + c * om;
        }
    }
    return start;
}

/* getoddvectors: src/skeweigen.jl:49-62 */
static void getoddvectors(double *Qodd, i64 ldo, const double *G, i64 n)
{
    i64 nn = n / 2 + 1;
    for (i64 i = 1; i <= n - 1; i += 2) {
        double c = G1(i), s = G1(i + 1);
        i64 ii = (i + 1) / 2;
        for (i64 j = 1; j <= nn; j++) {
            double sg = Qodd[(ii - 1) + (j - 1) * ldo], om = Qodd[(nn - 1) + (j - 1) * ldo];
            Qodd[(ii - 1) + (j - 1) * ldo] = c * sg + s * om;
            Qodd[(nn - 1) + (j - 1) * ldo] = -s * sg + c * om;
        }
    }
}

/* skewtrieigen_backup!: src/skeweigen.jl:470-490 (needs LAPACK dstev). */
static int skewtrieigen_backup(i64 n, const double *ev, double *vals, double *vr, double *vi)
{
    if (!B.dstev) return 2;
    double *d = (double *)calloc((size_t)n, sizeof(double));
    double *e = (double *)malloc(sizeof(double) * (size_t)(n > 1 ? n - 1 : 1));
    double *Z = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *work = (double *)malloc(sizeof(double) * (size_t)(2 * n + 2));
    for (i64 k = 0; k < n - 1; k++) e[k] = ev[k];
    char jobz = 'V';
    i64 info = 0;
    B.dstev(&jobz, &n, d, e, Z, &n, work, &info, 1);
    for (i64 k = 0; k < n; k++) vals[k] = -d[k];
    for (i64 k = 0; k < n * n; k++) vr[k] = vi[k] = 0.0;
    double c = 1.0;
    for (i64 j = 0; j < n; j++) {
        c = 1.0;
        for (i64 i = 0; i + 1 < n; i += 2) {
            vr[i + j * n] = Z[i + j * n] * c;
            vi[(i + 1) + j * n] = Z[(i + 1) + j * n] * c;
            c *= -1.0;
        }
    }
    if (n % 2 == 1)
        for (i64 j = 0; j < n; j++) vr[(n - 1) + j * n] = Z[(n - 1) + j * n] * c;
    free(d);
    free(e);
    free(Z);
    free(work);
    return (int)info;
}

/* skewtrieigen_divided! (variant 0, src/skeweigen.jl:322-457) and skewtrieigen_merged!
 * (variant 1, :195-320).  N = size; ev (N-1) is overwritten; outputs: vals (imaginary
 * parts), vr/vi = real/imaginary parts of the N x N eigenvector matrix (col-major). */
int orc_skewtrieigen_d(i64 N, double *ev, double *vals, double *vr, double *vi, int variant)
{
    double *backup = (double *)malloc(sizeof(double) * (size_t)(N > 1 ? N - 1 : 1));
    for (i64 k = 0; k < N - 1; k++) backup[k] = ev[k];
    for (i64 k = 0; k < N; k++) vals[k] = 0.0;
    for (i64 k = 0; k < N * N; k++) vr[k] = vi[k] = 0.0;
    i64 no = (N + 1) / 2, ne = N / 2;
    double *Qodd = (double *)calloc((size_t)(no * no > 0 ? no * no : 1), sizeof(double));
    double *Qeven = (double *)calloc((size_t)(ne * ne > 0 ? ne * ne : 1), sizeof(double));
    for (i64 k = 0; k < no; k++) Qodd[k + k * no] = 1.0;
    for (i64 k = 0; k < ne; k++) Qeven[k + k * ne] = 1.0;
    i64 n = N;
    double *G = NULL;
    if (n % 2 == 1) {
        n -= 1;
        G = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        reducetozero(ev, G, n);
    }
    double tol = 2.220446049250313e-16 * 10.0;
    i64 max_iter = 30 * n, iter = 0, count_static = 0, start = 1, halfN = n / 2;
    double mem = 1.0;
    int rc = 0;
    while (n > 2 && iter < max_iter) {
        start = implicitstep_vec(ev, Qeven, ne, Qodd, no, n - 1, halfN, start);
        while (n > 2 && fabs(E1(n - 1)) <= tol) {
            vals[n - 1] = 0.0;
            n -= 1;
        }
        while (n > 2 && fabs(E1(n - 2)) <= tol * fabs(E1(n - 1))) {
            vals[n - 2] = E1(n - 1);
            vals[n - 1] = -E1(n - 1);
            n -= 2;
        }
        if (variant == 0 && n > 2 && fabs(E1(n - 1) - mem) < tol) {
            vals[n - 2] = E1(n - 1);
            vals[n - 1] = -E1(n - 1);
            n -= 2;
        }
        if (start > n - 2) start = 1;
        if (n > 1 && fabs(mem - E1(n - 1)) < 0.0001 * fabs(E1(n - 1))) {
            count_static += 1;
            if (count_static > 4) {
                rc = skewtrieigen_backup(N, backup, vals, vr, vi);
                goto done;
            }
        } else {
            count_static = 0;
        }
        mem = (n > 1) ? E1(n - 1) : 0.0;
        iter += 1;
    }
    if (n > 0) {
        if (n == 2) {
            vals[0] = E1(1);
            vals[1] = -E1(1);
        } else {
            vals[0] = 0.0;
        }
        if (N % 2 == 1) getoddvectors(Qodd, no, G, N - 1);
        double s2 = 1.0 / sqrt(2.0);
#define VR(r, c) vr[((r) - 1) + ((c) - 1) * N]
#define VI(r, c) vi[((r) - 1) + ((c) - 1) * N]
#define QO(r, c) Qodd[((r) - 1) + ((c) - 1) * no]
#define QE(r, c) Qeven[((r) - 1) + ((c) - 1) * ne]
        i64 i = 1, countzeros = 0;
        while (i <= N) {
            i64 ii = (i + 1) / 2;
            if (vals[i - 1] == 0.0) {
                if (countzeros % 2 == 0) {
                    for (i64 j = 1; j <= N - 1; j += 2) VR(j, i) = QO((j + 1) / 2, ii);
                    if (N % 2 == 1) VR(N, i) = QO(halfN + 1, ii);
                } else {
                    for (i64 j = 1; j <= N - 1; j += 2) VR(j + 1, i) = QE((j + 1) / 2, ii);
                }
                countzeros += 1;
                i += 1;
            } else {
                if (countzeros % 2 == 1) {
                    for (i64 j = 1; j <= N - 1; j += 2) {
                        i64 jj = (j + 1) / 2;
                        VI(j, i) = -s2 * QO(jj, ii + 1);
                        VR(j + 1, i) = s2 * QE(jj, ii);
                        VR(j, i + 1) = -s2 * QO(jj, ii + 1);
                        VI(j + 1, i + 1) = s2 * QE(jj, ii);
                    }
                    if (N % 2 == 1) {
                        if (variant == 0) { /* skeweigen.jl:423-424 */
                            VR(N, i) = -s2 * QO(halfN + 1, ii + 1);
                            VI(N, i + 1) = -s2 * QO(halfN + 1, ii + 1);
                        } else { /* skeweigen.jl:287-288 */
                            VI(N, i) = -s2 * QO(halfN + 1, ii + 1);
                            VR(N, i + 1) = -s2 * QO(halfN + 1, ii + 1);
                        }
                    }
                } else {
                    for (i64 j = 1; j <= N - 1; j += 2) {
                        i64 jj = (j + 1) / 2;
                        VR(j, i) = s2 * QO(jj, ii);
                        VI(j + 1, i) = -s2 * QE(jj, ii);
                        VI(j, i + 1) = s2 * QO(jj, ii);
                        VR(j + 1, i + 1) = -s2 * QE(jj, ii);
                    }
                    if (N % 2 == 1) {
                        VR(N, i) = s2 * QO(halfN + 1, ii);
                        VI(N, i + 1) = s2 * QO(halfN + 1, ii);
                    }
                }
                i += 2;
            }
        }
        if (N % 2 == 1)
            for (i64 j = 1; j <= N; j += 2) VR(j, N) = QO((j + 1) / 2, halfN + 1);
#undef VR
#undef VI
#undef QO
#undef QE
    } else if (n == 0) {
        if (variant == 1)
            for (i64 k = 0; k < no; k++) vr[k + k * N] = 1.0; /* complex.(Qodd) */
    } else {
        rc = 1;
    }
done:
    free(backup);
    free(Qodd);
    free(Qeven);
    free(G);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Pfaffian path: src/pfaffian.jl:74-94, 167-191                               */
/* ------------------------------------------------------------------------- */
/* logabspfaffian of the real skew tridiagonal with subdiagonal ev (N-1). */
void orc_logabspfaffian_tri_d(i64 N, const double *ev, double *logpf, double *sgn)
{
    if (N % 2 == 1) {
        *logpf = -INFINITY;
        *sgn = 0.0;
        return;
    }
    double l = 0.0, s = 1.0;
    for (i64 i = 0; i + 1 < N; i += 2) { /* 1-based odd i */
        l += log(fabs(ev[i]));
        double m = -ev[i];
        s *= (m > 0) - (m < 0);
    }
    *logpf = l;
    *sgn = s;
}

/* _logabspfaffian!(::SkewHermitian{<:Real}): A is overwritten by the factors. */
void orc_logabspfaffian_d(i64 n, double *A, i64 lda, double *logpf, double *sgn)
{
    if (n % 2 == 1) {
        *logpf = -INFINITY;
        *sgn = 0.0;
        return;
    }
    double *tau = (double *)calloc((size_t)(n > 1 ? n - 1 : 1), sizeof(double));
    double *eta = (double *)calloc((size_t)(n > 1 ? n - 1 : 1), sizeof(double));
    orc_skewblockedhess_d(n, A, lda, tau, eta, 0);
    orc_logabspfaffian_tri_d(n, eta, logpf, sgn);
    /* sign(det(H.Q)) = (-1)^(number of nonzero tau) for real reflectors */
    i64 nz = 0;
    for (i64 k = 0; k < n - 1; k++) nz += (tau[k] != 0.0);
    if (nz % 2 == 1) *sgn = -*sgn;
    free(tau);
    free(eta);
}

/* _pfaffian!(::SkewHermitian{<:Real}) */
double orc_pfaffian_d(i64 n, double *A, i64 lda)
{
    if (n % 2 == 1) return 0.0;
    double *tau = (double *)calloc((size_t)(n > 1 ? n - 1 : 1), sizeof(double));
    double *eta = (double *)calloc((size_t)(n > 1 ? n - 1 : 1), sizeof(double));
    orc_skewblockedhess_d(n, A, lda, tau, eta, 0);
    double pf = 1.0;
    for (i64 i = 0; i + 1 < n; i += 2) pf *= -eta[i];
    i64 nz = 0;
    for (i64 k = 0; k < n - 1; k++) nz += (tau[k] != 0.0);
    free(tau);
    free(eta);
    return (nz % 2 == 1) ? -pf : pf;
}

/* ------------------------------------------------------------------------- */
/* Complex skew-symmetric Pfaffian: Parlett-Reid elimination with pivoting on   */
/* the largest |A[r,k]| (src/pfaffian.jl:95-137 _pfaffian!, :193-233            */
/* _logabspfaffian! for AbstractMatrix{<:Complex}); the full square is updated, */
/* as the reference does.                                                       */
/* ------------------------------------------------------------------------- */
/* Complex arithmetic of the elimination, written out so that its rounding is fixed (products
 * and sums rounded separately; this file is built with -ffp-contract=off): the product as
 * (ar br - ai bi, ar bi + ai br), the quotient by Smith's method (scaled by the larger component
 * of the divisor).  Julia's own complex division is a scaled variant of the same method; the two
 * can differ in the last bit. */
static zc zmul(zc a, zc b)
{
    const double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
    return CMPLX(ar * br - ai * bi, ar * bi + ai * br);
}
static zc zdiv(zc a, zc b)
{
    const double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
    if (fabs(br) >= fabs(bi)) {
        const double r = bi / br, d = br + bi * r;
        return CMPLX((ar + ai * r) / d, (ai - ar * r) / d);
    }
    const double r = br / bi, d = br * r + bi;
    return CMPLX((ar * r + ai) / d, (ai * r - ar) / d);
}

/* Step k (0-based, even): returns 0 if the pivot A[k+1,k] is zero after the swap, 1 after a
 * step without a swap, 2 with one; *piv = A[k,k+1]. */
static int parlett_reid_step_z(i64 n, zc *A, i64 lda, i64 k, zc *piv, zc *tau)
{
    /* kp = k + argmaxabs(A[k+1:end, k]): the first index of the largest modulus (:107); as in
     * Julia's argmax, NaN ranks above every number */
    i64 kp = k + 1;
    double best = -1.0;
    for (i64 r = k + 1; r < n; r++) {
        const double a = cabs(A[r + k * lda]);
        if (!isnan(best) && (isnan(a) || a > best)) {
            best = a;
            kp = r;
        }
    }
    int swapped = 0;
    if (kp != k + 1) { /* rows k+1 <-> kp, then columns, over l >= k (:108-115) */
        for (i64 l = k; l < n; l++) {
            const zc t = A[(k + 1) + l * lda];
            A[(k + 1) + l * lda] = A[kp + l * lda];
            A[kp + l * lda] = t;
        }
        for (i64 l = k; l < n; l++) {
            const zc t = A[l + (k + 1) * lda];
            A[l + (k + 1) * lda] = A[l + kp * lda];
            A[l + kp * lda] = t;
        }
        swapped = 1;
    }
    if (A[(k + 1) + k * lda] == 0) return 0; /* (:119, :133) */
    const zc a = A[k + (k + 1) * lda];
    const i64 m = n - k - 2;
    for (i64 c = 0; c < m; c++) tau[c] = zdiv(A[k + (k + 2 + c) * lda], a); /* Gauss vector (:120) */
    for (i64 l1 = 0; l1 < m; l1++) /* rank-2 update of the trailing square (:123-129) */
        for (i64 l2 = 0; l2 < m; l2++) {
            const zc d = zmul(tau[l2], A[(k + 2 + l1) + (k + 1) * lda]) - zmul(tau[l1], A[(k + 2 + l2) + (k + 1) * lda]);
            A[(k + 2 + l2) + (k + 2 + l1) * lda] += d;
        }
    *piv = a;
    return 1 + swapped;
}

/* _pfaffian!(A::AbstractMatrix{<:Complex}) (src/pfaffian.jl:97-137); A is overwritten. */
zc orc_pfaffian_z(i64 n, zc *A, i64 lda)
{
    if (n % 2 == 1) return 0;
    zc *tau = (zc *)malloc((size_t)(n > 2 ? n : 2) * sizeof(zc));
    zc pf = 1;
    for (i64 k = 0; k + 1 < n; k += 2) {
        zc piv = 0;
        const int st = parlett_reid_step_z(n, A, lda, k, &piv, tau);
        if (st == 2) pf = -pf;
        if (st == 0) {
            free(tau);
            return 0;
        }
        pf = zmul(pf, piv);
    }
    free(tau);
    return pf;
}

/* _logabspfaffian!(A::AbstractMatrix{<:Complex}) (src/pfaffian.jl:193-233): log|Pf| and the
 * unit cis(phase), phase = sum of angle(A[k,k+1]) + pi per pivot swap; odd n or a zero pivot give
 * (-Inf, 0).  A is overwritten. */
void orc_logabspfaffian_z(i64 n, zc *A, i64 lda, double *logpf, zc *phase_out)
{
    if (n % 2 == 1) {
        *logpf = -INFINITY;
        *phase_out = 0;
        return;
    }
    zc *tau = (zc *)malloc((size_t)(n > 2 ? n : 2) * sizeof(zc));
    double lp = 0.0, ph = 0.0;
    for (i64 k = 0; k + 1 < n; k += 2) {
        zc piv = 0;
        const int st = parlett_reid_step_z(n, A, lda, k, &piv, tau);
        if (st == 2) ph += 3.14159265358979323846; /* pi, Float64(pi) */
        if (st == 0) {
            free(tau);
            *logpf = -INFINITY;
            *phase_out = 0;
            return;
        }
        lp += log(cabs(piv));
        ph += carg(piv);
    }
    free(tau);
    *logpf = lp;
    *phase_out = cexp(I * ph);
}

/* orc_pfaffian_z with the result written through a pointer (plain-C callers: ctypes) */
void orc_pfaffian_z_out(i64 n, zc *A, i64 lda, zc *out) { *out = orc_pfaffian_z(n, A, lda); }

/* ------------------------------------------------------------------------------------------
 * Skew-Cholesky (P. Benner et al., "Cholesky-like factorizations of skew-symmetric matrices",
 * ETNA 11 (2000)): R^T J R = A[p, p] with R upper triangular and J = JMatrix{+1} (the 2 x 2
 * blocks [0 1; -1 0] on the diagonal).  Restates _skewchol!(A::SkewHermitian{<:Real}, pivot) of
 * src/cholesky.jl:28-56 (NoPivot) and :58-107 (RowMaximum).  B (n x n, full storage) is
 * overwritten as in the reference: its upper triangle is R.  p receives the permutation, 0-based.
 *
 * Step j (0-based pivot rows i0 = 2j, i1 = 2j + 1), old rows b0 = B[i0, c], b1 = B[i1, c]
 * (c > i1), r = sqrt(B[i0, i1]) and q = -1/r:
 *   rows   B[i0, c] = b1 * q,  B[i1, c] = (-b0) * q          (:94-97; J2 = [0 1; -1 0])
 *   update B[a, b] -= n0[a] * n1[b] - n1[a] * n0[b]          (:98-99, mul! with alpha -1)
 * The two products of the update are rounded separately (no fused multiply-add), so the
 * trailing square stays exactly skew-symmetric and its diagonal exactly zero; the reference's
 * BLAS gemm may fuse, which changes results at the rounding level only.
 * RowMaximum: the pivot is findmax over the full trailing square (:67-71): the largest value,
 * the first in column-major order among equal ones (NaN above all numbers, as Julia's findmax);
 * a value below tol = 1e-15 * ||B||_F ends the factorization, leaving the trailing square as
 * it is (:71).  Its row index moves to i0 and its column index to i1 by symmetric swaps of whole
 * rows and columns (:73-90).  NoPivot (:28-56): s = sign(B[i0, i1]), r = sqrt(|B[i0, i1]|),
 * B[i1, i1] = s r and the first new row is (s b1) * q.
 * ------------------------------------------------------------------------------------------ */
static void chol_swap(i64 n, double *B, i64 ldb, i64 x, i64 y)
{
    for (i64 t = 0; t < n; t++) { /* columns x <-> y (:77-79) */
        const double v = B[t + x * ldb];
        B[t + x * ldb] = B[t + y * ldb];
        B[t + y * ldb] = v;
    }
    for (i64 t = 0; t < n; t++) { /* rows x <-> y (:80-82) */
        const double v = B[x + t * ldb];
        B[x + t * ldb] = B[y + t * ldb];
        B[y + t * ldb] = v;
    }
}

/* Julia's isless ordering for findmax: NaN ranks above every number */
static int chol_greater(double a, double b) { return isnan(b) ? 0 : (isnan(a) || a > b); }

/* returns the number of 2 x 2 steps completed */
i64 orc_skewchol_d(i64 n, double *B, i64 ldb, int pivot, i64 *p)
{
    for (i64 i = 0; i < n; i++) p[i] = i;
    if (n <= 1) return 0; /* :31, :62 */
    double tol = 0.0;
    if (pivot) { /* tol = 1e-15 * norm(B) (:60), column-major sum of squares */
        double s = 0.0;
        for (i64 c = 0; c < n; c++)
            for (i64 r = 0; r < n; r++) s += B[r + c * ldb] * B[r + c * ldb];
        tol = 1e-15 * sqrt(s);
    }
    double *n0 = (double *)malloc((size_t)n * sizeof(double));
    double *n1 = (double *)malloc((size_t)n * sizeof(double));
    i64 steps = 0;
    for (i64 j = 0; j < n / 2; j++) {
        const i64 i0 = 2 * j, i1 = 2 * j + 1;
        double s = 1.0;
        if (pivot) {
            i64 ii = i0, jj = i0; /* findmax(B[i0:n, i0:n]), column-major, first on ties */
            double best = B[i0 + i0 * ldb];
            for (i64 c = i0; c < n; c++)
                for (i64 r = i0; r < n; r++)
                    if (chol_greater(B[r + c * ldb], best)) {
                        best = B[r + c * ldb];
                        ii = r;
                        jj = c;
                    }
            if (fabs(B[ii + jj * ldb]) < tol) break; /* :71 */
            const i64 kk = jj == i0 ? ii : jj;      /* :73 */
            if (ii != i0) {
                const i64 t = p[ii];
                p[ii] = p[i0];
                p[i0] = t;
                chol_swap(n, B, ldb, ii, i0);
            }
            if (kk != i1) {
                const i64 t = p[kk];
                p[kk] = p[i1];
                p[i1] = t;
                chol_swap(n, B, ldb, kk, i1);
            }
        }
        const double v = B[i0 + i1 * ldb];
        double r;
        if (pivot) {
            r = sqrt(v);
        } else {
            s = v > 0 ? 1.0 : v < 0 ? -1.0 : v; /* Julia's sign: +-0 and NaN pass through */
            r = sqrt(fabs(v));
        }
        const double q = -1.0 / r;
        for (i64 c = i1 + 1; c < n; c++) {
            const double b0 = B[i0 + c * ldb], b1 = B[i1 + c * ldb];
            n0[c] = (pivot ? b1 : s * b1) * q;
            n1[c] = (-b0) * q;
        }
        B[i0 + i0 * ldb] = r;
        B[i1 + i1 * ldb] = pivot ? r : s * r;
        B[i0 + i1 * ldb] = 0.0;
        for (i64 c = i1 + 1; c < n; c++) {
            B[i0 + c * ldb] = n0[c];
            B[i1 + c * ldb] = n1[c];
        }
        for (i64 b = i1 + 1; b < n; b++)
            for (i64 a = i1 + 1; a < n; a++) {
                const double x = n0[a] * n1[b], y = n1[a] * n0[b];
                B[a + b * ldb] -= x - y;
            }
        steps++;
    }
    free(n0);
    free(n1);
    return steps;
}
