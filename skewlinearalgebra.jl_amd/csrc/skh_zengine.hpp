// skh_zengine.hpp -- complex skew-Hermitian tridiagonalization (ComplexF64) on gfx950.
//
// Reference: the complex branches of skewblockedhess! (src/hessenberg.jl:92-131 with :96-104,
// :123-126, :161 and the tail skewhess! :70-90): reflector (householder!, :26-44) with
// beta = (Re alpha > 0 ? -1 : +1) * hypot(|alpha|, ||x2||) real, tau = 1 - alpha/beta complex,
// v = [1; x2/(alpha - beta)]; w = tau A22 v, w += (tau w^H v / 2) v; A22 += v w^H - w v^H.  The
// length-1 reflector of the last column is a full reflector in the complex case (it makes
// eta[n-2] real and stores the unit in the factor).  T keeps the imaginary diagonal.
//
// This engine applies the reflectors one column at a time (unblocked), in cooperative persistent
// launches of G co-resident workgroups, two grid barriers per column, one launch per segment of
// columns over which the 64x64 tile grid of the trailing block is the same (k + 1 within one
// 64-row block): each workgroup then updates and re-reads the same tiles for the whole launch,
// so tile data stays inside the workgroup (plain cached accesses), and everything that crosses
// workgroups (partials, the row pass's outputs, column k+1 and the diagonal of the tile update)
// goes through agent-scope atomic accesses and the fence-free barrier.  Per reflector k (v_k from x_k = A[k+1:, k], scalars computed redundantly by every
// workgroup from the same partials, bit-identical):
//   P(k): tiles of the lower triangle (diagonal included) of A[k+1:, k+1:], split into G
//         contiguous ranges: apply the previous reflector, a += v_{k-1} w_{k-1}^H - w_{k-1} v_{k-1}^H
//         (stored back), then from the updated tile accumulate y = A22 v_k (row partials per
//         strip, column partials per tile) and the real m with v_k^H A22 v_k = i m;
//   R(k): rows in 64-row chunks (chunk q owned by workgroup q mod G): y_r from the partials,
//         w_r = tau y_r + c v_r with c = tau (w'^H v)/2 = -|tau|^2 (i m)/2 (the reference's
//         correction in closed form), the factor v_r into A[r, k], the diagonal (k+1, k+1) update,
//         and the next column x_r = A[r, k+1] + v_r conj(w_{k+1}) - w_r conj(v_{k+1}) with its
//         norm partials.
// Every reduction has a fixed order (the results are deterministic).
#pragma once
#include "skh_device.hpp"

namespace skh {
namespace zc {

struct cd {
    double re, im;
};

__device__ __forceinline__ cd mk(double r, double i) { return cd{r, i}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd operator*(cd a, cd b) { return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cd conj(cd a) { return cd{a.re, -a.im}; }
__device__ __forceinline__ double norm2(cd a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ cd scal(double s, cd a) { return cd{s * a.re, s * a.im}; }
// a * conj(b)
__device__ __forceinline__ cd mulc(cd a, cd b) { return cd{a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im}; }

__device__ __forceinline__ cd ld(const double2* p) { const double2 t = *p; return cd{t.x, t.y}; }
__device__ __forceinline__ void st(double2* p, cd v) { *p = make_double2(v.re, v.im); }

// Data written by one workgroup and read by another within a launch goes through agent-scope
// atomic accesses (coherent across the XCD L2s; see ld_co/st_co), two doubles per complex entry.
__device__ __forceinline__ cd ldc(const double2* p)
{
    const double* q = reinterpret_cast<const double*>(p);
    return cd{ld_co(q), ld_co(q + 1)};
}
__device__ __forceinline__ void stc(double2* p, cd v)
{
    double* q = reinterpret_cast<double*>(p);
    st_co(q, v.re);
    st_co(q + 1, v.im);
}

// Reflector scalars (complex householder!): alpha = x[0], xn2 = ||x[1:]||^2.
struct ZScal {
    cd tau, scale;
    double beta, unit;
    cd corr;   // unit - scale * alpha: v = scale * x + corr * e_{j+1} for the unscaled x (x_{j+1} = alpha)
    cd q;      // x^H A22 e_{j+1} (the blocked panel kernel's m from the unscaled matvec)
};

__device__ __forceinline__ ZScal zreflector(cd alpha, double xn2)
{
    ZScal s;
    const double xnorm = sqrt(xn2);
    if (xnorm != 0.0 || alpha.re != 0.0 || alpha.im != 0.0) {
        const double b = ((alpha.re > 0.0) ? -1.0 : 1.0) * hypot(hypot(alpha.re, alpha.im), xnorm);
        s.beta = b;
        s.tau = cd{1.0 - alpha.re / b, -alpha.im / b};
        // scale = 1 / (alpha - b)
        const cd d = cd{alpha.re - b, alpha.im};
        const double den = d.re * d.re + d.im * d.im;
        s.scale = cd{d.re / den, -d.im / den};
        s.unit = 1.0;
    } else {
        s.beta = 0.0;
        s.tau = cd{0.0, 0.0};
        s.scale = cd{0.0, 0.0};
        s.unit = 0.0;
    }
    // b is real with the sign opposite to Re alpha and |b| >= |alpha|, so |alpha - b|^2 >= |alpha|^2 + |b|^2
    // and |scale * alpha| <= 1/sqrt(2): no cancellation
    s.corr = cd{s.unit - (s.scale.re * alpha.re - s.scale.im * alpha.im), -(s.scale.re * alpha.im + s.scale.im * alpha.re)};
    s.q = cd{0.0, 0.0};
    return s;
}

// Workspace of the complex engine (device).
struct ZWork {
    double2* Prow;   // [slots][ldp] row partials of the matvec
    double2* Pcol;   // [nt][ldp] column partials
    int64_t ldp;
    double2* Wv;     // [n] w of the previous reflector
    double* Pnrm;    // [G] norm partials of the next column
    double* Pmu;     // [G] partials of m (v^H A22 v = i m)
    double2* alpha;  // [1] head of the next column
};

// v_k at index c from its raw column entry x = A[c, k] (x_k is kept unscaled in A[:, k] until R(k)):
// unit at k+1, scale * x below, zero outside (k+1, n).  The entry is loaded unconditionally from a
// valid address and masked here: a use of a load right after it makes the compiler wait for it.
__device__ __forceinline__ cd vfrom_raw(int64_t n, int64_t k, int64_t c, cd x, const ZScal& s)
{
    if (c <= k || c >= n) return cd{0.0, 0.0};
    if (c == k + 1) return cd{s.unit, 0.0};
    return s.scale * x;
}

struct ZShared {
    cd red[4][TB];
    double sred[4];
    cd wspec;
    double mu;
};

// P(k): update with the previous reflector (upd) and the matvec partials of v_k.
__device__ void zpass(double2* __restrict__ A, int64_t lda, int64_t n, int64_t k, bool upd, const ZScal& s,
                      const ZWork& w, ZShared& sh, int g, int G)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int p = lane & 15, cg = lane >> 4;
    const int64_t s0 = k + 1;
    const int64_t nt = cdiv(n, TB);
    const int64_t J0 = s0 / TB;
    const int64_t ns = nt - J0;
    const int64_t T = ns * (ns + 1) / 2;
    const int64_t Ge = (G < T) ? G : T;
    double mu = 0.0;
    if (g < Ge) {
        const int64_t tlo = ((int64_t)g * T) / Ge, thi = ((int64_t)(g + 1) * T) / Ge;
        int64_t sI = strip_of(tlo);
        int64_t I = J0 + sI, J = J0 + (tlo - sI * (sI + 1) / 2);
        int64_t cur_slot = g - tile_owner(sI * (sI + 1) / 2, T, Ge);
        const double2* vprev = upd ? A + (k - 1) * lda : A;   // v_{k-1} (factor stored in R(k-1))
        cd racc[4], rv[4], rw[4], rn[4];
        int64_t rows[4];
        // every load of a step is issued before any is used, from valid addresses (index s0 stands
        // in for one outside [s0, n)), and masked afterwards
        auto set_rows = [&](int64_t II) {
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                rows[rr] = II * TB + 32 * (rr >> 1) + 2 * p + (rr & 1);
                const int64_t r = rows[rr];
                const int64_t rc = (r >= s0 && r < n) ? r : s0;
                rv[rr] = ldc(vprev + rc);
                rw[rr] = ldc(w.Wv + rc);
                rn[rr] = ldc(A + rc + k * lda);
            }
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int64_t r = rows[rr];
                if (!(upd && r >= s0 && r < n)) rv[rr] = rw[rr] = cd{0.0, 0.0};
                rn[rr] = vfrom_raw(n, k, r, rn[rr], s);
                racc[rr] = cd{0.0, 0.0};
            }
        };
        set_rows(I);
        for (int64_t t = tlo; t < thi; t++) {
            const int64_t cb = J * TB + 16 * wave + 4 * cg;
            cd a[4][4], cv[4], cw[4], cn[4], cacc[4];
#pragma unroll
            for (int cc = 0; cc < 4; cc++) {
                const int64_t c = cb + cc;
                const int64_t ci = (c >= s0 && c < n) ? c : s0;
                cv[cc] = ldc(vprev + ci);
                cw[cc] = ldc(w.Wv + ci);
                cn[cc] = ldc(A + ci + k * lda);
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const int64_t r = rows[rr];
                    a[rr][cc] = ld(A + ((r < n) ? r : s0) + ci * lda);
                }
            }
#pragma unroll
            for (int cc = 0; cc < 4; cc++) {
                const int64_t c = cb + cc;
                const bool cin = c >= s0 && c < n;
                if (!(upd && cin)) cv[cc] = cw[cc] = cd{0.0, 0.0};
                cn[cc] = vfrom_raw(n, k, c, cn[cc], s);
                cacc[cc] = cd{0.0, 0.0};
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const int64_t r = rows[rr];
                    if (!(cin && r >= c && r < n)) a[rr][cc] = cd{0.0, 0.0};
                }
            }
            if (upd) {
#pragma unroll
                for (int cc = 0; cc < 4; cc++) {
                    const int64_t c = cb + cc;
#pragma unroll
                    for (int rr = 0; rr < 4; rr++) {
                        const int64_t r = rows[rr];
                        if (c >= s0 && c < n && r >= c && r < n) {
                            a[rr][cc] = a[rr][cc] + mulc(rv[rr], cw[cc]) - mulc(rw[rr], cv[cc]);
                            // column s0 (read by the row pass) and the diagonal (read by the row
                            // pass and the last reflector) leave this workgroup; the rest of the
                            // tile is read again only by it, later in the same launch
                            if (c == s0 || r == c) stc(A + r + c * lda, a[rr][cc]);
                            else st(A + r + c * lda, a[rr][cc]);
                        }
                    }
                }
            }
#pragma unroll
            for (int cc = 0; cc < 4; cc++) {
                const int64_t c = cb + cc;
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const int64_t r = rows[rr];
                    const cd x = a[rr][cc];   // zero outside the stored lower triangle
                    if (r > c) {
                        racc[rr] = racc[rr] + x * cn[cc];
                        cacc[cc] = cacc[cc] - conj(x) * rn[rr];   // y_c += A[c, r] v_r, A[c, r] = -conj(a)
                        // m += 2 Im(conj(v_r) a_rc v_c)
                        const cd q = mulc(x * cn[cc], rn[rr]);   // a v_c conj(v_r)
                        mu += 2.0 * q.im;
                    } else if (r == c && c >= s0 && r < n) {
                        racc[rr] = racc[rr] + x * rn[rr];
                        mu += x.im * norm2(rn[rr]);
                    }
                }
            }
#pragma unroll
            for (int cc = 0; cc < 4; cc++) {
                const cd x = cd{row16_sum(cacc[cc].re), row16_sum(cacc[cc].im)};
                const int64_t c = cb + cc;
                if (p == 0 && c >= s0 && c < n) stc(w.Pcol + I * w.ldp + c, x);
            }
            const bool more = t + 1 < thi;
            const int64_t In = (J < I) ? I : I + 1;
            const int64_t Jn = (J < I) ? J + 1 : J0;
            if (!more || In != I) {
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    racc[rr] = cd{xor32_sum(xor16_sum(racc[rr].re)), xor32_sum(xor16_sum(racc[rr].im))};
                }
                if (cg == 0) {
                    sh.red[wave][2 * p] = racc[0];
                    sh.red[wave][2 * p + 1] = racc[1];
                    sh.red[wave][32 + 2 * p] = racc[2];
                    sh.red[wave][33 + 2 * p] = racc[3];
                }
                __syncthreads();
                if (wave == 0) {
                    const int64_t r = I * TB + lane;
                    if (r < n)
                        stc(w.Prow + cur_slot * w.ldp + r,
                            ((sh.red[0][lane] + sh.red[1][lane]) + (sh.red[2][lane] + sh.red[3][lane])));
                }
                __syncthreads();
                cur_slot = 0;
                if (more) set_rows(In);
            }
            I = In;
            J = Jn;
        }
    }
    mu = wave_sum(mu);
    if (lane == 0) sh.sred[wave] = mu;
    __syncthreads();
    if (tid == 0) st_co(w.Pmu + g, (sh.sred[0] + sh.sred[1]) + (sh.sred[2] + sh.sred[3]));
    __syncthreads();
}

// number of row-partial slots of strip index sb (from J0) with ns strips over G ranks
__device__ __forceinline__ int64_t zslots(int64_t sb, int64_t ns, int64_t G) { return strip_slots(sb, ns, G); }

// y_r for one row r of strip b (b0 = first strip): row slots + column partials of strips >= b.
// Wave-split sums (slots and strips by wave), combined by the caller in a fixed order.
// The row's partials are numbered q = 0 .. nsl + nt - b - 1 (row slots first, then the column
// partials of strips b ..); wave w takes q = w, w + 4, ... in batches of 8 loads, so a row costs a
// few memory round trips instead of one per partial.
__device__ __forceinline__ cd zrow_partial(const ZWork& w, int64_t r, int64_t b, int64_t b0, int64_t nt, int64_t G,
                                           int wave)
{
    constexpr int ZB = 8;
    cd acc = cd{0.0, 0.0};
    const int64_t nsl = zslots(b - b0, nt - b0, G);
    const int64_t nall = nsl + (nt - b);
    for (int64_t q0 = wave; q0 < nall; q0 += 4 * ZB) {
        cd buf[ZB];
#pragma unroll
        for (int u = 0; u < ZB; u++) {
            const int64_t q = q0 + 4 * u;
            buf[u] = (q < nsl) ? ldc(w.Prow + q * w.ldp + r)
                               : ((q < nall) ? ldc(w.Pcol + (b + q - nsl) * w.ldp + r) : cd{0.0, 0.0});
        }
#pragma unroll
        for (int u = 0; u < ZB; u++) acc = acc + buf[u];
    }
    return acc;
}

// R(k): finalise w_k, store the factor column k, update the diagonal (k+1, k+1), form the next
// column x_{k+1} (rows >= k+2) and its norm partials.
__device__ void zrows(double2* __restrict__ A, int64_t lda, int64_t n, int64_t k, const ZScal& s, const ZWork& w,
                      ZShared& sh, int g, int G, int Gfull)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t s0 = k + 1;
    const int64_t nt = cdiv(n, TB);
    const int64_t b0 = s0 / TB;
    const int64_t ns = nt - b0;
    const int64_t T = ns * (ns + 1) / 2;
    const int64_t Ge = (G < T) ? G : T;
    // m = sum of the pass partials (every workgroup wrote one), fixed order
    if (wave == 0) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = ld_co(w.Pmu + ((lane + 64 * u < Gfull) ? lane + 64 * u : 0));
        double m = 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++) m += (lane + 64 * u < Gfull) ? v[u] : 0.0;
        for (int q = lane + 512; q < Gfull; q += 64) m += ld_co(w.Pmu + q);
        m = wave_sum(m);
        if (lane == 0) sh.mu = m;
    }
    // special row k+1: y_{k+1}, redundantly in every workgroup with one fixed order
    {
        cd acc = zrow_partial(w, s0, b0, b0, nt, Ge, wave);
        sh.red[wave][lane] = (lane == 0) ? acc : cd{0.0, 0.0};
    }
    __syncthreads();
    // c = -|tau|^2 (i m) / 2 = (0, -|tau|^2 m / 2)
    const cd c = cd{0.0, -0.5 * norm2(s.tau) * sh.mu};
    const cd yspec = (sh.red[0][0] + sh.red[1][0]) + (sh.red[2][0] + sh.red[3][0]);
    const cd vspec = cd{s.unit, 0.0};
    const cd wspec = s.tau * yspec + c * vspec;
    __syncthreads();
    double nrm = 0.0;
    for (int64_t q = g; q < ns; q += G) {
        const int64_t b = b0 + q;
        const int64_t r = b * TB + lane;
        const bool rin = r >= s0 && r < n;
        cd part = rin ? zrow_partial(w, r, b, b0, nt, Ge, wave) : cd{0.0, 0.0};
        sh.red[wave][lane] = part;
        __syncthreads();
        if (wave == 0 && rin) {
            // loads first (column k; the diagonal or column k+1), then the arithmetic
            const cd xr = ldc(A + r + k * lda);
            const cd o = ldc(A + r + ((r == s0) ? r : k + 1) * lda);
            const cd y = (sh.red[0][lane] + sh.red[1][lane]) + (sh.red[2][lane] + sh.red[3][lane]);
            const cd vr = vfrom_raw(n, k, r, xr, s);
            const cd wr = (r == s0) ? wspec : (s.tau * y + c * vr);
            stc(w.Wv + r, wr);
            stc(A + r + k * lda, vr);   // factor (unit at k+1)
            if (r == s0) {
                // diagonal (k+1, k+1): += v conj(w) - w conj(v)
                stc(A + r + r * lda, o + mulc(vr, wr) - mulc(wr, vr));
            } else {
                // next column: x_r = A[r, k+1] + v_r conj(w_{k+1}) - w_r conj(v_{k+1})
                const cd x = o + mulc(vr, wspec) - mulc(wr, vspec);
                stc(A + r + (k + 1) * lda, x);
                if (r == k + 2) stc(w.alpha, x);
                else nrm += norm2(x);
            }
        }
        __syncthreads();
    }
    if (wave == 0) {
        nrm = wave_sum(nrm);
        if (lane == 0) st_co(w.Pnrm + g, nrm);
    }
}

// Last reflector (k = n-2, one thread): store its unit in the factor and apply the previous
// reflector (n-3) to the remaining diagonal entry (n-1, n-1), the part of the trailing block that
// no pass covers.
__device__ inline void zlast(double2* __restrict__ A, int64_t lda, int64_t n, const ZWork& w, const ZScal& s)
{
    stc(A + (n - 1) + (n - 2) * lda, cd{s.unit, 0.0});
    if (n >= 3) {
        const cd v = ldc(A + (n - 1) + (n - 3) * lda);
        const cd wv = ldc(w.Wv + (n - 1));
        const cd d = ldc(A + (n - 1) + (n - 1) * lda);
        stc(A + (n - 1) + (n - 1) * lda, d + mulc(v, wv) - mulc(wv, v));
    }
}

// Initial norm partials of column k (rows >= k + 2) and its head, for a start with no pending
// reflector (column 0, or the first column after the blocked panels).
__device__ void zinit(const double2* __restrict__ A, int64_t lda, int64_t n, const ZWork& w, int g, int G, int64_t k)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ double red[4];
    double nrm = 0.0;
    const double2* col = A + k * lda;
    for (int64_t r = k + 2 + (int64_t)g * 256 + tid; r < n; r += (int64_t)G * 256) nrm += norm2(ld(col + r));
    nrm = wave_sum(nrm);
    if (lane == 0) red[wave] = nrm;
    __syncthreads();
    if (tid == 0) {
        st_co(w.Pnrm + g, (red[0] + red[1]) + (red[2] + red[3]));
        if (g == 0) stc(w.alpha, k + 1 < n ? ld(col + k + 1) : cd{0.0, 0.0});
    }
    __syncthreads();
}

// Scalars of reflector k from the norm partials (every workgroup, fixed order).
__device__ ZScal zscalars(const ZWork& w, int Gfull, ZShared& sh)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (wave == 0) {
        // up to 512 partials: all loads issued before the sum (a sum between them serialises them)
        double v[8];
        const cd al = ldc(w.alpha);
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = ld_co(w.Pnrm + ((lane + 64 * u < Gfull) ? lane + 64 * u : 0));
        double a = 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++) a += (lane + 64 * u < Gfull) ? v[u] : 0.0;
        for (int q = lane + 512; q < Gfull; q += 64) a += ld_co(w.Pnrm + q);
        a = wave_sum(a);
        if (lane == 0) {
            sh.sred[0] = a;
            sh.wspec = al;
        }
    }
    __syncthreads();
    const ZScal s = zreflector(sh.wspec, sh.sred[0]);
    __syncthreads();
    (void)tid;
    return s;
}

}  // namespace zc
}  // namespace skh
