// skh_zpanel.hpp -- blocked complex skew-Hermitian tridiagonalization (ComplexF64): the panel
// kernel's device code.
//
// Reference: skewlatrd! with its complex branches (src/hessenberg.jl:92-131, :96-104, :123-126):
// per column j of a panel (local index i = j - c0, reflectors 0..i-1 of the panel pending):
//   x = A[j:, j] + V conj(W[j, :]) - W conj(V[j, :])        (column update, diagonal included)
//   reflector (householder!, :26-44) of x[j+1:]: tau complex, beta real, v = [1; x / (alpha - beta)]
//   y = A22 v + V (W^H v) - W (V^H v),  w = tau y,  w += (tau w^H v / 2) v
// and after each panel A22 += V W^H - W V^H (:158-166, the complex MFMA update in skh_complex.hip).
// The correction `w += (tau w^H v / 2) v` is taken in closed form, as in the unblocked engine:
// v^H y = i m is imaginary (A22 + V W^H - W V^H is skew-Hermitian), so it is c v with
// c = -|tau|^2 (i m) / 2 and m = Im(v^H A22 v) + 2 Im(t2^H t1), t1 = W^H v, t2 = V^H v.
//
// The structure is the real panel kernel's (skh_device.hpp): persistent launch (all co-resident) per
// panel, no grid barrier (every hand-over value sits in a slot holding the empty marker until it is
// stored, and is polled; the same copies, refill rules and drains), a row pass per 64-row block
// and a phase B per column with the skew matvec over 256 x 16 tiles of the trailing lower
// triangle.  What the complex case adds:
//   * the diagonal of the trailing block (imaginary, not zero) belongs to the matvec: the tile
//     holding (c, c) adds A[c,c] v_c to row c and Im(A[c,c]) |v_c|^2 to m;
//   * the matvec also produces the per-workgroup partials of m, polled by every row pass;
//   * the matvec runs on the unscaled column x (x_{j+1} = alpha), as in the real kernel: with
//     v = s x + c e_{j+1} (s = 1/(alpha - beta), c = unit - s alpha) the row pass forms
//     A22 v = s A22 x + c A22 e_{j+1} and m = |s|^2 Im(x^H A22 x) + 2 Im(conj(s) c q)
//     + |c|^2 Im(A[j+1, j+1]) with q = x^H A22 e_{j+1} from per-block partials of the row pass;
//   * A is not written on its diagonal during a launch: the panel columns' final diagonal goes to
//     Dfin and is copied into A after the launch (the special row of w reads A[j, j] as it was at
//     the launch's start);
//   * a complex slot is two 8-byte stores; a consumer waits until both halves are filled (both
//     were emptied, so no half can be stale).
// One workgroup per CU (a tile is 16 complex entries per lane), so a workgroup owns one 64-row
// block (n <= 64 x grid); the host keeps larger matrices on the unblocked engine.
#pragma once
#include "skh_device.hpp"
#include "skh_zengine.hpp"

namespace skh {
namespace zp {

using zc::cd;
using zc::ZScal;

constexpr int ZNB = 32;   // largest panel width
constexpr int ZKQ = 8;    // panel columns per wave (ZNB / 4)

struct ZPWork {
    double2* W;      // n x ZNB (ldw): w of the panel's reflectors (empty marker until stored)
    double2* V2;     // n x ZNB (ldw): the panel's reflectors, for the row hand-over
    int64_t ldw;
    double* Pnorm;   // [nt] per copy: per-block sums of |x|^2 of the reflector tail
    double2* PS;     // [nt][2 ZNB] per copy: per-block W^H x and V^H x (rows >= j+2)
    double2* X;      // [ld] per copy: x_j, the updated column j below the diagonal
    double2* Pq;     // [nt] per copy: per-block sums of conj(x_r) A[r, j+1], rows >= j+1
    int64_t pn_stride, ps_stride, x_stride, pq_stride;   // copy strides (column mod 4)
    double2* Tv;     // [2 ZNB] per copy: W^H x, V^H x sums (unscaled; the consumer scales, adds the head)
    double* Pmu;     // [grid] per copy: per-workgroup Im(v^H A22 v) partials
    double2* Prow;   // [nseg][ldp] per copy: row partials of the matvec
    double2* Pcol;   // [nts][ldp] per copy: column partials
    int64_t ldp;
    int64_t tv_stride, mu_stride, pr_stride, pc_stride;   // copy strides (column mod 3)
    double2* Dfin;   // [n] final diagonal entries of the panel columns
};

__device__ __forceinline__ double* pnorm_of(const ZPWork& w, int64_t j) { return w.Pnorm + own_copy(j) * w.pn_stride; }
__device__ __forceinline__ double2* ps_of(const ZPWork& w, int64_t j) { return w.PS + own_copy(j) * w.ps_stride; }
__device__ __forceinline__ double2* x_of(const ZPWork& w, int64_t j) { return w.X + own_copy(j) * w.x_stride; }
__device__ __forceinline__ double2* pq_of(const ZPWork& w, int64_t j) { return w.Pq + own_copy(j) * w.pq_stride; }
__device__ __forceinline__ double2* tv_of(const ZPWork& w, int64_t j) { return w.Tv + hand_copy(j) * w.tv_stride; }
__device__ __forceinline__ double* pmu_of(const ZPWork& w, int64_t j) { return w.Pmu + hand_copy(j) * w.mu_stride; }
__device__ __forceinline__ double2* prow_of(const ZPWork& w, int64_t j) { return w.Prow + hand_copy(j) * w.pr_stride; }
__device__ __forceinline__ double2* pcol_of(const ZPWork& w, int64_t j) { return w.Pcol + hand_copy(j) * w.pc_stride; }

__device__ __forceinline__ cd czero() { return cd{0.0, 0.0}; }
__device__ __forceinline__ cd cld(const double2* p) { return zc::ldc(p); }
__device__ __forceinline__ void cst(double2* p, cd v) { zc::stc(p, v); }
__device__ __forceinline__ bool cfilled(cd x) { return filled(x.re) && filled(x.im); }
__device__ __forceinline__ void cempty(double2* p) { cst(p, cd{empty_slot(), empty_slot()}); }
__device__ __forceinline__ cd cwave_sum(cd x) { return cd{wave_sum(x.re), wave_sum(x.im)}; }
__device__ __forceinline__ cd cmul(cd a, cd b) { return zc::operator*(a, b); }
__device__ __forceinline__ cd cadd(cd a, cd b) { return zc::operator+(a, b); }
__device__ __forceinline__ cd csub(cd a, cd b) { return zc::operator-(a, b); }
// a * conj(b)
__device__ __forceinline__ cd cmulc(cd a, cd b) { return zc::mulc(a, b); }
__device__ __forceinline__ cd cconj(cd a) { return cd{a.re, -a.im}; }

struct ZPShared {
    cd wrow[ZNB];        // row j of W (k < i), w_{j-1}[j] at i-1
    cd vrow[ZNB];        // row j of V
    cd t[2 * ZNB];       // t1 = W^H v, t2 = V^H v of column j-1
    union {
        struct {
            cd red[4][TB];
            cd red2[4][TB];
        };
        cd mred[16][MR + 2];   // matvec row sums of one strip per (wave, column group)
    };
    cd vstrip[MR];
    cd wspec;
    cd cfix;             // c of column j-1 (the closed-form correction)
    double mu[4];
    double psx[4][32];   // per wave: the 32 totals of the W^H x / V^H x reduction
    ZScal cur;           // scalars of the column finalised by the last phase B
};

template <int KQ>
struct ZRowRegs {
    cd* v;   // LDS [4 KQ][64]
    cd* w;
    cd xprev, xnext;
};

// does rank g own a t element live at local column i (elements e = off + k, off 0 or ZNB, owned by
// wave e mod 4 of rank (e / 4) mod G)
__device__ __forceinline__ bool zt_live(int g, int G, int i)
{
    for (int q = 4 * g; q < 2 * ZNB; q += 4 * G)
        if ((q % ZNB) < i) return true;
    return false;
}

// ---------------------------------------------------------------------------------------
// Row pass of column j (rows of this workgroup's 64-row block): C part finishes w_{j-1}, A part
// updates column j (diagonal into Dfin, x_j below it) and produces its norm and W^H x / V^H x
// partials.  Same hand-over rules as the real row pass.
// ---------------------------------------------------------------------------------------
template <int KQ>
__device__ bool zrow_pass(double2* __restrict__ A, int64_t lda, int64_t n, int64_t c0, int64_t j, int i, bool doC,
                          bool doA, const ZPWork& w, int g, int G, ZPShared& sh, ZRowRegs<KQ>& rr, const PanelMap& pm,
                          int* err)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = wave_id();
    const int64_t jm = j - 1;
    const int im = i - 1;
    const ZScal cs = sh.cur;
    double2* __restrict__ Wm = w.W;
    double2* __restrict__ V2 = w.V2;
    const int64_t ldw = w.ldw;
    const int64_t b0 = j / TB;
    const int64_t nt = cdiv(n, TB);
    const int64_t b = c0 / TB + g;        // this workgroup's block (one per workgroup)
    if (b >= nt || b < b0) return true;
    const int k0 = wave * KQ, k1 = min(i, k0 + KQ);
    const int kc = doC ? min(k1, im) : k0;
    constexpr int PB = 8;
    const TallGrid& tg = pm.tg;
    const int64_t r = b * TB + lane;
    const bool rin = r < n;
    const bool rowC = doC && rin && r >= j;
    // ---- 0. empty the block's slots of column j + 2 ----
    if (doA) {
        if (wave == 0) {
            if (rin) cempty(x_of(w, j + 2) + r);
            if (lane == 0) st_co(pnorm_of(w, j + 2) + b, empty_slot());
            if (lane == 1) cempty(pq_of(w, j + 2) + b);
        }
        if (lane < KQ) {
            double2* ps2 = ps_of(w, j + 2) + b * (2 * ZNB) + k0 + lane;
            cempty(ps2);
            cempty(ps2 + ZNB);
        }
    }
    // ---- 1. registers: the block's rows of V and W, x of column j-1, old column j ----
    cd vreg[KQ], wreg[KQ];
#pragma unroll
    for (int u = 0; u < KQ; u++) {
        vreg[u] = rr.v[(wave * KQ + u) * 64 + lane];
        wreg[u] = rr.w[(wave * KQ + u) * 64 + lane];
    }
    const cd xo = rr.xprev, xold = rr.xnext;
    // column j + 1 from its diagonal (as at the launch's start): q's partial here, the next pass's input
    const cd xn1 = (doA && rin && r >= j + 1 && j + 1 < n) ? zc::ld(A + r + (j + 1) * lda) : czero();
    cd pp = czero();
    cd tv1 = czero(), tv2 = czero(), wraw = czero(), vraw = czero(), sp[4] = {czero(), czero(), czero(), czero()};
    double pmu = 0.0;
    cd ajj = czero();
    if (doC) {
        const int S = (int)((b * TB) / MR);
        const int nsl = (int)pm.nslot_own;
        const int nts = (int)tg.nts, ldp = (int)w.ldp;
        const int rr32 = (int)(rowC ? r : j);
        const int S0 = (int)(j / MR), j32 = (int)j;
        const double2* prowm = prow_of(w, jm);
        const double2* pcolm = pcol_of(w, jm);
        const double2* tvm = tv_of(w, jm);
        const double* pmum = pmu_of(w, jm);
        const int kt = (tid < im) ? tid : 0;
        const int Ge = (int)pm.Ge;
        const int gm = (tid < Ge) ? tid : 0;
        ajj = zc::ld(A + j + j * lda);   // A[j, j] as at the launch's start (not written during it)
        bool alive = true;
        for (unsigned spins = 0;; spins++) {
            // every load of a round first, then the checks (see the real row pass)
            if (im > 0) {
                tv1 = cld(tvm + kt);
                tv2 = cld(tvm + ZNB + kt);
                wraw = cld(Wm + j + kt * ldw);
                vraw = cld(V2 + j + kt * ldw);
            }
            pmu = ld_co(pmum + gm);
            if (wave == 0) {
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int I = S0 + lane + 64 * u;
                    sp[u] = cld(pcolm + ((I < nts ? I : S0) * ldp + j32));
                }
            }
            cd sb[8], buf[PB];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int sl = wave + 4 * u;
                sb[u] = cld(prowm + ((sl < nsl ? sl : 0) * ldp + rr32));
            }
#pragma unroll
            for (int u = 0; u < PB; u++) {
                const int I = S + wave + 4 * u;
                buf[u] = cld(pcolm + ((I < nts ? I : S) * ldp + rr32));
            }
            unsigned miss = 0;
            if (tid < im) {
                if (!(cfilled(tv1) && cfilled(tv2))) miss |= 1;
                if (!(cfilled(wraw) && cfilled(vraw))) miss |= 2;
            }
            if (tid < Ge && !filled(pmu)) miss |= 1;
            if (wave == 0) {
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (S0 + lane + 64 * u < nts && !cfilled(sp[u])) miss |= 4;
            }
            pp = czero();
            if (rowC) {
                cd s8 = czero(), s4[4] = {czero(), czero(), czero(), czero()};
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    if (wave + 4 * u >= nsl) sb[u] = czero();
                    if (!cfilled(sb[u])) miss |= 8;
                    s8 = cadd(s8, sb[u]);
                }
#pragma unroll
                for (int u = 0; u < PB; u++) {
                    if (S + wave + 4 * u >= nts) buf[u] = czero();
                    if (!cfilled(buf[u])) miss |= 16;
                    s4[u & 3] = cadd(s4[u & 3], buf[u]);
                }
                pp = cadd(s8, cadd(cadd(s4[0], s4[1]), cadd(s4[2], s4[3])));
                // rare extra batches: more than 32 row slots, more than 32 strips below the block
                for (int s0 = wave + 32; s0 < nsl; s0 += 32) {
                    cd xb[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) xb[u] = cld(prowm + ((s0 + 4 * u < nsl ? s0 + 4 * u : 0) * ldp + rr32));
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        if (s0 + 4 * u >= nsl) xb[u] = czero();
                        if (!cfilled(xb[u])) miss |= 8;
                        pp = cadd(pp, xb[u]);
                    }
                }
                for (int I0 = S + wave + 4 * PB; I0 < nts; I0 += 4 * PB) {
                    cd xb[PB];
#pragma unroll
                    for (int u = 0; u < PB; u++) xb[u] = cld(pcolm + ((I0 + 4 * u < nts ? I0 + 4 * u : S) * ldp + rr32));
#pragma unroll
                    for (int u = 0; u < PB; u++) {
                        if (I0 + 4 * u >= nts) xb[u] = czero();
                        if (!cfilled(xb[u])) miss |= 16;
                        pp = cadd(pp, xb[u]);
                    }
                }
            }
            if (__all(miss == 0)) break;
            if (!poll_more(spins, err, poll_code(POLL_ROW, j))) {
                poll_report(err, miss, r);
                alive = false;
                break;
            }
        }
        drain();
        if (__syncthreads_or(!alive)) return false;
        if (wave == 0) {
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (S0 + lane + 64 * u >= nts) sp[u] = czero();
        }
        // m of column j-1: the matvec partials, in a fixed order
        const double ms = wave_sum((tid < Ge) ? pmu : 0.0);
        if (lane == 0) sh.mu[wave] = ms;
    }
    // ---- 2. shared inputs into LDS, the special row of w_{j-1} ----
    if (doC && tid < im) {
        // t of column j-1 = unit * conj(row j of W, V) + scale * (sums over rows > j of W^H x, V^H x)
        sh.t[tid] = cadd(cd{cs.unit * wraw.re, -cs.unit * wraw.im}, cmul(cs.scale, tv1));
        sh.t[ZNB + tid] = cadd(cd{cs.unit * vraw.re, -cs.unit * vraw.im}, cmul(cs.scale, tv2));
        sh.wrow[tid] = wraw;
        sh.vrow[tid] = vraw;
    }
    __syncthreads();
    if (doC && wave == 0) {
        cd acc = czero();
#pragma unroll
        for (int u = 0; u < 4; u++) acc = cadd(acc, sp[u]);
        // y_j: column partials, the diagonal A[j, j] v_j (v_j = unit), the panel corrections
        cd cor = czero();
        double tt = 0.0;   // Im(t2^H t1) over the panel
        for (int k = lane; k < im; k += 64) {
            cor = cadd(cor, csub(cmul(sh.vrow[k], sh.t[k]), cmul(sh.wrow[k], sh.t[ZNB + k])));
            tt += cmulc(sh.t[k], sh.t[ZNB + k]).im;   // t1 conj(t2), its imaginary part
        }
        // the column partials are of the unscaled x
        acc = cwave_sum(cadd(cmul(cs.scale, acc), cor));
        tt = wave_sum(tt);
        if (lane == 0) {
            // Im(v^H A22 v) for v = s x + c e_{j}: |s|^2 Im(x^H A22 x) + 2 Im(conj(s) c q) + |c|^2 Im(A[j, j])
            const double s2 = cs.scale.re * cs.scale.re + cs.scale.im * cs.scale.im;
            const double c2 = cs.corr.re * cs.corr.re + cs.corr.im * cs.corr.im;
            const cd scq = cmul(cconj(cs.scale), cmul(cs.corr, cs.q));
            const double m = s2 * ((sh.mu[0] + sh.mu[1]) + (sh.mu[2] + sh.mu[3])) + 2.0 * scq.im + c2 * ajj.im + 2.0 * tt;
            const double t2 = cs.tau.re * cs.tau.re + cs.tau.im * cs.tau.im;
            const cd c = cd{0.0, -0.5 * t2 * m};
            const cd yj = cadd(acc, cd{ajj.re * cs.unit, ajj.im * cs.unit});
            const cd ws = cadd(cmul(cs.tau, yj), cd{c.re * cs.unit, c.im * cs.unit});
            sh.wspec = ws;
            sh.cfix = c;
            sh.wrow[im] = ws;
            sh.vrow[im] = cd{cs.unit, 0.0};
        }
    }
    // ---- C part: w_{j-1} on this block; with it, the A part's terms of the reflectors before the
    // newest one (k < ka), which need nothing from the C part: one barrier publishes both sums.  The
    // LDS operands are read in groups of 4 k before use and masked by select (entries past the live
    // ones may hold anything) ----
    const bool live = doA && rin && r >= j;
    const int ka = doC ? im : i;
    {
        constexpr int KG = 4;
        cd c4 = czero(), p4 = czero();
#pragma unroll
        for (int u0 = 0; u0 < KQ; u0 += KG) {
            cd ta[KG], tb[KG], wa[KG], va[KG];
#pragma unroll
            for (int q = 0; q < KG; q++) {
                ta[q] = sh.t[k0 + u0 + q];
                tb[q] = sh.t[ZNB + k0 + u0 + q];
                wa[q] = sh.wrow[k0 + u0 + q];
                va[q] = sh.vrow[k0 + u0 + q];
            }
#pragma unroll
            for (int q = 0; q < KG; q++) {
                const int u = u0 + q;
                if (doC) {
                    const cd term = csub(cmul(vreg[u], ta[q]), cmul(wreg[u], tb[q]));
                    c4 = cadd(c4, (rowC && k0 + u < kc) ? term : czero());
                }
                if (doA) {
                    const cd term = csub(cmulc(vreg[u], wa[q]), cmulc(wreg[u], va[q]));
                    p4 = cadd(p4, (live && k0 + u < ka) ? term : czero());
                }
            }
        }
        // y_r = s (A22 x)_r + c A[r, j] (xold = A[r, j] as at the launch's start); pp is this wave's part
        // of the partials, the c term is wave 0's
        if (doC) sh.red[wave][lane] = rowC ? cadd(cadd(cmul(cs.scale, pp), (wave == 0) ? cmul(cs.corr, xold) : czero()), c4) : czero();
        if (doA) sh.red2[wave][lane] = p4;
    }
    __syncthreads();
    cd wr = czero(), vr = czero();
    if (rowC) {
        const cd sacc = cadd(cadd(sh.red[0][lane], sh.red[1][lane]), cadd(sh.red[2][lane], sh.red[3][lane]));
        vr = (r == j) ? cd{cs.unit, 0.0} : cmul(cs.scale, xo);
        wr = (r == j) ? sh.wspec : cadd(cmul(cs.tau, sacc), cmul(sh.cfix, vr));
        if (wave == 0) {
            cst(Wm + r + im * ldw, wr);
            cst(V2 + r + im * ldw, vr);
            cst(A + r + jm * lda, vr);
        }
    }
    if (doC) {
#pragma unroll
        for (int u = 0; u < KQ; u++) {
            if (k0 + u == im) {
                vreg[u] = vr;
                wreg[u] = wr;
            }
        }
    }
    cd x = czero();
    if (doA) {
        // ---- A part: column j (rows >= j) updated with the panel's reflectors 0..i-1: the sums over
        // k < ka from before the barrier, plus the newest reflector's term
        // V[r,im] conj(W[j,im]) - W[r,im] conj(V[j,im]) ----
        if (live) {
            x = cadd(xold, cadd(cadd(sh.red2[0][lane], sh.red2[1][lane]), cadd(sh.red2[2][lane], sh.red2[3][lane])));
            if (doC) x = cadd(x, csub(cmulc(vr, sh.wspec), cmulc(wr, cd{cs.unit, 0.0})));
        }
        if (wave == 0 && live) {
            if (r == j) w.Dfin[j] = make_double2(x.re, x.im);   // final diagonal entry of column j
            else cst(x_of(w, j) + r, x);
        }
        const bool tail = rin && r >= j + 2;
        if (wave == 0) {
            const double ss = wave_sum(tail ? zc::norm2(x) : 0.0);
            if (lane == 0) st_co(pnorm_of(w, j) + b, ss);
        }
        if (wave == 1) {   // the block's part of q = x^H A22 e_{j+1} (rows >= j + 1)
            const cd qb = cwave_sum((live && r >= j + 1) ? cmulc(xn1, x) : czero());
            if (lane == 0) cst(pq_of(w, j) + b, qb);
        }
        // W^H x and V^H x of the block's rows >= j+2: values 0..7 Re / 8..15 Im of conj(W) x,
        // 16..23 Re / 24..31 Im of conj(V) x (one k-group of KQ = 8 per wave)
        const cd xt = tail ? x : czero();
        double2* psb = ps_of(w, j) + b * (2 * ZNB);
        double v32[32];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const bool ok = u < KQ && k0 + u < k1;
            const cd pw = ok ? cmulc(xt, wreg[u < KQ ? u : 0]) : czero();   // x conj(W)
            const cd pv = ok ? cmulc(xt, vreg[u < KQ ? u : 0]) : czero();
            v32[u] = pw.re;
            v32[8 + u] = pw.im;
            v32[16 + u] = pv.re;
            v32[24 + u] = pv.im;
        }
        const double tot = wave_sum32(v32);
        // lane l holds the total of value (l >> 1) & 31; the Re and Im halves of one entry sit 8
        // values apart, so they are paired through this wave's LDS row (no other wave touches it)
        sh.psx[wave][(lane >> 1) & 31] = tot;
        __builtin_amdgcn_wave_barrier();
        if (lane < 16) {
            const int h = lane >> 3, u = lane & 7;   // h 0: W part, 1: V part
            if (k0 + u < k1) cst(psb + (h ? ZNB : 0) + k0 + u, cd{sh.psx[wave][16 * h + u], sh.psx[wave][16 * h + 8 + u]});
        }
    }
    // keep the block's rows for the next column of the panel; prefetch its input column
    if (doC) {
#pragma unroll
        for (int u = 0; u < KQ; u++) {
            if (k0 + u == im) {
                rr.v[(wave * KQ + u) * 64 + lane] = vreg[u];
                rr.w[(wave * KQ + u) * 64 + lane] = wreg[u];
            }
        }
    }
    rr.xprev = x;
    rr.xnext = xn1;
    __syncthreads();
    return true;
}

// local row of entry k of a lane's tile: rows 32 (k>>1) + 16 (k&1) + p, so that the 16 lanes of a
// column read 16 consecutive entries (256 bytes) per load
__device__ __forceinline__ int ztile_row(int k, int p) { return 32 * (k >> 1) + 16 * (k & 1) + p; }

// raw tile entries of one lane (16 complex, column c), unconditional and in bounds; masked at use
__device__ __forceinline__ void ztall_load(const double2* __restrict__ A, int64_t lda, int64_t n, int64_t I, int64_t c,
                                           int p, cd a[16])
{
    const double2* col = A + ((c < n) ? c : n - 1) * lda + I * MR;
    const int rn = (int)(n - I * MR);
#pragma unroll
    for (int h = 0; h < 8; h++) {
        const int r0 = ztile_row(2 * h, p), r1 = ztile_row(2 * h + 1, p);
        const double2 t0 = col[(r0 < rn) ? r0 : -(int)(I * MR)];
        const double2 t1 = col[(r1 < rn) ? r1 : -(int)(I * MR)];
        a[2 * h] = cd{t0.x, t0.y};
        a[2 * h + 1] = cd{t1.x, t1.y};
    }
}

// unscaled x_c of column j on the rows of a strip (rows past j + 1; row j + 1 meets only masked columns)
__device__ __forceinline__ cd zxraw(const double2* xj, int64_t n, int64_t j, int64_t c)
{
    return (c > j + 1 && c < n) ? cld(xj + c) : czero();
}

// unscaled x_c of column j for a tile column: alpha at c = j + 1, 0 outside (j, n)
__device__ __forceinline__ cd zxcol(const double2* xj, int64_t n, int64_t j, int64_t c)
{
    return (c > j && c < n) ? cld(xj + c) : czero();
}

// ---------------------------------------------------------------------------------------
// Phase B of column j: t, the matvec y = A22 x over the trailing block's lower triangle with its
// diagonal, on the unscaled x (row sums A[r,c] x_c for r >= c, column sums -conj(A[r,c]) x_r for
// r > c, m partials sum 2 Im(conj(x_r) A[r,c] x_c) (r > c) + Im(A[c,c]) |x_c|^2), then the scalars.
// The row pass forms A22 v = s A22 x + c A22 e_{j+1} (v = s x + c e_{j+1}) and m from the same parts.
// ---------------------------------------------------------------------------------------
__device__ bool zphase_b(const double2* __restrict__ A, int64_t lda, int64_t n, int64_t c0, int64_t j, int i,
                         const ZPWork& w, ZPShared& sh, double2* tau_out, double* eta_out, int g, int G,
                         const PanelMap& pm, bool rown, int* err)
{
    const bool mv = g < pm.Ge;
    if (!(mv || rown || zt_live(g, G, i))) return true;
    const int tid = threadIdx.x, lane = tid & 63, wave = wave_id();
    const int p = lane & 15, cg = lane >> 4;
    const int64_t b0 = j / TB;
    const int64_t nt = cdiv(n, TB);
    const int64_t nq = nt - b0;
    const TallGrid& tg = pm.tg;
    const int64_t tlo = pm.tlo, thi = pm.thi;
    int64_t I = pm.I, J = pm.J, cur_slot = pm.slot, c = 0;
    cd a[16];
    if (mv) {
        c = J * MC + 4 * wave + cg;
        ztall_load(A, lda, n, I, c, p, a);
    }
    const double* pnj = pnorm_of(w, j);
    const double2* pqj = pq_of(w, j);
    const double2* psj = ps_of(w, j);
    const double2* xj = x_of(w, j);
    const int e0 = 4 * g + wave;
    const bool t0 = e0 < 2 * ZNB && (e0 % ZNB) < i;
    const int64_t rs = I * MR + tid;
    const bool in0 = lane < nq, in1 = lane + 64 < nq;
    const int64_t q0 = b0 + (in0 ? lane : 0), q1 = b0 + (in1 ? lane + 64 : 0);
    const bool xcin = mv && c > j && c < n, xsin = mv && rs > j + 1 && rs < n;
    const double2* pxc = xj + (xcin ? c : j + 1);
    const double2* pxs = xj + (xsin ? rs : j + 1);
    const int64_t e0c = t0 ? e0 : 0;
    cd ps0 = czero(), ps1 = czero(), xc = czero(), xs = czero();
    bool alive = true;
    // ---- poll this wave's t partials and the first tile's x (the matvec runs on the unscaled x) ----
    for (unsigned spins = 0;; spins++) {
        if (t0) {
            ps0 = cld(psj + q0 * (2 * ZNB) + e0c);
            ps1 = cld(psj + q1 * (2 * ZNB) + e0c);
        }
        if (mv) {
            xc = cld(pxc);
            xs = cld(pxs);
        }
        unsigned miss = 0;
        if (t0 && ((in0 && !cfilled(ps0)) || (in1 && !cfilled(ps1)))) miss |= 128;
        if (xcin && !cfilled(xc)) miss |= 256;
        if (xsin && !cfilled(xs)) miss |= 512;
        if (__all(miss == 0)) break;
        if (!poll_more(spins, err, poll_code(POLL_B, j))) {
            poll_report(err, miss, (miss & 256) ? c : rs);
            alive = false;
            break;
        }
    }
    if (!in0) ps0 = czero();
    if (!in1) ps1 = czero();
    if (!xcin) xc = czero();
    if (!xsin) xs = czero();
    drain();
    if (mv) sh.vstrip[tid] = xs;
    // ---- (1) t of the live elements of this wave (unscaled: the consumer applies scale) ----
    for (int e = e0; e < 2 * ZNB && alive; e += 4 * G) {
        if ((e % ZNB) >= i) continue;   // wave-uniform
        cd acc;
        if (e == e0) {
            acc = cadd(ps0, ps1);
        } else {
            acc = czero();
            for (int64_t q = lane; q < 128 && q < nq; q += 64) {
                const double2* pq = psj + (b0 + q) * (2 * ZNB) + e;
                cd v = cld(pq);
                v.re = settle(v.re, reinterpret_cast<const double*>(pq), err, poll_code(POLL_SETTLE, j));
                v.im = settle(v.im, reinterpret_cast<const double*>(pq) + 1, err, poll_code(POLL_SETTLE, j));
                acc = cadd(acc, v);
            }
        }
        for (int64_t q = lane + 128; q < nq; q += 64) {
            const double2* pq = psj + (b0 + q) * (2 * ZNB) + e;
            cd v = cld(pq);
            v.re = settle(v.re, reinterpret_cast<const double*>(pq), err, poll_code(POLL_SETTLE, j));
            v.im = settle(v.im, reinterpret_cast<const double*>(pq) + 1, err, poll_code(POLL_SETTLE, j));
            acc = cadd(acc, v);
        }
        acc = cwave_sum(acc);
        if (lane == 0) cst(tv_of(w, j) + e, acc);
    }
    if (__syncthreads_or(!alive)) return false;   // x of the first strip is in LDS
    // ---- (2) matvec on the unscaled x ----
    if (mv) {
        cd racc[16];
#pragma unroll
        for (int k = 0; k < 16; k++) racc[k] = czero();
        double mu = 0.0;
        cd vc = xc;   // alpha at j + 1, 0 outside (j, n)
        for (int64_t t = tlo; t < thi; t++) {
            const bool more = t + 1 < thi;
            const bool last_in_strip = (J == MPS * I + MPS - 1);
            const int64_t In = last_in_strip ? I + 1 : I;
            const int64_t Jn = last_in_strip ? tg.Jlo : J + 1;
            const int64_t cn = Jn * MC + 4 * wave + cg;
            cd an[16], xcn = czero(), xsn = czero();
            if (more) {
                xcn = zxcol(xj, n, j, cn);
                ztall_load(A, lda, n, In, cn, p, an);
                if (In != I) xsn = zxraw(xj, n, j, In * MR + tid);
            }
            cd cp = czero();
            {
                const bool cok = c > j && c < n;
                const int64_t lo = cok ? c - I * MR : (int64_t)MR, hi = n - I * MR;
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    const int rl = ztile_row(k, p);
                    const bool below = rl > lo && rl < hi;
                    const bool diag = cok && rl == lo;
                    const cd av = (below || diag) ? a[k] : czero();
                    racc[k] = cadd(racc[k], cmul(av, vc));
                    const cd vr = sh.vstrip[rl];
                    if (below) {
                        cp = cadd(cp, cmul(cconj(av), vr));              // conj(A[r,c]) v_r
                        mu += 2.0 * cmulc(cmul(av, vc), vr).im;          // Im(conj(v_r) A[r,c] v_c)
                    } else if (diag) {
                        mu += av.im * zc::norm2(vc);
                    }
                }
            }
            cp = cd{row16_sum(cp.re), row16_sum(cp.im)};
            if (p == 0 && c > j && c < n) cst(pcol_of(w, j) + I * w.ldp + c, cd{-cp.re, -cp.im});
            if (!more || In != I) {
                // flush: each (wave, column group) writes its 16 rows to its own LDS row, thread t sums
                // the 16 values of strip row t in a fixed order (an LDS transposition)
                {
                    cd* mr = sh.mred[wave * 4 + cg];
#pragma unroll
                    for (int k = 0; k < 16; k++) {
                        mr[ztile_row(k, p)] = racc[k];
                        racc[k] = czero();
                    }
                }
                __syncthreads();
                {
                    const int64_t r = I * MR + tid;
                    cd q4[4];
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        q4[q] = cadd(cadd(sh.mred[4 * q][tid], sh.mred[4 * q + 1][tid]), cadd(sh.mred[4 * q + 2][tid], sh.mred[4 * q + 3][tid]));
                    if (r < n) cst(prow_of(w, j) + cur_slot * w.ldp + r, cadd(cadd(q4[0], q4[1]), cadd(q4[2], q4[3])));
                    if (more) {
                        const double2* px = xj + In * MR + tid;
                        xsn.re = settle(xsn.re, reinterpret_cast<const double*>(px), err, poll_code(POLL_SETTLE, j));
                        xsn.im = settle(xsn.im, reinterpret_cast<const double*>(px) + 1, err, poll_code(POLL_SETTLE, j));
                        sh.vstrip[tid] = xsn;
                    }
                }
                __syncthreads();
                cur_slot = 0;
            }
            if (more) {
#pragma unroll
                for (int k = 0; k < 16; k++) a[k] = an[k];
                const double2* px = xj + cn;
                xcn.re = settle(xcn.re, reinterpret_cast<const double*>(px), err, poll_code(POLL_SETTLE, j));
                xcn.im = settle(xcn.im, reinterpret_cast<const double*>(px) + 1, err, poll_code(POLL_SETTLE, j));
                vc = xcn;
            }
            I = In;
            J = Jn;
            c = cn;
        }
        // this workgroup's partial of m (of the unscaled x)
        mu = wave_sum(mu);
        if (lane == 0) sh.mu[wave] = mu;
        __syncthreads();
        if (tid == 0) st_co(pmu_of(w, j) + g, (sh.mu[0] + sh.mu[1]) + (sh.mu[2] + sh.mu[3]));
    }
    // ---- (3) scalars of column j after the matvec, which needs none of them (the next row pass's
    // inputs, with q = x^H A22 e_{j+1} from the block partials).  Having seen every block's norm
    // partial of column j tells this rank that every row pass of column j has finished, which the slot
    // emptying below relies on ----
    if (wave == 0) {
        double pn0 = 0.0, pn1 = 0.0;
        cd araw = czero(), pq0 = czero(), pq1 = czero();
        bool ok = alive;
        for (unsigned spins = 0; ok; spins++) {
            pn0 = ld_co(pnj + q0);
            pn1 = ld_co(pnj + q1);
            pq0 = cld(pqj + q0);
            pq1 = cld(pqj + q1);
            araw = cld(xj + (j + 1));
            unsigned miss = 0;
            if ((in0 && !(filled(pn0) && cfilled(pq0))) || (in1 && !(filled(pn1) && cfilled(pq1)))) miss |= 32;
            if (!cfilled(araw)) miss |= 64;
            if (__all(miss == 0)) break;
            if (!poll_more(spins, err, poll_code(POLL_B, j))) {
                poll_report(err, miss, j + 1);
                ok = false;
            }
        }
        if (ok) {
            double acc = (in0 ? pn0 : 0.0) + (in1 ? pn1 : 0.0);
            cd qa = cadd(in0 ? pq0 : czero(), in1 ? pq1 : czero());
            for (int64_t q = lane + 128; q < nq; q += 64) {
                acc += settle(ld_co(pnj + b0 + q), pnj + b0 + q, err, poll_code(POLL_SETTLE, j));
                const double2* pp = pqj + b0 + q;
                cd v = cld(pp);
                v.re = settle(v.re, reinterpret_cast<const double*>(pp), err, poll_code(POLL_SETTLE, j));
                v.im = settle(v.im, reinterpret_cast<const double*>(pp) + 1, err, poll_code(POLL_SETTLE, j));
                qa = cadd(qa, v);
            }
            acc = wave_sum(acc);
            qa = cwave_sum(qa);
            if (lane == 0) {
                ZScal sc = zc::zreflector(araw, acc);
                sc.q = qa;
                sh.cur = sc;
                if (g == 0) {
                    tau_out[j] = make_double2(sc.tau.re, sc.tau.im);
                    eta_out[j] = sc.beta;
                }
            }
        }
        alive = ok;
    }
    if (__syncthreads_or(!alive)) return false;
    // column j - 1's t was consumed by the row passes of column j, all finished (above)
    if (lane == 0)
        for (int e = e0; e < 2 * ZNB; e += 4 * G) cempty(tv_of(w, j + 2) + e);
    if (!mv) return true;
    // ---- empty this rank's matvec hand-over slots of column j + 2 (copy (j + 2) mod 3 = (j - 1) mod 3)
    // after its own outputs of column j are on their way, off the path from the poll to the matvec:
    // column j - 1's values there were consumed by the row passes of column j, which all finished
    // before their norm partials of column j, which wave 0 saw before the barrier above; the
    // next phase B drains these stores before storing anything of column j + 1 ----
    {
        double2* prn = prow_of(w, j + 2);
        double2* pcn = pcol_of(w, j + 2);
        int64_t It = pm.I, Jt = pm.J, slot = pm.slot;
        for (int64_t t = tlo; t < thi; t++) {
            if (wave == 0 && lane < MC) {
                const int64_t cc = Jt * MC + lane;
                if (cc > j + 2 && cc < n) cempty(pcn + It * w.ldp + cc);
            }
            const bool lis = (Jt == MPS * It + MPS - 1);
            if (lis || t + 1 == thi) {
                const int64_t rr = It * MR + tid;
                if (rr < n) cempty(prn + slot * w.ldp + rr);
                slot = 0;
            }
            if (lis) {
                It++;
                Jt = tg.Jlo;
            } else {
                Jt++;
            }
        }
        if (tid == 0) st_co(pmu_of(w, j + 2) + g, empty_slot());
    }
    __syncthreads();
    return true;
}

}  // namespace zp
}  // namespace skh
