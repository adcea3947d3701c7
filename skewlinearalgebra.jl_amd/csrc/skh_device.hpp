// skh_device.hpp -- device building blocks of the blocked skew tridiagonalization (gfx950, FP64).
//
// The factorization follows the blocked Householder reduction of SkewLinearAlgebra.jl
// (src/hessenberg.jl:92-174): panels of nb columns, each column computing the reflector of
// the virtually updated column, w = tau * (A22 v + V (W^T v) - W (V^T v)), and after each
// panel the skew rank-2k trailing update A22 += V2 W2^T - W2 V2^T (lower triangle only).
//
// Every phase below is written for a team of G workgroups (rank g in [0, G)); the large
// engine runs it with G = the whole (co-resident) grid and a grid barrier between phases.
//
// Memory: A column-major, lda even, base 16-byte aligned.  Only the strictly lower triangle
// of the trailing block is read (the real skew diagonal is zero, the upper half is -A^T).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "skh_sync.hpp"

namespace skh {

constexpr int WG = 256;       // threads per workgroup (4 waves of 64)
constexpr int TB = 64;        // matvec tile edge and row-chunk height
constexpr int NBMAX = 128;    // largest panel width

typedef double v4d __attribute__((ext_vector_type(4)));

// Device workspace of one factorization (one matrix).  Every value one workgroup of the panel
// kernel hands to another travels through a slot that holds the empty marker until the value is
// stored (see EMPTY_BITS), so consumers poll their inputs instead of meeting at grid barriers:
//   row pass of column j -> phase B of column j:  Pnorm, PS and X of column j (OWN_COPIES copies,
//                                                  one per column mod 4) and the rows of W and V2;
//   phase B of column j -> row pass of column j+1: Tv, Prow and Pcol of column j (HAND_COPIES
//                                                  copies, one per column mod 3).
// A slot of column j + OWN_COPIES (j + HAND_COPIES) is refilled with the empty marker by its own
// producer once every consumer of column j's value is known to be done (row_pass and phase_b
// state when).
constexpr int OWN_COPIES = 4, HAND_COPIES = 3;
struct Work {
    double* W;       // n x NBMAX (ldw)   w of the panel's reflectors (empty marker where not stored yet)
    double* V2;      // n x NBMAX (ldw)   the panel's reflectors (as stored in A), for the row hand-over
    int64_t ldw;
    double* Pnorm;   // [nt]              per-block sums of squares of the reflector tail
    double* PS;      // [nt][2*NBMAX]     per-block W^T x and V^T x partials
    double* X;       // [ld]              x_j, the updated column j (unscaled reflector)
    int64_t ostride; // Pnorm, PS and X of column j at + (j mod 4) * ostride
    double* Tv;      // [2*NBMAX]         partial t1 = W^T x, t2 = V^T x (unscaled; the consumer scales, adds the head)
    double* Prow;    // [nseg][ldp]       row partials of the skew matvec (one slot per rank per strip)
    double* Pcol;    // [nts][ldp]        column partials of the skew matvec
    int64_t ldp;
    int64_t hstride; // Tv, Prow and Pcol of column j at + (j mod 3) * hstride
};

__device__ __forceinline__ int64_t own_copy(int64_t j) { return j & (OWN_COPIES - 1); }
__device__ __forceinline__ int64_t hand_copy(int64_t j) { return j % HAND_COPIES; }
__device__ __forceinline__ double* pnorm_of(const Work& w, int64_t j) { return w.Pnorm + own_copy(j) * w.ostride; }
__device__ __forceinline__ double* ps_of(const Work& w, int64_t j) { return w.PS + own_copy(j) * w.ostride; }
__device__ __forceinline__ double* x_of(const Work& w, int64_t j) { return w.X + own_copy(j) * w.ostride; }
__device__ __forceinline__ double* tv_of(const Work& w, int64_t j) { return w.Tv + hand_copy(j) * w.hstride; }
__device__ __forceinline__ double* prow_of(const Work& w, int64_t j) { return w.Prow + hand_copy(j) * w.hstride; }
__device__ __forceinline__ double* pcol_of(const Work& w, int64_t j) { return w.Pcol + hand_copy(j) * w.hstride; }

// Per-column scalars kept in LDS across the phases.  The matvec runs on the unscaled column x
// (x[j+1] = alpha at the reflector's unit position), so its consumer forms
// A v = scale * (A x) + corr * A[:, j+1] with corr = unit - scale * alpha.
struct ColScal {
    double tau, scale, unit, beta, corr;
};

// ---------------------------------------------------------------------------------------
// Cross-lane sums on the VALU: DPP lane patterns inside rows of 16 lanes (quad_perm xor 1 and
// xor 2, row_half_mirror i <-> 7-i, row_mirror i <-> 15-i) and the gfx950 permlane swaps across
// rows (odd <-> even rows of 16, upper <-> lower 32 lanes).  A DPP wave sum takes about 80 ns
// against about 200 ns for the ds_bpermute butterfly (measured on MI355X).  Every step adds a
// lane's value and its partner's, so partner lanes end with bit-identical totals.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned lo32(double x) { return (unsigned)__builtin_bit_cast(unsigned long long, x); }
__device__ __forceinline__ unsigned hi32(double x) { return (unsigned)(__builtin_bit_cast(unsigned long long, x) >> 32); }
__device__ __forceinline__ double mk64(unsigned lo, unsigned hi)
{
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

// x of the lane selected by the DPP pattern CTRL (within each row of 16 lanes)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)lo32(x), CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)hi32(x), CTRL, 0xF, 0xF, false);
    return mk64(lo, hi);
}
constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1,0,3,2]
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2,3,0,1]
constexpr int DPP_HALF_MIRROR = 0x141; // lane i <-> 7 - i within each 8
constexpr int DPP_MIRROR = 0x140;      // lane i <-> 15 - i within each 16

// permlane swaps of (a, b): for the halving exchanges below, lanes of the first row (or half)
// of each pair end with (a', b') = (own a, partner's a), lanes of the second with
// (partner's b, own b), so a' + b' is the kept half of the pair summed over both lanes.
__device__ __forceinline__ void swap16(double a, double b, double& ra, double& rb)
{
    const auto l = __builtin_amdgcn_permlane16_swap(lo32(a), lo32(b), false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(hi32(a), hi32(b), false, false);
    ra = mk64(l[0], h[0]);
    rb = mk64(l[1], h[1]);
}
__device__ __forceinline__ void swap32(double a, double b, double& ra, double& rb)
{
    const auto l = __builtin_amdgcn_permlane32_swap(lo32(a), lo32(b), false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi32(a), hi32(b), false, false);
    ra = mk64(l[0], h[0]);
    rb = mk64(l[1], h[1]);
}

// x + x of the lane 16 (32) away: every lane of a row pair (half pair) gets the same sum
__device__ __forceinline__ double xor16_sum(double x)
{
    double a, b;
    swap16(x, x, a, b);
    return a + b;
}
__device__ __forceinline__ double xor32_sum(double x)
{
    double a, b;
    swap32(x, x, a, b);
    return a + b;
}

// sum over each row of 16 lanes, in every lane of the row
__device__ __forceinline__ double row16_sum(double x)
{
    x += dpp_d<DPP_XOR1>(x);
    x += dpp_d<DPP_XOR2>(x);
    x += dpp_d<DPP_HALF_MIRROR>(x);
    x += dpp_d<DPP_MIRROR>(x);
    return x;
}

// sum over the 64 lanes of the wave, in every lane
__device__ __forceinline__ double wave_sum(double x)
{
    return xor32_sum(xor16_sum(row16_sum(x)));
}

// One halving step: lanes with `hi` keep the upper half of their values, the others the lower,
// and add the partner lane's copy of the kept half (partner selected by the DPP pattern).
template <int CTRL, int HALF>
__device__ __forceinline__ void halve_dpp(double v[], bool hi)
{
#pragma unroll
    for (int i = 0; i < HALF; i++) {
        const double lo_v = v[i], hi_v = v[i + HALF];
        const double keep = hi ? hi_v : lo_v;
        const double send = hi ? lo_v : hi_v;
        v[i] = keep + dpp_d<CTRL>(send);
    }
}

// Sum each of the 32 values v[0..31] over the 64 lanes of the wave (halving exchange: the
// upper/lower 32 lanes, the odd/even rows, then mirror, half mirror and xor-2 partners inside
// the rows, and a last xor-1 add).  Afterwards lane l holds the total of value (l >> 1) & 31
// (lanes l and l ^ 1 hold the same total).
__device__ __forceinline__ double wave_sum32(double v[32])
{
#pragma unroll
    for (int i = 0; i < 16; i++) {
        double a, b;
        swap32(v[i], v[i + 16], a, b);
        v[i] = a + b;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        double a, b;
        swap16(v[i], v[i + 8], a, b);
        v[i] = a + b;
    }
    const unsigned l = threadIdx.x;
    halve_dpp<DPP_MIRROR, 4>(v, (l & 8) != 0);
    halve_dpp<DPP_HALF_MIRROR, 2>(v, (l & 4) != 0);
    halve_dpp<DPP_XOR2, 1>(v, (l & 2) != 0);
    return v[0] + dpp_d<DPP_XOR1>(v[0]);
}

__device__ __forceinline__ int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// index of this thread's wave in the workgroup, as a wave-uniform (scalar) value
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// Phase clock of the panel kernel (workgroup 0, thread 0; clk == nullptr when off): slots
// [0, 13) accumulate 100 MHz wall-clock ticks, PH_INNER / PH_OUTER hold the last stamps of the
// inner (sub-phase) and outer (phase) clocks.
constexpr int PH_SLOTS = 16, PH_INNER = 14, PH_OUTER = 15;
__device__ __forceinline__ void clk_begin(unsigned long long* clk)
{
    if (clk) clk[PH_INNER] = wall_clock64();
}
__device__ __forceinline__ void clk_inner(unsigned long long* clk, int k)
{
    if (clk) {
        const unsigned long long t = wall_clock64();
        clk[k] += t - clk[PH_INNER];
        clk[PH_INNER] = t;
    }
}
__device__ __forceinline__ void clk_outer(unsigned long long* clk, int k)
{
    if (clk) {
        const unsigned long long t = wall_clock64();
        clk[k] += t - clk[PH_OUTER];
        clk[PH_OUTER] = t;
    }
}

// Data one workgroup of the panel kernel writes and another reads within the same launch (the
// partial sums, the current column x_j, the rows of V and W) moves through agent-scope relaxed
// atomic accesses (global_load/global_store with sc1): coherent across the XCDs' L2 caches
// on their own, so the grid barrier needs no release/acquire fence, whose L2 writeback and
// invalidate, issued by every workgroup, cost 20 us per barrier at 512 workgroups against
// 3.4 us without them (measured on MI355X).  Tile data of the trailing block is never written
// during a launch and keeps plain (cached, vectorised) loads.
__device__ __forceinline__ double ld_co(const double* p)
{
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_co(double* p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Hand-over slots hold their values exactly.  An empty slot holds the signalling NaN EMPTY, a bit
// pattern no arithmetic result has (FP64 arithmetic returns quiet NaNs), so a consumer polling a
// slot sees EMPTY or the value of its own column, never an older one, as long as the producer
// empties the slot after the older value's last reader and that store completes before any reader
// of the new value starts polling.  The stores complete because every wave executes
// `s_waitcnt vmcnt(0)` (drain) when it starts polling its own next inputs, and the order follows
// from the data: a consumer only polls column j's slots after it has seen a value that the
// producer stored after its drain.
constexpr unsigned long long EMPTY_BITS = 0x7FF0DEADBEEF0001ull;
__device__ __forceinline__ double empty_slot() { return __builtin_bit_cast(double, EMPTY_BITS); }
__device__ __forceinline__ bool filled(double x) { return __builtin_bit_cast(unsigned long long, x) != EMPTY_BITS; }
__device__ __forceinline__ void drain() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// One more round of a poll loop?  False (the launch is abandoned) after POLL_LIMIT rounds (about
// a second; a round is at least one memory round trip), which flags the error, or once another
// workgroup has flagged one (looked at every 64 rounds).  The first wait to run out records where
// (poll_code) in the flag.
constexpr unsigned POLL_LIMIT = 1u << 20;
enum PollSite { POLL_ROW = 1, POLL_B = 2, POLL_SETTLE = 3 };
__device__ __forceinline__ int poll_code(int site, int64_t j)
{
    return 1 | (site << 1) | ((int)(j & 0xFFFF) << 3) | ((int)(blockIdx.x & 0x3FF) << 19);
}
__device__ __forceinline__ bool poll_more(unsigned spins, int* err, int code)
{
    if (spins > POLL_LIMIT) {
        int expect = 0;
        __hip_atomic_compare_exchange_strong(err, &expect, code, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        return false;
    }
    if ((spins & 63u) == 63u && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
// s_sleep units (64 clocks) between poll rounds: fewer rounds while waiting leaves more of the memory
// system to the producers (16 measured 0.6% faster than 1 at n = 8192)
#ifndef SKH_POLL_SLEEP
#define SKH_POLL_SLEEP 16
#endif
#ifdef SKH_POLL_FAST
    if (spins < SKH_POLL_FAST) __builtin_amdgcn_s_sleep(1);
    else
#endif
    __builtin_amdgcn_s_sleep(SKH_POLL_SLEEP);
    return true;
}

// what an abandoned poll still missed (err[1]: OR of the missing input kinds, err[2]: 1 + an index)
__device__ __forceinline__ void poll_report(int* err, unsigned miss, int64_t idx)
{
    if (miss) {
        __hip_atomic_fetch_or(err + 1, (int)miss, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int expect = 0;
        __hip_atomic_compare_exchange_strong(err + 2, &expect, (int)(idx + 1), __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    }
}

// x (already loaded from p) once it is filled: polls p again in the rare case that the producer's
// store is still on its way (0 if the launch is abandoned)
__device__ __forceinline__ double settle(double x, const double* p, int* err, int code)
{
    if (__builtin_expect(!filled(x), 0)) {
        for (unsigned spins = 0;; spins++) {
            if (!poll_more(spins, err, code)) return 0.0;
            x = ld_co(p);
            if (filled(x)) break;
        }
    }
    return x;
}

// ---------------------------------------------------------------------------------------
// Reflector scalars from alpha = x[0] and the sum of squares of x[1:] (householder!,
// src/hessenberg.jl:26-44).  beta = (alpha > 0 ? -1 : +1) * hypot(alpha, xnorm),
// tau = 1 - alpha/beta, v = [1; x[1:]/(alpha - beta)]; all-zero x gives tau = beta = 0, v = 0.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ ColScal reflector_scalars(double alpha, double xn2)
{
    ColScal s;
    double xnorm = sqrt(xn2);
    if (xnorm != 0.0 || alpha != 0.0) {
        double b = ((alpha > 0.0) ? -1.0 : 1.0) * hypot(alpha, xnorm);
        s.beta = b;
        s.tau = 1.0 - alpha / b;
        s.scale = 1.0 / (alpha - b);
        s.unit = 1.0;
    } else {
        s.beta = 0.0;
        s.tau = 0.0;
        s.scale = 0.0;
        s.unit = 0.0;
    }
    s.corr = s.unit - s.scale * alpha;   // |scale * alpha| <= 1/2: no cancellation
    return s;
}

// Ownership of the matvec tiles (strip-major order over the lower triangle): T tiles split
// into G <= T contiguous ranges [floor(g T / G), floor((g + 1) T / G)); the owner of tile t is
// the largest g with floor(g T / G) <= t, i.e. floor(((t + 1) G - 1) / T).  32-bit arithmetic:
// the callers keep T * G < 2^32 (the panel engine checks it; 5e6 tiles x 512 workgroups covers
// every matrix that fits in HBM).
__host__ __device__ inline int64_t tile_owner(int64_t t, int64_t T, int64_t G)
{
    return (int64_t)(((uint32_t)(t + 1) * (uint32_t)G - 1u) / (uint32_t)T);
}

// first tile of rank g's range
__host__ __device__ inline int64_t range_lo(int64_t g, int64_t T, int64_t G)
{
    return (int64_t)((uint32_t)g * (uint32_t)T / (uint32_t)G);
}

__host__ __device__ inline int64_t strip_of(int64_t t)
{
    int64_t s = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((s + 1) * (s + 2) / 2 <= t) s++;
    while (s > 0 && s * (s + 1) / 2 > t) s--;
    return s;
}

// number of row-partial slots of strip s (0-based from J0) when ns strips are split over G ranks
// (at most T ranks take part, so every taking part rank has a non-empty range)
__host__ __device__ inline int64_t strip_slots(int64_t s, int64_t ns, int64_t G)
{
    const int64_t T = ns * (ns + 1) / 2;
    if (G > T) G = T;
    const int64_t first = s * (s + 1) / 2;
    return tile_owner(first + s, T, G) - tile_owner(first, T, G) + 1;
}

// ---------------------------------------------------------------------------------------
// Tiles of the panel kernel's matvec: MR x MC (256 rows x 16 columns), strip-major.  Strip I
// holds rows [MR I, MR I + MR); for the trailing block starting at s0 its column blocks are
// J = s0/MC .. (MR I + MR)/MC - 1.  Tall tiles make every column segment a lane group reads
// 2 KB long (DRAM row locality), and leave few strips for the column partials of a row.
// ---------------------------------------------------------------------------------------
constexpr int MR = 256;
constexpr int MC = 16;
constexpr int MPS = MR / MC;   // column blocks per strip row band

struct TallGrid {
    int64_t s0, Jlo, I0, nts, T;
    int64_t nblk;   // 64-row blocks from the one holding row s0 - 1 to the end (the panel's row-block owners)
};

__host__ __device__ inline int64_t tall_first(const TallGrid& tg, int64_t I)
{
    // tiles of strips I0 .. I-1: sum (MPS I' + MPS - Jlo)
    const int64_t k = I - tg.I0;
    return MPS * (k * (tg.I0 + I - 1) / 2) + k * (MPS - tg.Jlo);
}

__host__ __device__ inline TallGrid tall_grid(int64_t n, int64_t s0)
{
    TallGrid tg;
    tg.s0 = s0;
    tg.Jlo = s0 / MC;
    tg.I0 = s0 / MR;
    tg.nts = (n + MR - 1) / MR;
    tg.T = tall_first(tg, tg.nts);
    tg.nblk = (n + TB - 1) / TB - (s0 - 1) / TB;
    return tg;
}

// Weighted split of the T matvec tiles over Ge ranks: the first O ranks (the panel's row-block owners,
// which also run the row pass and start their matvec after it) take WNUM / WDEN of the others' share,
// so that they do not finish it last.  Rank g's range starts at floor(T vpos(g) / Lv) with vpos(g) =
// WNUM min(g, O) + WDEN max(g - O, 0) and Lv = vpos(Ge).  Applied only while T >= 2 Ge (with
// WNUM / WDEN >= 1/2 every range then holds at least one tile), else the split is uniform.  64-bit
// arithmetic (computed once per launch).  5/8 measured best at n = 8192 (186.6 ms; 3/4 187.3, 7/8 189.1,
// 1/1 193.0).
#ifndef SKH_OWNER_WNUM
#define SKH_OWNER_WNUM 5
#endif
#ifndef SKH_OWNER_WDEN
#define SKH_OWNER_WDEN 8
#endif
static_assert(2 * SKH_OWNER_WNUM >= SKH_OWNER_WDEN && SKH_OWNER_WNUM <= SKH_OWNER_WDEN,
              "owner weight must lie in [1/2, 1]: a smaller one can leave an owner without tiles");
struct TileSplit {
    uint64_t T, Ge, O, wn, wd, Lv;
};
__host__ __device__ inline TileSplit tile_split(const TallGrid& tg, int64_t Ge)
{
    TileSplit ts;
    ts.T = (uint64_t)tg.T;
    ts.Ge = (uint64_t)Ge;
    ts.O = (uint64_t)(tg.nblk < Ge ? tg.nblk : Ge);
    const bool w = tg.T >= 2 * Ge;
    ts.wn = w ? SKH_OWNER_WNUM : 1;
    ts.wd = w ? SKH_OWNER_WDEN : 1;
    ts.Lv = ts.wn * ts.O + ts.wd * (ts.Ge - ts.O);
    return ts;
}
__host__ __device__ inline int64_t split_lo(const TileSplit& ts, int64_t g)
{
    const uint64_t ug = (uint64_t)g;
    const uint64_t v = ug <= ts.O ? ts.wn * ug : ts.wn * ts.O + ts.wd * (ug - ts.O);
    return (int64_t)(ts.T * v / ts.Lv);
}
// the rank whose range holds tile t: the largest g with T vpos(g) <= (t + 1) Lv - 1
__host__ __device__ inline int64_t split_owner(const TileSplit& ts, int64_t t)
{
    const uint64_t X = ((uint64_t)t + 1) * ts.Lv - 1;
    const uint64_t g1 = X / (ts.T * ts.wn);
    if (g1 < ts.O) return (int64_t)g1;
    const uint64_t g2 = ts.O + (X - ts.T * ts.wn * ts.O) / (ts.T * ts.wd);
    return (int64_t)(g2 < ts.Ge - 1 ? g2 : ts.Ge - 1);
}

// strip holding tile t: tall_first is 8 k^2 + (16 I0 + 8 - Jlo) k in k = I - I0, inverted in
// floating point and corrected to the exact integer
__host__ __device__ inline int64_t tall_strip_of(const TallGrid& tg, int64_t t)
{
    const double B = (double)(MPS * tg.I0 + MPS / 2 - tg.Jlo);
    int64_t k = (int64_t)((sqrt(B * B + 4.0 * (MPS / 2) * (double)t) - B) / MPS);
    if (k < 0) k = 0;
    if (k > tg.nts - 1 - tg.I0) k = tg.nts - 1 - tg.I0;
    while (k > 0 && tall_first(tg, tg.I0 + k) > t) k--;
    while (tg.I0 + k + 1 < tg.nts && tall_first(tg, tg.I0 + k + 1) <= t) k++;
    return tg.I0 + k;
}

// row-partial slots of strip I (ranks whose tile range touches it) with G ranks
__host__ __device__ inline int64_t tall_slots(const TallGrid& tg, int64_t I, int64_t G)
{
    if (G > tg.T) G = tg.T;
    const TileSplit ts = tile_split(tg, G);
    const int64_t f = tall_first(tg, I);
    const int64_t cnt = MPS * I + MPS - tg.Jlo;
    return split_owner(ts, f + cnt - 1) - split_owner(ts, f) + 1;
}

// Matvec tile map of one panel launch: the tiles of the trailing block of the panel's first column
// (s0 = c0 + 1) split into Ge contiguous ranges, one per rank, fixed for the whole launch.  A later
// column j of the panel keeps the same tiles: its columns <= j are masked to zero when loaded, so a
// tile left of the trailing block adds nothing, and every rank still writes the row slot of each
// strip it touches.  Computed once per launch instead of once per column.
struct PanelMap {
    TallGrid tg;
    int64_t Ge;                      // ranks taking part in the matvec (G capped at the tile count)
    int64_t tlo, thi, I, J, slot;    // this rank's tile range, its first tile, that tile's row slot
    int64_t nslot_own;               // row slots of the strip holding the rank's own row block
};

__device__ inline PanelMap panel_map(int64_t n, int64_t c0, int g, int G)
{
    PanelMap pm;
    pm.tg = tall_grid(n, c0 + 1);
    pm.Ge = (G < pm.tg.T) ? G : pm.tg.T;
    pm.tlo = pm.thi = pm.I = pm.J = pm.slot = 0;
    if (g < pm.Ge) {
        const TileSplit ts = tile_split(pm.tg, pm.Ge);
        pm.tlo = split_lo(ts, g);
        pm.thi = split_lo(ts, g + 1);
        pm.I = tall_strip_of(pm.tg, pm.tlo);
        pm.J = pm.tg.Jlo + (pm.tlo - tall_first(pm.tg, pm.I));
        pm.slot = g - split_owner(ts, tall_first(pm.tg, pm.I));
    }
    const int64_t S = ((c0 / TB + g) * TB) / MR;
    pm.nslot_own = (S >= pm.tg.I0 && S < pm.tg.nts) ? tall_slots(pm.tg, S, G) : 0;
    return pm;
}

// ---------------------------------------------------------------------------------------
// Row pass: finalises w of column j-1 (doC) and/or applies the panel column update to
// column j and produces its norm and W^T x / V^T x partials (doA).
//   i = j - c0 = number of previous panel columns that update column j.
// Rows are processed in absolute 64-row blocks starting at block j/64; block b is owned by
// rank (b - c0/64) mod G in both uses, so the row data a rank writes in C(j-1) is the data it
// reads in A(j).  Row j of w_{j-1} (needed by everyone as W[j, i-1]) is computed redundantly by
// every rank with one fixed summation order, and the owner stores that same value.
// Hand-over: the block's norm partial (slot b), its W^T x / V^T x partials (row b) and its rows of
// x_j go to copy j mod 4 of Pnorm / PS / X, and its rows of the new W and V columns to W and V2,
// all polled by their consumers.  At its start the pass empties the block's slots of column j + 2
// (copy (j + 2) mod 4 = (j - 2) mod 4): every phase B that read column j - 2 had finished before
// the owners of all blocks stored their norm partials of column j - 1, which this rank's phase B
// of column j - 1 has seen; the next pass's drain completes those stores before this rank stores
// anything of column j + 1, and readers of column j + 2 poll only after seeing column j + 1.
// ---------------------------------------------------------------------------------------
// MROWS: rows of the matvec flush buffer: 16 (one per wave and column group, the LDS transposition
// of the row sums; the 32-wide instance, where the LDS fits) or 4 (one per wave, after a
// cross-lane reduction over the column groups)
template <int MROWS>
struct alignas(16) RowShared {
    double wrow[NBMAX];
    double vrow[NBMAX];
    double t[2 * NBMAX];
    union {
        struct {
            double red[4][TB];    // row pass reductions
            double red2[4][TB];
        };
        double mred[MROWS][MR + 4];  // matvec row sums of one strip (other phase)
    };
    double vstrip[MR];    // v on the rows of the matvec strip being processed
    double wspec;
    ColScal cur;     // scalars of the column finalised by the last phase B
};

// Row data a workgroup keeps for the whole panel: its own 64-row block (absolute block
// c0/64 + g) of V and W in LDS (wave w holds k in [w*KQ, (w+1)*KQ)), and in registers the last x
// it computed (column j-1, needed as the unscaled reflector of the next C part) and the
// prefetched column j+1 (the next A part's input).  Every new V/W column entry of the block is
// computed by this workgroup, so the block's rows never have to be reloaded from memory.
template <int KQ>
struct RowRegs {
    double* v;   // LDS [4 * KQ][64]: (wave * KQ + u) * 64 + lane
    double* w;
    double xprev, xnext;
};

// KQ = k-group per wave (panel width <= 4*KQ).  The rows of V and W used by a lane are held in
// 2*KQ registers, shared by the C part (correction of w_{j-1}, k < i-1) and the A part (update of
// column j, k < i).  Every global load a chunk needs (row and column partials of the matvec;
// for blocks other than the workgroup's own, also its rows of V and W and the old entries of
// columns j-1 and j) and the shared inputs (t, row j of V and W, the special row's partials) are
// issued together and polled together, so a row pass costs one memory round trip once its inputs
// are there.  Returns false when the launch is abandoned (a poll ran out, or another workgroup
// flagged an error).
template <int KQ, bool MULTI, int MROWS>
__device__ bool row_pass(double* __restrict__ A, int64_t lda, int64_t n, int64_t c0, int64_t j, int i,
                         bool doC, bool doA, const Work& w, int g, int G, RowShared<MROWS>& sh, RowRegs<KQ>& rr,
                         const PanelMap& pm, int* err, unsigned long long* clk)
{
    clk_begin(clk);
    const int tid = threadIdx.x, lane = tid & 63, wave = wave_id();
    const int64_t jm = j - 1;         // column finalised by the C part
    const int im = i - 1;             // its local index
    const ColScal cs = sh.cur;        // scalars of column jm (valid when doC)
    double* __restrict__ Wm = w.W;
    double* __restrict__ V2 = w.V2;
    const int64_t ldw = w.ldw;
    const int64_t b0 = j / TB;        // block of row j
    const int64_t nt = cdiv(n, TB);
    const int64_t bown = c0 / TB + g; // this workgroup's own block
    const int k0 = wave * KQ, k1 = min(i, k0 + KQ);
    const int kc = doC ? min(k1, im) : k0;   // C correction uses k in [k0, kc)
    constexpr int PB = 8;             // column partials per lane per batch (4 waves: 32 strips)
    const TallGrid& tg = pm.tg;       // tiles of the matvec that produced the partials
    bool shared_done = false;

    for (int64_t b = bown; b < nt; b += (MULTI ? G : nt)) {
        if (b < b0) continue;         // block entirely above row j (wave-uniform)
        const bool first = !shared_done;
        shared_done = true;
        const bool own = !MULTI || (b == bown);
        const int64_t r = b * TB + lane;
        const bool rin = r < n;
        const bool rowC = doC && rin && r >= j;
        // ---- 0. empty the block's slots of column j + 2 ----
        if (doA) {
            const double e = empty_slot();
            if (wave == 0) {
                if (rin) st_co(x_of(w, j + 2) + r, e);
                if (lane == 0) st_co(pnorm_of(w, j + 2) + b, e);
            }
            if (lane < KQ) {
                double* ps2 = ps_of(w, j + 2) + b * (2 * NBMAX) + k0 + lane;
                st_co(ps2, e);
                st_co(ps2 + NBMAX, e);
            }
        }
        // ---- 1. every global load of this step, into registers ----
        double tv1 = 0.0, tv2 = 0.0, wraw = 0.0, vraw = 0.0, sprow = 0.0, sp[4] = {0.0, 0.0, 0.0, 0.0};
        double vreg[KQ], wreg[KQ], xo, xold;
        if (own) {
#pragma unroll
            for (int u = 0; u < KQ; u++) {
                vreg[u] = rr.v[(wave * KQ + u) * 64 + lane];
                wreg[u] = rr.w[(wave * KQ + u) * 64 + lane];
            }
            xo = rr.xprev;
            xold = rr.xnext;
        } else {
#pragma unroll
            for (int u = 0; u < KQ; u++) {
                const int k = k0 + u;
                const bool ok = rin && r >= j && k < k1 && !(doC && k == im);
                vreg[u] = ok ? ld_co(A + r + (c0 + k) * lda) : 0.0;
                wreg[u] = ok ? ld_co(Wm + r + k * ldw) : 0.0;
            }
            xo = rowC ? ld_co(x_of(w, jm) + r) : 0.0;         // x of column j-1 (unscaled)
            xold = ((doA || doC) && rin && r > j) ? A[r + j * lda] : 0.0;
        }
        double pp = 0.0;
        if (doC) {
            // t and the special row's column partials (phase B of column j-1), row j of W and V
            // (earlier row passes) and this row's matvec partials, polled until all have arrived.
            // Every load of a round is issued before any is looked at (atomic loads keep their
            // order, and a check between them would make the compiler wait for each); lanes that
            // need a value load a valid slot instead and ignore it.  32-bit element offsets:
            // (slots + strips) * ldp stays far below 2^31.
            const int S = (int)((b * TB) / MR);    // matvec strip holding this chunk
            const int nsl = (int)(own ? pm.nslot_own : tall_slots(tg, S, G));
            const int nts = (int)tg.nts, ldp = (int)w.ldp;
            const int rr32 = (int)(rowC ? r : j);
            const int S0 = (int)(j / MR), j32 = (int)j;
            const double* prowm = prow_of(w, jm);
            const double* pcolm = pcol_of(w, jm);
            const double* tvm = tv_of(w, jm);
            const bool ga = first && im > 0;      // wave-uniform
            const int kt = (tid < im) ? tid : 0;
            const bool spw = first && wave == 0;  // wave-uniform
            bool alive = true;
            for (unsigned spins = 0;; spins++) {
                if (ga) {
                    tv1 = ld_co(tvm + kt);
                    tv2 = ld_co(tvm + NBMAX + kt);
                    wraw = ld_co(Wm + j + kt * ldw);
                    vraw = ld_co(V2 + j + kt * ldw);
                }
                if (spw) {
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int I = S0 + lane + 64 * u;
                        sp[u] = ld_co(pcolm + ((I < nts ? I : S0) * ldp + j32));
                    }
                }
                double sb[8], buf[PB];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int sl = wave + 4 * u;
                    sb[u] = ld_co(prowm + ((sl < nsl ? sl : 0) * ldp + rr32));
                }
#pragma unroll
                for (int u = 0; u < PB; u++) {
                    const int I = S + wave + 4 * u;
                    buf[u] = ld_co(pcolm + ((I < nts ? I : S) * ldp + rr32));
                }
                unsigned miss = 0;
                if (ga && tid < im) {
                    if (!(filled(tv1) && filled(tv2))) miss |= 1;
                    if (!(filled(wraw) && filled(vraw))) miss |= 2;
                }
                if (spw) {
#pragma unroll
                    for (int u = 0; u < 4; u++)
                        if (S0 + lane + 64 * u < nts && !filled(sp[u])) miss |= 4;
                }
                pp = 0.0;
                if (rowC) {
                    // first batches of row slots and strips, in a fixed order ahead of the extra ones
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        if (wave + 4 * u >= nsl) sb[u] = 0.0;
                        if (!filled(sb[u])) miss |= 8;
                    }
                    double p4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int u = 0; u < PB; u++) {
                        if (S + wave + 4 * u >= nts) buf[u] = 0.0;
                        if (!filled(buf[u])) miss |= 16;
                        p4[u & 3] += buf[u];
                    }
                    pp = (((sb[0] + sb[1]) + (sb[2] + sb[3])) + ((sb[4] + sb[5]) + (sb[6] + sb[7]))) +
                         ((p4[0] + p4[1]) + (p4[2] + p4[3]));
                }
                // rare extra batches: strips beyond the first 256 for the special row (n > 65536),
                // more than 32 row slots, more than 64 strips below the chunk (n > 16384)
                sprow = 0.0;
                if (spw) {
                    for (int I = S0 + lane + 256; I < nts; I += 64) {
                        const double x = ld_co(pcolm + (I * ldp + j32));
                        if (!filled(x)) miss |= 4;
                        sprow += x;
                    }
                }
                if (rowC) {
                    for (int s0 = wave + 32; s0 < nsl; s0 += 32) {
                        double xb[8];
#pragma unroll
                        for (int u = 0; u < 8; u++) xb[u] = ld_co(prowm + ((s0 + 4 * u < nsl ? s0 + 4 * u : 0) * ldp + rr32));
#pragma unroll
                        for (int u = 0; u < 8; u++) {
                            if (s0 + 4 * u >= nsl) xb[u] = 0.0;
                            if (!filled(xb[u])) miss |= 8;
                        }
                        pp += ((xb[0] + xb[1]) + (xb[2] + xb[3])) + ((xb[4] + xb[5]) + (xb[6] + xb[7]));
                    }
                    for (int I0 = S + wave + 4 * PB; I0 < nts; I0 += 4 * PB) {
                        double xb[PB];
#pragma unroll
                        for (int u = 0; u < PB; u++) xb[u] = ld_co(pcolm + ((I0 + 4 * u < nts ? I0 + 4 * u : S) * ldp + rr32));
                        double p4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                        for (int u = 0; u < PB; u++) {
                            if (I0 + 4 * u >= nts) xb[u] = 0.0;
                            if (!filled(xb[u])) miss |= 16;
                            p4[u & 3] += xb[u];
                        }
                        pp += (p4[0] + p4[1]) + (p4[2] + p4[3]);
                    }
                }
                if (__all(miss == 0)) break;
                if (!poll_more(spins, err, poll_code(POLL_ROW, j))) {
                    poll_report(err, miss, r);
                    alive = false;
                    break;
                }
            }
            // this wave's earlier stores (the empty markers above and of the last pass, phase B's
            // partials) have completed before anything of column j is stored
            drain();
            if (__syncthreads_or(!alive)) return false;
            if (spw) {
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (S0 + lane + 64 * u >= nts) sp[u] = 0.0;
            }
        }
        // ---- 2. shared inputs into LDS (first chunk), the special row of w_{j-1} ----
        if (first) {
            if (doC && tid < im) {
                // t of column j-1 = unit * (row j of W, V) + scale * (sums over rows > j of W^T x, V^T x)
                sh.t[tid] = cs.unit * wraw + cs.scale * tv1;
                sh.t[NBMAX + tid] = cs.unit * vraw + cs.scale * tv2;
                sh.wrow[tid] = wraw;
                sh.vrow[tid] = vraw;
            }
            __syncthreads();
            clk_inner(clk, 4);
            if (doC && wave == 0) {
                // (A v)_j = -sum_{r > j} A[r, j] v_r: column partials of the unscaled x, times scale
                double acc = sprow;
#pragma unroll
                for (int u = 0; u < 4; u++) acc += sp[u];
                acc *= cs.scale;
                for (int k = lane; k < im; k += 64) acc += sh.vrow[k] * sh.t[k] - sh.wrow[k] * sh.t[NBMAX + k];
                acc = wave_sum(acc);
                if (lane == 0) {
                    const double ws = cs.tau * acc;
                    sh.wspec = ws;
                    sh.wrow[im] = ws;
                    sh.vrow[im] = cs.unit;
                }
            }
        }
        // ---- C part: w_{j-1} on this chunk; with it, the A part's terms of the reflectors before the
        // newest one (k < ka), which need nothing from the C part: one barrier publishes both sums ----
        // Every LDS operand of the wave's k-group is read first and masked at use (a read under a
        // per-term condition would be waited for on its own); entries past the live ones may hold
        // anything, and the select drops them.
        const bool live = doA && rin && r > j;
        const int ka = doC ? im : i;
        {
            // the 32-wide instance reads all its operands at once; wider ones in groups of 4 k (all of
            // them would not fit in registers)
            constexpr int KG = KQ <= 8 ? KQ : 4;
            double c4[4] = {0.0, 0.0, 0.0, 0.0}, p4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int u0 = 0; u0 < KQ; u0 += KG) {
                double ta[KG], tb[KG], wa[KG], va[KG];
#pragma unroll
                for (int q = 0; q < KG; q++) {
                    ta[q] = sh.t[k0 + u0 + q];
                    tb[q] = sh.t[NBMAX + k0 + u0 + q];
                    wa[q] = sh.wrow[k0 + u0 + q];
                    va[q] = sh.vrow[k0 + u0 + q];
                }
#pragma unroll
                for (int q = 0; q < KG; q++) {
                    const int u = u0 + q;
                    if (doC) {
                        const double term = vreg[u] * ta[q] - wreg[u] * tb[q];
                        c4[u & 3] += (rowC && k0 + u < kc) ? term : 0.0;
                    }
                    if (doA) {
                        const double term = vreg[u] * wa[q] - wreg[u] * va[q];
                        p4[u & 3] += (live && k0 + u < ka) ? term : 0.0;
                    }
                }
            }
            // (A v)_r = scale * (A x)_r + corr * A[r, j] (x_j = alpha where v_j = unit; xold = A[r, j], 0
            // for r <= j); pp is this wave's part of the partials, the corr term is wave 0's
            if (doC)
                sh.red[wave][lane] = rowC ? (cs.scale * pp + (wave == 0 ? cs.corr * xold : 0.0)) + ((c4[0] + c4[1]) + (c4[2] + c4[3]))
                                          : 0.0;
            if (doA) sh.red2[wave][lane] = (p4[0] + p4[1]) + (p4[2] + p4[3]);
        }
        __syncthreads();
        double wr = 0.0, vr = 0.0;
        if (rowC) {
            const double sacc = sh.red[0][lane] + sh.red[1][lane] + sh.red[2][lane] + sh.red[3][lane];
            wr = (r == j) ? sh.wspec : cs.tau * sacc;
            vr = (r == j) ? cs.unit : cs.scale * xo;
            if (wave == 0) {
                st_co(Wm + r + im * ldw, wr);
                st_co(V2 + r + im * ldw, vr);
                st_co(A + r + jm * lda, vr);
            }
        }
        if (doC) {
            // the block's rows of the new V/W column (zero above row j)
#pragma unroll
            for (int u = 0; u < KQ; u++) {
                if (k0 + u == im) {
                    vreg[u] = vr;
                    wreg[u] = wr;
                }
            }
        }
        clk_inner(clk, 5);
        double x = 0.0;
        if (doA) {
            // ---- A part: column j updated with the panel's reflectors 0..i-1: the sums over k < ka from
            // before the barrier, plus the newest reflector's term (V[r,im] W[j,im] - W[r,im] V[j,im]) ----
            if (live) {
                x = xold + (sh.red2[0][lane] + sh.red2[1][lane] + sh.red2[2][lane] + sh.red2[3][lane]);
                if (doC) x += vr * sh.wspec - wr * cs.unit;
            }
            if (live && wave == 0) st_co(x_of(w, j) + r, x);   // column j keeps its old entries until v
            const bool tail = rin && r >= j + 2;
            if (wave == 0) {
                const double ss = wave_sum(tail ? x * x : 0.0);
                if (lane == 0) st_co(pnorm_of(w, j) + b, ss);
            }
            // S partials of column j: sum over the chunk's rows r >= j+2 of W[r,k] x_r, V[r,k] x_r
            const double xt = tail ? x : 0.0;
            double* psb = ps_of(w, j) + b * (2 * NBMAX);
            if (KQ <= 16) {
                // values 0..15: W[r, k0+u] x_r, 16..31: V[r, k0+u] x_r (zero past the wave's group)
                double v32[32];
#pragma unroll
                for (int u = 0; u < 16; u++) {
                    v32[u] = (u < KQ && k0 + u < k1) ? wreg[u < KQ ? u : 0] * xt : 0.0;
                    v32[16 + u] = (u < KQ && k0 + u < k1) ? vreg[u < KQ ? u : 0] * xt : 0.0;
                }
                const double tot = wave_sum32(v32);
                const int idx = (lane >> 1) & 31;
                const int u = idx & 15;
                if ((lane & 1) == 0 && k0 + u < k1) st_co(psb + (idx < 16 ? 0 : NBMAX) + k0 + u, tot);
            } else {
#pragma unroll
                for (int u = 0; u < KQ; u++) {
                    if (k0 + u < k1) {   // wave-uniform
                        const double s1 = wave_sum(wreg[u] * xt);
                        const double s2 = wave_sum(vreg[u] * xt);
                        if (lane == 0) {
                            st_co(psb + k0 + u, s1);
                            st_co(psb + NBMAX + k0 + u, s2);
                        }
                    }
                }
            }
        }
        clk_inner(clk, 6);
        if (own) {
            // keep the block's rows for the next column of the panel; prefetch its input column
            if (doC) {   // only column im is new
#pragma unroll
                for (int u = 0; u < KQ; u++) {
                    if (k0 + u == im) {
                        rr.v[(wave * KQ + u) * 64 + lane] = vreg[u];
                        rr.w[(wave * KQ + u) * 64 + lane] = wreg[u];
                    }
                }
            }
            rr.xprev = x;
            rr.xnext = (doA && rin && r > j + 1 && j + 1 < n) ? A[r + (j + 1) * lda] : 0.0;
        }
        __syncthreads();
        clk_inner(clk, 7);
    }
    return true;
}

// ---------------------------------------------------------------------------------------
// Phase B of column j, once the row pass's outputs of column j it needs are in:
//   (1) after the matvec (2): the sums over rows >= j+2 of W^T x and V^T x of column j from the
//       per-block partials (unscaled), element e of t (W part e < NBMAX, V part NBMAX + k) by wave
//       e mod 4 of rank (e / 4) mod G for the whole launch, live while k < i (the consumer scales
//       them and adds unit * row j+1);
//   (2) the skew matvec partials of column j over the strictly lower triangle of
//       A[j+1:n, j+1:n] on MR x MC tiles (strip-major), split into G contiguous tile ranges, one
//       per rank, on the unscaled x (x_{j+1} = alpha): the consumer forms
//       A v = scale * A x + corr * A[:, j+1], so the matvec waits for no reduction.  Per tile:
//       wave w -> columns 4w..4w+3 (lane >> 4), lane & 15 = p -> row pairs
//       {32 h + 2p, 32 h + 2p + 1}, h = 0..7 (16 entries per lane, read as 16-byte pairs); the
//       next tile's loads are in flight while the current one is reduced.
//         Pcol[I][c]    = -sum_{r in strip I} A[r,c] x_r                     (one entry per tile)
//         Prow[slot][r] =  sum over the rank's tiles of strip I of A[r,c] x_c,
//                          slot = rank - (owner of the strip's first tile)
//   (3) after the matvec, the reflector scalars beta, tau, 1/(alpha - beta), corr from the block
//       norm partials (the next row pass's inputs), then the slot emptying.
// The first tile is loaded before the poll (tile data is not written while any rank can still read
// it: a column c > j is stored as a factor in row pass c + 1, after every rank's matvec of column
// j has been consumed); the poll of the first tile's x entries is one memory round trip once they
// are stored.
// Only ranks with a part in column j take it (tiles, a live t element, or rows of w_j to finish
// in the next row pass): every value such a rank reads is consumed by a row pass, which is what
// lets producers refill the slots with the empty marker safely.
// ---------------------------------------------------------------------------------------
// Raw tile entries of one lane: rows I*MR + 32 h + 2p (+1), column c, as 16-byte pairs.  The loads
// are unconditional and always in bounds (a row or column past the matrix reads row 0 or column
// n-1; lda is even and >= n, so the pair at an even row < n stays inside its column), and nothing
// is masked here: a mask right after a load makes the compiler wait for it, which would serialise
// the prefetch with the reduction of the current tile.  tall_ok masks at use.
__device__ __forceinline__ void tall_load(const double* __restrict__ A, int64_t lda, int64_t n, int64_t I, int64_t c,
                                          int p, double a[16])
{
    const double* col = A + ((c < n) ? c : n - 1) * lda + I * MR;
    const int rn = (int)(n - I * MR), rb = 2 * p;   // rows relative to the strip
#pragma unroll
    for (int h = 0; h < 8; h++) {
        const int r = rb + 32 * h;
        const double2 t = *reinterpret_cast<const double2*>(col + ((r < rn) ? r : -(int)(I * MR)));
        a[2 * h] = t.x;
        a[2 * h + 1] = t.y;
    }
}

// entry k of a lane's tile (row I*MR + 32 (k>>1) + 2p + (k&1), column c) lies in the strictly lower
// triangle of the trailing block of column j: lo = c - I*MR and hi = n - I*MR bound the local row
__device__ __forceinline__ bool tall_ok(int k, int p, int lo, int hi)
{
    const int rl = 32 * (k >> 1) + 2 * p + (k & 1);
    return rl > lo && rl < hi;
}

// unscaled x_c of column j (x_j in xj) on the rows of a strip: rows past j + 1 (row j + 1 meets
// only columns <= j, which are masked)
__device__ __forceinline__ double xraw(const double* __restrict__ xj, int64_t n, int64_t j, int64_t c)
{
    return (c > j + 1 && c < n) ? ld_co(xj + c) : 0.0;
}

// unscaled x_c of column j for a tile column: alpha at c = j + 1, 0 outside (j, n)
__device__ __forceinline__ double xcol(const double* __restrict__ xj, int64_t n, int64_t j, int64_t c)
{
    return (c > j && c < n) ? ld_co(xj + c) : 0.0;
}

// does rank g own a t element that is live at local column i?
__device__ __forceinline__ bool t_live(int g, int G, int i)
{
    for (int q = 4 * g; q < 2 * NBMAX; q += 4 * G)
        if ((q % NBMAX) < i) return true;
    return false;
}

// RES: the rank's first matvec tile (tile tlo, the same for every column of the launch) is kept in
// LDS for the whole launch (rt: WG x 8 double2, thread tid's 16 entries at [h * WG + tid], loaded
// once by tile_keep), so each column streams one tile fewer per rank from memory; the loads of the
// rank's second tile then go out before the poll, as the first tile's did.
// RES: thread tid's 16 entries of the rank's first matvec tile into rt (read back only by the same
// thread, so no barrier); the tile data is the trailing block at the launch's start, which the matvec
// of every column of the launch reads (columns <= j are masked at use)
__device__ __forceinline__ void tile_keep(const double* __restrict__ A, int64_t lda, int64_t n, const PanelMap& pm,
                                          int g, double2* rt)
{
    if (g >= pm.Ge) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = wave_id();
    double a[16];
    tall_load(A, lda, n, pm.I, pm.J * MC + 4 * wave + (lane >> 4), lane & 15, a);
#pragma unroll
    for (int h = 0; h < 8; h++) rt[h * WG + tid] = make_double2(a[2 * h], a[2 * h + 1]);
}

template <int MROWS, bool RES>
__device__ bool phase_b(const double* __restrict__ A, int64_t lda, int64_t n, int64_t c0, int64_t j, int i,
                        const Work& w, RowShared<MROWS>& sh, double* tau_out, double* eta_out, int g, int G,
                        const PanelMap& pm, bool rown, const double2* rt, int* err, unsigned long long* clk)
{
    const bool mv = g < pm.Ge;
    if (!(mv || rown || t_live(g, G, i))) return true;
    clk_begin(clk);
    const int tid = threadIdx.x, lane = tid & 63, wave = wave_id();
    const int p = lane & 15, cg = lane >> 4;
    const int64_t b0 = j / TB;
    const int64_t nt = cdiv(n, TB);
    const int64_t nq = nt - b0;
    const TallGrid& tg = pm.tg;
    const int64_t tlo = pm.tlo, thi = pm.thi;
    int64_t I = pm.I, J = pm.J, cur_slot = pm.slot, c = 0;
    double a[16], an[16];
    if (mv) {
        c = J * MC + 4 * wave + cg;
        if constexpr (RES) {
            if (tlo + 1 < thi) {
                const bool lis = (J == MPS * I + MPS - 1);
                // wave-uniform, and opaque per column: keeps the compiler from hoisting this address math
                // (lane-dependent, 64-bit) out of the column loop into registers it would have to spill
                int I2 = __builtin_amdgcn_readfirstlane((int)(lis ? I + 1 : I));
                int J2 = __builtin_amdgcn_readfirstlane((int)(lis ? tg.Jlo : J + 1));
                asm volatile("" : "+s"(I2), "+s"(J2));
                tall_load(A, lda, n, I2, J2 * MC + 4 * wave + cg, p, an);
            }
        } else {
            tall_load(A, lda, n, I, c, p, a);
        }
    }
    clk_inner(clk, 11);
    // ---- poll what the row pass of column j produced: this wave's t partials, the first tile's x ----
    const double* pnj = pnorm_of(w, j);
    const double* psj = ps_of(w, j);
    const double* xj = x_of(w, j);
    const int e0 = 4 * g + wave;                                 // this wave's first t element
    const bool t0 = e0 < 2 * NBMAX && (e0 % NBMAX) < i;          // wave-uniform
    const int64_t rs = I * MR + tid;                             // this thread's row of the first strip
    double ps0 = 0.0, ps1 = 0.0, xc = 0.0, xs = 0.0;
    // every load of a round is issued before any is looked at (the loads are atomic, so the compiler
    // keeps their order and would otherwise wait for each before issuing the next); out-of-range
    // lanes load a valid slot and ignore it
    const bool in0 = lane < nq, in1 = lane + 64 < nq;
    const int64_t q0 = b0 + (in0 ? lane : 0), q1 = b0 + (in1 ? lane + 64 : 0);
    const bool xcin = mv && c > j && c < n, xsin = mv && rs > j + 1 && rs < n;
    const double* pxc = xj + (xcin ? c : j + 1);
    const double* pxs = xj + (xsin ? rs : j + 1);
    const int64_t e0c = t0 ? e0 : 0;
    // t is formed after the matvec (its partials not polled here), so the ranks that own t elements
    // (also row-block owners) do not wait for every block's row pass before their matvec (n = 8192:
    // 186.8 -> 185.7 ms, nb = 64: 213.6 -> 207.2 ms); SKH_T_AFTER=0 builds the form before it
#ifndef SKH_T_AFTER
#define SKH_T_AFTER 1
#endif
    constexpr bool TA = SKH_T_AFTER;
    const bool tp = t0 && !TA;
    bool alive = true;
    for (unsigned spins = 0;; spins++) {
        if (tp) {   // wave-uniform branches: no per-lane exec mask around the loads
            ps0 = ld_co(psj + q0 * (2 * NBMAX) + e0c);
            ps1 = ld_co(psj + q1 * (2 * NBMAX) + e0c);
        }
        if (mv) {
            xc = ld_co(pxc);
            xs = ld_co(pxs);
        }
        unsigned miss = 0;
        if (tp && ((in0 && !filled(ps0)) || (in1 && !filled(ps1)))) miss |= 128;
        if (xcin && !filled(xc)) miss |= 256;
        if (xsin && !filled(xs)) miss |= 512;
        if (__all(miss == 0)) break;
        if (!poll_more(spins, err, poll_code(POLL_B, j))) {
            poll_report(err, miss, (miss & 256) ? c : rs);
            alive = false;
            break;
        }
    }
    if (!in0) ps0 = 0.0;
    if (!in1) ps1 = 0.0;
    if (!xcin) xc = 0.0;
    if (!xsin) xs = 0.0;
    // the empty markers this rank stored in the last phase B have completed (they must before any
    // value of column j + 1 leaves this rank)
    drain();
    clk_inner(clk, 12);
    if (mv) sh.vstrip[tid] = xs;
    // ---- (2) t of the live elements of this wave (unscaled: the consumer applies scale) ----
    auto form_t = [&]() {
    for (int e = e0; e < 2 * NBMAX && alive; e += 4 * G) {
        if ((e % NBMAX) >= i) continue;   // wave-uniform
        double acc;
        if (e == e0 && !TA) {
            acc = ps0 + ps1;
        } else {
            acc = 0.0;
            for (int64_t q = lane; q < 128 && q < nq; q += 64) {
                const double* pq = psj + (b0 + q) * (2 * NBMAX) + e;
                acc += settle(ld_co(pq), pq, err, poll_code(POLL_SETTLE, j));
            }
        }
        for (int64_t q = lane + 128; q < nq; q += 64) {
            const double* pq = psj + (b0 + q) * (2 * NBMAX) + e;
            acc += settle(ld_co(pq), pq, err, poll_code(POLL_SETTLE, j));
        }
        acc = wave_sum(acc);
        if (lane == 0) st_co(tv_of(w, j) + e, acc);
    }
    };
    if (!TA) form_t();
    clk_inner(clk, 9);
    if (__syncthreads_or(!alive)) return false;   // x of the first strip is in LDS
    // ---- (3) matvec ----
    if (mv) {
        if constexpr (RES) {
#pragma unroll
            for (int h = 0; h < 8; h++) {
                const double2 t2 = rt[h * WG + tid];
                a[2 * h] = t2.x;
                a[2 * h + 1] = t2.y;
            }
        }
        double racc[16];
#pragma unroll
        for (int k = 0; k < 16; k++) racc[k] = 0.0;
        double vc = xc;   // unscaled x of the lane's column (alpha at j + 1, 0 outside (j, n))
        for (int64_t t = tlo; t < thi; t++) {
            const bool more = t + 1 < thi;
            const bool last_in_strip = (J == MPS * I + MPS - 1);
            const int64_t In = last_in_strip ? I + 1 : I;
            const int64_t Jn = last_in_strip ? tg.Jlo : J + 1;
            const int64_t cn = Jn * MC + 4 * wave + cg;
            double xcn = 0.0, xsn = 0.0;
            if (more) {
                xcn = xcol(xj, n, j, cn);
                if (!RES || t != tlo) tall_load(A, lda, n, In, cn, p, an);   // (RES: tile tlo + 1 is loaded)
                if (In != I) xsn = xraw(xj, n, j, In * MR + tid);
            }
            double cp = 0.0;
            {
                // v of the lane's 16 strip rows, all read before use (eight 16-byte LDS reads); the
                // column sum in four chains (a single chain of 16 dependent FMAs is a latency chain)
                double vsr[16];
#pragma unroll
                for (int h = 0; h < 8; h++) {
                    const double2 t2 = *reinterpret_cast<const double2*>(&sh.vstrip[32 * h + 2 * p]);
                    vsr[2 * h] = t2.x;
                    vsr[2 * h + 1] = t2.y;
                }
                // columns outside (j, n) have vc = 0, but their raw entries are masked all the same
                // (0 * Inf or 0 * NaN of an unused entry must not reach the sums)
                const int lo = (c > j && c < n) ? (int)(c - I * MR) : MR, hi = (int)min<int64_t>(n - I * MR, MR + 1);
                double cq[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    const double av = tall_ok(k, p, lo, hi) ? a[k] : 0.0;
                    racc[k] += av * vc;
                    cq[k & 3] += av * vsr[k];
                }
                cp = (cq[0] + cq[1]) + (cq[2] + cq[3]);
            }
            cp = row16_sum(cp);
            if (p == 0 && c > j && c < n) st_co(pcol_of(w, j) + I * w.ldp + c, -cp);
            if (!more || In != I) {
                // flush this rank's row sums of strip I into one slot row
                if constexpr (MROWS == 16) {
                    // each (wave, column group) writes its 16 rows (8 row pairs) to its own LDS row; thread
                    // t then sums the 16 values of strip row t in a fixed order (an LDS transposition
                    // instead of 16 cross-lane reduction chains)
                    double* mr = sh.mred[wave * 4 + cg];
#pragma unroll
                    for (int h = 0; h < 8; h++) {
                        *reinterpret_cast<double2*>(mr + 32 * h + 2 * p) = make_double2(racc[2 * h], racc[2 * h + 1]);
                        racc[2 * h] = 0.0;
                        racc[2 * h + 1] = 0.0;
                    }
                } else {
                    // each row's sum over the wave's 4 column groups (lanes p, p + 16, p + 32, p + 48) by
                    // halving exchanges: 8 swaps between the wave's halves, then 4 between rows of 16 lanes;
                    // lane (cg, p) ends with the totals of values 4 cg .. 4 cg + 3, i.e. rows
                    // 64 cg + 32 h + 2 p + e (h, e in {0, 1}), one LDS row per wave
                    double v[16];
#pragma unroll
                    for (int k = 0; k < 16; k++) {
                        v[k] = racc[k];
                        racc[k] = 0.0;
                    }
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        double a0, b0v;
                        swap32(v[q], v[q + 8], a0, b0v);
                        v[q] = a0 + b0v;
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        double a0, b0v;
                        swap16(v[q], v[q + 4], a0, b0v);
                        v[q] = a0 + b0v;
                    }
                    *reinterpret_cast<double2*>(&sh.mred[wave][64 * cg + 2 * p]) = make_double2(v[0], v[1]);
                    *reinterpret_cast<double2*>(&sh.mred[wave][64 * cg + 32 + 2 * p]) = make_double2(v[2], v[3]);
                }
                __syncthreads();   // also: every wave is done reading vstrip of strip I
                {
                    const int64_t r = I * MR + tid;
                    double tot;
                    if constexpr (MROWS == 16) {
                        double q4[4];
#pragma unroll
                        for (int q = 0; q < 4; q++)
                            q4[q] = (sh.mred[4 * q][tid] + sh.mred[4 * q + 1][tid]) + (sh.mred[4 * q + 2][tid] + sh.mred[4 * q + 3][tid]);
                        tot = (q4[0] + q4[1]) + (q4[2] + q4[3]);
                    } else {
                        tot = (sh.mred[0][tid] + sh.mred[1][tid]) + (sh.mred[2][tid] + sh.mred[3][tid]);
                    }
                    if (r < n) st_co(prow_of(w, j) + cur_slot * w.ldp + r, tot);
                    if (more) sh.vstrip[tid] = settle(xsn, xj + In * MR + tid, err, poll_code(POLL_SETTLE, j));
                }
                __syncthreads();
                cur_slot = 0;   // this rank owns the first tile of every later strip it touches
            }
            if (more) {
#pragma unroll
                for (int k = 0; k < 16; k++) a[k] = an[k];
                vc = settle(xcn, xj + cn, err, poll_code(POLL_SETTLE, j));
            }
            I = In;
            J = Jn;
            c = cn;
        }
        clk_inner(clk, 10);
    }
    if (TA) form_t();
    // ---- (1) the reflector scalars of column j, after the matvec, which needs none of them: from the
    // block norm partials, formed by every rank in the same fixed order (wave 0, strided lanes, a
    // butterfly), so all ranks hold bit-identical values.  Having seen every block's norm partial of
    // column j also tells this rank that every row pass of column j has finished (the slot emptying
    // below relies on it) ----
    if (wave == 0) {
        double pn0 = 0.0, pn1 = 0.0, araw = 0.0;
        bool ok = alive;
        for (unsigned spins = 0; ok; spins++) {
            pn0 = ld_co(pnj + q0);
            pn1 = ld_co(pnj + q1);
            araw = ld_co(xj + (j + 1));
            unsigned miss = 0;
            if ((in0 && !filled(pn0)) || (in1 && !filled(pn1))) miss |= 32;
            if (!filled(araw)) miss |= 64;
            if (__all(miss == 0)) break;
            if (!poll_more(spins, err, poll_code(POLL_B, j))) {
                poll_report(err, miss, j + 1);
                ok = false;
            }
        }
        if (ok) {
            double acc = (in0 ? pn0 : 0.0) + (in1 ? pn1 : 0.0);
            for (int64_t q = lane + 128; q < nq; q += 64) {
                const double x = ld_co(pnj + b0 + q);
                acc += settle(x, pnj + b0 + q, err, poll_code(POLL_SETTLE, j));
            }
            acc = wave_sum(acc);
            const ColScal sc = reflector_scalars(araw, acc);
            if (lane == 0) {
                sh.cur = sc;
                if (g == 0) {
                    tau_out[j] = sc.tau;
                    eta_out[j] = sc.beta;
                }
            }
        }
        alive = ok;
    }
    if (__syncthreads_or(!alive)) return false;
    clk_inner(clk, 8);
    // ---- empty this rank's t slots of column j + 2 (copy (j + 2) mod 3 = (j - 1) mod 3): column
    // j - 1's t was consumed by the row passes of column j, all finished (above) ----
    if (lane == 0)
        for (int e = e0; e < 2 * NBMAX; e += 4 * G) st_co(tv_of(w, j + 2) + e, empty_slot());
    if (!mv) return true;
    // ---- empty this rank's matvec hand-over slots of column j + 2 (copy (j + 2) mod 3 = (j - 1) mod 3),
    // after its own partials of column j are on their way, off the path from the poll to the matvec.
    // Column j - 1's values there were consumed by the row passes of column j, which all finished
    // before their norm partials of column j, which wave 0 saw before the barrier above.  The next
    // phase B drains these stores before it stores anything of column j + 1, and a reader polls
    // column j + 2's slots only after a row pass of column j + 2 consumed column j + 1's partials. ----
    {
        double* prn = prow_of(w, j + 2);
        double* pcn = pcol_of(w, j + 2);
        const double e = empty_slot();
        int64_t It = pm.I, Jt = pm.J, slot = pm.slot;
        for (int64_t t = tlo; t < thi; t++) {
            if (wave == 0 && lane < MC) {
                const int64_t cc = Jt * MC + lane;
                if (cc > j + 2 && cc < n) st_co(pcn + It * w.ldp + cc, e);
            }
            const bool lis = (Jt == MPS * It + MPS - 1);
            if (lis || t + 1 == thi) {
                const int64_t r = It * MR + tid;
                if (r < n) st_co(prn + slot * w.ldp + r, e);
                slot = 0;
            }
            if (lis) {
                It++;
                Jt = tg.Jlo;
            } else {
                Jt++;
            }
        }
    }
    return true;
}

// ---------------------------------------------------------------------------------------
// Skew rank-2k trailing update on FP64 MFMA (src/hessenberg.jl:158-166):
//   A[r,c] += sum_k L[r,k] R[c,k]   for r > c >= e,
//   L = [V U], R = [U -V], V[:,k] = A[:, c0+k] (reflector k, its unit element in place),
//   U[:,k] = W[:,k], k < nbp.
// One wave owns a 32x32 output tile (rows r0.., cols q0..) and runs without LDS or
// barriers: its operands stream from L2 straight into the MFMA operand registers, one k-step
// ahead, through a uniform base pointer plus a 32-bit lane offset.  Independent waves drift
// out of phase, so one wave's C traffic overlaps another's MFMAs (an LDS-staged 128x128
// workgroup tile moves its C in lockstep and was 1.3-1.8x slower at n = 8192).
// v_mfma_f64_16x16x4f64 (lane layout measured on gfx950 by k_mfma_probe): a-operand lane
// l = (row l%16, k l/16); b-operand lane l = (k l/16, col l%16); accumulator item q of lane
// l = (row 4q + l/16, col l%16).  Each 16x16 block is computed transposed, D = R_blk L_blk^T,
// so item q of lane l is C[r + l%16, c + 4q + l/16]: one item of a wave is four 128-byte
// column segments.
// ---------------------------------------------------------------------------------------
#ifndef SKH_R2K_TR
#define SKH_R2K_TR 32
#endif
// 32 x 32 wave tiles: four per 64 x 64 block, about 2.6 tiles per wave slot over an n = 8192
// factorization's launches and 4 waves per SIMD (32 x 64: 15.14 ms of rank-2k per factorization,
// 32 x 32: 14.40 ms, 16 x 64: 18.21 ms)
#ifndef SKH_R2K_TC
#define SKH_R2K_TC 32
#endif
constexpr int R2K_BLK = 64;          // edge of a tile block (the unit of the supertile order)
constexpr int R2K_TR = SKH_R2K_TR;   // rows of a wave tile
constexpr int R2K_TC = SKH_R2K_TC;   // columns of a wave tile
constexpr int R2K_BI = R2K_TR / 16;  // 16-row MFMA blocks of a wave tile
constexpr int R2K_BJ = R2K_TC / 16;  // 16-column MFMA blocks
constexpr int R2K_WPB = (R2K_BLK / R2K_TR) * (R2K_BLK / R2K_TC);   // wave tiles per tile block
constexpr int R2K_SUPER = 16;  // tile blocks per supertile edge (the L2 reuse unit)

// lane offsets (doubles) of the operand rows: V and W columns are at uniform offsets on top
struct R2kRows {
    uint32_t lv[R2K_BI], lu[R2K_BI];   // L rows r0 + 16 bi + lane%16, V and W layouts
    uint32_t rv[R2K_BJ], ru[R2K_BJ];   // R rows q0 + 16 bj + lane%16
};

// k-step s of 4: h = 0 gives L = V, R = W; h = 1 gives L = W, R = -V.  Lanes with
// 4 ks + lane/16 >= nbp read nothing (zero operands pad nbp to a multiple of 4).
__device__ __forceinline__ void r2k_operands(const double* __restrict__ A, int64_t lda, const double* __restrict__ U,
                                             int64_t ldu, int64_t c0, int nbp, int nkh, int s, int lq,
                                             const R2kRows& o, double lf[R2K_BI], double rf[R2K_BJ])
{
    const int h = s >= nkh;
    const int ks = s - h * nkh;
    const bool ok = ks * 4 + lq < nbp;
    const double* pv = A + (c0 + ks * 4) * lda;
    const double* pu = U + (int64_t)ks * 4 * ldu;
    if (h == 0) {
#pragma unroll
        for (int b = 0; b < R2K_BI; b++) lf[b] = ok ? pv[o.lv[b]] : 0.0;
#pragma unroll
        for (int b = 0; b < R2K_BJ; b++) rf[b] = ok ? pu[o.ru[b]] : 0.0;
    } else {
#pragma unroll
        for (int b = 0; b < R2K_BI; b++) lf[b] = ok ? pu[o.lu[b]] : 0.0;
#pragma unroll
        for (int b = 0; b < R2K_BJ; b++) rf[b] = ok ? -pv[o.rv[b]] : 0.0;
    }
}

__device__ __forceinline__ void r2k_mfma(v4d acc[R2K_BI][R2K_BJ], const double lf[R2K_BI], const double rf[R2K_BJ])
{
#pragma unroll
    for (int bi = 0; bi < R2K_BI; bi++)
#pragma unroll
        for (int bj = 0; bj < R2K_BJ; bj++)
            acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(rf[bj], lf[bi], acc[bi][bj], 0, 0, 0);
}

// One wave: rows [r0, r0+32) x cols [q0, q0+64) of the update, q0 <= r0, stored where
// r > c >= e and r < n.  Operand rows past n are clamped to n-1 (their products land in
// rows or columns that are not stored); rows and columns below e are computed and dropped.
__device__ __forceinline__ void rank2k_wave(double* __restrict__ A, int64_t lda, const double* __restrict__ U,
                                            int64_t ldu, int64_t n, int64_t e, int64_t c0, int nbp, int64_t r0,
                                            int64_t q0)
{
    const int lane = threadIdx.x & 63, lc = lane & 15, lq = lane >> 4;
    R2kRows o;
#pragma unroll
    for (int b = 0; b < R2K_BI; b++) {
        const int64_t rl = min(r0 + b * 16 + lc, n - 1);
        o.lv[b] = (uint32_t)(rl + lq * lda);
        o.lu[b] = (uint32_t)(rl + lq * ldu);
    }
#pragma unroll
    for (int b = 0; b < R2K_BJ; b++) {
        const int64_t rr = min(q0 + b * 16 + lc, n - 1);
        o.rv[b] = (uint32_t)(rr + lq * lda);
        o.ru[b] = (uint32_t)(rr + lq * ldu);
    }
    // C tile: uniform base + lane offset; item (bi, bj, q) at bi*16 + (bj*16 + 4q)*lda
    double* C = A + r0 + q0 * lda;
    const uint32_t co = (uint32_t)(lc + lq * lda);
    const bool inner = (q0 + R2K_TC - 1 < r0) && (r0 + R2K_TR - 1 < n) && (q0 >= e);
    v4d acc[R2K_BI][R2K_BJ];
    if (inner) {
#pragma unroll
        for (int bi = 0; bi < R2K_BI; bi++)
#pragma unroll
            for (int bj = 0; bj < R2K_BJ; bj++) {
                v4d c;
#pragma unroll
                for (int q = 0; q < 4; q++) c[q] = C[co + bi * 16 + (bj * 16 + 4 * q) * lda];
                acc[bi][bj] = c;
            }
    } else {
#pragma unroll
        for (int bi = 0; bi < R2K_BI; bi++)
#pragma unroll
            for (int bj = 0; bj < R2K_BJ; bj++) {
                const int64_t row = r0 + lc + bi * 16;
                v4d c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int64_t col = q0 + lq + bj * 16 + 4 * q;
                    if (row < n && col >= e && col < row) c[q] = C[co + bi * 16 + (bj * 16 + 4 * q) * lda];
                }
                acc[bi][bj] = c;
            }
    }
    const int nkh = (nbp + 3) / 4, S = 2 * nkh;
    double la[R2K_BI], ra[R2K_BJ], lb[R2K_BI], rb[R2K_BJ];
    r2k_operands(A, lda, U, ldu, c0, nbp, nkh, 0, lq, o, la, ra);
    for (int s = 0; s < S; s += 2) {
        if (s + 1 < S) r2k_operands(A, lda, U, ldu, c0, nbp, nkh, s + 1, lq, o, lb, rb);
        r2k_mfma(acc, la, ra);
        if (s + 1 >= S) break;
        if (s + 2 < S) r2k_operands(A, lda, U, ldu, c0, nbp, nkh, s + 2, lq, o, la, ra);
        r2k_mfma(acc, lb, rb);
    }
    if (inner) {
#pragma unroll
        for (int bi = 0; bi < R2K_BI; bi++)
#pragma unroll
            for (int bj = 0; bj < R2K_BJ; bj++)
#pragma unroll
                for (int q = 0; q < 4; q++) C[co + bi * 16 + (bj * 16 + 4 * q) * lda] = acc[bi][bj][q];
        return;
    }
#pragma unroll
    for (int bi = 0; bi < R2K_BI; bi++) {
        const int64_t row = r0 + lc + bi * 16;
#pragma unroll
        for (int bj = 0; bj < R2K_BJ; bj++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int64_t col = q0 + lq + bj * 16 + 4 * q;
                if (row < n && col >= e && col < row) C[co + bi * 16 + (bj * 16 + 4 * q) * lda] = acc[bi][bj][q];
            }
    }
}

// Triangle index t -> (i, j), j <= i, row-major: row i holds i + 1 entries.
__host__ __device__ inline void tri_ij(int64_t t, int64_t& i, int64_t& j)
{
    int64_t x = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (x * (x + 1) / 2 > t) x--;
    while ((x + 1) * (x + 2) / 2 <= t) x++;
    i = x;
    j = t - x * (x + 1) / 2;
}

// Lower triangle of nt x nt tile blocks in supertile order: supertiles of R2K_SUPER x
// R2K_SUPER blocks in row-major triangle order, blocks row-major inside each supertile, so
// that consecutive indices (one XCD's share of the launch) reuse the same operand rows in L2.
__host__ __device__ inline void super_ij(int64_t t, int64_t nt, int64_t& i, int64_t& j)
{
    const int64_t S = R2K_SUPER;
    for (int64_t I = 0;; I++) {
        const int64_t rI = (nt - I * S < S) ? nt - I * S : S;   // block rows in supertile row I
        const int64_t full = I * rI * S;                          // blocks left of the diagonal supertile
        const int64_t cnt = full + rI * (rI + 1) / 2;
        if (t < cnt) {
            if (t < full) {
                const int64_t J = t / (rI * S), u = t - J * rI * S;
                i = I * S + u / S;
                j = J * S + u % S;
            } else {
                int64_t a, b;
                tri_ij(t - full, a, b);
                i = I * S + a;
                j = I * S + b;
            }
            return;
        }
        t -= cnt;
    }
}

// log|Pf| and sign from eta, tau of an even-n factorization (src/pfaffian.jl:170-178):
// logpf = sum_{odd i} log|ev_i|, sign = prod sign(-ev_i) * (-1)^{#tau != 0}.
__device__ void logpf_reduce(const double* eta, const double* tau, int64_t n, double* out_lp, double* out_sg)
{
    __shared__ double s_l[4];
    __shared__ int s_c[4][3];
    const int tid = threadIdx.x, lane = tid & 63, wave = wave_id();
    double l = 0.0;
    int nneg = 0, nzero = 0, ntau = 0;
    for (int64_t i = tid; i < n - 1; i += WG) {
        if ((i & 1) == 0) {
            const double e = eta[i];
            l += log(fabs(e));
            nneg += (e > 0.0);   // sign(-e) < 0
            nzero += (e == 0.0);
        }
        ntau += (tau[i] != 0.0);
    }
    l = wave_sum(l);
    double dn = wave_sum((double)nneg), dz = wave_sum((double)nzero), dt = wave_sum((double)ntau);
    if (lane == 0) {
        s_l[wave] = l;
        s_c[wave][0] = (int)dn;
        s_c[wave][1] = (int)dz;
        s_c[wave][2] = (int)dt;
    }
    __syncthreads();
    if (tid == 0) {
        double L = s_l[0] + s_l[1] + s_l[2] + s_l[3];
        int cn = s_c[0][0] + s_c[1][0] + s_c[2][0] + s_c[3][0];
        int cz = s_c[0][1] + s_c[1][1] + s_c[2][1] + s_c[3][1];
        int ct = s_c[0][2] + s_c[1][2] + s_c[2][2] + s_c[3][2];
        *out_lp = L;
        *out_sg = (cz > 0) ? 0.0 : (((cn + ct) & 1) ? -1.0 : 1.0);
    }
    __syncthreads();
}


}  // namespace skh
