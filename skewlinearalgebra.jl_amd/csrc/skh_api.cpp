// skh_api.cpp -- the C ABI of libskewhip.so (declared in include/skewhip.h).
//
// Host-memory entry points copy in, run the device engine on the library's stream and copy
// back; device entry points run on the caller's stream.  Workspaces are cached per device
// and only grow.  All entry points are serialised by one mutex.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <rocsolver/rocsolver.h>

#include "../../include/skewhip.h"
#include "skh_internal.h"

namespace {

std::mutex g_mu;
thread_local std::string g_err;
bool g_inited = false;
int g_device = 0;
int g_num_cu = 0;
hipStream_t g_stream = nullptr;
int g_noalloc_status = 0;   // result of looking up the fixed scratch at init (hipError_t)

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    void* get(size_t want)
    {
        if (want > bytes) {
            if (p) (void)hipFree(p);
            p = nullptr;
            bytes = 0;
            if (hipMalloc(&p, want) != hipSuccess) {
                p = nullptr;
                return nullptr;
            }
            bytes = want;
        }
        return p;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

Buf g_mat, g_ws, g_vec, g_ctl, g_bin, g_bout, g_q, g_ev;
rocblas_handle g_blas = nullptr;

// Waits for a stream when it goes out of scope.  Host memory that asynchronous copies on that
// stream read or write (local vectors, staging areas, the caller's arrays) must outlive those
// copies on every return path, error returns included: declare the guard after such buffers, so
// that it is destroyed, and the stream drained, before they are.
struct StreamDrain {
    hipStream_t st;
    explicit StreamDrain(hipStream_t s) : st(s) {}
    StreamDrain(const StreamDrain&) = delete;
    StreamDrain& operator=(const StreamDrain&) = delete;
    ~StreamDrain() { (void)hipStreamSynchronize(st); }
};

// roctx range around the host-side enqueue of an engine or a launch (rocprofv3 --marker-trace shows
// them beside the kernels; without a profiler a push / pop costs a call into the roctx library).
struct Range {
    explicit Range(const char* fmt, ...) __attribute__((format(printf, 2, 3)))
    {
        char buf[128];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        roctxRangePushA(buf);
    }
    Range(const Range&) = delete;
    Range& operator=(const Range&) = delete;
    ~Range() { roctxRangePop(); }
};

int set_err(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) return set_err(SKH_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

int ensure_init_locked(int device)
{
    if (g_inited) return 0;
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    if (count <= 0) return set_err(SKH_ERR_HIP, "no HIP device visible");
    if (device < 0 || device >= count) return set_err(-1, "device %d out of range (have %d)", device, count);
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    g_num_cu = prop.multiProcessorCount;
    int coop = 0;
    HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device));
    if (!coop) return set_err(SKH_ERR_UNSUPPORTED, "device does not support cooperative launch");
    HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    // device address of the allocation-free entry point's scratch, with `device` current; a failure
    // is reported by that entry point only
    g_noalloc_status = skh_internal_noalloc_setup(device);
    g_device = device;
    g_inited = true;
    return 0;
}

// autotuned panel width per size (refreshed from measurements; see the bench/tuning notes)
int64_t pick_nb(int64_t n, const skh_opts* o)
{
    if (o && o->nb > 0) return std::min<int64_t>(o->nb, skh::NBMAX);
    // measured on MI355X (whole factorizations): nb = 32 is faster up to n = 8192 (30.4 / 74.3 /
    // 249 ms at n = 2048 / 4096 / 8192 against 32.9 / 78.1 / 258 ms at 64); at n = 16384 the two
    // tie (1649 ms) and 64 runs the trailing update closer to the matrix-core peak
    return n < 12288 ? 32 : 64;
}

// compact engine limits: even n up to this size (one workgroup per matrix)
constexpr int64_t CMP_NMAX = 1024;
// blocked one-workgroup engine: even n up to this size; panel width 8 (n <= 512) or 16 (n <= 256)
constexpr int64_t BLK_NMAX = 512;

// (kernel NMAX, panel width) of the blocked engine for matrices up to nmax, panel width nb (0 = 8)
bool blocked_instance(int64_t nmax, int64_t nb, int* kmax, int* knb)
{
    const int w = nb == 16 ? 16 : 8;
    if (nb != 0 && nb != 8 && nb != 16) return false;
    if (nmax <= 256) {
        *kmax = 256;
        *knb = w;
        return true;
    }
    if (nmax <= 512) {
        *kmax = 512;
        *knb = w;
        return true;
    }
    return false;
}
// automatic choice for a single matrix: one workgroup (the blocked engine, k_cblock) up to this size, the
// large engine above.  Measured device log|Pf| calls (MI355X): n = 128 0.71 ms against 1.47 ms, n = 256
// 2.23 against 3.02, n = 320 3.64 against 3.62, n = 384 5.38 against 4.37, n = 512 10.4 against 5.94
// (batches keep one workgroup per matrix to n = 1024: there every workgroup has a matrix of its own)
constexpr int64_t CMP_AUTO = 320;

int engine_of(int64_t n, const skh_opts* o)
{
    int64_t f = o ? (o->flags & 0xff) : 0;
    if (f == SKH_ENGINE_LARGE || f == SKH_ENGINE_TEAM1 || f == SKH_ENGINE_BLOCKED) return (int)f;
    return (n % 2 == 0 && n <= CMP_AUTO) ? SKH_ENGINE_TEAM1 : SKH_ENGINE_LARGE;
}

// Optional per-kernel timing (SKH_FLAG_TIMING): event pairs around every panel and rank-2k
// launch of the last large-engine factorization, with each launch's algorithmic work.
struct KTimer {
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;        // 0 = panel, 1 = rank-2k
    std::vector<double> work;     // bytes (panel) or flops (rank-2k)
    size_t used = 0;
    bool on = false;
    bool phases = false;          // SKH_FLAG_PHASES: the panel kernel's phase clock as well
    int begin(hipStream_t st)
    {
        if (used + 2 > ev.size()) {
            for (int k = 0; k < 64; k++) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return 1;
                ev.push_back(e);
            }
        }
        return (int)hipEventRecord(ev[used], st);
    }
    int end(hipStream_t st, int k, double w)
    {
        int rc = (int)hipEventRecord(ev[used + 1], st);
        kind.push_back(k);
        work.push_back(w);
        used += 2;
        return rc;
    }
    void reset()
    {
        used = 0;
        kind.clear();
        work.clear();
    }
    void release()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        ev.clear();
        reset();
    }
};
KTimer g_timer;

// control block: [0] barrier counter, [1] error flag, [2] batch queue
struct Ctl {
    unsigned* bar;
    int* err;
    unsigned long long* ph;   // phase clock of the panel kernel (4 counters)
    unsigned* cqueue;         // compact engine: work queue and error flag (own cache lines)
    int* cerr;
};

// words [0, 288): barrier counters (9 lines of 32 words); [320]: error flag of the panel
// kernel; [384, 410): phase clocks (13 x u64); [448]: compact queue; [480]: compact error flag
constexpr size_t CTL_BYTES = 2048;

void ctl_at(unsigned* p, Ctl& c)
{
    c.bar = p;
    c.err = (int*)(p + 320);
    c.ph = (unsigned long long*)(p + 384);
    c.cqueue = p + 448;
    c.cerr = (int*)(p + 480);
}

int get_ctl(Ctl& c)
{
    unsigned* p = (unsigned*)g_ctl.get(CTL_BYTES);
    if (!p) return set_err(SKH_ERR_HIP, "hipMalloc of control block failed");
    ctl_at(p, c);
    return 0;
}

// ----------------------------------------------------------------------------------------
// Large engine: persistent panel kernel (every workgroup resident) + MFMA rank-2k update per panel.
// The workspace and control block are the library's cached buffers, or (ws_ext, ctl_ext) given by
// the caller: ws_cap doubles and CTL_BYTES, nothing allocated then.
// ----------------------------------------------------------------------------------------
int sytrd_large(int64_t n, double* dA, int64_t lda, double* dtau, double* deta, int64_t nb, hipStream_t st,
                double* ws_ext = nullptr, size_t ws_cap = 0, unsigned* ctl_ext = nullptr)
{
    Range rg("skh sytrd n=%lld nb=%lld", (long long)n, (long long)nb);
    Ctl ctl;
    if (ctl_ext) {
        ctl_at(ctl_ext, ctl);
    } else {
        int rc = get_ctl(ctl);
        if (rc) return rc;
    }
    // the matvec reads 16-byte row pairs at even rows < n: an even leading dimension keeps them
    // inside their column
    if ((lda & 1) || lda < n) return set_err(-3, "lda must be >= n and even");
    const int64_t ld = (n + 1) & ~int64_t(1);
    const int64_t nt = (n + skh::TB - 1) / skh::TB + 1;
    const int64_t nts = (n + skh::MR - 1) / skh::MR + 1;
    int per_cu0 = 0;
    HIPCHK((hipError_t)skh_internal_panel_occupancy((int)nb, &per_cu0));
    // row-partial slots of one matvec strip: at most one per workgroup of the grid
    const int64_t nseg = (int64_t)std::max(1, std::min(per_cu0, 2)) * g_num_cu + 2;
    // W | V2 (two pairs when double-buffered) | (Pnorm | PS | X) x 4 | (Tv | Prow | Pcol) x 3: the
    // copies of the hand-over slots by column mod 4 (written by the row passes) and mod 3 (written by
    // phase B).  Up to nb = 64 the panels alternate between two W | V2 pairs, so the rank-2k launch of
    // one panel (which reads its W) can fill the next panel's slots with the empty marker itself.
    const bool dbuf = nb <= 64;
    const int64_t wpair = 2 * ld * nb;
    const int64_t ostride = (nt + nt * 2 * skh::NBMAX + ld + 1) & ~int64_t(1);
    const int64_t hsize = 2 * skh::NBMAX + nseg * ld + nts * ld;
    size_t doubles = (size_t)((dbuf ? 2 : 1) * wpair + skh::OWN_COPIES * ostride + skh::HAND_COPIES * hsize + 64);
    double* ws;
    if (ws_ext) {
        if (doubles > ws_cap)
            return set_err(SKH_ERR_UNSUPPORTED, "n = %lld, nb = %lld needs %zu doubles of workspace, the scratch "
                           "holds %zu", (long long)n, (long long)nb, doubles, ws_cap);
        ws = ws_ext;
    } else {
        ws = (double*)g_ws.get(doubles * sizeof(double));
        if (!ws) return set_err(SKH_ERR_HIP, "hipMalloc of %zu bytes of workspace failed", doubles * 8);
    }
    skh::Work w;
    w.W = ws;
    w.V2 = w.W + ld * nb;
    w.ldw = ld;
    w.Pnorm = ws + (dbuf ? 2 : 1) * wpair;
    w.PS = w.Pnorm + nt;
    w.X = w.PS + nt * 2 * skh::NBMAX;
    w.ostride = ostride;
    w.Tv = w.Pnorm + skh::OWN_COPIES * ostride;
    w.Prow = w.Tv + 2 * skh::NBMAX;
    w.ldp = ld;
    w.Pcol = w.Prow + nseg * ld;
    w.hstride = hsize;

    int per_cu = 0;
    HIPCHK((hipError_t)skh_internal_panel_occupancy((int)nb, &per_cu));
    per_cu = std::max(1, std::min(per_cu, 2));
    const int grid = per_cu * g_num_cu;
    if ((double)skh::tall_grid(n, 1).T * (double)grid >= 4294967296.0)
        return set_err(SKH_ERR_UNSUPPORTED, "n = %lld too large for the 32-bit tile ownership", (long long)n);

    // target matvec tiles per workgroup (tuning knob SKH_PANEL_TPW; measured default) and an
    // optional cap on the grid (SKH_PANEL_GRID_CAP, for tests of workgroups owning several blocks)
    static const int64_t tpw = [] {
        const char* e = getenv("SKH_PANEL_TPW");
        const long v = e ? atol(e) : 0;
        return (int64_t)(v > 0 ? v : 1);
    }();
    const int64_t gcap = [] {
        const char* e = getenv("SKH_PANEL_GRID_CAP");
        const long v = e ? atol(e) : 0;
        return (int64_t)(v > 0 ? v : 1 << 30);
    }();
    // grid of the panel starting at c0, sized to its first matvec: about `tpw` tiles per workgroup, and
    // at least one workgroup per 64-row block of the panel (each keeps its block's rows)
    auto panel_grid = [&](int64_t c0) {
        const skh::TallGrid tg = skh::tall_grid(n, c0 + 1);
        const int64_t blocks = (n + skh::TB - 1) / skh::TB - c0 / skh::TB;
        const int64_t want = std::max<int64_t>(std::max<int64_t>(8, blocks), (tg.T + tpw - 1) / tpw);
        return (int)std::min<int64_t>(std::min<int64_t>(grid, want), gcap);
    };
    // the panel's W | V2 pair
    auto panel_w = [&](int64_t c0) {
        skh::Work pw = w;
        if (dbuf && ((c0 / nb) & 1)) {
            pw.W = ws + wpair;
            pw.V2 = pw.W + ld * nb;
        }
        return pw;
    };
    // the slots polled before any producer of the panel's launch has refilled them start empty: the
    // hand-over copies of the launch's first two columns (t, rows >= c0 of the row-partial slots in
    // use and of the column partials), the row-pass copies of its first two columns, and the panel's
    // columns of W and V2
    auto panel_fill = [&](int64_t c0, int nbp, int pgrid, const skh::Work& pw, FillSegs& fs) {
        const skh::TallGrid tg = skh::tall_grid(n, c0 + 1);
        int64_t maxslot = 1;
        for (int64_t I = tg.I0; I < tg.nts; I++) maxslot = std::max(maxslot, skh::tall_slots(tg, I, pgrid));
        int ns = 0;
        auto seg = [&](double* p, int64_t rows, int64_t ldr, int64_t lo, int64_t hi) {
            fs.p[ns] = p;
            fs.rows[ns] = rows;
            fs.ld[ns] = ldr;
            fs.lo[ns] = lo;
            fs.hi[ns] = hi;
            ns++;
        };
        for (int64_t cc = c0; cc < c0 + 2; cc++) {
            double* h = w.Tv + (cc % skh::HAND_COPIES) * hsize;
            seg(h, 1, 0, 0, 2 * skh::NBMAX);
            seg(h + 2 * skh::NBMAX, maxslot, ld, c0, n);
            seg(h + 2 * skh::NBMAX + nseg * ld, tg.nts, ld, c0, n);
            seg(w.Pnorm + (cc & (skh::OWN_COPIES - 1)) * ostride, 1, 0, 0, ostride);
        }
        // t of the third copy too: an element that has no part in the launch's first columns
        // empties its slots only from the column it becomes live
        seg(w.Tv + ((c0 + 2) % skh::HAND_COPIES) * hsize, 1, 0, 0, 2 * skh::NBMAX);
        seg(pw.W, nbp, ld, c0, n);
        seg(pw.V2, nbp, ld, c0, n);
        return ns;
    };
    // error flag and phase clocks start at zero
    HIPCHK((hipError_t)skh_internal_zero_u32(ctl.bar, 512, st));
    for (int64_t c0 = 0; c0 <= n - 3; c0 += nb) {
        const int nbp = (int)std::min<int64_t>(nb, n - 2 - c0);
        PanelLaunch pl;
        pl.A = dA;
        pl.lda = lda;
        pl.n = n;
        pl.c0 = c0;
        pl.nbp = nbp;
        pl.w = panel_w(c0);
        pl.tau = dtau;
        pl.eta = deta;
        pl.err = ctl.err;
        pl.grid = panel_grid(c0);
        pl.nbmax = (int)nb;
        {
            const int64_t blocks = (n + skh::TB - 1) / skh::TB - c0 / skh::TB;
            pl.multi = (int64_t)pl.grid < blocks ? 1 : 0;
        }
        pl.ph = g_timer.phases ? ctl.ph : nullptr;
        if (c0 == 0 || !dbuf) {   // otherwise the previous rank-2k launch did it
            FillSegs fs;
            const int ns = panel_fill(c0, nbp, pl.grid, pl.w, fs);
            HIPCHK((hipError_t)skh_internal_fill_empty(&fs, ns, st));
        }
        const bool tm = g_timer.on;
        Range rp("skh panel + rank2k c0=%lld", (long long)c0);   // both launches of this panel
        if (tm) HIPCHK((hipError_t)g_timer.begin(st));
        HIPCHK((hipError_t)skh_internal_launch_panel(&pl, st));
        if (tm) {
            // algorithmic bytes: strictly lower triangle of each column's trailing block, once
            double bytes = 0.0;
            for (int64_t j = c0; j < c0 + nbp; j++) {
                const double m = (double)(n - j - 1);
                bytes += 8.0 * m * (m - 1.0) / 2.0;
            }
            HIPCHK((hipError_t)g_timer.end(st, 0, bytes));
        }
        const int64_t e = c0 + nbp;
        // the next panel's slots, filled by extra workgroups of this rank-2k launch
        FillSegs fnext;
        int nsnext = 0;
        const int64_t c1 = c0 + nb;
        if (dbuf && c1 <= n - 3) {
            const int nbp1 = (int)std::min<int64_t>(nb, n - 2 - c1);
            nsnext = panel_fill(c1, nbp1, panel_grid(c1), panel_w(c1), fnext);
        }
        if (tm) HIPCHK((hipError_t)g_timer.begin(st));
        HIPCHK((hipError_t)skh_internal_launch_rank2k(dA, lda, pl.w.W, w.ldw, n, e, c0, nbp, nsnext ? &fnext : nullptr,
                                                      nsnext, st));
        if (tm) {
            const double m = (double)(n - e);
            HIPCHK((hipError_t)g_timer.end(st, 1, 2.0 * m * (m - 1.0) * nbp));
        }
    }
    HIPCHK((hipError_t)skh_internal_launch_finish(dA, lda, n, dtau, deta, st));
    return 0;
}

// ----------------------------------------------------------------------------------------
// Batched factorizations.  Matrices with even 2 <= n <= CMP_NMAX go through the compact engine
// (one workgroup per matrix, one launch for all of them, largest first); larger even ones
// through the large engine one after another; odd and empty ones get their closed-form result
// in the compact launch.  With dfac, each factorization runs in place in dfac (input copied
// in first; offsets as the input) and tau|eta go to dtaueta + toffs[b].
// ----------------------------------------------------------------------------------------
Buf g_cslot, g_big, g_bctl, g_cwbuf;

double watchdog_seconds()
{
    const char* e = getenv("SKH_WATCHDOG_SECONDS");
    double s = e ? atof(e) : 30.0;
    return s > 0 ? s : 0.0;
}

int check_barrier_flag();

int batched_run(int64_t batch, const int64_t* ns, const int64_t* offs, const double* dA, double* dlogpf,
                double* dsgn, const skh_opts* opts, hipStream_t st, double* dfac = nullptr, double* dtaueta = nullptr,
                const int64_t* htoffs = nullptr)
{
    if (batch == 0) return 0;
    Range rg("skh logpf batched %lld matrices", (long long)batch);
    Ctl ctl;
    int rc = get_ctl(ctl);
    if (rc) return rc;
    std::vector<int64_t> small, big;
    for (int64_t b = 0; b < batch; b++) {
        if (ns[b] > CMP_NMAX && ns[b] % 2 == 0) big.push_back(b);
        else small.push_back(b);
    }
    std::stable_sort(small.begin(), small.end(), [&](int64_t a, int64_t b) { return ns[a] > ns[b]; });

    // large matrices: one large-engine factorization each, then log|Pf| from eta, tau on the device
    for (int64_t b : big) {
        const int64_t n = ns[b];
        double* M;
        double* te;
        if (dfac) {
            M = dfac + offs[b];
            te = dtaueta + htoffs[b];
        } else {
            double* w = (double*)g_big.get((size_t)(n * n + 2 * n) * sizeof(double));
            if (!w) return set_err(SKH_ERR_HIP, "hipMalloc of %lld doubles failed", (long long)(n * n + 2 * n));
            M = w;
            te = w + n * n;
        }
        if (M != dA + offs[b])
            HIPCHK(hipMemcpyAsync(M, dA + offs[b], (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, st));
        g_timer.reset();
        g_timer.on = false;
        g_timer.phases = false;
        rc = sytrd_large(n, M, n, te, te + n, pick_nb(n, opts), st);
        if (rc) return rc;
        HIPCHK((hipError_t)skh_internal_logpf_from(te + n, te, n, dlogpf + b, dsgn + b, st));
    }
    if (!big.empty()) {
        HIPCHK(hipStreamSynchronize(st));
        rc = check_barrier_flag();
        if (rc) return rc;
    }
    if (small.empty()) return 0;

    int64_t nmax = 2;
    for (int64_t b : small) nmax = std::max(nmax, ns[b]);
    // the blocked engine on request, and by default for batches whose members all fit it; the
    // unblocked one (SKH_ENGINE_TEAM1) on request and for larger members
    const int64_t eflag = opts ? (opts->flags & 0xff) : 0;
    const bool blocked = eflag == SKH_ENGINE_BLOCKED || (eflag != SKH_ENGINE_TEAM1 && nmax <= BLK_NMAX);
    int kmax = nmax <= 512 ? 512 : 1024, knb = 0;
    if (blocked && !blocked_instance(nmax, eflag == SKH_ENGINE_BLOCKED ? opts->nb : 0, &kmax, &knb))
        return set_err(SKH_ERR_UNSUPPORTED, "the blocked engine takes even n <= %lld with panel width 8, n <= 256 "
                       "with width 16 (largest n %lld, nb %lld)", (long long)BLK_NMAX, (long long)nmax,
                       (long long)opts->nb);
    int per_cu = 0;
    if (blocked) HIPCHK((hipError_t)skh_internal_cblock_occupancy(kmax, knb, &per_cu));
    else HIPCHK((hipError_t)skh_internal_compact_occupancy(kmax, &per_cu));
    per_cu = std::max(1, per_cu);
    const int64_t count = (int64_t)small.size();
    const int grid = (int)std::min<int64_t>((int64_t)per_cu * g_num_cu, count);
    const int64_t slot = ((nmax * nmax + 2 * kmax) + 31) & ~int64_t(31);
    double* slots = nullptr;
    if (!dfac) {
        slots = (double*)g_cslot.get((size_t)slot * grid * sizeof(double));
        if (!slots) return set_err(SKH_ERR_HIP, "hipMalloc of compact work slots failed");
    }
    // control arrays: ns, offs, items[, toffs] (device copies)
    const int64_t nctl = 2 * batch + count + (dtaueta ? batch : 0);
    const size_t ctl_bytes = sizeof(int64_t) * nctl;
    int64_t* dctl = (int64_t*)g_bctl.get(ctl_bytes);
    if (!dctl) return set_err(SKH_ERR_HIP, "hipMalloc of batched control arrays failed");
    std::vector<int64_t> h(nctl);
    StreamDrain drain(st);   // h is read by the upload below
    std::copy(ns, ns + batch, h.begin());
    std::copy(offs, offs + batch, h.begin() + batch);
    std::copy(small.begin(), small.end(), h.begin() + 2 * batch);
    if (dtaueta) std::copy(htoffs, htoffs + batch, h.begin() + 2 * batch + count);
    HIPCHK(hipMemcpyAsync(dctl, h.data(), ctl_bytes, hipMemcpyHostToDevice, st));
    HIPCHK((hipError_t)skh_internal_zero_u32(ctl.cqueue, 64, st));
    CompactArgs a;
    a.A = dA;
    a.ns = dctl;
    a.offs = dctl + batch;
    a.items = dctl + 2 * batch;
    a.count = count;
    a.slots = slots;
    a.slot_stride = slot;
    a.queue = ctl.cqueue;
    a.err = ctl.cerr;
    a.budget = (unsigned long long)(watchdog_seconds() * 1e8);
    a.logpf = dlogpf;
    a.sgn = dsgn;
    a.fac = dfac;
    a.taueta = dtaueta;
    a.toffs = dtaueta ? dctl + 2 * batch + count : nullptr;
    a.wbuf = nullptr;
    a.wstride = 0;
    if (blocked) {
        a.wstride = (int64_t)kmax * knb;
        a.wbuf = (double*)g_cwbuf.get((size_t)a.wstride * grid * sizeof(double));
        if (!a.wbuf) return set_err(SKH_ERR_HIP, "hipMalloc of the blocked engine's panel workspaces failed");
        HIPCHK((hipError_t)skh_internal_launch_cblock(&a, kmax, knb, grid, st));
    } else {
        HIPCHK((hipError_t)skh_internal_launch_compact(&a, kmax, grid, st));
    }
    // the host copy of the control arrays must outlive the async copy; the flag is read after it
    HIPCHK(hipStreamSynchronize(st));
    int herr = 0;
    HIPCHK(hipMemcpy(&herr, ctl.cerr, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) return set_err(SKH_ERR_BARRIER, "compact engine stopped by its watchdog after %.0f s", watchdog_seconds());
    return 0;
}

#define BLASCHK(expr)                                                                         \
    do {                                                                                      \
        rocblas_status _s = (expr);                                                           \
        if (_s != rocblas_status_success) return set_err(SKH_ERR_HIP, "%s failed: %s", #expr, \
                                                         rocblas_status_to_string(_s));       \
    } while (0)

int blas_on(hipStream_t st)
{
    if (!g_blas) BLASCHK(rocblas_create_handle(&g_blas));
    BLASCHK(rocblas_set_stream(g_blas, st));
    BLASCHK(rocblas_set_pointer_mode(g_blas, rocblas_pointer_mode_host));
    return 0;
}

// Q (n x n, ldq) = H_0 ... H_{n-2} from factors F and tau (device), blocked compact WY.
int orgtr_locked(int64_t n, const double* dF, int64_t ldf, const double* dtau, double* dQ, int64_t ldq, hipStream_t st)
{
    if (n <= 0) return 0;
    Range rg("skh orgtr n=%lld", (long long)n);
    if (n > 2147483647LL / 2) return set_err(SKH_ERR_UNSUPPORTED, "n too large for the BLAS interface");
    HIPCHK((hipError_t)skh_internal_eye(dQ, ldq, n, st));
    if (n <= 1) return 0;
    int rc = blas_on(st);
    if (rc) return rc;
    const int64_t nb = 64;
    const int64_t nref = n - 1;
    // workspace: V (n x nb) | Y (nb x n) | Z (nb x n) | Gm (nb x nb) | T (nb x nb)
    const size_t doubles = (size_t)(n * nb + 2 * nb * n + 2 * nb * nb + 64);
    double* ws = (double*)g_ws.get(doubles * sizeof(double));
    if (!ws) return set_err(SKH_ERR_HIP, "hipMalloc of orgtr workspace failed");
    double* V = ws;
    double* Y = V + n * nb;
    double* Z = Y + nb * n;
    double* Gm = Z + nb * n;
    double* T = Gm + nb * nb;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    const int64_t last = ((nref - 1) / nb) * nb;
    for (int64_t c0 = last; c0 >= 0; c0 -= nb) {
        const int nbp = (int)std::min<int64_t>(nb, nref - c0);
        const int64_t m = n - c0 - 1;
        HIPCHK((hipError_t)skh_internal_form_v(dF, ldf, n, c0, nbp, V, m, st));
        BLASCHK(rocblas_dgemm(g_blas, rocblas_operation_transpose, rocblas_operation_none, nbp, nbp, (int)m, &one, V,
                              (int)m, V, (int)m, &zero, Gm, nbp));
        HIPCHK((hipError_t)skh_internal_larft(Gm, nbp, dtau + c0, nbp, T, nbp, st));
        double* Q22 = dQ + (c0 + 1) + (c0 + 1) * ldq;
        BLASCHK(rocblas_dgemm(g_blas, rocblas_operation_transpose, rocblas_operation_none, nbp, (int)m, (int)m, &one, V,
                              (int)m, Q22, (int)ldq, &zero, Y, nbp));
        BLASCHK(rocblas_dgemm(g_blas, rocblas_operation_none, rocblas_operation_none, nbp, (int)m, nbp, &one, T, nbp, Y,
                              nbp, &zero, Z, nbp));
        BLASCHK(rocblas_dgemm(g_blas, rocblas_operation_none, rocblas_operation_none, (int)m, (int)m, nbp, &mone, V,
                              (int)m, Z, nbp, &one, Q22, (int)ldq));
    }
    return 0;
}


// ----------------------------------------------------------------------------------------
// Complex (ComplexF64) path: cooperative unblocked engine, explicit Q, eigen via the real
// symmetric tridiagonal D^H (iT) D (D = diag(i^r)) and a back-transform by Q.
// ----------------------------------------------------------------------------------------
Buf g_zws, g_zq, g_zc, g_zx, g_zbws;

// columns the unblocked engine finishes after the blocked panels (and the size below which it
// takes the whole matrix)
constexpr int64_t ZTAIL = 128, ZBLOCKED_MIN_N = 512;

// Blocked complex panels (skewlatrd! + the complex rank-2k update) for columns [0, kb_end); the
// trailing matrix A[kb_end:, kb_end:] is fully updated on return.  grid >= the 64-row blocks of the
// matrix (checked by the caller).
int zsytrd_blocked(int64_t n, double2* dA, int64_t lda, double2* dtau, double* deta, int64_t kb_end, int grid, Ctl& ctl,
                   hipStream_t st)
{
    Range rg("skh zsytrd blocked n=%lld", (long long)n);
    const int64_t nb = skh::zp::ZNB;
    const int64_t ld = (n + 1) & ~int64_t(1);
    const int64_t nt = (n + skh::TB - 1) / skh::TB + 1;
    const int64_t ntp = (nt + 1) & ~int64_t(1);
    const int64_t nts = (n + skh::MR - 1) / skh::MR + 1;
    const int64_t nseg = grid + 2;
    const int64_t gp = (grid + 1) & ~int64_t(1);
    // (W | V2) x 2 | (Pnorm | PS | X | Pq) x 4 | (Tv | Pmu | Prow | Pcol) x 3 | Dfin, in doubles; the two
    // W | V2 pairs alternate by panel parity, so each rank-2k launch (reading its panel's W) can fill
    // the next panel's slots
    const int64_t wpair = 2 * 2 * ld * nb;
    const int64_t ostride = ntp + 2 * nt * 2 * nb + 2 * ld + 2 * ntp;
    const int64_t hstride = 2 * 2 * nb + gp + 2 * nseg * ld + 2 * nts * ld;
    const size_t doubles = (size_t)(2 * wpair + skh::OWN_COPIES * ostride + skh::HAND_COPIES * hstride + 2 * ld + 64);
    double* ws = (double*)g_zbws.get(doubles * sizeof(double));
    if (!ws) return set_err(SKH_ERR_HIP, "hipMalloc of %zu bytes of workspace failed", doubles * 8);
    skh::zp::ZPWork w;
    w.W = (double2*)ws;
    w.V2 = w.W + ld * nb;
    w.ldw = ld;
    double* own0 = ws + 2 * wpair;
    w.Pnorm = own0;
    w.PS = (double2*)(own0 + ntp);
    w.X = (double2*)(own0 + ntp + 2 * nt * 2 * nb);
    w.Pq = (double2*)(own0 + ntp + 2 * nt * 2 * nb + 2 * ld);
    w.pn_stride = ostride;
    w.ps_stride = ostride / 2;
    w.x_stride = ostride / 2;
    w.pq_stride = ostride / 2;
    double* hand0 = own0 + skh::OWN_COPIES * ostride;
    w.Tv = (double2*)hand0;
    w.Pmu = hand0 + 2 * 2 * nb;
    w.Prow = (double2*)(hand0 + 2 * 2 * nb + gp);
    w.Pcol = w.Prow + nseg * ld;
    w.ldp = ld;
    w.tv_stride = hstride / 2;
    w.mu_stride = hstride;
    w.pr_stride = hstride / 2;
    w.pc_stride = hstride / 2;
    w.Dfin = (double2*)(hand0 + skh::HAND_COPIES * hstride);
    auto zgrid = [&](int64_t c0) {
        const skh::TallGrid tg = skh::tall_grid(n, c0 + 1);
        const int64_t blocks = (n + skh::TB - 1) / skh::TB - c0 / skh::TB;
        return (int)std::min<int64_t>(grid, std::max<int64_t>(std::max<int64_t>(8, blocks), tg.T));
    };
    auto zpanel_w = [&](int64_t c0) {
        skh::zp::ZPWork pw = w;
        if ((c0 / nb) & 1) {
            pw.W = (double2*)(ws + wpair);
            pw.V2 = pw.W + ld * nb;
        }
        return pw;
    };
    // the slots polled before any producer of the panel's launch refills them start empty
    auto zpanel_fill = [&](int64_t c0, int nbp, int g, const skh::zp::ZPWork& pw, FillSegs& fs) {
        const skh::TallGrid tg = skh::tall_grid(n, c0 + 1);
        int64_t maxslot = 1;
        for (int64_t I = tg.I0; I < tg.nts; I++) maxslot = std::max(maxslot, skh::tall_slots(tg, I, g));
        int ns = 0;
        auto seg = [&](double* p, int64_t rows, int64_t ldr, int64_t lo, int64_t hi) {
            fs.p[ns] = p;
            fs.rows[ns] = rows;
            fs.ld[ns] = ldr;
            fs.lo[ns] = lo;
            fs.hi[ns] = hi;
            ns++;
        };
        for (int64_t cc = c0; cc < c0 + 2; cc++) {
            double* h = hand0 + (cc % skh::HAND_COPIES) * hstride;
            seg(h, 1, 0, 0, 2 * 2 * nb + gp);                                   // t and m partials
            seg(h + 2 * 2 * nb + gp, maxslot, 2 * ld, 2 * c0, 2 * n);           // row partials
            seg(h + 2 * 2 * nb + gp + 2 * nseg * ld, tg.nts, 2 * ld, 2 * c0, 2 * n);   // column partials
            seg(own0 + (cc & (skh::OWN_COPIES - 1)) * ostride, 1, 0, 0, ostride);
        }
        seg(hand0 + ((c0 + 2) % skh::HAND_COPIES) * hstride, 1, 0, 0, 2 * 2 * nb);
        seg((double*)pw.W, nbp, 2 * ld, 2 * c0, 2 * n);
        seg((double*)pw.V2, nbp, 2 * ld, 2 * c0, 2 * n);
        return ns;
    };
    for (int64_t c0 = 0; c0 < kb_end; c0 += nb) {
        const int nbp = (int)std::min<int64_t>(nb, kb_end - c0);
        const int g = zgrid(c0);
        const skh::zp::ZPWork pw = zpanel_w(c0);
        if (c0 == 0) {   // later panels' slots are filled by the rank-2k launch before them
            FillSegs fs;
            const int ns = zpanel_fill(c0, nbp, g, pw, fs);
            HIPCHK((hipError_t)skh_internal_fill_empty(&fs, ns, st));
        }
        HIPCHK((hipError_t)skh_internal_launch_zpanel(dA, lda, n, c0, nbp, &pw, dtau, deta, ctl.err, g, st));
        FillSegs fnext;
        int nsnext = 0;
        const int64_t c1 = c0 + nb;
        if (c1 < kb_end) {
            const int nbp1 = (int)std::min<int64_t>(nb, kb_end - c1);
            nsnext = zpanel_fill(c1, nbp1, zgrid(c1), zpanel_w(c1), fnext);
        }
        HIPCHK((hipError_t)skh_internal_launch_zrank2k(dA, lda, pw.W, pw.ldw, n, c0 + nbp, c0, nbp, nsnext ? &fnext : nullptr,
                                                       nsnext, st));
    }
    // the panel columns' final diagonal entries into A (no later launch reads them)
    HIPCHK((hipError_t)skh_internal_launch_zdfin(dA, lda, w.Dfin, 0, (int)kb_end, st));
    return 0;
}

// engine: SKH_ENGINE_TEAM1 keeps the whole matrix on the unblocked engine, SKH_ENGINE_LARGE asks
// for the blocked panels, AUTO takes them from n = ZBLOCKED_MIN_N
int zsytrd_device(int64_t n, double2* dA, int64_t lda, double2* dtau, double* deta, double* ddvim, hipStream_t st,
                  int engine = SKH_ENGINE_AUTO)
{
    if (n <= 0) return 0;
    Range rg("skh zsytrd n=%lld", (long long)n);
    if (n >= 2) {
        Ctl ctl;
        int rc = get_ctl(ctl);
        if (rc) return rc;
        HIPCHK((hipError_t)skh_internal_zero_u32(ctl.bar, 512, st));
        // blocked panels first (one workgroup per CU, one 64-row block each)
        int64_t kb = 0;
        if (engine != SKH_ENGINE_TEAM1 && (engine == SKH_ENGINE_LARGE || n >= ZBLOCKED_MIN_N)) {
            int zpc = 0;
            HIPCHK((hipError_t)skh_internal_zpanel_occupancy(&zpc));
            const int zgrid = std::max(1, std::min(zpc, 1)) * g_num_cu;
            const int64_t blocks = (n + skh::TB - 1) / skh::TB;
            if (zpc >= 1 && blocks <= zgrid && n > ZTAIL + skh::zp::ZNB) {
                kb = ((n - ZTAIL) / skh::zp::ZNB) * skh::zp::ZNB;
                rc = zsytrd_blocked(n, dA, lda, dtau, deta, kb, zgrid, ctl, st);
                if (rc) return rc;
            }
        }
        const int64_t ldp = (n + 1) & ~int64_t(1);
        const int64_t nt = (n + skh::TB - 1) / skh::TB + 1;
        int per_cu = 0;
        HIPCHK((hipError_t)skh_internal_zfused_occupancy(&per_cu));
        per_cu = std::max(1, std::min(per_cu, 2));
        const int64_t ns0 = (n - kb + skh::TB - 1) / skh::TB + 1;
        const int64_t tiles0 = ns0 * (ns0 + 1) / 2;
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)per_cu * g_num_cu, tiles0));
        const int64_t gpad = (grid + 1) & ~1;
        const size_t doubles = (size_t)(4 * nt * ldp + 2 * ldp + 2 * gpad + 2 + 64);
        double* ws = (double*)g_zws.get(doubles * sizeof(double));
        if (!ws) return set_err(SKH_ERR_HIP, "hipMalloc of %zu bytes of workspace failed", doubles * 8);
        skh::zc::ZWork w;
        w.Prow = (double2*)ws;
        w.Pcol = w.Prow + nt * ldp;
        w.ldp = ldp;
        w.Wv = w.Pcol + nt * ldp;
        w.Pnrm = (double*)(w.Wv + ldp);
        w.Pmu = w.Pnrm + gpad;
        w.alpha = (double2*)(w.Pmu + gpad);
        // after the blocked panels nothing is pending: w of the "previous reflector" is zero
        if (kb > 0) HIPCHK((hipError_t)skh_internal_zero_u32((unsigned*)w.Wv, (int)(ldp * 4), st));
        // one launch per segment of columns whose trailing block starts in the same 64-row block
        // (the tile grid, and so each workgroup's tiles, stay fixed within a launch)
        for (int64_t k0 = kb; k0 <= n - 2;) {
            const int64_t ke = std::min<int64_t>(n - 1, ((k0 + 1) / skh::TB + 1) * skh::TB - 1);
            HIPCHK((hipError_t)skh_internal_launch_zfused(dA, lda, n, &w, dtau, deta, ctl.bar, ctl.err, grid, k0, ke,
                                                          (k0 == kb && kb > 0) ? 1 : 0, st));
            HIPCHK((hipError_t)skh_internal_zero_u32(ctl.bar, 288, st));
            k0 = ke;
        }
    }
    HIPCHK((hipError_t)skh_internal_zdiag(dA, lda, n, ddvim, st));
    return 0;
}

// Q (n x n, ldq) = H_0 ... H_{n-2} from complex factors and tau (device), blocked compact WY.
int zorgtr_locked(int64_t n, const double2* dF, int64_t ldf, const double2* dtau, double2* dQ, int64_t ldq,
                  hipStream_t st)
{
    if (n <= 0) return 0;
    Range rg("skh zorgtr n=%lld", (long long)n);
    if (n > 2147483647LL / 2) return set_err(SKH_ERR_UNSUPPORTED, "n too large for the BLAS interface");
    HIPCHK((hipError_t)skh_internal_zeye(dQ, ldq, n, st));
    if (n <= 1) return 0;
    int rc = blas_on(st);
    if (rc) return rc;
    const int64_t nb = 64;
    const int64_t nref = n - 1;
    const size_t cplx = (size_t)(n * nb + 2 * nb * n + 2 * nb * nb + 64);
    double2* ws = (double2*)g_ws.get(cplx * sizeof(double2));
    if (!ws) return set_err(SKH_ERR_HIP, "hipMalloc of orgtr workspace failed");
    double2* V = ws;
    double2* Y = V + n * nb;
    double2* Z = Y + nb * n;
    double2* Gm = Z + nb * n;
    double2* T = Gm + nb * nb;
    const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0}, mone = {-1.0, 0.0};
    auto rz = [](double2* p) { return (rocblas_double_complex*)p; };
    auto rzc = [](const double2* p) { return (const rocblas_double_complex*)p; };
    const int64_t last = ((nref - 1) / nb) * nb;
    for (int64_t c0 = last; c0 >= 0; c0 -= nb) {
        const int nbp = (int)std::min<int64_t>(nb, nref - c0);
        const int64_t m = n - c0 - 1;
        HIPCHK((hipError_t)skh_internal_zform_v(dF, ldf, n, c0, nbp, V, m, st));
        BLASCHK(rocblas_zgemm(g_blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, nbp, nbp, (int)m,
                              &one, rz(V), (int)m, rz(V), (int)m, &zero, rz(Gm), nbp));
        HIPCHK((hipError_t)skh_internal_zlarft(Gm, nbp, dtau + c0, nbp, T, nbp, st));
        double2* Q22 = dQ + (c0 + 1) + (c0 + 1) * ldq;
        BLASCHK(rocblas_zgemm(g_blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, nbp, (int)m,
                              (int)m, &one, rzc(V), (int)m, rzc(Q22), (int)ldq, &zero, rz(Y), nbp));
        BLASCHK(rocblas_zgemm(g_blas, rocblas_operation_none, rocblas_operation_none, nbp, (int)m, nbp, &one,
                              rzc(T), nbp, rzc(Y), nbp, &zero, rz(Z), nbp));
        BLASCHK(rocblas_zgemm(g_blas, rocblas_operation_none, rocblas_operation_none, (int)m, (int)m, nbp, &mone,
                              rzc(V), (int)m, rzc(Z), nbp, &one, rz(Q22), (int)ldq));
    }
    return 0;
}

// Column blocks between pageable host memory (leading dimension lda) and the device (leading dimension
// ldd), in bytes, as one contiguous copy: direct when the leading dimensions agree, else packed or
// unpacked through a host staging vector.  (hipMemcpy2DAsync from pageable memory delivered wrong data
// to the kernels here when run under rocprofv3's kernel tracing; contiguous copies did not.)
// upload_cols leaves the copy in flight: `stage` must outlive it (the caller synchronizes first).
hipError_t upload_cols(void* d, size_t ldd, const void* h, size_t lda, size_t rows, int64_t ncols,
                       std::vector<char>& stage, hipStream_t st)
{
    if (ncols <= 0) return hipSuccess;
    if (lda == ldd) return hipMemcpyAsync(d, h, ldd * (ncols - 1) + rows, hipMemcpyHostToDevice, st);
    stage.assign(ldd * ncols, 0);
    for (int64_t c = 0; c < ncols; c++) std::memcpy(stage.data() + c * ldd, (const char*)h + c * lda, rows);
    return hipMemcpyAsync(d, stage.data(), ldd * ncols, hipMemcpyHostToDevice, st);
}

// the reverse; complete on return (it synchronizes the stream)
hipError_t download_cols(void* h, size_t lda, const void* d, size_t ldd, size_t rows, int64_t ncols, hipStream_t st)
{
    if (ncols <= 0) return hipSuccess;
    if (lda == ldd) {
        const hipError_t e = hipMemcpyAsync(h, d, ldd * (ncols - 1) + rows, hipMemcpyDeviceToHost, st);
        return e != hipSuccess ? e : hipStreamSynchronize(st);
    }
    std::vector<char> stage(ldd * ncols);
    hipError_t e = hipMemcpyAsync(stage.data(), d, ldd * ncols, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    for (int64_t c = 0; c < ncols; c++) std::memcpy((char*)h + c * lda, stage.data() + c * ldd, rows);
    return hipSuccess;
}

// host copy-in of an n x n complex matrix (interleaved pairs, lda) into a device buffer with ld n
double2* zupload(int64_t n, const double* A, int64_t lda, hipStream_t st, int* rc)
{
    double2* dA = (double2*)g_mat.get((size_t)n * n * sizeof(double2));
    if (!dA) {
        *rc = set_err(SKH_ERR_HIP, "hipMalloc failed for n = %lld", (long long)n);
        return nullptr;
    }
    std::vector<char> stage;
    hipError_t e = upload_cols(dA, n * sizeof(double2), A, lda * sizeof(double2), n * sizeof(double2), n, stage, st);
    if (e == hipSuccess && !stage.empty()) e = hipStreamSynchronize(st);   // stage ends here
    if (e != hipSuccess) {
        *rc = set_err(SKH_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
        return nullptr;
    }
    *rc = 0;
    return dA;
}

// ----------------------------------------------------------------------------------------
// Gradient of log|Pf|: grad_b = -1/2 A_b^{-1} = -1/2 (K - K^T), K = Q_E C^{-1} Q_O^T, grouped by size
// (kernels in skh_grad.hip).  Everything the device work needs (control arrays, tau | eta, the
// workspaces of the largest size group) is prepared once in a GradPlan, so an evaluation only
// enqueues kernels: no host copy, synchronisation, allocation or memset (capturable in a graph).
// The gradient buffer first receives the factors from the forward kernel (same offsets as the
// input), then the gradient.
// ----------------------------------------------------------------------------------------
constexpr int64_t GRAD_NB = 32;   // reflectors per compact-WY block of the Q formation

struct GradChunk {
    int64_t n;     // size of every member
    int cnt;       // members
    int64_t at;    // goff[cnt] | toff[cnt] at dgrp + at
};

struct GradPlan {
    int64_t batch = 0;
    std::vector<int64_t> ns, offs;   // host copies: sizes and element offsets (input and gradient)
    std::vector<int64_t> big;        // even n > CMP_NMAX: forward by the large engine, one at a time
    std::vector<int64_t> toffs;      // tau | eta offsets (2 max(n, 1) doubles per member)
    int64_t count = 0;               // members of the compact launch
    int grid = 0, kmax = 512;
    int knb = 0;                     // blocked engine panel width for the forward (0: unblocked engine)
    double* wbuf = nullptr;          // its W workspaces (kmax * knb doubles per workgroup)
    std::vector<GradChunk> chunks;
    int64_t nodd = 0, odd_at = 0;    // odd members: goff[nodd] | n[nodd] at dgrp + odd_at
    int64_t* dctl = nullptr;         // ns | offs | items | toffs of the compact launch
    int64_t* dgrp = nullptr;         // per chunk goff | toff, then the odd list
    unsigned* ctl = nullptr;         // [0] queue, [32] error flag (256 bytes)
    double* dte = nullptr;           // tau | eta of every member
    double* ws = nullptr;            // Q | K | Zt | Y | Z | V | Gm | T of the largest chunk
    double flops = 0.0;              // forward (4/3) n^3 + backward (7/3) n^3 per even member
};

void grad_plan_free(GradPlan* p)
{
    if (!p) return;
    for (void* q : {(void*)p->dctl, (void*)p->dgrp, (void*)p->ctl, (void*)p->dte, (void*)p->ws, (void*)p->wbuf})
        if (q) (void)hipFree(q);
    delete p;
}

// doubles of one chunk's workspaces
size_t grad_ws_doubles(int64_t n, int64_t cnt)
{
    const int64_t nn = n * n, h = n / 2, NB = GRAD_NB;
    return (size_t)cnt * (size_t)(2 * nn + n * h + 2 * NB * n + n * NB + 2 * NB * NB);
}

// allow_big: members with even n > CMP_NMAX are factored by the large engine (not capturable)
int grad_plan_build(int64_t batch, const int64_t* n, const int64_t* offset, bool allow_big, GradPlan** out)
{
    *out = nullptr;
    GradPlan* p = new GradPlan();
    p->batch = batch;
    p->ns.assign(n, n + batch);
    p->offs.assign(offset, offset + batch);
    p->toffs.resize(batch);
    int64_t tt = 0;
    for (int64_t b = 0; b < batch; b++) {
        p->toffs[b] = tt;
        tt += 2 * std::max<int64_t>(n[b], 1);
        if (n[b] % 2 == 0 && n[b] > 0)
            p->flops += (11.0 / 3.0) * (double)n[b] * (double)n[b] * (double)n[b];
    }
    std::vector<int64_t> small;
    for (int64_t b = 0; b < batch; b++) {
        if (n[b] > CMP_NMAX && n[b] % 2 == 0) {
            if (!allow_big) {
                grad_plan_free(p);
                return set_err(SKH_ERR_UNSUPPORTED, "plans take even n <= %lld (n[%lld] = %lld)", (long long)CMP_NMAX,
                               (long long)b, (long long)n[b]);
            }
            p->big.push_back(b);
        } else {
            small.push_back(b);
        }
    }
    std::stable_sort(small.begin(), small.end(), [&](int64_t a, int64_t b) { return n[a] > n[b]; });
    p->count = (int64_t)small.size();
    int64_t nmax = 2;
    for (int64_t b : small) nmax = std::max(nmax, n[b]);
    p->kmax = nmax <= 512 ? 512 : 1024;
    {   // the forward on the blocked engine (panel width 8) when every member fits it
        int bk = 0, bnb = 0;
        if (nmax <= BLK_NMAX && blocked_instance(nmax, 0, &bk, &bnb)) {
            p->kmax = bk;
            p->knb = bnb;
        }
    }
    int per_cu = 0;
    const int occ = p->knb ? skh_internal_cblock_occupancy(p->kmax, p->knb, &per_cu)
                           : skh_internal_compact_occupancy(p->kmax, &per_cu);
    if (occ != 0) {
        grad_plan_free(p);
        return set_err(SKH_ERR_HIP, "occupancy query failed");
    }
    p->grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)std::max(1, per_cu) * g_num_cu, p->count));
    // size groups (ascending n), chunked so that Q and K stay below about 1.5 GB each
    std::vector<int64_t> idx;
    std::vector<int64_t> odd;
    for (int64_t b = 0; b < batch; b++) {
        if (n[b] % 2 == 1) odd.push_back(b);
        else if (n[b] > 0) idx.push_back(b);
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return n[a] < n[b]; });
    std::vector<int64_t> g;   // goff | toff per chunk, then the odd list
    size_t wsmax = 0;
    for (size_t gi = 0; gi < idx.size();) {
        const int64_t m = n[idx[gi]];
        size_t gj = gi;
        while (gj < idx.size() && n[idx[gj]] == m) gj++;
        const int64_t per = std::max<int64_t>(1, std::min<int64_t>(4096, (int64_t)(1.5e9 / (8.0 * m * m))));
        for (size_t q0 = gi; q0 < gj; q0 += per) {
            const int cnt = (int)std::min<int64_t>(per, (int64_t)(gj - q0));
            p->chunks.push_back(GradChunk{m, cnt, (int64_t)g.size()});
            for (int i = 0; i < cnt; i++) g.push_back(offset[idx[q0 + i]]);
            for (int i = 0; i < cnt; i++) g.push_back(p->toffs[idx[q0 + i]]);
            wsmax = std::max(wsmax, grad_ws_doubles(m, cnt));
        }
        gi = gj;
    }
    p->nodd = (int64_t)odd.size();
    p->odd_at = (int64_t)g.size();
    for (int64_t b : odd) g.push_back(offset[b]);
    for (int64_t b : odd) g.push_back(n[b]);
    // compact control: ns | offs | items | toffs
    std::vector<int64_t> h(4 * batch);
    std::copy(n, n + batch, h.begin());
    std::copy(offset, offset + batch, h.begin() + batch);
    std::copy(small.begin(), small.end(), h.begin() + 2 * batch);
    std::copy(p->toffs.begin(), p->toffs.end(), h.begin() + 3 * batch);
    bool ok = hipMalloc((void**)&p->dctl, sizeof(int64_t) * h.size()) == hipSuccess &&
              hipMalloc((void**)&p->dgrp, sizeof(int64_t) * std::max<size_t>(g.size(), 1)) == hipSuccess &&
              hipMalloc((void**)&p->ctl, 256) == hipSuccess &&
              hipMalloc((void**)&p->dte, sizeof(double) * (size_t)(tt + 16)) == hipSuccess &&
              hipMalloc((void**)&p->ws, sizeof(double) * std::max<size_t>(wsmax, 1)) == hipSuccess &&
              (p->knb == 0 || hipMalloc((void**)&p->wbuf, sizeof(double) * (size_t)p->kmax * p->knb * p->grid) ==
                                  hipSuccess);
    if (!ok) {
        grad_plan_free(p);
        return set_err(SKH_ERR_HIP, "hipMalloc of the gradient plan's buffers failed (%zu doubles of workspace)", wsmax);
    }
    if (hipMemcpy(p->dctl, h.data(), sizeof(int64_t) * h.size(), hipMemcpyHostToDevice) != hipSuccess ||
        (!g.empty() && hipMemcpy(p->dgrp, g.data(), sizeof(int64_t) * g.size(), hipMemcpyHostToDevice) != hipSuccess)) {
        grad_plan_free(p);
        return set_err(SKH_ERR_HIP, "upload of the gradient plan's control arrays failed");
    }
    *out = p;
    return 0;
}

// Enqueue one evaluation on st.  Members with even n > CMP_NMAX (device entry only) are factored
// first by the large engine, in place in dgrad.
int grad_enqueue(const GradPlan& p, const double* dA, double* dlogpf, double* dsgn, double* dgrad, int grid,
                 hipStream_t st)
{
    for (int64_t b : p.big) {
        const int64_t n = p.ns[b];
        double* M = dgrad + p.offs[b];
        double* te = p.dte + p.toffs[b];
        HIPCHK(hipMemcpyAsync(M, dA + p.offs[b], (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, st));
        g_timer.reset();
        g_timer.on = false;
        g_timer.phases = false;
        int rc = sytrd_large(n, M, n, te, te + n, pick_nb(n, nullptr), st);
        if (rc) return rc;
        HIPCHK((hipError_t)skh_internal_logpf_from(te + n, te, n, dlogpf + b, dsgn + b, st));
    }
    if (p.count > 0) {
        HIPCHK((hipError_t)skh_internal_zero_u32(p.ctl, 64, st));
        CompactArgs a;
        a.A = dA;
        a.ns = p.dctl;
        a.offs = p.dctl + p.batch;
        a.items = p.dctl + 2 * p.batch;
        a.count = p.count;
        a.slots = nullptr;
        a.slot_stride = 0;
        a.queue = p.ctl;
        a.err = (int*)(p.ctl + 32);
        a.budget = (unsigned long long)(watchdog_seconds() * 1e8);
        a.logpf = dlogpf;
        a.sgn = dsgn;
        a.fac = dgrad;
        a.taueta = p.dte;
        a.toffs = p.dctl + 3 * p.batch;
        a.wbuf = p.wbuf;
        a.wstride = (int64_t)p.kmax * p.knb;
        if (p.knb > 0) HIPCHK((hipError_t)skh_internal_launch_cblock(&a, p.kmax, p.knb, grid, st));
        else HIPCHK((hipError_t)skh_internal_launch_compact(&a, p.kmax, grid, st));
    }
    if (p.nodd > 0)
        HIPCHK((hipError_t)skh_internal_grad_zero(dgrad, p.dgrp + p.odd_at, p.dgrp + p.odd_at + p.nodd, (int)p.nodd, st));
    const int64_t NB = GRAD_NB;
    for (const GradChunk& c : p.chunks) {
        const int64_t n = c.n, nn = n * n, h = n / 2;
        const int cnt = c.cnt;
        const int64_t* goff = p.dgrp + c.at;
        const int64_t* toff = goff + cnt;
        double* Q = p.ws;
        double* K = Q + (int64_t)cnt * nn;
        double* Zt = K + (int64_t)cnt * nn;
        double* Y = Zt + (int64_t)cnt * n * h;
        double* Z = Y + (int64_t)cnt * NB * n;
        double* V = Z + (int64_t)cnt * NB * n;
        double* Gm = V + (int64_t)cnt * n * NB;
        double* T = Gm + (int64_t)cnt * NB * NB;
        const int64_t vstride = n * NB, gstride = NB * NB, ystride = NB * n;
        HIPCHK((hipError_t)skh_internal_grad_eye(Q, nn, n, cnt, st));
        const int64_t nref = n - 1;
        for (int64_t c0 = ((nref - 1) / NB) * NB; c0 >= 0; c0 -= NB) {
            const int nbp = (int)std::min<int64_t>(NB, nref - c0);
            const int64_t m = n - c0 - 1;
            double* Q22 = Q + (c0 + 1) + (c0 + 1) * n;
            HIPCHK((hipError_t)skh_internal_grad_form_v(dgrad, goff, n, c0, nbp, V, vstride, cnt, st));
            HIPCHK((hipError_t)skh_internal_bgemm(1, 0, nbp, nbp, m, 1.0, V, m, vstride, V, m, vstride, 0.0, Gm, nbp,
                                                  gstride, cnt, st));
            HIPCHK((hipError_t)skh_internal_larft_batched(Gm, gstride, p.dte, toff, c0, nbp, T, gstride, cnt, st));
            HIPCHK((hipError_t)skh_internal_bgemm(1, 0, nbp, m, m, 1.0, V, m, vstride, Q22, n, nn, 0.0, Y, nbp, ystride,
                                                  cnt, st));
            HIPCHK((hipError_t)skh_internal_bgemm(0, 0, nbp, m, nbp, 1.0, T, nbp, gstride, Y, nbp, ystride, 0.0, Z, nbp,
                                                  ystride, cnt, st));
            HIPCHK((hipError_t)skh_internal_bgemm(0, 0, m, m, nbp, -1.0, V, m, vstride, Z, nbp, ystride, 1.0, Q22, n, nn,
                                                  cnt, st));
        }
        // Zt = (C^{-1} Q_O^T)^T, K = Q_E Zt^T, grad = -1/2 (K - K^T)
        HIPCHK((hipError_t)skh_internal_grad_bidiag(Q, nn, p.dte, toff, n, Zt, n * h, cnt, st));
        HIPCHK((hipError_t)skh_internal_bgemm(0, 1, n, n, h, 1.0, Q + n, 2 * n, nn, Zt, n, n * h, 0.0, K, n, nn, cnt,
                                              st));
        HIPCHK((hipError_t)skh_internal_grad_skew_out(K, nn, dgrad, goff, n, -0.5, cnt, st));
    }
    return 0;
}

// the device entry's plan, rebuilt when the batch's sizes or offsets change
GradPlan* g_gradplan = nullptr;

int grad_run(int64_t batch, const int64_t* ns, const int64_t* offs, const double* dA, double* dlogpf, double* dsgn,
             double* dgrad, const skh_opts* opts, hipStream_t st)
{
    (void)opts;
    if (batch == 0) return 0;
    Range rg("skh logpf grad batched %lld matrices", (long long)batch);
    GradPlan* p = g_gradplan;
    if (!p || p->batch != batch || !std::equal(ns, ns + batch, p->ns.begin()) ||
        !std::equal(offs, offs + batch, p->offs.begin())) {
        if (p) {
            HIPCHK(hipStreamSynchronize(st));   // the previous plan's work may still be running on st
            grad_plan_free(p);
            g_gradplan = nullptr;
        }
        int rc = grad_plan_build(batch, ns, offs, true, &p);
        if (rc) return rc;
        g_gradplan = p;
    }
    int rc = grad_enqueue(*p, dA, dlogpf, dsgn, dgrad, p->grid, st);
    if (rc) return rc;
    if (!p->big.empty()) {
        HIPCHK(hipStreamSynchronize(st));
        rc = check_barrier_flag();
        if (rc) return rc;
    }
    return 0;
}

// the error words of a control block (flag, inputs missing, index + 1) as a status
int barrier_status(const int hv[3])
{
    const int herr = hv[0];
    if (herr) {
        if ((herr & 1) && (herr >> 1) != 0) {
            static const char* site[] = {"?", "row pass", "phase B", "phase B column entry"};
            return set_err(SKH_ERR_BARRIER, "a bounded wait of the panel kernel ran out (%s, column %d mod 65536, "
                           "workgroup %d; inputs missing 0x%x, index %d); if other work shares the GPU, "
                           "SKH_PANEL_COOPERATIVE=1 keeps the panel's workgroups co-resident", site[(herr >> 1) & 3],
                           (herr >> 3) & 0xFFFF, (herr >> 19) & 0x3FF, hv[1], hv[2] - 1);
        }
        return set_err(SKH_ERR_BARRIER, "a bounded wait of a cooperative kernel ran out (lost workgroup)");
    }
    return 0;
}

int check_barrier_flag()
{
    Ctl ctl;
    int rc = get_ctl(ctl);
    if (rc) return rc;
    int hv[3] = {0, 0, 0};
    HIPCHK(hipMemcpy(hv, ctl.err, sizeof(hv), hipMemcpyDeviceToHost));
    return barrier_status(hv);
}

int sytrd_device_locked(int64_t n, double* dA, int64_t lda, double* dtau, double* deta, hipStream_t st,
                        const skh_opts* opts)
{
    if (n <= 1) return 0;
    const int64_t nb = pick_nb(n, opts);
    g_timer.reset();
    g_timer.on = opts && (opts->flags & SKH_FLAG_TIMING);
    g_timer.phases = opts && (opts->flags & SKH_FLAG_PHASES);
    return sytrd_large(n, dA, lda, dtau, deta, nb, st);
}

}  // namespace


// ----------------------------------------------------------------------------------------
// Batched log|Pf| plans: sizes fixed at creation; exec enqueues without allocating,
// synchronising or copying from the host, so it can be captured into a hipGraph.
// ----------------------------------------------------------------------------------------
// grids of up to this many workgroups per CU are timed by the plan tuner
constexpr int PLAN_TUNE_PER_CU = 4;
// W workspace of one workgroup of the blocked engine, enough for every instance (NMAX x NB)
constexpr int64_t PLAN_WSTRIDE = 512 * 16;

struct skh_plan_s {
    int64_t batch = 0;
    int64_t count = 0;
    int64_t nmax = 0;
    int kmax = 512;
    int grid = 0;
    int64_t slot = 0;
    int64_t* dctl = nullptr;   // ns | offs | items
    double* slots = nullptr;
    unsigned* ctl = nullptr;   // [0] queue, [32] error flag
    double flops = 0.0;
    int tuned = 0;
    GradPlan* g = nullptr;     // gradient plans: everything above except grid / kmax / count lives here
    int knb = 0;               // blocked engine panel width (0: the unblocked engine)
    double* wbuf = nullptr;    // blocked engine: W workspace per workgroup (PLAN_WSTRIDE doubles each)
};

namespace {

int plan_enqueue(skh_plan p, const double* dA, double* dlogpf, double* dsgn, hipStream_t st, int grid, int kmax,
                 int knb)
{
    Range rg("skh logpf plan %lld matrices", (long long)p->batch);
    HIPCHK((hipError_t)skh_internal_zero_u32(p->ctl, 64, st));
    CompactArgs a;
    a.A = dA;
    a.ns = p->dctl;
    a.offs = p->dctl + p->batch;
    a.items = p->dctl + 2 * p->batch;
    a.count = p->count;
    a.slots = p->slots;
    a.slot_stride = p->slot;
    a.queue = p->ctl;
    a.err = (int*)(p->ctl + 32);
    a.budget = (unsigned long long)(watchdog_seconds() * 1e8);
    a.logpf = dlogpf;
    a.sgn = dsgn;
    a.fac = nullptr;
    a.taueta = nullptr;
    a.toffs = nullptr;
    a.wbuf = p->wbuf;
    a.wstride = PLAN_WSTRIDE;
    if (knb > 0) HIPCHK((hipError_t)skh_internal_launch_cblock(&a, kmax, knb, grid, st));
    else HIPCHK((hipError_t)skh_internal_launch_compact(&a, kmax, grid, st));
    return 0;
}

void plan_free(skh_plan p)
{
    if (!p) return;
    if (p->g) {
        grad_plan_free(p->g);
        delete p;
        return;
    }
    if (p->dctl) (void)hipFree(p->dctl);
    if (p->wbuf) (void)hipFree(p->wbuf);
    if (p->slots) (void)hipFree(p->slots);
    if (p->ctl) (void)hipFree(p->ctl);
    delete p;
}

// ---- complex skew-symmetric Pfaffians (Parlett-Reid, one workgroup per matrix) ----
Buf g_zpslot, g_zpctl, g_zpin, g_zpout;
constexpr size_t ZPF_SLOT_BUDGET = size_t(4) << 30;   // bytes of working copies in flight
constexpr int64_t ZPF_COOP_MIN = 512;                  // even members this large use the whole GPU
constexpr int64_t ZPF_ONE_WG_MAX = 2048;               // largest n of the one-workgroup kernel
constexpr int64_t ZPF_NMAX = 8192;                     // largest n accepted

int zpf_check(int64_t batch, const int64_t* n, const int64_t* offset, const void* A, const void* o1, const void* o2,
              bool dev)
{
    if (batch < 0) return set_err(-1, "batch < 0");
    if (batch > 0 && (!n || !offset || !A || !o1 || !o2)) return set_err(-2, "NULL argument");
    const int64_t nmax = ZPF_NMAX;
    for (int64_t b = 0; b < batch; b++) {
        if (n[b] < 0) return set_err(-2, "n[%lld] < 0", (long long)b);
        if (n[b] > nmax) return set_err(-2, "n[%lld] = %lld exceeds %lld", (long long)b, (long long)n[b], (long long)nmax);
        if (offset[b] < 0) return set_err(-3, "offset[%lld] < 0", (long long)b);
    }
    if (dev && batch > 0 && (((uintptr_t)A) & 15)) return set_err(-4, "A must be 16-byte aligned");
    return 0;
}

// dA: matrices at complex offsets offs[b]; mode 0: Pf into do1 (2 doubles each); mode 1: log|Pf|
// into do1 and cis(phase) into do2.  Host n/offs; enqueues on st and waits for the control copy.
int zpf_run(int64_t batch, const int64_t* ns, const int64_t* offs, const double2* dA, int mode, double* do1,
            double* do2, hipStream_t st)
{
    if (batch == 0) return 0;
    std::vector<int64_t> order;
    std::vector<int64_t> big;
    for (int64_t b = 0; b < batch; b++) {
        if (ns[b] % 2 == 0 && ns[b] >= ZPF_COOP_MIN) big.push_back(b);
        else if (ns[b] % 2 == 1 && ns[b] > ZPF_ONE_WG_MAX) {   // odd: Pf = 0 (log: -Inf, cis 0)
            static const double zero2[2] = {0.0, 0.0}, minf = -INFINITY;
            HIPCHK(hipMemcpyAsync(do1 + (mode == 0 ? 2 * b : b), mode == 0 ? zero2 : &minf,
                                  (mode == 0 ? 2 : 1) * sizeof(double), hipMemcpyHostToDevice, st));
            if (mode == 1) HIPCHK(hipMemcpyAsync(do2 + 2 * b, zero2, 2 * sizeof(double), hipMemcpyHostToDevice, st));
        } else {
            order.push_back(b);
        }
    }
    // large even members: the large engine over the whole GPU, one after another
    if (!big.empty()) {
        Ctl ctl;
        int rc = get_ctl(ctl);
        if (rc) return rc;
        int per_cu = 0;
        HIPCHK((hipError_t)skh_internal_zpfaff_coop_occupancy(&per_cu));
        per_cu = std::max(1, std::min(per_cu, 2));
        for (int64_t b : big) {
            const int64_t n = ns[b], nt = (n + 63) / 64;
            const int grid = (int)std::min<int64_t>((int64_t)per_cu * g_num_cu, nt * (nt + 1) / 2);
            double2* work = (double2*)g_zpslot.get((size_t)(n * n + 3 * n + grid) * sizeof(double2));
            if (!work) return set_err(SKH_ERR_HIP, "hipMalloc of the Pfaffian work copy failed");
            HIPCHK((hipError_t)skh_internal_zero_u32(ctl.bar, 512, st));
            HIPCHK((hipError_t)skh_internal_zpfaff_coop(dA + offs[b], work, n, mode, do1 + (mode == 0 ? 2 * b : b),
                                                        mode == 1 ? do2 + 2 * b : nullptr, work + n * n,
                                                        (double*)(work + n * n + 3 * n), ctl.bar, ctl.err, grid, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        rc = check_barrier_flag();
        if (rc) return rc;
    }
    if (order.empty()) {
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return ns[a] > ns[b]; });
    const int64_t nmax = std::max<int64_t>(ns[order[0]], 2);
    const int64_t count = (int64_t)order.size();
    int per_cu = 0;
    HIPCHK((hipError_t)skh_internal_zpfaff_occupancy(nmax, &per_cu));
    // two matrices per CU once members exceed n = 128: fewer working copies in flight per XCD
    // (4096 matrices n in {32..256}: 12.5 ms against 13.1 ms at the full occupancy of 3)
    per_cu = std::max(1, nmax > 128 ? std::min(per_cu, 2) : per_cu);
    const int64_t slot = (nmax * nmax + 15) & ~int64_t(15);   // complex elements
    const int64_t by_mem = std::max<int64_t>(1, (int64_t)(ZPF_SLOT_BUDGET / ((size_t)slot * sizeof(double2))));
    const int grid = (int)std::min<int64_t>({(int64_t)per_cu * g_num_cu, count, by_mem});
    double2* slots = (double2*)g_zpslot.get((size_t)slot * grid * sizeof(double2));
    int64_t* dctl = (int64_t*)g_zpctl.get((2 * batch + count) * sizeof(int64_t));
    if (!slots || !dctl) return set_err(SKH_ERR_HIP, "hipMalloc of Pfaffian work slots failed");
    std::vector<int64_t> h(2 * batch + count);
    StreamDrain drain(st);   // h is read by the upload below
    std::copy(ns, ns + batch, h.begin());
    std::copy(offs, offs + batch, h.begin() + batch);
    std::copy(order.begin(), order.end(), h.begin() + 2 * batch);
    HIPCHK(hipMemcpyAsync(dctl, h.data(), h.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK((hipError_t)skh_internal_zpfaff(dA, dctl, dctl + batch, dctl + 2 * batch, count, nmax, slots, slot, grid,
                                           mode, do1, do2, st));
    // the host copy of the control arrays must outlive the async copy
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int zpf_host(int64_t batch, const int64_t* n, const int64_t* offset, const double* A, int mode, double* o1, double* o2)
{
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = g_stream;
    StreamDrain drain(st);   // the caller's arrays are copied asynchronously
    std::vector<int64_t> offs2(batch);
    int64_t tot = 0;
    for (int64_t b = 0; b < batch; b++) {
        offs2[b] = tot;
        tot += n[b] * n[b];
    }
    double2* dA = (double2*)g_zpin.get((size_t)std::max<int64_t>(tot, 1) * sizeof(double2));
    double* dout = (double*)g_zpout.get((size_t)4 * batch * sizeof(double));
    if (!dA || !dout) return set_err(SKH_ERR_HIP, "hipMalloc of batched buffers failed");
    for (int64_t b = 0; b < batch; b++)
        HIPCHK(hipMemcpyAsync(dA + offs2[b], A + 2 * offset[b], (size_t)n[b] * n[b] * sizeof(double2),
                              hipMemcpyHostToDevice, st));
    rc = zpf_run(batch, n, offs2.data(), dA, mode, dout, dout + 2 * batch, st);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(o1, dout, (mode == 0 ? 2 : 1) * batch * sizeof(double), hipMemcpyDeviceToHost, st));
    if (mode == 1) HIPCHK(hipMemcpyAsync(o2, dout + 2 * batch, 2 * batch * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// ---- eigenvector rotations of the skew-tridiagonal QR on the GPU ----
Buf g_rotq, g_rotk, g_rotcs, g_rotseg;
Buf g_fvec, g_fpiv, g_fz;   // matrix functions: eigenvalue vectors, LU pivots, complex work matrices
Buf g_stq;                  // symmetric tridiagonal fallback: eigenvectors, diagonal, off-diagonal, info
constexpr int64_t ROT_GPU_MIN = 256;   // eigen(A) applies the QR rotations on the GPU from this n

std::vector<int64_t> chain_starts(const int32_t* k, int64_t cnt)
{
    std::vector<int64_t> s;
    for (int64_t t = 0; t < cnt; t++)
        if (t == 0 || k[t] != k[t - 1] + 1) s.push_back(t);
    s.push_back(cnt);
    return s;
}

// skh_rot_apply_fn: Qo, Qe (host, still the identity) receive the logged rotations, applied on the GPU
int rot_apply_gpu(void*, int64_t nrows, const int32_t* ko, const double* cso, int64_t nro, double* Qo, int64_t no,
                  const int32_t* ke, const double* cse, int64_t nre, double* Qe, int64_t ne)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = g_stream;
    const std::vector<int64_t> so = chain_starts(ko, nro), se = chain_starts(ke, nre);
    StreamDrain drain(st);   // so, se and the caller's arrays are copied asynchronously
    double* dq = (double*)g_rotq.get((size_t)(no * no + ne * ne + 1) * sizeof(double));
    int32_t* dk = (int32_t*)g_rotk.get((size_t)(nro + nre + 1) * sizeof(int32_t));
    double* dcs = (double*)g_rotcs.get((size_t)(2 * (nro + nre) + 1) * sizeof(double));
    int64_t* dseg = (int64_t*)g_rotseg.get((so.size() + se.size()) * sizeof(int64_t));
    if (!dq || !dk || !dcs || !dseg) return set_err(SKH_ERR_HIP, "hipMalloc of rotation buffers failed");
    double* dQo = dq;
    double* dQe = dq + no * no;
    HIPCHK(hipMemcpyAsync(dQo, Qo, (size_t)no * no * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dQe, Qe, (size_t)ne * ne * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dk, ko, (size_t)nro * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dk + nro, ke, (size_t)nre * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dcs, cso, (size_t)2 * nro * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dcs + 2 * nro, cse, (size_t)2 * nre * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dseg, so.data(), so.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dseg + so.size(), se.data(), se.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK((hipError_t)skh_internal_rot_chains(dQo, no, dk, dcs, dseg, (int64_t)so.size() - 1, dQe, ne, dk + nro,
                                               dcs + 2 * nro, dseg + so.size(), (int64_t)se.size() - 1, nrows, st));
    HIPCHK(hipMemcpyAsync(Qo, dQo, (size_t)no * no * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(Qe, dQe, (size_t)ne * ne * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// skh_symtri_fn: eigenpairs of SymTridiagonal(0, e) by rocSOLVER's divide and conquer on the GPU, the
// fallback of the eigenvector QR when its sweeps stagnate (the reference calls LAPACK there,
// src/skeweigen.jl:470-490); eigenvalues ascending, as LAPACK's
int symtri_gpu_diag(int64_t n, const double* diag, const double* e, double* d, double* Z);
int symtri_gpu(void*, int64_t n, const double* e, double* d, double* Z) { return symtri_gpu_diag(n, nullptr, e, d, Z); }

// eigenpairs of SymTridiagonal(diag (0 when NULL), e) on the GPU: d ascending, Z n x n
int symtri_gpu_diag(int64_t n, const double* diag, const double* e, double* d, double* Z)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    if (n > 46340) return set_err(SKH_ERR_UNSUPPORTED, "n too large for the solver interface");
    hipStream_t st = g_stream;
    double* dC = (double*)g_stq.get((size_t)(n * n + 2 * n + 2) * sizeof(double));
    if (!dC) return set_err(SKH_ERR_HIP, "hipMalloc failed (tridiagonal fallback)");
    double* dD = dC + n * n;
    double* dE = dD + n;
    int* dinfo = (int*)(dE + n);
    const std::vector<double> zeros(diag ? 0 : n, 0.0);
    StreamDrain drain(st);   // zeros and the caller's arrays are copied asynchronously
    HIPCHK(hipMemcpyAsync(dD, diag ? diag : zeros.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
    if (n > 1) HIPCHK(hipMemcpyAsync(dE, e, (n - 1) * sizeof(double), hipMemcpyHostToDevice, st));
    rc = blas_on(st);
    if (rc) return rc;
    BLASCHK(rocsolver_dstedc(g_blas, rocblas_evect_tridiagonal, (rocblas_int)n, dD, dE, dC, (rocblas_int)n, dinfo));
    int info = 0;
    HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(d, dD, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(Z, dC, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (info != 0) return set_err(SKH_ERR_NOCONV, "tridiagonal divide and conquer failed (info %d)", info);
    return 0;
}

// ---- skew-Cholesky: one workgroup per matrix, or the whole GPU for one large matrix ----
Buf g_chctl, g_chin, g_chout, g_chp, g_chws;
constexpr int64_t CHOL_COOP_MIN = 512;   // matrices this large get the cooperative kernel

int chol_check(int64_t batch, const int64_t* n, const int64_t* offset, const void* A, const void* R, const void* p,
               int pivot)
{
    if (batch < 0) return set_err(-1, "batch < 0");
    if (batch > 0 && (!n || !offset || !A || !R || !p)) return set_err(-2, "NULL argument");
    const int64_t nmax = skh_internal_skewchol_nmax();
    for (int64_t b = 0; b < batch; b++) {
        if (n[b] < 0) return set_err(-2, "n[%lld] < 0", (long long)b);
        if (n[b] > nmax) return set_err(-2, "n[%lld] = %lld exceeds %lld", (long long)b, (long long)n[b], (long long)nmax);
        if (offset[b] < 0) return set_err(-3, "offset[%lld] < 0", (long long)b);
    }
    if (pivot != 0 && pivot != 1) return set_err(-8, "pivot must be 0 (NoPivot) or 1 (RowMaximum)");
    return 0;
}

// device arrays; host n/offs.  p of matrix b at dp + (n[0] + ... + n[b-1]).
int chol_run(int64_t batch, const int64_t* ns, const int64_t* offs, const double* dA, double* dR, int64_t* dp,
             int64_t* dsteps, int pivot, hipStream_t st)
{
    if (batch == 0) return 0;
    std::vector<int64_t> h(4 * batch);
    StreamDrain drain(st);   // h is read by the upload below
    int64_t* order = h.data() + 3 * batch;
    int64_t pacc = 0;
    for (int64_t b = 0; b < batch; b++) {
        h[b] = ns[b];
        h[batch + b] = offs[b];
        h[2 * batch + b] = pacc;
        pacc += ns[b];
        order[b] = b;
    }
    std::stable_sort(order, order + batch, [&](int64_t a, int64_t b) { return ns[a] > ns[b]; });
    // large members: one cooperative launch each over the whole GPU
    int64_t nbig = 0;
    while (nbig < batch && ns[order[nbig]] >= CHOL_COOP_MIN) nbig++;
    if (nbig > 0) {
        Ctl ctl;
        int rc = get_ctl(ctl);
        if (rc) return rc;
        int per_cu = 0;
        HIPCHK((hipError_t)skh_internal_skewchol_coop_occupancy(&per_cu));
        per_cu = std::max(1, std::min(per_cu, 2));
        for (int64_t q = 0; q < nbig; q++) {
            const int64_t b = order[q], n = ns[b];
            const int64_t nt = (n + 63) / 64;
            const int grid = (int)std::min<int64_t>((int64_t)per_cu * g_num_cu, nt * (nt + 1) / 2);
            double* ws = (double*)g_chws.get((size_t)(4 * n + 3 * grid + 1) * sizeof(double));
            if (!ws) return set_err(SKH_ERR_HIP, "hipMalloc of skew-Cholesky workspace failed");
            int64_t* st1 = dsteps ? dsteps + b : (int64_t*)(ws + 4 * n + 3 * grid);
            HIPCHK((hipError_t)skh_internal_zero_u32(ctl.bar, 512, st));
            HIPCHK((hipError_t)skh_internal_skewchol_coop(dA + offs[b], dR + offs[b], n, pivot, dp + h[2 * batch + b],
                                                          st1, ws, ws + 4 * n, ctl.bar, ctl.err, grid, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        rc = check_barrier_flag();
        if (rc) return rc;
        if (nbig == batch) return 0;
    }
    const int64_t nmax = std::max<int64_t>(ns[order[nbig]], 2);
    int per_cu = 0;
    HIPCHK((hipError_t)skh_internal_skewchol_occupancy(nmax, &per_cu));
    per_cu = std::max(1, per_cu);
    const int grid = (int)std::min<int64_t>((int64_t)per_cu * g_num_cu, batch - nbig);
    int64_t* dctl = (int64_t*)g_chctl.get(h.size() * sizeof(int64_t));
    if (!dctl) return set_err(SKH_ERR_HIP, "hipMalloc of skew-Cholesky control arrays failed");
    HIPCHK(hipMemcpyAsync(dctl, h.data(), h.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK((hipError_t)skh_internal_skewchol(dA, dR, dctl, dctl + batch, dctl + 2 * batch, dctl + 3 * batch + nbig,
                                             batch - nbig, nmax, pivot, dp, dsteps, grid, st));
    // the host copy of the control arrays must outlive the async copy
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int chol_host(int64_t batch, const int64_t* n, const int64_t* offset, const double* A, double* R, int64_t* p,
              int64_t* steps, int pivot)
{
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = g_stream;
    StreamDrain drain(st);   // the caller's arrays are copied asynchronously
    std::vector<int64_t> offs2(batch);
    int64_t tot = 0, ptot = 0;
    for (int64_t b = 0; b < batch; b++) {
        offs2[b] = tot;
        tot += n[b] * n[b];
        ptot += n[b];
    }
    double* dA = (double*)g_chin.get((size_t)std::max<int64_t>(tot, 1) * sizeof(double));
    double* dR = (double*)g_chout.get((size_t)std::max<int64_t>(tot, 1) * sizeof(double));
    int64_t* dp = (int64_t*)g_chp.get((size_t)(ptot + batch + 1) * sizeof(int64_t));
    if (!dA || !dR || !dp) return set_err(SKH_ERR_HIP, "hipMalloc of skew-Cholesky buffers failed");
    int64_t* dsteps = dp + ptot;
    for (int64_t b = 0; b < batch; b++)
        HIPCHK(hipMemcpyAsync(dA + offs2[b], A + offset[b], (size_t)n[b] * n[b] * sizeof(double), hipMemcpyHostToDevice,
                              st));
    rc = chol_run(batch, n, offs2.data(), dA, dR, dp, dsteps, pivot, st);
    if (rc) return rc;
    for (int64_t b = 0; b < batch; b++)
        HIPCHK(hipMemcpyAsync(R + offset[b], dR + offs2[b], (size_t)n[b] * n[b] * sizeof(double), hipMemcpyDeviceToHost,
                              st));
    HIPCHK(hipMemcpyAsync(p, dp, (size_t)ptot * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    if (steps) HIPCHK(hipMemcpyAsync(steps, dsteps, (size_t)batch * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

}  // namespace

extern "C" {

const char* skh_version(void) { return "skewhip 0.1.0 (gfx950, FP64)"; }

const char* skh_last_error(void) { return g_err.c_str(); }

int skh_init(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_inited && device != g_device) return set_err(SKH_ERR_UNSUPPORTED, "already initialised on device %d", g_device);
    return ensure_init_locked(device);
}

int skh_finalize(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_inited) return 0;
    (void)hipStreamSynchronize(g_stream);
    for (Buf* b : {&g_mat, &g_ws, &g_vec, &g_ctl, &g_bin, &g_bout, &g_q, &g_ev, &g_cslot, &g_big, &g_bctl, &g_cwbuf, &g_zws, &g_zq, &g_zc, &g_zx, &g_zbws, &g_zpslot, &g_zpctl, &g_zpin,
                   &g_zpout, &g_chctl, &g_chin, &g_chout, &g_chp,
                   &g_chws, &g_rotq, &g_rotk, &g_rotcs, &g_rotseg, &g_fvec, &g_fpiv, &g_fz, &g_stq})
        b->release();
    g_timer.release();
    grad_plan_free(g_gradplan);
    g_gradplan = nullptr;
    if (g_blas) (void)rocblas_destroy_handle(g_blas);
    g_blas = nullptr;
    (void)hipStreamDestroy(g_stream);
    g_stream = nullptr;
    g_inited = false;
    return 0;
}

int skh_sytrd_d_device(int64_t n, double* dA, int64_t lda, double* dtau, double* deta, void* stream,
                       const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && !dA) return set_err(-2, "A is NULL");
    if (lda < std::max<int64_t>(n, 1) || (lda & 1)) return set_err(-3, "lda must be >= n and even");
    if (n > 1 && (!dtau || !deta)) return set_err(-4, "tau/eta NULL");
    if (((uintptr_t)dA) & 15) return set_err(-2, "A must be 16-byte aligned");
    if (opts && (opts->nb < 0 || opts->nb > skh::NBMAX)) return set_err(-7, "nb out of range");
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (n <= 1) return 0;
    skh_opts o = opts ? *opts : skh_opts{0, 0};
    o.flags = SKH_ENGINE_LARGE | (o.flags & (SKH_FLAG_TIMING | SKH_FLAG_PHASES));
    return sytrd_device_locked(n, dA, lda, dtau, deta, st, &o);
}

int skh_sytrd_d_device_noalloc(int64_t n, double* dA, int64_t lda, double* dtau, double* deta, void* stream,
                               const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n < 0) return set_err(-1, "n < 0");
    if (n > SKH_NOALLOC_MAX_N)
        return set_err(-1, "n = %lld exceeds the fixed scratch of this entry point (n <= %d)", (long long)n,
                       SKH_NOALLOC_MAX_N);
    if (n > 0 && !dA) return set_err(-2, "A is NULL");
    if (lda < std::max<int64_t>(n, 1) || (lda & 1)) return set_err(-3, "lda must be >= n and even");
    if (n > 1 && (!dtau || !deta)) return set_err(-4, "tau/eta NULL");
    if (((uintptr_t)dA) & 15) return set_err(-2, "A must be 16-byte aligned");
    if (opts && (opts->nb < 0 || opts->nb > skh::NBMAX)) return set_err(-7, "nb out of range");
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    if (n <= 1) return 0;
    unsigned* ctl = nullptr;
    double* ws = nullptr;
    size_t cap = 0;
    if (g_noalloc_status || skh_internal_noalloc_scratch(g_device, &ctl, &ws, &cap))
        return set_err(SKH_ERR_HIP, "the fixed scratch could not be located on device %d: %s", g_device,
                       hipGetErrorString((hipError_t)(g_noalloc_status ? g_noalloc_status : hipErrorNotInitialized)));
    g_timer.reset();
    g_timer.on = false;
    g_timer.phases = false;
    return sytrd_large(n, dA, lda, dtau, deta, pick_nb(n, opts), (hipStream_t)stream, ws, cap, ctl);
}

int skh_device_status(void* stream)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    rc = check_barrier_flag();
    if (rc) return rc;
    unsigned* ctl = nullptr;
    double* ws = nullptr;
    size_t cap = 0;
    if (!g_noalloc_status && !skh_internal_noalloc_scratch(g_device, &ctl, &ws, &cap)) {
        Ctl c;
        ctl_at(ctl, c);
        int hv[3] = {0, 0, 0};
        HIPCHK(hipMemcpy(hv, c.err, sizeof(hv), hipMemcpyDeviceToHost));
        rc = barrier_status(hv);
        if (rc) return rc;
    }
    return 0;
}

int skh_sytrd_d(int64_t n, double* A, int64_t lda, double* tau, double* eta, int64_t* ntau_nonzero,
                const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && !A) return set_err(-2, "A is NULL");
    if (lda < std::max<int64_t>(n, 1)) return set_err(-3, "lda < n");
    if (n > 1 && (!tau || !eta)) return set_err(-4, "tau/eta NULL");
    if (opts && (opts->nb < 0 || opts->nb > skh::NBMAX)) return set_err(-7, "nb out of range");
    if (ntau_nonzero) *ntau_nonzero = 0;
    if (n <= 1) return 0;
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = g_stream;
    const int64_t ld = (n + 1) & ~int64_t(1);
    double* dA = (double*)g_mat.get((size_t)ld * n * sizeof(double));
    double* dv = (double*)g_vec.get((size_t)2 * n * sizeof(double));
    if (!dA || !dv) return set_err(SKH_ERR_HIP, "hipMalloc failed for n = %lld", (long long)n);
    double* dtau = dv;
    double* deta = dv + n;
    std::vector<char> stage;
    StreamDrain drain(st);   // stage and the caller's arrays are copied asynchronously
    HIPCHK(upload_cols(dA, ld * sizeof(double), A, lda * sizeof(double), n * sizeof(double), n, stage, st));
    if (!stage.empty()) HIPCHK(hipStreamSynchronize(st));   // every return path below outlives the copy
    if (engine_of(n, opts) == SKH_ENGINE_LARGE) {
        g_timer.reset();
        g_timer.on = false;
        g_timer.phases = false;
        rc = sytrd_large(n, dA, ld, dtau, deta, pick_nb(n, opts), st);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(st));
        rc = check_barrier_flag();
        if (rc) return rc;
    } else {
        // one-workgroup engines, in place (ld = n for even n); tau | eta land in dv
        if (n % 2 == 1 || n > CMP_NMAX)
            return set_err(SKH_ERR_UNSUPPORTED, "the compact engine takes even n <= %lld", (long long)CMP_NMAX);
        double* dout = (double*)g_bout.get(2 * sizeof(double));
        if (!dout) return set_err(SKH_ERR_HIP, "hipMalloc failed");
        int64_t nn = n, off = 0, toff = 0;
        rc = batched_run(1, &nn, &off, dA, dout, dout + 1, opts, st, dA, dv, &toff);
        if (rc) return rc;
    }
    HIPCHK(download_cols(A, lda * sizeof(double), dA, ld * sizeof(double), n * sizeof(double), n, st));
    HIPCHK(hipMemcpyAsync(tau, dtau, (n - 1) * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(eta, deta, (n - 1) * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (ntau_nonzero) {
        int64_t c = 0;
        for (int64_t k = 0; k < n - 1; k++) c += (tau[k] != 0.0);
        *ntau_nonzero = c;
    }
    return 0;
}

int skh_logpf_d(int64_t n, const double* A, int64_t lda, double* logpf, double* sgn, const skh_opts* opts)
{
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && !A) return set_err(-2, "A is NULL");
    if (lda < std::max<int64_t>(n, 1)) return set_err(-3, "lda < n");
    if (!logpf || !sgn) return set_err(-4, "output NULL");
    if (n % 2 == 1) {
        *logpf = -INFINITY;
        *sgn = 0.0;
        return 0;
    }
    if (n == 0) {
        *logpf = 0.0;
        *sgn = 1.0;
        return 0;
    }
    if (engine_of(n, opts) == SKH_ENGINE_TEAM1) {
        int64_t nn = n, off = 0;
        std::vector<double> packed((size_t)n * n);
        for (int64_t c = 0; c < n; c++) std::memcpy(&packed[c * n], A + c * lda, n * sizeof(double));
        return skh_logpf_batched_d(1, &nn, &off, packed.data(), logpf, sgn, opts);
    }
    std::vector<double> F((size_t)n * n), tau(n - 1), eta(n - 1);
    for (int64_t c = 0; c < n; c++) std::memcpy(&F[c * n], A + c * lda, n * sizeof(double));
    int64_t nz = 0;
    int rc = skh_sytrd_d(n, F.data(), n, tau.data(), eta.data(), &nz, opts);
    if (rc) return rc;
    // src/pfaffian.jl:171-178, in the reference's summation order
    double l = 0.0, s = 1.0;
    for (int64_t i = 0; i + 1 < n; i += 2) {
        l += std::log(std::fabs(eta[i]));
        const double m = -eta[i];
        s *= (m > 0) - (m < 0);
    }
    *logpf = l;
    *sgn = (nz % 2) ? -s : s;
    return 0;
}

int skh_logpf_d_device(int64_t n, const double* dA, int64_t lda, double* dlogpf, double* dsgn, void* stream,
                       const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && !dA) return set_err(-2, "A is NULL");
    if (lda < std::max<int64_t>(n, 1)) return set_err(-3, "lda < n");
    if (!dlogpf || !dsgn) return set_err(-4, "output NULL");
    if (((uintptr_t)dA) & 15) return set_err(-2, "A must be 16-byte aligned");
    if (opts && (opts->nb < 0 || opts->nb > skh::NBMAX)) return set_err(-7, "nb out of range");
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int eng = engine_of(n, opts);
    const bool one_wg = eng == SKH_ENGINE_TEAM1 || eng == SKH_ENGINE_BLOCKED;   // one workgroup per matrix
    if (n % 2 == 1 || n <= 1 || (one_wg && n <= CMP_NMAX)) {
        // one workgroup (or the closed form for odd and empty n), on a packed copy when lda != n
        const double* src = dA;
        if (lda != n && n > 1) {
            double* P = (double*)g_big.get((size_t)n * n * sizeof(double));
            if (!P) return set_err(SKH_ERR_HIP, "hipMalloc of %lld doubles failed", (long long)(n * n));
            HIPCHK(hipMemcpy2DAsync(P, n * sizeof(double), dA, lda * sizeof(double), n * sizeof(double), n,
                                    hipMemcpyDeviceToDevice, st));
            src = P;
        }
        int64_t nn = n, off = 0;
        return batched_run(1, &nn, &off, src, dlogpf, dsgn, opts, st);
    }
    if (one_wg) return set_err(SKH_ERR_UNSUPPORTED, "the one-workgroup engines take n <= %lld", (long long)CMP_NMAX);
    // large engine on a working copy (even leading dimension), log|Pf| and sign from eta, tau on the device
    const int64_t ld = (n + 1) & ~int64_t(1);
    double* w = (double*)g_big.get((size_t)(ld * n + 2 * n) * sizeof(double));
    if (!w) return set_err(SKH_ERR_HIP, "hipMalloc of %lld doubles failed", (long long)(ld * n + 2 * n));
    double* te = w + ld * n;
    HIPCHK(hipMemcpy2DAsync(w, ld * sizeof(double), dA, lda * sizeof(double), n * sizeof(double), n,
                            hipMemcpyDeviceToDevice, st));
    g_timer.reset();
    g_timer.on = false;
    g_timer.phases = false;
    rc = sytrd_large(n, w, ld, te, te + n, pick_nb(n, opts), st);
    if (rc) return rc;
    HIPCHK((hipError_t)skh_internal_logpf_from(te + n, te, n, dlogpf, dsgn, st));
    return 0;
}

int skh_logpf_batched_d_device(int64_t batch, const int64_t* n, const int64_t* offset, const double* dA,
                               double* dlogpf, double* dsgn, void* stream, const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (batch < 0) return set_err(-1, "batch < 0");
    if (batch > 0 && (!n || !offset || !dA || !dlogpf || !dsgn)) return set_err(-2, "NULL argument");
    for (int64_t b = 0; b < batch; b++) {
        if (n[b] < 0) return set_err(-2, "n[%lld] < 0", (long long)b);
        if (offset[b] < 0 || (offset[b] & 1)) return set_err(-3, "offset[%lld] must be >= 0 and even", (long long)b);
    }
    if (opts && (opts->nb < 0 || opts->nb > skh::NBMAX)) return set_err(-8, "nb out of range");
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    return batched_run(batch, n, offset, dA, dlogpf, dsgn, opts, st);
}

int skh_logpf_batched_d(int64_t batch, const int64_t* n, const int64_t* offset, const double* Apacked,
                        double* logpf, double* sgn, const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (batch < 0) return set_err(-1, "batch < 0");
    if (batch == 0) return 0;
    if (!n || !offset || !Apacked || !logpf || !sgn) return set_err(-2, "NULL argument");
    for (int64_t b = 0; b < batch; b++) {
        if (n[b] < 0) return set_err(-2, "n[%lld] < 0", (long long)b);
        if (offset[b] < 0) return set_err(-3, "offset[%lld] < 0", (long long)b);
    }
    if (opts && (opts->nb < 0 || opts->nb > skh::NBMAX)) return set_err(-7, "nb out of range");
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = g_stream;
    StreamDrain drain(st);   // the caller's arrays are copied asynchronously
    // repack with even offsets (16-byte aligned matrices)
    std::vector<int64_t> offs2(batch);
    int64_t tot2 = 0;
    for (int64_t b = 0; b < batch; b++) {
        offs2[b] = tot2;
        tot2 += (n[b] * n[b] + 1) & ~int64_t(1);
    }
    double* dA = (double*)g_bin.get(std::max<int64_t>(tot2, 1) * sizeof(double));
    double* dout = (double*)g_bout.get(2 * batch * sizeof(double));
    if (!dA || !dout) return set_err(SKH_ERR_HIP, "hipMalloc of batched buffers failed");
    for (int64_t b = 0; b < batch; b++)
        HIPCHK(hipMemcpyAsync(dA + offs2[b], Apacked + offset[b], n[b] * n[b] * sizeof(double), hipMemcpyHostToDevice, st));
    rc = batched_run(batch, n, offs2.data(), dA, dout, dout + batch, opts, st);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(logpf, dout, batch * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(sgn, dout + batch, batch * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int skh_orgtr_d_device(int64_t n, const double* dF, int64_t ldf, const double* dtau, double* dQ, int64_t ldq,
                       void* stream)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && (!dF || !dQ)) return set_err(-2, "NULL matrix");
    if (ldf < std::max<int64_t>(n, 1)) return set_err(-3, "ldf < n");
    if (n > 1 && !dtau) return set_err(-4, "tau NULL");
    if (ldq < std::max<int64_t>(n, 1)) return set_err(-6, "ldq < n");
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    return orgtr_locked(n, dF, ldf, dtau, dQ, ldq, (hipStream_t)stream);
}

// skeweigen! (src/eigen.jl:75-97) of the real skew A times `scale` (1 or 2, exact), the eigenvectors left
// on the device.  The host part (tridiagonalization call, QR sweeps) runs unlocked; on success lk holds g_mu,
// vals (host) holds the imaginary parts of the eigenvalues and *base a buffer of `slots` >= 5 n x n
// matrices with Q Qr in slot EV_R and Q Qim in slot EV_I (slots 1-3 and from 5 on are free).
}  // extern "C"

namespace {
constexpr int EV_I = 0, EV_R = 4;
int eigen_d_dev(int64_t n, const double* A, int64_t lda, double scale, double* vals, const skh_opts* opts, int slots,
                std::unique_lock<std::mutex>& lk, double** base)
{
    const size_t nn = (size_t)n * n;
    std::vector<double> F(nn), tau(std::max<int64_t>(n - 1, 1)), eta(std::max<int64_t>(n - 1, 1));
    for (int64_t c = 0; c < n; c++)
        for (int64_t r = 0; r < n; r++) F[c * n + r] = A[c * lda + r] * scale;
    // implicit QR with Givens accumulation (skewtrieigen_divided!, src/eigen.jl:87): the sweeps on
    // the host, their rotations applied to Qodd / Qeven on the GPU from n = 256
    std::vector<double> hr(nn), hi(nn);
    if (n == 1) {   // src/eigen.jl:77-79
        vals[0] = 0.0;
        hr[0] = 1.0;
        hi[0] = 0.0;
    } else {
        int rc = skh_sytrd_d(n, F.data(), n, tau.data(), eta.data(), nullptr, opts);
        if (rc) return rc;
        rc = skh_internal_skewtrieigen(n, eta.data(), vals, hr.data(), hi.data(), 0,
                                       n >= ROT_GPU_MIN ? rot_apply_gpu : nullptr, symtri_gpu, nullptr);
        if (rc == 2) return set_err(SKH_ERR_NOCONV, "skewtrieigen did not converge");
        if (rc) return rc < 0 ? set_err(rc, "skewtrieigen failed (%d)", rc) : rc;
    }
    lk = std::unique_lock<std::mutex>(g_mu);
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = g_stream;
    StreamDrain drain(st);   // F, tau, hr and hi are copied asynchronously
    double* d = (double*)g_q.get((size_t)slots * nn * sizeof(double) + 2 * n * sizeof(double));
    if (!d) return set_err(SKH_ERR_HIP, "hipMalloc failed for eigen n = %lld", (long long)n);
    *base = d;
    if (n == 1) {
        HIPCHK(hipMemcpyAsync(d + EV_R * nn, hr.data(), sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d + EV_I * nn, hi.data(), sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    double* dF = d;
    double* dQ = d + nn;
    double* dR = d + 2 * nn;
    double* dI = d + 3 * nn;
    double* dtau = d + slots * nn;
    HIPCHK(hipMemcpyAsync(dF, F.data(), nn * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dtau, tau.data(), (n - 1) * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dR, hr.data(), nn * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dI, hi.data(), nn * sizeof(double), hipMemcpyHostToDevice, st));
    rc = orgtr_locked(n, dF, n, dtau, dQ, n, st);
    if (rc) return rc;
    rc = blas_on(st);
    if (rc) return rc;
    // vectorsreal = Q * Qr, vectorsim = Q * Qim (src/eigen.jl:88-89); the factors in slot 0 are spent
    const double one = 1.0, zero = 0.0;
    BLASCHK(rocblas_dgemm(g_blas, rocblas_operation_none, rocblas_operation_none, (int)n, (int)n, (int)n, &one, dQ,
                          (int)n, dR, (int)n, &zero, d + EV_R * nn, (int)n));
    BLASCHK(rocblas_dgemm(g_blas, rocblas_operation_none, rocblas_operation_none, (int)n, (int)n, (int)n, &one, dQ,
                          (int)n, dI, (int)n, &zero, d + EV_I * nn, (int)n));
    // the uploads above read F, tau, hr and hi, which end with this function
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}
}  // namespace

extern "C" {

int skh_eigen_d(int64_t n, const double* A, int64_t lda, double* vals, double* vr, double* vi, const skh_opts* opts)
{
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && (!A || !vals || !vr || !vi)) return set_err(-2, "NULL argument");
    if (lda < std::max<int64_t>(n, 1)) return set_err(-3, "lda < n");
    if (n == 0) return 0;
    if (n == 1) {   // src/eigen.jl:77-79
        vals[0] = 0.0;
        vr[0] = 1.0;
        vi[0] = 0.0;
        return 0;
    }
    std::unique_lock<std::mutex> lk;
    double* d = nullptr;
    int rc = eigen_d_dev(n, A, lda, 1.0, vals, opts, 5, lk, &d);
    if (rc) return rc;
    const size_t nn = (size_t)n * n;
    HIPCHK(hipMemcpyAsync(vr, d + EV_R * nn, nn * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipMemcpyAsync(vi, d + EV_I * nn, nn * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return 0;
}

int skh_logpf_grad_batched_d_device(int64_t batch, const int64_t* n, const int64_t* offset, const double* dA,
                                    double* dlogpf, double* dsgn, double* dgrad, void* stream, const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (batch < 0) return set_err(-1, "batch < 0");
    if (batch > 0 && (!n || !offset || !dA || !dlogpf || !dsgn || !dgrad)) return set_err(-2, "NULL argument");
    for (int64_t b = 0; b < batch; b++) {
        if (n[b] < 0) return set_err(-2, "n[%lld] < 0", (long long)b);
        if (offset[b] < 0 || (offset[b] & 1)) return set_err(-3, "offset[%lld] must be >= 0 and even", (long long)b);
        if (n[b] > 46340) return set_err(-2, "n[%lld] too large", (long long)b);
    }
    if (opts && (opts->nb < 0 || opts->nb > skh::NBMAX)) return set_err(-9, "nb out of range");
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    return grad_run(batch, n, offset, dA, dlogpf, dsgn, dgrad, opts, (hipStream_t)stream);
}

int skh_logpf_grad_batched_d(int64_t batch, const int64_t* n, const int64_t* offset, const double* Apacked,
                             double* logpf, double* sgn, double* grad, const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (batch < 0) return set_err(-1, "batch < 0");
    if (batch == 0) return 0;
    if (!n || !offset || !Apacked || !logpf || !sgn || !grad) return set_err(-2, "NULL argument");
    for (int64_t b = 0; b < batch; b++) {
        if (n[b] < 0) return set_err(-2, "n[%lld] < 0", (long long)b);
        if (offset[b] < 0) return set_err(-3, "offset[%lld] < 0", (long long)b);
        if (n[b] > 46340) return set_err(-2, "n[%lld] too large", (long long)b);
    }
    if (opts && (opts->nb < 0 || opts->nb > skh::NBMAX)) return set_err(-8, "nb out of range");
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = g_stream;
    StreamDrain drain(st);   // the caller's arrays are copied asynchronously
    std::vector<int64_t> offs2(batch);
    int64_t tot2 = 0;
    for (int64_t b = 0; b < batch; b++) {
        offs2[b] = tot2;
        tot2 += (n[b] * n[b] + 1) & ~int64_t(1);
    }
    double* dA = (double*)g_bin.get(std::max<int64_t>(tot2, 1) * sizeof(double));
    double* dG = (double*)g_mat.get(std::max<int64_t>(tot2, 1) * sizeof(double));
    double* dout = (double*)g_bout.get(2 * batch * sizeof(double));
    if (!dA || !dG || !dout) return set_err(SKH_ERR_HIP, "hipMalloc of batched buffers failed");
    for (int64_t b = 0; b < batch; b++)
        HIPCHK(hipMemcpyAsync(dA + offs2[b], Apacked + offset[b], n[b] * n[b] * sizeof(double), hipMemcpyHostToDevice, st));
    rc = grad_run(batch, n, offs2.data(), dA, dout, dout + batch, dG, opts, st);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(logpf, dout, batch * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(sgn, dout + batch, batch * sizeof(double), hipMemcpyDeviceToHost, st));
    for (int64_t b = 0; b < batch; b++)
        HIPCHK(hipMemcpyAsync(grad + offset[b], dG + offs2[b], n[b] * n[b] * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int skh_diag_timings(double* out10)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!out10) return set_err(-1, "NULL");
    double* out6 = out10;
    for (int k = 0; k < 10; k++) out10[k] = 0.0;
    if (g_timer.used == 0) return 0;
    HIPCHK(hipEventSynchronize(g_timer.ev[g_timer.used - 1]));
    {
        Ctl ctl;
        int rc = get_ctl(ctl);
        if (rc) return rc;
        unsigned long long ph[4];
        HIPCHK(hipMemcpy(ph, ctl.ph, sizeof(ph), hipMemcpyDeviceToHost));
        int khz = 0;
        HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g_device));
        for (int k = 0; k < 4; k++) out10[6 + k] = khz > 0 ? (double)ph[k] / (double)khz : 0.0;
    }
    for (size_t q = 0; q < g_timer.kind.size(); q++) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, g_timer.ev[2 * q], g_timer.ev[2 * q + 1]));
        const int k = g_timer.kind[q];
        out6[2 * k] += ms;
        out6[2 * k + 1] += 1.0;
        out6[4 + k] += g_timer.work[q];
    }
    return 0;
}

int skh_diag_phase_clock(double* out, int count)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!out) return set_err(-1, "NULL");
    if (count < 0 || count > 13) return set_err(-2, "count must be in [0, 13]");
    if (g_timer.used) HIPCHK(hipEventSynchronize(g_timer.ev[g_timer.used - 1]));
    Ctl ctl;
    int rc = get_ctl(ctl);
    if (rc) return rc;
    unsigned long long ph[13];
    HIPCHK(hipMemcpy(ph, ctl.ph, sizeof(ph), hipMemcpyDeviceToHost));
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g_device));
    for (int k = 0; k < count; k++) out[k] = khz > 0 ? (double)ph[k] / (double)khz : 0.0;
    return 0;
}

int skh_diag_mfma_probe(double* out768)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!out768) return set_err(-1, "NULL");
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    double* d = (double*)g_vec.get(768 * sizeof(double));
    if (!d) return set_err(SKH_ERR_HIP, "hipMalloc failed");
    HIPCHK((hipError_t)skh_internal_mfma_probe(d, g_stream));
    HIPCHK(hipMemcpyAsync(out768, d, 768 * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return 0;
}

int skh_logpf_batched_plan_create(int64_t batch, const int64_t* n, const int64_t* offset, const skh_opts* opts,
                                  skh_plan* plan)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (batch < 1) return set_err(-1, "batch must be >= 1");
    if (!n || !offset) return set_err(-2, "NULL argument");
    if (!plan) return set_err(-5, "plan is NULL");
    *plan = nullptr;
    int64_t nmax = 2;
    for (int64_t b = 0; b < batch; b++) {
        if (n[b] < 0) return set_err(-2, "n[%lld] < 0", (long long)b);
        if (n[b] > CMP_NMAX && n[b] % 2 == 0)
            return set_err(SKH_ERR_UNSUPPORTED, "plans take even n <= %lld (n[%lld] = %lld); use "
                           "skh_logpf_batched_d_device", (long long)CMP_NMAX, (long long)b, (long long)n[b]);
        if (offset[b] < 0 || (offset[b] & 1)) return set_err(-3, "offset[%lld] must be >= 0 and even", (long long)b);
        if (n[b] % 2 == 0) nmax = std::max(nmax, n[b]);
    }
    (void)opts;
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    skh_plan p = new skh_plan_s();
    p->batch = batch;
    p->count = batch;
    p->nmax = nmax;
    p->kmax = nmax <= 512 ? 512 : 1024;
    // untuned: the blocked engine (panel width 8) when every member fits it, else the unblocked one
    int bk = 0, bnb = 0;
    if (nmax <= BLK_NMAX && blocked_instance(nmax, 0, &bk, &bnb)) {
        p->kmax = bk;
        p->knb = bnb;
    }
    int per_cu = 0;
    hipError_t e = p->knb ? (hipError_t)skh_internal_cblock_occupancy(p->kmax, p->knb, &per_cu)
                          : (hipError_t)skh_internal_compact_occupancy(p->kmax, &per_cu);
    if (e != hipSuccess) { plan_free(p); return set_err(SKH_ERR_HIP, "occupancy query failed"); }
    per_cu = std::max(1, per_cu);
    p->grid = (int)std::min<int64_t>((int64_t)per_cu * g_num_cu, batch);
    p->slot = ((nmax * nmax + 2 * 1024) + 31) & ~int64_t(31);   // fits every kernel instance
    std::vector<int64_t> order(batch);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return n[a] > n[b]; });
    std::vector<int64_t> h(3 * batch);
    std::copy(n, n + batch, h.begin());
    std::copy(offset, offset + batch, h.begin() + batch);
    std::copy(order.begin(), order.end(), h.begin() + 2 * batch);
    for (int64_t b = 0; b < batch; b++) p->flops += 4.0 / 3.0 * (double)n[b] * (double)n[b] * (double)n[b];
    // work slots for the largest grid the tuner tries (a persistent queue kernel: workgroups beyond the
    // resident ones start when others finish and leave once the queue is empty)
    const int64_t max_grid = (int64_t)std::max(per_cu, PLAN_TUNE_PER_CU) * g_num_cu;
    if (hipMalloc((void**)&p->dctl, sizeof(int64_t) * 3 * batch) != hipSuccess ||
        hipMalloc((void**)&p->slots, sizeof(double) * (size_t)p->slot * (size_t)std::min<int64_t>(max_grid, batch)) !=
            hipSuccess ||
        hipMalloc((void**)&p->ctl, 256) != hipSuccess ||
        hipMalloc((void**)&p->wbuf, sizeof(double) * (size_t)PLAN_WSTRIDE * (size_t)std::min<int64_t>(max_grid, batch)) !=
            hipSuccess) {
        plan_free(p);
        return set_err(SKH_ERR_HIP, "hipMalloc of the plan's buffers failed");
    }
    if (hipMemcpy(p->dctl, h.data(), sizeof(int64_t) * 3 * batch, hipMemcpyHostToDevice) != hipSuccess) {
        plan_free(p);
        return set_err(SKH_ERR_HIP, "upload of the plan's control arrays failed");
    }
    *plan = p;
    return 0;
}

int skh_logpf_batched_plan_exec(skh_plan plan, const double* dA, double* dlogpf, double* dsgn, void* stream)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan) return set_err(-1, "plan is NULL");
    if (plan->g) return set_err(-1, "a gradient plan runs through skh_logpf_grad_batched_plan_exec");
    if (!dA || !dlogpf || !dsgn) return set_err(-2, "NULL device pointer");
    return plan_enqueue(plan, dA, dlogpf, dsgn, (hipStream_t)stream, plan->grid, plan->kmax, plan->knb);
}

int skh_logpf_grad_batched_plan_create(int64_t batch, const int64_t* n, const int64_t* offset, const skh_opts* opts,
                                       skh_plan* plan)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (batch < 1) return set_err(-1, "batch must be >= 1");
    if (!n || !offset) return set_err(-2, "NULL argument");
    if (!plan) return set_err(-5, "plan is NULL");
    *plan = nullptr;
    for (int64_t b = 0; b < batch; b++) {
        if (n[b] < 0) return set_err(-2, "n[%lld] < 0", (long long)b);
        if (offset[b] < 0 || (offset[b] & 1)) return set_err(-3, "offset[%lld] must be >= 0 and even", (long long)b);
    }
    (void)opts;
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    GradPlan* g = nullptr;
    rc = grad_plan_build(batch, n, offset, false, &g);
    if (rc) return rc;
    skh_plan p = new skh_plan_s();
    p->g = g;
    p->batch = batch;
    p->count = g->count;
    p->kmax = g->kmax;
    p->grid = g->grid;
    p->ctl = g->ctl;
    p->flops = g->flops;
    *plan = p;
    return 0;
}

int skh_logpf_grad_batched_plan_exec(skh_plan plan, const double* dA, double* dlogpf, double* dsgn, double* dgrad,
                                     void* stream)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan || !plan->g) return set_err(-1, "plan is NULL or not a gradient plan");
    if (!dA || !dlogpf || !dsgn || !dgrad) return set_err(-2, "NULL device pointer");
    if ((((uintptr_t)dA) | ((uintptr_t)dgrad)) & 15) return set_err(-2, "dA and dgrad must be 16-byte aligned");
    return grad_enqueue(*plan->g, dA, dlogpf, dsgn, dgrad, plan->grid, (hipStream_t)stream);
}

int skh_logpf_batched_plan_tune(skh_plan plan, const double* dA, double* dlogpf, double* dsgn, void* stream,
                                double* best_ms)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan) return set_err(-1, "plan is NULL");
    if (plan->g) return set_err(-1, "gradient plans are not tuned by this call");
    if (!dA || !dlogpf || !dsgn) return set_err(-2, "NULL device pointer");
    hipStream_t st = (hipStream_t)stream;
    // candidates: the unblocked engine (kernel NMAX 512 or 1024) and the blocked engine (panel width
    // 8 or 16 where an instance exists), each at 1..PLAN_TUNE_PER_CU workgroups per CU
    struct Cand { int grid, kmax, knb; };
    std::vector<Cand> cands;
    auto add = [&](int kmax, int knb) {
        for (int per = 1; per <= PLAN_TUNE_PER_CU; per++) {
            const int g = (int)std::min<int64_t>((int64_t)per * g_num_cu, plan->count);
            bool dup = false;
            for (auto& c : cands) dup |= (c.grid == g && c.kmax == kmax && c.knb == knb);
            if (!dup) cands.push_back({g, kmax, knb});
        }
    };
    for (int k : {512, 1024})
        if (k >= plan->nmax) add(k, 0);
    for (int w : {8, 16}) {
        int bk = 0, bnb = 0;
        if (plan->nmax <= BLK_NMAX && blocked_instance(plan->nmax, w, &bk, &bnb) && !(bk == 512 && bnb == 16))
            add(bk, bnb);
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float best = 1e30f;
    Cand pick{plan->grid, plan->kmax, plan->knb};
    for (const Cand& c : cands) {
        int rc = plan_enqueue(plan, dA, dlogpf, dsgn, st, c.grid, c.kmax, c.knb);   // warm-up
        if (rc) return rc;
        HIPCHK(hipEventRecord(e0, st));
        for (int r = 0; r < 3; r++) {
            rc = plan_enqueue(plan, dA, dlogpf, dsgn, st, c.grid, c.kmax, c.knb);
            if (rc) return rc;
        }
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        ms /= 3.f;
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    plan->grid = pick.grid;
    plan->kmax = pick.kmax;
    plan->knb = pick.knb;
    plan->tuned = 1;
    if (best_ms) *best_ms = best;
    return 0;
}

int skh_plan_config(skh_plan plan, int64_t* out4)
{
    if (!plan || !out4) return set_err(-1, "NULL argument");
    const bool blk = plan->g ? plan->g->knb > 0 : plan->knb > 0;
    out4[0] = blk ? SKH_ENGINE_BLOCKED : SKH_ENGINE_TEAM1;
    out4[1] = plan->g ? plan->g->kmax : plan->kmax;
    out4[2] = plan->g ? plan->g->knb : plan->knb;
    out4[3] = plan->g ? plan->g->grid : plan->grid;
    return 0;
}

int skh_plan_set_config(skh_plan plan, const int64_t* in4)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan || !in4) return set_err(-1, "NULL argument");
    if (plan->g) return set_err(-1, "gradient plans take their configuration at creation");
    const int64_t eng = in4[0], kmax = in4[1], knb = in4[2], grid = in4[3];
    const int64_t max_grid = (int64_t)PLAN_TUNE_PER_CU * g_num_cu;
    if (grid < 1 || grid > std::min<int64_t>(max_grid, plan->count)) return set_err(-2, "grid out of range");
    if (eng == SKH_ENGINE_BLOCKED) {
        int bk = 0, bnb = 0;
        if (plan->nmax > BLK_NMAX || !blocked_instance(plan->nmax, knb, &bk, &bnb) || bk != kmax)
            return set_err(-2, "no blocked instance with NMAX %lld and panel width %lld for this batch",
                           (long long)kmax, (long long)knb);
        plan->knb = (int)knb;
    } else if (eng == SKH_ENGINE_TEAM1) {
        if ((kmax != 512 && kmax != 1024) || kmax < plan->nmax) return set_err(-2, "kernel NMAX out of range");
        plan->knb = 0;
    } else {
        return set_err(-2, "engine must be SKH_ENGINE_TEAM1 or SKH_ENGINE_BLOCKED");
    }
    plan->kmax = (int)kmax;
    plan->grid = (int)grid;
    plan->tuned = 1;
    return 0;
}

int skh_plan_status(skh_plan plan, void* stream)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!plan) return set_err(-1, "plan is NULL");
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    int herr = 0;
    HIPCHK(hipMemcpy(&herr, plan->ctl + 32, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) return set_err(SKH_ERR_BARRIER, "compact engine stopped by its watchdog");
    return 0;
}

int skh_plan_info(skh_plan plan, int64_t* out4)
{
    if (!plan || !out4) return set_err(-1, "NULL argument");
    out4[0] = plan->grid;
    out4[1] = plan->kmax;
    out4[2] = plan->tuned;
    out4[3] = plan->count;
    return 0;
}

int skh_plan_destroy(skh_plan plan)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (plan) (void)hipDeviceSynchronize();
    plan_free(plan);
    return 0;
}

int skh_sytrd_z_device(int64_t n, double* dA, int64_t lda, double* dtau, double* deta, double* ddvim, void* stream,
                       const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && !dA) return set_err(-2, "A is NULL");
    if (lda < std::max<int64_t>(n, 1)) return set_err(-3, "lda < n");
    if (n > 1 && (!dtau || !deta)) return set_err(-4, "tau/eta NULL");
    if (n > 0 && !ddvim) return set_err(-6, "dvim NULL");
    if (((uintptr_t)dA) & 15) return set_err(-2, "A must be 16-byte aligned");
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    return zsytrd_device(n, (double2*)dA, lda, (double2*)dtau, deta, ddvim, (hipStream_t)stream,
                         opts ? (int)(opts->flags & 0xFF) : SKH_ENGINE_AUTO);
}

int skh_sytrd_z(int64_t n, double* A, int64_t lda, double* tau, double* eta, double* dvim, const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && !A) return set_err(-2, "A is NULL");
    if (lda < std::max<int64_t>(n, 1)) return set_err(-3, "lda < n");
    if (n > 1 && (!tau || !eta)) return set_err(-4, "tau/eta NULL");
    if (n > 0 && !dvim) return set_err(-6, "dvim NULL");
    if (n == 0) return 0;
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = g_stream;
    StreamDrain drain(st);   // the caller's arrays are copied asynchronously
    double2* dA = zupload(n, A, lda, st, &rc);
    if (!dA) return rc;
    double* dv = (double*)g_vec.get((size_t)(4 * n + 4) * sizeof(double));
    if (!dv) return set_err(SKH_ERR_HIP, "hipMalloc failed");
    double2* dtau = (double2*)dv;
    double* deta = dv + 2 * n;
    double* ddvim = deta + n;
    rc = zsytrd_device(n, dA, n, dtau, deta, ddvim, st, opts ? (int)(opts->flags & 0xFF) : SKH_ENGINE_AUTO);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(st));
    rc = check_barrier_flag();
    if (rc) return rc;
    HIPCHK(download_cols(A, lda * sizeof(double2), dA, n * sizeof(double2), n * sizeof(double2), n, st));
    if (n > 1) {
        HIPCHK(hipMemcpyAsync(tau, dtau, (n - 1) * sizeof(double2), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(eta, deta, (n - 1) * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipMemcpyAsync(dvim, ddvim, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int skh_orgtr_z_device(int64_t n, const double* dF, int64_t ldf, const double* dtau, double* dQ, int64_t ldq,
                       void* stream)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && (!dF || !dQ)) return set_err(-2, "NULL matrix");
    if (ldf < std::max<int64_t>(n, 1)) return set_err(-3, "ldf < n");
    if (n > 1 && !dtau) return set_err(-4, "tau NULL");
    if (ldq < std::max<int64_t>(n, 1)) return set_err(-6, "ldq < n");
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    return zorgtr_locked(n, (const double2*)dF, ldf, (const double2*)dtau, (double2*)dQ, ldq, (hipStream_t)stream);
}

// eigenvalues of the complex skew-Hermitian A: GPU tridiagonalization, then the eigenvalues of
// the real symmetric tridiagonal D^H (iT) D (diagonal -dvim, off-diagonal eta) on the host
int skh_eigvals_z(int64_t n, const double* A, int64_t lda, double* vals, const skh_opts* opts)
{
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && (!A || !vals)) return set_err(-2, "NULL argument");
    if (lda < std::max<int64_t>(n, 1)) return set_err(-3, "lda < n");
    if (n == 0) return 0;
    std::vector<double> F((size_t)2 * n * n), tau(2 * n), eta(n), dvim(n);
    for (int64_t c = 0; c < n; c++) std::memcpy(&F[2 * c * n], A + 2 * c * lda, 2 * n * sizeof(double));
    int rc = skh_sytrd_z(n, F.data(), n, tau.data(), eta.data(), dvim.data(), opts);
    if (rc) return rc;
    std::vector<double> d(n);
    for (int64_t k = 0; k < n; k++) d[k] = -dvim[k];
    if (skh_internal_symtri_eigvals(n, d.data(), eta.data()) != 0)
        return set_err(SKH_ERR_NOCONV, "tridiagonal QL did not converge");
    // eigenvalues of A are -i lambda for the ascending eigenvalues lambda of iA (src/eigen.jl:19-22)
    for (int64_t k = 0; k < n; k++) vals[k] = -d[k];
    return 0;
}

// eigen-decomposition of the complex skew-Hermitian A times `scale` (1 or 2, exact) (src/eigen.jl:101-106):
// vals = -lambda (imaginary parts of the eigenvalues) for ascending lambda of iA, eigenvectors X = Q D Z
// left on the device in *dV (ld n).  The caller holds g_mu.
}  // extern "C"

namespace {
int eigen_z_dev_locked(int64_t n, const double* A, int64_t lda, double scale, double* vals, double2** dV)
{
    int rc = ensure_init_locked(g_device);
    if (rc) return rc;
    hipStream_t st = g_stream;
    double2* dA = zupload(n, A, lda, st, &rc);
    if (!dA) return rc;
    if (scale != 1.0) HIPCHK((hipError_t)skh_internal_fun_scale(2 * n * n, (double*)dA, scale, st));
    double* dv = (double*)g_vec.get((size_t)(4 * n + 4) * sizeof(double));
    double2* dQ = (double2*)g_zq.get((size_t)n * n * sizeof(double2));
    double* dC = (double*)g_zc.get((size_t)(n * n + 2 * n + 2) * sizeof(double));
    double2* dX = (double2*)g_zx.get((size_t)n * n * sizeof(double2));
    if (!dv || !dQ || !dC || !dX) return set_err(SKH_ERR_HIP, "hipMalloc failed (eigen buffers)");
    double2* dtau = (double2*)dv;
    double* deta = dv + 2 * n;
    double* ddvim = deta + n;
    rc = zsytrd_device(n, dA, n, dtau, deta, ddvim, st);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(st));
    rc = check_barrier_flag();
    if (rc) return rc;
    rc = zorgtr_locked(n, dA, n, dtau, dQ, n, st);
    if (rc) return rc;
    // tridiagonal eigenpairs on the GPU (divide and conquer)
    std::vector<double> eta(n), dvim(n), d(n);
    StreamDrain drain(st);   // eta, dvim and d are copied asynchronously
    if (n > 1) HIPCHK(hipMemcpyAsync(eta.data(), deta, (n - 1) * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(dvim.data(), ddvim, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int64_t k = 0; k < n; k++) d[k] = -dvim[k];
    double* dD = dC + n * n;
    double* dE = dD + n;
    int* dinfo = (int*)(dE + n);
    HIPCHK(hipMemcpyAsync(dD, d.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
    if (n > 1) HIPCHK(hipMemcpyAsync(dE, eta.data(), (n - 1) * sizeof(double), hipMemcpyHostToDevice, st));
    rc = blas_on(st);
    if (rc) return rc;
    BLASCHK(rocsolver_dstedc(g_blas, rocblas_evect_tridiagonal, (rocblas_int)n, dD, dE, dC, (rocblas_int)n, dinfo));
    int info = 0;
    HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(d.data(), dD, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (info != 0) return set_err(SKH_ERR_NOCONV, "tridiagonal divide and conquer failed (info %d)", info);
    HIPCHK((hipError_t)skh_internal_zphase(dC, n, dX, st));   // dX = D Z (reuses dX as Zc)
    // vectors = Q (D Z), into dA (no longer needed)
    const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
    BLASCHK(rocblas_zgemm(g_blas, rocblas_operation_none, rocblas_operation_none, (int)n, (int)n, (int)n, &one,
                          (const rocblas_double_complex*)dQ, (int)n, (const rocblas_double_complex*)dX, (int)n, &zero,
                          (rocblas_double_complex*)dA, (int)n));
    for (int64_t k = 0; k < n; k++) vals[k] = -d[k];
    *dV = dA;
    return 0;
}
}  // namespace

extern "C" {

int skh_eigen_z(int64_t n, const double* A, int64_t lda, double* vals, double* V, const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && (!A || !vals || !V)) return set_err(-2, "NULL argument");
    if (lda < std::max<int64_t>(n, 1)) return set_err(-3, "lda < n");
    (void)opts;
    if (n == 0) return 0;
    if (n > 46340) return set_err(SKH_ERR_UNSUPPORTED, "n too large for the solver interface");
    double2* dV = nullptr;
    int rc = eigen_z_dev_locked(n, A, lda, 1.0, vals, &dV);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(V, dV, (size_t)n * n * sizeof(double2), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    return 0;
}

int skh_pfaffian_z_batched(int64_t batch, const int64_t* n, const int64_t* offset, const double* Apacked, double* pf,
                           const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    (void)opts;
    int rc = zpf_check(batch, n, offset, Apacked, pf, pf, false);
    if (rc || batch == 0) return rc;
    return zpf_host(batch, n, offset, Apacked, 0, pf, nullptr);
}

int skh_pfaffian_z_batched_device(int64_t batch, const int64_t* n, const int64_t* offset, const double* dA,
                                  double* dpf, void* stream, const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    (void)opts;
    int rc = zpf_check(batch, n, offset, dA, dpf, dpf, true);
    if (rc || batch == 0) return rc;
    rc = ensure_init_locked(g_device);
    if (rc) return rc;
    return zpf_run(batch, n, offset, (const double2*)dA, 0, dpf, nullptr, (hipStream_t)stream);
}

int skh_logabspfaffian_z_batched(int64_t batch, const int64_t* n, const int64_t* offset, const double* Apacked,
                                 double* logpf, double* phase, const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    (void)opts;
    int rc = zpf_check(batch, n, offset, Apacked, logpf, phase, false);
    if (rc || batch == 0) return rc;
    return zpf_host(batch, n, offset, Apacked, 1, logpf, phase);
}

int skh_logabspfaffian_z_batched_device(int64_t batch, const int64_t* n, const int64_t* offset, const double* dA,
                                        double* dlogpf, double* dphase, void* stream, const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    (void)opts;
    int rc = zpf_check(batch, n, offset, dA, dlogpf, dphase, true);
    if (rc || batch == 0) return rc;
    rc = ensure_init_locked(g_device);
    if (rc) return rc;
    return zpf_run(batch, n, offset, (const double2*)dA, 1, dlogpf, dphase, (hipStream_t)stream);
}

int skh_skewchol_d_batched(int64_t batch, const int64_t* n, const int64_t* offset, const double* Apacked,
                           double* Rpacked, int64_t* p, int64_t* steps, int pivot, const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    (void)opts;
    int rc = chol_check(batch, n, offset, Apacked, Rpacked, p, pivot);
    if (rc || batch == 0) return rc;
    return chol_host(batch, n, offset, Apacked, Rpacked, p, steps, pivot);
}

int skh_skewchol_d_batched_device(int64_t batch, const int64_t* n, const int64_t* offset, const double* dA,
                                  double* dR, int64_t* dp, int64_t* dsteps, int pivot, void* stream,
                                  const skh_opts* opts)
{
    std::lock_guard<std::mutex> lk(g_mu);
    (void)opts;
    int rc = chol_check(batch, n, offset, dA, dR, dp, pivot);
    if (rc || batch == 0) return rc;
    rc = ensure_init_locked(g_device);
    if (rc) return rc;
    return chol_run(batch, n, offset, dA, dR, dp, dsteps, pivot, (hipStream_t)stream);
}

int skh_skewchol_d(int64_t n, const double* A, int64_t lda, double* R, int64_t ldr, int64_t* p, int pivot,
                   int64_t* steps, const skh_opts* opts)
{
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && (!A || !R || !p)) return set_err(-2, "NULL argument");
    if (lda < std::max<int64_t>(n, 1)) return set_err(-3, "lda < n");
    if (ldr < std::max<int64_t>(n, 1)) return set_err(-5, "ldr < n");
    if (pivot != 0 && pivot != 1) return set_err(-7, "pivot must be 0 (NoPivot) or 1 (RowMaximum)");
    if (n > skh_internal_skewchol_nmax()) return set_err(-1, "n = %lld exceeds %d", (long long)n,
                                                         skh_internal_skewchol_nmax());
    if (steps) *steps = 0;
    if (n == 0) return 0;
    std::vector<double> a((size_t)n * n), r((size_t)n * n);
    for (int64_t c = 0; c < n; c++) std::memcpy(&a[c * n], A + c * lda, n * sizeof(double));
    const int64_t off = 0;
    int rc = skh_skewchol_d_batched(1, &n, &off, a.data(), r.data(), p, steps, pivot, opts);
    if (rc) return rc;
    for (int64_t c = 0; c < n; c++) std::memcpy(R + c * ldr, &r[c * n], n * sizeof(double));
    return 0;
}

int skh_skewtrieigen_d_gpu(int64_t N, double* ev, double* vals, double* vr, double* vi, int variant)
{
    if (N < 0) return set_err(-1, "n < 0");
    if (N > 1 && !ev) return set_err(-2, "ev NULL");
    if (N > 0 && (!vals || !vr || !vi)) return set_err(-3, "output NULL");
    if (variant != 0 && variant != 1) return set_err(-6, "variant must be 0 or 1");
    const int rc = skh_internal_skewtrieigen(N, ev, vals, vr, vi, variant, rot_apply_gpu, symtri_gpu, nullptr);
    if (rc == 2) return set_err(SKH_ERR_NOCONV, "tridiagonal QR did not converge");
    return rc;
}

int skh_eigvals_range_d(int64_t n, const double* A, int64_t lda, int64_t il, int64_t iu, double vl, double vh,
                        int by_value, double* vals, int64_t* m, const skh_opts* opts)
{
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && !A) return set_err(-2, "A is NULL");
    if (lda < std::max<int64_t>(n, 1)) return set_err(-3, "lda < n");
    if (!vals || !m) return set_err(-9, "output NULL");
    *m = 0;
    if (n == 0) return 0;
    if (!by_value && (il < 1 || iu > n || il > iu + 1)) return set_err(-4, "irange out of 1:n");
    if (n == 1) {
        // src/eigen.jl:58,67: the single entry, unless (vl, vh) excludes it
        const double a = A[0];
        if (!by_value || (0.0 > vl && 0.0 < vh)) {
            vals[0] = a;
            *m = 1;
            return 0;
        }
    }
    std::vector<double> F((size_t)n * n), tau(n), eta(n);
    for (int64_t c = 0; c < n; c++) std::memcpy(&F[c * n], A + c * lda, n * sizeof(double));
    int64_t nz = 0;
    int rc = n > 1 ? skh_sytrd_d(n, F.data(), n, tau.data(), eta.data(), &nz, opts) : 0;
    if (rc) return rc;
    rc = skh_skewtrieigvals_range_d(n, eta.data(), il, iu, vl, vh, by_value, vals, m);
    if (rc == 2) return set_err(SKH_ERR_NOCONV, "tridiagonal QL did not converge");
    return rc;
}

// ----------------------------------------------------------------------------------------
// Matrix functions of a skew matrix (src/exp.jl:3-296) from its eigen-decomposition: the eigenvectors
// stay on the device, their columns are scaled by a function of the eigenvalues, the products
// V D V^T (real parts and imaginary parts separately, as the reference) / V D V^H run in rocBLAS,
// tan / cot / tanh / coth solve with rocSOLVER's LU, and the symmetrizations run in skh_fun.hip.
// ----------------------------------------------------------------------------------------
}  // extern "C"

namespace {
constexpr int FUN_SLOTS = 8;   // real case: EV_I, T1, T2, P1, EV_R, P2, and two for the complex cis output

bool fun_valid(int fn) { return fn >= SKH_FN_EXP && fn <= SKH_FN_COTH; }
bool fun_expish(int fn)
{
    return fn == SKH_FN_EXP || fn == SKH_FN_LOG || fn == SKH_FN_SINH || fn == SKH_FN_COSH || fn == SKH_FN_TANH ||
           fn == SKH_FN_COTH;
}
double fun_scale(int fn) { return fn == SKH_FN_TANH || fn == SKH_FN_COTH ? 2.0 : 1.0; }   // exp(2A)

// L <- LU(L), R <- L \ R (n x n, ld n); an exactly zero pivot is Julia's SingularException
int lu_solve_locked(int64_t n, double* L, double* R, bool cplx, hipStream_t st)
{
    rocblas_int* piv = (rocblas_int*)g_fpiv.get((size_t)(n + 1) * sizeof(rocblas_int));
    if (!piv) return set_err(SKH_ERR_HIP, "hipMalloc failed (pivots)");
    rocblas_int* dinfo = piv + n;
    const rocblas_int m = (rocblas_int)n;
    if (cplx)
        BLASCHK(rocsolver_zgetrf(g_blas, m, m, (rocblas_double_complex*)L, m, piv, dinfo));
    else
        BLASCHK(rocsolver_dgetrf(g_blas, m, m, L, m, piv, dinfo));
    rocblas_int info = 0;
    HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (info > 0) return set_err(SKH_ERR_SINGULAR, "SingularException(%d): the matrix to divide by is singular", info);
    if (cplx)
        BLASCHK(rocsolver_zgetrs(g_blas, rocblas_operation_none, m, m, (rocblas_double_complex*)L, m, piv,
                                 (rocblas_double_complex*)R, m));
    else
        BLASCHK(rocsolver_dgetrs(g_blas, rocblas_operation_none, m, m, L, m, piv, R, m));
    return 0;
}

// the rest of fn after E = exp(A) (or exp(2A)) in E, with W a free n x n work matrix
int fun_from_exp_locked(int64_t n, int fn, bool cplx, double* E, double* W, double* out1, hipStream_t st)
{
    const size_t bytes = (size_t)n * n * (cplx ? 2 : 1) * sizeof(double);
    if (fn == SKH_FN_SINH) HIPCHK((hipError_t)skh_internal_fun_sympart(n, E, cplx, 0, st));   // skewhermitian!(exp(A))
    if (fn == SKH_FN_COSH) HIPCHK((hipError_t)skh_internal_fun_sympart(n, E, cplx, 2, st));   // hermitian!(exp(A))
    if (fn == SKH_FN_TANH || fn == SKH_FN_COTH) {
        // tanh: skewhermitian!((E + I) \ (E - I)); coth: skewhermitian!((E - I) \ (E + I))  (src/exp.jl:270-284)
        const double sl = fn == SKH_FN_TANH ? 1.0 : -1.0;
        HIPCHK(hipMemcpyAsync(W, E, bytes, hipMemcpyDeviceToDevice, st));
        HIPCHK((hipError_t)skh_internal_fun_adddiag(n, W, cplx, sl, st));
        HIPCHK((hipError_t)skh_internal_fun_adddiag(n, E, cplx, -sl, st));
        int rc = lu_solve_locked(n, W, E, cplx, st);
        if (rc) return rc;
        HIPCHK((hipError_t)skh_internal_fun_sympart(n, E, cplx, 0, st));
    }
    HIPCHK(hipMemcpyAsync(out1, E, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// S (raw sin) and C (raw cos) to the outputs of sin / cos / sincos / tan / cot
int fun_from_sincos_locked(int64_t n, int fn, bool cplx, double* S, double* C, double* out1, double* out2,
                           hipStream_t st)
{
    const size_t bytes = (size_t)n * n * (cplx ? 2 : 1) * sizeof(double);
    const bool need_s = fn != SKH_FN_COS, need_c = fn != SKH_FN_SIN;
    if (need_s) HIPCHK((hipError_t)skh_internal_fun_sympart(n, S, cplx, 0, st));   // skewhermitian!(Sin)
    if (need_c) HIPCHK((hipError_t)skh_internal_fun_sympart(n, C, cplx, 1, st));   // Hermitian(Cos)
    double* res = fn == SKH_FN_COS ? C : S;
    if (fn == SKH_FN_TAN || fn == SKH_FN_COT) {
        // tan: skewhermitian!(C \ S); cot: skewhermitian!(S \ C)  (src/exp.jl:265-278)
        double* L = fn == SKH_FN_TAN ? C : S;
        res = fn == SKH_FN_TAN ? S : C;
        int rc = lu_solve_locked(n, L, res, cplx, st);
        if (rc) return rc;
        HIPCHK((hipError_t)skh_internal_fun_sympart(n, res, cplx, 0, st));
    }
    HIPCHK(hipMemcpyAsync(out1, res, bytes, hipMemcpyDeviceToHost, st));
    if (fn == SKH_FN_SINCOS) HIPCHK(hipMemcpyAsync(out2, C, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// real case: eigenvalues i vals[j], eigenvectors Vr + i Vi in slots EV_R / EV_I of base
int fun_real_locked(int64_t n, int fn, const double* vals, double* base, double* out1, double* out2)
{
    hipStream_t st = g_stream;
    const size_t nn = (size_t)n * n;
    double* Vr = base + EV_R * nn;
    double* Vi = base + EV_I * nn;
    double* T1 = base + 1 * nn;
    double* T2 = base + 2 * nn;
    double* P1 = base + 3 * nn;
    double* P2 = base + 5 * nn;
    double* W = base + 6 * nn;   // slots 6 and 7: work matrix, or the complex cis output
    std::vector<double> a(n), b(n);
    const bool ex = fun_expish(fn);
    for (int64_t j = 0; j < n; j++) {
        if (fn == SKH_FN_LOG) {   // log|lambda| and sign(lambda) pi/2 (src/exp.jl:66-69)
            if (vals[j] == 0.0) return set_err(SKH_ERR_DOMAIN, "Logarithm of a singular matrix doesn't exist");
            a[j] = std::log(std::fabs(vals[j]));
            b[j] = (vals[j] > 0.0 ? 1.0 : -1.0) * M_PI / 2;
        } else if (ex) {          // sincos(lambda) (src/exp.jl:19-21)
            a[j] = std::cos(vals[j]);
            b[j] = std::sin(vals[j]);
        } else {                  // exp(-lambda): the eigenvalues of iA (src/exp.jl:111, :134, :177)
            a[j] = std::exp(-vals[j]);
        }
    }
    double* dv = (double*)g_fvec.get((size_t)2 * n * sizeof(double));
    if (!dv) return set_err(SKH_ERR_HIP, "hipMalloc failed (eigenvalue vectors)");
    HIPCHK(hipMemcpyAsync(dv, a.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dv + n, b.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));   // a and b end with this function, on every return path
    HIPCHK((hipError_t)skh_internal_fun_colmix(n, Vr, Vi, dv, dv + n, ex ? 0 : 1, T1, T2, st));
    int rc = blas_on(st);
    if (rc) return rc;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    const int m = (int)n;
    const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
    if (ex) {
        // (Qr C - Qim S) Qr^T + (Qr S + Qim C) Qim^T  (src/exp.jl:26-33; log: :76-83)
        BLASCHK(rocblas_dgemm(g_blas, N, T, m, m, m, &one, T1, m, Vr, m, &zero, P1, m));
        BLASCHK(rocblas_dgemm(g_blas, N, T, m, m, m, &one, T2, m, Vi, m, &one, P1, m));
        if (fn == SKH_FN_EXP || fn == SKH_FN_LOG) {
            HIPCHK(hipMemcpyAsync(out1, P1, nn * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            return 0;
        }
        return fun_from_exp_locked(n, fn, false, P1, P2, out1, st);
    }
    // E = exp(-lambda): Cos = Qr E Qr^T + Qim E Qim^T, Sin = Qim E Qr^T - Qr E Qim^T (src/exp.jl:168-174,
    // :219-225, :234-241); cis = Cos + i Sin as a Hermitian matrix (src/exp.jl:98-109)
    const bool need_s = fn != SKH_FN_COS, need_c = fn != SKH_FN_SIN;
    if (need_c) {
        BLASCHK(rocblas_dgemm(g_blas, N, T, m, m, m, &one, T1, m, Vr, m, &zero, P2, m));
        BLASCHK(rocblas_dgemm(g_blas, N, T, m, m, m, &one, T2, m, Vi, m, &one, P2, m));
    }
    if (need_s) {
        BLASCHK(rocblas_dgemm(g_blas, N, T, m, m, m, &one, T2, m, Vr, m, &zero, P1, m));
        BLASCHK(rocblas_dgemm(g_blas, N, T, m, m, m, &mone, T1, m, Vi, m, &one, P1, m));
    }
    if (fn == SKH_FN_CIS) {
        HIPCHK((hipError_t)skh_internal_fun_cis_out(n, P2, P1, W, st));
        HIPCHK(hipMemcpyAsync(out1, W, 2 * nn * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    return fun_from_sincos_locked(n, fn, false, P1, P2, out1, out2, st);
}

// complex case: eigenvalues i vals[j], eigenvectors V (device, ld n)
int fun_complex_locked(int64_t n, int fn, const double* vals, const double2* V, double* out1, double* out2)
{
    hipStream_t st = g_stream;
    const size_t nn = (size_t)n * n;
    double2* Z = (double2*)g_fz.get(3 * nn * sizeof(double2));
    double2* dv = (double2*)g_fvec.get((size_t)2 * n * sizeof(double2));
    if (!Z || !dv) return set_err(SKH_ERR_HIP, "hipMalloc failed (complex work matrices)");
    double2* Tm = Z;
    double2* P1 = Z + nn;
    double2* P2 = Z + 2 * nn;
    const bool ex = fun_expish(fn);
    std::vector<double2> d1(n), d2(n);
    for (int64_t j = 0; j < n; j++) {
        if (fn == SKH_FN_LOG) {   // log.(values), values = i vals (src/exp.jl:87-90)
            const std::complex<double> z = std::log(std::complex<double>(0.0, vals[j]));
            d1[j] = make_double2(z.real(), z.imag());
        } else if (ex) {          // exp.(values) (src/exp.jl:39-41)
            d1[j] = make_double2(std::cos(vals[j]), std::sin(vals[j]));
        } else {                  // exp(-imag(values)), exp(imag(values)) (src/exp.jl:113-118, :145-151)
            d1[j] = make_double2(std::exp(-vals[j]), 0.0);
            d2[j] = make_double2(std::exp(vals[j]), 0.0);
        }
    }
    HIPCHK(hipMemcpyAsync(dv, d1.data(), n * sizeof(double2), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dv + n, d2.data(), n * sizeof(double2), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));   // d1 and d2 end with this function, on every return path
    int rc = blas_on(st);
    if (rc) return rc;
    const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
    const int m = (int)n;
    const rocblas_operation N = rocblas_operation_none, C = rocblas_operation_conjugate_transpose;
    auto rz = [](const double2* p) { return (const rocblas_double_complex*)p; };
    // P = (V diag(d)) V^H
    auto vdvh = [&](const double2* d, double2* P) -> int {
        HIPCHK((hipError_t)skh_internal_fun_zcolscale(n, (const double*)V, (const double*)d, (double*)Tm, st));
        BLASCHK(rocblas_zgemm(g_blas, N, C, m, m, m, &one, rz(Tm), m, rz(V), m, &zero, (rocblas_double_complex*)P, m));
        return 0;
    };
    rc = vdvh(dv, P1);
    if (rc) return rc;
    if (ex) {
        if (fn == SKH_FN_EXP || fn == SKH_FN_LOG) {
            HIPCHK(hipMemcpyAsync(out1, P1, nn * sizeof(double2), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            return 0;
        }
        return fun_from_exp_locked(n, fn, true, (double*)P1, (double*)P2, out1, st);
    }
    if (fn == SKH_FN_CIS) {   // Hermitian(V e^{-mu} V^H)
        HIPCHK((hipError_t)skh_internal_fun_sympart(n, (double*)P1, 1, 1, st));
        HIPCHK(hipMemcpyAsync(out1, P1, nn * sizeof(double2), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    rc = vdvh(dv + n, P2);
    if (rc) return rc;
    const int mode = fn == SKH_FN_COS ? 0 : fn == SKH_FN_SIN ? 1 : 2;
    HIPCHK((hipError_t)skh_internal_fun_zsincos(n, (double*)P1, (double*)P2, mode, st));   // P1 <- Sin, P2 <- Cos
    return fun_from_sincos_locked(n, fn, true, (double*)P1, (double*)P2, out1, out2, st);
}

int fun_args(int64_t n, const void* A, int64_t lda, int fn, const double* out1, const double* out2)
{
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 0 && !A) return set_err(-2, "A is NULL");
    if (lda < std::max<int64_t>(n, 1)) return set_err(-3, "lda < n");
    if (!fun_valid(fn)) return set_err(-4, "unknown matrix function %d", fn);
    if (n > 0 && !out1) return set_err(-5, "out1 is NULL");
    if (n > 0 && fn == SKH_FN_SINCOS && !out2) return set_err(-6, "out2 is NULL (sincos has two outputs)");
    if (n > 46340) return set_err(SKH_ERR_UNSUPPORTED, "n too large for the BLAS interface");
    return 0;
}
}  // namespace

extern "C" {

int skh_skewfun_d(int64_t n, const double* A, int64_t lda, int fn, double* out1, double* out2, const skh_opts* opts)
{
    int rc = fun_args(n, A, lda, fn, out1, out2);
    if (rc || n == 0) return rc;
    if (fn == SKH_FN_LOG && n % 2) return set_err(SKH_ERR_DOMAIN, "Logarithm of a singular matrix doesn't exist");
    std::vector<double> vals(n);
    std::unique_lock<std::mutex> lk;
    double* base = nullptr;
    rc = eigen_d_dev(n, A, lda, fun_scale(fn), vals.data(), opts, FUN_SLOTS, lk, &base);
    if (rc) return rc;
    StreamDrain drain(g_stream);   // out1 / out2 are written asynchronously
    return fun_real_locked(n, fn, vals.data(), base, out1, out2);
}

int skh_skewfun_tri_d(int64_t n, const double* ev, int fn, double* out1, double* out2)
{
    int rc = fun_args(n, n > 1 ? (const void*)ev : (const void*)out1, std::max<int64_t>(n, 1), fn, out1, out2);
    if (rc || n == 0) return rc;
    if (fn == SKH_FN_LOG && n % 2) return set_err(SKH_ERR_DOMAIN, "Logarithm of a singular matrix doesn't exist");
    // eigen!(SkewHermTridiagonal) = skewtrieigen_merged! (src/tridiag.jl:506); rotations on the GPU from n = 256
    const size_t nn = (size_t)n * n;
    std::vector<double> e(std::max<int64_t>(n - 1, 1)), vals(n), hr(nn), hi(nn);
    const double sc = fun_scale(fn);
    for (int64_t k = 0; k + 1 < n; k++) e[k] = ev[k] * sc;
    if (n == 1) {
        vals[0] = 0.0;
        hr[0] = 1.0;
    } else {
        rc = skh_internal_skewtrieigen(n, e.data(), vals.data(), hr.data(), hi.data(), 1,
                                       n >= ROT_GPU_MIN ? rot_apply_gpu : nullptr, symtri_gpu, nullptr);
        if (rc == 2) return set_err(SKH_ERR_NOCONV, "tridiagonal QR did not converge");
        if (rc) return rc < 0 ? set_err(rc, "skewtrieigen failed (%d)", rc) : rc;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    rc = ensure_init_locked(g_device);
    if (rc) return rc;
    StreamDrain drain(g_stream);   // hr and hi are copied asynchronously; fun_real_locked may return early
    double* base = (double*)g_q.get((size_t)FUN_SLOTS * nn * sizeof(double));
    if (!base) return set_err(SKH_ERR_HIP, "hipMalloc failed for n = %lld", (long long)n);
    HIPCHK(hipMemcpyAsync(base + EV_R * nn, hr.data(), nn * sizeof(double), hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipMemcpyAsync(base + EV_I * nn, hi.data(), nn * sizeof(double), hipMemcpyHostToDevice, g_stream));
    return fun_real_locked(n, fn, vals.data(), base, out1, out2);
}

int skh_skewfun_z(int64_t n, const double* A, int64_t lda, int fn, double* out1, double* out2, const skh_opts* opts)
{
    (void)opts;
    int rc = fun_args(n, A, lda, fn, out1, out2);
    if (rc || n == 0) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    std::vector<double> vals(n);
    double2* dV = nullptr;
    rc = ensure_init_locked(g_device);
    if (rc) return rc;
    StreamDrain drain(g_stream);   // the caller's arrays are copied asynchronously
    rc = eigen_z_dev_locked(n, A, lda, fun_scale(fn), vals.data(), &dV);
    if (rc) return rc;
    return fun_complex_locked(n, fn, vals.data(), dV, out1, out2);
}

int skh_skewtrieigvals_z(int64_t n, const double* ev, const double* dvim, double* mu)
{
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 1 && !ev) return set_err(-2, "ev is NULL");
    if (n > 0 && !mu) return set_err(-4, "mu is NULL");
    if (n == 0) return 0;
    std::vector<double> e(std::max<int64_t>(n - 1, 1)), q(2 * n);
    skh_internal_to_symtri_z(n, ev, dvim, mu, e.data(), q.data());
    const int rc = skh_internal_symtri_eigvals(n, mu, e.data());
    if (rc == 2) return set_err(SKH_ERR_NOCONV, "symmetric tridiagonal QL did not converge");
    return rc;
}

int skh_skewtrieigen_z(int64_t n, const double* ev, const double* dvim, double* mu, double* V)
{
    if (n < 0) return set_err(-1, "n < 0");
    if (n > 1 && !ev) return set_err(-2, "ev is NULL");
    if (n > 0 && (!mu || !V)) return set_err(-4, "output is NULL");
    if (n == 0) return 0;
    std::vector<double> d(n), e(std::max<int64_t>(n - 1, 1)), q(2 * n), Z((size_t)n * n);
    skh_internal_to_symtri_z(n, ev, dvim, d.data(), e.data(), q.data());
    const int rc = symtri_gpu_diag(n, d.data(), e.data(), mu, Z.data());
    if (rc) return rc;
    // Vec = Diagonal(Q) * vectors (src/tridiag.jl:511-513)
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < n; i++) {
            const double z = Z[i + j * n];
            V[2 * (i + j * n)] = q[2 * i] * z;
            V[2 * (i + j * n) + 1] = q[2 * i + 1] * z;
        }
    return 0;
}

}  // extern "C"
