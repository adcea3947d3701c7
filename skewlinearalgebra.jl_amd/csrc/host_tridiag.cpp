// host_tridiag.cpp -- host implicit QR for the real skew-symmetric tridiagonal (kept on the CPU:
// O(n^2), sequential).  Same deflation rules, shifts and output order as the reference
// (src/skeweigen.jl), so its eigenvalue order and eigenvector conventions match it:
//   * Ward-Gray elimination of the zero eigenvalue for odd n     (reducetozero, :11-47)
//   * shifted Givens bulge chase, shift ev[n-1]^2                 (implicitstep_*, :73-103, :158-193)
//   * absolute deflation 10*eps, relative 2x2 split, stagnation detection -> symmetric tridiagonal
//     fallback (:105-156, :195-320, :322-457, :463-490).
// The fallback solver here is an implicit QL with Wilkinson shifts (the classical tql2 scheme),
// sorted ascending, where the reference calls LAPACK stev/stegr.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

// Reference arithmetic rounds every product and sum (no fused multiply-add contraction).
#pragma clang fp contract(off)

namespace {

constexpr double kEps = 2.220446049250313e-16;

struct Rot {
    double c, s;
};

inline Rot givens(double a, double b)
{
    const double nm = std::hypot(a, b);
    if (nm == 0.0) return {1.0, 0.0};
    return {a / nm, b / nm};
}

// Rotations of the QR sweeps logged for a device to apply (instead of applying them on the host):
// [0] the chain on Qodd, [1] the chain on Qeven; rotation t acts on columns (k - 1, k), 0-based.
struct RotLog {
    std::vector<int32_t> k[2];
    std::vector<double> cs[2];   // c, s interleaved
};

// 1-based view over a contiguous array
struct V1 {
    double* p;
    double& operator()(int64_t k) const { return p[k - 1]; }
};

void zero_eigenvalue_sweep(V1 ev, V1 G, int64_t n)
{
    if (n == 0) return;
    double bulge = 0.0;
    if (n > 2) {
        const double al = ev(n - 2), be = ev(n - 1), ga = ev(n);
        const Rot r = givens(-be, ga);
        G(n - 1) = r.c;
        G(n) = -r.s;
        ev(n - 2) *= r.c;
        ev(n - 1) = r.c * be - r.s * ga;
        ev(n) = 0.0;
        bulge = -r.s * al;
        for (int64_t i = n - 2; i >= 4; i -= 2) {
            const double a2 = ev(i - 2), b2 = ev(i - 1);
            const Rot q = givens(-b2, bulge);
            G(i - 1) = q.c;
            G(i) = -q.s;
            ev(i - 2) *= q.c;
            ev(i - 1) = q.c * b2 - q.s * bulge;
            bulge = -q.s * a2;
        }
    }
    const double be = (n == 2) ? ev(2) : bulge;
    const double al = ev(1);
    const Rot r = givens(-al, be);
    G(1) = r.c;
    G(2) = -r.s;
    ev(1) = r.c * al - r.s * be;
    if (n == 2) ev(2) = r.s * al + be * r.c;
}

// One shifted sweep over ev[start..n]; Q-rotations applied when qodd/qeven are given.
// novec semantics (early return on an interior split) when early_return is set.
int64_t qr_sweep(V1 ev, int64_t n, int64_t start, double tol, bool early_return, double* qeven, int64_t lde,
                 double* qodd, int64_t ldo, int64_t nrows, RotLog* log = nullptr)
{
    double bulge = 0.0;
    const double shift = ev(n) * ev(n);
    for (int64_t i = start; i <= n - 1; i++) {
        const double al = (i > start) ? ev(i - 1) : 0.0;
        const double be = ev(i), ga = ev(i + 1);
        const double x1 = -al * al - be * be + shift;
        const double x2 = -al * bulge + be * ga;
        const Rot r = (i > start) ? givens(al, bulge) : givens(x1, x2);
        if (i > 1) ev(i - 1) = r.c * al + r.s * bulge;
        ev(i) = r.c * be - r.s * ga;
        ev(i + 1) = r.s * be + r.c * ga;
        if (i < n - 1) {
            const double ze = ev(i + 2);
            ev(i + 2) *= r.c;
            bulge = r.s * ze;
            if (std::fabs(bulge) < tol && std::fabs(ev(i)) < tol) {
                if (early_return) return i + 1;
                start = i + 1;
            }
        }
        if (log) {
            const int w = (i % 2 == 1) ? 0 : 1;
            log->k[w].push_back((int32_t)((i + 1) / 2));
            log->cs[w].push_back(r.c);
            log->cs[w].push_back(r.s);
        } else if (qodd) {
            double* Q = (i % 2 == 1) ? qodd : qeven;
            const int64_t ld = (i % 2 == 1) ? ldo : lde;
            const int64_t k = (i + 1) / 2;
            double* ca = Q + (k - 1) * ld;
            double* cb = Q + k * ld;
            for (int64_t j = 0; j < nrows; j++) {
                const double sg = ca[j], om = cb[j];
                ca[j] = r.c * sg + r.s * om;
                cb[j] = -r.s * sg + r.c * om;
            }
        }
    }
    return start;
}

// Symmetric tridiagonal (diagonal d, off-diagonal e) eigenpairs by implicit QL with Wilkinson
// shifts; eigenvalues ascending; Z (n x n, column-major) receives eigenvectors when non-null.
bool symtri_ql(int64_t n, std::vector<double>& d, std::vector<double> e, double* Z)
{
    if (Z) {
        std::fill(Z, Z + n * n, 0.0);
        for (int64_t k = 0; k < n; k++) Z[k + k * n] = 1.0;
    }
    e.resize(n, 0.0);
    for (int64_t l = 0; l < n; l++) {
        int iter = 0;
        int64_t m;
        do {
            for (m = l; m < n - 1; m++) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= kEps * dd) break;
            }
            if (m != l) {
                if (iter++ == 60) return false;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
                double s = 1.0, c = 1.0, p = 0.0;
                int64_t i;
                for (i = m - 1; i >= l; i--) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    if (Z) {
                        for (int64_t k = 0; k < n; k++) {
                            f = Z[k + (i + 1) * n];
                            Z[k + (i + 1) * n] = s * Z[k + i * n] + c * f;
                            Z[k + i * n] = c * Z[k + i * n] - s * f;
                        }
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    // ascending sort (with vectors)
    std::vector<int64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return d[a] < d[b]; });
    std::vector<double> d2(n);
    for (int64_t k = 0; k < n; k++) d2[k] = d[idx[k]];
    d = d2;
    if (Z) {
        std::vector<double> Z2(Z, Z + n * n);
        for (int64_t k = 0; k < n; k++) std::memcpy(Z + k * n, Z2.data() + idx[k] * n, sizeof(double) * n);
    }
    return true;
}

}  // namespace

extern "C" int skh_skewtrieigvals_d(int64_t N, double* evp, double* vals)
{
    if (N < 0) return -1;
    if (N > 1 && !evp) return -2;
    if (N > 0 && !vals) return -3;
    std::fill(vals, vals + N, 0.0);
    V1 ev{evp};
    int64_t n = N;
    std::vector<double> Gv;
    if (n % 2 == 1) {
        n -= 1;
        Gv.assign(std::max<int64_t>(n, 1), 0.0);
        zero_eigenvalue_sweep(ev, V1{Gv.data()}, n);
    }
    const double tol = kEps * 10.0;
    const int64_t max_iter = 30 * n;
    int64_t iter = 0, count_static = 0, start = 1;
    double mem = 1.0;
    while (n > 2 && iter < max_iter) {
        start = qr_sweep(ev, n - 1, start, kEps, true, nullptr, 0, nullptr, 0, 0);
        while (n > 2 && std::fabs(ev(n - 1)) <= tol) {
            vals[n - 1] = 0.0;
            n -= 1;
        }
        while (n > 2 && std::fabs(ev(n - 2)) <= tol * std::fabs(ev(n - 1))) {
            vals[n - 2] = ev(n - 1);
            vals[n - 1] = -ev(n - 1);
            n -= 2;
        }
        if (start > n - 2) start = 1;
        if (n > 1 && std::fabs(mem - ev(n - 1)) < 0.0001 * std::fabs(ev(n - 1))) {
            if (++count_static > 4) {
                std::vector<double> d(n, 0.0), e(evp, evp + (n - 1));
                if (!symtri_ql(n, d, e, nullptr)) return 2;
                for (int64_t k = 0; k < n; k++) vals[k] = -d[k];
                return 0;
            }
        } else {
            count_static = 0;
        }
        mem = (n > 1) ? ev(n - 1) : 0.0;
        iter += 1;
    }
    if (n == 2) {
        vals[0] = ev(1);
        vals[1] = -ev(1);
        return 0;
    }
    if (n == 1 || n == 0) {
        if (n == 1) vals[0] = 0.0;
        return 0;
    }
    return 2;
}

// Qo (no x no) and Qe (ne x ne), column-major; the rotations act on rows 0 .. nrows-1 only (the
// last row of Qo for odd N belongs to the zero eigenvalue and is left to getoddvectors).
typedef int (*skh_rot_apply_fn)(void* ctx, int64_t nrows, const int32_t* ko, const double* cso, int64_t nro,
                                double* Qo, int64_t no, const int32_t* ke, const double* cse, int64_t nre, double* Qe,
                                int64_t ne);
typedef int (*skh_symtri_fn)(void* ctx, int64_t n, const double* e, double* d, double* Z);

// skewtrieigen_divided! / merged!.  With `apply` set, the sweeps' rotations are logged and handed
// to it once the iteration has converged (it must leave Qo and Qe as the host would have them);
// getoddvectors and the assembly of vr, vi then run here as usual.  With `symtri` set, the
// fallback taken when the sweeps stagnate (the reference's eigen!(SymTridiagonal), LAPACK there)
// calls it instead of the host QL, whose eigenvector accumulation costs O(n^3) scalar work.
extern "C" int skh_internal_skewtrieigen(int64_t N, double* evp, double* vals, double* vr, double* vi, int variant,
                                         skh_rot_apply_fn apply, skh_symtri_fn symtri, void* ctx)
{
    if (N < 0) return -1;
    if (N > 1 && !evp) return -2;
    if (N > 0 && (!vals || !vr || !vi)) return -3;
    if (variant != 0 && variant != 1) return -6;
    std::vector<double> backup(evp, evp + std::max<int64_t>(N - 1, 0));
    std::fill(vals, vals + N, 0.0);
    std::fill(vr, vr + N * N, 0.0);
    std::fill(vi, vi + N * N, 0.0);
    const int64_t no = (N + 1) / 2, ne = N / 2;
    std::vector<double> Qo(std::max<int64_t>(no * no, 1), 0.0), Qe(std::max<int64_t>(ne * ne, 1), 0.0);
    for (int64_t k = 0; k < no; k++) Qo[k + k * no] = 1.0;
    for (int64_t k = 0; k < ne; k++) Qe[k + k * ne] = 1.0;
    V1 ev{evp};
    int64_t n = N;
    std::vector<double> Gv;
    if (n % 2 == 1) {
        n -= 1;
        Gv.assign(std::max<int64_t>(n, 1), 0.0);
        zero_eigenvalue_sweep(ev, V1{Gv.data()}, n);
    }
    const double tol = kEps * 10.0;
    const int64_t max_iter = 30 * n, halfN = n / 2;
    int64_t iter = 0, count_static = 0, start = 1;
    double mem = 1.0;
    RotLog log;
    RotLog* lp = apply ? &log : nullptr;
    while (n > 2 && iter < max_iter) {
        start = qr_sweep(ev, n - 1, start, tol, false, Qe.data(), ne, Qo.data(), no, halfN, lp);
        while (n > 2 && std::fabs(ev(n - 1)) <= tol) {
            vals[n - 1] = 0.0;
            n -= 1;
        }
        while (n > 2 && std::fabs(ev(n - 2)) <= tol * std::fabs(ev(n - 1))) {
            vals[n - 2] = ev(n - 1);
            vals[n - 1] = -ev(n - 1);
            n -= 2;
        }
        // divided variant only (src/skeweigen.jl:355-358); mem starts at 1 as in the reference
        if (variant == 0 && n > 2 && std::fabs(ev(n - 1) - mem) < tol) {
            vals[n - 2] = ev(n - 1);
            vals[n - 1] = -ev(n - 1);
            n -= 2;
        }
        if (start > n - 2) start = 1;
        if (n > 1 && std::fabs(mem - ev(n - 1)) < 0.0001 * std::fabs(ev(n - 1))) {
            if (++count_static > 4) {
                // symmetric tridiagonal fallback on the original ev (src/skeweigen.jl:470-490)
                std::vector<double> d(N, 0.0), Z(N * N);
                if (symtri) {
                    const int rc = symtri(ctx, N, backup.data(), d.data(), Z.data());
                    if (rc) return rc;
                } else if (!symtri_ql(N, d, backup, Z.data())) {
                    return 2;
                }
                std::fill(vr, vr + N * N, 0.0);
                std::fill(vi, vi + N * N, 0.0);
                for (int64_t k = 0; k < N; k++) vals[k] = -d[k];
                double c = 1.0;
                for (int64_t j = 0; j < N; j++) {
                    c = 1.0;
                    for (int64_t i = 0; i + 1 < N; i += 2) {
                        vr[i + j * N] = Z[i + j * N] * c;
                        vi[(i + 1) + j * N] = Z[(i + 1) + j * N] * c;
                        c *= -1.0;
                    }
                }
                if (N % 2 == 1)
                    for (int64_t j = 0; j < N; j++) vr[(N - 1) + j * N] = Z[(N - 1) + j * N] * c;
                return 0;
            }
        } else {
            count_static = 0;
        }
        mem = (n > 1) ? ev(n - 1) : 0.0;
        iter += 1;
    }
    if (n == 0) {
        if (variant == 1)
            for (int64_t k = 0; k < no; k++) vr[k + k * N] = 1.0;
        return 0;
    }
    if (n > 2) return 2;
    if (n == 2) {
        vals[0] = ev(1);
        vals[1] = -ev(1);
    } else {
        vals[0] = 0.0;
    }
    if (apply) {
        const int rc = apply(ctx, halfN, log.k[0].data(), log.cs[0].data(), (int64_t)log.k[0].size(), Qo.data(), no,
                             log.k[1].data(), log.cs[1].data(), (int64_t)log.k[1].size(), Qe.data(), ne);
        if (rc) return rc;
    }
    if (N % 2 == 1) {
        // getoddvectors (src/skeweigen.jl:49-62)
        const int64_t m = N - 1, nn = m / 2 + 1;
        for (int64_t i = 1; i <= m - 1; i += 2) {
            const double c = Gv[i - 1], s = Gv[i];
            const int64_t ii = (i + 1) / 2;
            for (int64_t j = 1; j <= nn; j++) {
                double& a = Qo[(ii - 1) + (j - 1) * no];
                double& b = Qo[(nn - 1) + (j - 1) * no];
                const double sg = a, om = b;
                a = c * sg + s * om;
                b = -s * sg + c * om;
            }
        }
    }
    const double s2 = 1.0 / std::sqrt(2.0);
    auto QO = [&](int64_t r, int64_t c) { return Qo[(r - 1) + (c - 1) * no]; };
    auto QE = [&](int64_t r, int64_t c) { return Qe[(r - 1) + (c - 1) * ne]; };
    auto VR = [&](int64_t r, int64_t c) -> double& { return vr[(r - 1) + (c - 1) * N]; };
    auto VI = [&](int64_t r, int64_t c) -> double& { return vi[(r - 1) + (c - 1) * N]; };
    int64_t i = 1, countzeros = 0;
    while (i <= N) {
        const int64_t ii = (i + 1) / 2;
        if (vals[i - 1] == 0.0) {
            if (countzeros % 2 == 0) {
                for (int64_t j = 1; j <= N - 1; j += 2) VR(j, i) = QO((j + 1) / 2, ii);
                if (N % 2 == 1) VR(N, i) = QO(halfN + 1, ii);
            } else {
                for (int64_t j = 1; j <= N - 1; j += 2) VR(j + 1, i) = QE((j + 1) / 2, ii);
            }
            countzeros += 1;
            i += 1;
        } else {
            if (countzeros % 2 == 1) {
                for (int64_t j = 1; j <= N - 1; j += 2) {
                    const int64_t jj = (j + 1) / 2;
                    VI(j, i) = -s2 * QO(jj, ii + 1);
                    VR(j + 1, i) = s2 * QE(jj, ii);
                    VR(j, i + 1) = -s2 * QO(jj, ii + 1);
                    VI(j + 1, i + 1) = s2 * QE(jj, ii);
                }
                if (N % 2 == 1) {
                    if (variant == 0) {   // divided stores these two in the other parts (:423-424)
                        VR(N, i) = -s2 * QO(halfN + 1, ii + 1);
                        VI(N, i + 1) = -s2 * QO(halfN + 1, ii + 1);
                    } else {
                        VI(N, i) = -s2 * QO(halfN + 1, ii + 1);
                        VR(N, i + 1) = -s2 * QO(halfN + 1, ii + 1);
                    }
                }
            } else {
                for (int64_t j = 1; j <= N - 1; j += 2) {
                    const int64_t jj = (j + 1) / 2;
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
        for (int64_t j = 1; j <= N; j += 2) VR(j, N) = QO((j + 1) / 2, halfN + 1);
    return 0;
}

// to_symtridiagonal (src/tridiag.jl:558-583) of the complex SkewHermTridiagonal(ev, dvim): V = i ev, made
// real by diagonal phases Q (Q[0] = 1, Q[i+1] = V[i] / |V[i]|, V[i] = |V[i]|, V[i+1] *= Q[i+1]), so that
// i A = Q S Q^H with S = SymTridiagonal(-dvim (0 without dvim), real(V)).  ev: n - 1 interleaved complex
// values; d, e: n and n - 1 reals out; q: n interleaved complex values out.
extern "C" void skh_internal_to_symtri_z(int64_t n, const double* ev, const double* dvim, double* d, double* e,
                                         double* q)
{
    for (int64_t i = 0; i < n; i++) d[i] = dvim ? -dvim[i] : 0.0;
    if (n <= 0) return;
    q[0] = 1.0;
    q[1] = 0.0;
    if (n == 1) return;
    std::vector<double> vr(n - 1), vi(n - 1);
    for (int64_t k = 0; k + 1 < n; k++) {   // V .*= 1im
        vr[k] = -ev[2 * k + 1];
        vi[k] = ev[2 * k];
    }
    for (int64_t i = 0; i + 1 < n; i++) {
        const double nm = std::hypot(vr[i], vi[i]);
        const double qr = vr[i] / nm, qi = vi[i] / nm;
        q[2 * (i + 1)] = qr;
        q[2 * (i + 1) + 1] = qi;
        vr[i] = nm;
        vi[i] = 0.0;
        if (i + 2 < n) {   // V[i+1] *= Q[i+1]
            const double a = vr[i + 1], b = vi[i + 1];
            vr[i + 1] = a * qr - b * qi;
            vi[i + 1] = a * qi + b * qr;
        }
    }
    for (int64_t k = 0; k + 1 < n; k++) e[k] = vr[k];
}

extern "C" int skh_internal_symtri_eigvals(int64_t n, double* d, const double* e)
{
    if (n <= 0) return 0;
    std::vector<double> dv(d, d + n), ev(e, e + (n > 1 ? n - 1 : 0));
    if (!symtri_ql(n, dv, ev, nullptr)) return 2;
    std::copy(dv.begin(), dv.end(), d);
    return 0;
}

// Eigenvalues of a real SkewHermTridiagonal(ev) by index or value range, as the reference's
// skewtrieigvals!(S, irange) / (S, vl, vh) (src/tridiag.jl:525-537): the eigenvalues of the
// symmetric tridiagonal H = SymTridiagonal(0, ev), ascending, either with 1-based indices
// il..iu or in (-vh, -vl] (LAPACK stebz's half-open interval), each negated.  vals receives
// *m values (room for n).
extern "C" int skh_skewtrieigvals_range_d(int64_t N, const double* ev, int64_t il, int64_t iu, double vl, double vh,
                                          int by_value, double* vals, int64_t* m)
{
    if (N < 0) return -1;
    if (N > 1 && !ev) return -2;
    if (!vals || !m) return -8;
    *m = 0;
    if (N == 0) return 0;
    if (!by_value && (il < 1 || iu > N || il > iu + 1)) return -3;
    std::vector<double> d(N, 0.0);
    if (N > 1) {
        std::vector<double> e(ev, ev + (N - 1));
        if (!symtri_ql(N, d, e, nullptr)) return 2;
    }
    int64_t k = 0;
    for (int64_t i = 0; i < N; i++) {
        const bool take = by_value ? (d[i] > -vh && d[i] <= -vl) : (i + 1 >= il && i + 1 <= iu);
        if (take) vals[k++] = -d[i];
    }
    *m = k;
    return 0;
}

extern "C" int skh_skewtrieigen_d(int64_t N, double* evp, double* vals, double* vr, double* vi, int variant)
{
    return skh_internal_skewtrieigen(N, evp, vals, vr, vi, variant, nullptr, nullptr, nullptr);
}
