// esc_capi.cpp -- implementation of the C ABI declared in include/esc.h.
//
// Host orchestration only: device tables, workspaces, the Lanczos loop and the
// FP64 small solves.  All vector work runs in the HIP kernels of
// esc_kernels.hip; nothing here falls back to a CPU path.
#include "esc_ctx.h"
#include "esc_tridiag.h"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

using namespace esc_impl;

namespace esc_impl {
thread_local std::string g_err;
}  // namespace esc_impl

namespace esc {
void set_error(const char *msg) { esc_impl::g_err = msg; }
void clear_error() { esc_impl::g_err.clear(); }
}  // namespace esc

namespace esc_impl {

// ------------------------------------------------------------------------
// FP64 symmetric tridiagonal eigensolver: implicit QL with Wilkinson-type
// shifts (Golub & Van Loan, Matrix Computations, sec. 8.3), eigenvectors
// accumulated by Givens rotations, results sorted ascending.
// ------------------------------------------------------------------------
int tridiag_ql(int n, const double *a, const double *b, double *evals, double *Z) {
    std::vector<double> d(a, a + n), e(n, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = b[i + 1];
    const bool vec = Z != nullptr;
    std::vector<double> V;
    if (vec) {
        V.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;   // column-major, V[col*n + row]
    }
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= std::numeric_limits<double>::epsilon() * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) return -1;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool early = false;
            for (; i >= l; --i) {
                const double f = s * e[i], bb = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * bb;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - bb;
                if (vec) {
                    double *ci = &V[(size_t)i * n], *cj = &V[(size_t)(i + 1) * n];
                    for (int k = 0; k < n; ++k) {
                        const double t = cj[k];
                        cj[k] = s * ci[k] + c * t;
                        ci[k] = c * ci[k] - s * t;
                    }
                }
            }
            if (early) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    for (int i = 0; i < n; ++i) evals[i] = d[idx[i]];
    if (vec)
        for (int j = 0; j < n; ++j) std::memcpy(Z + (size_t)j * n, &V[(size_t)idx[j] * n], sizeof(double) * n);
    return 0;
}

// All eigenpairs of T (esc_tridiag.h: Sturm bisection, inverse iteration
// orthogonalised within clusters of close eigenvalues), on the device (one
// workgroup) or the host; both compute the same bits.  Z column-major,
// column j = eigenvector j.
void tridiag_eigenpairs(int n, const double *a, const double *b, double *ev, double *Z, bool on_device,
                        hipStream_t s) {
    if (n < 1 || n > esc::kMaxTridiag) throw EscError(ESC_ERR_ARG, "tridiagonal size must be in [1, 4096]");
    if (!on_device) {
        double lo, hi, piv;
        if (!esc::td_range(n, a, b, &lo, &hi, &piv)) throw EscError(ESC_ERR_NUMERIC, "non-finite tridiagonal matrix");
        for (int k = 0; k < n; ++k) ev[k] = n == 1 ? a[0] : esc::td_kth(n, a, b, k, lo, hi, piv);
        const double tn = esc::td_eig_norm(n, a, b);
        std::vector<double> work(5 * (size_t)n);
        for (int j = 0; j < n; ++j)
            if (esc::td_eig_cluster_start(ev, j, tn)) esc::td_eig_vectors(n, a, b, ev, j, tn, work.data(), Z);
        return;
    }
    const size_t nd = 2 * (size_t)n + (size_t)n + (size_t)n * n + (size_t)esc::kTridiagThreads * 5 * n + 1;
    double *d = nullptr;
    if (hipMalloc(&d, nd * sizeof(double)) != hipSuccess) throw EscError(ESC_ERR_ALLOC, "tridiagonal workspace");
    struct Free { double *p; ~Free() { (void)hipFree(p); } } fr{d};
    double *da = d, *db = da + n, *dev = db + n, *dz = dev + n, *dw = dz + (size_t)n * n;
    int *dst = reinterpret_cast<int *>(dw + (size_t)esc::kTridiagThreads * 5 * n);
    std::vector<double> bb(b, b + n);
    bb[0] = 0.0;
    hip_check(hipMemcpyAsync(da, a, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s), "hipMemcpy");
    hip_check(hipMemcpyAsync(db, bb.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, s), "hipMemcpy");
    hip_check(esc::launch_tridiag_eig(n, da, db, dev, dz, dw, dst, s), "tridiagonal kernel");
    int st = 0;
    hip_check(hipMemcpyAsync(&st, dst, sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpy");
    hip_check(hipMemcpyAsync(ev, dev, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpy");
    hip_check(hipMemcpyAsync(Z, dz, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpy");
    hip_check(hipStreamSynchronize(s), "sync");
    if (st) throw EscError(ESC_ERR_NUMERIC, "non-finite tridiagonal matrix");
}

// Lowest eigenvalue only (the per-iteration convergence test): Sturm-count
// bisection (esc::td_lowest), O(n) per halving instead of the O(n^2) of a QL
// solve; the batched solves run the same routine on the device.
double tridiag_lowest(int n, const double *a, const double *b) {
    bool ok = false;
    const double e = esc::td_lowest(n, a, b, &ok);
    if (!ok) throw EscError(ESC_ERR_NUMERIC, "non-finite tridiagonal matrix");
    return e;
}

}  // namespace esc_impl

namespace esc_impl {
// One line on stderr per finished stage of a long call (a continued fraction,
// a ground-state cycle) when the environment sets ESC_VERBOSE to anything but
// "0" (read at every call, so a caller may switch it on and off); silent
// otherwise.
void progress(const char *fmt, ...) {
    const char *v = std::getenv("ESC_VERBOSE");
    if (v == nullptr || v[0] == '\0' || std::strcmp(v, "0") == 0) return;
    static const auto t0 = std::chrono::steady_clock::now();
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "esc: [%.1f ms] %s\n", ms, buf);
    std::fflush(stderr);
}
}  // namespace esc_impl


namespace esc_impl {

void build_tables(esc_ctx *c) {
    const esc_model &m = c->model;
    esc::HTables &t = c->htab;
    std::memset(&t, 0, sizeof(t));
    const int N = m.n_sites;
    t.n_sites = (uint32_t)N;
    for (int k = 0; k < esc::kMaxSites; ++k) t.xpart[k] = t.ypart[k] = -1;
    // per-family bond coefficients (gather form of src/Hamiltonian.h:113-158;
    // J and Gamma are extensions)
    const double J4 = 0.25 * m.J, G4 = 0.25 * m.Gamma;
    t.p0[0] = 0.25 * m.K[0] + J4; t.p1[0] = -J4;                t.q[0] = 0.0;
    t.p0[1] = J4;                 t.p1[1] = -0.25 * m.K[1] - J4; t.q[1] = 0.0;
    t.p0[2] = J4;                 t.p1[2] = -J4;                t.q[2] = -G4;
    t.kz4 = 0.25 * m.K[2];
    t.j4 = J4;
    t.bz = m.B[2];
    t.bx2 = 0.5 * m.B[0];
    t.by2 = 0.5 * m.B[1];
    t.g4 = G4;
    uint32_t nb = 0, nz = 0, nj = 0;
    for (int f = 0; f < 3; ++f) {
        const bool offdiag = (f == 0 && (m.K[0] != 0.0 || m.J != 0.0)) || (f == 1 && (m.K[1] != 0.0 || m.J != 0.0)) ||
                             (f == 2 && (m.J != 0.0 || m.Gamma != 0.0));
        for (int b = 0; b < m.n_bonds[f]; ++b) {
            const int i = c->bonds[f][2 * b], j = c->bonds[f][2 * b + 1];
            const uint32_t packed = (uint32_t)i | ((uint32_t)j << 8);
            if (offdiag) {
                if (nb >= (uint32_t)esc::kMaxBonds) throw EscError(ESC_ERR_UNSUPPORTED, "too many bonds");
                t.bond[nb++] = packed | ((uint32_t)f << 16);
            }
            if (f == 2) t.zdiag[nz++] = packed;
            if (m.J != 0.0) {
                if (nj >= (uint32_t)esc::kMaxBonds) throw EscError(ESC_ERR_UNSUPPORTED, "too many bonds");
                t.jdiag[nj++] = packed;
            }
            if (m.Gamma != 0.0 && f < 2) {
                int32_t *part = f == 0 ? t.xpart : t.ypart;
                if (part[i] >= 0 || part[j] >= 0)
                    throw EscError(ESC_ERR_UNSUPPORTED, "Gamma needs at most one x and one y bond per site");
                part[i] = j;
                part[j] = i;
            }
        }
    }
    t.n_bonds = nb;
    t.n_zdiag = nz;
    t.n_jdiag = nj;
    t.has_gamma_flip = m.Gamma != 0.0 ? 1u : 0u;
    t.n_flip = (m.B[0] != 0.0 || m.B[1] != 0.0 || m.Gamma != 0.0) ? (uint32_t)N : 0u;
}

// Tables of the tiled kernel for tile size TB (see esc::GTables).
bool kz_sum_exact(double kz4, int nz) {
    // every partial sum +-kz4 * m, |m| <= nz, must be exactly representable
    for (int m = 1; m <= nz; ++m) {
        const double p = kz4 * (double)m;
        if (std::fma(kz4, (double)m, -p) != 0.0) return false;
    }
    return true;
}

struct HostTerm {
    uint32_t mask;
    int b0, b1, b2;
    double coef[8][2];
    bool cplx;
};

// all off-diagonal terms of H in the 8-entry-table form (values as HTables)
std::vector<HostTerm> host_terms(const esc_ctx *c) {
    const esc::HTables &h = c->htab;
    std::vector<HostTerm> out;
    auto push = [&](uint32_t mask, int b0, int b1, int b2, const double (*coef)[2]) {
        // a pair that sits in several families (the chain's XYZ bonds) is one
        // gather with the summed table
        for (HostTerm &o : out)
            if (o.mask == mask && o.b0 == b0 && o.b1 == b1 && o.b2 == b2) {
                for (int e = 0; e < 8; ++e) {
                    o.coef[e][0] += coef[e][0];
                    o.coef[e][1] += coef[e][1];
                    if (o.coef[e][1] != 0.0) o.cplx = true;
                }
                return;
            }
        HostTerm t;
        t.mask = mask;
        t.b0 = b0; t.b1 = b1; t.b2 = b2;
        t.cplx = false;
        for (int e = 0; e < 8; ++e) {
            t.coef[e][0] = coef[e][0];
            t.coef[e][1] = coef[e][1];
            if (coef[e][1] != 0.0) t.cplx = true;
        }
        out.push_back(t);
    };
    for (uint32_t b = 0; b < h.n_bonds; ++b) {          // p0 + p1 s_i s_j + i q (s_i + s_j)
        const uint32_t p = h.bond[b];
        const int i = p & 0xff, j = (p >> 8) & 0xff, f = p >> 16;
        double coef[8][2];
        for (int e = 0; e < 8; ++e) {
            const double si = (e & 1) ? -1.0 : 1.0, sj = (e & 2) ? -1.0 : 1.0;
            coef[e][0] = h.p0[f] + h.p1[f] * (si * sj);
            coef[e][1] = h.q[f] * (si + sj);
        }
        push((1u << i) | (1u << j), i, j, i, coef);
    }
    for (uint32_t k = 0; k < h.n_flip; ++k) {            // bx2 + g4 s_q - i s_k (by2 + g4 s_p)
        const int xp = h.has_gamma_flip ? h.xpart[k] : -1, yp = h.has_gamma_flip ? h.ypart[k] : -1;
        double coef[8][2];
        for (int e = 0; e < 8; ++e) {
            const double sk = (e & 1) ? -1.0 : 1.0, sp = (e & 2) ? -1.0 : 1.0, sq = (e & 4) ? -1.0 : 1.0;
            double re = h.bx2, im = h.by2;
            if (xp >= 0) im += h.g4 * sp;
            if (yp >= 0) re += h.g4 * sq;
            coef[e][0] = re;
            coef[e][1] = -sk * im;
        }
        push(1u << k, (int)k, xp >= 0 ? xp : (int)k, yp >= 0 ? yp : (int)k, coef);
    }
    return out;
}

// Tile sites of one pass: the six lane sites plus TB - 6 sites of the window
// W, added one at a time so that as many of the given masks as possible lie
// inside the tile (ties: the site most masks touch, then the lowest site).
uint32_t greedy_tile(int N, uint32_t W, const std::vector<uint32_t> &masks, int TB) {
    uint32_t T = 63u;
    while (__builtin_popcount(T) < TB) {
        int best = -1;
        long best_key = -1;
        for (int s = 6; s < N; ++s) {
            if (!((W >> s) & 1u) || ((T >> s) & 1u)) continue;
            const uint32_t nt = T | (1u << s);
            long inside = 0, touch = 0;
            for (uint32_t m : masks) {
                inside += (m & ~nt) == 0;
                touch += (m >> s) & 1u;
            }
            const long key = inside * 1024 + touch;
            if (key > best_key) { best_key = key; best = s; }
        }
        if (best < 0) break;
        T |= 1u << best;
    }
    return T;
}

// A window of w sites holding the tile T and as many of the masks as fit
// (highest masks first), filled up with the lowest free sites.
uint32_t grow_window(int N, uint32_t T, const std::vector<uint32_t> &masks, int w) {
    uint32_t W = T;
    std::vector<uint32_t> order(masks);
    std::sort(order.begin(), order.end(), [](uint32_t a, uint32_t b) { return a > b; });
    for (uint32_t m : order)
        if (__builtin_popcount(W | m) <= w) W |= m;
    for (int b = 0; b < N && __builtin_popcount(W) < w; ++b) W |= 1u << b;
    return W;
}

// Passes of the windowed tiled kernel for tile bits TB, window width w and at
// most P planned passes (w >= n_sites: one pass over one window = the whole
// vector).  Planning: pass 1 takes the low w sites as its window, later
// passes are centred on the terms not yet covered; each pass picks the tile
// sites that make the most of its terms "near" (partner inside the LDS tile).
// Terms that are near in no planned pass are applied as "far" terms (partner
// read from the L2-resident window) by a pass whose window holds them; terms
// that fit no window open further passes.  P = 5 selects the cover-first
// planner below instead: windows chosen to hold all terms in as few passes as
// the window width allows, tiles second.  P = 6 is the same plan with the
// windows run in reverse order: the first window (the low sites, one
// contiguous block per window) runs last, where the output is re-read, and
// the first tile is chosen for the window that holds the fewest terms.
constexpr int kCoverFirst = 5;   // pass digit of the cover-first planner
constexpr int kCoverLast = 6;    // cover-first, windows in reverse order

// A kernel configuration is 100000 * P + 10000 * V + 100 * TB + W: P planned
// passes (kCoverFirst: the cover-first planner), 2^V waves per workgroup,
// 2^TB-row tiles, 2^W-row windows; 0 = one row per lane.
// Configurations 7 and 8 in the pass digit run the cover-first plan (5) or
// its reverse (6) with the generated sub-cube kernels of esc_cube.cpp; the
// millions digit is then k (2^k rows per lane).
constexpr int kCubeFirst = 7;
constexpr int kCubeLast = 8;
int cfg_passes(int config) { return (config / 100000) % 10; }
int cfg_regs(int config) { return (config / 1000000) % 10; }
// the layout digit of the generated kernels (0..3): bit 0 = over row pairs
// (complex64: 16-byte lane accesses), bit 1 = lanes on six contiguous sites
bool cfg_paired(int config) { return ((config / 10000000) % 10) & 1; }
bool cfg_contiguous(int config) { return ((config / 10000000) % 10) & 2; }
bool cfg_cube(int config) { return cfg_passes(config) == kCubeFirst || cfg_passes(config) == kCubeLast; }
int cfg_waves(int config) { return 1 << ((config / 10000) % 10); }
int cfg_tile_bits(int config) { return (config / 100) % 100; }
int cfg_window(int config) { return config % 100; }
int make_config(int P, int waves, int tb, int w, int k = 0, bool paired = false, bool contiguous = false) {
    int v = 0;
    while ((1 << v) < waves) ++v;
    return 10000000 * ((paired ? 1 : 0) | (contiguous ? 2 : 0)) + 1000000 * k + 100000 * P + 10000 * v + 100 * tb + w;
}

// One planned pass: window sites W, tile sites T (lane sites 0..5 included)
// and the terms (indices into host_terms) it applies from the tile (near) and
// from the rest of the window (far).
struct Plan {
    uint32_t W, T;
    std::vector<size_t> near, far;
};

std::vector<Plan> plan_windows(const esc_ctx *c, const std::vector<HostTerm> &terms, int TB, int w, int P) {
    const int N = c->n_sites;
    if (w > N) w = N;
    if (TB > w || TB < 7 || N - w > 16 || w - TB > 16) throw EscError(ESC_ERR_UNSUPPORTED, "tile/window shape");
    const uint32_t all = (N >= 32) ? 0xffffffffu : ((1u << N) - 1u);
    std::vector<Plan> plan;
    std::vector<size_t> rem(terms.size());
    for (size_t t = 0; t < terms.size(); ++t) rem[t] = t;
    auto masks_of = [&](const std::vector<size_t> &ids) {
        std::vector<uint32_t> m;
        for (size_t t : ids) m.push_back(terms[t].mask);
        return m;
    };
    auto take_near = [&](Plan &p) {
        std::vector<size_t> keep;
        for (size_t t : rem) ((terms[t].mask & ~p.T) == 0 ? p.near : keep).push_back(t);
        rem.swap(keep);
    };
    if (w >= N) P = 1;
    if (P == kCoverFirst || P == kCoverLast) {
        // Windows first: each holds as many of the terms no earlier window
        // holds as possible (grown site by site from the lane sites, then
        // single-site swaps while they gain), until every term lies in some
        // window.  Then each window's tile makes the most of its terms near
        // that no earlier tile has, and every term goes to the first pass in
        // which it is near, else (far) to the window holding it with the
        // fewest far terms so far.
        auto held = [&](uint32_t W, const std::vector<size_t> &ids) {
            long n = 0;
            for (size_t t : ids) n += (terms[t].mask & ~W) == 0;
            return n;
        };
        std::vector<size_t> left(rem);
        std::vector<uint32_t> wins;
        while (!left.empty()) {
            uint32_t W = 63u;
            while (__builtin_popcount(W) < w) {
                int best = -1;
                long bk = -1;
                for (int st = 6; st < N; ++st) {
                    if ((W >> st) & 1u) continue;
                    long touch = 0;
                    for (size_t t : left) touch += (terms[t].mask >> st) & 1u;
                    const long key = held(W | (1u << st), left) * 1024 + touch;
                    if (key > bk) { bk = key; best = st; }
                }
                W |= 1u << best;
            }
            long val = held(W, left);
            for (bool gain = true; gain;) {
                gain = false;
                for (int a = 6; a < N && !gain; ++a) {
                    if (!((W >> a) & 1u)) continue;
                    for (int b = 6; b < N && !gain; ++b) {
                        if ((W >> b) & 1u) continue;
                        const uint32_t W2 = (W & ~(1u << a)) | (1u << b);
                        const long v = held(W2, left);
                        if (v > val) { W = W2; val = v; gain = true; }
                    }
                }
            }
            if (val == 0) throw EscError(ESC_ERR_UNSUPPORTED, "window construction made no progress");
            wins.push_back(W);
            std::vector<size_t> keep;
            for (size_t t : left)
                if ((terms[t].mask & ~W) != 0) keep.push_back(t);
            left.swap(keep);
            if (wins.size() > 8) throw EscError(ESC_ERR_UNSUPPORTED, "too many window passes");
        }
        if (P == kCoverLast) std::reverse(wins.begin(), wins.end());
        std::vector<char> is_near(terms.size(), 0);
        for (uint32_t W : wins) {
            std::vector<uint32_t> cand;
            for (size_t t = 0; t < terms.size(); ++t)
                if (!is_near[t] && (terms[t].mask & ~W) == 0) cand.push_back(terms[t].mask);
            Plan p;
            p.W = W;
            p.T = greedy_tile(N, W, cand, TB);
            for (size_t t = 0; t < terms.size(); ++t)
                if ((terms[t].mask & ~p.T) == 0) is_near[t] = 1;
            plan.push_back(std::move(p));
        }
        for (size_t t = 0; t < terms.size(); ++t) {
            Plan *dst = nullptr;
            for (auto &p : plan)
                if ((terms[t].mask & ~p.T) == 0) { dst = &p; break; }
            if (dst) { dst->near.push_back(t); continue; }
            for (auto &p : plan)
                if ((terms[t].mask & ~p.W) == 0 && (!dst || p.far.size() < dst->far.size())) dst = &p;
            dst->far.push_back(t);
        }
        rem.clear();
    }
    while (plan.empty() || (!rem.empty() && (int)plan.size() < P)) {
        Plan p;
        const std::vector<uint32_t> rm = masks_of(rem);
        if (plan.empty()) {
            p.W = (w >= N) ? all : ((1u << w) - 1u);
        } else {
            p.W = grow_window(N, greedy_tile(N, all, rm, TB), rm, w);
        }
        p.T = greedy_tile(N, p.W, rm, TB);
        take_near(p);
        plan.push_back(std::move(p));
    }
    {   // far terms: the window with the fewest far terms so far
        std::vector<size_t> keep;
        for (size_t t : rem) {
            Plan *best = nullptr;
            for (auto &p : plan)
                if ((terms[t].mask & ~p.W) == 0 && (!best || p.far.size() < best->far.size())) best = &p;
            if (best) best->far.push_back(t);
            else keep.push_back(t);
        }
        rem.swap(keep);
    }
    while (!rem.empty()) {   // leftovers: further passes
        Plan p;
        const std::vector<uint32_t> rm = masks_of(rem);
        p.W = grow_window(N, 63u, rm, w);
        p.T = greedy_tile(N, p.W, rm, TB);
        take_near(p);
        std::vector<size_t> keep;
        for (size_t t : rem) ((terms[t].mask & ~p.W) == 0 ? p.far : keep).push_back(t);
        if (p.near.empty() && p.far.empty()) throw EscError(ESC_ERR_UNSUPPORTED, "window construction made no progress");
        rem.swap(keep);
        plan.push_back(std::move(p));
        if (plan.size() > 8) throw EscError(ESC_ERR_UNSUPPORTED, "too many window passes");
    }
    return plan;
}

// window bit order of a pass: lane sites, the other tile sites, then the rest
// of the window (wb); the complement (window id bits) ascending (cb)
void window_order(int N, const Plan &pl, std::vector<int> &wb, std::vector<int> &cb) {
    wb.clear();
    cb.clear();
    for (int b = 0; b < 6; ++b) wb.push_back(b);
    for (int b = 6; b < N; ++b)
        if ((pl.T >> b) & 1u) wb.push_back(b);
    for (int b = 6; b < N; ++b)
        if (((pl.W & ~pl.T) >> b) & 1u) wb.push_back(b);
    for (int b = 0; b < N; ++b)
        if (!((pl.W >> b) & 1u)) cb.push_back(b);
}

std::vector<esc::GTables> build_passes(const esc_ctx *c, int TB, int w, int P, int WPB,
                                       const std::vector<HostTerm> *custom = nullptr) {
    const int N = c->n_sites;
    if (w > N) w = N;
    if (TB - 6 > 7 || WPB < 1 || ((1 << (TB - 6)) % WPB) != 0) throw EscError(ESC_ERR_UNSUPPORTED, "tile/wave shape");
    esc::HTables h = c->htab;
    if (custom) {   // an operator without a diagonal: every diagonal coupling zero
        h.kz4 = h.j4 = h.bz = 0.0;
        h.n_zdiag = h.n_jdiag = 0;
    }
    const std::vector<HostTerm> terms = custom ? *custom : host_terms(c);
    const std::vector<Plan> plan = plan_windows(c, terms, TB, w, P);

    std::vector<esc::GTables> passes;
    bool first = true;
    for (const Plan &pl : plan) {
        std::vector<int> wb, cb;
        window_order(N, pl, wb, cb);
        const int ww = (int)wb.size();
        std::vector<size_t> mine(pl.near);
        mine.insert(mine.end(), pl.far.begin(), pl.far.end());
        esc::GTables g;
        std::memset(&g, 0, sizeof(g));
        g.n_sites = (uint32_t)N;
        g.tile_bits = (uint32_t)TB;
        g.waves = (uint32_t)WPB;
        g.first_pass = first ? 1u : 0u;
        g.n_windows = 1u << (N - ww);
        g.blocks_per_window = 1u << (ww - TB);
        g.xcd_windows = (g.n_windows % 8u == 0u) ? 1u : 0u;
        auto scat = [](const std::vector<int> &bits, int from, uint32_t v) {
            uint32_t r = 0;
            for (int q = 0; from + q < (int)bits.size() && (v >> q); ++q)
                if ((v >> q) & 1u) r |= 1u << bits[from + q];
            return r;
        };
        for (uint32_t gi = 0; gi < (1u << (TB - 6)); ++gi) g.tscat[gi] = scat(wb, 6, gi);
        for (uint32_t v = 0; v < 256; ++v) {
            g.bscat[0][v] = scat(wb, TB, v);
            g.bscat[1][v] = scat(wb, TB + 8, v);
            g.wscat[0][v] = scat(cb, 0, v);
            g.wscat[1][v] = scat(cb, 8, v);
        }
        // tile-group sites: group index bit q sits on site wb[6 + q]
        const int ngb = TB - 6, ngrp = 1 << ngb, ng = ngrp / WPB;
        uint32_t gsites = 0;
        int gidx[32];
        for (int b = 0; b < 32; ++b) gidx[b] = -1;
        for (int q = 0; q < ngb; ++q) {
            gsites |= 1u << wb[6 + q];
            gidx[wb[6 + q]] = q;
        }
        std::vector<esc::GTerm> cls[esc::kNumClasses];
        for (size_t t : mine) {
            const HostTerm &ht = terms[t];
            esc::GTerm gt;
            std::memset(&gt, 0, sizeof(gt));
            gt.mask = ht.mask;
            gt.mlo = ht.mask & 63u;
            gt.bits = (uint32_t)ht.b0 | ((uint32_t)ht.b1 << 8) | ((uint32_t)ht.b2 << 16);
            std::memcpy(gt.coef, ht.coef, sizeof(gt.coef));
            for (int e = 0; e < 8; ++e)
                for (int q = 0; q < 2; ++q) gt.coef32[e][q] = (float)ht.coef[e][q];
            const bool far = (ht.mask & ~(63u | gsites)) != 0;
            gt.mhib = ht.mask & ~(63u | gsites);
            uint32_t mg = 0;     // mask on the group index
            for (int q = 0; q < ngb; ++q)
                if ((ht.mask >> wb[6 + q]) & 1u) mg |= 1u << q;
            const int sel[3] = {ht.b0, ht.b1, ht.b2};
            bool lane = false;
            for (int q = 0; q < 3; ++q) {
                if (sel[q] < 6) {
                    gt.lp_pos[q] = (uint32_t)sel[q];
                    gt.lp_w[q] = 1;
                    lane = true;
                } else if (gidx[sel[q]] < 0) {
                    gt.bsel[q] = (uint32_t)sel[q] | (1u << 16);
                }
            }
            for (int wv = 0; wv < WPB; ++wv) {
                for (int k = 0; k < ng; ++k) {
                    const uint32_t gg = (uint32_t)(wv + WPB * k), slot = (uint32_t)(wv * ng + k);
                    gt.gofs[slot] = far ? g.tscat[gg ^ mg] : ((gg ^ mg) << 6);
                    uint32_t up = 0;
                    for (int q = 0; q < 3; ++q)
                        if (sel[q] >= 6 && gidx[sel[q]] >= 0) up |= ((gg >> gidx[sel[q]]) & 1u) << q;
                    gt.ugrp[slot] = up * (c->c64() ? 8u : 16u);   // bytes into the table
                }
            }
            cls[(ht.cplx ? 1 : 0) | (far ? 2 : 0) | (lane ? 4 : 0)].push_back(gt);
        }
        size_t n = 0;
        for (int q = 0; q < esc::kNumClasses; ++q) {
            g.cls_begin[q] = (uint32_t)n;
            if (n + cls[q].size() > (size_t)esc::kMaxTerms) throw EscError(ESC_ERR_UNSUPPORTED, "too many terms");
            for (auto &t : cls[q]) {
                if (q >= 4) t.ct_base = (uint32_t)(n - g.cls_begin[4]) * 8u;
                g.term[n++] = t;
            }
        }
        g.cls_begin[esc::kNumClasses] = (uint32_t)n;
        // diagonal (used by the first pass only)
        g.kz4 = h.kz4;
        g.j4 = h.j4;
        g.bz = h.bz;
        g.nz = h.n_zdiag;
        g.nj = h.n_jdiag;
        g.diag_closed = kz_sum_exact(h.kz4, (int)h.n_zdiag) ? 1u : 0u;
        auto groups = [&](const uint32_t *list, uint32_t nl, uint32_t *gd, uint32_t *gm, uint32_t &ng) {
            ng = 0;
            for (uint32_t b = 0; b < nl; ++b) {
                uint32_t i = list[b] & 0xff, j = (list[b] >> 8) & 0xff;
                if (i > j) std::swap(i, j);
                const uint32_t d = j - i;
                uint32_t q = 0;
                while (q < ng && gd[q] != d) ++q;
                if (q == ng) {
                    if (ng >= (uint32_t)esc::kMaxDistGroups) { g.diag_closed = 0; return; }
                    gd[ng] = d;
                    gm[ng] = 0;
                    ++ng;
                }
                if (gm[q] & (1u << i)) { g.diag_closed = 0; return; }   // two bonds on one lower site
                gm[q] |= 1u << i;
            }
        };
        groups(h.zdiag, h.n_zdiag, g.zgrp_d, g.zgrp_m, g.n_zgrp);
        groups(h.jdiag, h.n_jdiag, g.jgrp_d, g.jgrp_m, g.n_jgrp);
        passes.push_back(g);
        first = false;
    }
    passes.back().last_pass = 1u;
    return passes;
}

// The generated sub-cube kernels (esc_cube.cpp) for configuration cfg: the
// cover-first plan of the generic tiled kernel seeds the sub-cube planner, its
// terms in bond / flip form.  Returns false (with the reason) for models they
// do not cover: a pair in several bond families (the chain's XYZ bonds).  A
// Kitaev-z diagonal whose closed form would round differently is summed bond
// by bond in the reference's order.
bool cube_spec(const esc_ctx *c, int cfg, esc::CubeModel &m, std::vector<esc::CubePass> &passes, std::string *why,
               bool do_plan = true) {
    const esc::HTables &h = c->htab;
    const int N = c->n_sites;
    const std::vector<HostTerm> terms = host_terms(c);
    std::vector<esc::CubeTerm> ct;
    for (uint32_t b = 0; b < h.n_bonds; ++b) {
        const uint32_t p = h.bond[b];
        esc::CubeTerm t{};
        t.kind = 0;
        t.i = (int)(p & 0xff);
        t.j = (int)((p >> 8) & 0xff);
        t.f = (int)(p >> 16);
        t.mask = (1u << t.i) | (1u << t.j);
        t.k = t.xp = t.yp = -1;
        ct.push_back(t);
    }
    for (uint32_t k = 0; k < h.n_flip; ++k) {
        esc::CubeTerm t{};
        t.kind = 1;
        t.k = (int)k;
        t.xp = h.has_gamma_flip ? h.xpart[k] : -1;
        t.yp = h.has_gamma_flip ? h.ypart[k] : -1;
        t.mask = 1u << k;
        t.f = t.i = t.j = -1;
        ct.push_back(t);
    }
    if (ct.size() != terms.size()) {
        if (why) *why = "a pair sits in several bond families";
        return false;
    }
    const int TB = cfg_tile_bits(cfg), w = cfg_window(cfg), P = cfg_passes(cfg) - 2, WPB = cfg_waves(cfg);
    const std::vector<Plan> plan = plan_windows(c, terms, TB, w, P);
    const std::vector<esc::GTables> g = build_passes(c, TB, w, P, 4);
    std::memset(&m, 0, sizeof(m));
    if (!g[0].diag_closed) {
        if (h.n_zdiag > 64 || h.n_jdiag > 96) {
            if (why) *why = "too many diagonal bonds for the bond-by-bond sum";
            return false;
        }
        m.seq_diag = true;
        m.n_zpair = h.n_zdiag;
        m.n_jpair = h.n_jdiag;
        for (uint32_t q = 0; q < h.n_zdiag; ++q) m.zpair[q] = h.zdiag[q];
        for (uint32_t q = 0; q < h.n_jdiag; ++q) m.jpair[q] = h.jdiag[q];
    }
    m.n_sites = N;
    m.c64 = c->c64();
    for (int f = 0; f < 3; ++f) m.bond_cplx[f] = h.q[f] != 0.0;
    m.flip_cplx = h.by2 != 0.0 || h.g4 != 0.0;
    m.flip_partners = h.has_gamma_flip != 0;
    m.nz = g[0].nz;
    m.nj = g[0].nj;
    m.n_zgrp = g[0].n_zgrp;
    m.n_jgrp = g[0].n_jgrp;
    for (int q = 0; q < 32; ++q) {
        m.zgrp_d[q] = g[0].zgrp_d[q];
        m.zgrp_m[q] = g[0].zgrp_m[q];
        m.jgrp_d[q] = g[0].jgrp_d[q];
        m.jgrp_m[q] = g[0].jgrp_m[q];
    }
    m.reg_sites = cfg_regs(cfg);
    m.paired = cfg_paired(cfg);
    m.contiguous = cfg_contiguous(cfg);
    if (m.paired && !m.c64) {
        if (why) *why = "row pairs are a complex64 layout";
        return false;
    }
    m.waves = WPB;
    if (!do_plan) return true;
    // the generic kernel's cover-first tiles seed the sub-cube planner, with
    // at most w - TB far sites per pass
    std::vector<uint32_t> tiles;
    for (const Plan &pl : plan) tiles.push_back(pl.T);
    if (!esc::cube_plan(m, ct, TB, tiles, std::max(0, std::min(w, N) - TB), passes)) {
        if (why) *why = "no sub-cube plan places every term";
        return false;
    }
    return true;
}

// coupling scalars of the generated kernels (esc::CubeScalar order)
void cube_scalars(const esc::HTables &h, double *cf) {
    for (int f = 0; f < 3; ++f) {
        cf[3 * f] = h.p0[f] + h.p1[f];
        cf[3 * f + 1] = h.p0[f] + h.p1[f] * -1.0;
        cf[3 * f + 2] = h.q[f] * 2.0;
    }
    cf[esc::kCubeRp] = h.bx2 + h.g4;
    cf[esc::kCubeRm] = h.bx2 + h.g4 * -1.0;
    cf[esc::kCubeIp] = h.by2 + h.g4;
    cf[esc::kCubeIm] = h.by2 + h.g4 * -1.0;
    cf[esc::kCubeKz4] = h.kz4;
    cf[esc::kCubeJ4] = h.j4;
    cf[esc::kCubeBz] = h.bz;
}

// Pauli-string table for one operator (src/Observables.h:161-389,429-521)
void add_pair(esc::CurrentTable &t, std::initializer_list<esc::OpFactor> a, double wa,
              std::initializer_list<esc::OpFactor> b, double wb) {
    // weights are real here; complex scale factors are applied by scale_table
    if (t.n_pairs >= esc::kMaxCurrentPairs) throw EscError(ESC_ERR_UNSUPPORTED, "too many current terms");
    esc::CurrentPair &p = t.pair[t.n_pairs++];
    std::memset(&p, 0, sizeof(p));
    p.na = (int)a.size();
    p.nb = (int)b.size();
    int q = 0;
    for (auto &f : a) p.fa[q++] = f;
    q = 0;
    for (auto &f : b) p.fb[q++] = f;
    p.wa_re = wa;
    p.wa_im = 0.0;
    p.wb_re = wb;
    p.wb_im = 0.0;
}

// append src scaled by the complex factor (cr + i ci) to dst
void append_scaled(esc::CurrentTable &dst, const esc::CurrentTable &src, double cr, double ci) {
    for (int k = 0; k < src.n_pairs; ++k) {
        if (dst.n_pairs >= esc::kMaxCurrentPairs) throw EscError(ESC_ERR_UNSUPPORTED, "too many current terms");
        esc::CurrentPair p = src.pair[k];
        const double ar = p.wa_re, ai = p.wa_im, br = p.wb_re, bi = p.wb_im;
        if (!(cr == 1.0 && ci == 0.0)) {
            p.wa_re = cr * ar - ci * ai;
            p.wa_im = cr * ai + ci * ar;
            p.wb_re = cr * br - ci * bi;
            p.wb_im = cr * bi + ci * br;
        }
        dst.pair[dst.n_pairs++] = p;
    }
}

int nb_of(const esc_ctx *c, int site, int col) {
    if (c->neigh.empty()) throw EscError(ESC_ERR_ARG, "current operators need the neighbour table (model.neigh)");
    const int j = c->neigh[(size_t)site * 3 + col];
    if (j < 0 || j >= c->n_sites) throw EscError(ESC_ERR_ARG, "site " + std::to_string(site) + " has no neighbour in column " + std::to_string(col));
    return j;
}

// Factor lists are in application order: the factor acting first on the ket
// comes first (ApplyOp chains, e.g. src/Observables.h:181-183).
void add_site_current(const esc_ctx *c, esc::CurrentTable &t, int op, int si) {
    const double Kx = c->model.K[0], Ky = c->model.K[1], Kz = c->model.K[2];
    enum { X = 0, Y = 1, Z = 2 };
    switch (op) {
    case ESC_OP_SX: add_pair(t, {{X, si}}, 1.0, {}, 0.0); break;
    case ESC_OP_SY: add_pair(t, {{Y, si}}, 1.0, {}, 0.0); break;
    case ESC_OP_SZ: add_pair(t, {{Z, si}}, 1.0, {}, 0.0); break;
    case ESC_OP_JSZ: {  // Kyy Sx_i Sy_y - Kxx Sy_i Sx_x   (:181-198)
        const int xn = nb_of(c, si, 2), yn = nb_of(c, si, 1);
        (void)nb_of(c, si, 0);
        add_pair(t, {{Y, yn}, {X, si}}, Ky, {{X, xn}, {Y, si}}, Kx);
        break;
    }
    case ESC_OP_JSY: {  // Kxx Sz_i Sx_x - Kzz Sx_i Sz_z   (:222-239)
        const int xn = nb_of(c, si, 2), zn = nb_of(c, si, 0);
        (void)nb_of(c, si, 1);
        add_pair(t, {{X, xn}, {Z, si}}, Kx, {{Z, zn}, {X, si}}, Kz);
        break;
    }
    case ESC_OP_JSX: {  // Kzz Sy_i Sz_z - Kyy Sz_i Sy_y   (:263-280)
        const int yn = nb_of(c, si, 1), zn = nb_of(c, si, 0);
        (void)nb_of(c, si, 2);
        add_pair(t, {{Z, zn}, {Y, si}}, Kz, {{Y, yn}, {Z, si}}, Ky);
        break;
    }
    case ESC_OP_JEX: {  // (KxKz) Sx_i Sy_ix Sz_ixpz - (KxKy) Sx_i Sz_ix Sy_ixpy   (:362-388)
        const int ix = nb_of(c, si, 2);
        const int ixpy = nb_of(c, ix, 1), ixpz = nb_of(c, ix, 0);
        add_pair(t, {{Z, ixpz}, {Y, ix}, {X, si}}, Kx * Kz, {{Y, ixpy}, {Z, ix}, {X, si}}, Kx * Ky);
        break;
    }
    case ESC_OP_JEY: {  // (KxKy) Sy_i Sz_iy Sx_iypx - (KyKz) Sy_i Sx_iy Sz_iypz   (:325-353)
        const int iy = nb_of(c, si, 1);
        const int iypx = nb_of(c, iy, 2), iypz = nb_of(c, iy, 0);
        add_pair(t, {{X, iypx}, {Z, iy}, {Y, si}}, Kx * Ky, {{Z, iypz}, {X, iy}, {Y, si}}, Ky * Kz);
        break;
    }
    default: throw EscError(ESC_ERR_ARG, "unknown operator " + std::to_string(op));
    }
}

void host_current_table(const esc_ctx *c, int op, int site, esc::CurrentTable *t) {
    std::memset(t, 0, sizeof(esc::CurrentTable));
    if (site == ESC_SITE_TOTAL) {
        if (op <= ESC_OP_SZ) throw EscError(ESC_ERR_ARG, "single-site operators need a site");
        // Js totals over all sites (:664-699); Jex over even, Jey over odd sites (:448-470)
        const int start = op == ESC_OP_JEY ? 1 : 0, step = (op == ESC_OP_JEX || op == ESC_OP_JEY) ? 2 : 1;
        for (int s = start; s < c->n_sites; s += step) add_site_current(c, *t, op, s);
    } else {
        if (site < 0 || site >= c->n_sites) throw EscError(ESC_ERR_ARG, "site out of range");
        add_site_current(c, *t, op, site);
    }
}

esc::CurrentTable *current_table(esc_ctx *c, int op, int site) {
    const int64_t key = (int64_t)op * 1000 + (site + 1);
    auto it = c->d_cur.find(key);
    if (it != c->d_cur.end()) return it->second;
    auto t = std::make_unique<esc::CurrentTable>();
    host_current_table(c, op, site, t.get());
    esc::CurrentTable *d = nullptr;
    if (hipMalloc(&d, sizeof(esc::CurrentTable)) != hipSuccess) throw EscError(ESC_ERR_ALLOC, "table allocation failed");
    hip_check(hipMemcpy(d, t.get(), sizeof(esc::CurrentTable), hipMemcpyHostToDevice), "hipMemcpy");
    c->d_cur[key] = d;
    return d;
}

// One Pauli string of a current table as a gather term of the tiled kernel:
// y[r] += w phase(r) x[r ^ m], m the sites of its Sx / Sy factors, phase(r)
// the product of the factors applied first to last as k_current applies
// them, tabulated over the bits of r on the sites of its Sy / Sz factors (at
// most three).  The entries are exact (w times +-2^-n or +-i 2^-n).  False
// for strings with no flip or more than three phase sites.
bool string_gather_term(const esc::OpFactor *f, int n, double wr, double wi, HostTerm &t) {
    uint32_t m = 0;
    int sel[3], ns = 0;
    for (int q = 0; q < n; ++q) {
        if (f[q].op != 2) m ^= 1u << f[q].site;
        if (f[q].op != 0) {
            bool have = false;
            for (int u = 0; u < ns; ++u) have = have || sel[u] == f[q].site;
            if (!have) {
                if (ns == 3) return false;
                sel[ns++] = f[q].site;
            }
        }
    }
    if (m == 0) return false;
    if (ns == 0) sel[ns++] = __builtin_ctz(m);
    while (ns < 3) { sel[ns] = sel[0]; ++ns; }
    t.mask = m;
    t.b0 = sel[0]; t.b1 = sel[1]; t.b2 = sel[2];
    t.cplx = false;
    for (int e = 0; e < 8; ++e) {
        uint32_t r = 0;
        for (int q = 2; q >= 0; --q) r = (r & ~(1u << sel[q])) | ((uint32_t)((e >> q) & 1) << sel[q]);
        uint32_t k = r ^ m;
        double cr = 1.0, ci = 0.0;
        for (int q = 0; q < n; ++q) {
            const uint32_t site = (uint32_t)f[q].site, bit = (k >> site) & 1u;
            if (f[q].op == 0) {
                cr *= 0.5; ci *= 0.5; k ^= 1u << site;
            } else if (f[q].op == 1) {
                const double h = bit ? -0.5 : 0.5;
                const double nr = -ci * h, ni = cr * h;
                cr = nr; ci = ni; k ^= 1u << site;
            } else {
                const double h = bit ? -0.5 : 0.5;
                cr *= h; ci *= h;
            }
        }
        t.coef[e][0] = wr * cr - wi * ci;
        t.coef[e][1] = wr * ci + wi * cr;
        if (t.coef[e][1] != 0.0) t.cplx = true;
    }
    return true;
}

// the value of a gather term's coefficient for row r
const double *term_value(const HostTerm &t, uint32_t r) {
    const int e = (int)((r >> t.b0) & 1u) | (int)(((r >> t.b1) & 1u) << 1) | (int)(((r >> t.b2) & 1u) << 2);
    return t.coef[e];
}

// All strings of a current table as gather terms, terms with the same mask
// merged while their phase sites number at most three.  False when some
// string does not fit (then k_current applies the table).
bool current_gather_terms(const esc::CurrentTable &ct, std::vector<HostTerm> &out) {
    out.clear();
    for (int p = 0; p < ct.n_pairs; ++p) {
        const esc::CurrentPair &cp = ct.pair[p];
        for (int side = 0; side < 2; ++side) {
            const int n = side ? cp.nb : cp.na;
            if (n == 0) continue;
            HostTerm t;
            const double wr = side ? -cp.wb_re : cp.wa_re, wi = side ? -cp.wb_im : cp.wa_im;
            if (!string_gather_term(side ? cp.fb : cp.fa, n, wr, wi, t)) return false;
            bool merged = false;
            for (HostTerm &o : out) {
                if (o.mask != t.mask) continue;
                int u[3], nu = 0;
                for (int site : {o.b0, o.b1, o.b2, t.b0, t.b1, t.b2}) {
                    bool have = false;
                    for (int v = 0; v < nu; ++v) have = have || u[v] == site;
                    if (!have) {
                        if (nu == 3) { nu = 4; break; }
                        u[nu++] = site;
                    }
                }
                if (nu > 3) continue;
                while (nu < 3) { u[nu] = u[0]; ++nu; }
                HostTerm mt = o;
                mt.b0 = u[0]; mt.b1 = u[1]; mt.b2 = u[2];
                mt.cplx = false;
                for (int e = 0; e < 8; ++e) {
                    uint32_t r = 0;
                    for (int q = 2; q >= 0; --q) r = (r & ~(1u << u[q])) | ((uint32_t)((e >> q) & 1) << u[q]);
                    const double *a = term_value(o, r), *b = term_value(t, r);
                    mt.coef[e][0] = a[0] + b[0];
                    mt.coef[e][1] = a[1] + b[1];
                    if (mt.coef[e][1] != 0.0) mt.cplx = true;
                }
                o = mt;
                merged = true;
                break;
            }
            if (!merged) out.push_back(t);
        }
    }
    return !out.empty();
}

uint64_t table_hash(const esc::CurrentTable &t) {
    uint64_t h = 1469598103934665603ull;
    const unsigned char *p = reinterpret_cast<const unsigned char *>(&t.pair[0]);
    const size_t n = (size_t)std::max(0, t.n_pairs) * sizeof(esc::CurrentPair);
    for (size_t i = 0; i < n; ++i) { h ^= p[i]; h *= 1099511628211ull; }
    h ^= (uint64_t)(uint32_t)t.n_pairs;
    h *= 1099511628211ull;
    return h;
}

void apply_current_table(esc_ctx *c, const esc::CurrentTable &ht, const esc::CurrentTable *dt, const void *x, void *y,
                         hipStream_t s) {
    if (x == y) throw EscError(ESC_ERR_ARG, "operator input and output must differ");
    const int N = c->n_sites;
    const esc_ctx::CurPass *cp = nullptr;
    if (N >= 17) {
        const uint64_t key = table_hash(ht);
        auto it = c->d_curpass.find(key);
        if (it == c->d_curpass.end()) {
            const auto t0 = std::chrono::steady_clock::now();
            esc_ctx::CurPass e;
            std::vector<HostTerm> terms;
            if (current_gather_terms(ht, terms)) {
                // the generic kernel's shapes for this size and precision
                const bool c64 = c->c64();
                std::vector<int> pref;
                if (N <= 18) pref.push_back(make_config(1, 8, 10, N));
                else if (N <= 22) pref.push_back(c64 ? make_config(kCoverLast, 16, 12, 17) : make_config(kCoverFirst, 16, 11, 17));
                else pref.push_back(c64 ? make_config(kCoverFirst, 8, 12, 18) : make_config(kCoverFirst, 16, 12, 18));
                pref.push_back(make_config(1, 8, 11, N));
                pref.push_back(make_config(1, 4, 10, N));
                for (int cfg : pref) {
                    const int tb = cfg_tile_bits(cfg), waves = cfg_waves(cfg);
                    if (N < tb || !esc::tile_shape_supported(c64, tb, waves)) continue;
                    std::vector<esc::GTables> host;
                    try {
                        host = build_passes(c, tb, cfg_window(cfg), cfg_passes(cfg), waves, &terms);
                    } catch (const EscError &) {
                        continue;
                    }
                    bool fits = true;
                    for (const auto &g : host)
                        fits = fits && esc::tile_lds_bytes(c64, tb, (int)g.cls_begin[4],
                                                           (int)g.cls_begin[esc::kNumClasses]) <= esc::kMaxTileLds;
                    if (!fits) continue;
                    for (const auto &g : host) {
                        esc::GTables *d = nullptr;
                        if (hipMalloc(&d, sizeof(esc::GTables)) != hipSuccess) {
                            for (auto &q : e.pass) (void)hipFree(q.d);
                            throw EscError(ESC_ERR_ALLOC, "table allocation failed");
                        }
                        e.pass.push_back({d, (int)g.cls_begin[4], (int)g.cls_begin[esc::kNumClasses]});
                        hip_check(hipMemcpy(d, &g, sizeof(esc::GTables), hipMemcpyHostToDevice), "hipMemcpy");
                    }
                    e.config = cfg;
                    break;
                }
            }
            progress("operator table of %d pairs: %d gather terms, %s (%.1f ms)", ht.n_pairs, (int)terms.size(),
                     e.config ? "tiled kernel" : "k_current",
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
            it = c->d_curpass.emplace(key, std::move(e)).first;
        }
        if (it->second.config) cp = &it->second;
    }
    if (cp) {
        const int tb = cfg_tile_bits(cp->config), waves = cfg_waves(cp->config);
        for (const auto &p : cp->pass)
            hip_check(esc::launch_hpsi_tile(c->c64(), p.d, c->d_htab, tb, waves, p.n_uniform, p.n_terms, x, y, nullptr,
                                            c->dim, s),
                      "tiled operator launch");
        return;
    }
    const esc::CurrentTable *d = dt;
    if (!d) {
        if (!c->d_combo && hipMalloc(&c->d_combo, sizeof(esc::CurrentTable)) != hipSuccess)
            throw EscError(ESC_ERR_ALLOC, "table allocation failed");
        hip_check(hipMemcpyAsync(c->d_combo, &ht, sizeof(esc::CurrentTable), hipMemcpyHostToDevice, s), "table");
        hip_check(hipStreamSynchronize(s), "sync");   // the host table may go out of scope
        d = c->d_combo;
    }
    hip_check(c->c64() ? esc::launch_current_c64(d, x, y, c->dim, s) : esc::launch_current_c128(d, x, y, c->dim, s),
              "current launch");
}

void check_ctx(const esc_ctx *c) {
    if (!c) throw EscError(ESC_ERR_ARG, "null context");
}
void check_ptr(const void *p, const char *name) {
    if (!p) throw EscError(ESC_ERR_ARG, std::string("null pointer: ") + name);
}

void sync_get(hipStream_t s, void *dst, const void *src, size_t n) {
    hip_check(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

void device_dot(esc_ctx *c, const void *x, const void *y, double out[2], hipStream_t s) {
    c->ensure_partials(2 * esc::kRedBlocks);
    c->ensure_result(4);
    hip_check(esc::launch_dot(c->c64(), x, y, c->dim, c->d_partials, c->d_result, s), "dot");
    sync_get(s, out, c->d_result, 2 * sizeof(double));
}

double device_norm2(esc_ctx *c, const void *x, hipStream_t s) {
    c->ensure_partials(2 * esc::kRedBlocks);
    c->ensure_result(4);
    hip_check(esc::launch_norm2(c->c64(), x, c->dim, c->d_partials, c->d_result, s), "norm2");
    double r;
    sync_get(s, &r, c->d_result, sizeof(double));
    return r;
}

const std::vector<esc_ctx::PassDev> &passes_for(esc_ctx *c, int config) {
    auto it = c->d_pass.find(config);
    if (it != c->d_pass.end()) return it->second;
    const std::vector<esc::GTables> host =
        build_passes(c, cfg_tile_bits(config), cfg_window(config), cfg_passes(config), cfg_waves(config));
    std::vector<esc_ctx::PassDev> dev;
    for (const auto &g : host) {
        esc::GTables *d = nullptr;
        if (hipMalloc(&d, sizeof(esc::GTables)) != hipSuccess) {
            for (auto &p : dev) (void)hipFree(p.d);
            throw EscError(ESC_ERR_ALLOC, "table allocation failed");
        }
        hip_check(hipMemcpy(d, &g, sizeof(esc::GTables), hipMemcpyHostToDevice), "hipMemcpy");
        dev.push_back({d, (int)g.cls_begin[4], (int)g.cls_begin[esc::kNumClasses]});
    }
    return c->d_pass[config] = dev;
}

// the generated kernels of a cube configuration, compiled (or taken from the
// caches) on first use; throws with the reason when they cannot be built
esc::CubeProgram &cube_for(esc_ctx *c, int config) {
    auto it = c->cube.find(config);
    if (it != c->cube.end()) return *it->second;
    auto f = c->cube_fail.find(config);
    if (f != c->cube_fail.end()) throw EscError(ESC_ERR_UNSUPPORTED, f->second);
    esc::CubeModel m;
    std::vector<esc::CubePass> passes;
    std::string why;
    std::shared_ptr<esc::CubeProgram> prog;
    if (cube_spec(c, config, m, passes, &why)) prog = esc::cube_program(m, passes, &why);
    if (!prog) {
        c->cube_fail[config] = "generated H.psi kernels unavailable: " + why;
        throw EscError(ESC_ERR_UNSUPPORTED, c->cube_fail[config]);
    }
    if (!c->d_cube_cf) {
        double cf[esc::kCubeScalars];
        cube_scalars(c->htab, cf);
        if (hipMalloc(&c->d_cube_cf, sizeof(cf)) != hipSuccess) throw EscError(ESC_ERR_ALLOC, "table allocation failed");
        hip_check(hipMemcpy(c->d_cube_cf, cf, sizeof(cf), hipMemcpyHostToDevice), "hipMemcpy");
    }
    c->cube[config] = prog;
    return *prog;
}

// Compile the generated programs of several configurations into the process
// and disk caches on up to 16 host threads, so that a tuner loading them one
// after another does not wait for one compilation after another.  Failures
// are left for cube_for to report.
void cube_compile_parallel(const esc_ctx *c, const std::vector<int> &configs) {
    std::vector<int> todo;
    for (int cfg : configs)
        if (cfg > 0 && cfg_cube(cfg) && !c->cube.count(cfg) && !c->cube_fail.count(cfg)) todo.push_back(cfg);
    if (todo.size() < 2) return;
    host_parallel_for((int)todo.size(), 1, [&](int b, int e) {
        for (int q = b; q < e; ++q) {
            try {
                esc::CubeModel m;
                std::vector<esc::CubePass> passes;
                std::string why;
                if (cube_spec(c, todo[q], m, passes, &why)) (void)esc::cube_precompile(m, passes, &why);
            } catch (...) {
            }
        }
    });
}

// whether every pass of a configuration fits the LDS of one CU (cached)
bool tile_config_fits(const esc_ctx *c, int config) {
    auto it = c->tile_fits.find(config);
    if (it != c->tile_fits.end()) return it->second;
    if (cfg_cube(config)) {
        esc::CubeModel m;
        std::vector<esc::CubePass> passes;
        bool ok = false;
        try {
            ok = cube_spec(c, config, m, passes, nullptr, false);
        } catch (const EscError &) {
            ok = false;
        }
        const int TB = cfg_tile_bits(config), k = cfg_regs(config), waves = cfg_waves(config);
        const int P = cfg_paired(config) ? 1 : 0;
        ok = ok && k >= 1 && TB - 6 - k >= 0 && waves <= 16 && ((1 << (TB - 6 - k)) % waves) == 0 &&
             ((1 << (TB - 6 - P)) % waves) == 0 && ((size_t)1 << TB) * c->esz <= 131072;
        c->tile_fits[config] = ok;
        return ok;
    }
    bool ok = true;
    try {
        for (const auto &g : build_passes(c, cfg_tile_bits(config), cfg_window(config), cfg_passes(config),
                                          cfg_waves(config)))
            if (esc::tile_lds_bytes(c->c64(), (int)g.tile_bits, (int)g.cls_begin[4],
                                    (int)g.cls_begin[esc::kNumClasses]) > esc::kMaxTileLds)
                ok = false;
    } catch (const EscError &) {
        ok = false;
    }
    c->tile_fits[config] = ok;
    return ok;
}

// generated sub-cube configurations: 2^12- and 2^13-row tiles, 2^3 or 2^4
// rows per lane, 4 or 8 waves, the cover-first windows of the generic kernel
// in both orders; complex64 with and without row pairs; from 25 sites on
// (vectors of 256 MiB and more) the contiguous lane layout besides the free
// one.  ESC_NO_JIT=1 leaves them out.
std::vector<int> cube_candidates(const esc_ctx *c) {
    std::vector<int> v;
    const char *nj = std::getenv("ESC_NO_JIT");
    if (nj && nj[0] && std::strcmp(nj, "0") != 0) return v;
    const int N = c->n_sites;
    if (N < 13) return v;
    const int w = N <= 18 ? N : (N <= 22 ? 17 : 18);
    for (int tb : {12, 13}) {
        if (tb > w) continue;
        for (int k : {3, 4})
            for (int waves : {4, 8})
                for (int P : {kCubeFirst, kCubeLast})
                    for (bool paired : {false, true})
                        for (bool contiguous : {false, true}) {
                            if ((w == N && P == kCubeLast) || (paired && !c->c64()) || (contiguous && N < 25))
                                continue;
                            const int cfg = make_config(P, waves, tb, w, k, paired, contiguous);
                            if (tile_config_fits(c, cfg)) v.push_back(cfg);
                        }
    }
    return v;
}

// kernel configurations (make_config; window >= N: one pass), 0 = one row per
// lane
std::vector<int> tile_candidates(const esc_ctx *c) {
    std::vector<int> v;
    const int N = c->n_sites;
    const std::vector<int> windows = c->c64() ? std::vector<int>{N, 17, 18, 19} : std::vector<int>{N, 16, 17, 18};
    for (int waves : {4, 8, 16}) {
        for (int tb : {10, 11, 12, 13}) {
            if (N < tb || !esc::tile_shape_supported(c->c64(), tb, waves)) continue;
            for (int w : windows) {
                const int ww = w > N ? N : w;
                if (ww < tb || N - ww > 16 || ww - tb > 16) continue;
                for (int P : {1, 2, 3, kCoverFirst, kCoverLast}) {
                    if ((ww == N) != (P == 1)) continue;
                    const int cfg = make_config(P, waves, tb, ww);
                    if (std::find(v.begin(), v.end(), cfg) == v.end() && tile_config_fits(c, cfg)) v.push_back(cfg);
                }
            }
        }
    }
    for (int cfg : cube_candidates(c))
        if (std::find(v.begin(), v.end(), cfg) == v.end()) v.push_back(cfg);
    return v;
}

// Before tuning: the shapes measured fastest on the K-J-Gamma-h honeycomb.
// Up to 16 sites (c128 up to 18 on the tiled side) the vector stays in L2 and
// one row per lane wins (N = 16: 8.8 against 15.9 us for 2^11-row tiles);
// 18 sites c128: a single pass over 2^10-row tiles; 19-22 sites: cover-first
// over 2^17-row windows, 16 waves (c128 2^11-row tiles; c64 2^12-row tiles,
// windows in reverse order); from 23 sites: cover-first over 2^18-row windows
// with 2^12-row tiles (16 waves for c128, which moves the least HBM traffic of
// the equally fast shapes at N = 24, 8 waves for c64).  A shape the context
// cannot run falls back to a single pass over 2^11- or 2^10-row tiles.
// From 19 sites the generated sub-cube kernels come first (N = 24: 0.29 ms
// against 0.52 ms for the best generic shape in complex128, 0.15 against
// 0.32 ms in complex64; N = 20: 18.8 against 40.1 us and 13.9 against
// 28.8 us): complex128 2^3 rows per lane in 8-wave workgroups on 2^12-row
// tiles (N = 24: 289 us with free lane sites, 307 us with contiguous ones);
// complex64 over row pairs, with 2^3 rows per lane in 8-wave workgroups on
// 2^12-row tiles up to 22 sites, 2^4 rows per lane in 4-wave workgroups on
// 2^12-row tiles with the windows in reverse order up to 26 sites (N = 24:
// 154 against 162 us for 8 waves on 2^13-row tiles), and 2^3 rows per lane
// in 8-wave workgroups on 2^13-row tiles, windows in reverse order, beyond
// (N = 32: 66.6 ms; 72 ms without row pairs with contiguous lanes, 84 ms
// with 2^4 rows per lane).  A model they do not cover, or a missing run-time
// compiler, falls through to the generic shapes.
int default_tile_bits(esc_ctx *c) {
    const int N = c->n_sites;
    const bool c64 = c->c64();
    if (N <= 16 || (N <= 18 && c64)) return 0;
    std::vector<int> pref;
    if (N >= 19) {
        const int w = N <= 22 ? 17 : 18;
        if (c64 && N <= 22) pref.push_back(make_config(kCubeFirst, 8, 12, w, 3, true));
        else if (c64 && N <= 26) pref.push_back(make_config(kCubeLast, 4, 12, w, 4, true));
        else if (c64) pref.push_back(make_config(kCubeLast, 8, 13, w, 3, true));
        else pref.push_back(make_config(kCubeFirst, 8, 12, w, 3));
    }
    if (N <= 18) pref.push_back(make_config(1, 8, 10, N));
    else if (N <= 22) pref.push_back(c64 ? make_config(kCoverLast, 16, 12, 17) : make_config(kCoverFirst, 16, 11, 17));
    else pref.push_back(c64 ? make_config(kCoverFirst, 8, 12, 18) : make_config(kCoverFirst, 16, 12, 18));
    pref.push_back(make_config(1, 8, 11, N));
    pref.push_back(make_config(1, 4, 10, N));
    const std::vector<int> cand = tile_candidates(c);
    for (int cfg : pref) {
        if (std::find(cand.begin(), cand.end(), cfg) == cand.end()) continue;
        if (cfg_cube(cfg)) {
            try {
                cube_for(c, cfg);
            } catch (const EscError &) {
                continue;
            }
        }
        return cfg;
    }
    return 0;
}

void launch_hpsi_with(esc_ctx *c, int config, const void *in, void *out, hipStream_t s, double *dot_part) {
    if (config > 0 && cfg_cube(config)) {
        esc::CubeProgram &p = cube_for(c, config);
        hip_check(esc::cube_launch(p, c->d_cube_cf, in, out, dot_part, c->dim, s), "generated hpsi launch");
    } else if (config > 0) {
        const int waves = cfg_waves(config), tb = cfg_tile_bits(config);
        for (const auto &p : passes_for(c, config))
            hip_check(esc::launch_hpsi_tile(c->c64(), p.d, c->d_htab, tb, waves, p.n_uniform, p.n_terms, in, out,
                                            dot_part, c->dim, s),
                      "hpsi tile launch");
    } else {
        hip_check(c->c64() ? esc::launch_hpsi_c64(c->d_htab, in, out, c->dim, s)
                           : esc::launch_hpsi_c128(c->d_htab, in, out, c->dim, s), "hpsi launch");
    }
}

// H.psi on the device with the context's kernel choice
void run_hpsi(esc_ctx *c, const void *in, void *out, hipStream_t s) {
    if (c->tile_bits > 0) {
        launch_hpsi_with(c, c->tile_bits, in, out, s);
    } else {
        hip_check(c->c64() ? esc::launch_hpsi_c64(c->d_htab, in, out, c->dim, s)
                           : esc::launch_hpsi_c128(c->d_htab, in, out, c->dim, s), "hpsi launch");
    }
}

// out = H in and <in|out> into c->d_result[0..1] (FP64): the tiled kernel's
// last pass forms the per-tile partial sums from its staged input rows, so no
// separate pass over the two vectors is needed
void run_hpsi_dot(esc_ctx *c, const void *in, void *out, hipStream_t s) {
    c->ensure_result(8);
    if (c->tile_bits > 0) {
        const uint64_t ntiles = c->dim >> cfg_tile_bits(c->tile_bits);
        c->ensure_partials(2 * ntiles + 2 * esc::kRedBlocks);
        launch_hpsi_with(c, c->tile_bits, in, out, s, c->d_partials);
        hip_check(esc::launch_reduce(c->d_partials, (int)ntiles, 2, c->d_result, s), "reduce");
    } else {
        run_hpsi(c, in, out, s);
        c->ensure_partials(2 * esc::kRedBlocks);
        hip_check(esc::launch_dot(c->c64(), in, out, c->dim, c->d_partials, c->d_result, s), "dot");
    }
}

// classical Gram-Schmidt pass against V_0..V_{k-1}
void cgs_pass(esc_ctx *c, int k, void *w, hipStream_t s) {
    c->ensure_partials((size_t)2 * k * 256);
    c->ensure_result((size_t)2 * k + 8);
    hip_check(esc::launch_multidot(c->c64(), c->vec(0), k, w, c->dim, c->d_partials, c->d_result, s), "multidot");
    hip_check(esc::launch_multi_axpy(c->c64(), c->vec(0), k, c->d_result, -1.0, w, c->dim, false, s), "multi_axpy");
}

// Reorthogonalise the unit vector w against V_0..V_{k-1} and normalise it
// (continued fractions with reorthogonalisation): one classical Gram-Schmidt
// pass, and a second one only when the first removed more than half of w's
// squared norm (|w| < 1/sqrt 2 afterwards: the "twice is enough" test of
// Daniel, Gragg, Kaufman and Stewart).  With reorthogonalisation at every
// step the first pass removes components at the rounding level and the second
// is skipped; after a breakdown (w mostly inside the basis) both run.
void reorthogonalize(esc_ctx *c, int k, void *w, hipStream_t s) {
    const bool c64 = c->c64();
    cgs_pass(c, k, w, s);
    hip_check(esc::launch_norm2(c64, w, c->dim, c->d_partials, c->d_result + 4, s), "norm2");
    double n2;
    sync_get(s, &n2, c->d_result + 4, sizeof(double));
    if (!(n2 >= 0.5)) {
        cgs_pass(c, k, w, s);
        hip_check(esc::launch_norm2(c64, w, c->dim, c->d_partials, c->d_result + 4, s), "norm2");
    }
    hip_check(esc::launch_normalize(c64, w, c->d_result + 4, c->dim, s), "normalize");
}

// Three-vector Lanczos without reorthogonalisation from phi/|phi| (work holds
// three device vectors distinct from phi).  alpha[0..it-1], beta[0..it].
// The plain recursion steps are enqueued without a host round trip (scalars
// and the step index on the device, see k_rec_step); the host reads the
// scalars back every kRecChunk steps and applies the exit rules to every step
// in order, so the result is that of a check after each step.  At most
// kRecChunk - 1 steps past the exit are computed and dropped (small systems,
// one launch per chunk: kRecChunkSmall - 1).
//
// Device layout inside the context's result buffer: [0..1] the H.psi dot,
// [2] |w|^2, [4..5] rec, [6] the step index (uint32), [8 ..] 3 doubles per
// step (history), then 2 doubles per step (replay coefficients).
constexpr size_t kRecOff = 4, kStepOff = 6, kHistOff = 8;
constexpr int kRecChunk = 16, kRecChunkSmall = 64;

// Runs the recursion in the three rotating vectors buf: step k maps u_{k-1},
// u_k, u_{k+1} to buf[k % 3], buf[(k + 1) % 3], buf[(k + 2) % 3].  With coef
// (device, 2 doubles per step) every step first accumulates acc += coef_k
// c_k u_k.
struct RecRunner {
    esc_ctx *c;
    hipStream_t s;
    void *buf[3];
    const double *coef;   // device replay coefficients or null
    void *acc;
    uint32_t cap;         // history capacity in steps
    int k = 0;            // next step

    RecRunner(esc_ctx *cx, hipStream_t st, void *b0, void *b1, void *b2, uint32_t capacity,
              const double *d_coef = nullptr, void *d_acc = nullptr)
        : c(cx), s(st), buf{b0, b1, b2}, coef(d_coef), acc(d_acc), cap(capacity) {}
    double *rec() const { return c->d_result + kRecOff; }
    uint32_t *step() const { return reinterpret_cast<uint32_t *>(c->d_result + kStepOff); }
    double *hist() const { return c->d_result + kHistOff; }

    void start(double c0) {
        hip_check(esc::launch_rec_init(rec(), step(), c0, s), "recursion start");
        k = 0;
    }
    void accumulate(int kk) {
        hip_check(esc::launch_axpy_rec_step(c->c64(), coef, step(), rec(), buf[(kk + 1) % 3], acc, c->dim, s),
                  "accumulate");
    }
    void enqueue(int kk) {
        if (coef) accumulate(kk);
        const void *u = buf[(kk + 1) % 3];
        void *w = buf[(kk + 2) % 3];
        run_hpsi_dot(c, u, w, s);
        hip_check(esc::launch_lanczos_update_rec(c->c64(), w, u, kk > 0 ? buf[kk % 3] : nullptr, rec(), c->d_result,
                                                 hist(), step(), c->dim, c->d_partials, s),
                  "update");
        hip_check(esc::launch_rec_step(rec(), c->d_result, hist(), step(), cap, c->d_partials, s), "recursion scalars");
    }
    // small systems with the one-row-per-lane H.psi run whole chunks of steps
    // in one single-workgroup launch
    bool small() const { return c->tile_bits == 0 && c->dim <= esc::kSmallRecMaxDim; }
    // up to kRecChunk (small systems: kRecChunkSmall) more steps, stopping at
    // `target`; returns the number of steps done so far.  A single-workgroup
    // chunk ends early after an exhausted step, which the caller's exit rule
    // meets first.
    int advance(int target) {
        const int stop = std::min(target, k + (small() ? kRecChunkSmall : kRecChunk));
        if (small()) {
            if (stop > k)
                hip_check(esc::launch_rec_small(c->c64(), c->d_htab, buf[0], buf[1], buf[2], rec(), step(), hist(), cap,
                                                stop - k, coef, acc, c->dim, s),
                          "small recursion");
            k = std::max(k, stop);
            return k;
        }
        for (; k < stop; ++k) enqueue(k);
        return k;
    }
};

// hist entries [from, to) of the recursion, copied to the host
std::vector<double> rec_history(esc_ctx *c, int from, int to, hipStream_t s) {
    std::vector<double> h(3 * (size_t)(to - from));
    if (to > from) sync_get(s, h.data(), c->d_result + kHistOff + 3 * (size_t)from, h.size() * sizeof(double));
    return h;
}

void cf_run(esc_ctx *ctx, const void *d_phi, int steps, void *work[3], double *alpha, double *beta, int *iters,
            double *norm2, hipStream_t s, bool reortho = false) {
    check_ptr(alpha, "alpha");
    check_ptr(beta, "beta");
    check_ptr(iters, "iters");
    check_ptr(norm2, "norm2");
    if (steps < 1 || steps > 100000) throw EscError(ESC_ERR_ARG, "steps out of range");
    const bool c64 = ctx->c64();
    const uint64_t dim = ctx->dim;
    ctx->ensure_partials(2 * esc::kRedBlocks);
    ctx->ensure_result(8);
    const double n2 = device_norm2(ctx, d_phi, s);
    if (!(n2 > 0.0) || !std::isfinite(n2)) throw EscError(ESC_ERR_NUMERIC, "start vector has zero or non-finite norm");
    *norm2 = n2;
    // with reortho the vectors v_0..v_steps are kept in the context's vectors
    // and every new one is orthogonalised against all before it (two
    // classical Gram-Schmidt passes)
    if (reortho) {
        ctx->ensure_vecs((size_t)steps + 1);
        ctx->basis_kept = false;
        ctx->ensure_partials((size_t)2 * (steps + 2) * 256 + 2 * esc::kRedBlocks);
        ctx->ensure_result((size_t)2 * (steps + 2) + 8);
    }
    beta[0] = 0.0;
    int k = 0;
    if (!reortho) {
        // The recursion keeps unnormalised vectors u_k with v_k = c_k u_k
        // (c_0 = 1/|phi|, c_k = 1/beta_k): H v_k = c_k H u_k and alpha_k =
        // c_k^2 <u_k|H u_k> (formed by the H.psi kernel itself), and the
        // update w = c_k H u_k - alpha_k c_k u_k - beta_k c_{k-1} u_{k-1}
        // becomes u_{k+1}.  No separate dot or scaling pass per step, and the
        // scalars stay on the device between the checks.
        ctx->ensure_result(kHistOff + 3 * ((size_t)steps + 1));
        hip_check(hipMemcpyAsync(work[1], d_phi, dim * ctx->esz, hipMemcpyDeviceToDevice, s), "copy");
        RecRunner run(ctx, s, work[0], work[1], work[2], (uint32_t)steps);
        run.start(1.0 / std::sqrt(n2));
        int checked = 0;
        while (checked < steps) {
            k = run.advance(steps) - 1;
            const std::vector<double> h = rec_history(ctx, checked, k + 1, s);
            for (int j = checked; j <= k; ++j) {
                const double ar = h[3 * (size_t)(j - checked)], b = h[3 * (size_t)(j - checked) + 2];
                if (!std::isfinite(ar)) throw EscError(ESC_ERR_NUMERIC, "non-finite alpha");
                alpha[j] = ar;
                beta[j + 1] = b;
                if (!std::isfinite(b)) throw EscError(ESC_ERR_NUMERIC, "non-finite beta");
                if (b <= 1e-13 * std::max(1.0, std::fabs(ar))) {   // Krylov space exhausted: the fraction terminates
                    beta[j + 1] = 0.0;
                    *iters = j + 1;
                    return;
                }
            }
            checked = k + 1;
        }
        *iters = steps;
        return;
    }
    void *vp = nullptr, *v = ctx->vec(0), *w = ctx->vec(1);
    hip_check(hipMemcpyAsync(v, d_phi, dim * ctx->esz, hipMemcpyDeviceToDevice, s), "copy");
    hip_check(esc::launch_scale(c64, v, 1.0 / std::sqrt(n2), dim, s), "scale");
    for (; k < steps; ++k) {
        vp = k > 0 ? ctx->vec((size_t)k - 1) : nullptr;
        v = ctx->vec((size_t)k);
        w = ctx->vec((size_t)k + 1);
        run_hpsi(ctx, v, w, s);
        double a[2];
        hip_check(esc::launch_dot(c64, v, w, dim, ctx->d_partials, ctx->d_result, s), "dot");
        sync_get(s, a, ctx->d_result, 2 * sizeof(double));
        if (!std::isfinite(a[0])) throw EscError(ESC_ERR_NUMERIC, "non-finite alpha");
        alpha[k] = a[0];
        hip_check(esc::launch_lanczos_update(c64, w, v, k > 0 ? vp : nullptr, ctx->d_result, beta[k], dim,
                                             ctx->d_partials, ctx->d_result + 2, s), "update");
        double b2;
        sync_get(s, &b2, ctx->d_result + 2, sizeof(double));
        const double b = std::sqrt(b2);
        beta[k + 1] = b;
        if (!std::isfinite(b)) throw EscError(ESC_ERR_NUMERIC, "non-finite beta");
        if (b <= 1e-13 * std::max(1.0, std::fabs(a[0]))) {   // Krylov space exhausted: the fraction terminates
            beta[k + 1] = 0.0;
            ++k;
            break;
        }
        hip_check(esc::launch_scale(c64, w, 1.0 / b, dim, s), "scale");
        if (k > 0) reorthogonalize(ctx, k + 1, w, s);
    }
    *iters = k;
}

// Lanczos without reorthogonalisation from the unit vector `start`, in the
// context's vectors 0..2, with the exit rule of esc_lanczos (|dE0| < eps or
// `steps` iterations; also when the Krylov space is exhausted).  Returns the
// number n of alphas.  With resid_floor > 0 it also stops once the lowest
// Ritz pair's residual bound beta_n |z_{n-1}| falls to resid_floor |T|.  With
// coef != nullptr the same recursion is replayed (deterministic kernels: the
// same vectors bit for bit) for exactly n_fixed vectors, accumulating
// acc += coef[k] v_k (complex coefficients).
int plain_recursion(esc_ctx *c, const void *start, int steps, double eps, const std::complex<double> *coef,
                    int n_fixed, void *acc, std::vector<double> &ar, std::vector<double> &bt, hipStream_t s,
                    double resid_floor = 0.0) {
    const uint64_t dim = c->dim;
    c->ensure_partials(2 * esc::kRedBlocks);
    const int cap = coef ? n_fixed : steps;
    c->ensure_result(kHistOff + 3 * ((size_t)cap + 1) + 2 * (size_t)n_fixed);
    // unnormalised vectors u_k with v_k = cs_k u_k (cs_0 = 1, cs_k = 1/beta_k):
    // alpha_k = cs_k^2 <u_k|H u_k> comes from the H.psi kernel, and the update
    // w = cs_k H u_k - alpha_k cs_k u_k - beta_k cs_{k-1} u_{k-1} is u_{k+1};
    // the scalars stay on the device (RecRunner)
    hip_check(hipMemcpyAsync(c->vec(1), start, dim * c->esz, hipMemcpyDeviceToDevice, s), "copy start");
    ar.clear();
    bt.assign(1, 0.0);
    if (coef) {
        double *d_coef = c->d_result + kHistOff + 3 * ((size_t)cap + 1);
        std::vector<double> h(2 * (size_t)n_fixed);
        for (int j = 0; j < n_fixed; ++j) {
            h[2 * (size_t)j] = coef[j].real();
            h[2 * (size_t)j + 1] = coef[j].imag();
        }
        hip_check(hipMemcpyAsync(d_coef, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s), "coef");
        RecRunner run(c, s, c->vec(0), c->vec(1), c->vec(2), (uint32_t)cap, d_coef, acc);
        run.start(1.0);
        while (run.k < n_fixed - 1) run.advance(n_fixed - 1);
        run.accumulate(n_fixed - 1);   // the last vector: accumulate only
        // the replay ran as long as the recorded recursion did
        const std::vector<double> hh = rec_history(c, 0, n_fixed - 1, s);   // also orders the upload above
        for (int j = 0; j + 1 < n_fixed; ++j) {
            const double a0 = hh[3 * (size_t)j], b = hh[3 * (size_t)j + 2];
            ar.push_back(a0);
            bt.push_back(b);
            if (!std::isfinite(b) || b <= 1e-13 * std::max(1.0, std::fabs(a0)))
                throw EscError(ESC_ERR_NUMERIC, "replayed recursion ended early");
        }
        return n_fixed;
    }
    RecRunner run(c, s, c->vec(0), c->vec(1), c->vec(2), (uint32_t)cap);
    run.start(1.0);
    double E0 = 1.0, tnorm = 0.0;
    std::vector<double> z, work;
    int checked = 0;
    for (;;) {
        const int k = run.advance(steps) - 1;
        const std::vector<double> h = rec_history(c, checked, k + 1, s);
        for (int j = checked; j <= k; ++j) {
            const double a0 = h[3 * (size_t)(j - checked)], a1 = h[3 * (size_t)(j - checked) + 1];
            const double b = h[3 * (size_t)(j - checked) + 2];
            if (!std::isfinite(a0) || !std::isfinite(a1)) throw EscError(ESC_ERR_NUMERIC, "non-finite alpha");
            ar.push_back(a0);
            if (!std::isfinite(b)) throw EscError(ESC_ERR_NUMERIC, "non-finite beta");
            bt.push_back(b);
            const bool exhausted = b <= 1e-13 * std::max(1.0, std::fabs(a0));
            const int n = j + 1;
            bool done = exhausted || n == steps;
            if (n >= 2 && eps > 0.0) {   // with eps <= 0 the rule cannot fire
                const double D0 = tridiag_lowest(n, ar.data(), bt.data());
                if (std::fabs(E0 - D0) < eps) done = true;
                E0 = D0;
            }
            // residual floor: the lowest Ritz pair's residual |H y - theta y| =
            // beta_n |z_{n-1}| has reached what the vector storage can hold
            tnorm = std::max(tnorm, std::fabs(a0) + std::fabs(bt[n - 1]) + std::fabs(b));
            if (!done && resid_floor > 0.0 && n >= 2) {
                bool ok = false;
                const double th = esc::td_lowest(n, ar.data(), bt.data(), &ok);
                if (ok) {
                    z.resize(n);
                    work.resize(5 * (size_t)n);
                    esc::td_vector(n, ar.data(), bt.data(), th, work.data(), z.data());
                    if (b * std::fabs(z[n - 1]) <= resid_floor * tnorm) done = true;
                }
            }
            if (done) return n;
        }
        checked = k + 1;
    }
}

// ---- spectra from continued fractions (shared by one context and batches) ----

// column pairs of the output tables (src/dynamics.h:318-328,427-435); index
// nc is the total channel
std::vector<std::pair<int, int>> spectrum_pairs(bool js, int nc) {
    const int T = nc;
    if (js) return {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}, {T, T}};
    return {{0, 0}, {1, 1}, {0, 1}, {1, 0}, {T, T}};
}

// One continued-fraction run: phi = (sum_a coef_a J_a)|gs>.  kind 0: a
// diagonal channel (a), 1: the total channel, 2 / 3: the polarisation runs
// J_a + J_b and J_a + i J_b of a cross channel (a, b).

std::vector<CfRun> cf_runs(bool js, int nc) {
    std::vector<CfRun> r;
    for (int a = 0; a < nc; ++a) {
        std::vector<cplx> c(nc, cplx(0.0, 0.0));
        c[a] = 1.0;
        r.push_back({c, 0, a, a});
    }
    r.push_back({std::vector<cplx>(nc, cplx(1.0, 0.0)), 1, nc, nc});
    std::set<std::pair<int, int>> done;
    for (auto pr : spectrum_pairs(js, nc)) {
        const int a = pr.first, b = pr.second;
        if (a == b || done.count({a, b}) || done.count({b, a})) continue;
        done.insert({a, b});
        std::vector<cplx> c1(nc, cplx(0.0, 0.0)), ci(nc, cplx(0.0, 0.0));
        c1[a] = 1.0; c1[b] = 1.0;
        ci[a] = 1.0; ci[b] = cplx(0.0, 1.0);
        if (nc != 2) r.push_back({c1, 2, a, b});   // with two components J_a + J_b is the total
        r.push_back({ci, 3, a, b});
    }
    return r;
}

// norm2 <phi|(z - H)^-1|phi>, z = omega + e0 - i eta, from the recursion
// coefficients, on every omega (esc::td_cf, the routine the batched solves
// run on the device); the omegas run on the host threads.
std::vector<cplx> cf_eval(int it, const double *al, const double *be, double n2, double e0, double eta,
                          const double *omega, int nw) {
    std::vector<cplx> f(nw);
    host_parallel_for(nw, 128, [&](int q0, int q1) {
        for (int q = q0; q < q1; ++q) {
            double fr, fi;
            esc::td_cf(it, al, 1, be, n2, omega[q] + e0, -eta, &fr, &fi);
            f[q] = cplx(fr, fi);
        }
    });
    return f;
}

// I_ab(w) before the elastic subtraction from the fractions of cf_runs
std::map<std::pair<int, int>, std::vector<cplx>> cf_assemble(bool js, int nc, const std::vector<CfRun> &runs,
                                                             const std::vector<std::vector<cplx>> &f, int nw) {
    std::map<std::pair<int, int>, std::vector<cplx>> F;
    const int T = nc;
    std::map<std::pair<int, int>, const std::vector<cplx> *> s1, si;
    for (size_t r = 0; r < runs.size(); ++r) {
        const CfRun &c = runs[r];
        if (c.kind == 0 || c.kind == 1) F[{c.a, c.b}] = f[r];
        else if (c.kind == 2) s1[{c.a, c.b}] = &f[r];
        else si[{c.a, c.b}] = &f[r];
    }
    for (auto &kv : si) {
        const int a = kv.first.first, b = kv.first.second;
        // polarisation: f_ab = 1/2 [(S1 - faa - fbb) - i (Si - faa - fbb)],
        //               f_ba = 1/2 [(S1 - faa - fbb) + i (Si - faa - fbb)]
        const std::vector<cplx> &S1 = (nc == 2) ? F[{T, T}] : *s1[{a, b}];
        const std::vector<cplx> &Si = *kv.second;
        std::vector<cplx> fab(nw), fba(nw);
        for (int q = 0; q < nw; ++q) {
            const cplx u = S1[q] - F[{a, a}][q] - F[{b, b}][q];
            const cplx v = Si[q] - F[{a, a}][q] - F[{b, b}][q];
            fab[q] = 0.5 * (u - cplx(0.0, 1.0) * v);
            fba[q] = 0.5 * (u + cplx(0.0, 1.0) * v);
        }
        F[{a, b}] = fab;
        F[{b, a}] = fba;
    }
    (void)js;
    return F;
}

// output rows: elastic subtraction conj(<J_a>) <J_b> / (w - i eta), none for
// JeJe yx (the reference subtracts an uninitialised value there)
void write_spectrum_table(bool js, int nc, const std::vector<cplx> &ex,
                          std::map<std::pair<int, int>, std::vector<cplx>> &F, const double *omega, int nw,
                          double eta, double *table) {
    const int ncols = js ? 14 : 5;
    const auto pairs = spectrum_pairs(js, nc);
    for (int q = 0; q < nw; ++q) {
        double *row = table + (size_t)q * (1 + ncols);
        row[0] = omega[q];
        const cplx dn1(omega[q], -eta);
        int col = 1;
        for (auto pr : pairs) {
            cplx v = F[pr][q];
            const bool yx = !js && pr.first == 1 && pr.second == 0;
            if (!yx) v = v - std::conj(ex[pr.first]) * ex[pr.second] / dn1;
            if (js) { row[col++] = v.real(); row[col++] = v.imag(); }
            else row[col++] = v.imag();
        }
    }
}


// LocalDynSzSz (src/dynamics.h:64-217): per site i the spin components
// S^a_i (a = x, y, z) and the rotated S+e3_i, S-e3_i (src/Observables.h:
// 539-589, the reference's six-digit coefficients) applied to |gs>; columns
// Im of I_xx, I_yy, I_zz, I_xy, I_xz, I_yz summed over sites, each site with
// its elastic term conj(<S^a_i>) <S^b_i> / (w - i eta), then the two rotated
// channels with <S-e3_i><S+e3_i> subtracted as written.  Ritz mode sums over
// the kept Lanczos basis; CF mode runs 11 continued fractions per site.
const cplx kSpe3[3] = {cplx(-0.408248, 0.707107), cplx(-0.408248, -0.707107), cplx(0.816497, 0.0)};
const cplx kSme3[3] = {cplx(-0.408248, -0.707107), cplx(-0.408248, 0.707107), cplx(0.816497, 0.0)};

void local_szsz_spectra(esc_ctx *ctx, int method, bool reortho, const void *d_gs, double e0, int cf_steps,
                        double eta, const double *omega, int nw, double *table, double *expect, hipStream_t s) {
    const int N = ctx->n_sites;
    const bool cf = method == ESC_SPEC_CF;
    const bool cf_in_vecs = cf && !ctx->basis_kept;
    if (cf_in_vecs) ctx->ensure_vecs(3);
    ctx->ensure_aux(cf && !cf_in_vecs ? 4 : 1);
    ctx->ensure_partials(2 * esc::kRedBlocks);
    ctx->ensure_result(8);
    if (!ctx->d_combo && hipMalloc(&ctx->d_combo, sizeof(esc::CurrentTable)) != hipSuccess)
        throw EscError(ESC_ERR_ALLOC, "table allocation failed");
    // 8 channels: xx yy zz xy xz yz pp mm
    std::vector<std::vector<cplx>> acc(8, std::vector<cplx>(nw, cplx(0.0, 0.0)));
    cplx sum_ex[3] = {0.0, 0.0, 0.0};
    void *work[3] = {cf_in_vecs ? ctx->vec(0) : ctx->aux(1), cf_in_vecs ? ctx->vec(1) : ctx->aux(2),
                     cf_in_vecs ? ctx->vec(2) : ctx->aux(3)};
    std::vector<double> al(cf_steps + 2), be(cf_steps + 3);
    for (int si = 0; si < N; ++si) {
        esc::CurrentTable comp[3];
        for (int c = 0; c < 3; ++c) host_current_table(ctx, ESC_OP_SX + c, si, &comp[c]);
        auto apply = [&](const cplx *coef) {   // aux(0) = (sum_c coef_c S^c_i)|gs>
            esc::CurrentTable t;
            std::memset(&t, 0, sizeof(t));
            for (int c = 0; c < 3; ++c)
                if (coef[c] != cplx(0.0, 0.0)) append_scaled(t, comp[c], coef[c].real(), coef[c].imag());
            hip_check(hipMemcpy(ctx->d_combo, &t, sizeof(t), hipMemcpyHostToDevice), "hipMemcpy");
            hip_check(ctx->c64() ? esc::launch_current_c64(ctx->d_combo, d_gs, ctx->aux(0), ctx->dim, s)
                                 : esc::launch_current_c128(ctx->d_combo, d_gs, ctx->aux(0), ctx->dim, s),
                      "current launch");
        };
        const cplx unit[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
        cplx ex[3];
        std::vector<std::vector<cplx>> O(3);
        for (int a = 0; a < 3; ++a) {
            apply(unit[a]);
            double d[2];
            device_dot(ctx, d_gs, ctx->aux(0), d, s);
            ex[a] = cplx(d[0], d[1]);
            sum_ex[a] += ex[a];
            if (!cf) {   // O(n, a) = sum_m Z(m, n) <v_m|S^a_i gs>
                const int n = ctx->basis_n;
                ctx->ensure_partials((size_t)2 * n * 256);
                ctx->ensure_result((size_t)2 * n + 8);
                hip_check(esc::launch_multidot(ctx->c64(), ctx->vec(0), n, ctx->aux(0), ctx->dim, ctx->d_partials,
                                               ctx->d_result, s), "multidot");
                std::vector<double> tv(2 * (size_t)n);
                sync_get(s, tv.data(), ctx->d_result, sizeof(double) * 2 * n);
                O[a].assign(n, cplx(0.0, 0.0));
                for (int k = 0; k < n; ++k) {
                    cplx x(0.0, 0.0);
                    for (int m = 0; m < n; ++m) x += ctx->Z[(size_t)k * n + m] * cplx(tv[2 * m], tv[2 * m + 1]);
                    O[a][k] = x;
                }
            }
        }
        cplx exp_ = 0.0, exm = 0.0;
        for (int c = 0; c < 3; ++c) { exp_ += kSpe3[c] * ex[c]; exm += kSme3[c] * ex[c]; }
        const int pr[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
        std::vector<std::vector<cplx>> F(8, std::vector<cplx>(nw));
        if (!cf) {
            const int n = ctx->basis_n;
            std::vector<cplx> Op(n), Om(n);
            for (int k = 0; k < n; ++k) {
                Op[k] = kSpe3[0] * O[0][k] + kSpe3[1] * O[1][k] + kSpe3[2] * O[2][k];
                Om[k] = kSme3[0] * O[0][k] + kSme3[1] * O[1][k] + kSme3[2] * O[2][k];
            }
            const double Eg = ctx->evals[0];
            for (int q = 0; q < nw; ++q) {
                cplx v[8] = {};
                for (int k = 0; k < n; ++k) {
                    const cplx den(omega[q] - (ctx->evals[k] - Eg), -eta);
                    for (int c = 0; c < 6; ++c) v[c] += std::conj(O[pr[c][0]][k]) * O[pr[c][1]][k] / den;
                    v[6] += std::norm(Op[k]) / den;
                    v[7] += std::norm(Om[k]) / den;
                }
                for (int c = 0; c < 8; ++c) F[c][q] = v[c];
            }
        } else {
            auto run = [&](const cplx *coef) {
                apply(coef);
                int it = 0;
                double n2 = 0.0;
                cf_run(ctx, ctx->aux(0), cf_steps, work, al.data(), be.data(), &it, &n2, s, reortho);
                return cf_eval(it, al.data(), be.data(), n2, e0, eta, omega, nw);
            };
            for (int a = 0; a < 3; ++a) F[a] = run(unit[a]);
            F[6] = run(kSpe3);
            F[7] = run(kSme3);
            for (int c = 3; c < 6; ++c) {   // cross channels by polarisation
                const int a = pr[c][0], b = pr[c][1];
                cplx c1[3] = {0.0, 0.0, 0.0}, ci[3] = {0.0, 0.0, 0.0};
                c1[a] = 1.0; c1[b] = 1.0;
                ci[a] = 1.0; ci[b] = cplx(0.0, 1.0);
                const std::vector<cplx> S1 = run(c1), Si = run(ci);
                for (int q = 0; q < nw; ++q) {
                    const cplx u = S1[q] - F[a][q] - F[b][q];
                    const cplx v = Si[q] - F[a][q] - F[b][q];
                    F[c][q] = 0.5 * (u - cplx(0.0, 1.0) * v);
                }
            }
        }
        for (int q = 0; q < nw; ++q) {
            const cplx dn1(omega[q], -eta);
            for (int c = 0; c < 6; ++c) acc[c][q] += F[c][q] - std::conj(ex[pr[c][0]]) * ex[pr[c][1]] / dn1;
            acc[6][q] += F[6][q] - exm * exp_ / dn1;
            acc[7][q] += F[7][q] - exp_ * exm / dn1;
        }
    }
    for (int q = 0; q < nw; ++q) {
        double *row = table + (size_t)q * 9;
        row[0] = omega[q];
        for (int c = 0; c < 8; ++c) row[1 + c] = acc[c][q].imag();
    }
    if (expect)
        for (int a = 0; a < 3; ++a) { expect[2 * a] = sum_ex[a].real(); expect[2 * a + 1] = sum_ex[a].imag(); }
}

// <psi_k|vec> for every Ritz state psi_k = sum_m Z(m, k) V_m of the kept basis
std::vector<cplx> ritz_amplitudes(esc_ctx *ctx, const void *d_vec, hipStream_t s) {
    const int n = ctx->basis_n;
    ctx->ensure_partials((size_t)2 * n * 256);
    ctx->ensure_result((size_t)2 * n + 8);
    hip_check(esc::launch_multidot(ctx->c64(), ctx->vec(0), n, d_vec, ctx->dim, ctx->d_partials, ctx->d_result, s),
              "multidot");
    std::vector<double> t(2 * (size_t)n);
    sync_get(s, t.data(), ctx->d_result, sizeof(double) * 2 * n);
    std::vector<cplx> out(n);
    for (int k = 0; k < n; ++k) {
        cplx x(0.0, 0.0);
        for (int m = 0; m < n; ++m) x += ctx->Z[(size_t)k * n + m] * cplx(t[2 * m], t[2 * m + 1]);
        out[k] = x;
    }
    return out;
}

// invComplex(x, y) = (x + i y)/(x^2 + y^2) = 1/(x - i y)  (src/dynamics.h:34-39)
cplx inv_complex(double x, double y) {
    const double d = x * x + y * y;
    return cplx(x / d, y / d);
}

// LocalDynRIXS (src/dynamics.h:465-562).  Intermediate states n belong to
// H_int, the Hamiltonian without the bonds of the core-hole site c
// (src/Hamiltonian.h:57-91); final states m to H.  With
// inv_n = invComplex(0 - E_n, Gamma) the reference forms
//   A_X(m, n) = <n|gs> <m|X_c|n> inv_n,  X = 1, S^x_c, S^y_c, S^z_c,
//   I_X(w)   = sum_{m >= 1} |sum_n A_X(m, n)|^2 eta / ((w - (E_m - Eg))^2 + eta^2).
// The n sum is one vector, chi = sum_n |n> <n|gs> inv_n, so that
// sum_n A_X(m, n) = <m|X_c chi>: one multi-axpy over the intermediate basis and
// one multi-dot per X over the final basis, in place of |m| x |n| x 4 dot
// products.  Ritz mode: n and m run over the Ritz states of the two kept bases,
// Eg = the lowest Ritz value of H.  CF mode: chi from the Krylov space of H_int
// started at |gs> (replayed recursion of cf_steps steps), I_X from the
// continued fraction of H started at X_c chi at z = w + e0 - i eta, minus the
// m = 0 term |<gs|X_c chi>|^2 eta / (w^2 + eta^2).  Table rows: w, Re I_1,
// Re I_x, Re I_y, Re I_z, then the imaginary parts (zero: each I_X is a sum
// of |.|^2 times a real Lorentzian).
void rixs_spectra(esc_ctx *g, esc_ctx *q, int method, int site, const void *d_gs, double e0, int cf_steps,
                  double gamma, double eta, const double *omega, int nw, double *table, hipStream_t s) {
    const bool cf = method == ESC_SPEC_CF;
    const uint64_t dim = g->dim;
    g->ensure_aux(cf && g->basis_kept ? 5 : 2);
    g->ensure_partials(2 * esc::kRedBlocks);
    g->ensure_result(8);
    void *chi = g->aux(0), *phi = g->aux(1);
    if (!cf) {
        const int nq = q->basis_n;
        const std::vector<cplx> n0 = ritz_amplitudes(q, d_gs, s);   // <n|gs>
        std::vector<cplx> w(nq);
        for (int n = 0; n < nq; ++n) w[n] = n0[n] * inv_complex(0.0 - q->evals[n], gamma);
        std::vector<double> c(2 * (size_t)nq);
        for (int m = 0; m < nq; ++m) {
            cplx x(0.0, 0.0);
            for (int n = 0; n < nq; ++n) x += q->Z[(size_t)n * nq + m] * w[n];
            c[2 * m] = x.real();
            c[2 * m + 1] = x.imag();
        }
        q->ensure_result((size_t)2 * nq + 8);
        hip_check(hipMemcpyAsync(q->d_result, c.data(), sizeof(double) * c.size(), hipMemcpyHostToDevice, s), "coef");
        hip_check(esc::launch_multi_axpy(q->c64(), q->vec(0), nq, q->d_result, 1.0, chi, dim, true, s), "chi");
        hip_check(hipStreamSynchronize(s), "sync");
    } else {
        // chi = (-H_int - i gamma)^-1 |gs> in the Krylov space of H_int from
        // |gs>: with T = Z diag(theta) Z^T,
        // chi = |gs| sum_k v_k sum_j Z(k, j) Z(0, j) inv(theta_j)
        q->ensure_vecs(4);
        q->basis_kept = false;
        q->ensure_partials(2 * esc::kRedBlocks);
        q->ensure_result(8);
        void *start = q->vec(3);
        hip_check(hipMemcpyAsync(start, d_gs, dim * q->esz, hipMemcpyDeviceToDevice, s), "copy gs");
        const double n2 = device_norm2(q, start, s);
        if (!(n2 > 0.0) || !std::isfinite(n2)) throw EscError(ESC_ERR_NUMERIC, "ground state has zero or non-finite norm");
        hip_check(esc::launch_scale(q->c64(), start, 1.0 / std::sqrt(n2), dim, s), "scale");
        std::vector<double> ar, bt;
        const int n = plain_recursion(q, start, cf_steps, 0.0, nullptr, 0, nullptr, ar, bt, s);
        std::vector<double> ev(n), Z((size_t)n * n);
        tridiag_eigenpairs(n, ar.data(), bt.data(), ev.data(), Z.data(), true, s);
        std::vector<cplx> y(n, cplx(0.0, 0.0));
        for (int j = 0; j < n; ++j) {
            const cplx f = std::sqrt(n2) * Z[(size_t)j * n] * inv_complex(0.0 - ev[j], gamma);
            for (int k = 0; k < n; ++k) y[k] += Z[(size_t)j * n + k] * f;
        }
        hip_check(hipMemsetAsync(chi, 0, dim * g->esz, s), "clear");
        std::vector<double> ar2, bt2;
        plain_recursion(q, start, cf_steps, 0.0, y.data(), n, chi, ar2, bt2, s);
    }
    const int ops[4] = {-1, ESC_OP_SX, ESC_OP_SY, ESC_OP_SZ};
    std::vector<double> al(cf_steps + 2), be(cf_steps + 3);
    if (cf && !g->basis_kept) g->ensure_vecs(3);
    void *work[3] = {nullptr, nullptr, nullptr};
    if (cf)
        for (int k = 0; k < 3; ++k) work[k] = g->basis_kept ? g->aux(2 + k) : g->vec(k);
    for (int x = 0; x < 4; ++x) {
        const void *src = chi;
        if (ops[x] >= 0) {
            esc::CurrentTable *t = current_table(g, ops[x], site);
            hip_check(g->c64() ? esc::launch_current_c64(t, chi, phi, dim, s) : esc::launch_current_c128(t, chi, phi, dim, s),
                      "site operator");
            src = phi;
        }
        std::vector<double> I(nw, 0.0);
        if (!cf) {
            const std::vector<cplx> F = ritz_amplitudes(g, src, s);   // sum_n A_X(m, n)
            const double Eg = g->evals[0];
            for (int qw = 0; qw < nw; ++qw) {
                double acc = 0.0;
                for (int m = 1; m < g->basis_n; ++m) {
                    const double xm = omega[qw] - (g->evals[m] - Eg);
                    acc += std::norm(F[m]) * (eta / (xm * xm + eta * eta));
                }
                I[qw] = acc;
            }
        } else {
            double d[2];
            device_dot(g, d_gs, src, d, s);
            const double el = d[0] * d[0] + d[1] * d[1];
            int it = 0;
            double n2 = 0.0;
            cf_run(g, src, cf_steps, work, al.data(), be.data(), &it, &n2, s);
            const std::vector<cplx> f = cf_eval(it, al.data(), be.data(), n2, e0, eta, omega, nw);
            for (int qw = 0; qw < nw; ++qw) I[qw] = f[qw].imag() - el * eta / (omega[qw] * omega[qw] + eta * eta);
        }
        for (int qw = 0; qw < nw; ++qw) {
            double *row = table + (size_t)qw * 9;
            row[0] = omega[qw];
            row[1 + x] = I[qw];
            row[5 + x] = 0.0;
        }
    }
}

}  // namespace esc_impl

extern "C" {

const char *esc_last_error(void) { return esc_impl::g_err.c_str(); }
int esc_version(void) { return 100; }

namespace esc_impl {
// validated copy of the model and its host tables; no device work
void ctx_init_host(esc_ctx *c, const esc_model *model, int precision) {
        check_ptr(model, "model");
        const int N = model->n_sites;
        if (N < 2 || N > esc::kMaxSites) throw EscError(ESC_ERR_ARG, "n_sites must be in [2, 32]");
        if (precision != ESC_C128 && precision != ESC_C64) throw EscError(ESC_ERR_ARG, "precision must be ESC_C128 or ESC_C64");
        for (double v : {model->K[0], model->K[1], model->K[2], model->J, model->Gamma, model->B[0], model->B[1], model->B[2]})
            if (!std::isfinite(v)) throw EscError(ESC_ERR_NUMERIC, "non-finite coupling");
        c->n_sites = N;
        c->precision = precision;
        c->dim = (uint64_t)1 << N;
        c->esz = precision == ESC_C64 ? 8 : 16;
        c->model = *model;
        for (int f = 0; f < 3; ++f) {
            const int nb = model->n_bonds[f];
            if (nb < 0) throw EscError(ESC_ERR_ARG, "negative bond count");
            if (nb > 0) check_ptr(model->bonds[f], "bonds");
            c->bonds[f].assign(model->bonds[f], model->bonds[f] + 2 * nb);
            for (int b = 0; b < nb; ++b) {
                const int i = c->bonds[f][2 * b], j = c->bonds[f][2 * b + 1];
                if (i < 0 || j < 0 || i >= N || j >= N || i == j) throw EscError(ESC_ERR_ARG, "invalid bond");
            }
            c->model.bonds[f] = nullptr;
        }
        if (model->neigh) c->neigh.assign(model->neigh, model->neigh + 3 * N);
        c->model.neigh = nullptr;
        build_tables(c);
}
}  // namespace esc_impl

int esc_ctx_create(const esc_model *model, int precision, esc_ctx **out) {
    return guarded([&] {
        check_ptr(out, "out");
        *out = nullptr;
        auto c = std::make_unique<esc_ctx>();
        ctx_init_host(c.get(), model, precision);
        hip_check(hipGetDevice(&c->device), "hipGetDevice");
        hip_check(esc::scratch_init(), "no-allocation workspace address");
        if (hipMalloc(&c->d_htab, sizeof(esc::HTables)) != hipSuccess) throw EscError(ESC_ERR_ALLOC, "table allocation failed");
        hip_check(hipMemcpy(c->d_htab, &c->htab, sizeof(esc::HTables), hipMemcpyHostToDevice), "hipMemcpy");
        c->tile_bits = default_tile_bits(c.get());
        *out = c.release();
    });
}

int esc_kernel_precompile(const esc_model *model, int precision, int config, char *source, size_t cap,
                          size_t *len) {
    return guarded([&] {
        esc_ctx c;
        ctx_init_host(&c, model, precision);
        if (!cfg_cube(config)) throw EscError(ESC_ERR_ARG, "not a generated-kernel configuration");
        esc::CubeModel m;
        std::vector<esc::CubePass> passes;
        std::string why;
        if (!tile_config_fits(&c, config) || !cube_spec(&c, config, m, passes, &why))
            throw EscError(ESC_ERR_UNSUPPORTED, "configuration does not apply to this model" +
                                                    (why.empty() ? std::string() : ": " + why));
        const std::string src = esc::cube_source(m, passes);
        if (len) *len = src.size();
        if (source && cap) {
            const size_t n = std::min(cap - 1, src.size());
            std::memcpy(source, src.data(), n);
            source[n] = '\0';
        }
        if (!esc::cube_precompile(m, passes, &why)) throw EscError(ESC_ERR_UNSUPPORTED, why);
    });
}

int esc_kernel_plan(const esc_model *model, int precision, int config, int64_t *passes, int cap_passes,
                    int *n_passes, int32_t *term_pass, uint32_t *term_mask, int cap_terms, int *n_terms) {
    return guarded([&] {
        check_ptr(n_passes, "n_passes");
        check_ptr(n_terms, "n_terms");
        esc_ctx c;
        ctx_init_host(&c, model, precision);
        if (config <= 0) throw EscError(ESC_ERR_ARG, "configuration 0 (one row per lane) has no passes");
        const std::vector<HostTerm> terms = host_terms(&c);
        struct Row { uint32_t W, T; int nnear, nfar; size_t lds; bool diag; };
        std::vector<Row> rows;
        std::vector<int> owner(terms.size(), -1);
        auto own = [&](uint32_t mask, int p) {
            for (size_t t = 0; t < terms.size(); ++t)
                if (terms[t].mask == mask) {
                    if (owner[t] >= 0) throw EscError(ESC_ERR_NUMERIC, "a term sits in two passes");
                    owner[t] = p;
                    return;
                }
            throw EscError(ESC_ERR_NUMERIC, "a pass applies a term the model does not have");
        };
        if (cfg_cube(config)) {
            esc::CubeModel m;
            std::vector<esc::CubePass> cp;
            std::string why;
            if (!tile_config_fits(&c, config) || !cube_spec(&c, config, m, cp, &why))
                throw EscError(ESC_ERR_UNSUPPORTED, "configuration does not apply to this model" +
                                                        (why.empty() ? std::string() : ": " + why));
            for (size_t p = 0; p < cp.size(); ++p) {
                Row r{0u, 0u, 0, 0, ((size_t)1 << cp[p].tile_bits) * c.esz, cp[p].diag};
                for (size_t i = 0; i < cp[p].wb.size(); ++i) {
                    r.W |= 1u << cp[p].wb[i];
                    if ((int)i < cp[p].tile_bits) r.T |= 1u << cp[p].wb[i];
                }
                for (const esc::CubeTerm &t : cp[p].terms) {
                    ((t.mask & ~r.T) ? r.nfar : r.nnear)++;
                    own(t.mask, (int)p);
                }
                rows.push_back(r);
            }
        } else {
            const int TB = cfg_tile_bits(config), w = cfg_window(config), P = cfg_passes(config);
            if (!tile_config_fits(&c, config)) throw EscError(ESC_ERR_UNSUPPORTED, "configuration does not fit");
            const std::vector<Plan> plan = plan_windows(&c, terms, TB, w, P);
            const std::vector<esc::GTables> g = build_passes(&c, TB, w, P, cfg_waves(config));
            if (g.size() != plan.size()) throw EscError(ESC_ERR_NUMERIC, "plan and tables disagree");
            for (size_t p = 0; p < plan.size(); ++p) {
                Row r{plan[p].W, plan[p].T, (int)plan[p].near.size(), (int)plan[p].far.size(),
                      esc::tile_lds_bytes(c.c64(), TB, (int)g[p].cls_begin[4], (int)g[p].cls_begin[esc::kNumClasses]),
                      g[p].first_pass != 0};
                for (size_t t : plan[p].near) own(terms[t].mask, (int)p);
                for (size_t t : plan[p].far) own(terms[t].mask, (int)p);
                rows.push_back(r);
            }
        }
        *n_passes = (int)rows.size();
        *n_terms = (int)terms.size();
        for (int p = 0; p < (int)rows.size() && passes && p < cap_passes; ++p) {
            int64_t *o = passes + 6 * (size_t)p;
            o[0] = rows[p].W; o[1] = rows[p].T; o[2] = rows[p].nnear; o[3] = rows[p].nfar;
            o[4] = (int64_t)rows[p].lds; o[5] = rows[p].diag ? 1 : 0;
        }
        for (int t = 0; t < (int)terms.size() && t < cap_terms; ++t) {
            if (term_pass) term_pass[t] = owner[t];
            if (term_mask) term_mask[t] = terms[t].mask;
        }
    });
}

int esc_ctx_destroy(esc_ctx *ctx) {
    return guarded([&] { delete ctx; });
}

int esc_ctx_info(const esc_ctx *ctx, int64_t *dim, int32_t *elem_bytes, int32_t *precision) {
    return guarded([&] {
        check_ctx(ctx);
        if (dim) *dim = (int64_t)ctx->dim;
        if (elem_bytes) *elem_bytes = (int32_t)ctx->esz;
        if (precision) *precision = ctx->precision;
    });
}

int esc_hpsi(esc_ctx *ctx, const void *d_in, void *d_out, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_in, "d_in");
        check_ptr(d_out, "d_out");
        if (d_in == d_out) throw EscError(ESC_ERR_ARG, "d_in and d_out must not overlap");
        run_hpsi(ctx, d_in, d_out, (hipStream_t)stream);
    });
}

int esc_apply_op(esc_ctx *ctx, int op, int site, const void *d_in, void *d_out, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_in, "d_in");
        check_ptr(d_out, "d_out");
        if (d_in == d_out) throw EscError(ESC_ERR_ARG, "d_in and d_out must not overlap");
        esc::CurrentTable *t = current_table(ctx, op, site);
        hipStream_t s = (hipStream_t)stream;
        auto ht = std::make_unique<esc::CurrentTable>();
        host_current_table(ctx, op, site, ht.get());
        apply_current_table(ctx, *ht, t, d_in, d_out, s);
    });
}

int esc_start_vector(esc_ctx *ctx, uint64_t seed, int kind, void *d_out, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_out, "d_out");
        hipStream_t s = (hipStream_t)stream;
        if (kind == ESC_START_REF_UNIFORM) {
            // mt19937_64(seed) -> uniform_real_distribution<double>(0,1) as
            // generate_canonical with one 64-bit draw: double(x) * 2^-64
            // (libc++ convention, no clamp); re drawn before im.
            // Serial on the host as in the reference: bounded to 2^28 rows.
            if (ctx->n_sites > 28)
                throw EscError(ESC_ERR_UNSUPPORTED, "the reference (mt19937_64) start vector is limited to N <= 28; "
                                                    "use the Gaussian start vector");
            const uint64_t dim = ctx->dim;
            std::vector<double> h(2 * dim);
            std::mt19937_64 g(seed);
            const double inv = 1.0 / 18446744073709551616.0;
            double n2 = 0.0;
            for (uint64_t i = 0; i < 2 * dim; ++i) {
                const double u = (double)g() * inv;
                h[i] = -1.0 + u * 2.0;
            }
            for (uint64_t i = 0; i < 2 * dim; ++i) n2 += h[i] * h[i];
            const double nrm = std::sqrt(n2);
            for (auto &x : h) x /= nrm;    // Eigen normalize(): divide by the 2-norm
            void *tmp = nullptr;
            if (hipMalloc(&tmp, 2 * dim * sizeof(double)) != hipSuccess) throw EscError(ESC_ERR_ALLOC, "staging allocation failed");
            hipError_t e = hipMemcpyAsync(tmp, h.data(), 2 * dim * sizeof(double), hipMemcpyHostToDevice, s);
            if (e == hipSuccess) e = esc::launch_convert_from_c128(ctx->c64(), tmp, d_out, dim, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            (void)hipFree(tmp);
            hip_check(e, "start vector upload");
        } else if (kind == ESC_START_GAUSSIAN) {
            hip_check(esc::launch_gaussian(ctx->c64(), seed, d_out, ctx->dim, s), "gaussian");
            ctx->ensure_partials(2 * esc::kRedBlocks);
            ctx->ensure_result(4);
            hip_check(esc::launch_norm2(ctx->c64(), d_out, ctx->dim, ctx->d_partials, ctx->d_result, s), "norm2");
            hip_check(esc::launch_normalize(ctx->c64(), d_out, ctx->d_result, ctx->dim, s), "normalize");
        } else {
            throw EscError(ESC_ERR_ARG, "unknown start-vector kind");
        }
    });
}

int esc_dot(esc_ctx *ctx, const void *d_x, const void *d_y, double *out, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_x, "d_x");
        check_ptr(d_y, "d_y");
        check_ptr(out, "out");
        device_dot(ctx, d_x, d_y, out, (hipStream_t)stream);
    });
}

int esc_dot_noalloc(esc_ctx *ctx, const void *d_x, const void *d_y, double *d_out, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_x, "d_x");
        check_ptr(d_y, "d_y");
        check_ptr(d_out, "d_out");
        hip_check(esc::launch_dot_noalloc(ctx->c64(), d_x, d_y, ctx->dim, d_out, (hipStream_t)stream), "dot");
    });
}

int esc_axpby(esc_ctx *ctx, const double *a, const void *d_x, const double *b, void *d_y, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(a, "a");
        check_ptr(b, "b");
        check_ptr(d_x, "d_x");
        check_ptr(d_y, "d_y");
        hip_check(esc::launch_axpby(ctx->c64(), a[0], a[1], d_x, b[0], b[1], d_y, ctx->dim, (hipStream_t)stream), "axpby");
    });
}

int esc_upload(esc_ctx *ctx, const double *host, void *d_out, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(host, "host");
        check_ptr(d_out, "d_out");
        hipStream_t s = (hipStream_t)stream;
        if (!ctx->c64()) {
            hip_check(hipMemcpyAsync(d_out, host, ctx->dim * 16, hipMemcpyHostToDevice, s), "upload");
            hip_check(hipStreamSynchronize(s), "sync");
            return;
        }
        void *tmp = nullptr;
        if (hipMalloc(&tmp, ctx->dim * 16) != hipSuccess) throw EscError(ESC_ERR_ALLOC, "staging allocation failed");
        hipError_t e = hipMemcpyAsync(tmp, host, ctx->dim * 16, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = esc::launch_convert_from_c128(true, tmp, d_out, ctx->dim, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(tmp);
        hip_check(e, "upload");
    });
}

int esc_download(esc_ctx *ctx, const void *d_in, double *host, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(host, "host");
        check_ptr(d_in, "d_in");
        hipStream_t s = (hipStream_t)stream;
        if (!ctx->c64()) {
            hip_check(hipMemcpyAsync(host, d_in, ctx->dim * 16, hipMemcpyDeviceToHost, s), "download");
            hip_check(hipStreamSynchronize(s), "sync");
            return;
        }
        void *tmp = nullptr;
        if (hipMalloc(&tmp, ctx->dim * 16) != hipSuccess) throw EscError(ESC_ERR_ALLOC, "staging allocation failed");
        hipError_t e = esc::launch_convert_to_c128(true, d_in, tmp, ctx->dim, s);
        if (e == hipSuccess) e = hipMemcpyAsync(host, tmp, ctx->dim * 16, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(tmp);
        hip_check(e, "download");
    });
}

int esc_lanczos(esc_ctx *ctx, const void *d_v0, int steps, double eps, int mode, double *alpha, double *beta,
                int *iters, double *evals, double *Z, void *d_gs, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_v0, "d_v0");
        check_ptr(alpha, "alpha");
        check_ptr(beta, "beta");
        check_ptr(iters, "iters");
        if (steps < 2 || steps > 4096) throw EscError(ESC_ERR_ARG, "steps must be in [2, 4096]");
        if (mode < ESC_LANC_ONLYEGS || mode > ESC_LANC_PLAIN) throw EscError(ESC_ERR_ARG, "unknown Lanczos mode");
        const bool reortho = mode == ESC_LANC_GS || mode == ESC_LANC_ALL;
        if ((mode == ESC_LANC_GS || mode == ESC_LANC_ALL) && !d_gs) throw EscError(ESC_ERR_ARG, "d_gs required for GS/All");
        hipStream_t s = (hipStream_t)stream;
        const bool c64 = ctx->c64();
        const uint64_t dim = ctx->dim;
        // storage: reortho modes keep V_0..V_steps; the others rotate three
        // vectors.  The new vector is formed in place in the V_{k+1} slot.
        const size_t nvec = reortho ? (size_t)steps + 1 : 3;
        ctx->ensure_vecs(nvec);
        ctx->basis_kept = false;
        ctx->ensure_partials((size_t)2 * (steps + 2) * 256 + 2 * esc::kRedBlocks);
        ctx->ensure_result((size_t)2 * (steps + 2) + 8);
        double *d_alpha = ctx->d_result;   // alpha lives in result[0..1] during updates
        auto V = [&](int m) -> void * { return reortho ? ctx->vec((size_t)m) : ctx->vec((size_t)(m % 3)); };

        hip_check(hipMemcpyAsync(V(0), d_v0, dim * ctx->esz, hipMemcpyDeviceToDevice, s), "copy v0");
        std::vector<double> ar, ai, bt;
        bt.push_back(0.0);
        double E0 = 1.0, D0 = 0.0;
        int it = 0;
        auto step = [&](int k) {
            // w = H V_k; alpha_k = <V_k|w>; w -= alpha_k V_k + beta_k V_{k-1}
            void *w = V(k + 1);
            run_hpsi(ctx, V(k), w, s);
            double a[2];
            hip_check(esc::launch_dot(c64, V(k), w, dim, ctx->d_partials, d_alpha, s), "dot");
            sync_get(s, a, d_alpha, 2 * sizeof(double));
            if (!std::isfinite(a[0]) || !std::isfinite(a[1])) throw EscError(ESC_ERR_NUMERIC, "non-finite alpha");
            ar.push_back(a[0]);
            ai.push_back(a[1]);
            hip_check(esc::launch_lanczos_update(c64, w, V(k), k > 0 ? V(k - 1) : nullptr, d_alpha, bt[k], dim,
                                                 ctx->d_partials, ctx->d_result + 2, s), "update");
            double n2;
            sync_get(s, &n2, ctx->d_result + 2, sizeof(double));
            const double b = std::sqrt(n2);
            if (!std::isfinite(b)) throw EscError(ESC_ERR_NUMERIC, "non-finite beta");
            // Invariant subspace: beta at rounding level of |alpha|.  With full
            // reorthogonalisation the normalised remainder is re-orthogonalised
            // into a fresh direction and the run continues (as the reference
            // does); without it the recursion would continue on noise.
            if (b == 0.0 || (!reortho && b <= 1e-13 * std::max(1.0, std::fabs(a[0]))))
                throw EscError(ESC_ERR_BREAKDOWN, "Lanczos breakdown: beta ~ 0 at step " + std::to_string(k));
            bt.push_back(b);
            hip_check(esc::launch_scale(c64, w, 1.0 / b, dim, s), "scale");
            if (reortho && k > 0) {   // (src/DynLanczosSolver.h:105-119), no reorthogonalisation at k = 0
                // two classical passes always: the orthogonality of the
                // reference's modified Gram-Schmidt, on which its Ritz spectra
                // (and their reproduction) depend
                cgs_pass(ctx, k + 1, w, s);
                cgs_pass(ctx, k + 1, w, s);
                hip_check(esc::launch_norm2(c64, w, dim, ctx->d_partials, ctx->d_result + 4, s), "norm2");
                hip_check(esc::launch_normalize(c64, w, ctx->d_result + 4, dim, s), "normalize");
            }
        };
        step(0);
        bool done = false;
        while (!done) {
            ++it;
            step(it);
            D0 = tridiag_lowest((int)ar.size(), ar.data(), bt.data());
            const double diff = std::fabs(E0 - D0);
            if (diff < eps) done = true;
            E0 = D0;
            if (it == steps - 1) done = true;
        }
        const int n = it + 1;
        *iters = n;
        for (int k = 0; k < n; ++k) alpha[k] = ar[k];
        for (int k = 0; k <= n; ++k) beta[k] = bt[k];
        ctx->Z.assign((size_t)n * n, 0.0);
        std::vector<double> ev(n);
        tridiag_eigenpairs(n, ar.data(), bt.data(), ev.data(), ctx->Z.data(), true, s);
        ctx->z_n = n;
        ctx->evals = ev;
        if (evals) std::memcpy(evals, ev.data(), sizeof(double) * n);
        if (Z) std::memcpy(Z, ctx->Z.data(), sizeof(double) * n * n);
        if (reortho) {
            // |gs> = sum_m Z(m,0) V_m  (src/DynLanczosSolver.h:213-223 for k = 0)
            std::vector<double> c(2 * (size_t)n);
            for (int m = 0; m < n; ++m) { c[2 * m] = ctx->Z[m]; c[2 * m + 1] = 0.0; }
            hip_check(hipMemcpyAsync(ctx->d_result, c.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, s), "coef");
            hip_check(esc::launch_multi_axpy(c64, ctx->vec(0), n, ctx->d_result, 1.0, d_gs, dim, true, s), "gs");
            hip_check(hipStreamSynchronize(s), "sync");
            ctx->basis_kept = mode == ESC_LANC_ALL;
            ctx->basis_n = n;
        }
    });
}

// Solver=ED: H as a dense column-major matrix in the context's basis
// storage, column k = H e_k from the context's H.psi kernel (every element is
// the kernel's own, so bit-exact), then all eigenpairs by rocSOLVER's
// divide-and-conquer zheevd on the device.  The eigenvectors stay in place as
// the kept basis with an identity rotation, so everything that reads the
// Ritz states of an ESC_LANC_ALL run reads the eigenstates instead
// (src/Exactdiag.h:36-81, src/DynLanczosSolver.h:269-283).  The reference's
// matrix is never zero-initialised before the += (src/Exactdiag.h:42); here it
// is.
int esc_exact_diag(esc_ctx *ctx, double *evals, void *d_gs, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(evals, "evals");
        if (ctx->c64()) throw EscError(ESC_ERR_ARG, "Solver=ED runs in complex128 (Precision=c128)");
        if (ctx->n_sites > kMaxEdSites)
            throw EscError(ESC_ERR_ARG, "Solver=ED needs NumberofSites <= " + std::to_string(kMaxEdSites) +
                                            " (dense 2^N x 2^N matrix); use the Lanczos solver");
        hipStream_t s = (hipStream_t)stream;
        const uint64_t dim = ctx->dim;
        const size_t vb = dim * ctx->esz;
        ctx->ensure_vecs(dim);
        ctx->basis_kept = false;
        ctx->ensure_aux(1);
        const std::complex<double> one(1.0, 0.0);
        for (uint64_t k = 0; k < dim; ++k) {
            hip_check(hipMemsetAsync(ctx->aux(0), 0, vb, s), "hipMemsetAsync");
            hip_check(hipMemcpyAsync((char *)ctx->aux(0) + k * ctx->esz, &one, sizeof(one), hipMemcpyHostToDevice, s),
                      "unit vector");
            run_hpsi(ctx, ctx->aux(0), ctx->vec(k), s);
        }
        double *d_work = nullptr;
        rocblas_int *d_info = nullptr;
        rocblas_handle h = nullptr;
        auto release = [&] {
            if (h) (void)rocblas_destroy_handle(h);
            if (d_work) (void)hipFree(d_work);
            if (d_info) (void)hipFree(d_info);
        };
        try {
            if (hipMalloc(&d_work, 2 * dim * sizeof(double)) != hipSuccess ||
                hipMalloc(&d_info, sizeof(rocblas_int)) != hipSuccess)
                throw EscError(ESC_ERR_ALLOC, "ED workspace allocation failed");
            if (rocblas_create_handle(&h) != rocblas_status_success || rocblas_set_stream(h, s) != rocblas_status_success)
                throw EscError(ESC_ERR_HIP, "rocBLAS handle creation failed");
            const rocblas_status st =
                rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)dim,
                                 reinterpret_cast<rocblas_double_complex *>(ctx->vec(0)), (rocblas_int)dim, d_work,
                                 d_work + dim, d_info);
            if (st != rocblas_status_success)
                throw EscError(ESC_ERR_NUMERIC, "rocsolver_zheevd failed (status " + std::to_string((int)st) + ")");
            rocblas_int info = 0;
            std::vector<double> ev(dim);
            sync_get(s, &info, d_info, sizeof(info));
            if (info != 0) throw EscError(ESC_ERR_NUMERIC, "zheevd: " + std::to_string(info) + " off-diagonal elements did not converge");
            sync_get(s, ev.data(), d_work, dim * sizeof(double));
            release();
            h = nullptr; d_work = nullptr; d_info = nullptr;
            for (double e : ev)
                if (!std::isfinite(e)) throw EscError(ESC_ERR_NUMERIC, "non-finite eigenvalue");
            std::memcpy(evals, ev.data(), dim * sizeof(double));
            const int n = (int)dim;
            ctx->Z.assign((size_t)n * n, 0.0);
            for (int k = 0; k < n; ++k) ctx->Z[(size_t)k * n + k] = 1.0;
            ctx->z_n = n;
            ctx->evals = ev;
            ctx->basis_n = n;
            ctx->basis_kept = true;
            if (d_gs) hip_check(hipMemcpyAsync(d_gs, ctx->vec(0), vb, hipMemcpyDeviceToDevice, s), "copy gs");
            hip_check(hipStreamSynchronize(s), "sync");
        } catch (...) {
            release();
            throw;
        }
    });
}

int esc_ground_state(esc_ctx *ctx, const void *d_v0, int steps, double eps, int restarts, double *e0, double *resid,
                     int *iters, void *d_gs, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_v0, "d_v0");
        check_ptr(d_gs, "d_gs");
        if (d_v0 == d_gs) throw EscError(ESC_ERR_ARG, "d_v0 and d_gs must not overlap");
        if (steps < 2 || steps > 4096) throw EscError(ESC_ERR_ARG, "steps must be in [2, 4096]");
        if (restarts < 0 || restarts > 64) throw EscError(ESC_ERR_ARG, "restarts must be in [0, 64]");
        hipStream_t s = (hipStream_t)stream;
        const bool c64 = ctx->c64();
        const uint64_t dim = ctx->dim;
        ctx->ensure_vecs(4);
        ctx->basis_kept = false;
        void *start = ctx->vec(3);
        hip_check(hipMemcpyAsync(start, d_v0, dim * ctx->esz, hipMemcpyDeviceToDevice, s), "copy v0");
        const double n0 = device_norm2(ctx, start, s);
        if (!(n0 > 0.0) || !std::isfinite(n0)) throw EscError(ESC_ERR_NUMERIC, "start vector has zero or non-finite norm");
        hip_check(esc::launch_scale(c64, start, 1.0 / std::sqrt(n0), dim, s), "scale");
        long total = 0;
        for (int cyc = 0; cyc <= restarts; ++cyc) {
            std::vector<double> ar, bt;
            // restarts stop once the Ritz residual reaches the storage's
            // precision (a few units of 2^-24 or 2^-53 of |T|): the vector
            // cannot hold a better state
            const double floor = cyc > 0 ? 4.0 * (c64 ? 5.96e-8 : 1.11e-16) : 0.0;
            const int n = plain_recursion(ctx, start, steps, eps, nullptr, 0, nullptr, ar, bt, s, floor);
            total += n;
            progress("ground state cycle %d: recursion of %d steps", cyc, n);
            // the lowest Ritz pair (the routine the batched ground states run on
            // the device): its eigenvector gives the replay coefficients
            bool ok = false;
            const double lam = esc::td_lowest(n, ar.data(), bt.data(), &ok);
            if (!ok) throw EscError(ESC_ERR_NUMERIC, "non-finite tridiagonal matrix");
            std::vector<double> z(n), work(5 * (size_t)n);
            esc::td_vector(n, ar.data(), bt.data(), lam, work.data(), z.data());
            std::vector<std::complex<double>> coef(z.begin(), z.end());
            hip_check(hipMemsetAsync(d_gs, 0, dim * ctx->esz, s), "clear");
            std::vector<double> ar2, bt2;
            plain_recursion(ctx, start, steps, eps, coef.data(), n, d_gs, ar2, bt2, s, 0.0);
            total += n - 1;
            progress("ground state cycle %d: replayed", cyc);
            const double ng = device_norm2(ctx, d_gs, s);
            hip_check(esc::launch_scale(c64, d_gs, 1.0 / std::sqrt(ng), dim, s), "scale");
            if (cyc < restarts) hip_check(hipMemcpyAsync(start, d_gs, dim * ctx->esz, hipMemcpyDeviceToDevice, s), "copy");
        }
        // Rayleigh quotient and residual |H gs - e gs| in FP64 reductions
        void *w = ctx->vec(0);
        run_hpsi(ctx, d_gs, w, s);
        ++total;
        double d[2];
        device_dot(ctx, d_gs, w, d, s);
        hip_check(esc::launch_axpby(c64, -d[0], 0.0, d_gs, 1.0, 0.0, w, dim, s), "residual");
        const double r2 = device_norm2(ctx, w, s);
        if (e0) *e0 = d[0];
        if (resid) *resid = std::sqrt(r2);
        if (iters) *iters = (int)total;
    });
}

int esc_two_point(esc_ctx *ctx, int op_a, int op_b, const void *d_gs, double *corr, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_gs, "d_gs");
        check_ptr(corr, "corr");
        if (op_a < ESC_OP_SX || op_a > ESC_OP_JEY || op_b < ESC_OP_SX || op_b > ESC_OP_JEY)
            throw EscError(ESC_ERR_ARG, "unknown operator");
        hipStream_t s = (hipStream_t)stream;
        const int N = ctx->n_sites;
        // a_i = A_i|gs> kept for all i (the context's vectors), b_j = B_j|gs>
        // one at a time; Corr(i, j) = <gs|A_i B_j|gs> = <a_i|b_j> (every
        // operator here is Hermitian: Pauli strings with real weights)
        ctx->ensure_vecs((size_t)N);
        ctx->basis_kept = false;
        ctx->ensure_aux(1);
        ctx->ensure_partials((size_t)2 * N * 256 + 2 * esc::kRedBlocks);
        ctx->ensure_result((size_t)2 * N + 8);
        auto apply = [&](int op, int site, void *out) {
            esc::CurrentTable *t = current_table(ctx, op, site);
            hip_check(ctx->c64() ? esc::launch_current_c64(t, d_gs, out, ctx->dim, s)
                                 : esc::launch_current_c128(t, d_gs, out, ctx->dim, s), "operator");
        };
        for (int i = 0; i < N; ++i) apply(op_a, i, ctx->vec((size_t)i));
        std::vector<double> tv(2 * (size_t)N);
        for (int j = 0; j < N; ++j) {
            apply(op_b, j, ctx->aux(0));
            hip_check(esc::launch_multidot(ctx->c64(), ctx->vec(0), N, ctx->aux(0), ctx->dim, ctx->d_partials,
                                           ctx->d_result, s), "multidot");
            sync_get(s, tv.data(), ctx->d_result, sizeof(double) * 2 * N);
            // the reference fills si >= sj and mirrors it (src/Observables.h:142-151)
            for (int i = j; i < N; ++i) {
                for (int q = 0; q < 2; ++q) {
                    corr[2 * ((size_t)i * N + j) + q] = tv[2 * i + q];
                    corr[2 * ((size_t)j * N + i) + q] = tv[2 * i + q];
                }
            }
        }
    });
}

int esc_ritz_overlaps(esc_ctx *ctx, const void *d_vec, int n_states, double *out, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_vec, "d_vec");
        check_ptr(out, "out");
        if (!ctx->basis_kept) throw EscError(ESC_ERR_ARG, "no Lanczos basis kept (run esc_lanczos with ESC_LANC_ALL)");
        const int n = ctx->basis_n;
        if (n_states < 0 || n_states > n) throw EscError(ESC_ERR_ARG, "n_states out of range");
        hipStream_t s = (hipStream_t)stream;
        ctx->ensure_partials((size_t)2 * n * 256);
        ctx->ensure_result((size_t)2 * n + 8);
        hip_check(esc::launch_multidot(ctx->c64(), ctx->vec(0), n, d_vec, ctx->dim, ctx->d_partials, ctx->d_result, s), "multidot");
        std::vector<double> t(2 * (size_t)n);
        sync_get(s, t.data(), ctx->d_result, sizeof(double) * 2 * n);
        for (int k = 0; k < n_states; ++k) {
            double re = 0.0, im = 0.0;
            for (int m = 0; m < n; ++m) {
                const double z = ctx->Z[(size_t)k * n + m];
                re += z * t[2 * m];
                im += z * t[2 * m + 1];
            }
            out[2 * k] = re;
            out[2 * k + 1] = im;
        }
    });
}

int esc_ritz_vector(esc_ctx *ctx, int k, void *d_out, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_out, "d_out");
        if (!ctx->basis_kept) throw EscError(ESC_ERR_ARG, "no Lanczos basis kept (run esc_lanczos with ESC_LANC_ALL)");
        const int n = ctx->basis_n;
        if (k < 0 || k >= n) throw EscError(ESC_ERR_ARG, "state index out of range");
        hipStream_t s = (hipStream_t)stream;
        std::vector<double> c(2 * (size_t)n);
        for (int m = 0; m < n; ++m) { c[2 * m] = ctx->Z[(size_t)k * n + m]; c[2 * m + 1] = 0.0; }
        ctx->ensure_result((size_t)2 * n + 8);
        hip_check(hipMemcpyAsync(ctx->d_result, c.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, s), "coef");
        hip_check(esc::launch_multi_axpy(ctx->c64(), ctx->vec(0), n, ctx->d_result, 1.0, d_out, ctx->dim, true, s), "ritz");
        hip_check(hipStreamSynchronize(s), "sync");
    });
}

int esc_lanczos_cf(esc_ctx *ctx, const void *d_phi, int steps, double *alpha, double *beta, int *iters,
                   double *norm2, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_phi, "d_phi");
        ctx->ensure_vecs(3);
        ctx->basis_kept = false;
        void *work[3] = {ctx->vec(0), ctx->vec(1), ctx->vec(2)};
        cf_run(ctx, d_phi, steps, work, alpha, beta, iters, norm2, (hipStream_t)stream);
    });
}

int esc_dynamics(esc_ctx *ctx, int kind, int method, const void *d_gs, double e0, int cf_steps, double eta,
                 const double *omega, int nw, double *table, double *expect, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_gs, "d_gs");
        check_ptr(omega, "omega");
        check_ptr(table, "table");
        if (kind != ESC_DYN_JSJS && kind != ESC_DYN_JEJE && kind != ESC_DYN_LOCALSZSZ)
            throw EscError(ESC_ERR_ARG, "unknown dynamics kind");
        if (method != ESC_SPEC_RITZ && method != ESC_SPEC_CF && method != ESC_SPEC_CF_REORTH)
            throw EscError(ESC_ERR_ARG, "unknown spectral method");
        const bool reortho = method == ESC_SPEC_CF_REORTH;
        if (reortho) method = ESC_SPEC_CF;
        if (nw < 1) throw EscError(ESC_ERR_ARG, "nw must be >= 1");
        if (method == ESC_SPEC_RITZ && !ctx->basis_kept)
            throw EscError(ESC_ERR_ARG, "Ritz spectra need the basis of an ESC_LANC_ALL run (LancType=All)");
        if (method == ESC_SPEC_CF && (cf_steps < 1 || cf_steps > (reortho ? 4096 : 100000)))
            throw EscError(ESC_ERR_ARG, "cf_steps out of range");
        hipStream_t s = (hipStream_t)stream;
        if (kind == ESC_DYN_LOCALSZSZ) {
            local_szsz_spectra(ctx, method, reortho, d_gs, e0, cf_steps, eta, omega, nw, table, expect, s);
            return;
        }
        const bool js = kind == ESC_DYN_JSJS;
        const std::vector<int> ops = js ? std::vector<int>{ESC_OP_JSX, ESC_OP_JSY, ESC_OP_JSZ}
                                        : std::vector<int>{ESC_OP_JEX, ESC_OP_JEY};
        const int nc = (int)ops.size();
        using cplx = std::complex<double>;
        // continued fractions run in the context's Lanczos vectors unless those
        // hold a kept basis (then in separate work vectors): at N = 32 the
        // ground state, J|gs> and three work vectors are all that is resident
        const bool cf_in_vecs = method == ESC_SPEC_CF && !ctx->basis_kept;
        if (cf_in_vecs) ctx->ensure_vecs(3);
        ctx->ensure_aux(method == ESC_SPEC_CF && !cf_in_vecs ? 4 : 1);
        ctx->ensure_partials(2 * esc::kRedBlocks);
        ctx->ensure_result(8);
        std::vector<esc::CurrentTable> tabs(nc);
        for (int a = 0; a < nc; ++a) host_current_table(ctx, ops[a], ESC_SITE_TOTAL, &tabs[a]);
        if (!ctx->d_combo && hipMalloc(&ctx->d_combo, sizeof(esc::CurrentTable)) != hipSuccess)
            throw EscError(ESC_ERR_ALLOC, "table allocation failed");
        // phi = (sum_a c_a J_a) |gs> into aux(0), from one combined gather table
        auto apply_combo = [&](const std::vector<cplx> &coef) {
            esc::CurrentTable t;
            std::memset(&t, 0, sizeof(t));
            for (int a = 0; a < nc; ++a)
                if (coef[a] != cplx(0.0, 0.0)) append_scaled(t, tabs[a], coef[a].real(), coef[a].imag());
            hip_check(hipMemcpy(ctx->d_combo, &t, sizeof(t), hipMemcpyHostToDevice), "hipMemcpy");
            apply_current_table(ctx, t, ctx->d_combo, d_gs, ctx->aux(0), s);
        };
        auto unit = [&](int a) {
            std::vector<cplx> c(nc, cplx(0.0, 0.0));
            c[a] = 1.0;
            return c;
        };
        // <J_a> = <gs|J_a|gs>  (src/dynamics.h:250-253,370-372); t = sum of components
        std::vector<cplx> ex(nc + 1, cplx(0.0, 0.0));
        std::vector<std::vector<cplx>> O(nc + 1);
        const int nst = method == ESC_SPEC_RITZ ? ctx->basis_n : 0;
        for (int a = 0; a < nc; ++a) {
            apply_combo(unit(a));
            double d[2];
            device_dot(ctx, d_gs, ctx->aux(0), d, s);
            ex[a] = cplx(d[0], d[1]);
            ex[nc] += ex[a];
            if (method == ESC_SPEC_RITZ) {
                // O(n,a) = <psi_n|J_a gs> = sum_m Z(m,n) <v_m|J_a gs>  (src/dynamics.h:255-263,375-383)
                const int n = ctx->basis_n;
                ctx->ensure_partials((size_t)2 * n * 256);
                ctx->ensure_result((size_t)2 * n + 8);
                hip_check(esc::launch_multidot(ctx->c64(), ctx->vec(0), n, ctx->aux(0), ctx->dim, ctx->d_partials,
                                               ctx->d_result, s), "multidot");
                std::vector<double> tv(2 * (size_t)n);
                sync_get(s, tv.data(), ctx->d_result, sizeof(double) * 2 * n);
                O[a].assign(nst, cplx(0.0, 0.0));
                for (int k = 0; k < nst; ++k) {
                    cplx acc(0.0, 0.0);
                    for (int m = 0; m < n; ++m) acc += ctx->Z[(size_t)k * n + m] * cplx(tv[2 * m], tv[2 * m + 1]);
                    O[a][k] = acc;
                }
            }
        }
        if (expect)
            for (int a = 0; a <= nc; ++a) { expect[2 * a] = ex[a].real(); expect[2 * a + 1] = ex[a].imag(); }
        // spectra I_ab(w) before the elastic subtraction
        std::map<std::pair<int, int>, std::vector<cplx>> F;
        const int T = nc;   // index of the total channel
        std::vector<std::pair<int, int>> pairs;
        if (js) pairs = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}, {T, T}};
        else pairs = {{0, 0}, {1, 1}, {0, 1}, {1, 0}, {T, T}};
        if (method == ESC_SPEC_RITZ) {
            O[T].assign(nst, cplx(0.0, 0.0));
            for (int a = 0; a < nc; ++a)
                for (int k = 0; k < nst; ++k) O[T][k] += O[a][k];
            const double Eg = ctx->evals[0];
            for (auto pr : pairs) {
                std::vector<cplx> f(nw);
                for (int q = 0; q < nw; ++q) {
                    cplx acc(0.0, 0.0);
                    for (int k = 0; k < nst; ++k) {
                        const cplx den(omega[q] - (ctx->evals[k] - Eg), -eta);
                        acc = acc + std::conj(O[pr.first][k]) * O[pr.second][k] / den;
                    }
                    f[q] = acc;
                }
                F[pr] = f;
            }
        } else {
            void *work[3] = {cf_in_vecs ? ctx->vec(0) : ctx->aux(1), cf_in_vecs ? ctx->vec(1) : ctx->aux(2),
                             cf_in_vecs ? ctx->vec(2) : ctx->aux(3)};
            std::vector<double> al(cf_steps + 1), be(cf_steps + 2);
            const std::vector<CfRun> runs = cf_runs(js, nc);
            std::vector<std::vector<cplx>> f;
            for (const CfRun &r : runs) {
                apply_combo(r.coef);
                int it = 0;
                double n2 = 0.0;
                cf_run(ctx, ctx->aux(0), cf_steps, work, al.data(), be.data(), &it, &n2, s, reortho);
                f.push_back(cf_eval(it, al.data(), be.data(), n2, e0, eta, omega, nw));
                progress("continued fraction %zu of %zu: %d steps", f.size(), runs.size(), it);
            }
            F = cf_assemble(js, nc, runs, f, nw);
        }
        write_spectrum_table(js, nc, ex, F, omega, nw, eta, table);
    });
}

int esc_rixs(esc_ctx *gs_ctx, esc_ctx *int_ctx, int method, int site, const void *d_gs, double e0, int cf_steps,
             double gamma, double eta, const double *omega, int nw, double *table, void *stream) {
    return guarded([&] {
        check_ctx(gs_ctx);
        check_ctx(int_ctx);
        check_ptr(d_gs, "d_gs");
        check_ptr(omega, "omega");
        check_ptr(table, "table");
        if (gs_ctx->dim != int_ctx->dim || gs_ctx->precision != int_ctx->precision)
            throw EscError(ESC_ERR_ARG, "the two contexts must have the same size and precision");
        if (site < 0 || site >= gs_ctx->n_sites) throw EscError(ESC_ERR_ARG, "site out of range");
        if (method != ESC_SPEC_RITZ && method != ESC_SPEC_CF) throw EscError(ESC_ERR_ARG, "method must be Ritz or CF");
        if (!(gamma > 0.0) || !std::isfinite(gamma)) throw EscError(ESC_ERR_ARG, "gamma must be positive");
        if (!(eta > 0.0) || !std::isfinite(eta)) throw EscError(ESC_ERR_ARG, "eta must be positive");
        if (nw < 1) throw EscError(ESC_ERR_ARG, "nw must be >= 1");
        if (method == ESC_SPEC_RITZ && (!gs_ctx->basis_kept || !int_ctx->basis_kept))
            throw EscError(ESC_ERR_ARG, "Ritz RIXS needs the kept bases of ESC_LANC_ALL runs on both contexts");
        if (method == ESC_SPEC_CF && (cf_steps < 2 || cf_steps > 4096)) throw EscError(ESC_ERR_ARG, "cf_steps out of range");
        rixs_spectra(gs_ctx, int_ctx, method, site, d_gs, e0, cf_steps, gamma, eta, omega, nw, table,
                     (hipStream_t)stream);
    });
}

int esc_szsz(esc_ctx *ctx, const void *d_gs, double eta, const double *omega, int nw, double *table, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_gs, "d_gs");
        check_ptr(omega, "omega");
        check_ptr(table, "table");
        if (!ctx->basis_kept) throw EscError(ESC_ERR_ARG, "SzSz needs the kept basis of an ESC_LANC_ALL run");
        if (!(eta > 0.0) || !std::isfinite(eta)) throw EscError(ESC_ERR_ARG, "eta must be positive");
        if (nw < 1) throw EscError(ESC_ERR_ARG, "nw must be >= 1");
        hipStream_t s = (hipStream_t)stream;
        const int N = ctx->n_sites, n = ctx->basis_n;
        ctx->ensure_aux(1);
        // O_i(n) = <n|S^z_i gs> over the Ritz states of the kept basis
        std::vector<std::vector<cplx>> O(N);
        for (int i = 0; i < N; ++i) {
            esc::CurrentTable *t = current_table(ctx, ESC_OP_SZ, i);
            hip_check(ctx->c64() ? esc::launch_current_c64(t, d_gs, ctx->aux(0), ctx->dim, s)
                                 : esc::launch_current_c128(t, d_gs, ctx->aux(0), ctx->dim, s), "S^z");
            O[i] = ritz_amplitudes(ctx, ctx->aux(0), s);
        }
        // DynSzSz (src/dynamics.h:566-611): Re(<S^z_j gs|n><n|S^z_i gs>) times
        // x/(x^2 + eta^2) and eta/(x^2 + eta^2), x = w - (E_n - Eg), Eg = the
        // lowest Ritz value; table[w][i][j] = (real, imaginary) sums
        const double Eg = ctx->evals[0];
        std::vector<double> tmp((size_t)N * N * n);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < n; ++k) tmp[((size_t)i * N + j) * n + k] = (std::conj(O[j][k]) * O[i][k]).real();
        for (int qw = 0; qw < nw; ++qw)
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) {
                    double re = 0.0, im = 0.0;
                    const double *tp = &tmp[((size_t)i * N + j) * n];
                    for (int k = 0; k < n; ++k) {
                        const double x = omega[qw] - (ctx->evals[k] - Eg);
                        re += tp[k] * x / (x * x + eta * eta);
                        im += tp[k] * eta / (x * x + eta * eta);
                    }
                    double *out = table + (((size_t)qw * N + i) * N + j) * 2;
                    out[0] = re;
                    out[1] = im;
                }
    });
}

int esc_autotune(esc_ctx *ctx, const void *d_in, void *d_out, int reps, int *tile_bits, double *ms, void *stream) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(d_in, "d_in");
        check_ptr(d_out, "d_out");
        if (d_in == d_out) throw EscError(ESC_ERR_ARG, "d_in and d_out must not overlap");
        if (reps < 1) reps = 1;
        hipStream_t s = (hipStream_t)stream;
        std::vector<int> cand = tile_candidates(ctx);
        cand.push_back(0);   // one-row-per-lane kernel
        cube_compile_parallel(ctx, cand);
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        int best = ctx->tile_bits;
        float best_ms = 1e30f;
        for (int tb : cand) {
            try {
                launch_hpsi_with(ctx, tb, d_in, d_out, s);   // warm-up (also builds tables)
            } catch (const EscError &e) {
                if (e.code != ESC_ERR_UNSUPPORTED) throw;
                continue;   // a generated configuration with no plan or program for this model
            }
            hip_check(hipEventRecord(e0, s), "hipEventRecord");
            for (int r = 0; r < reps; ++r) launch_hpsi_with(ctx, tb, d_in, d_out, s);
            hip_check(hipEventRecord(e1, s), "hipEventRecord");
            hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
            float t = 0.0f;
            hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
            t /= (float)reps;
            if (t < best_ms) { best_ms = t; best = tb; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ctx->tile_bits = best;
        if (tile_bits) *tile_bits = best;
        if (ms) *ms = best_ms;
    });
}

int esc_kernel_configs(const esc_ctx *ctx, int32_t *configs, int32_t cap, int32_t *n) {
    return guarded([&] {
        check_ctx(ctx);
        check_ptr(n, "n");
        std::vector<int> v = tile_candidates(ctx);
        v.insert(v.begin(), 0);
        *n = (int32_t)v.size();
        if (configs)
            for (int32_t i = 0; i < cap && i < (int32_t)v.size(); ++i) configs[i] = v[i];
    });
}

int esc_set_tile_bits(esc_ctx *ctx, int tile_bits) {
    return guarded([&] {
        check_ctx(ctx);
        if (tile_bits != 0) {
            // any generated-kernel configuration that plans for this context,
            // else one of the candidates
            const auto cand = tile_candidates(ctx);
            const bool cube_ok = cfg_cube(tile_bits) && tile_bits < 40000000 && tile_config_fits(ctx, tile_bits);
            if (!cube_ok && std::find(cand.begin(), cand.end(), tile_bits) == cand.end())
                throw EscError(ESC_ERR_ARG, "tile size not available for this context");
        }
        ctx->tile_bits = tile_bits;
    });
}

int esc_tridiag_eig(int n, const double *a, const double *b, double *evals, double *Z) {
    return guarded([&] {
        if (n < 1 || n > 4096) throw EscError(ESC_ERR_ARG, "n must be in [1, 4096]");
        check_ptr(a, "a");
        check_ptr(b, "b");
        check_ptr(evals, "evals");
        if (tridiag_ql(n, a, b, evals, Z) != 0) throw EscError(ESC_ERR_NUMERIC, "tridiagonal QL did not converge");
    });
}

int esc_tridiag_eigenpairs(int n, const double *a, const double *b, double *evals, double *Z, int on_device) {
    return guarded([&] {
        check_ptr(a, "a");
        check_ptr(b, "b");
        check_ptr(evals, "evals");
        check_ptr(Z, "Z");
        if (n < 1 || n > esc::kMaxTridiag) throw EscError(ESC_ERR_ARG, "n must be in [1, 4096]");
        tridiag_eigenpairs(n, a, b, evals, Z, on_device != 0, nullptr);
    });
}

int esc_tridiag_lowest(int n, const double *a, const double *b, double *e0, double *z, int on_device) {
    return guarded([&] {
        if (n < 1 || n > 4096) throw EscError(ESC_ERR_ARG, "n must be in [1, 4096]");
        check_ptr(a, "a");
        check_ptr(b, "b");
        check_ptr(e0, "e0");
        // b[0] is not part of T; a zero b[k] (k >= 1) would end the recursion
        // there, so the routines see the first n rows only when none is zero
        for (int k = 1; k < n; ++k)
            if (b[k] == 0.0) throw EscError(ESC_ERR_ARG, "b[k] == 0 for some 1 <= k < n: split the matrix first");
        if (!on_device) {
            bool ok = false;
            *e0 = esc::td_lowest(n, a, b, &ok);
            if (!ok) throw EscError(ESC_ERR_NUMERIC, "non-finite tridiagonal matrix");
            if (z) {
                std::vector<double> work(5 * (size_t)n);
                esc::td_vector(n, a, b, *e0, work.data(), z);
            }
            return;
        }
        // one instance of the batched kernel: alpha as (re, im) pairs, beta
        // with beta[n] = 0 so that the recursion's length is n
        std::vector<double> a2(2 * (size_t)n, 0.0), bt((size_t)n + 1, 0.0);
        for (int k = 0; k < n; ++k) a2[2 * (size_t)k] = a[k];
        for (int k = 1; k < n; ++k) bt[k] = b[k];
        bt[0] = 1.0;
        double *d = nullptr;
        const size_t nd = 2 * (size_t)n + ((size_t)n + 1) + 6 * (size_t)n + (size_t)n + 1;
        if (hipMalloc(&d, nd * sizeof(double) + sizeof(int)) != hipSuccess) throw EscError(ESC_ERR_ALLOC, "allocation failed");
        struct Free { double *p; ~Free() { (void)hipFree(p); } } fr{d};
        double *da = d, *db = da + 2 * n, *dw = db + n + 1, *dz = dw + 6 * (size_t)n, *de = dz + n;
        int *di = (int *)(de + 1);
        hip_check(hipMemcpy(da, a2.data(), a2.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(db, bt.data(), bt.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(esc::launch_batch_tridiag(1, da, n, db, n + 1, n, dw, dz, n, de, di, nullptr), "tridiagonal kernel");
        hip_check(hipMemcpy(e0, de, sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        if (!std::isfinite(*e0)) throw EscError(ESC_ERR_NUMERIC, "non-finite tridiagonal matrix");
        if (z) hip_check(hipMemcpy(z, dz, (size_t)n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

int esc_cf_eval(int n, const double *a, const double *b, double norm2, double e0, double eta, const double *omega,
                int nw, double *out, int on_device) {
    return guarded([&] {
        if (n < 0 || n > 4096) throw EscError(ESC_ERR_ARG, "n must be in [0, 4096]");
        if (nw < 1) throw EscError(ESC_ERR_ARG, "nw must be >= 1");
        check_ptr(a, "a");
        check_ptr(b, "b");
        check_ptr(omega, "omega");
        check_ptr(out, "out");
        if (!on_device) {
            const std::vector<cplx> f = cf_eval(norm2 > 0.0 ? n : 0, a, b, norm2, e0, eta, omega, nw);
            std::memcpy(out, f.data(), (size_t)nw * sizeof(cplx));
            return;
        }
        // one instance of the batched kernel; its length is n through a zero
        // beta[n] and nonzero stand-ins for zeros inside the first n
        const int st = std::max(n, 1);
        std::vector<double> a2(2 * (size_t)st, 0.0), bt((size_t)st + 1, 0.0);
        for (int k = 0; k < n; ++k) a2[2 * (size_t)k] = a[k];
        for (int k = 1; k < n; ++k) bt[k] = b[k];
        for (int k = 1; k < n; ++k)
            if (bt[k] == 0.0) throw EscError(ESC_ERR_ARG, "b[k] == 0 for some 1 <= k < n: the fraction ends there");
        bt[0] = 1.0;
        const double n2 = n > 0 ? norm2 : 0.0;
        double *d = nullptr;
        const size_t nd = 2 * (size_t)st + ((size_t)st + 1) + 2 + (size_t)nw + 2 * (size_t)nw;
        if (hipMalloc(&d, nd * sizeof(double)) != hipSuccess) throw EscError(ESC_ERR_ALLOC, "allocation failed");
        struct Free { double *p; ~Free() { (void)hipFree(p); } } fr{d};
        double *da = d, *db = da + 2 * st, *dn = db + st + 1, *de = dn + 1, *dw = de + 1, *df = dw + nw;
        hip_check(hipMemcpy(da, a2.data(), a2.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(db, bt.data(), bt.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(dn, &n2, sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(de, &e0, sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(dw, omega, (size_t)nw * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(esc::launch_batch_cf(1, da, st, db, st + 1, st, dn, de, eta, dw, nw, df, nullptr), "fraction kernel");
        hip_check(hipMemcpy(out, df, 2 * (size_t)nw * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

int esc_cf_spectrum(int n, const double *a, const double *b, double norm2, double e0, double eta,
                    const double *omega, int nw, const double *elastic, double *out) {
    return guarded([&] {
        if (n < 1) throw EscError(ESC_ERR_ARG, "n must be >= 1");
        check_ptr(a, "a");
        check_ptr(b, "b");
        check_ptr(omega, "omega");
        check_ptr(out, "out");
        for (int q = 0; q < nw; ++q) {
            const std::complex<double> z(omega[q] + e0, -eta);
            std::complex<double> g(0.0, 0.0);
            for (int k = n - 1; k >= 0; --k) {
                const std::complex<double> tail = (k + 1 < n) ? b[k + 1] * b[k + 1] * g : std::complex<double>(0.0, 0.0);
                g = 1.0 / (z - a[k] - tail);
            }
            std::complex<double> r = norm2 * g;
            if (elastic) r -= std::complex<double>(elastic[0], elastic[1]) / std::complex<double>(omega[q], -eta);
            out[2 * q] = r.real();
            out[2 * q + 1] = r.imag();
        }
    });
}

int esc_lehmann_spectrum(int n, const double *evals, const double *Oa, const double *Ob, const double *elastic,
                         double eta, const double *omega, int nw, double *out) {
    return guarded([&] {
        if (n < 1) throw EscError(ESC_ERR_ARG, "n must be >= 1");
        check_ptr(evals, "evals");
        check_ptr(Oa, "Oa");
        check_ptr(Ob, "Ob");
        check_ptr(omega, "omega");
        check_ptr(out, "out");
        const double Eg = evals[0];
        for (int q = 0; q < nw; ++q) {
            std::complex<double> acc(0.0, 0.0);
            for (int k = 0; k < n; ++k) {
                const std::complex<double> den(omega[q] - (evals[k] - Eg), -eta);
                const std::complex<double> num = std::conj(std::complex<double>(Oa[2 * k], Oa[2 * k + 1])) *
                                                 std::complex<double>(Ob[2 * k], Ob[2 * k + 1]);
                acc = acc + num / den;
            }
            if (elastic) acc = acc - std::complex<double>(elastic[0], elastic[1]) / std::complex<double>(omega[q], -eta);
            out[2 * q] = acc.real();
            out[2 * q + 1] = acc.imag();
        }
    });
}

}  // extern "C"
