// esc_capi.cpp -- implementation of the C ABI declared in include/esc.h.
//
// Host orchestration only: contexts and workspaces, the kernel choice and
// autotuner, the Lanczos loop and the FP64 small solves (the model's tables
// and pass plans are in esc_plan.cpp, the spectra in esc_spectra.cpp).  All vector work runs in the HIP kernels of
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
// one, without row pairs.  ESC_NO_JIT=1 leaves them out.
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
                            // contiguous lanes only without row pairs (with them: 86 against 63 ms at
                            // 32 sites)
                            if ((w == N && P == kCubeLast) || (paired && !c->c64()) || (contiguous && (N < 25 || paired)))
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
// From 19 sites the generated sub-cube kernels come first (N = 24: 0.27 ms
// against 0.52 ms for the best generic shape in complex128, 0.15 against
// 0.32 ms in complex64; N = 20: 18.8 against 40.1 us and 13.9 against
// 28.8 us): complex128 2^3 rows per lane in 8-wave workgroups on 2^12-row
// tiles (N = 24: 289 us with free lane sites, 307 us with contiguous ones);
// complex64 over row pairs, with 2^3 rows per lane in 8-wave workgroups on
// 2^12-row tiles up to 22 sites, 2^4 rows per lane in 4-wave workgroups on
// 2^12-row tiles with the windows in reverse order up to 26 sites (N = 24:
// 154 against 162 us for 8 waves on 2^13-row tiles), and 2^3 rows per lane
// in 8-wave workgroups on 2^13-row tiles, windows in reverse order, beyond
// (N = 32: 61.3 ms; with one search per plan 66.6 ms, 72 ms without row
// pairs with contiguous lanes, 84 ms with 2^4 rows per lane).  A model they do not cover, or a missing run-time
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
            double *norm2, hipStream_t s, bool reortho) {
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
                    double resid_floor) {
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
