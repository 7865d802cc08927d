// esc_ctx.h -- the device context behind the C ABI (esc_ctx) and the host
// helpers its source files share: esc_capi.cpp (contexts, kernel choice,
// Lanczos, small solves, the entry points), esc_plan.cpp (term tables, pass
// plans, kernel configurations, operator tables), esc_spectra.cpp (spectra
// from a ground state), esc_batch.cpp (lockstep batches) and esc_grad.cpp
// (spectrum gradients).  Internal to the library.
#pragma once
#include "esc.h"
#include "esc_cube.h"
#include "esc_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace esc_impl {

using cplx = std::complex<double>;

// last error text of the calling thread (esc_last_error)
extern thread_local std::string g_err;

struct EscError : std::runtime_error {
    int code;
    EscError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw EscError(ESC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int guarded(F &&f) {
    try {
        f();
        g_err.clear();
        return ESC_OK;
    } catch (const EscError &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = "host allocation failed";
        return ESC_ERR_ALLOC;
    } catch (const std::exception &e) {
        g_err = e.what();
        return ESC_ERR_NUMERIC;
    }
}

// one stderr line per finished stage when ESC_VERBOSE is set
void progress(const char *fmt, ...);

// dense diagonalisation (Solver=ED) up to 2^13 states: 1 GiB of matrix
constexpr int kMaxEdSites = 13;

}  // namespace esc_impl

// ------------------------------------------------------------------------
// Context
// ------------------------------------------------------------------------
struct esc_ctx {
    int n_sites = 0;
    int precision = ESC_C128;
    uint64_t dim = 0;
    size_t esz = 16;
    int device = 0;
    esc_model model{};
    std::vector<int32_t> bonds[3];
    std::vector<int32_t> neigh;
    esc::HTables htab{};
    esc::HTables *d_htab = nullptr;
    std::map<int64_t, esc::CurrentTable *> d_cur;   // cached per (op, site)
    double *d_partials = nullptr;                   // reduction workspace
    size_t partials_cap = 0;
    double *d_result = nullptr;                     // small device results
    size_t result_cap = 0;
    void *d_vecs = nullptr;                         // Lanczos basis / work vectors
    size_t vecs_cap = 0;                            // in vectors
    int basis_n = 0;                                // vectors valid in the kept basis
    std::vector<double> Z;                          // last tridiagonal eigenvectors
    std::vector<double> evals;                      // last Ritz values (ascending)
    int z_n = 0;
    bool basis_kept = false;
    void *d_aux = nullptr;                          // dynamics work vectors
    size_t aux_cap = 0;
    esc::CurrentTable *d_combo = nullptr;           // combined-operator table
    struct PassDev {
        esc::GTables *d;
        int n_uniform, n_terms;
    };
    std::map<int, std::vector<PassDev>> d_pass;     // windowed tiled kernel per configuration
    // operator tables applied by the windowed tiled kernel, per table content
    // (hash of the host table): configuration and passes, or no entry = k_current
    struct CurPass {
        int config = 0;
        std::vector<PassDev> pass;
    };
    std::map<uint64_t, CurPass> d_curpass;
    mutable std::map<int, bool> tile_fits;           // configuration fits the LDS
    int tile_bits = 0;                              // kernel configuration: 0 = one row per lane,
                                                    // else 100 * tile bits + window bits
    double *d_cube_cf = nullptr;                    // coupling scalars of the generated kernels
    std::map<int, std::shared_ptr<esc::CubeProgram>> cube;   // generated kernels per configuration
    mutable std::map<int, std::string> cube_fail;   // configurations whose generation failed, and why

    bool c64() const { return precision == ESC_C64; }

    ~esc_ctx() {
        if (d_htab) (void)hipFree(d_htab);
        for (auto &kv : d_cur) (void)hipFree(kv.second);
        if (d_partials) (void)hipFree(d_partials);
        if (d_result) (void)hipFree(d_result);
        if (d_vecs) (void)hipFree(d_vecs);
        if (d_aux) (void)hipFree(d_aux);
        if (d_combo) (void)hipFree(d_combo);
        if (d_cube_cf) (void)hipFree(d_cube_cf);
        for (auto &kv : d_pass)
            for (auto &p : kv.second) (void)hipFree(p.d);
        for (auto &kv : d_curpass)
            for (auto &p : kv.second.pass) (void)hipFree(p.d);
    }
    void ensure_aux(size_t nvec) {
        if (nvec <= aux_cap) return;
        if (d_aux) esc_impl::hip_check(hipFree(d_aux), "hipFree");
        d_aux = nullptr;
        aux_cap = 0;
        if (hipMalloc(&d_aux, nvec * dim * esz) != hipSuccess)
            throw esc_impl::EscError(ESC_ERR_ALLOC, "cannot allocate " + std::to_string(nvec) + " work vectors on the device");
        aux_cap = nvec;
    }
    void *aux(size_t m) { return (char *)d_aux + m * dim * esz; }

    void ensure_partials(size_t n) {
        if (n <= partials_cap) return;
        if (d_partials) esc_impl::hip_check(hipFree(d_partials), "hipFree");
        d_partials = nullptr;
        if (hipMalloc(&d_partials, n * sizeof(double)) != hipSuccess) throw esc_impl::EscError(ESC_ERR_ALLOC, "workspace allocation failed");
        partials_cap = n;
    }
    void ensure_result(size_t n) {
        if (n <= result_cap) return;
        if (d_result) esc_impl::hip_check(hipFree(d_result), "hipFree");
        d_result = nullptr;
        if (hipMalloc(&d_result, n * sizeof(double)) != hipSuccess) throw esc_impl::EscError(ESC_ERR_ALLOC, "workspace allocation failed");
        result_cap = n;
    }
    void ensure_vecs(size_t nvec) {
        if (nvec <= vecs_cap) return;
        if (d_vecs) esc_impl::hip_check(hipFree(d_vecs), "hipFree");
        d_vecs = nullptr;
        vecs_cap = 0;
        basis_kept = false;
        if (hipMalloc(&d_vecs, nvec * dim * esz) != hipSuccess)
            throw esc_impl::EscError(ESC_ERR_ALLOC, "cannot allocate " + std::to_string(nvec) + " vectors of " +
                                              std::to_string(dim * esz) + " bytes on the device");
        vecs_cap = nvec;
    }
    void *vec(size_t m) { return (char *)d_vecs + m * dim * esz; }
};

namespace esc_impl {

// argument and workspace helpers
void check_ctx(const esc_ctx *c);
void check_ptr(const void *p, const char *name);
void sync_get(hipStream_t s, void *dst, const void *src, size_t n);
void device_dot(esc_ctx *c, const void *x, const void *y, double out[2], hipStream_t s);
double device_norm2(esc_ctx *c, const void *x, hipStream_t s);

// H.psi kernel configurations
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
inline int cfg_passes(int config) { return (config / 100000) % 10; }
inline int cfg_regs(int config) { return (config / 1000000) % 10; }
// the layout digit of the generated kernels (0..3): bit 0 = over row pairs
// (complex64: 16-byte lane accesses), bit 1 = lanes on six contiguous sites
inline bool cfg_paired(int config) { return ((config / 10000000) % 10) & 1; }
inline bool cfg_contiguous(int config) { return ((config / 10000000) % 10) & 2; }
inline bool cfg_cube(int config) { return cfg_passes(config) == kCubeFirst || cfg_passes(config) == kCubeLast; }
inline int cfg_waves(int config) { return 1 << ((config / 10000) % 10); }
inline int cfg_tile_bits(int config) { return (config / 100) % 100; }
inline int cfg_window(int config) { return config % 100; }
inline int make_config(int P, int waves, int tb, int w, int k = 0, bool paired = false, bool contiguous = false) {
    int v = 0;
    while ((1 << v) < waves) ++v;
    return 10000000 * ((paired ? 1 : 0) | (contiguous ? 2 : 0)) + 1000000 * k + 100000 * P + 10000 * v + 100 * tb + w;
}

// the model on the device (esc_plan.cpp): term tables into c->htab and the
// device copies
void build_tables(esc_ctx *c);
// an off-diagonal term of H in the 8-entry-table form: partner r ^ mask,
// coefficient coef[index of bits b0, b1, b2 of r] (re, im)
struct HostTerm {
    uint32_t mask;
    int b0, b1, b2;
    double coef[8][2];
    bool cplx;
};
std::vector<HostTerm> host_terms(const esc_ctx *c);
// One planned pass: window sites W, tile sites T (lane sites 0..5 included)
// and the terms (indices into host_terms) it applies from the tile (near) and
// from the rest of the window (far).
struct Plan {
    uint32_t W, T;
    std::vector<size_t> near, far;
};
std::vector<Plan> plan_windows(const esc_ctx *c, const std::vector<HostTerm> &terms, int TB, int w, int P);
// the tiled kernel's passes of a configuration (custom: other terms than H's)
std::vector<esc::GTables> build_passes(const esc_ctx *c, int TB, int w, int P, int WPB,
                                       const std::vector<HostTerm> *custom = nullptr);
// the generated kernels' model and passes of a configuration (do_plan false:
// the model only); false with *why when the configuration does not apply
bool cube_spec(const esc_ctx *c, int cfg, esc::CubeModel &m, std::vector<esc::CubePass> &passes, std::string *why,
               bool do_plan = true);
// coupling scalars of the generated kernels (esc::CubeScalar order)
void cube_scalars(const esc::HTables &h, double *cf);

// H.psi with a given kernel configuration / the context's choice
// the generated program of a configuration (compiled or fetched from the
// caches), its scalars uploaded to c->d_cube_cf
esc::CubeProgram &cube_for(esc_ctx *c, int config);
std::vector<int> cube_candidates(const esc_ctx *c);
void launch_hpsi_with(esc_ctx *c, int config, const void *in, void *out, hipStream_t s, double *dot_part = nullptr);
void run_hpsi(esc_ctx *c, const void *in, void *out, hipStream_t s);

// one classical Gram-Schmidt pass of w against the first k kept vectors
void cgs_pass(esc_ctx *c, int k, void *w, hipStream_t s);

// operator tables (Pauli strings)
void host_current_table(const esc_ctx *c, int op, int site, esc::CurrentTable *t);
esc::CurrentTable *current_table(esc_ctx *c, int op, int site);
// y = O x for the operator of host table ht (device copy dt, may be null):
// from 17 sites the windowed tiled kernel with the table's Pauli strings as
// gather terms, else (or when a string does not fit that form) k_current
void apply_current_table(esc_ctx *c, const esc::CurrentTable &ht, const esc::CurrentTable *dt, const void *x, void *y,
                         hipStream_t s);
void append_scaled(esc::CurrentTable &dst, const esc::CurrentTable &src, double cr, double ci);

// FP64 small solves
// fn(begin, end) over contiguous chunks of [0, n) on up to 16 host threads
// (the caller takes the first chunk), at least `grain` items per chunk
template <typename F>
void host_parallel_for(int n, int grain, const F &fn) {
    const unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min(hw ? hw : 1u, 16u);
    nt = std::max(1, std::min(nt, n / std::max(1, grain)));
    if (nt <= 1) {
        fn(0, n);
        return;
    }
    const int chunk = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    try {
        for (int t = 1; t < nt; ++t) {
            const int b = t * chunk, e = std::min(n, b + chunk);
            if (b < e) th.emplace_back([&fn, b, e] { fn(b, e); });
        }
    } catch (...) {
        for (auto &x : th) x.join();
        throw;
    }
    fn(0, std::min(n, chunk));
    for (auto &x : th) x.join();
}

int tridiag_ql(int n, const double *a, const double *b, double *evals, double *Z);
// all eigenpairs of a tridiagonal matrix, on the host or the device
void tridiag_eigenpairs(int n, const double *a, const double *b, double *ev, double *Z, bool on_device,
                        hipStream_t s);
// one continued-fraction recursion from d_phi (work: three device vectors);
// alpha[0..iters-1], beta[0..iters], |phi|^2
void cf_run(esc_ctx *ctx, const void *d_phi, int steps, void *work[3], double *alpha, double *beta, int *iters,
            double *norm2, hipStream_t s, bool reortho = false);
// the three-vector Lanczos recursion from start with esc_lanczos's exit
// rule (returns the number of alphas); with coef, the same recursion
// replayed for n_fixed vectors, accumulating acc += coef[k] v_k
int plain_recursion(esc_ctx *c, const void *start, int steps, double eps, const std::complex<double> *coef,
                    int n_fixed, void *acc, std::vector<double> &ar, std::vector<double> &bt, hipStream_t s,
                    double resid_floor = 0.0);
double tridiag_lowest(int n, const double *a, const double *b);

// spectra from continued fractions: one run per channel / polarisation pair
struct CfRun {
    std::vector<cplx> coef;
    int kind, a, b;
};
std::vector<std::pair<int, int>> spectrum_pairs(bool js, int nc);
std::vector<CfRun> cf_runs(bool js, int nc);
std::vector<cplx> cf_eval(int it, const double *al, const double *be, double n2, double e0, double eta,
                          const double *omega, int nw);
std::map<std::pair<int, int>, std::vector<cplx>> cf_assemble(bool js, int nc, const std::vector<CfRun> &runs,
                                                             const std::vector<std::vector<cplx>> &f, int nw);
void write_spectrum_table(bool js, int nc, const std::vector<cplx> &ex,
                          std::map<std::pair<int, int>, std::vector<cplx>> &F, const double *omega, int nw,
                          double eta, double *table);
// LocalSzSz tables (ESC_DYN_LOCALSZSZ), the Ritz amplitudes <psi_k|vec> of
// the kept basis, and the RIXS state sums (esc_spectra.cpp)
void local_szsz_spectra(esc_ctx *ctx, int method, bool reortho, const void *d_gs, double e0, int cf_steps,
                        double eta, const double *omega, int nw, double *table, double *expect, hipStream_t s);
std::vector<cplx> ritz_amplitudes(esc_ctx *ctx, const void *d_vec, hipStream_t s);
void rixs_spectra(esc_ctx *g, esc_ctx *q, int method, int site, const void *d_gs, double e0, int cf_steps,
                  double gamma, double eta, const double *omega, int nw, double *table, hipStream_t s);

}  // namespace esc_impl
