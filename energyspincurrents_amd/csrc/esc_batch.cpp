// esc_batch.cpp -- many independent systems solved in lockstep (esc_batch_*):
// one Lanczos step of all instances per replayed HIP graph.
#include "esc_ctx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <string>
#include <vector>

using namespace esc_impl;

// ------------------------------------------------------------------------
// Batches of independent systems (different sizes and couplings) solved in
// lockstep: every Lanczos step of all instances is one captured HIP graph
// (the instances' H.psi launches plus four batched kernels for the scalars),
// and the scalars never leave the device until the run ends.
// ------------------------------------------------------------------------
struct esc_batch {
    int n = 0;
    bool c64 = false;
    size_t esz = 16;
    std::vector<esc_ctx *> ctx;
    std::vector<void *> mem;              // per instance: slots 0..2, start, gs, phi
    std::vector<double *> dotbuf;         // per instance: tiled H.psi dot partials, or null
    std::vector<esc::BatchLane> lanes[3]; // per rotation phase
    esc::BatchLane *d_lanes[3] = {nullptr, nullptr, nullptr};
    uint32_t total_blocks = 0;
    double *d_partials = nullptr, *d_alpha = nullptr, *d_beta = nullptr, *d_coef = nullptr;
    uint32_t *d_step = nullptr;
    int cap = 0;                          // step capacity of alpha / beta / coef (graphs bake it)
    hipStream_t s = nullptr;
    hipGraph_t graph[2][3] = {};
    hipGraphExec_t exec[2][3] = {};
    std::vector<double> e0;
    bool have_gs = false;

    void *slot(int i, int q) const { return (char *)mem[i] + (size_t)q * ctx[i]->dim * esz; }
    void drop_graphs() {
        for (int a = 0; a < 2; ++a)
            for (int p = 0; p < 3; ++p) {
                if (exec[a][p]) (void)hipGraphExecDestroy(exec[a][p]);
                if (graph[a][p]) (void)hipGraphDestroy(graph[a][p]);
                exec[a][p] = nullptr;
                graph[a][p] = nullptr;
            }
    }
    ~esc_batch() {
        if (s) (void)hipStreamSynchronize(s);
        drop_graphs();
        for (void *m : mem) if (m) (void)hipFree(m);
        for (double *d : dotbuf) if (d) (void)hipFree(d);
        for (auto *d : d_lanes) if (d) (void)hipFree(d);
        for (void *p : {(void *)d_partials, (void *)d_alpha, (void *)d_beta, (void *)d_coef, (void *)d_step})
            if (p) (void)hipFree(p);
        for (esc_ctx *c : ctx) delete c;
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

void batch_check(const esc_batch *b) {
    if (!b) throw EscError(ESC_ERR_ARG, "null batch");
}

void batch_capacity(esc_batch *b, int steps) {
    if (steps <= b->cap) return;
    b->drop_graphs();
    for (double **p : {&b->d_alpha, &b->d_beta, &b->d_coef}) {
        if (*p) hip_check(hipFree(*p), "hipFree");
        *p = nullptr;
    }
    const size_t n = (size_t)b->n;
    if (hipMalloc(&b->d_alpha, n * steps * 2 * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->d_beta, n * (steps + 1) * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->d_coef, n * steps * sizeof(double)) != hipSuccess)
        throw EscError(ESC_ERR_ALLOC, "batch scalar allocation failed");
    b->cap = steps;
}

// Instances with a tiled H.psi get their dot <v|w> from the kernel's per-tile
// partials (one buffer each, sized for the current kernel choice), so the
// batched dot pass skips them.  Called when the kernel choices are set
// (creation, autotuning), with no graph holding the old lanes.
void batch_setup_dots(esc_batch *b) {
    b->dotbuf.resize(b->n, nullptr);
    for (int i = 0; i < b->n; ++i) {
        esc_ctx *c = b->ctx[i];
        if (b->dotbuf[i]) hip_check(hipFree(b->dotbuf[i]), "hipFree");
        b->dotbuf[i] = nullptr;
        uint32_t ndot = 0;
        if (c->tile_bits > 0) {
            ndot = (uint32_t)(c->dim >> cfg_tile_bits(c->tile_bits));
            if (hipMalloc(&b->dotbuf[i], 2 * (size_t)ndot * sizeof(double)) != hipSuccess)
                throw EscError(ESC_ERR_ALLOC, "batch dot buffer allocation failed");
        }
        for (int p = 0; p < 3; ++p) {
            b->lanes[p][i].dotp = b->dotbuf[i];
            b->lanes[p][i].ndot = ndot;
        }
    }
    for (int p = 0; p < 3; ++p)
        hip_check(hipMemcpy(b->d_lanes[p], b->lanes[p].data(), sizeof(esc::BatchLane) * b->n, hipMemcpyHostToDevice),
                  "hipMemcpy");
}

// the launches of step k (phase k % 3) for all instances
void batch_step_launches(esc_batch *b, int phase, bool accum) {
    const int lda = b->cap, ldb = b->cap + 1, ldc = b->cap;
    if (accum)
        hip_check(esc::launch_batch_step(b->c64, esc::kBatchAccum, b->d_lanes[phase], b->n, b->total_blocks,
                                         b->d_partials, b->d_alpha, lda, b->d_beta, ldb, b->d_coef, ldc, b->d_step,
                                         b->s), "batch accumulate");
    for (int i = 0; i < b->n; ++i) {
        const esc::BatchLane &l = b->lanes[phase][i];
        if (l.ndot > 0) launch_hpsi_with(b->ctx[i], b->ctx[i]->tile_bits, l.v, l.w, b->s, b->dotbuf[i]);
        else run_hpsi(b->ctx[i], l.v, l.w, b->s);
    }
    for (int piece : {esc::kBatchDot, esc::kBatchUpdate})
        hip_check(esc::launch_batch_step(b->c64, piece, b->d_lanes[phase], b->n, b->total_blocks, b->d_partials,
                                         b->d_alpha, lda, b->d_beta, ldb, b->d_coef, ldc, b->d_step, b->s),
                  "batch step");
}

// `steps` lockstep Lanczos steps of every instance from the unit vectors in
// slot 1 (slot 0 is cleared); with accum, acc += coef[i][k] v_k into gs.
// Returns alpha (real parts, n x steps) and beta (n x (steps + 1)).
void batch_run(esc_batch *b, int steps, bool use_graph, const std::vector<double> *coef, std::vector<double> &alpha,
               std::vector<double> &beta) {
    batch_capacity(b, steps);
    const bool accum = coef != nullptr;
    for (int i = 0; i < b->n; ++i) {
        hip_check(hipMemsetAsync(b->slot(i, 0), 0, b->ctx[i]->dim * b->esz, b->s), "clear");
        if (accum) hip_check(hipMemsetAsync(b->slot(i, 4), 0, b->ctx[i]->dim * b->esz, b->s), "clear");
    }
    hip_check(hipMemsetAsync(b->d_beta, 0, (size_t)b->n * (b->cap + 1) * sizeof(double), b->s), "clear");
    hip_check(hipMemsetAsync(b->d_step, 0, sizeof(uint32_t), b->s), "clear");
    if (accum) {
        std::vector<double> c((size_t)b->n * b->cap, 0.0);
        for (int i = 0; i < b->n; ++i)
            for (int k = 0; k < steps; ++k) c[(size_t)i * b->cap + k] = (*coef)[(size_t)i * steps + k];
        hip_check(hipMemcpyAsync(b->d_coef, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice, b->s), "coef");
    }
    for (int k = 0; k < steps; ++k) {
        const int p = k % 3;
        if (!use_graph) {
            batch_step_launches(b, p, accum);
            continue;
        }
        if (!b->exec[accum][p]) {
            hip_check(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal), "begin capture");
            try {
                batch_step_launches(b, p, accum);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(b->s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hip_check(hipStreamEndCapture(b->s, &b->graph[accum][p]), "end capture");
            hip_check(hipGraphInstantiate(&b->exec[accum][p], b->graph[accum][p], nullptr, nullptr, 0), "instantiate");
        }
        hip_check(hipGraphLaunch(b->exec[accum][p], b->s), "graph launch");
    }
    std::vector<double> a2((size_t)b->n * b->cap * 2), bt((size_t)b->n * (b->cap + 1));
    hip_check(hipMemcpyAsync(a2.data(), b->d_alpha, a2.size() * sizeof(double), hipMemcpyDeviceToHost, b->s), "alpha");
    hip_check(hipMemcpyAsync(bt.data(), b->d_beta, bt.size() * sizeof(double), hipMemcpyDeviceToHost, b->s), "beta");
    hip_check(hipStreamSynchronize(b->s), "sync");
    alpha.assign((size_t)b->n * steps, 0.0);
    beta.assign((size_t)b->n * (steps + 1), 0.0);
    for (int i = 0; i < b->n; ++i) {
        for (int k = 0; k < steps; ++k) alpha[(size_t)i * steps + k] = a2[2 * ((size_t)i * b->cap + k)];
        for (int k = 0; k <= steps; ++k) beta[(size_t)i * (steps + 1) + k] = bt[(size_t)i * (b->cap + 1) + k];
    }
}

// length of instance i's recursion: up to and including the first step whose
// beta_{k+1} came out zero (exhausted Krylov space)
int batch_length(const std::vector<double> &beta, int i, int steps) {
    for (int k = 0; k < steps; ++k)
        if (beta[(size_t)i * (steps + 1) + k + 1] == 0.0) return k + 1;
    return steps;
}

// copy x / |x| of instance i into its slot 1 (norm2 returned)
double batch_load_unit(esc_batch *b, int i, const void *x) {
    esc_ctx *c = b->ctx[i];
    const double n2 = device_norm2(c, x, b->s);
    void *dst = b->slot(i, 1);
    hip_check(hipMemcpyAsync(dst, x, c->dim * b->esz, hipMemcpyDeviceToDevice, b->s), "copy");
    if (n2 > 0.0 && std::isfinite(n2)) hip_check(esc::launch_scale(b->c64, dst, 1.0 / std::sqrt(n2), c->dim, b->s), "scale");
    return n2;
}

}  // namespace

extern "C" {

int esc_batch_create(int n, const esc_model *models, int precision, esc_batch **out) {
    return guarded([&] {
        check_ptr(models, "models");
        check_ptr(out, "out");
        *out = nullptr;
        if (n < 1 || n > 1024) throw EscError(ESC_ERR_ARG, "n must be in [1, 1024]");
        auto b = std::make_unique<esc_batch>();
        b->n = n;
        b->c64 = precision == ESC_C64;
        b->esz = b->c64 ? 8 : 16;
        hip_check(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking), "hipStreamCreate");
        for (int i = 0; i < n; ++i) {
            esc_ctx *c = nullptr;
            const int rc = esc_ctx_create(&models[i], precision, &c);
            if (rc != ESC_OK) throw EscError(rc, "instance " + std::to_string(i) + ": " + esc_last_error());
            b->ctx.push_back(c);
            void *m = nullptr;
            if (hipMalloc(&m, 6 * c->dim * b->esz) != hipSuccess)
                throw EscError(ESC_ERR_ALLOC, "instance " + std::to_string(i) + ": cannot allocate its 6 vectors");
            b->mem.push_back(m);
        }
        uint32_t blk = 0;
        for (int p = 0; p < 3; ++p) b->lanes[p].resize(n);
        for (int i = 0; i < n; ++i) {
            const uint64_t dim = b->ctx[i]->dim;
            uint64_t nb = (dim + 16 * (uint64_t)256 - 1) / (16 * (uint64_t)256);
            nb = std::max<uint64_t>(1, std::min<uint64_t>(nb, 1024));
            for (int p = 0; p < 3; ++p) {
                esc::BatchLane &l = b->lanes[p][i];
                l.vp = b->slot(i, p);
                l.v = b->slot(i, (p + 1) % 3);
                l.w = b->slot(i, (p + 2) % 3);
                l.acc = b->slot(i, 4);
                l.dim = dim;
                l.blk0 = blk;
                l.nblk = (uint32_t)nb;
                l.dotp = nullptr;
                l.ndot = 0;
            }
            blk += (uint32_t)nb;
        }
        b->total_blocks = blk;
        for (int p = 0; p < 3; ++p) {
            if (hipMalloc(&b->d_lanes[p], sizeof(esc::BatchLane) * n) != hipSuccess)
                throw EscError(ESC_ERR_ALLOC, "batch table allocation failed");
            hip_check(hipMemcpy(b->d_lanes[p], b->lanes[p].data(), sizeof(esc::BatchLane) * n, hipMemcpyHostToDevice),
                      "hipMemcpy");
        }
        if (hipMalloc(&b->d_partials, (size_t)blk * 2 * sizeof(double)) != hipSuccess ||
            hipMalloc(&b->d_step, sizeof(uint32_t)) != hipSuccess)
            throw EscError(ESC_ERR_ALLOC, "batch workspace allocation failed");
        batch_setup_dots(b.get());
        b->e0.assign(n, 0.0);
        *out = b.release();
    });
}

int esc_batch_destroy(esc_batch *b) {
    return guarded([&] { delete b; });
}

int esc_batch_ground_state(esc_batch *b, uint64_t seed, int steps, int use_graph, double *e0, int *iters,
                           double *alpha, double *beta) {
    return guarded([&] {
        batch_check(b);
        if (steps < 2 || steps > 4096) throw EscError(ESC_ERR_ARG, "steps must be in [2, 4096]");
        b->have_gs = false;
        // start vectors (Gaussian, seed + i), kept in slot 3 for the replay;
        // one eager H.psi per instance builds its kernel tables before capture
        for (int i = 0; i < b->n; ++i) {
            esc_ctx *c = b->ctx[i];
            hip_check(esc::launch_gaussian(b->c64, seed + (uint64_t)i, b->slot(i, 3), c->dim, b->s), "gaussian");
            batch_load_unit(b, i, b->slot(i, 3));
            hip_check(hipMemcpyAsync(b->slot(i, 3), b->slot(i, 1), c->dim * b->esz, hipMemcpyDeviceToDevice, b->s), "copy");
            run_hpsi(c, b->slot(i, 1), b->slot(i, 2), b->s);
        }
        std::vector<double> al, bt;
        batch_run(b, steps, use_graph != 0, nullptr, al, bt);
        std::vector<double> coef((size_t)b->n * steps, 0.0);
        for (int i = 0; i < b->n; ++i) {
            const int n = batch_length(bt, i, steps);
            std::vector<double> ev(n), Z((size_t)n * n);
            if (tridiag_ql(n, al.data() + (size_t)i * steps, bt.data() + (size_t)i * (steps + 1), ev.data(), Z.data()) != 0)
                throw EscError(ESC_ERR_NUMERIC, "tridiagonal QL did not converge");
            b->e0[i] = ev[0];
            for (int k = 0; k < n; ++k) coef[(size_t)i * steps + k] = Z[k];
            if (e0) e0[i] = ev[0];
            if (iters) iters[i] = n;
        }
        if (alpha) std::memcpy(alpha, al.data(), al.size() * sizeof(double));
        if (beta) std::memcpy(beta, bt.data(), bt.size() * sizeof(double));
        // replay from the same starts, forming |gs> = sum_k Z(k,0) v_k in slot 4
        for (int i = 0; i < b->n; ++i)
            hip_check(hipMemcpyAsync(b->slot(i, 1), b->slot(i, 3), b->ctx[i]->dim * b->esz, hipMemcpyDeviceToDevice, b->s),
                      "copy");
        std::vector<double> al2, bt2;
        batch_run(b, steps, use_graph != 0, &coef, al2, bt2);
        for (int i = 0; i < b->n; ++i) {
            const double n2 = device_norm2(b->ctx[i], b->slot(i, 4), b->s);
            hip_check(esc::launch_scale(b->c64, b->slot(i, 4), 1.0 / std::sqrt(n2), b->ctx[i]->dim, b->s), "scale");
        }
        hip_check(hipStreamSynchronize(b->s), "sync");
        b->have_gs = true;
    });
}

int esc_batch_autotune(esc_batch *b, int reps, int *configs) {
    return guarded([&] {
        batch_check(b);
        b->drop_graphs();   // the captured steps bake each instance's kernel choice
        b->have_gs = false;
        for (int i = 0; i < b->n; ++i) {
            esc_ctx *c = b->ctx[i];
            hip_check(esc::launch_gaussian(b->c64, 0x5eedULL + (uint64_t)i, b->slot(i, 1), c->dim, b->s), "gaussian");
            int cfg = 0;
            double ms = 0.0;
            const int rc = esc_autotune(c, b->slot(i, 1), b->slot(i, 2), reps, &cfg, &ms, b->s);
            if (rc != ESC_OK) throw EscError(rc, "instance " + std::to_string(i) + ": " + esc_last_error());
            if (configs) configs[i] = cfg;
        }
        hip_check(hipStreamSynchronize(b->s), "sync");
        batch_setup_dots(b);
    });
}

int esc_batch_state(esc_batch *b, int i, void *d_out, void *stream) {
    return guarded([&] {
        batch_check(b);
        check_ptr(d_out, "d_out");
        if (i < 0 || i >= b->n) throw EscError(ESC_ERR_ARG, "instance out of range");
        if (!b->have_gs) throw EscError(ESC_ERR_ARG, "no ground states yet (esc_batch_ground_state)");
        hip_check(hipMemcpyAsync(d_out, b->slot(i, 4), b->ctx[i]->dim * b->esz, hipMemcpyDeviceToDevice,
                                 (hipStream_t)stream), "copy");
    });
}

int esc_batch_dynamics(esc_batch *b, int kind, int cf_steps, double eta, const double *omega, int nw, int use_graph,
                       double *tables, double *expect) {
    return guarded([&] {
        batch_check(b);
        check_ptr(omega, "omega");
        check_ptr(tables, "tables");
        if (!b->have_gs) throw EscError(ESC_ERR_ARG, "no ground states yet (esc_batch_ground_state)");
        if (kind != ESC_DYN_JSJS && kind != ESC_DYN_JEJE) throw EscError(ESC_ERR_ARG, "unknown dynamics kind");
        if (cf_steps < 2 || cf_steps > 4096) throw EscError(ESC_ERR_ARG, "cf_steps must be in [2, 4096]");
        if (nw < 1) throw EscError(ESC_ERR_ARG, "nw must be >= 1");
        const bool js = kind == ESC_DYN_JSJS;
        const std::vector<int> ops = js ? std::vector<int>{ESC_OP_JSX, ESC_OP_JSY, ESC_OP_JSZ}
                                        : std::vector<int>{ESC_OP_JEX, ESC_OP_JEY};
        const int nc = (int)ops.size(), ncols = js ? 14 : 5;
        const int n = b->n;
        std::vector<std::vector<esc::CurrentTable>> tabs(n, std::vector<esc::CurrentTable>(nc));
        for (int i = 0; i < n; ++i) {
            for (int a = 0; a < nc; ++a) host_current_table(b->ctx[i], ops[a], ESC_SITE_TOTAL, &tabs[i][a]);
            if (!b->ctx[i]->d_combo && hipMalloc(&b->ctx[i]->d_combo, sizeof(esc::CurrentTable)) != hipSuccess)
                throw EscError(ESC_ERR_ALLOC, "table allocation failed");
        }
        // phi_i = (sum_a c_a J_a)|gs_i> into slot 5
        auto apply = [&](int i, const std::vector<cplx> &coef) {
            esc_ctx *c = b->ctx[i];
            esc::CurrentTable t;
            std::memset(&t, 0, sizeof(t));
            for (int a = 0; a < nc; ++a)
                if (coef[a] != cplx(0.0, 0.0)) append_scaled(t, tabs[i][a], coef[a].real(), coef[a].imag());
            hip_check(hipMemcpyAsync(c->d_combo, &t, sizeof(t), hipMemcpyHostToDevice, b->s), "table");
            hip_check(b->c64 ? esc::launch_current_c64(c->d_combo, b->slot(i, 4), b->slot(i, 5), c->dim, b->s)
                             : esc::launch_current_c128(c->d_combo, b->slot(i, 4), b->slot(i, 5), c->dim, b->s),
                      "current launch");
            hip_check(hipStreamSynchronize(b->s), "sync");   // the host table is reused
        };
        std::vector<std::vector<cplx>> ex(n, std::vector<cplx>(nc + 1, cplx(0.0, 0.0)));
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < nc; ++a) {
                std::vector<cplx> c(nc, cplx(0.0, 0.0));
                c[a] = 1.0;
                apply(i, c);
                double d[2];
                device_dot(b->ctx[i], b->slot(i, 4), b->slot(i, 5), d, b->s);
                ex[i][a] = cplx(d[0], d[1]);
                ex[i][nc] += ex[i][a];
            }
        const std::vector<CfRun> runs = cf_runs(js, nc);
        std::vector<std::vector<std::vector<cplx>>> f(n);
        for (const CfRun &r : runs) {
            std::vector<double> n2(n);
            for (int i = 0; i < n; ++i) {
                apply(i, r.coef);
                n2[i] = batch_load_unit(b, i, b->slot(i, 5));
            }
            std::vector<double> al, bt;
            batch_run(b, cf_steps, use_graph != 0, nullptr, al, bt);
            for (int i = 0; i < n; ++i) {
                const int it = n2[i] > 0.0 ? batch_length(bt, i, cf_steps) : 0;
                f[i].push_back(cf_eval(it, al.data() + (size_t)i * cf_steps, bt.data() + (size_t)i * (cf_steps + 1),
                                       n2[i], b->e0[i], eta, omega, nw));
            }
        }
        for (int i = 0; i < n; ++i) {
            auto F = cf_assemble(js, nc, runs, f[i], nw);
            write_spectrum_table(js, nc, ex[i], F, omega, nw, eta, tables + (size_t)i * nw * (1 + ncols));
            if (expect)
                for (int a = 0; a <= nc; ++a) {
                    expect[2 * ((size_t)i * 4 + a)] = ex[i][a].real();
                    expect[2 * ((size_t)i * 4 + a) + 1] = ex[i][a].imag();
                }
        }
    });
}

}  // extern "C"
