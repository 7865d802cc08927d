// esc_batch.cpp -- many independent systems solved in lockstep (esc_batch_*):
// one Lanczos step of all instances per replayed HIP graph.
#include "esc_ctx.h"
#include "esc_tridiag.h"

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
    // device workspaces of the solves, grown on demand: the tridiagonal
    // solves, energies, lengths, kept alpha / beta rows, norms, start-vector
    // norms, expectation values, fractions, omegas and operator tables
    struct DevBuf {
        void *p = nullptr;
        size_t bytes = 0;
        void *get(size_t n) {
            if (n > bytes) {
                if (p) hip_check(hipFree(p), "hipFree");
                p = nullptr;
                bytes = 0;
                if (hipMalloc(&p, n) != hipSuccess) throw EscError(ESC_ERR_ALLOC, "batch workspace allocation failed");
                bytes = n;
            }
            return p;
        }
    };
    DevBuf work, e0_dev, iters_dev, keep_alpha, keep_beta, norms, n2_dev, ex_dev, f_dev, omega_dev, tabs_dev;
    // Instances that run the same generated program (same lattice, term
    // structure and configuration) share one launch per pass: grid y = member,
    // each with its own coupling scalars, vectors and dot partials.
    struct Group {
        std::vector<int> inst;
        esc::CubeProgram *prog = nullptr;
        uint64_t dim = 0;
        DevBuf ptrs;   // per rotation phase: ins, outs, dots (3 x 3 x members pointers)
        DevBuf cfs;    // members x kCubeScalars doubles
        const void *const *ins(int phase) { return (const void *const *)ptrs.p + (size_t)phase * 3 * inst.size(); }
        void *const *outs(int phase) { return (void *const *)ptrs.p + ((size_t)phase * 3 + 1) * inst.size(); }
        double *const *dots(int phase) { return (double *const *)ptrs.p + ((size_t)phase * 3 + 2) * inst.size(); }
    };
    std::vector<std::unique_ptr<Group>> groups;
    std::vector<int> group_of;            // instance -> group, or -1

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
        for (DevBuf *d : {&work, &e0_dev, &iters_dev, &keep_alpha, &keep_beta, &norms, &n2_dev, &ex_dev, &f_dev,
                          &omega_dev, &tabs_dev})
            if (d->p) (void)hipFree(d->p);
        for (auto &g : groups)
            for (DevBuf *d : {&g->ptrs, &g->cfs})
                if (d->p) (void)hipFree(d->p);
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

// The generated program of instance i's kernel choice, or null.
esc::CubeProgram *instance_program(esc_batch *b, int i) {
    esc_ctx *c = b->ctx[i];
    if (c->tile_bits <= 0 || !cfg_cube(c->tile_bits)) return nullptr;
    return &cube_for(c, c->tile_bits);
}

// Groups of two or more instances whose kernel choices are the same
// generated program (cube_program shares one program per source and device),
// with their per-phase pointer tables and coupling scalars on the device.
void batch_setup_groups(esc_batch *b) {
    for (auto &g : b->groups)
        for (esc_batch::DevBuf *d : {&g->ptrs, &g->cfs})
            if (d->p) hip_check(hipFree(d->p), "hipFree");
    b->groups.clear();
    b->group_of.assign(b->n, -1);
    std::map<std::pair<esc::CubeProgram *, uint64_t>, std::vector<int>> by;
    for (int i = 0; i < b->n; ++i)
        if (esc::CubeProgram *p = instance_program(b, i)) by[{p, b->ctx[i]->dim}].push_back(i);
    for (auto &kv : by) {
        if (kv.second.size() < 2) continue;
        auto g = std::make_unique<esc_batch::Group>();
        g->inst = kv.second;
        g->prog = kv.first.first;
        g->dim = kv.first.second;
        const size_t m = g->inst.size();
        std::vector<const void *> h(9 * m);
        for (int p = 0; p < 3; ++p)
            for (size_t q = 0; q < m; ++q) {
                const int i = g->inst[q];
                h[(size_t)p * 3 * m + q] = b->lanes[p][i].v;
                h[((size_t)p * 3 + 1) * m + q] = b->lanes[p][i].w;
                h[((size_t)p * 3 + 2) * m + q] = b->dotbuf[i];
            }
        hip_check(hipMemcpy(g->ptrs.get(h.size() * sizeof(void *)), h.data(), h.size() * sizeof(void *),
                            hipMemcpyHostToDevice), "hipMemcpy");
        double *cfs = (double *)g->cfs.get(m * esc::kCubeScalars * sizeof(double));
        for (size_t q = 0; q < m; ++q)
            hip_check(hipMemcpy(cfs + q * esc::kCubeScalars, b->ctx[g->inst[q]]->d_cube_cf,
                                esc::kCubeScalars * sizeof(double), hipMemcpyDeviceToDevice), "hipMemcpy");
        for (int i : g->inst) b->group_of[i] = (int)b->groups.size();
        b->groups.push_back(std::move(g));
    }
}

// Group tuning: instances whose generated programs coincide (same lattice and
// term structure: the first candidate configuration gives one program) are
// timed as one batched launch per candidate configuration, from their slot 1
// into slot 2; the fastest configuration replaces the members' own choices
// when it beats the sum of their best single-instance times.
void batch_tune_groups(esc_batch *b, int reps, const std::vector<double> &solo) {
    std::map<std::pair<esc::CubeProgram *, uint64_t>, std::vector<int>> cls;
    for (int i = 0; i < b->n; ++i) {
        const std::vector<int> cand = cube_candidates(b->ctx[i]);
        if (cand.empty()) continue;
        try {
            cls[{&cube_for(b->ctx[i], cand[0]), b->ctx[i]->dim}].push_back(i);
        } catch (const EscError &) {
            continue;   // no generated kernels for this instance
        }
    }
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    esc_batch::DevBuf ptrs, cfs;
    try {
        for (auto &kv : cls) {
            const std::vector<int> &mem = kv.second;
            const size_t m = mem.size();
            if (m < 2) continue;
            double best_ms = 0.0;
            for (int i : mem) best_ms += solo[i];
            int best = 0;
            std::vector<const void *> h(3 * m);
            for (size_t q = 0; q < m; ++q) {
                h[q] = b->slot(mem[q], 1);
                h[m + q] = b->slot(mem[q], 2);
                h[2 * m + q] = nullptr;
            }
            void **d = (void **)ptrs.get(h.size() * sizeof(void *));
            hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(void *), hipMemcpyHostToDevice), "hipMemcpy");
            double *dcf = (double *)cfs.get(m * esc::kCubeScalars * sizeof(double));
            for (int cfg : cube_candidates(b->ctx[mem[0]])) {
                esc::CubeProgram *p = nullptr;
                bool same = true;
                try {
                    for (int i : mem) {
                        esc::CubeProgram *pi = &cube_for(b->ctx[i], cfg);
                        if (!p) p = pi;
                        same = same && pi == p;
                    }
                } catch (const EscError &) {
                    same = false;
                }
                if (!same) continue;
                for (size_t q = 0; q < m; ++q)
                    hip_check(hipMemcpyAsync(dcf + q * esc::kCubeScalars, b->ctx[mem[q]]->d_cube_cf,
                                             esc::kCubeScalars * sizeof(double), hipMemcpyDeviceToDevice, b->s),
                              "copy");
                auto launch = [&] {
                    hip_check(esc::cube_launch_batched(*p, dcf, (const void *const *)d, (void *const *)d + m,
                                                       (double *const *)d + 2 * m, b->ctx[mem[0]]->dim, (int)m, b->s),
                              "generated hpsi launch (batched)");
                };
                launch();   // warm-up
                hip_check(hipEventRecord(e0, b->s), "hipEventRecord");
                for (int r = 0; r < reps; ++r) launch();
                hip_check(hipEventRecord(e1, b->s), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                float t = 0.0f;
                hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
                if ((double)t / reps < best_ms) {
                    best_ms = (double)t / reps;
                    best = cfg;
                }
            }
            if (best)
                for (int i : mem) b->ctx[i]->tile_bits = best;
        }
    } catch (...) {
        for (esc_batch::DevBuf *x : {&ptrs, &cfs})
            if (x->p) (void)hipFree(x->p);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    for (esc_batch::DevBuf *x : {&ptrs, &cfs})
        if (x->p) (void)hipFree(x->p);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
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
    batch_setup_groups(b);
}

// the launches of step k (phase k % 3) for all instances
void batch_step_launches(esc_batch *b, int phase, bool accum) {
    const int lda = b->cap, ldb = b->cap + 1, ldc = b->cap;
    if (accum)
        hip_check(esc::launch_batch_step(b->c64, esc::kBatchAccum, b->d_lanes[phase], b->n, b->total_blocks,
                                         b->d_partials, b->d_alpha, lda, b->d_beta, ldb, b->d_coef, ldc, b->d_step,
                                         b->s), "batch accumulate");
    for (auto &g : b->groups)
        hip_check(esc::cube_launch_batched(*g->prog, (const double *)g->cfs.p, g->ins(phase), g->outs(phase),
                                           g->dots(phase), g->dim, (int)g->inst.size(), b->s),
                  "generated hpsi launch (batched)");
    for (int i = 0; i < b->n; ++i) {
        if (b->group_of[i] >= 0) continue;
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
// slot 1 (slot 0 is cleared); with accum, acc += coef[i][k] v_k into gs, the
// coefficients read from d_coef (row i at i * cap).  alpha and beta stay in
// d_alpha / d_beta: nothing waits for the device.
void batch_run(esc_batch *b, int steps, bool use_graph, bool accum) {
    batch_capacity(b, steps);
    for (int i = 0; i < b->n; ++i) {
        hip_check(hipMemsetAsync(b->slot(i, 0), 0, b->ctx[i]->dim * b->esz, b->s), "clear");
        if (accum) hip_check(hipMemsetAsync(b->slot(i, 4), 0, b->ctx[i]->dim * b->esz, b->s), "clear");
    }
    hip_check(hipMemsetAsync(b->d_beta, 0, (size_t)b->n * (b->cap + 1) * sizeof(double), b->s), "clear");
    hip_check(hipMemsetAsync(b->d_step, 0, sizeof(uint32_t), b->s), "clear");
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
}

// host copies of alpha (real parts, n x steps) and beta (n x (steps + 1))
// from device rows of leading dimensions cap and cap + 1 (waits for the device)
void batch_read(esc_batch *b, const double *d_alpha, const double *d_beta, int steps, std::vector<double> &alpha,
                std::vector<double> &beta) {
    std::vector<double> a2((size_t)b->n * b->cap * 2), bt((size_t)b->n * (b->cap + 1));
    hip_check(hipMemcpyAsync(a2.data(), d_alpha, a2.size() * sizeof(double), hipMemcpyDeviceToHost, b->s), "alpha");
    hip_check(hipMemcpyAsync(bt.data(), d_beta, bt.size() * sizeof(double), hipMemcpyDeviceToHost, b->s), "beta");
    hip_check(hipStreamSynchronize(b->s), "sync");
    alpha.assign((size_t)b->n * steps, 0.0);
    beta.assign((size_t)b->n * (steps + 1), 0.0);
    for (int i = 0; i < b->n; ++i) {
        for (int k = 0; k < steps; ++k) alpha[(size_t)i * steps + k] = a2[2 * ((size_t)i * b->cap + k)];
        for (int k = 0; k <= steps; ++k) beta[(size_t)i * (steps + 1) + k] = bt[(size_t)i * (b->cap + 1) + k];
    }
}

// x / |x| of instance i into its slot 1, |x|^2 into the device double n2
// (x is copied unscaled when the norm is zero or not finite)
void batch_load_unit(esc_batch *b, int i, const void *x, double *n2) {
    esc_ctx *c = b->ctx[i];
    c->ensure_partials(2 * esc::kRedBlocks);
    hip_check(esc::launch_norm2(b->c64, x, c->dim, c->d_partials, n2, b->s), "norm2");
    void *dst = b->slot(i, 1);
    hip_check(hipMemcpyAsync(dst, x, c->dim * b->esz, hipMemcpyDeviceToDevice, b->s), "copy");
    hip_check(esc::launch_normalize(b->c64, dst, n2, c->dim, b->s), "normalize");
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
        const int n = b->n;
        batch_capacity(b, steps);
        double *norms = (double *)b->norms.get(2 * (size_t)n * sizeof(double));
        // start vectors (Gaussian, seed + i), kept in slot 3 for the replay;
        // one eager H.psi per instance builds its kernel tables before capture
        for (int i = 0; i < n; ++i) {
            esc_ctx *c = b->ctx[i];
            hip_check(esc::launch_gaussian(b->c64, seed + (uint64_t)i, b->slot(i, 3), c->dim, b->s), "gaussian");
            batch_load_unit(b, i, b->slot(i, 3), norms + i);
            hip_check(hipMemcpyAsync(b->slot(i, 3), b->slot(i, 1), c->dim * b->esz, hipMemcpyDeviceToDevice, b->s), "copy");
            run_hpsi(c, b->slot(i, 1), b->slot(i, 2), b->s);
        }
        batch_run(b, steps, use_graph != 0, false);
        // the recursion's rows are kept for the caller; the lowest eigenpair
        // of each T (Sturm bisection, inverse iteration) is solved on the
        // device straight into the replay coefficients
        const size_t na = (size_t)n * b->cap * 2 * sizeof(double), nb = (size_t)n * (b->cap + 1) * sizeof(double);
        double *ka = (double *)b->keep_alpha.get(na), *kb = (double *)b->keep_beta.get(nb);
        hip_check(hipMemcpyAsync(ka, b->d_alpha, na, hipMemcpyDeviceToDevice, b->s), "copy");
        hip_check(hipMemcpyAsync(kb, b->d_beta, nb, hipMemcpyDeviceToDevice, b->s), "copy");
        double *d_e0 = (double *)b->e0_dev.get((size_t)n * sizeof(double));
        int *d_it = (int *)b->iters_dev.get((size_t)n * sizeof(int));
        double *work = (double *)b->work.get((size_t)n * 6 * steps * sizeof(double));
        hip_check(esc::launch_batch_tridiag(n, b->d_alpha, b->cap, b->d_beta, b->cap + 1, steps, work, b->d_coef, b->cap,
                                            d_e0, d_it, b->s),
                  "batch tridiagonal");
        // replay from the same starts, forming |gs> = sum_k Z(k,0) v_k in slot 4
        for (int i = 0; i < n; ++i)
            hip_check(hipMemcpyAsync(b->slot(i, 1), b->slot(i, 3), b->ctx[i]->dim * b->esz, hipMemcpyDeviceToDevice, b->s),
                      "copy");
        batch_run(b, steps, use_graph != 0, true);
        for (int i = 0; i < n; ++i) {
            esc_ctx *c = b->ctx[i];
            c->ensure_partials(2 * esc::kRedBlocks);
            hip_check(esc::launch_norm2(b->c64, b->slot(i, 4), c->dim, c->d_partials, norms + n + i, b->s), "norm2");
            hip_check(esc::launch_normalize(b->c64, b->slot(i, 4), norms + n + i, c->dim, b->s), "normalize");
        }
        // the one wait of the call
        std::vector<int> it(n);
        b->e0.assign(n, 0.0);
        hip_check(hipMemcpyAsync(b->e0.data(), d_e0, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, b->s), "e0");
        hip_check(hipMemcpyAsync(it.data(), d_it, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, b->s), "iters");
        std::vector<double> al, bt;
        batch_read(b, ka, kb, steps, al, bt);
        for (int i = 0; i < n; ++i)
            if (!std::isfinite(b->e0[i]))
                throw EscError(ESC_ERR_NUMERIC, "instance " + std::to_string(i) + ": non-finite tridiagonal matrix");
        if (e0) std::memcpy(e0, b->e0.data(), (size_t)n * sizeof(double));
        if (iters) std::memcpy(iters, it.data(), (size_t)n * sizeof(int));
        if (alpha) std::memcpy(alpha, al.data(), al.size() * sizeof(double));
        if (beta) std::memcpy(beta, bt.data(), bt.size() * sizeof(double));
        b->have_gs = true;
    });
}

int esc_batch_autotune(esc_batch *b, int reps, int *configs) {
    return guarded([&] {
        batch_check(b);
        b->drop_graphs();   // the captured steps bake each instance's kernel choice
        b->have_gs = false;
        std::vector<double> solo(b->n, 0.0);
        for (int i = 0; i < b->n; ++i) {
            esc_ctx *c = b->ctx[i];
            hip_check(esc::launch_gaussian(b->c64, 0x5eedULL + (uint64_t)i, b->slot(i, 1), c->dim, b->s), "gaussian");
            int cfg = 0;
            double ms = 0.0;
            const int rc = esc_autotune(c, b->slot(i, 1), b->slot(i, 2), reps, &cfg, &ms, b->s);
            if (rc != ESC_OK) throw EscError(rc, "instance " + std::to_string(i) + ": " + esc_last_error());
            solo[i] = ms;
        }
        batch_tune_groups(b, reps, solo);
        if (configs)
            for (int i = 0; i < b->n; ++i) configs[i] = b->ctx[i]->tile_bits;
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
        const std::vector<CfRun> runs = cf_runs(js, nc);
        const int R = (int)runs.size();
        // every run's operator table for every instance, uploaded once
        std::vector<esc::CurrentTable> tabs((size_t)n * R);
        for (int i = 0; i < n; ++i) {
            std::vector<esc::CurrentTable> comp(nc);
            for (int a = 0; a < nc; ++a) host_current_table(b->ctx[i], ops[a], ESC_SITE_TOTAL, &comp[a]);
            for (int r = 0; r < R; ++r) {
                esc::CurrentTable &t = tabs[(size_t)i * R + r];
                std::memset(&t, 0, sizeof(t));
                for (int a = 0; a < nc; ++a) {
                    const cplx c = runs[r].coef[a];
                    if (c != cplx(0.0, 0.0)) append_scaled(t, comp[a], c.real(), c.imag());
                }
            }
        }
        auto *d_tabs = (esc::CurrentTable *)b->tabs_dev.get(tabs.size() * sizeof(esc::CurrentTable));
        hip_check(hipMemcpyAsync(d_tabs, tabs.data(), tabs.size() * sizeof(esc::CurrentTable), hipMemcpyHostToDevice,
                                 b->s), "tables");
        auto *d_omega = (double *)b->omega_dev.get((size_t)nw * sizeof(double));
        hip_check(hipMemcpyAsync(d_omega, omega, (size_t)nw * sizeof(double), hipMemcpyHostToDevice, b->s), "omega");
        auto *d_e0 = (double *)b->e0_dev.get((size_t)n * sizeof(double));
        hip_check(hipMemcpyAsync(d_e0, b->e0.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, b->s), "e0");
        auto *d_n2 = (double *)b->n2_dev.get((size_t)R * n * sizeof(double));
        auto *d_ex = (double *)b->ex_dev.get((size_t)n * nc * 2 * sizeof(double));
        void *d_f = b->f_dev.get((size_t)R * n * nw * 2 * sizeof(double));
        for (int r = 0; r < R; ++r) {
            // phi_i = (sum_a c_a J_a)|gs_i> into slot 5; the diagonal runs also
            // give <gs_i|J_a|gs_i>
            for (int i = 0; i < n; ++i) {
                esc_ctx *c = b->ctx[i];
                apply_current_table(c, tabs[(size_t)i * R + r], d_tabs + (size_t)i * R + r, b->slot(i, 4), b->slot(i, 5),
                                    b->s);
                if (runs[r].kind == 0) {
                    c->ensure_partials(2 * esc::kRedBlocks);
                    hip_check(esc::launch_dot(b->c64, b->slot(i, 4), b->slot(i, 5), c->dim, c->d_partials,
                                              d_ex + 2 * ((size_t)i * nc + runs[r].a), b->s),
                              "dot");
                }
                batch_load_unit(b, i, b->slot(i, 5), d_n2 + (size_t)r * n + i);
            }
            batch_run(b, cf_steps, use_graph != 0, false);
            hip_check(esc::launch_batch_cf(n, b->d_alpha, b->cap, b->d_beta, b->cap + 1, cf_steps,
                                           d_n2 + (size_t)r * n, d_e0, eta, d_omega, nw,
                                           (char *)d_f + (size_t)r * n * nw * 2 * sizeof(double), b->s),
                      "batch continued fractions");
        }
        // the one wait of the call: fractions and expectation values
        std::vector<cplx> fh((size_t)R * n * nw);
        std::vector<double> exh((size_t)n * nc * 2);
        hip_check(hipMemcpyAsync(fh.data(), d_f, fh.size() * sizeof(cplx), hipMemcpyDeviceToHost, b->s), "fractions");
        hip_check(hipMemcpyAsync(exh.data(), d_ex, exh.size() * sizeof(double), hipMemcpyDeviceToHost, b->s), "ex");
        hip_check(hipStreamSynchronize(b->s), "sync");
        for (int i = 0; i < n; ++i) {
            std::vector<cplx> ex(nc + 1, cplx(0.0, 0.0));
            for (int a = 0; a < nc; ++a) {
                ex[a] = cplx(exh[2 * ((size_t)i * nc + a)], exh[2 * ((size_t)i * nc + a) + 1]);
                ex[nc] += ex[a];
            }
            std::vector<std::vector<cplx>> f(R);
            for (int r = 0; r < R; ++r) {
                const cplx *src = fh.data() + ((size_t)r * n + i) * nw;
                f[r].assign(src, src + nw);
            }
            auto F = cf_assemble(js, nc, runs, f, nw);
            write_spectrum_table(js, nc, ex, F, omega, nw, eta, tables + (size_t)i * nw * (1 + ncols));
            if (expect)
                for (int a = 0; a <= nc; ++a) {
                    expect[2 * ((size_t)i * 4 + a)] = ex[a].real();
                    expect[2 * ((size_t)i * 4 + a) + 1] = ex[a].imag();
                }
        }
    });
}

}  // extern "C"
