// esc_grad.cpp -- gradients of a continued-fraction spectrum loss with
// respect to (K, J, Gamma, h) by a hand-written adjoint of the Lanczos
// recursion (esc_spectrum_grad).
#include "esc_ctx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <string>
#include <vector>

using namespace esc_impl;

// ------------------------------------------------------------------------
// Gradients of a spectrum loss with respect to (K, J, Gamma, h) by a
// hand-written adjoint of the ground-state problem, the Lanczos recursion
// from phi = O|gs> and the continued fraction, with the recursion vectors
// checkpointed every C steps and recomputed segment by segment.
// ------------------------------------------------------------------------
namespace {

struct DevVecs {
    std::vector<void *> p;
    size_t bytes = 0;
    explicit DevVecs(size_t b) : bytes(b) {}
    void *get() {
        void *x = nullptr;
        if (hipMalloc(&x, bytes) != hipSuccess) throw EscError(ESC_ERR_ALLOC, "gradient work vector allocation failed");
        p.push_back(x);
        return x;
    }
    ~DevVecs() {
        for (void *x : p) (void)hipFree(x);
    }
};

struct CtxPtr {
    esc_ctx *c = nullptr;
    ~CtxPtr() { delete c; }
};

void make_ctx(const esc_model &m, CtxPtr &out) {
    const int rc = esc_ctx_create(&m, ESC_C128, &out.c);
    if (rc != ESC_OK) throw EscError(rc, esc_last_error());
}

// The continued fraction G(w) = n2 g_0, g_k = 1 / (z - a_k - b_{k+1}^2 g_{k+1}),
// z = w + e0 - i eta, on every omega
void cf_values(int M, const double *a, const double *b, double n2, double e0, double eta, const double *omega, int nw,
               std::vector<cplx> &G) {
    G.assign(nw, cplx(0.0, 0.0));
    std::vector<cplx> g(std::max(M, 1));
    for (int q = 0; q < nw; ++q) {
        const cplx z(omega[q] + e0, -eta);
        for (int k = M - 1; k >= 0; --k) {
            const cplx tail = (k + 1 < M) ? b[k + 1] * b[k + 1] * g[k + 1] : cplx(0.0, 0.0);
            g[k] = 1.0 / (z - a[k] - tail);
        }
        G[q] = M > 0 ? n2 * g[0] : cplx(0.0, 0.0);
    }
}

// Derivatives of L with respect to a_k, b_k, n2 and e0, given the loss's
// sensitivity to the fraction on each omega as dL = Re sum_q w_q dG(w_q)
// (holomorphic chain rule per frequency)
void cf_adjoint(int M, const double *a, const double *b, double n2, double e0, double eta, const double *omega, int nw,
                const cplx *w, std::vector<double> &abar, std::vector<double> &bbar, double &n2bar, double &e0bar) {
    abar.assign(M, 0.0);
    bbar.assign(M + 1, 0.0);
    n2bar = 0.0;
    e0bar = 0.0;
    std::vector<cplx> g(std::max(M, 1));
    for (int q = 0; q < nw; ++q) {
        if (w[q] == cplx(0.0, 0.0)) continue;
        const cplx z(omega[q] + e0, -eta);
        for (int k = M - 1; k >= 0; --k) {
            const cplx tail = (k + 1 < M) ? b[k + 1] * b[k + 1] * g[k + 1] : cplx(0.0, 0.0);
            g[k] = 1.0 / (z - a[k] - tail);
        }
        cplx c = n2, dz(0.0, 0.0);
        for (int k = 0; k < M; ++k) {
            const cplx gk2 = g[k] * g[k];
            abar[k] += (w[q] * c * gk2).real();
            dz -= c * gk2;
            if (k + 1 < M) {
                bbar[k + 1] += (w[q] * c * 2.0 * b[k + 1] * g[k + 1] * gk2).real();
                c = c * b[k + 1] * b[k + 1] * gk2;
            }
        }
        if (M > 0) n2bar += (w[q] * g[0]).real();
        e0bar += (w[q] * dz).real();
    }
}

// The loss as a function of the runs' fractions G_r(w_q) and the expectation
// values ex_a = <gs|J_a|gs> of the base operators, with its sensitivities:
// dL = Re sum_{r,q} w[r][q] dG_r(w_q) + sum_a exbar[a] d(Re ex_a).
struct Upstream {
    double loss = 0.0;
    std::vector<std::vector<cplx>> w;
    std::vector<double> exbar;
};
using UpstreamFn = std::function<Upstream(const std::vector<std::vector<cplx>> &G, const std::vector<cplx> &ex)>;

// The adjoint shared by the entry points.  Base operators J_a = (op, osite);
// run r's operator O_r = sum_a run_coef[r][a] J_a (its continued fraction
// starts at O_r|gs>).  Gradient with respect to theta = (K, J, Gamma, h).
void grad_core(const esc_model &model, const std::vector<int> &base_ops, int osite,
               const std::vector<std::vector<cplx>> &run_coef, int gs_steps, int cf_steps, int checkpoint, double eta,
               const double *omega, int nw, const UpstreamFn &upstream, double *grad, double *e0_out, double *de0,
               hipStream_t s) {
    const int nb = (int)base_ops.size(), R = (int)run_coef.size();
    // currents (Js ~ K, Je ~ K^2) move with K; single-site operators do not
    bool kdep = false;
    for (int op : base_ops)
        if (op >= ESC_OP_JSX && op <= ESC_OP_JEY) kdep = true;

    // H and its parts dH/dK (along (1,1,1) in Kx,Ky,Kz), dH/dJ, dH/dGamma, dH/dh ([111] field)
    CtxPtr H, P[4], Oplus, Ominus;
    make_ctx(model, H);
    for (int p = 0; p < 4; ++p) {
        esc_model m = model;
        m.K[0] = m.K[1] = m.K[2] = (p == 0) ? 1.0 : 0.0;
        m.J = (p == 1) ? 1.0 : 0.0;
        m.Gamma = (p == 2) ? 1.0 : 0.0;
        const double hb = (p == 3) ? 1.0 / std::sqrt(3.0) : 0.0;
        m.B[0] = m.B[1] = m.B[2] = hb;
        make_ctx(m, P[p]);
    }
    if (kdep) {   // dO/dK = (O(K + 1) - O(K - 1)) / 2, exact for couplings of degree <= 2
        esc_model mp = model, mm = model;
        for (int f = 0; f < 3; ++f) { mp.K[f] += 1.0; mm.K[f] -= 1.0; }
        make_ctx(mp, Oplus);
        make_ctx(mm, Ominus);
    }
    esc_ctx *c = H.c;
    const uint64_t dim = c->dim;
    DevVecs V(dim * c->esz);

    // operator tables on the device: per context (H, K + 1, K - 1), the base
    // operators, then per run O_r and O_r^dagger
    const int nctx = kdep ? 3 : 1;
    esc_ctx *cx[3] = {c, kdep ? Oplus.c : nullptr, kdep ? Ominus.c : nullptr};
    const int ntab = nb + 2 * R;
    std::vector<esc::CurrentTable> host((size_t)nctx * ntab);
    for (int x = 0; x < nctx; ++x) {
        std::vector<esc::CurrentTable> base(nb);
        for (int a = 0; a < nb; ++a) host_current_table(cx[x], base_ops[a], osite, &base[a]);
        for (int a = 0; a < nb; ++a) host[(size_t)x * ntab + a] = base[a];
        for (int r = 0; r < R; ++r)
            for (int dag = 0; dag < 2; ++dag) {
                esc::CurrentTable &t = host[(size_t)x * ntab + nb + 2 * r + dag];
                std::memset(&t, 0, sizeof(t));
                for (int a = 0; a < nb; ++a) {
                    const cplx k = dag ? std::conj(run_coef[r][a]) : run_coef[r][a];
                    if (k != cplx(0.0, 0.0)) append_scaled(t, base[a], k.real(), k.imag());
                }
            }
    }
    DevVecs TB(host.size() * sizeof(esc::CurrentTable));
    auto *d_tab = static_cast<esc::CurrentTable *>(TB.get());
    hip_check(hipMemcpy(d_tab, host.data(), host.size() * sizeof(esc::CurrentTable), hipMemcpyHostToDevice), "tables");
    auto tab = [&](int x, int i) { return (size_t)x * ntab + i; };
    auto apply = [&](size_t t, const void *x, void *y) { apply_current_table(c, host[t], d_tab + t, x, y, s); };
    auto axpby = [&](cplx a, const void *x, cplx b, void *y) {
        hip_check(esc::launch_axpby(false, a.real(), a.imag(), x, b.real(), b.imag(), y, dim, s), "axpby");
    };
    auto dot = [&](const void *x, const void *y) {
        double d[2];
        device_dot(c, x, y, d, s);
        return cplx(d[0], d[1]);
    };

    // ground state (FP64, replayed recursion with one restart) and E0
    void *v0 = V.get(), *gs = V.get();
    hip_check(esc::launch_gaussian(false, 0x9e3779b97f4a7c15ULL, v0, dim, s), "gaussian");
    double e0 = 0.0, resid = 0.0;
    int it = 0;
    if (esc_ground_state(c, v0, gs_steps, 1e-14, 1, &e0, &resid, &it, gs, s) != ESC_OK)
        throw EscError(ESC_ERR_NUMERIC, std::string("ground state: ") + esc_last_error());
    progress("gradient: ground state, %d H.psi", it);
    void *t0 = V.get(), *t1 = V.get();
    void *cg_r = V.get(), *cg_p = V.get(), *cg_ap = V.get();
    // x = (H - e)^+ b on the complement of gs (b orthogonal to gs), by
    // conjugate gradients until the relative residual reaches tol, the
    // curvature test fails, or (stall > 0) stall iterations bring no
    // quartering of the best squared residual; returns the relative residual
    // reached
    auto cg = [&](double e, const void *b, void *x, double tol, int stall) {
        hip_check(hipMemsetAsync(x, 0, dim * c->esz, s), "clear");
        hip_check(hipMemcpyAsync(cg_r, b, dim * c->esz, hipMemcpyDeviceToDevice, s), "copy");
        hip_check(hipMemcpyAsync(cg_p, b, dim * c->esz, hipMemcpyDeviceToDevice, s), "copy");
        double rr = device_norm2(c, cg_r, s);
        const double bn = std::sqrt(rr);
        if (!(bn > 0.0)) return 0.0;
        int iter = 0, best_at = 0;
        double best = rr;
        for (; iter < 5000 && std::sqrt(rr) > tol * bn && (stall == 0 || iter - best_at < stall); ++iter) {
            run_hpsi(c, cg_p, cg_ap, s);
            axpby(-e, cg_p, 1.0, cg_ap);
            axpby(-dot(gs, cg_ap), gs, 1.0, cg_ap);
            const double pAp = dot(cg_p, cg_ap).real();
            if (!(pAp > 0.0)) break;   // no positive curvature left: the solve is as good as it gets
            const double a = rr / pAp;
            axpby(a, cg_p, 1.0, x);
            axpby(-a, cg_ap, 1.0, cg_r);
            const double rr2 = device_norm2(c, cg_r, s);
            axpby(1.0, cg_r, rr2 / rr, cg_p);
            rr = rr2;
            if (rr < 0.25 * best) {
                best = rr;
                best_at = iter + 1;
            }
        }
        progress("conjugate gradients: %d iterations, relative residual %.2e", iter, std::sqrt(rr) / bn);
        return std::sqrt(rr) / bn;
    };
    // residual correction of the ground state (Newton steps for the
    // eigenvector): r = (H - e0) gs, gs -= (H - e0)^+ r, renormalise
    for (int pass = 0; pass < 4; ++pass) {
        run_hpsi(c, gs, t0, s);
        e0 = dot(gs, t0).real();
        axpby(-e0, gs, 1.0, t0);
        resid = std::sqrt(device_norm2(c, t0, s));
        progress("gradient: Newton pass %d, E0 %.15f, residual %.3e", pass, e0, resid);
        if (resid < 1e-13 * std::max(1.0, std::fabs(e0))) break;
        cg(e0, t0, t1, 1e-12, 0);
        axpby(-1.0, t1, 1.0, gs);
        hip_check(esc::launch_scale(false, gs, 1.0 / std::sqrt(device_norm2(c, gs, s)), dim, s), "scale");
    }
    run_hpsi(c, gs, t0, s);
    e0 = dot(gs, t0).real();
    if (e0_out) *e0_out = e0;
    // Hellmann-Feynman: dE0/dtheta_p = <gs|H_p|gs>
    double hf[4];
    for (int p = 0; p < 4; ++p) {
        run_hpsi(P[p].c, gs, t0, s);
        hf[p] = dot(gs, t0).real();
        if (de0) de0[p] = hf[p];
    }
    // expectation values of the base operators
    std::vector<cplx> ex(nb);
    for (int a = 0; a < nb; ++a) {
        apply(tab(0, a), gs, t0);
        ex[a] = dot(gs, t0);
    }

    // Lanczos recursion from phi (normalised inside).  checkpoint > 0: three
    // rotating vectors, with store the pair (v_{k-1}, v_k) saved every C steps
    // (segments are recomputed in the reverse sweep).  checkpoint == 0: every
    // vector kept (in the context's Lanczos storage) and each new one
    // reorthogonalised against all before it, so the floating-point recursion
    // stays the exact-arithmetic one the adjoint differentiates
    // (reorthogonalisation changes nothing in exact arithmetic, so the adjoint
    // formulas are the same).  Deterministic: a second run repeats the first.
    const bool full = checkpoint == 0;
    const int C = full ? cf_steps : checkpoint;
    if (full) c->ensure_vecs((size_t)cf_steps + 1);
    c->ensure_partials(2 * esc::kRedBlocks);
    c->ensure_result(8);
    auto step = [&](const void *vp, const void *v, void *w, double bk, double *ak, double *bk1) {
        run_hpsi(c, v, w, s);
        double a[2];
        hip_check(esc::launch_dot(false, v, w, dim, c->d_partials, c->d_result, s), "dot");
        sync_get(s, a, c->d_result, 2 * sizeof(double));
        hip_check(esc::launch_lanczos_update(false, w, v, vp, c->d_result, bk, dim, c->d_partials, c->d_result + 2, s),
                  "update");
        double b2;
        sync_get(s, &b2, c->d_result + 2, sizeof(double));
        *ak = a[0];
        *bk1 = std::sqrt(b2);
        if (!std::isfinite(*ak) || !std::isfinite(*bk1)) throw EscError(ESC_ERR_NUMERIC, "non-finite recursion");
        if (*bk1 > 1e-13 * std::max(1.0, std::fabs(a[0])))
            hip_check(esc::launch_scale(false, w, 1.0 / *bk1, dim, s), "scale");
    };
    void *zero = V.get();
    hip_check(hipMemsetAsync(zero, 0, dim * c->esz, s), "clear");
    void *rot[3] = {zero, nullptr, nullptr};
    if (!full) {
        rot[0] = V.get();
        rot[1] = V.get();
        rot[2] = V.get();
    }
    std::vector<void *> ckp_pool, ckp;   // ckp[2j] = v_{jC-1} (or zero), ckp[2j+1] = v_{jC}
    auto forward = [&](const void *phi, double nphi, bool store, std::vector<double> &al, std::vector<double> &be) {
        c->basis_kept = false;
        al.clear();
        be.assign(1, 0.0);
        ckp.clear();
        void *first = full ? c->vec(0) : rot[1];
        hip_check(hipMemcpyAsync(first, phi, dim * c->esz, hipMemcpyDeviceToDevice, s), "copy");
        hip_check(esc::launch_scale(false, first, 1.0 / nphi, dim, s), "scale");
        if (!full) hip_check(hipMemsetAsync(rot[0], 0, dim * c->esz, s), "clear");
        for (int k = 0; k < cf_steps; ++k) {
            void *vp = full ? (k > 0 ? c->vec((size_t)k - 1) : zero) : rot[k % 3];
            void *v = full ? c->vec((size_t)k) : rot[(k + 1) % 3];
            void *w = full ? c->vec((size_t)k + 1) : rot[(k + 2) % 3];
            if (!full && store && k % C == 0) {
                const size_t n = ckp.size();
                while (ckp_pool.size() < n + 2) ckp_pool.push_back(V.get());
                hip_check(hipMemcpyAsync(ckp_pool[n], vp, dim * c->esz, hipMemcpyDeviceToDevice, s), "checkpoint");
                hip_check(hipMemcpyAsync(ckp_pool[n + 1], v, dim * c->esz, hipMemcpyDeviceToDevice, s), "checkpoint");
                ckp.push_back(ckp_pool[n]);
                ckp.push_back(ckp_pool[n + 1]);
            }
            double ak, bk1;
            step(k > 0 ? vp : zero, v, w, be[k], &ak, &bk1);
            if (full && k > 0 && bk1 > 1e-13 * std::max(1.0, std::fabs(ak))) {
                cgs_pass(c, k + 1, w, s);
                cgs_pass(c, k + 1, w, s);
                hip_check(esc::launch_norm2(false, w, dim, c->d_partials, c->d_result + 4, s), "norm2");
                hip_check(esc::launch_normalize(false, w, c->d_result + 4, dim, s), "normalize");
            }
            al.push_back(ak);
            be.push_back(bk1);
            if (!(bk1 > 1e-13 * std::max(1.0, std::fabs(ak)))) break;   // Krylov space exhausted
        }
    };

    // pass 1: every run's recursion and fraction
    std::vector<void *> phi(R);
    std::vector<double> n2(R);
    std::vector<std::vector<double>> AL(R), BE(R);
    std::vector<std::vector<cplx>> G(R);
    for (int r = 0; r < R; ++r) {
        phi[r] = V.get();
        apply(tab(0, nb + 2 * r), gs, phi[r]);
        n2[r] = device_norm2(c, phi[r], s);
        if (!(n2[r] > 0.0)) throw EscError(ESC_ERR_NUMERIC, "O|gs> vanishes");
        forward(phi[r], std::sqrt(n2[r]), false, AL[r], BE[r]);
        cf_values((int)AL[r].size(), AL[r].data(), BE[r].data(), n2[r], e0, eta, omega, nw, G[r]);
        progress("gradient: run %d, forward recursion, %d steps", r, (int)AL[r].size());
    }
    const Upstream up = upstream(G, ex);
    if ((int)up.w.size() != R || (int)up.exbar.size() != nb) throw EscError(ESC_ERR_ARG, "internal: upstream shape");

    double g[4] = {0.0, 0.0, 0.0, 0.0};
    double e0bar = 0.0;
    void *gsbar = V.get();
    hip_check(hipMemsetAsync(gsbar, 0, dim * c->esz, s), "clear");
    DevVecs GD(sizeof(double) * 8 * (size_t)std::max(cf_steps, 1));
    double *gdot = static_cast<double *>(GD.get());
    std::vector<void *> seg(C + 2);
    if (!full)
        for (auto &x : seg) x = V.get();
    void *vbar_next = V.get(), *vbar_k = V.get(), *vbar_prev = V.get(), *wbar = V.get(), *ubar = V.get();
    void *phibar = V.get();

    // pass 2: per run, the recursion again (stored) and its reverse sweep
    for (int r = 0; r < R; ++r) {
        bool any = false;
        for (const cplx &x : up.w[r]) any = any || x != cplx(0.0, 0.0);
        if (!any) continue;
        std::vector<double> al, be;
        forward(phi[r], std::sqrt(n2[r]), true, al, be);
        if (al != AL[r] || be != BE[r]) throw EscError(ESC_ERR_NUMERIC, "recomputed recursion differs");
        const int M = (int)al.size();
        std::vector<double> abar, bbar;
        double n2bar = 0.0, e0b = 0.0;
        cf_adjoint(M, al.data(), be.data(), n2[r], e0, eta, omega, nw, up.w[r].data(), abar, bbar, n2bar, e0b);
        e0bar += e0b;
        // the sweep's Re<u_bar|H_p v_k> go to device slots without a host
        // round trip and are added to g afterwards, in the sweep's order
        size_t ngdot = 0;
        c->ensure_partials(2 * esc::kRedBlocks);
        hip_check(hipMemsetAsync(vbar_next, 0, dim * c->esz, s), "clear");
        hip_check(hipMemsetAsync(vbar_k, 0, dim * c->esz, s), "clear");
        const int nseg = (M + C - 1) / C;
        for (int j = nseg - 1; j >= 0; --j) {
            const int k0 = j * C, k1 = std::min(M, k0 + C);
            // seg[0] = v_{k0-1}, seg[1 + i] = v_{k0 + i}, i = 0..k1-k0
            if (full) {   // all vectors are resident
                seg[0] = zero;
                for (int i = 0; i <= k1 - k0; ++i) seg[1 + i] = c->vec((size_t)(k0 + i));
            } else {
                hip_check(hipMemcpyAsync(seg[0], ckp[2 * j], dim * c->esz, hipMemcpyDeviceToDevice, s), "restore");
                hip_check(hipMemcpyAsync(seg[1], ckp[2 * j + 1], dim * c->esz, hipMemcpyDeviceToDevice, s), "restore");
                for (int k = k0; k < k1; ++k) {
                    double ak, bk1;
                    step(k > 0 ? seg[k - k0] : zero, seg[k - k0 + 1], seg[k - k0 + 2], be[k], &ak, &bk1);
                    if (ak != al[k] || bk1 != be[k + 1])
                        throw EscError(ESC_ERR_NUMERIC, "recomputed recursion differs");
                }
            }
            for (int k = k1 - 1; k >= k0; --k) {
                const void *vkm1 = k > 0 ? seg[k - k0] : zero, *vk = seg[k - k0 + 1], *vk1 = seg[k - k0 + 2];
                // w_bar = v_bar_{k+1} / b + (b_bar_{k+1} - Re<v_{k+1}|v_bar_{k+1}>/b) v_{k+1}
                hip_check(hipMemsetAsync(wbar, 0, dim * c->esz, s), "clear");
                const double b = be[k + 1];
                if (k + 1 < M && b > 0.0) {
                    const double pr = dot(vk1, vbar_next).real();
                    axpby(1.0 / b, vbar_next, 0.0, wbar);
                    axpby(bbar[k + 1] - pr / b, vk1, 1.0, wbar);
                }
                const double atot = abar[k] - dot(wbar, vk).real();
                // u_bar = w_bar + a_tot v_k
                hip_check(hipMemcpyAsync(ubar, wbar, dim * c->esz, hipMemcpyDeviceToDevice, s), "copy");
                axpby(atot, vk, 1.0, ubar);
                // v_bar_k += -a_k w_bar + a_tot H v_k + H u_bar
                run_hpsi(c, vk, t0, s);
                axpby(-al[k], wbar, 1.0, vbar_k);
                axpby(atot, t0, 1.0, vbar_k);
                run_hpsi(c, ubar, t1, s);
                axpby(1.0, t1, 1.0, vbar_k);
                // beta_k and v_{k-1}
                hip_check(hipMemsetAsync(vbar_prev, 0, dim * c->esz, s), "clear");
                if (k > 0) {
                    bbar[k] += -dot(wbar, vkm1).real();
                    axpby(-be[k], wbar, 0.0, vbar_prev);
                }
                // theta_bar_p += Re<u_bar|H_p v_k>
                for (int p = 0; p < 4; ++p) {
                    run_hpsi(P[p].c, vk, t0, s);
                    hip_check(esc::launch_dot(false, ubar, t0, dim, c->d_partials, gdot + 2 * ngdot++, s), "dot");
                }
                std::swap(vbar_next, vbar_k);   // v_bar_k is complete
                std::swap(vbar_k, vbar_prev);
            }
        }
        {
            std::vector<double> h(2 * ngdot);
            if (ngdot) sync_get(s, h.data(), gdot, h.size() * sizeof(double));
            for (size_t q = 0; q < ngdot; ++q) g[q % 4] += h[2 * q];
        }
        progress("gradient: run %d, reverse sweep", r);
        // vbar_next holds v_bar_0: phi_bar = (v_bar_0 - v_0 Re<v_0|v_bar_0>) / |phi| + 2 n2_bar phi
        const double nphi = std::sqrt(n2[r]);
        void *v0u = seg[1];   // the last segment restored was the first: v_0
        const double pr0 = dot(v0u, vbar_next).real();
        hip_check(hipMemsetAsync(phibar, 0, dim * c->esz, s), "clear");
        axpby(1.0 / nphi, vbar_next, 0.0, phibar);
        axpby(-pr0 / nphi, v0u, 1.0, phibar);
        axpby(2.0 * n2bar, phi[r], 1.0, phibar);
        // O_r's own dependence on K
        if (kdep) {
            apply(tab(1, nb + 2 * r), gs, t0);
            apply(tab(2, nb + 2 * r), gs, t1);
            axpby(-1.0, t1, 1.0, t0);
            g[0] += 0.5 * dot(phibar, t0).real();
        }
        // gs_bar += O_r^dagger phi_bar
        apply(tab(0, nb + 2 * r + 1), phibar, t0);
        axpby(1.0, t0, 1.0, gsbar);
    }
    progress("gradient: recursion terms %.10g %.10g %.10g %.10g", g[0], g[1], g[2], g[3]);
    // the expectation values: d Re<gs|J_a|gs> = 2 Re<J_a gs|d gs> + <gs|dJ_a/dK|gs> dK
    for (int a = 0; a < nb; ++a) {
        if (up.exbar[a] == 0.0) continue;
        apply(tab(0, a), gs, t0);
        axpby(2.0 * up.exbar[a], t0, 1.0, gsbar);
        if (kdep) {
            apply(tab(1, a), gs, t0);
            apply(tab(2, a), gs, t1);
            axpby(-1.0, t1, 1.0, t0);
            g[0] += up.exbar[a] * 0.5 * dot(gs, t0).real();
        }
    }
    // E0 through z: e0_bar dE0/dtheta_p (Hellmann-Feynman)
    for (int p = 0; p < 4; ++p) g[p] += e0bar * hf[p];
    progress("gradient: E0 terms %.10g %.10g %.10g %.10g", e0bar * hf[0], e0bar * hf[1], e0bar * hf[2], e0bar * hf[3]);
    // the ground state: mu = (H - E0)^+ P gs_bar by CG on the complement of
    // gs; theta_bar_p -= Re<mu|H_p gs>.  gs_bar is large (|gs_bar| / |grad|
    // reaches 1e9 at 120 steps, in directions far above E0) and its part of
    // the gradient small, so the solve runs to the rounding level: a residual
    // of 1e-11 |gs_bar| moved 120-step gradients by 1e-3 relative.
    const cplx pg = dot(gs, gsbar);
    axpby(-pg, gs, 1.0, gsbar);
    void *mu = V.get();
    if (cg(e0, gsbar, mu, 1e-15, 50) > 1e-8)
        throw EscError(ESC_ERR_NUMERIC, "CG with H - E0 did not converge (ground state not separated by a gap?)");
    double gpart[4];
    for (int p = 0; p < 4; ++p) {
        run_hpsi(P[p].c, gs, t0, s);
        gpart[p] = -dot(mu, t0).real();
        g[p] += gpart[p];
    }
    progress("gradient: ground-state terms %.10g %.10g %.10g %.10g", gpart[0], gpart[1], gpart[2], gpart[3]);
    hip_check(hipStreamSynchronize(s), "sync");
    progress("gradient: done");
    for (int p = 0; p < 4; ++p) grad[p] = g[p];
}

void check_grad_args(const esc_model *model, int gs_steps, int cf_steps, int checkpoint, double eta,
                     const double *omega, int nw, const double *target, const double *grad) {
    check_ptr(model, "model");
    check_ptr(omega, "omega");
    check_ptr(target, "target");
    check_ptr(grad, "grad");
    if (nw < 1) throw EscError(ESC_ERR_ARG, "nw must be >= 1");
    if (gs_steps < 2 || gs_steps > 4096 || cf_steps < 2 || cf_steps > 4096)
        throw EscError(ESC_ERR_ARG, "steps must be in [2, 4096]");
    if (checkpoint < 0) throw EscError(ESC_ERR_ARG, "checkpoint interval must be >= 0");
    if (!(eta > 0.0)) throw EscError(ESC_ERR_ARG, "eta must be > 0");
}

}  // namespace

extern "C" {

int esc_spectrum_grad(const esc_model *model, int op, int site, int gs_steps, int cf_steps, int checkpoint,
                      double eta, const double *omega, int nw, const double *target, double *loss, double *spectrum,
                      double *grad, double *e0_out, double *de0, void *stream) {
    return guarded([&] {
        check_grad_args(model, gs_steps, cf_steps, checkpoint, eta, omega, nw, target, grad);
        const bool current = op >= ESC_OP_JSX && op <= ESC_OP_JEY;
        if (!current && (op < ESC_OP_SX || op > ESC_OP_SZ)) throw EscError(ESC_ERR_ARG, "unknown operator");
        if (!current && (site < 0 || site >= model->n_sites)) throw EscError(ESC_ERR_ARG, "site out of range");
        // L = sum_q (Im G(w_q) - target_q)^2: dL = Re sum_q 2 (S_q - t_q)(-i) dG_q
        UpstreamFn up = [&](const std::vector<std::vector<cplx>> &G, const std::vector<cplx> &) {
            Upstream u;
            u.w.assign(1, std::vector<cplx>(nw));
            u.exbar.assign(1, 0.0);
            for (int q = 0; q < nw; ++q) {
                const double S = G[0][q].imag();
                if (spectrum) spectrum[q] = S;
                const double d = S - target[q];
                u.loss += d * d;
                u.w[0][q] = cplx(0.0, -2.0 * d);
            }
            if (loss) *loss = u.loss;
            return u;
        };
        grad_core(*model, {op}, current ? ESC_SITE_TOTAL : site, {{cplx(1.0, 0.0)}}, gs_steps, cf_steps, checkpoint,
                  eta, omega, nw, up, grad, e0_out, de0, (hipStream_t)stream);
    });
}

int esc_table_grad(const esc_model *model, int kind, int gs_steps, int cf_steps, int checkpoint, double eta,
                   const double *omega, int nw, const double *target, double *loss, double *table, double *grad,
                   double *e0_out, double *de0, void *stream) {
    return guarded([&] {
        check_grad_args(model, gs_steps, cf_steps, checkpoint, eta, omega, nw, target, grad);
        if (kind != ESC_DYN_JSJS && kind != ESC_DYN_JEJE) throw EscError(ESC_ERR_ARG, "kind must be JsJs or JeJe");
        const bool js = kind == ESC_DYN_JSJS;
        const std::vector<int> ops = js ? std::vector<int>{ESC_OP_JSX, ESC_OP_JSY, ESC_OP_JSZ}
                                        : std::vector<int>{ESC_OP_JEX, ESC_OP_JEY};
        const int nc = (int)ops.size(), ncols = js ? 14 : 5;
        const std::vector<CfRun> runs = cf_runs(js, nc);
        const int R = (int)runs.size();
        const auto pairs = spectrum_pairs(js, nc);
        // kappa[p][r]: the table's pair p as a linear combination of the runs'
        // fractions, read off the assembly itself with unit inputs
        std::vector<std::vector<cplx>> kappa(pairs.size(), std::vector<cplx>(R));
        for (int r = 0; r < R; ++r) {
            std::vector<std::vector<cplx>> f(R, std::vector<cplx>(1, cplx(0.0, 0.0)));
            f[r][0] = 1.0;
            auto F = cf_assemble(js, nc, runs, f, 1);
            for (size_t p = 0; p < pairs.size(); ++p) kappa[p][r] = F[pairs[p]][0];
        }
        std::vector<std::vector<cplx>> coef(R);
        for (int r = 0; r < R; ++r) coef[r] = runs[r].coef;
        std::vector<double> tab((size_t)nw * (1 + ncols));
        UpstreamFn up = [&](const std::vector<std::vector<cplx>> &G, const std::vector<cplx> &ex) {
            // the table exactly as the spectra entry points write it
            std::vector<cplx> e(nc + 1, cplx(0.0, 0.0));
            for (int a = 0; a < nc; ++a) {
                e[a] = ex[a];
                e[nc] += ex[a];
            }
            auto F = cf_assemble(js, nc, runs, G, nw);
            write_spectrum_table(js, nc, e, F, omega, nw, eta, tab.data());
            Upstream u;
            u.w.assign(R, std::vector<cplx>(nw, cplx(0.0, 0.0)));
            std::vector<double> eb(nc + 1, 0.0);
            for (int q = 0; q < nw; ++q) {
                const double *row = tab.data() + (size_t)q * (1 + ncols);
                const cplx zeta = 1.0 / cplx(omega[q], -eta);
                int col = 1;
                for (size_t p = 0; p < pairs.size(); ++p) {
                    const int a = pairs[p].first, b = pairs[p].second;
                    const bool elastic = js || !(a == 1 && b == 0);
                    for (int part = js ? 0 : 1; part < 2; ++part, ++col) {   // 0: real part, 1: imaginary part
                        const double d = row[col] - target[(size_t)q * ncols + (col - 1)];
                        u.loss += d * d;
                        const double dd = 2.0 * d;
                        for (int r = 0; r < R; ++r) u.w[r][q] += dd * (part ? cplx(0.0, -1.0) * kappa[p][r] : kappa[p][r]);
                        if (elastic) {
                            // the column subtracts Re/Im of e_a e_b zeta (the expectation values are real)
                            const cplx za = zeta * e[b].real(), zb = zeta * e[a].real();
                            eb[a] -= dd * (part ? za.imag() : za.real());
                            eb[b] -= dd * (part ? zb.imag() : zb.real());
                        }
                    }
                }
            }
            u.exbar.assign(nc, 0.0);
            for (int a = 0; a < nc; ++a) u.exbar[a] = eb[a] + eb[nc];
            if (loss) *loss = u.loss;
            return u;
        };
        grad_core(*model, ops, ESC_SITE_TOTAL, coef, gs_steps, cf_steps, checkpoint, eta, omega, nw, up, grad, e0_out,
                  de0, (hipStream_t)stream);
        if (table) std::memcpy(table, tab.data(), tab.size() * sizeof(double));
    });
}

}  // extern "C"
