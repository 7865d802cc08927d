// esc_spectra.cpp -- spectra from a ground state: the continued-fraction
// runs and their assembly into the JsJs / JeJe tables (shared by one context,
// batches and the gradients), LocalSzSz, and the RIXS state sums.
#include "esc_ctx.h"
#include "esc_tridiag.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <set>
#include <thread>
#include <vector>

using namespace esc_impl;

namespace esc_impl {

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
