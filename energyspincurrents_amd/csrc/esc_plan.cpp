// esc_plan.cpp -- what a model becomes on the device: the H.psi term tables,
// the window planner of the tiled kernel and the specification of the
// generated kernels (kernel configurations), and the operator tables of the
// site, spin-current and energy-current operators.
#include "esc_ctx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

using namespace esc_impl;

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
                                       const std::vector<HostTerm> *custom) {
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
               bool do_plan) {
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

}  // namespace esc_impl
