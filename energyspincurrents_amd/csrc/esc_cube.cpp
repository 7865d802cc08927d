// esc_cube.cpp -- run-time specialised H.psi passes (see esc_cube.h).
//
// Source generation, compilation with hiprtc for gfx950 (packed FP32 off, as
// for every other kernel of the library), a process cache and a disk cache of
// code objects, module loading and launches.
#include "esc_cube.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <thread>
#include <random>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <utility>

namespace esc {

namespace {

constexpr uint32_t kCubeMaxGrid = 1u << 22;
constexpr int kFarAhead = 1;   // far sub-cubes in flight ahead of their terms
// local searches per plan: several for complex128 and for complex64 beyond
// 26 sites (N = 24 c128: 291 -> 279 us; N = 32 c64: 66.7 -> 63.5 ms), one for
// complex64 up to 26 sites, whose single-search plans measured 1.5-3 %
// faster (N = 24: 155 against 157 us; N = 20: 13.6 against 14.1 us)
constexpr int kPlanStarts = 10;
int plan_starts(const CubeModel &m) { return (m.c64 && m.n_sites <= 26) ? 1 : kPlanStarts; }

// ---------------------------------------------------------------------------
// Layout of one pass
// ---------------------------------------------------------------------------
// Tile-index bit q holds window site wb[q].  Positions P..P+5 (P = 0, or 1
// for row pairs) are the lanes: lane bit i is site wb[P + i].  Lane bits
// 0..run-1 are the sites P..P+run-1 above the pair bit, so 2^run lanes read
// one run of rows: 64 bytes (run = 2 with 16-byte lane accesses, 3 with 8-byte
// ones) in the free layout, whose other lane sites are any tile sites, or the
// whole wave's 512 or 1024 bytes in the contiguous layout (run = 6).
// The register sites are k of the tile positions (slot bit b of a lane's
// sub-cube = position regpos[b]; with row pairs regpos[0] = 0, site 0, so a
// lane's slots 2y and 2y + 1 are adjacent rows and move as one 16-byte
// access), the others are group positions (bit q of the group index =
// position grppos[q]); far sites wb[TB..w-1] and window-id sites cb are fixed
// per workgroup.
enum Role { kLane, kReg, kGrp, kBase };

struct Layout {
    int N = 0, TB = 0, w = 0, k = 0, P = 0, run = 6;
    std::vector<int> wb, cb, regpos, grppos;
    int role[64];
    int idx[64];     // lane bit (kLane), slot bit (kReg) or group bit (kGrp) of a site

    void finish() {
        for (int s = 0; s < 64; ++s) { role[s] = kBase; idx[s] = -1; }
        for (int i = 0; i < 6; ++i) { role[wb[P + i]] = kLane; idx[wb[P + i]] = i; }
        for (int b = 0; b < (int)regpos.size(); ++b) { role[wb[regpos[b]]] = kReg; idx[wb[regpos[b]]] = b; }
        for (int q = 0; q < (int)grppos.size(); ++q) { role[wb[grppos[q]]] = kGrp; idx[wb[grppos[q]]] = q; }
    }
    // slot x -> tile-index offset and row bits
    uint32_t slot_off(uint32_t x) const {
        uint32_t o = 0;
        for (int b = 0; b < k; ++b)
            if ((x >> b) & 1u) o |= 1u << regpos[b];
        return o;
    }
    uint32_t slot_row(uint32_t x) const {
        uint32_t o = 0;
        for (int b = 0; b < k; ++b)
            if ((x >> b) & 1u) o |= 1u << wb[regpos[b]];
        return o;
    }
};

// The part of a mask that selects the partner sub-cube: lane bits, group
// index bits and far-site row bits.  rho = the register part (slot bits).
struct Outer {
    uint32_t lane = 0, grp = 0, far = 0;
    bool operator<(const Outer &o) const {
        return std::tie(lane, grp, far) < std::tie(o.lane, o.grp, o.far);
    }
    bool operator==(const Outer &o) const { return lane == o.lane && grp == o.grp && far == o.far; }
    bool own() const { return lane == 0 && grp == 0 && far == 0; }
};

// lane sites of the run: all six in the contiguous layout, else a 64-byte run
// (two sites of 16-byte lane accesses, three of 8-byte ones)
int lane_run(const CubeModel &m) { return m.contiguous ? 6 : (m.c64 && !m.paired ? 3 : 2); }

// row bits of the lane bits `lanes` of layout L
uint32_t lane_rows(const Layout &L, uint32_t lanes) {
    uint32_t r = 0;
    for (int i = 0; i < 6; ++i)
        if ((lanes >> i) & 1u) r |= 1u << L.wb[L.P + i];
    return r;
}

void split_mask(const Layout &L, uint32_t mask, Outer &o, uint32_t &rho) {
    o = Outer();
    rho = 0;
    for (int s = 0; s < L.N; ++s) {
        if (!((mask >> s) & 1u)) continue;
        switch (L.role[s]) {
        case kLane: o.lane |= 1u << L.idx[s]; break;
        case kReg: rho |= 1u << L.idx[s]; break;
        case kGrp: o.grp |= 1u << L.idx[s]; break;
        default: o.far |= 1u << s; break;
        }
    }
}

// Layout of a pass's tile sites (wb[0..TB-1], any order on entry): site 0
// below the lanes with row pairs; lanes = sites P..P+run-1 (the run) and
// 6 - run more; the register sites (site 0 and k - 1 others with row pairs);
// the rest are group sites.  `sel` picks the registers among `cand` (the tile
// sites that are neither site 0 with row pairs nor a run site); the lanes take
// the 6 - run lowest of the others (a lane or a group bit costs the same LDS
// load, so that choice does not change the plan's cost, and low lane sites
// keep a wave's accesses close together in memory).
void set_layout(Layout &L, const std::vector<int> &cand, uint32_t sel) {
    std::vector<int> regs, rest;
    for (size_t q = 0; q < cand.size(); ++q) ((sel >> q) & 1u ? regs : rest).push_back(cand[q]);
    std::vector<int> tile;
    if (L.P) tile.push_back(0);
    for (int i = 0; i < L.run; ++i) tile.push_back(L.P + i);
    for (int i = 0; i < 6 - L.run; ++i) tile.push_back(rest[i]);
    std::vector<int> others(regs.begin(), regs.end());
    others.insert(others.end(), rest.begin() + (6 - L.run), rest.end());
    std::sort(others.begin(), others.end());
    tile.insert(tile.end(), others.begin(), others.end());
    for (int q = 0; q < L.TB; ++q) L.wb[q] = tile[q];
    L.regpos.clear();
    L.grppos.clear();
    if (L.P) L.regpos.push_back(0);
    for (int q = 6 + L.P; q < L.TB; ++q)
        (std::find(regs.begin(), regs.end(), L.wb[q]) != regs.end() ? L.regpos : L.grppos).push_back(q);
    L.finish();
}

std::vector<int> layout_candidates(const Layout &L) {
    std::vector<int> cand;
    for (int q = 0; q < L.TB; ++q) {
        const int s = L.wb[q];
        if ((L.P && s == 0) || (s >= L.P && s < L.P + L.run)) continue;
        cand.push_back(s);
    }
    std::sort(cand.begin(), cand.end());
    return cand;
}

// The register choice that leaves the fewest distinct outer parts (far ones
// weighted 3x), then the most terms whose mask lies inside the registers,
// then the lowest sites.
void choose_registers(Layout &L, const std::vector<CubeTerm> &terms) {
    const std::vector<int> cand = layout_candidates(L);
    const int npos = (int)cand.size();
    long best = -1;
    uint32_t best_sel = 0;
    for (uint32_t sel = 0; sel < (1u << npos); ++sel) {
        if (__builtin_popcount(sel) != L.k - L.P) continue;
        Layout t = L;
        set_layout(t, cand, sel);
        std::vector<Outer> parts;
        long inside = 0, cost = 0;
        for (const CubeTerm &ct : terms) {
            Outer o;
            uint32_t rho;
            split_mask(t, ct.mask, o, rho);
            if (o.own()) ++inside;
            if (std::find(parts.begin(), parts.end(), o) == parts.end()) {
                parts.push_back(o);
                cost += o.far ? 3 : 1;   // a far sub-cube goes through L2 and the texture path
            }
        }
        const long key = (100000 - cost) * 1000 + inside;
        if (key > best) { best = key; best_sel = sel; }
    }
    set_layout(L, cand, best_sel);
}

// ---------------------------------------------------------------------------
// Pass planning
// ---------------------------------------------------------------------------
// Cost of one pass: every distinct partner sub-cube is one load per row, a
// far one (from L2, through the texture path) weighted three times a near
// one (from LDS); the register sites are chosen for the pass's terms.
long pass_cost(int N, int P, int run, uint32_t T, uint32_t F, int TB, int k,
               const std::vector<const CubeTerm *> &ts) {
    Layout L;
    L.N = N;
    L.TB = TB;
    L.k = k;
    L.P = P;
    L.run = run;
    L.wb.clear();
    for (int s = 0; s < N; ++s)
        if ((T >> s) & 1u) L.wb.push_back(s);
    for (int s = 0; s < N; ++s)
        if ((F >> s) & 1u) L.wb.push_back(s);
    L.w = (int)L.wb.size();
    std::vector<CubeTerm> terms;
    for (const CubeTerm *t : ts) terms.push_back(*t);
    choose_registers(L, terms);
    std::vector<Outer> parts;
    long cost = 0;
    for (const CubeTerm &ct : terms) {
        Outer o;
        uint32_t rho;
        split_mask(L, ct.mask, o, rho);
        if (std::find(parts.begin(), parts.end(), o) == parts.end()) {
            parts.push_back(o);
            cost += o.own() ? 0 : (o.far ? 3 : 1);
        }
    }
    return cost;
}

struct PlanEval {
    long cost = -1;                       // -1: some term has no pass
    std::vector<std::vector<int>> assign;  // term indices per pass
    std::vector<uint32_t> far;             // far sites per pass
};

// place every term given the tiles: near where possible (the pass whose
// current terms already touch the most of its sites), else far in the pass
// that needs the fewest new far sites (at most max_far per pass); then the
// summed pass costs
PlanEval evaluate(int N, int pair, int run, int TB, int k, int max_far, const std::vector<uint32_t> &T,
                  const std::vector<CubeTerm> &terms) {
    PlanEval ev;
    const size_t P = T.size();
    ev.assign.assign(P, {});
    ev.far.assign(P, 0u);
    std::vector<uint32_t> touched(P, 0u);
    std::vector<int> farq;
    for (size_t t = 0; t < terms.size(); ++t) {
        int best = -1, bk = -1;
        for (size_t p = 0; p < P; ++p) {
            if (terms[t].mask & ~T[p]) continue;
            const int key = __builtin_popcount(terms[t].mask & touched[p]);
            if (key > bk) { bk = key; best = (int)p; }
        }
        if (best < 0) { farq.push_back((int)t); continue; }
        ev.assign[best].push_back((int)t);
        touched[best] |= terms[t].mask;
    }
    // far terms: repeatedly place the (term, pass) pair that adds the fewest
    // far sites to its pass (ties: the pass with fewer terms)
    while (!farq.empty()) {
        int bt = -1, bp = -1, bk = 1 << 30;
        for (size_t q = 0; q < farq.size(); ++q) {
            const int t = farq[q];
            for (size_t p = 0; p < P; ++p) {
                const uint32_t nf = ev.far[p] | (terms[t].mask & ~T[p]);
                if (__builtin_popcount(nf) > max_far) continue;
                const int key = __builtin_popcount(nf & ~ev.far[p]) * 1024 + (int)ev.assign[p].size();
                if (key < bk) { bk = key; bt = (int)q; bp = (int)p; }
            }
        }
        if (bt < 0) return ev;
        const int t = farq[bt];
        ev.far[bp] |= terms[t].mask & ~T[bp];
        ev.assign[bp].push_back(t);
        farq.erase(farq.begin() + bt);
    }
    long cost = 0;
    for (size_t p = 0; p < P; ++p) {
        std::vector<const CubeTerm *> ts;
        for (int t : ev.assign[p]) ts.push_back(&terms[t]);
        cost += pass_cost(N, pair, run, T[p], ev.far[p], TB, k, ts);
    }
    ev.cost = cost;
    return ev;
}

// ---------------------------------------------------------------------------
// Boolean expressions of row bits
// ---------------------------------------------------------------------------
struct BE {
    bool is_const = false;
    int c = 0;
    std::string s;
};

BE be_const(int v) { BE b; b.is_const = true; b.c = v ? 1 : 0; b.s = v ? "true" : "false"; return b; }
BE be_var(const std::string &s) { BE b; b.s = s; return b; }

// bit of `site` in the row of slot x
BE site_bit(const Layout &L, int site, uint32_t x) {
    switch (L.role[site]) {
    case kLane: return be_var("L" + std::to_string(site));
    case kReg: return be_const((int)((x >> L.idx[site]) & 1u));
    case kGrp: return be_var("G" + std::to_string(L.idx[site]));
    default: return be_var("U" + std::to_string(site));
    }
}

BE be_eq(const BE &a, const BE &b) {
    if (a.is_const && b.is_const) return be_const(a.c == b.c);
    if (a.is_const) return a.c ? b : be_var("(!" + b.s + ")");
    if (b.is_const) return b.c ? a : be_var("(!" + a.s + ")");
    return be_var("(" + a.s + " == " + b.s + ")");
}

// a value: either the literal zero or an expression
struct Val {
    bool zero = false;
    std::string s;
};
Val val(const std::string &s) { Val v; v.s = s; return v; }
Val val_zero() { Val v; v.zero = true; v.s = "0"; return v; }

Val sel(const BE &c, const Val &a, const Val &b) {
    if (c.is_const) return c.c ? a : b;
    if (a.zero && b.zero) return val_zero();
    return val("(" + c.s + " ? " + a.s + " : " + b.s + ")");
}
Val neg(const Val &a) { return a.zero ? a : val("(-" + a.s + ")"); }

// coefficient of a term at slot x (real, imaginary)
void coefficient(const CubeModel &m, const Layout &L, const CubeTerm &t, uint32_t x, Val &cr, Val &ci) {
    if (t.kind == 0) {
        const BE bi = site_bit(L, t.i, x), bj = site_bit(L, t.j, x);
        const BE eq = be_eq(bi, bj);
        const std::string f = std::to_string(t.f);
        cr = sel(eq, val("A" + f), val("B" + f));
        if (m.bond_cplx[t.f]) {
            const Val q = val("Q" + f);
            ci = sel(eq, sel(bi, neg(q), q), val_zero());
        } else {
            ci = val_zero();
        }
    } else {
        cr = (m.flip_partners && t.yp >= 0) ? sel(site_bit(L, t.yp, x), val("Rm"), val("Rp")) : val("Rp");
        if (m.flip_cplx) {
            const Val mag = (m.flip_partners && t.xp >= 0) ? sel(site_bit(L, t.xp, x), val("Im"), val("Ip")) : val("Ip");
            ci = sel(site_bit(L, t.k, x), mag, neg(mag));
        } else {
            ci = val_zero();
        }
    }
}

std::string hex32(uint32_t v) {
    char b[16];
    std::snprintf(b, sizeof(b), "0x%xu", v);
    return b;
}

// scatter of the low bits of `var` onto the given row bits
std::string scatter(const std::string &var, const std::vector<int> &bits, int from, int count) {
    std::ostringstream o;
    o << "0u";
    for (int q = 0; q < count; ++q) o << " | (((" << var << " >> " << q << ") & 1u) << " << bits[from + q] << ")";
    return o.str();
}

// Diagonal of a lane's 2^k rows, added to the accumulators as d(r) psi(r)
// (after the previous passes' output when the pass is not the first).  The
// rows are R0 | X with X the slot's row bits, so for a bond-distance group
// (r ^ (r >> d)) & m = A ^ C with A = (R0 ^ (R0 >> d)) & m formed once per
// lane and C = (X ^ (X >> d)) & m a constant of the slot; groups whose masks
// are disjoint share one word, and popc(r) = popc(R0) + popc(X).  The counts
// are the closed form's integers, so d(r) is the same double.
void emit_diag(std::ostringstream &o, const CubeModel &m, const Layout &L, int S, bool first) {
    o << "    const unsigned R0 = base | grw | lrow;\n";
    auto acc = [&](int x, const std::string &d) {
        const std::string X = std::to_string(x);
        if (first)
            o << "      ar[" << X << "] = " << d << " * p[" << X << "].x; ai[" << X << "] = " << d << " * p[" << X << "].y;\n";
        else
            o << "      ar[" << X << "] = fma(" << d << ", p[" << X << "].x, ar[" << X << "]); ai[" << X << "] = fma(" << d
              << ", p[" << X << "].y, ai[" << X << "]);\n";
    };
    if (m.seq_diag) {
        for (int x = 0; x < S; ++x) {
            o << "    { const R d = (R)esc_diag(R0 | " << hex32(L.slot_row((uint32_t)x)) << ", cf[13], cf[14], cf[15]);\n";
            acc(x, "d");
            o << "    }\n";
        }
        return;
    }
    // pack groups with disjoint masks into words
    struct Word { uint32_t mask = 0; std::vector<std::pair<uint32_t, uint32_t>> grp; };
    auto pack = [](const uint32_t *gd, const uint32_t *gm, uint32_t ng) {
        std::vector<Word> w;
        for (uint32_t q = 0; q < ng; ++q) {
            size_t i = 0;
            while (i < w.size() && (w[i].mask & gm[q])) ++i;
            if (i == w.size()) w.emplace_back();
            w[i].mask |= gm[q];
            w[i].grp.push_back({gd[q], gm[q]});
        }
        return w;
    };
    const std::vector<Word> zw = pack(m.zgrp_d, m.zgrp_m, m.n_zgrp);
    const std::vector<Word> jw = m.nj ? pack(m.jgrp_d, m.jgrp_m, m.n_jgrp) : std::vector<Word>();
    auto word_expr = [](const std::string &r, const Word &w) {
        std::string e = "0u";
        for (const auto &g : w.grp)
            e += " | ((" + r + " ^ (" + r + " >> " + std::to_string(g.first) + ")) & " + hex32(g.second) + ")";
        return e;
    };
    auto word_const = [](uint32_t X, const Word &w) {
        uint32_t c = 0;
        for (const auto &g : w.grp) c |= (X ^ (X >> g.first)) & g.second;
        return c;
    };
    for (size_t i = 0; i < zw.size(); ++i) o << "    const unsigned zw" << i << " = " << word_expr("R0", zw[i]) << ";\n";
    for (size_t i = 0; i < jw.size(); ++i) o << "    const unsigned jw" << i << " = " << word_expr("R0", jw[i]) << ";\n";
    o << "    const unsigned pc0 = __popc(R0);\n";
    for (int x = 0; x < S; ++x) {
        const uint32_t X = L.slot_row((uint32_t)x);
        std::string cz = "0u", cj = "0u";
        for (size_t i = 0; i < zw.size(); ++i) cz += " + __popc(zw" + std::to_string(i) + " ^ " + hex32(word_const(X, zw[i])) + ")";
        for (size_t i = 0; i < jw.size(); ++i) cj += " + __popc(jw" + std::to_string(i) + " ^ " + hex32(word_const(X, jw[i])) + ")";
        o << "    { const R d = (R)esc_diag_c(" << cz << ", " << cj << ", pc0 + " << __builtin_popcount(X)
          << "u, cf[13], cf[14], cf[15]);\n";
        acc(x, "d");
        o << "    }\n";
    }
}

void emit_pass(std::ostringstream &o, const CubeModel &m, const CubePass &p, int index) {
    Layout L;
    L.N = m.n_sites;
    L.TB = p.tile_bits;
    L.w = (int)p.wb.size();
    L.k = m.reg_sites;
    L.wb = p.wb;
    L.cb = p.cb;
    L.P = m.paired ? 1 : 0;
    L.run = lane_run(m);
    choose_registers(L, p.terms);
    const int TB = L.TB, k = L.k, S = 1 << k, WPB = m.waves, P = L.P;
    const int ngb = TB - 6 - k;                 // group index bits
    const int ngrp_stage = 1 << (TB - 6 - P);   // groups of 64 lane accesses (2^P rows each) of the tile
    const int nsub = 1 << ngb;                  // sub-cubes per tile
    const int NG = nsub / WPB;                  // sub-cubes per wave
    const int nwin = 1 << (int)L.cb.size();
    const int bpw = 1 << (L.w - TB);
    const bool xcdw = nwin % 8 == 0;

    // group the terms by outer part: the own sub-cube first
    std::map<Outer, std::vector<std::pair<CubeTerm, uint32_t>>> groups;
    for (const CubeTerm &t : p.terms) {
        Outer ou;
        uint32_t rho;
        split_mask(L, t.mask, ou, rho);
        groups[ou].push_back({t, rho});
    }

    o << "// pass " << index << ": window";
    for (int s : L.wb) o << " " << s;
    o << "; register sites";
    for (int b = 0; b < k; ++b) o << " " << L.wb[L.regpos[b]];
    int nfar = 0, nown = 0;
    for (const auto &g : groups) {
        if (g.first.far) ++nfar;
        if (g.first.own()) nown = (int)g.second.size();
    }
    o << "; " << p.terms.size() << " terms (" << nown << " in registers) in " << groups.size()
      << " partner sub-cubes (" << nfar << " far)\n";
    for (const auto &g : groups) {
        o << "//   sub-cube lane " << hex32(g.first.lane) << " group " << hex32(g.first.grp) << " far "
          << hex32(g.first.far) << ":";
        for (const auto &pr : g.second)
            o << (pr.first.kind == 0 ? " b" + std::to_string(pr.first.i) + "-" + std::to_string(pr.first.j)
                                     : " f" + std::to_string(pr.first.k));
        o << "\n";
    }
    // the pass body (b = tile block index); the kernels below call it for one
    // system, or for system blockIdx.y of a batch of systems that share this
    // lattice and term structure (own couplings, vectors and partials each)
    o << "__device__ __forceinline__ void esc_cube_body" << index
      << "(const double *__restrict__ cf, const V2 *__restrict__ in, V2 *__restrict__ out, "
         "double *__restrict__ dot_part, const unsigned b) {\n";
    o << "  __shared__ __attribute__((aligned(16))) V2 tile[" << (1 << TB) << "];\n";
    o << "  __shared__ double red[2][" << WPB << "];\n";
    o << "  const unsigned lane = threadIdx.x & 63u;\n";
    // lane bits in tile-index units (positions P..P+5) and in row units
    o << "  const unsigned ln = lane << " << P << "u;\n";
    o << "  const unsigned lrow = " << scatter("lane", L.wb, P, 6) << ";\n";
    if (P) o << "  F4 *tile4 = reinterpret_cast<F4 *>(tile); (void)tile4;\n";
    o << "  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);\n";
    if (xcdw)
        o << "  const unsigned seq = b / 8u, wi = (seq / " << bpw << "u) * 8u + b % 8u, tl = seq % " << bpw << "u;\n";
    else
        o << "  const unsigned wi = b / " << bpw << "u, tl = b % " << bpw << "u;\n";
    o << "  const unsigned base = (" << scatter("wi", L.cb, 0, (int)L.cb.size()) << ") | ("
      << scatter("tl", L.wb, TB, L.w - TB) << ");\n";
    // staging: 64-row groups g = wave + WPB * i
    o << "#pragma unroll\n";
    o << "  for (unsigned i = 0; i < " << ngrp_stage / WPB << "u; ++i) {\n";
    o << "    const unsigned g = wave + " << WPB << "u * i;\n";
    if (P)
        o << "    tile4[(g << 6) | lane] = *reinterpret_cast<const F4 *>(in + (base | ("
          << scatter("g", L.wb, 6 + P, TB - 6 - P) << ") | lrow));\n";
    else
        o << "    tile[(g << 6) | lane] = in[base | (" << scatter("g", L.wb, 6, TB - 6) << ") | lrow];\n";
    o << "  }\n";
    o << "  __syncthreads();\n";
    o << "  const R A0 = (R)cf[0], B0 = (R)cf[1], Q0 = (R)cf[2], A1 = (R)cf[3], B1 = (R)cf[4], Q1 = (R)cf[5];\n";
    o << "  const R A2 = (R)cf[6], B2 = (R)cf[7], Q2 = (R)cf[8], Rp = (R)cf[9], Rm = (R)cf[10], Ip = (R)cf[11], "
         "Im = (R)cf[12];\n";
    o << "  (void)A0; (void)B0; (void)Q0; (void)A1; (void)B1; (void)Q1; (void)A2; (void)B2; (void)Q2;\n";
    o << "  (void)Rp; (void)Rm; (void)Ip; (void)Im;\n";
    for (int i = 0; i < 6; ++i)
        o << "  const bool L" << L.wb[P + i] << " = (lane >> " << i << ") & 1u; (void)L" << L.wb[P + i] << ";\n";
    for (int s = 0; s < L.N; ++s)
        if (L.role[s] == kBase)
            o << "  const bool U" << s << " = (base >> " << s << ") & 1u; (void)U" << s << ";\n";
    o << "  double dr = 0.0, di = 0.0;\n";
    o << "#pragma unroll 1\n";
    o << "  for (unsigned it = 0; it < " << NG << "u; ++it) {\n";
    o << "    const unsigned gi = wave + " << WPB << "u * it;\n";
    {
        // group index -> tile-index offset and row bits
        std::vector<int> gp(L.grppos), gr;
        for (int q : L.grppos) gr.push_back(L.wb[q]);
        o << "    const unsigned gix = " << scatter("gi", gp, 0, ngb) << ";\n";
        o << "    const unsigned grw = " << scatter("gi", gr, 0, ngb) << ";\n";
        for (int q = 0; q < ngb; ++q) o << "    const bool G" << q << " = (gi >> " << q << ") & 1u; (void)G" << q << ";\n";
    }
    // a sub-cube from the tile (tile index va of slot 0) or from memory (row
    // src + slot row + lo); with row pairs slots 2y, 2y + 1 are one 16-byte access
    auto lds_cube = [&](const std::string &name, const std::string &va) {
        for (int y = 0; y < S; y += 1 + P) {
            if (P)
                o << "      { const F4 t = tile4[(" << va << " + " << hex32(L.slot_off(y)) << ") >> 1]; " << name << "["
                  << y << "] = make_float2(t.x, t.y); " << name << "[" << y + 1 << "] = make_float2(t.z, t.w); }\n";
            else
                o << "      " << name << "[" << y << "] = tile[" << va << " + " << hex32(L.slot_off(y)) << "];\n";
        }
    };
    auto glb_cube = [&](const std::string &name, const std::string &src, const std::string &lo) {
        for (int y = 0; y < S; y += 1 + P) {
            if (P)
                o << "      { const F4 t = *reinterpret_cast<const F4 *>(" << src << " + " << hex32(L.slot_row(y)) << " + "
                  << lo << "); " << name << "[" << y << "] = make_float2(t.x, t.y); " << name << "[" << y + 1
                  << "] = make_float2(t.z, t.w); }\n";
            else
                o << "      " << name << "[" << y << "] = (" << src << " + " << hex32(L.slot_row(y)) << ")[" << lo << "];\n";
        }
    };
    o << "    R ar[" << S << "], ai[" << S << "];\n";
    o << "    V2 p[" << S << "];\n";
    o << "    {\n      const unsigned va = ln | gix;\n";
    lds_cube("p", "va");
    o << "    }\n";
    // accumulators: zero in the first pass; a later pass starts from the
    // previous passes' output (streamed loads, in flight while the first
    // terms run).  The pass that holds the diagonal adds d(r) psi(r) first.
    for (int x = 0; x < S; x += 1 + P) {
        if (p.first) {
            for (int y = x; y <= x + P; ++y) o << "    ar[" << y << "] = 0; ai[" << y << "] = 0;\n";
        } else if (P) {
            o << "    { const F4 v = __builtin_nontemporal_load(reinterpret_cast<const F4 *>(out + (base | grw | "
              << hex32(L.slot_row(x)) << " | lrow))); ar[" << x << "] = v.x; ai[" << x << "] = v.y; ar[" << x + 1
              << "] = v.z; ai[" << x + 1 << "] = v.w; }\n";
        } else {
            o << "    { const NV2 v = __builtin_nontemporal_load(reinterpret_cast<const NV2 *>(out + (base | grw | "
              << hex32(L.slot_row(x)) << " | lrow))); ar[" << x << "] = v.x; ai[" << x << "] = v.y; }\n";
        }
    }
    if (p.diag) emit_diag(o, m, L, S, p.first);
    auto emit_terms = [&](const std::vector<std::pair<CubeTerm, uint32_t>> &ts, const std::string &src) {
        for (const auto &pr : ts) {
            const CubeTerm &t = pr.first;
            const uint32_t rho = pr.second;
            o << "    // " << (t.kind == 0 ? "bond " + std::to_string(t.i) + "-" + std::to_string(t.j) + " family " +
                                                 std::to_string(t.f)
                                           : "flip " + std::to_string(t.k))
              << "\n";
            for (int x = 0; x < S; ++x) {
                Val cr, ci;
                coefficient(m, L, t, (uint32_t)x, cr, ci);
                const std::string q = src + "[" + std::to_string(x ^ (int)rho) + "]";
                const std::string X = std::to_string(x);
                o << "    ar[" << X << "] = fma(" << cr.s << ", " << q << ".x, ar[" << X << "]);";
                if (!ci.zero) o << " ar[" << X << "] = fma(-" << ci.s << ", " << q << ".y, ar[" << X << "]);";
                o << " ai[" << X << "] = fma(" << cr.s << ", " << q << ".y, ai[" << X << "]);";
                if (!ci.zero) o << " ai[" << X << "] = fma(" << ci.s << ", " << q << ".x, ai[" << X << "]);";
                o << "\n";
            }
        }
    };
    // partner sub-cube loads: near ones from the tile, far ones from the
    // window (L2), each into its own registers
    auto emit_load = [&](const Outer &ou, const std::string &name) {
        o << "    V2 " << name << "[" << S << "];\n";
        uint32_t gx = 0, grr = 0;
        for (int qq = 0; qq < ngb; ++qq)
            if ((ou.grp >> qq) & 1u) { gx |= 1u << L.grppos[qq]; grr |= 1u << L.wb[L.grppos[qq]]; }
        if (ou.far == 0) {
            o << "    { const unsigned va = (ln ^ " << hex32(ou.lane << P) << ") | (gix ^ " << hex32(gx) << ");\n";
            lds_cube(name, "va");
            o << "    }\n";
        } else {
            o << "    { const unsigned lo = lrow ^ " << hex32(lane_rows(L, ou.lane)) << ";\n";
            o << "      const V2 *src = in + ((base ^ " << hex32(ou.far) << ") | (grw ^ " << hex32(grr) << "));\n";
            glb_cube(name, "src", "lo");
            o << "    }\n";
        }
    };
    // order: far sub-cubes are issued kFarAhead near sub-cubes before their
    // terms run, so their L2 latency overlaps the LDS work
    std::vector<const std::pair<const Outer, std::vector<std::pair<CubeTerm, uint32_t>>> *> nearg, farg;
    for (const auto &g : groups) (g.first.far ? farg : nearg).push_back(&g);
    const int ahead = std::min<int>(kFarAhead, (int)farg.size());
    size_t next_far = 0, done_far = 0;
    std::vector<size_t> pending;
    auto issue_far = [&] {
        if (next_far < farg.size()) {
            emit_load(farg[next_far]->first, "fq" + std::to_string(next_far));
            pending.push_back(next_far++);
        }
    };
    auto retire_far = [&] {
        const size_t j = pending.front();
        pending.erase(pending.begin());
        emit_terms(farg[j]->second, "fq" + std::to_string(j));
        ++done_far;
        issue_far();
    };
    for (int j = 0; j < ahead; ++j) issue_far();
    for (size_t i = 0; i < nearg.size(); ++i) {
        const auto &g = *nearg[i];
        if (g.first.own()) {
            emit_terms(g.second, "p");
        } else {
            const std::string name = "nq" + std::to_string(i);
            emit_load(g.first, name);
            emit_terms(g.second, name);
        }
        // spread the far sub-cubes evenly over the near ones
        while (!pending.empty() && (done_far + 1) * nearg.size() <= (i + 1) * farg.size()) retire_far();
    }
    while (!pending.empty()) retire_far();
    // store (streamed) and the tile's share of <in|out> from the staged rows
    for (int x = 0; x < S; x += 1 + P) {
        if (P)
            o << "    { F4 v; v.x = ar[" << x << "]; v.y = ai[" << x << "]; v.z = ar[" << x + 1 << "]; v.w = ai[" << x + 1
              << "]; __builtin_nontemporal_store(v, reinterpret_cast<F4 *>(out + (base | grw | " << hex32(L.slot_row(x))
              << " | lrow))); }\n";
        else
            o << "    { NV2 v = {ar[" << x << "], ai[" << x << "]}; __builtin_nontemporal_store(v, reinterpret_cast<NV2 *>(out + "
              << "(base | grw | " << hex32(L.slot_row(x)) << " | lrow))); }\n";
    }
    if (p.last) {
        o << "    if (dot_part) {\n      const unsigned va = ln | gix;\n      V2 q[" << S << "];\n";
        lds_cube("q", "va");
        for (int x = 0; x < S; ++x)
            o << "      { const double xr = q[" << x << "].x, xi = q[" << x << "].y, yr = ar[" << x << "], yi = ai[" << x
              << "]; dr += xr * yr + xi * yi; di += xr * yi - xi * yr; }\n";
        o << "    }\n";
    }
    o << "  }\n";
    if (p.last) {
        o << "  if (dot_part) {\n";
        o << "#pragma unroll\n    for (int off = 32; off > 0; off >>= 1) { dr += __shfl_xor(dr, off); di += __shfl_xor(di, off); }\n";
        o << "    if (lane == 0) { red[0][wave] = dr; red[1][wave] = di; }\n";
        o << "    __syncthreads();\n";
        o << "    if (threadIdx.x == 0) {\n      double x = 0.0, y = 0.0;\n";
        o << "      for (int i = 0; i < " << WPB << "; ++i) { x += red[0][i]; y += red[1][i]; }\n";
        o << "      dot_part[2 * (size_t)b] = x; dot_part[2 * (size_t)b + 1] = y;\n    }\n  }\n";
    }
    o << "}\n\n";
    // the first of several passes (no previous output, no diagonal) is
    // compiled for four waves per SIMD, two 8-wave workgroups per CU (N = 24
    // c128: 137 -> 128 VGPRs, the pass 118 -> 106 us); the last pass keeps
    // its registers (capped it measured slower)
    const std::string cap = (p.first && !p.last) ? " __attribute__((amdgpu_waves_per_eu(4)))" : "";
    o << "extern \"C\" __global__ void __launch_bounds__(" << WPB * 64 << ")" << cap << " esc_cube_p" << index
      << "(const double *__restrict__ cf, const V2 *__restrict__ in, V2 *__restrict__ out, "
         "double *__restrict__ dot_part, unsigned block_base) {\n";
    o << "  esc_cube_body" << index << "(cf, in, out, dot_part, blockIdx.x + block_base);\n}\n\n";
    o << "extern \"C\" __global__ void __launch_bounds__(" << WPB * 64 << ")" << cap << " esc_cube_b" << index
      << "(const double *__restrict__ cfs, const V2 *const *__restrict__ ins, V2 *const *__restrict__ outs, "
         "double *const *__restrict__ dots, unsigned block_base) {\n";
    o << "  const unsigned inst = blockIdx.y;\n";
    o << "  esc_cube_body" << index << "(cfs + " << (int)kCubeScalars
      << "u * inst, ins[inst], outs[inst], dots[inst], blockIdx.x + block_base);\n}\n\n";
}

}  // namespace

bool cube_plan(const CubeModel &m, const std::vector<CubeTerm> &terms, int tile_bits, std::vector<uint32_t> tiles,
               int max_far, std::vector<CubePass> &passes) {
    const int N = m.n_sites, TB = tile_bits, k = m.reg_sites, P = m.paired ? 1 : 0, run = lane_run(m);
    const int nstarts = plan_starts(m);
    // every tile holds the sites below and of the lanes' run; a 64-byte run
    // also the next site, so its 128-byte line lies in the tile (sites 0..2
    // with 16-byte rows, else 0..3); a seed tile that gains one of them drops
    // its highest other sites
    const uint32_t fixed = run == 6 ? (1u << (6 + P)) - 1u : (1u << (P + run + 1)) - 1u;
    if (N < 6 + P || TB < 6 + P) return false;
    for (uint32_t &t : tiles) {
        t |= fixed;
        for (int s = N - 1; s >= 0 && __builtin_popcount(t) > TB; --s)
            if (!((fixed >> s) & 1u)) t &= ~(1u << s);
    }
    // plans are cached per lattice, term set, shape and seed tiles
    std::string key = std::to_string(N) + ":" + std::to_string(TB) + ":" + std::to_string(k) + ":" +
                      std::to_string(max_far) + ":" + std::to_string(P) + ":" + std::to_string(run) + ":" +
                      std::to_string(nstarts);
    for (uint32_t t : tiles) key += ":" + std::to_string(t);
    for (const CubeTerm &t : terms)
        key += "," + std::to_string(t.mask) + "/" + std::to_string(t.kind) + "/" + std::to_string(t.f) + "/" +
               std::to_string(t.xp) + "/" + std::to_string(t.yp);
    static std::mutex mu;
    static std::map<std::string, std::vector<CubePass>> cache;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) { passes = it->second; return !passes.empty(); }
    }
    // local search from several site orders (start 0: ascending sites, the
    // others fixed shuffles), each on its own thread; the plan of lowest cost
    // is kept (ties: the earliest start).  A single greedy search stops in
    // local minima that depend on the order it tries swaps in.
    struct Start {
        std::vector<uint32_t> tiles;
        PlanEval ev;
    };
    std::vector<Start> starts(nstarts);
    auto search = [&](int si) {
        std::vector<int> ord(N);
        for (int s = 0; s < N; ++s) ord[s] = s;
        if (si > 0) {
            std::mt19937 g((unsigned)si);
            std::shuffle(ord.begin(), ord.end(), g);
        }
        std::vector<uint32_t> t = tiles;
        PlanEval cur = evaluate(N, P, run, TB, k, max_far, t, terms);
        if (cur.cost < 0) { starts[si].ev = cur; return; }
        // hill climbing: swap one tile site of one pass for a site outside it
        for (bool gain = true; gain;) {
            gain = false;
            for (size_t p = 0; p < t.size(); ++p) {
                for (int ai = 0; ai < N; ++ai) {
                    const int a = ord[ai];
                    if (!((t[p] >> a) & 1u) || ((fixed >> a) & 1u)) continue;
                    for (int bi = 0; bi < N; ++bi) {
                        const int b = si > 0 ? ord[(bi + ai) % N] : bi;
                        if ((t[p] >> b) & 1u) continue;
                        std::vector<uint32_t> t2(t);
                        t2[p] = (t2[p] & ~(1u << a)) | (1u << b);
                        const PlanEval ev = evaluate(N, P, run, TB, k, max_far, t2, terms);
                        if (ev.cost >= 0 && ev.cost < cur.cost) {
                            t = t2;
                            cur = ev;
                            gain = true;
                            break;
                        }
                    }
                }
            }
        }
        starts[si].tiles = t;
        starts[si].ev = cur;
    };
    {
        std::vector<std::thread> th;
        for (int si = 1; si < nstarts; ++si) th.emplace_back(search, si);
        search(0);
        for (auto &x : th) x.join();
    }
    int best = -1;
    for (int si = 0; si < nstarts; ++si)
        if (starts[si].ev.cost >= 0 && (best < 0 || starts[si].ev.cost < starts[best].ev.cost)) best = si;
    if (best < 0) return false;
    tiles = starts[best].tiles;
    const PlanEval cur = starts[best].ev;
    passes.clear();
    for (size_t p = 0; p < tiles.size(); ++p) {
        if (cur.assign[p].empty()) continue;
        CubePass cp;
        for (int s = 0; s < N; ++s)
            if ((tiles[p] >> s) & 1u) cp.wb.push_back(s);
        for (int s = 0; s < N; ++s)
            if ((cur.far[p] >> s) & 1u) cp.wb.push_back(s);
        for (int s = 0; s < N; ++s)
            if (!((tiles[p] | cur.far[p]) >> s & 1u)) cp.cb.push_back(s);
        cp.tile_bits = TB;
        for (int t : cur.assign[p]) cp.terms.push_back(terms[t]);
        passes.push_back(std::move(cp));
    }
    for (size_t i = 0; i < passes.size(); ++i) {
        passes[i].first = i == 0;
        passes[i].last = i + 1 == passes.size();
        passes[i].diag = passes[i].last;
    }
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = passes;
    return !passes.empty();
}

std::string cube_source(const CubeModel &m, const std::vector<CubePass> &passes) {
    std::ostringstream o;
    o << "// generated H.psi passes: " << m.n_sites << " sites, " << (m.c64 ? "complex64" : "complex128")
      << ", 2^" << m.reg_sites << " rows per lane, " << m.waves << " waves per workgroup\n";
    o << "typedef " << (m.c64 ? "float" : "double") << " R;\n";
    o << "typedef " << (m.c64 ? "float2" : "double2") << " V2;\n";
    o << "typedef R NV2 __attribute__((ext_vector_type(2)));\n";
    if (m.paired) o << "typedef float F4 __attribute__((ext_vector_type(4)));\n";
    // diagonal: the closed form where every partial sum of the Kitaev-z part is
    // exact (it then equals the reference's bond-by-bond sum), else that sum
    // itself, z bonds then Heisenberg bonds in the reference's order
    if (m.seq_diag) {
        o << "__device__ __forceinline__ double esc_diag(unsigned r, double kz4, double j4, double bz) {\n";
        o << "#pragma clang fp contract(off)\n";
        o << "  double d = 0.0;\n";
        for (uint32_t q = 0; q < m.n_zpair; ++q)
            o << "  d += (((r >> " << (m.zpair[q] & 0xffu) << ") ^ (r >> " << ((m.zpair[q] >> 8) & 0xffu)
              << ")) & 1u) ? -kz4 : kz4;\n";
        for (uint32_t q = 0; q < m.n_jpair; ++q)
            o << "  d += (((r >> " << (m.jpair[q] & 0xffu) << ") ^ (r >> " << ((m.jpair[q] >> 8) & 0xffu)
              << ")) & 1u) ? -j4 : j4;\n";
        o << "  d = d + bz * (0.5 * (double)(" << m.n_sites << " - 2 * (int)__popc(r)));\n";
        o << "  (void)j4;\n  return d;\n}\n\n";
        for (size_t i = 0; i < passes.size(); ++i) emit_pass(o, m, passes[i], (int)i);
        return o.str();
    }
    // the closed form from the three counts: anti-aligned z bonds, anti-aligned
    // Heisenberg bonds, down spins (emit_diag forms them per row)
    o << "__device__ __forceinline__ double esc_diag_c(unsigned cz, unsigned cj, unsigned pc, double kz4, double j4, "
         "double bz) {\n";
    o << "#pragma clang fp contract(off)\n";
    o << "  double d = kz4 * (double)(" << (int)m.nz << " - 2 * (int)cz);\n";
    if (m.nj) o << "  d = d + j4 * (double)(" << (int)m.nj << " - 2 * (int)cj);\n";
    o << "  d = d + bz * (0.5 * (double)(" << m.n_sites << " - 2 * (int)pc));\n";
    o << "  (void)cj; (void)j4;\n  return d;\n}\n\n";
    for (size_t i = 0; i < passes.size(); ++i) emit_pass(o, m, passes[i], (int)i);
    return o.str();
}

// ---------------------------------------------------------------------------
// Compilation and caches
// ---------------------------------------------------------------------------
struct CubeProgram {
    // the code object stays alive as long as its module: the runtime and
    // tools attached to it (a profiler's code-object tracing) may read the
    // image passed to hipModuleLoadData after the call returns
    std::vector<char> image;
    hipModule_t module = nullptr;
    std::vector<hipFunction_t> fn, fnb;   // per pass: one system, a batch of systems
    int tile_bits = 0, threads = 0;
    ~CubeProgram() {
        if (module) (void)hipModuleUnload(module);
    }
};

namespace {

std::mutex g_mu;
std::map<std::string, std::vector<char>> g_code;                        // hash -> code object
std::map<std::pair<std::string, int>, std::weak_ptr<CubeProgram>> g_prog;  // (hash, device) -> program

const char *const kOpts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                             "-packed-fp32-ops"};

std::string key_of(const std::string &src) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const std::string &s) {
        for (unsigned char c : s) { h ^= c; h *= 1099511628211ull; }
    };
    mix(src);
    for (const char *opt : kOpts) mix(opt);
    int maj = 0, min = 0;
    if (hiprtcVersion(&maj, &min) == HIPRTC_SUCCESS) mix(std::to_string(maj) + "." + std::to_string(min));
    char b[32];
    std::snprintf(b, sizeof(b), "%016llx", (unsigned long long)h);
    return b;
}

// cache directory: ESC_JIT_CACHE, else jit_cache/ beside the library
std::string cache_dir() {
    if (const char *e = std::getenv("ESC_JIT_CACHE")) return e;
    Dl_info info;
    if (dladdr((const void *)&cube_source, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t sl = p.rfind('/');
        return (sl == std::string::npos ? std::string(".") : p.substr(0, sl)) + "/jit_cache";
    }
    return "";
}

bool read_file(const std::string &path, std::vector<char> &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return !out.empty();
}

void write_file(const std::string &dir, const std::string &name, const std::vector<char> &data) {
    if (dir.empty()) return;
    (void)mkdir(dir.c_str(), 0755);
    const std::string tmp = dir + "/" + name + ".tmp" + std::to_string((long)getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return;
        f.write(data.data(), (std::streamsize)data.size());
        if (!f) { std::remove(tmp.c_str()); return; }
    }
    if (std::rename(tmp.c_str(), (dir + "/" + name).c_str()) != 0) std::remove(tmp.c_str());
}

// code object of a source: process cache, disk cache, then hiprtc
bool code_for(const std::string &src, const std::string &key, std::vector<char> &code, std::string *err) {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_code.find(key);
        if (it != g_code.end()) { code = it->second; return true; }
    }
    const std::string dir = cache_dir();
    if (!dir.empty() && read_file(dir + "/" + key + ".co", code)) {
        std::lock_guard<std::mutex> lk(g_mu);
        g_code[key] = code;
        return true;
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "esc_cube.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        if (err) *err = "hiprtcCreateProgram failed";
        return false;
    }
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(kOpts) / sizeof(kOpts[0])), kOpts);
    if (rc != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        if (err) *err = std::string("hiprtc: ") + hiprtcGetErrorString(rc) + "\n" + log.substr(0, 4000);
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    code.assign(n, 0);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    write_file(dir, key + ".co", code);
    std::lock_guard<std::mutex> lk(g_mu);
    g_code[key] = code;
    return true;
}

bool check_model(const CubeModel &m, const std::vector<CubePass> &passes, std::string *err) {
    for (const CubePass &p : passes) {
        const size_t esz = m.c64 ? 8 : 16;
        const int P = m.paired ? 1 : 0;
        const int ngb = p.tile_bits - 6 - m.reg_sites;
        if (ngb < 0 || m.reg_sites < P || (1 << ngb) % m.waves != 0 || ((size_t)1 << p.tile_bits) * esz > 131072 ||
            p.tile_bits < 6 + P || (1 << (p.tile_bits - 6 - P)) % m.waves != 0 || (int)p.wb.size() < p.tile_bits) {
            if (err) *err = "tile shape does not fit the sub-cube kernel";
            return false;
        }
    }
    return !passes.empty();
}

}  // namespace

bool cube_precompile(const CubeModel &m, const std::vector<CubePass> &passes, std::string *err) {
    if (!check_model(m, passes, err)) return false;
    const std::string src = cube_source(m, passes);
    std::vector<char> code;
    return code_for(src, key_of(src), code, err);
}

std::shared_ptr<CubeProgram> cube_program(const CubeModel &m, const std::vector<CubePass> &passes,
                                          std::string *err) {
    if (!check_model(m, passes, err)) return nullptr;
    const std::string src = cube_source(m, passes);
    const std::string key = key_of(src);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        if (err) *err = "no HIP device";
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_prog.find({key, dev});
        if (it != g_prog.end())
            if (auto sp = it->second.lock()) return sp;
    }
    auto prog = std::make_shared<CubeProgram>();
    if (!code_for(src, key, prog->image, err)) return nullptr;
    if (hipModuleLoadData(&prog->module, prog->image.data()) != hipSuccess) {
        prog->module = nullptr;
        if (err) *err = "hipModuleLoadData failed for the generated H.psi kernels";
        return nullptr;
    }
    for (size_t i = 0; i < passes.size(); ++i) {
        hipFunction_t f;
        const std::string name = "esc_cube_p" + std::to_string(i);
        if (hipModuleGetFunction(&f, prog->module, name.c_str()) != hipSuccess) {
            if (err) *err = "generated kernel " + name + " missing";
            return nullptr;
        }
        prog->fn.push_back(f);
        const std::string bname = "esc_cube_b" + std::to_string(i);
        if (hipModuleGetFunction(&f, prog->module, bname.c_str()) != hipSuccess) {
            if (err) *err = "generated kernel " + bname + " missing";
            return nullptr;
        }
        prog->fnb.push_back(f);
    }
    prog->tile_bits = passes[0].tile_bits;
    prog->threads = m.waves * 64;
    std::lock_guard<std::mutex> lk(g_mu);
    g_prog[{key, dev}] = prog;
    return prog;
}

int cube_tile_bits(const CubeProgram &p) { return p.tile_bits; }

hipError_t cube_launch_batched(const CubeProgram &p, const double *d_cfs, const void *const *d_ins,
                               void *const *d_outs, double *const *d_dots, uint64_t dim, int ninst, hipStream_t s) {
    const uint64_t ntiles64 = dim >> p.tile_bits;
    if (ntiles64 == 0 || ntiles64 > 0x7fffffffull || ninst < 1 || ninst > 65535) return hipErrorInvalidValue;
    const uint32_t ntiles = (uint32_t)ntiles64;
    for (size_t i = 0; i < p.fnb.size(); ++i) {
        for (uint32_t base = 0; base < ntiles; base += kCubeMaxGrid) {
            const uint32_t n = (ntiles - base) < kCubeMaxGrid ? (ntiles - base) : kCubeMaxGrid;
            const double *a0 = d_cfs;
            const void *const *a1 = d_ins;
            void *const *a2 = d_outs;
            double *const *a3 = d_dots;
            uint32_t a4 = base;
            void *args[] = {&a0, &a1, &a2, &a3, &a4};
            const hipError_t e =
                hipModuleLaunchKernel(p.fnb[i], n, (unsigned)ninst, 1, (unsigned)p.threads, 1, 1, 0, s, args, nullptr);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

hipError_t cube_launch(const CubeProgram &p, const double *d_scalars, const void *in, void *out, double *dot_part,
                       uint64_t dim, hipStream_t s) {
    const uint64_t ntiles64 = dim >> p.tile_bits;
    if (ntiles64 == 0 || ntiles64 > 0x7fffffffull) return hipErrorInvalidValue;
    const uint32_t ntiles = (uint32_t)ntiles64;
    for (size_t i = 0; i < p.fn.size(); ++i) {
        double *dp = (i + 1 == p.fn.size()) ? dot_part : nullptr;
        for (uint32_t base = 0; base < ntiles; base += kCubeMaxGrid) {
            const uint32_t n = (ntiles - base) < kCubeMaxGrid ? (ntiles - base) : kCubeMaxGrid;
            const double *a0 = d_scalars;
            const void *a1 = in;
            void *a2 = out;
            double *a3 = dp;
            uint32_t a4 = base;
            void *args[] = {&a0, &a1, &a2, &a3, &a4};
            const hipError_t e = hipModuleLaunchKernel(p.fn[i], n, 1, 1, (unsigned)p.threads, 1, 1, 0, s, args, nullptr);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

}  // namespace esc
