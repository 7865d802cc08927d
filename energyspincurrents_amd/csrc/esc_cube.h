// esc_cube.h -- H.psi kernels specialised to one lattice and compiled at run
// time (hiprtc) for gfx950.
//
// A planned pass (window W of sites held in L2, tile T of sites staged in LDS,
// the same plan as the generic tiled kernel) is emitted as straight-line
// code: every lane owns a register sub-cube of 2^k rows over k tile sites, so
// terms whose mask lies inside those sites take their partners from the
// lane's own registers; every other term loads its partner sub-cube once per
// distinct "outer part" of its mask (the mask without the register sites),
// shared by all terms with that outer part.  All LDS and row offsets are
// compile-time immediates and every coefficient is a selection among a few
// coupling scalars loaded once per workgroup, so the term loop issues only
// partner loads and FMAs.  Couplings stay run-time arguments: one compiled
// program serves every coupling set of the same lattice and term structure.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace esc {

// One off-diagonal term of the honeycomb gather form (coefficients as the
// HTables comment of esc_internal.h):
//   bond (family f, sites i, j):   p0[f] + p1[f] s_i s_j + i q[f] (s_i + s_j)
//   flip (site k, partners xp, yp): bx2 + g4 s_yp - i s_k (by2 + g4 s_xp)
struct CubeTerm {
    uint32_t mask;
    int kind;          // 0 = bond, 1 = flip
    int f, i, j;       // bond
    int k, xp, yp;     // flip; -1 = no partner
};

struct CubePass {
    std::vector<int> wb;    // window order: lane sites 0..5, tile sites, far sites
    std::vector<int> cb;    // window-id sites, ascending
    int tile_bits;
    std::vector<CubeTerm> terms;   // near and far terms of the pass
    bool first, last;              // first: the output starts at zero; last: <in|out> partials
    bool diag;                     // this pass adds the diagonal
};

struct CubeModel {
    int n_sites;
    bool c64;
    bool bond_cplx[3];      // q[f] != 0
    bool flip_cplx;         // by2 != 0 or g4 != 0
    bool flip_partners;     // Gamma: the x / y partners select the flip coefficient
    // diagonal in closed form: z and Heisenberg bonds counted by distance groups
    uint32_t nz, nj, n_zgrp, n_jgrp;
    uint32_t zgrp_d[32], zgrp_m[32], jgrp_d[32], jgrp_m[32];
    // or bond by bond in the reference's order (when the closed form's
    // partial sums would round differently): pairs i | j << 8
    bool seq_diag;
    uint32_t n_zpair, n_jpair;
    uint32_t zpair[64], jpair[96];
    int reg_sites;          // k: rows per lane = 2^k
    int waves;              // waves per workgroup
    bool paired;            // row pairs: site 0 is a register site under lanes on sites 1..6,
                            // so the lane moves rows 2y, 2y + 1 as one 16-byte access (complex64)
    bool contiguous;        // lanes on the six sites above the pair bit (one run of 512 or 1024
                            // bytes per wave access); else only a 64-byte run is fixed and the
                            // planner picks the other lane sites
};

// coupling scalars of the kernels (double, device memory), in this order
enum CubeScalar {
    kCubeA0, kCubeB0, kCubeQ0,   // bond family 0: p0 + p1, p0 - p1, 2 q
    kCubeA1, kCubeB1, kCubeQ1,
    kCubeA2, kCubeB2, kCubeQ2,
    kCubeRp, kCubeRm,            // flip real part bx2 + g4, bx2 - g4
    kCubeIp, kCubeIm,            // flip imaginary magnitude by2 + g4, by2 - g4
    kCubeKz4, kCubeJ4, kCubeBz,  // diagonal
    kCubeScalars
};

struct CubeProgram;

// Passes of the generated kernels: starting from one tile per pass (sites,
// the lane sites 0..5 implied), local search over the tile sites and the
// placement of every term (near in a pass whose tile holds its mask, else
// far in a pass with at most max_far sites outside its tile) minimises the
// partner sub-cubes loaded per row, far ones weighted above near ones.  Each
// pass's window is its tile plus the far sites its terms use.  False if some
// term fits no pass.
bool cube_plan(const CubeModel &m, const std::vector<CubeTerm> &terms, int tile_bits, std::vector<uint32_t> tiles,
               int max_far, std::vector<CubePass> &passes);
// The generated source of all passes (for tests and diagnostics).
std::string cube_source(const CubeModel &m, const std::vector<CubePass> &passes);
// Compile (or fetch from the process and disk caches) and load on the current
// device.  Returns null and sets *err when the run-time compiler is missing or
// fails.
std::shared_ptr<CubeProgram> cube_program(const CubeModel &m, const std::vector<CubePass> &passes,
                                          std::string *err);
// Compile only (no device needed); the code object lands in the disk cache.
bool cube_precompile(const CubeModel &m, const std::vector<CubePass> &passes, std::string *err);
// All passes of one H.psi; dot_part (may be null): the last pass writes the
// per-tile partial sums of <in|out> at dot_part[2 * tile].
hipError_t cube_launch(const CubeProgram &p, const double *d_scalars, const void *in, void *out, double *dot_part,
                       uint64_t dim, hipStream_t s);
int cube_tile_bits(const CubeProgram &p);
// All passes for ninst systems of one program at once (grid y = system):
// d_cfs holds kCubeScalars coupling scalars per system, d_ins / d_outs /
// d_dots one device pointer per system (d_dots as dot_part above, entries
// may be null).
hipError_t cube_launch_batched(const CubeProgram &p, const double *d_cfs, const void *const *d_ins,
                               void *const *d_outs, double *const *d_dots, uint64_t dim, int ninst, hipStream_t s);

}  // namespace esc
