// esc_internal.h -- device tables and kernel launchers shared by the C ABI
// (esc_capi.cpp) and the HIP kernels (esc_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace esc {

constexpr int kMaxSites = 32;
constexpr int kMaxBonds = 96;      // off-diagonal two-site terms
constexpr int kMaxCurrentPairs = 96;

// Matrix-element tables of H in gather form.  Row r, s_k = +1 if bit k of r
// is 0 (spin up), sigma_ij = s_i s_j.
//   diagonal:  sum over zdiag (+-kz4 by sigma, in order), then over jdiag
//              (+-j4), then bz * (N - 2 popcount(r)) / 2
//   bond b:    H[r, r ^ (e_i|e_j)] = p0[f] + p1[f] sigma_ij + i q[f] (s_i + s_j)
//   flip k:    H[r, r ^ e_k] = bx2 + g4 s_{ypart(k)} - i s_k (by2 + g4 s_{xpart(k)})
struct HTables {
    uint32_t n_sites;
    uint32_t n_bonds;       // off-diagonal bond terms
    uint32_t n_zdiag;       // Kitaev z bonds on the diagonal (reference order)
    uint32_t n_jdiag;       // bonds with a Heisenberg diagonal part (0 if J == 0)
    uint32_t n_flip;        // 0 (no single-flip terms) or n_sites
    uint32_t has_gamma_flip;
    uint32_t bond[kMaxBonds];     // i | j << 8 | family << 16
    uint32_t zdiag[kMaxBonds];    // i | j << 8
    uint32_t jdiag[kMaxBonds];    // i | j << 8
    int32_t xpart[kMaxSites];     // x-bond partner per site, -1 = none
    int32_t ypart[kMaxSites];     // y-bond partner per site, -1 = none
    double p0[3], p1[3], q[3];
    double kz4, j4, bz;
    double bx2, by2, g4;
};

// Tiled gather form (k_hpsi_tile).  Every off-diagonal term is
//   H[r, r ^ mask] = coef[ bit(r,b0) | bit(r,b1) << 1 | bit(r,b2) << 2 ]
// with an 8-entry complex table built on the host (unused selector bits repeat
// b0).  Terms are grouped by class = cplx | far << 1 | lane << 2, so the
// terms whose selector bits are all >= 6 (uniform across a wave) come first:
//   cplx  some table entry has a non-zero imaginary part
//   far   mask has bits at or above the tile bits (partner outside the LDS tile)
//   lane  a selector bit lies below bit 6 (coefficient differs across a wave)
constexpr int kMaxTerms = 128;
constexpr int kMaxDistGroups = 32;
constexpr int kNumClasses = 8;

constexpr int kMaxGroups = 128;    // row groups of 64 per tile (tiles up to 2^13 rows)

// One off-diagonal term of a pass: row r gathers coef[sel(r)] * psi[r ^ mask]
// where sel(r) collects the row bits at the selector sites b0, b1, b2
// (selector bit j = bit b_j of r).  The selector bits are split by where
// they live: lane sites (per lane, lp_pos/lp_w as bit-field extracts),
// tile-group sites (per row group, ugrp) and the remaining sites of the row
// (per tile, bsel as s_bfe operands offset | width << 16 on the tile base).
struct GTerm {
    uint32_t mask;          // partner offset in row space
    uint32_t mlo;           // mask & 63 (lane part)
    uint32_t mhib;          // mask bits outside the lane and tile-group sites
    uint32_t ct_base;       // lane-dependent terms: first entry of the table in LDS
    uint32_t lp_pos[3];     // lane selector bit positions (width 0: not a lane site)
    uint32_t lp_w[3];
    uint32_t bsel[3];       // base selector extracts
    uint32_t bits;          // b0 | b1 << 8 | b2 << 16
    uint32_t pad[2];
    double coef[8][2];
    float coef32[8][2];     // the same table rounded to float (complex64 vectors)
    // per row-group slot j = wave * (groups per wave) + k of the workgroup:
    // near terms: tile element index of the partner group's first row;
    // far terms: partner row bits on the tile-group sites
    uint32_t gofs[kMaxGroups];
    uint32_t ugrp[kMaxGroups];  // selector bits taken from the tile-group sites, as a byte
                                // offset into the term's coefficient table
};

// One pass of the windowed tiled H.psi.  The pass owns a window: an ordered
// set of w row bits (bits 0..5 first).  Rows are enumerated as
//   r = wscat(window id) | bscat(tile within window) | tscat[g] | lane
// where tscat places the row-group index g of a tile on window bits 6..TB-1,
// bscat the tile index on window bits TB..w-1 and wscat the window id on the
// bits outside the window (each as two byte-indexed lookup tables).  Terms
// whose mask lies inside the window are applied by this pass; the first pass
// also adds the diagonal and overwrites the output, later passes accumulate.
struct GTables {
    uint32_t n_sites, tile_bits, waves;
    uint32_t first_pass;            // 1: acc = diag * psi, else acc = out
    uint32_t last_pass;             // 1: the output rows are final (the fused <psi|H psi> is formed here)
    uint32_t n_windows, blocks_per_window, xcd_windows;
    uint32_t tscat[128];
    uint32_t bscat[2][256];
    uint32_t wscat[2][256];
    uint32_t cls_begin[kNumClasses + 1];
    // diagonal: closed form when the Kitaev-z sum is exact in any order
    uint32_t diag_closed;           // 1: popcount form, 0: sequential (HTables path)
    uint32_t n_zgrp, n_jgrp, nz, nj;
    uint32_t zgrp_d[kMaxDistGroups], zgrp_m[kMaxDistGroups];
    uint32_t jgrp_d[kMaxDistGroups], jgrp_m[kMaxDistGroups];
    double kz4, j4, bz;
    GTerm term[kMaxTerms];
};

// One factor of a Pauli-string term: op 0 = Sx, 1 = Sy, 2 = Sz at site.
struct OpFactor {
    int32_t op;
    int32_t site;
};

// A current (or single-site) operator applied in gather form.  Each pair
// contributes (wa * A - wb * B) psi with A, B products of up to three
// single-site factors, applied first-to-last (factors[0] acts first), and
// complex weights wa = wa_re + i wa_im (likewise wb).
struct CurrentPair {
    int32_t na, nb;             // number of factors (nb = 0: no B term)
    OpFactor fa[3], fb[3];
    double wa_re, wa_im, wb_re, wb_im;
};

struct CurrentTable {
    int32_t n_pairs;
    CurrentPair pair[kMaxCurrentPairs];
};

// Last-error slot of the C ABI (thread-local), shared by all host files.
void set_error(const char *msg);
void clear_error();

// Launchers (esc_kernels.hip).  dim = 2^N.  All return hipError_t of the launch.
// Tiled H.psi: 2^tile_bits rows per workgroup staged in LDS; h is the
// sequential-diagonal table used when g->diag_closed == 0.
// n_uniform = number of terms in classes 0..3, n_terms = all terms (for the
// LDS size).
// dot_part (may be null): the last pass also writes, per tile, the partial
// sums (re, im) of <in|out> over the tile's rows, at dot_part[2 * tile].
hipError_t launch_hpsi_tile(bool c64, const GTables *d_g, const HTables *d_h, int tile_bits, int waves,
                            int n_uniform, int n_terms, const void *in, void *out, double *dot_part, uint64_t dim,
                            hipStream_t s);
// tile shapes the tiled kernel is compiled for: 2^10..2^12 rows (2^13 for
// c64), 4, 8 or 16 waves per workgroup, 2 to 8 row groups of 64 per wave
bool tile_shape_supported(bool c64, int tile_bits, int waves);
// dynamic LDS of one tiled workgroup: the tile, the uniform coefficients per
// row group, the lane-dependent tables and their per-lane offsets
size_t tile_lds_bytes(bool c64, int tile_bits, int n_uniform, int n_terms);
constexpr size_t kMaxTileLds = 160 * 1024 - 1024;  // 160 KiB per CU minus the static row-group and sum tables
hipError_t launch_hpsi_c128(const HTables *d_tab, const void *in, void *out, uint64_t dim, hipStream_t s);
hipError_t launch_hpsi_c64(const HTables *d_tab, const void *in, void *out, uint64_t dim, hipStream_t s);
hipError_t launch_current_c128(const CurrentTable *d_tab, const void *in, void *out, uint64_t dim, hipStream_t s);
hipError_t launch_current_c64(const CurrentTable *d_tab, const void *in, void *out, uint64_t dim, hipStream_t s);

// Reductions write nblk partial results to `partials`, then a single-block
// pass writes the final FP64 result to `result`.  Deterministic.
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

// result[0..1] = <x|y>
hipError_t launch_dot(bool c64, const void *x, const void *y, uint64_t dim, double *partials, double *result, hipStream_t s);
// y = a x + b y (complex a, b held on the host)
hipError_t launch_axpby(bool c64, double ar, double ai, const void *x, double br, double bi, void *y, uint64_t dim, hipStream_t s);
// w = w - alpha v - beta u ; result[0] = |w|^2     (alpha complex, beta real)
hipError_t launch_lanczos_update(bool c64, void *w, const void *v, const void *u, const double *d_alpha,
                                 double beta, uint64_t dim, double *partials, double *result, hipStream_t s);
// x *= 1 / sqrt(result[0]) with result read on the device (no host round trip);
// x is left as it is when result[0] is not positive and finite
hipError_t launch_normalize(bool c64, void *x, const double *d_norm2, uint64_t dim, hipStream_t s);
// x *= scale (real)
hipError_t launch_scale(bool c64, void *x, double scale, uint64_t dim, hipStream_t s);
// result[q] = sum over p of partials[p * nvals + q], q < nvals, in a fixed order
hipError_t launch_reduce(const double *partials, int npart, int nvals, double *result, hipStream_t s);
// The plain recursion with device-side scalars (rec = {c_k, c_{k-1}}, the
// step index k in *step, dres = the H.psi dot and |w|^2, hist = 3 doubles per
// step); see k_rec_step.  Nothing in a step depends on host values.
hipError_t launch_rec_init(double *rec, uint32_t *step, double c0, hipStream_t s);
// w = c H u - alpha c u - beta_k c' u_prev from rec, dres[0..1] and
// hist[3 (k - 1) + 2] (u null at k = 0); kRedBlocks partials of |w|^2
hipError_t launch_lanczos_update_rec(bool c64, void *w, const void *v, const void *u, const double *rec,
                                     const double *dres, const double *hist, const uint32_t *step, uint64_t dim,
                                     double *partials, hipStream_t s);
// dres[2] = |w|^2 from those partials, hist[3k .. 3k+2] = (alpha re,
// alpha im, beta) if k < cap, then rec and *step advance one step
hipError_t launch_rec_step(double *rec, double *dres, double *hist, uint32_t *step, uint32_t cap,
                           const double *partials, hipStream_t s);
// Up to nsteps steps of the same recursion in one workgroup for dim <=
// kSmallRecMaxDim (one-row-per-lane H.psi from t; buffers b0..b2 rotate as u_k
// = b[(k+1) % 3], w = b[(k+2) % 3], u_{k-1} = b[k % 3]; with coef, acc +=
// coef_k c_k u_k before step k).  Ends with rec, *step and hist as after as
// many k_rec_step launches; stops after an exhausted step.
constexpr uint64_t kSmallRecMaxDim = 1u << 11;   // at most 2 rows per lane
hipError_t launch_rec_small(bool c64, const HTables *t, void *b0, void *b1, void *b2, double *rec, uint32_t *step,
                            double *hist, uint32_t cap, int nsteps, const double *coef, void *acc, uint64_t dim,
                            hipStream_t s);
// y += (coef[2k] + i coef[2k+1]) rec[0] x, k = *step
hipError_t launch_axpy_rec_step(bool c64, const double *coef, const uint32_t *step, const double *rec, const void *x,
                                void *y, uint64_t dim, hipStream_t s);
// The no-allocation dot: the same kernels as launch_dot with the partial sums
// in the library's fixed per-GPU array; result is caller device memory.
// scratch_init resolves that array's address on the current device (once per
// device; also done on the first launch_dot_noalloc there).
constexpr size_t kScratchBytes = 2 * kRedBlocks * sizeof(double);
hipError_t scratch_init();
hipError_t launch_dot_noalloc(bool c64, const void *x, const void *y, uint64_t dim, double *result, hipStream_t s);
// result[0] = |x|^2
hipError_t launch_norm2(bool c64, const void *x, uint64_t dim, double *partials, double *result, hipStream_t s);
// coef[m] = <V_m | w> for m < k (V_m = base + m * dim), FP64; result has 2k doubles
hipError_t launch_multidot(bool c64, const void *base, int k, const void *w, uint64_t dim, double *partials,
                           double *result, hipStream_t s);
// w = beta_w * w + sum_m c[m] V_m   (c on the device, complex FP64)
hipError_t launch_multi_axpy(bool c64, const void *base, int k, const double *d_coef, double sign, void *w,
                             uint64_t dim, bool overwrite, hipStream_t s);
// Philox4x32-10 Gaussian start vector, N(0, 1/2) per component
hipError_t launch_gaussian(bool c64, uint64_t seed, void *out, uint64_t dim, hipStream_t s);

// One instance of a batched Lanczos step: its three rotating vectors for this
// step (vp = v_{k-1}, v = v_k, w = H v_k -> v_{k+1}), the accumulation target
// of a replay, and its blocks in the flattened grid.
struct BatchLane {
    void *vp, *v, *w, *acc;
    uint64_t dim;
    uint32_t blk0, nblk;
    // tiled H.psi: the dot <v|w> comes from the kernel's ndot per-tile
    // partials at dotp (2 doubles each), and the batched dot pass skips the
    // lane; ndot = 0 otherwise
    const double *dotp;
    uint32_t ndot;
};
// The batched recursion keeps unnormalised vectors u_k (v_k = u_k / beta_k):
constexpr int kBatchDot = 0;      // alpha[i][k] = <v_k|H v_k>
constexpr int kBatchUpdate = 1;   // w = H v_k - alpha v_k - beta_k v_{k-1} (unscaled); beta[i][k+1] = |w|; k += 1
constexpr int kBatchAccum = 2;    // acc += coef[i][k] v_k
hipError_t launch_batch_step(bool c64, int piece, const BatchLane *d_lanes, int n, uint32_t total_blocks,
                             double *partials, double *alpha, int lda, double *beta, int ldb, const double *coef,
                             int ldc, uint32_t *step, hipStream_t s);
// One thread per instance i < n: the length of its recorded recursion, the
// lowest eigenvalue of its tridiagonal matrix (alpha[2 (i lda + k)] real
// parts, beta[i ldb + k]) into e0[i], its length into iters[i], and the unit
// eigenvector into coef[i ldc + k] (zero beyond the length).  work: 6 steps
// doubles per instance.
hipError_t launch_batch_tridiag(int n, const double *alpha, int lda, const double *beta, int ldb, int steps,
                                double *work, double *coef, int ldc, double *e0, int *iters, hipStream_t s);
// The continued fraction of every instance i < n on every omega q < nw, as
// complex doubles f[i nw + q], from the same alpha / beta rows, the norms
// n2[i] of the recursions' start vectors and the energies e0[i].
hipError_t launch_batch_cf(int n, const double *alpha, int lda, const double *beta, int ldb, int steps,
                           const double *n2, const double *e0, double eta, const double *omega, int nw, void *f,
                           hipStream_t s);
// All eigenpairs of the tridiagonal matrix a[0..n-1], b[1..n-1] (device
// arrays): ascending eigenvalues ev[n], column j of Z (n x n, column-major)
// the unit eigenvector j; work holds kTridiagThreads * 5 * n doubles;
// status[0] = 1 for a non-finite matrix.  Sturm bisection and inverse
// iteration, orthogonalised within clusters of close eigenvalues.
constexpr int kMaxTridiag = 4096;
constexpr int kTridiagThreads = 256;
hipError_t launch_tridiag_eig(int n, const double *a, const double *b, double *ev, double *Z, double *work,
                              int *status, hipStream_t s);
// convert FP64 complex device buffer to c64 (or copy for c128)
hipError_t launch_convert_from_c128(bool c64, const void *src, void *dst, uint64_t dim, hipStream_t s);
hipError_t launch_convert_to_c128(bool c64, const void *src, void *dst, uint64_t dim, hipStream_t s);

}  // namespace esc
