// esc_kernels.hip -- HIP kernels for gfx950 (MI355X): matrix-free H.psi in gather
// form, current operators, and the FP64 reductions of the Lanczos recursion.
//
// Vectors are interleaved complex (double2 for c128, float2 for c64) indexed by
// the basis state r (bit k = site k, bit 0 = spin up).  Every output row is
// owned by exactly one lane, so no atomics are used anywhere.
#include "esc_internal.h"
#include "esc_tridiag.h"

#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>

namespace esc {
namespace {

constexpr int kThreads = 256;
constexpr int kNumXcd = 8;

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup g runs on
// XCD g % 8).  Remap so that each XCD walks one contiguous range of rows: the
// partner rows r ^ m of masks below that range's size then stay in the same
// XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t g, uint32_t nblk) {
    if (nblk % kNumXcd != 0) return g;
    uint32_t per = nblk / kNumXcd;
    return (g % kNumXcd) * per + g / kNumXcd;
}

template <typename R> struct Vec2;
template <> struct Vec2<double> { using T = double2; };
template <> struct Vec2<float> { using T = float2; };

__device__ __forceinline__ double sgn(uint32_t r, uint32_t k) { return ((r >> k) & 1u) ? -1.0 : 1.0; }

// ---------------------------------------------------------------------------
// H.psi, one row per lane.  The diagonal is accumulated in FP64 in the
// reference order (z bonds, then Heisenberg bonds, then Bzz * Sz_total), so for
// c128 every matrix element equals the reference's bit for bit.
// ---------------------------------------------------------------------------
// NR rows at a time, the term loop outermost (unrolled by 4), so 4 NR
// independent gathers are in flight; per row the arithmetic is that of one
// row.  NR <= 2 fits the registers of a 1024-lane workgroup without spills.
template <typename R, int NR>
__device__ __forceinline__ void hpsi_rows(const HTables *__restrict__ T, const typename Vec2<R>::T *__restrict__ in,
                                          const uint32_t (&r)[NR], typename Vec2<R>::T (&o)[NR]) {
    using V2 = typename Vec2<R>::T;
    double d[NR];
    {
        // no FMA contraction: Bzz * Sz_total is rounded before it is added, as
        // the reference's separate complex operators do (src/Hamiltonian.h:374)
#pragma clang fp contract(off)
#pragma unroll
        for (int q = 0; q < NR; ++q) d[q] = 0.0;
        const uint32_t nz = T->n_zdiag;
        const double kz4 = T->kz4;
        for (uint32_t b = 0; b < nz; ++b) {
            const uint32_t p = T->zdiag[b];
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                const uint32_t x = ((r[q] >> (p & 0xffu)) ^ (r[q] >> ((p >> 8) & 0xffu))) & 1u;
                d[q] += x ? -kz4 : kz4;
            }
        }
        const uint32_t nj = T->n_jdiag;
        const double j4 = T->j4;
        for (uint32_t b = 0; b < nj; ++b) {
            const uint32_t p = T->jdiag[b];
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                const uint32_t x = ((r[q] >> (p & 0xffu)) ^ (r[q] >> ((p >> 8) & 0xffu))) & 1u;
                d[q] += x ? -j4 : j4;
            }
        }
        // __popc returns unsigned: keep the difference signed
#pragma unroll
        for (int q = 0; q < NR; ++q) d[q] += T->bz * (0.5 * (double)((int)T->n_sites - 2 * (int)__popc(r[q])));
    }
    R ar[NR], ai[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const V2 v0 = in[r[q]];
        const R dd = (R)d[q];
        ar[q] = dd * v0.x;
        ai[q] = dd * v0.y;
    }

    const uint32_t nb = T->n_bonds;
#pragma unroll 4
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t p = T->bond[b];
        const uint32_t i = p & 0xffu, j = (p >> 8) & 0xffu, f = p >> 16;
        const uint32_t m = (1u << i) | (1u << j);
        V2 x[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) x[q] = in[r[q] ^ m];
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const R si = ((r[q] >> i) & 1u) ? R(-1) : R(1);
            const R sj = ((r[q] >> j) & 1u) ? R(-1) : R(1);
            const R cr = (R)T->p0[f] + (R)T->p1[f] * (si * sj);
            const R ci = (R)T->q[f] * (si + sj);
            ar[q] += cr * x[q].x - ci * x[q].y;
            ai[q] += cr * x[q].y + ci * x[q].x;
        }
    }

    const uint32_t nf = T->n_flip;
    const R bx2 = (R)T->bx2, by2 = (R)T->by2, g4 = (R)T->g4;
    const bool gam = T->has_gamma_flip != 0;
#pragma unroll 4
    for (uint32_t k = 0; k < nf; ++k) {
        V2 x[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) x[q] = in[r[q] ^ (1u << k)];
        const int32_t xp = gam ? T->xpart[k] : -1, yp = gam ? T->ypart[k] : -1;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const R sk = ((r[q] >> k) & 1u) ? R(-1) : R(1);
            R re = bx2, im = by2;
            if (gam) {
                if (xp >= 0) im += g4 * (((r[q] >> xp) & 1u) ? R(-1) : R(1));
                if (yp >= 0) re += g4 * (((r[q] >> yp) & 1u) ? R(-1) : R(1));
            }
            im = -sk * im;
            ar[q] += re * x[q].x - im * x[q].y;
            ai[q] += re * x[q].y + im * x[q].x;
        }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        o[q].x = ar[q];
        o[q].y = ai[q];
    }
}

template <typename R>
__device__ __forceinline__ typename Vec2<R>::T hpsi_row(const HTables *__restrict__ T,
                                                        const typename Vec2<R>::T *__restrict__ in, uint32_t r) {
    const uint32_t rr[1] = {r};
    typename Vec2<R>::T o[1];
    hpsi_rows<R, 1>(T, in, rr, o);
    return o[0];
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_hpsi(const HTables *__restrict__ T,
                                                   const typename Vec2<R>::T *__restrict__ in,
                                                   typename Vec2<R>::T *__restrict__ out, uint64_t dim,
                                                   uint32_t nblk_total, uint32_t block_base) {
    const uint32_t g = xcd_block(blockIdx.x + block_base, nblk_total);
    const uint64_t r64 = (uint64_t)g * kThreads + threadIdx.x;
    if (r64 >= dim) return;
    out[r64] = hpsi_row<R>(T, in, (uint32_t)r64);
}

// ---------------------------------------------------------------------------
// Tiled H.psi.  A workgroup owns 2^TB consecutive rows (the tile), stages them
// in LDS, and each wave walks row groups of 64 consecutive rows: lane = bits
// 0..5, everything above is uniform across the wave (r_hi).  Per term:
//  * partner: tile-local masks read LDS, others global memory through a
//    uniform base pointer (in + (r_hi ^ mask_hi)) plus a lane offset;
//  * coefficient: 8-entry table indexed by three bits of the row; for terms
//    whose selector bits are all >= 6 the index is uniform (scalar unit);
//  * real-coefficient terms use two FMAs, complex ones four.
// Output rows are written once each (no atomics).
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ double hpsi_diag(const GTables *__restrict__ G, const HTables *__restrict__ H,
                                            uint32_t r) {
#pragma clang fp contract(off)
    double d;
    if (G->diag_closed) {
        // Kitaev-z and Heisenberg sums by counting anti-aligned bonds; used only
        // when +-Kz/4 partial sums are exact, so the value equals the
        // reference's bond-by-bond accumulation bit for bit
        uint32_t cz = 0;
        const uint32_t nzg = G->n_zgrp;
        for (uint32_t q = 0; q < nzg; ++q) cz += __popc((r ^ (r >> G->zgrp_d[q])) & G->zgrp_m[q]);
        d = G->kz4 * (double)((int)G->nz - 2 * (int)cz);
        if (G->nj) {
            uint32_t cj = 0;
            const uint32_t njg = G->n_jgrp;
            for (uint32_t q = 0; q < njg; ++q) cj += __popc((r ^ (r >> G->jgrp_d[q])) & G->jgrp_m[q]);
            d = d + G->j4 * (double)((int)G->nj - 2 * (int)cj);
        }
    } else {
        d = 0.0;
        const uint32_t nz = H->n_zdiag;
        const double kz4 = H->kz4;
        for (uint32_t b = 0; b < nz; ++b) {
            const uint32_t p = H->zdiag[b];
            d += (((r >> (p & 0xffu)) ^ (r >> ((p >> 8) & 0xffu))) & 1u) ? -kz4 : kz4;
        }
        const uint32_t nj = H->n_jdiag;
        const double j4 = H->j4;
        for (uint32_t b = 0; b < nj; ++b) {
            const uint32_t p = H->jdiag[b];
            d += (((r >> (p & 0xffu)) ^ (r >> ((p >> 8) & 0xffu))) & 1u) ? -j4 : j4;
        }
    }
    d = d + G->bz * (0.5 * (double)((int)G->n_sites - 2 * (int)__popc(r)));
    return d;
}

// coefficient table of a term in the vector precision
template <typename R> __device__ __forceinline__ const unsigned char *coef_table(const GTerm &T);
template <> __device__ __forceinline__ const unsigned char *coef_table<double>(const GTerm &T) {
    return reinterpret_cast<const unsigned char *>(T.coef);
}
template <> __device__ __forceinline__ const unsigned char *coef_table<float>(const GTerm &T) {
    return reinterpret_cast<const unsigned char *>(T.coef32);
}

// selector bits of a term that sit on tile-base or window-id sites
__device__ __forceinline__ uint32_t base_selector(const GTerm &T, uint32_t base) {
    uint32_t sel = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) sel |= ((base >> (T.bsel[q] & 0xffffu)) & (T.bsel[q] >> 16)) << q;
    return sel;
}

// Terms of one class for all NG row groups of a wave.  Row group k of wave w
// is g = w + WPB * k, rows r = base | tscat[g] | lane; its per-term constants
// sit in slot w * NG + k of the term record.
template <typename R, int TB, int WPB, int CLS>
__device__ __forceinline__ void hpsi_class(const GTables *__restrict__ G, uint32_t base, uint32_t wave,
                                           uint32_t lane, const typename Vec2<R>::T *__restrict__ in,
                                           const typename Vec2<R>::T *s_in, const typename Vec2<R>::T *s_ct,
                                           const uint32_t *s_lp, R *ar, R *ai) {
    using V2 = typename Vec2<R>::T;
    constexpr int NGRP = 1 << (TB - 6), NG = NGRP / WPB;
    constexpr bool CPLX = (CLS & 1) != 0, FAR = (CLS & 2) != 0, LANE = (CLS & 4) != 0;
    const uint32_t t0 = G->cls_begin[CLS], t1 = G->cls_begin[CLS + 1];
    const uint32_t nU = G->cls_begin[4];
    const uint32_t slot0 = wave * (uint32_t)NG;
    // two terms per iteration: their loads are in flight together
#pragma unroll 2
    for (uint32_t t = t0; t < t1; ++t) {
        const GTerm &T = G->term[t];
        // lane part of the partner row; far terms also fix the partner's
        // sites outside the tile for the whole tile
        const uint32_t loff = lane ^ T.mlo;
        const V2 *src = FAR ? in + (size_t)(base ^ T.mhib) : s_in;
        // lane-dependent coefficients: byte offset of this lane's entry in the
        // term's LDS table for row group 0 of the tile (set up per workgroup);
        // wave-uniform ones: the tile's entry of the term's table in memory,
        // read through the scalar unit
        const uint32_t lpc = LANE ? s_lp[(t - nU) * 64u + lane] : 0u;
        const unsigned char *utab = LANE ? nullptr : coef_table<R>(T) + base_selector(T, base) * (uint32_t)sizeof(V2);
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            const uint32_t j = slot0 + (uint32_t)k;
            const V2 x = src[T.gofs[j] + loff];
            V2 c;
            if (LANE) {
                c = *reinterpret_cast<const V2 *>(reinterpret_cast<const unsigned char *>(s_ct) + lpc + T.ugrp[j]);
            } else {
                c = *reinterpret_cast<const V2 *>(utab + T.ugrp[j]);
            }
            // accumulate with fused multiply-adds (4 FMAs per complex term)
            if (CPLX) {
                ar[k] = fma(c.x, x.x, ar[k]);
                ar[k] = fma(-c.y, x.y, ar[k]);
                ai[k] = fma(c.x, x.y, ai[k]);
                ai[k] = fma(c.y, x.x, ai[k]);
            } else {
                ar[k] = fma(c.x, x.x, ar[k]);
                ai[k] = fma(c.x, x.y, ai[k]);
            }
        }
    }
}

// (a, b) summed over the NT threads of a workgroup in a fixed order; thread 0
// writes the two sums to dst
template <int NT>
__device__ __forceinline__ void group_sum2_store(double a, double b, double *dst) {
    __shared__ double sh[2][NT / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if ((threadIdx.x & 63u) == 0) {
        sh[0][threadIdx.x >> 6] = a;
        sh[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0;
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) {
            x += sh[0][i];
            y += sh[1][i];
        }
        dst[0] = x;
        dst[1] = y;
    }
}

template <typename R, int TB, int WPB>
__global__ __launch_bounds__(WPB * 64) void k_hpsi_tile(const GTables *__restrict__ G, const HTables *__restrict__ H,
                                                        const typename Vec2<R>::T *__restrict__ in,
                                                        typename Vec2<R>::T *__restrict__ out,
                                                        double *__restrict__ dot_part, uint32_t nblk_total,
                                                        uint32_t block_base) {
    using V2 = typename Vec2<R>::T;
    constexpr int NGRP = 1 << (TB - 6), NG = NGRP / WPB;
    constexpr int NT = WPB * 64;
    constexpr uint32_t NROW = 1u << TB;
    extern __shared__ unsigned char esc_lds[];
    V2 *s_in = reinterpret_cast<V2 *>(esc_lds);                 // the tile
    const uint32_t nU = G->cls_begin[4], nT = G->cls_begin[kNumClasses];
    V2 *s_ct = s_in + NROW;                                     // lane-term tables [nT - nU][8]
    uint32_t *s_lp = reinterpret_cast<uint32_t *>(s_ct + (nT - nU) * 8u);   // [nT - nU][64] byte offsets
    __shared__ uint32_t s_rhi[NGRP];                            // first row of each row group
    // block -> (window, tile within window); with xcd_windows every window
    // is walked by one XCD (workgroups are dealt round-robin over 8 XCDs)
    const uint32_t b = blockIdx.x + block_base;
    const uint32_t bpw = G->blocks_per_window;
    uint32_t wi, tl;
    if (G->xcd_windows) {
        const uint32_t seq = b / kNumXcd;
        wi = (seq / bpw) * kNumXcd + b % kNumXcd;
        tl = seq % bpw;
    } else {
        wi = b / bpw;
        tl = b % bpw;
    }
    const uint32_t base = G->wscat[0][wi & 0xffu] | G->wscat[1][(wi >> 8) & 0xffu] | G->bscat[0][tl & 0xffu] |
                          G->bscat[1][(tl >> 8) & 0xffu];
    for (uint32_t g = threadIdx.x; g < (uint32_t)NGRP; g += NT) s_rhi[g] = base | G->tscat[g];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < NROW; i += NT) s_in[i] = in[s_rhi[i >> 6] | (i & 63u)];
    for (uint32_t e = threadIdx.x; e < (nT - nU) * 8u; e += NT) {
        V2 c;
        c.x = (R)G->term[nU + e / 8u].coef[e % 8u][0];
        c.y = (R)G->term[nU + e / 8u].coef[e % 8u][1];
        s_ct[e] = c;
    }
    // lane-dependent terms: table entry of each lane for row group 0, from the
    // selector bits on lane sites (per lane) and on tile-base sites (per tile);
    // the row group adds its own selector bits (ugrp, in bytes)
    for (uint32_t e = threadIdx.x; e < (nT - nU) * 64u; e += NT) {
        const GTerm &T = G->term[nU + e / 64u];
        const uint32_t ln = e % 64u;
        uint32_t sel = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            sel |= ((ln >> T.lp_pos[q]) & T.lp_w[q]) << q;
            sel |= ((base >> (T.bsel[q] & 0xffffu)) & (T.bsel[q] >> 16)) << q;
        }
        s_lp[e] = (T.ct_base + sel) * (uint32_t)sizeof(V2);
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    R ar[NG], ai[NG];
    if (G->first_pass) {
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            const uint32_t g = wave + (uint32_t)WPB * (uint32_t)k;
            const V2 v0 = s_in[(g << 6) | lane];
            const R dd = (R)hpsi_diag<R>(G, H, s_rhi[g] | lane);
            ar[k] = dd * v0.x;
            ai[k] = dd * v0.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            // streamed like the store below: keeps the window rows in L2
            typedef R nvec2 __attribute__((ext_vector_type(2)));
            const nvec2 o = __builtin_nontemporal_load(
                reinterpret_cast<const nvec2 *>(out + (s_rhi[wave + (uint32_t)WPB * (uint32_t)k] | lane)));
            ar[k] = o.x;
            ai[k] = o.y;
        }
    }
    hpsi_class<R, TB, WPB, 0>(G, base, wave, lane, in, s_in, s_ct, s_lp, ar, ai);
    hpsi_class<R, TB, WPB, 1>(G, base, wave, lane, in, s_in, s_ct, s_lp, ar, ai);
    hpsi_class<R, TB, WPB, 2>(G, base, wave, lane, in, s_in, s_ct, s_lp, ar, ai);
    hpsi_class<R, TB, WPB, 3>(G, base, wave, lane, in, s_in, s_ct, s_lp, ar, ai);
    hpsi_class<R, TB, WPB, 4>(G, base, wave, lane, in, s_in, s_ct, s_lp, ar, ai);
    hpsi_class<R, TB, WPB, 5>(G, base, wave, lane, in, s_in, s_ct, s_lp, ar, ai);
    hpsi_class<R, TB, WPB, 6>(G, base, wave, lane, in, s_in, s_ct, s_lp, ar, ai);
    hpsi_class<R, TB, WPB, 7>(G, base, wave, lane, in, s_in, s_ct, s_lp, ar, ai);
    // streamed result: a non-temporal store keeps it from displacing the
    // input rows that other tiles still gather from
    typedef R nvec2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        const uint32_t r = s_rhi[wave + (uint32_t)WPB * (uint32_t)k] | lane;
        nvec2 o = {ar[k], ai[k]};
        __builtin_nontemporal_store(o, reinterpret_cast<nvec2 *>(out + r));
    }
    // <in|out> over the tile (the Lanczos alpha), from the staged input rows
    // and the stored values, accumulated in FP64
    if (dot_part != nullptr && G->last_pass) {
        double dr = 0.0, di = 0.0;
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            const V2 x = s_in[((wave + (uint32_t)WPB * (uint32_t)k) << 6) | lane];
            const double xr = x.x, xi = x.y, yr = ar[k], yi = ai[k];
            dr += xr * yr + xi * yi;
            di += xr * yi - xi * yr;
        }
        group_sum2_store<WPB * 64>(dr, di, dot_part + 2 * (size_t)(blockIdx.x + block_base));
    }
}

// ---------------------------------------------------------------------------
// Pauli-string operators (single-site S^a, spin and energy currents).
// Row r gathers from c = r ^ M where M flips the sites of the Sx/Sy factors;
// the coefficient follows the factor sequence exactly as the reference applies
// it (first factor acts first on c), so products are exact and the only
// rounding is the coupling weight times the value.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void string_term(const OpFactor *f, int n, uint32_t r, double w_re, double w_im,
                                            const typename Vec2<R>::T *__restrict__ in, R &outr, R &outi) {
    uint32_t m = 0;
    for (int t = 0; t < n; ++t)
        if (f[t].op != 2) m ^= 1u << f[t].site;
    const uint32_t c = r ^ m;
    uint32_t k = c;
    // coefficient = w * (cr + i ci), built from exact factors +-1/2, +-i/2
    double cr = 1.0, ci = 0.0;
    for (int t = 0; t < n; ++t) {
        const uint32_t s = f[t].site;
        const uint32_t bit = (k >> s) & 1u;
        if (f[t].op == 0) {
            cr *= 0.5; ci *= 0.5; k ^= 1u << s;
        } else if (f[t].op == 1) {
            const double h = bit ? -0.5 : 0.5;   // multiply by i*h
            const double nr = -ci * h, ni = cr * h;
            cr = nr; ci = ni; k ^= 1u << s;
        } else {
            const double h = bit ? -0.5 : 0.5;
            cr *= h; ci *= h;
        }
    }
    const R wr = (R)(w_re * cr - w_im * ci), wi = (R)(w_re * ci + w_im * cr);
    const typename Vec2<R>::T x = in[c];
    outr = wr * x.x - wi * x.y;
    outi = wr * x.y + wi * x.x;
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_current(const CurrentTable *__restrict__ T,
                                                      const typename Vec2<R>::T *__restrict__ in,
                                                      typename Vec2<R>::T *__restrict__ out, uint64_t dim,
                                                      uint32_t nblk_total, uint32_t block_base) {
    const uint32_t g = xcd_block(blockIdx.x + block_base, nblk_total);
    const uint64_t r64 = (uint64_t)g * kThreads + threadIdx.x;
    if (r64 >= dim) return;
    const uint32_t r = (uint32_t)r64;
    R accr = 0, acci = 0;
    const int np = T->n_pairs;
    for (int p = 0; p < np; ++p) {
        const CurrentPair &cp = T->pair[p];
        R ar, ai;
        string_term<R>(cp.fa, cp.na, r, cp.wa_re, cp.wa_im, in, ar, ai);
        if (cp.nb > 0) {
            R br, bi;
            string_term<R>(cp.fb, cp.nb, r, cp.wb_re, cp.wb_im, in, br, bi);
            ar = ar - br;
            ai = ai - bi;
        }
        accr += ar;
        acci += ai;
    }
    typename Vec2<R>::T o;
    o.x = accr;
    o.y = acci;
    out[r] = o;
}

// ---------------------------------------------------------------------------
// Deterministic FP64 block reduction helpers
// ---------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double *dst) {
    __shared__ double sh[NV][kThreads];
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int q = 0; q < NV; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) dst[q] = sh[q][0];
    }
}

// final pass: nvals independent sums, each of npart partials laid out
// partials[part * nvals + q]; one block per group of values
__global__ __launch_bounds__(kThreads) void k_reduce_final(const double *__restrict__ partials, int npart, int nvals,
                                                           double *__restrict__ result) {
    const int q = blockIdx.x;
    double v[1] = {0.0};
    for (int p = threadIdx.x; p < npart; p += kThreads) v[0] += partials[(size_t)p * nvals + q];
    block_reduce_store<1>(v, result + q);
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_dot_partial(const typename Vec2<R>::T *__restrict__ x,
                                                          const typename Vec2<R>::T *__restrict__ y, uint64_t dim,
                                                          double *__restrict__ partials) {
    double v[2] = {0.0, 0.0};
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim; i += stride) {
        const double xr = x[i].x, xi = x[i].y, yr = y[i].x, yi = y[i].y;
        v[0] += xr * yr + xi * yi;
        v[1] += xr * yi - xi * yr;
    }
    block_reduce_store<2>(v, partials + 2 * blockIdx.x);
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_norm2_partial(const typename Vec2<R>::T *__restrict__ x, uint64_t dim,
                                                            double *__restrict__ partials) {
    double v[1] = {0.0};
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim; i += stride) {
        const double xr = x[i].x, xi = x[i].y;
        v[0] += xr * xr + xi * xi;
    }
    block_reduce_store<1>(v, partials + blockIdx.x);
}

template <typename R>
__device__ __forceinline__ void axpby_body(double ar, double ai, const typename Vec2<R>::T *__restrict__ x,
                                           double br, double bi, typename Vec2<R>::T *__restrict__ y, uint64_t dim) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim; i += stride) {
        const double xr = x[i].x, xi = x[i].y, yr = y[i].x, yi = y[i].y;
        typename Vec2<R>::T o;
        o.x = (R)((ar * xr - ai * xi) + (br * yr - bi * yi));
        o.y = (R)((ar * xi + ai * xr) + (br * yi + bi * yr));
        y[i] = o;
    }
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_axpby(double ar, double ai, const typename Vec2<R>::T *__restrict__ x,
                                                    double br, double bi, typename Vec2<R>::T *__restrict__ y,
                                                    uint64_t dim) {
    axpby_body<R>(ar, ai, x, br, bi, y, dim);
}


// w = w - alpha v - beta u  (as V2 - alpha*V1 - beta*V0, src/DynLanczosSolver.h:96),
// partial |w|^2 of the updated w
template <typename R>
__global__ __launch_bounds__(kThreads) void k_lanczos_update(typename Vec2<R>::T *__restrict__ w,
                                                             const typename Vec2<R>::T *__restrict__ v,
                                                             const typename Vec2<R>::T *__restrict__ u,
                                                             const double *__restrict__ d_alpha, double beta,
                                                             uint64_t dim, double *__restrict__ partials) {
    const double a_r = d_alpha[0], a_i = d_alpha[1];
    double acc[1] = {0.0};
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim; i += stride) {
        const double vr = v[i].x, vi = v[i].y;
        double wr = w[i].x - (a_r * vr - a_i * vi);
        double wi = w[i].y - (a_r * vi + a_i * vr);
        if (u != nullptr) {
            wr -= beta * (double)u[i].x;
            wi -= beta * (double)u[i].y;
        }
        typename Vec2<R>::T o;
        o.x = (R)wr;
        o.y = (R)wi;
        w[i] = o;
        acc[0] += (double)o.x * o.x + (double)o.y * o.y;
    }
    block_reduce_store<1>(acc, partials + blockIdx.x);
}

// w = a w + b v + e u (a, e real, b complex; u may be null), partial |w|^2
template <typename R, bool NT = false>
__device__ __forceinline__ void lanczos_update_body(typename Vec2<R>::T *__restrict__ w,
                                                    const typename Vec2<R>::T *__restrict__ v,
                                                    const typename Vec2<R>::T *__restrict__ u, double a, double br,
                                                    double bi, double e, uint64_t dim, double *__restrict__ partials) {
    double acc[1] = {0.0};
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    const uint64_t first = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if constexpr (NT) {
        static_assert(sizeof(R) == 8, "streamed update: complex128 only");
        typedef double D2 __attribute__((ext_vector_type(2)));
        const D2 *v2 = (const D2 *)v, *u2 = (const D2 *)u;
        D2 *w2 = (D2 *)w;
        for (uint64_t i = first; i < dim; i += stride) {
            const D2 x = __builtin_nontemporal_load(v2 + i), y = __builtin_nontemporal_load(w2 + i);
            double wr = a * y.x + (br * x.x - bi * x.y);
            double wi = a * y.y + (br * x.y + bi * x.x);
            if (u != nullptr) {
                const D2 z = __builtin_nontemporal_load(u2 + i);
                wr += e * z.x;
                wi += e * z.y;
            }
            D2 o;
            o.x = wr; o.y = wi;
            __builtin_nontemporal_store(o, w2 + i);
            acc[0] += wr * wr + wi * wi;
        }
    } else if constexpr (sizeof(R) == 4) {
        // complex64: two elements per lane (16-byte accesses); dim is a power
        // of two >= 2
        const float4 *v4 = (const float4 *)v, *u4 = (const float4 *)u;
        float4 *w4 = (float4 *)w;
        for (uint64_t i = first; i < dim / 2; i += stride) {
            const float4 x = v4[i], y = w4[i];
            double r0 = a * (double)y.x + (br * (double)x.x - bi * (double)x.y);
            double i0 = a * (double)y.y + (br * (double)x.y + bi * (double)x.x);
            double r1 = a * (double)y.z + (br * (double)x.z - bi * (double)x.w);
            double i1 = a * (double)y.w + (br * (double)x.w + bi * (double)x.z);
            if (u != nullptr) {
                const float4 z = u4[i];
                r0 += e * (double)z.x;
                i0 += e * (double)z.y;
                r1 += e * (double)z.z;
                i1 += e * (double)z.w;
            }
            const float4 o = make_float4((float)r0, (float)i0, (float)r1, (float)i1);
            w4[i] = o;
            acc[0] += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
            acc[0] += (double)o.z * (double)o.z + (double)o.w * (double)o.w;
        }
    } else {
        for (uint64_t i = first; i < dim; i += stride) {
            const double vr = v[i].x, vi = v[i].y;
            double wr = a * (double)w[i].x + (br * vr - bi * vi);
            double wi = a * (double)w[i].y + (br * vi + bi * vr);
            if (u != nullptr) {
                wr += e * (double)u[i].x;
                wi += e * (double)u[i].y;
            }
            typename Vec2<R>::T o;
            o.x = (R)wr;
            o.y = (R)wi;
            w[i] = o;
            const double sr = o.x, si = o.y;
            acc[0] += sr * sr + si * si;
        }
    }
    block_reduce_store<1>(acc, partials + blockIdx.x);
}

// ---------------------------------------------------------------------------
// The plain Lanczos recursion with its scalars kept on the device, so a step
// needs no host round trip.
// The kept vector is u_k with v_k = c_k u_k; rec[0] = c_k, rec[1] = c_{k-1},
// and the step index k sits in device memory (rec_step); dres[0..1] =
// <u_k|H u_k> (the fused dot of the H.psi kernel), dres[2] = |u_{k+1}|^2;
// hist[3k .. 3k+2] = (Re alpha_k, Im alpha_k, beta_{k+1}).  Same arithmetic
// as a host-driven step: alpha_k = c_k^2 <u_k|H u_k>, u_{k+1} = c_k H u_k -
// alpha_k c_k u_k - beta_k c_{k-1} u_{k-1}, beta_{k+1} = |u_{k+1}|, c_{k+1} =
// 1 / beta_{k+1}.
// ---------------------------------------------------------------------------
__global__ void k_rec_init(double *__restrict__ rec, uint32_t *__restrict__ step, double c0) {
    rec[0] = c0;
    rec[1] = 0.0;
    step[0] = 0;
}

// NT: streamed (non-temporal) loads and stores, used for complex128, where
// they measured faster (about 15 % at 2^24 rows); complex64 keeps the cached
// accesses, which measured faster there and at 2^32 rows (a recursion step
// 0.259 against 0.268 ms and 93.8 against 95.6 ms)
template <typename R, bool NT>
__global__ __launch_bounds__(kThreads) void k_lanczos_update_rec(typename Vec2<R>::T *__restrict__ w,
                                                                 const typename Vec2<R>::T *__restrict__ v,
                                                                 const typename Vec2<R>::T *__restrict__ u,
                                                                 const double *__restrict__ rec,
                                                                 const double *__restrict__ dres,
                                                                 const double *__restrict__ hist,
                                                                 const uint32_t *__restrict__ step, uint64_t dim,
                                                                 double *__restrict__ partials) {
    const double c = rec[0], cp = rec[1];
    const double ar = c * c * dres[0], ai = c * c * dres[1];
    // beta_k of the previous step (u is null at k = 0)
    const double e = (u != nullptr) ? -hist[3 * (size_t)(step[0] - 1) + 2] * cp : 0.0;
    lanczos_update_body<R, NT>(w, v, u, c, -ar * c, -ai * c, e, dim, partials);
}

// one block: dres[2] = |u_{k+1}|^2 from the update's partials (as
// k_reduce_final sums them), then the scalars of the step
__global__ __launch_bounds__(kThreads) void k_rec_step(double *__restrict__ rec, double *__restrict__ dres,
                                                       double *__restrict__ hist, uint32_t *__restrict__ step,
                                                       uint32_t cap, const double *__restrict__ partials,
                                                       int npart) {
    double v[1] = {0.0};
    for (int p = threadIdx.x; p < npart; p += kThreads) v[0] += partials[p];
    block_reduce_store<1>(v, dres + 2);
    if (threadIdx.x != 0) return;
    const uint32_t k = step[0];
    const double c = rec[0];
    const double b = sqrt(dres[2]);
    if (k < cap) {   // steps past the recorded capacity are never read
        hist[3 * (size_t)k] = c * c * dres[0];
        hist[3 * (size_t)k + 1] = c * c * dres[1];
        hist[3 * (size_t)k + 2] = b;
    }
    rec[1] = c;
    rec[0] = 1.0 / b;
    step[0] = k + 1;
}

// y += (coef[2k] + i coef[2k+1]) c_k x for the step k held on the device
template <typename R>
__global__ __launch_bounds__(kThreads) void k_axpy_rec_step(const double *__restrict__ coef,
                                                            const uint32_t *__restrict__ step,
                                                            const double *__restrict__ rec,
                                                            const typename Vec2<R>::T *__restrict__ x,
                                                            typename Vec2<R>::T *__restrict__ y, uint64_t dim) {
    const uint32_t k = step[0];
    const double c = rec[0];
    axpby_body<R>(coef[2 * (size_t)k] * c, coef[2 * (size_t)k + 1] * c, x, 1.0, 0.0, y, dim);
}

// ---------------------------------------------------------------------------
// The same recursion for small systems, up to `nsteps` steps in one
// workgroup.  In the multi-kernel form a step costs five launches (H.psi, the
// dot and its reduction, the update, the scalars), and below a few thousand
// rows those launches, not the rows, set the time.  Here one workgroup of
// kSmallThreads lanes walks the rows of every pass and keeps the scalars in
// registers; rec, *step and hist end with the values k_rec_step would leave,
// so the two forms can follow each other.  After the update pass every wave
// waits for its stores and invalidates the CU's L1 (agent-scope fence) before
// the barrier, so the next step's gathers read the new rows.  The loop stops
// after a step whose beta is not above 1e-13 max(1, |alpha|) (the host's
// exhaustion rule; a non-finite beta stops it too).
// ---------------------------------------------------------------------------
constexpr int kSmallThreads = 1024;
constexpr int kSmallWaves = kSmallThreads / 64;

// sum over the workgroup in a fixed order (lanes by xor shuffle, then the
// waves' sums in wave order); every lane returns the total
template <int NV>
__device__ __forceinline__ void small_block_sum(double (&v)[NV], double (&sh)[NV][kSmallWaves]) {
#pragma unroll
    for (int q = 0; q < NV; ++q)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) sh[q][threadIdx.x >> 6] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double t = 0.0;
        for (int wv = 0; wv < kSmallWaves; ++wv) t += sh[q][wv];
        v[q] = t;
    }
    __syncthreads();   // sh is written again by the next sum
}

template <typename R, int NR>
__global__ __launch_bounds__(kSmallThreads) void k_rec_small(const HTables *__restrict__ T, typename Vec2<R>::T *b0,
                                                             typename Vec2<R>::T *b1, typename Vec2<R>::T *b2,
                                                             double *__restrict__ rec, uint32_t *__restrict__ step,
                                                             double *__restrict__ hist, uint32_t cap, int nsteps,
                                                             const double *__restrict__ coef,
                                                             typename Vec2<R>::T *acc, uint32_t dim) {
    using V2 = typename Vec2<R>::T;
    __shared__ double sh_dot[2][kSmallWaves];
    __shared__ double sh_n2[1][kSmallWaves];
    V2 *const buf[3] = {b0, b1, b2};
    uint32_t k = step[0];
    double c = rec[0], cp = rec[1];
    double beta = k > 0 ? hist[3 * (size_t)(k - 1) + 2] : 0.0;
    for (int it = 0; it < nsteps; ++it) {
        V2 *const u = buf[(k + 1) % 3];
        V2 *const w = buf[(k + 2) % 3];
        const V2 *const up = k > 0 ? buf[k % 3] : nullptr;
        if (coef != nullptr) {   // acc += (coef[2k] + i coef[2k+1]) c_k u_k, as k_axpy_rec_step
            const double ar = coef[2 * (size_t)k] * c, ai = coef[2 * (size_t)k + 1] * c;
            for (uint32_t i = threadIdx.x; i < dim; i += kSmallThreads) {
                const double xr = u[i].x, xi = u[i].y, yr = acc[i].x, yi = acc[i].y;
                V2 o;
                o.x = (R)((ar * xr - ai * xi) + yr);
                o.y = (R)((ar * xi + ai * xr) + yi);
                acc[i] = o;
            }
        }
        // w = H u_k and <u_k|H u_k>: lane t owns rows t + q kSmallThreads,
        // q < NR (dim = NR kSmallThreads when NR > 1), and reads back only
        // its own rows of w
        double d[2] = {0.0, 0.0};
        if (NR > 1 || threadIdx.x < dim) {
            uint32_t rr[NR];
#pragma unroll
            for (int q = 0; q < NR; ++q) rr[q] = threadIdx.x + (uint32_t)q * kSmallThreads;
            V2 o[NR];
            hpsi_rows<R, NR>(T, u, rr, o);
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                w[rr[q]] = o[q];
                const double xr = u[rr[q]].x, xi = u[rr[q]].y, yr = o[q].x, yi = o[q].y;
                d[0] += xr * yr + xi * yi;
                d[1] += xr * yi - xi * yr;
            }
        }
        small_block_sum<2>(d, sh_dot);
        // u_{k+1} = c_k H u_k - alpha_k c_k u_k - beta_k c_{k-1} u_{k-1}
        const double ar = c * c * d[0], ai = c * c * d[1];
        const double br = -ar * c, bi = -ai * c, e = up != nullptr ? -beta * cp : 0.0;
        double n2[1] = {0.0};
        for (uint32_t i = threadIdx.x; i < dim; i += kSmallThreads) {
            const double vr = u[i].x, vi = u[i].y;
            double wr = c * (double)w[i].x + (br * vr - bi * vi);
            double wi = c * (double)w[i].y + (br * vi + bi * vr);
            if (up != nullptr) {
                wr += e * (double)up[i].x;
                wi += e * (double)up[i].y;
            }
            V2 o;
            o.x = (R)wr;
            o.y = (R)wi;
            w[i] = o;
            n2[0] += (double)o.x * o.x + (double)o.y * o.y;
        }
        small_block_sum<1>(n2, sh_n2);
        const double b = sqrt(n2[0]);
        if (threadIdx.x == 0 && k < cap) {
            hist[3 * (size_t)k] = ar;
            hist[3 * (size_t)k + 1] = ai;
            hist[3 * (size_t)k + 2] = b;
        }
        cp = c;
        c = 1.0 / b;
        beta = b;
        ++k;
        __threadfence();
        __syncthreads();
        if (!(b > 1e-13 * fmax(1.0, fabs(ar)))) break;   // the same value in every lane
    }
    if (threadIdx.x == 0) {
        rec[0] = c;
        rec[1] = cp;
        step[0] = k;
    }
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_normalize(typename Vec2<R>::T *__restrict__ x,
                                                        const double *__restrict__ d_norm2, uint64_t dim) {
    const double n2 = d_norm2[0];
    if (!(n2 > 0.0) || !(n2 < INFINITY)) return;   // a zero or non-finite norm leaves x as it is
    const double s = 1.0 / sqrt(n2);
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim; i += stride) {
        typename Vec2<R>::T o = x[i];
        o.x = (R)(s * o.x);
        o.y = (R)(s * o.y);
        x[i] = o;
    }
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_scale(typename Vec2<R>::T *__restrict__ x, double s, uint64_t dim) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim; i += stride) {
        typename Vec2<R>::T o = x[i];
        o.x = (R)(s * o.x);
        o.y = (R)(s * o.y);
        x[i] = o;
    }
}

// coef[m] = <V_m|w> partials: grid (nchunk, ceil(k/kMultiVecs)); kMultiVecs
// basis vectors per block, so w is read ceil(k/kMultiVecs) times
constexpr int kMultiChunks = 256;
constexpr int kMultiVecs = 4;
template <typename R>
__global__ __launch_bounds__(kThreads) void k_multidot_partial(const typename Vec2<R>::T *__restrict__ base, int k,
                                                               const typename Vec2<R>::T *__restrict__ w,
                                                               uint64_t dim, double *__restrict__ partials) {
    constexpr int MV = kMultiVecs;
    const int m0 = blockIdx.y * MV;
    double v[2 * MV];
#pragma unroll
    for (int q = 0; q < 2 * MV; ++q) v[q] = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim; i += stride) {
        const double wr = w[i].x, wi = w[i].y;
#pragma unroll
        for (int q = 0; q < MV; ++q) {
            if (m0 + q < k) {
                const typename Vec2<R>::T x = base[(uint64_t)(m0 + q) * dim + i];
                v[2 * q] += (double)x.x * wr + (double)x.y * wi;
                v[2 * q + 1] += (double)x.x * wi - (double)x.y * wr;
            }
        }
    }
    // partial layout: partials[chunk * (2k) + 2m + {0,1}]
    __shared__ double sh[2 * MV][kThreads];
#pragma unroll
    for (int q = 0; q < 2 * MV; ++q) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int q = 0; q < 2 * MV; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x < 2 * MV) {
        const int q = threadIdx.x / 2;
        if (m0 + q < k) partials[(size_t)blockIdx.x * (2 * k) + 2 * (m0 + q) + (threadIdx.x & 1)] = sh[threadIdx.x][0];
    }
}

// w = (overwrite ? 0 : w) + sign * sum_m c[m] V_m
template <typename R>
__global__ __launch_bounds__(kThreads) void k_multi_axpy(const typename Vec2<R>::T *__restrict__ base, int k,
                                                         const double *__restrict__ c, double sign,
                                                         typename Vec2<R>::T *__restrict__ w, uint64_t dim,
                                                         int overwrite) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim; i += stride) {
        double sr = 0.0, si = 0.0;
        for (int m = 0; m < k; ++m) {
            const typename Vec2<R>::T x = base[(uint64_t)m * dim + i];
            const double cr = c[2 * m], ci = c[2 * m + 1];
            sr += cr * x.x - ci * x.y;
            si += cr * x.y + ci * x.x;
        }
        double wr = 0.0, wi = 0.0;
        if (!overwrite) {
            wr = w[i].x;
            wi = w[i].y;
        }
        typename Vec2<R>::T o;
        o.x = (R)(wr + sign * sr);
        o.y = (R)(wi + sign * si);
        w[i] = o;
    }
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed by the seed,
// counter = element index; two 53-bit uniforms -> Box-Muller, N(0, 1/2).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int rnd = 0; rnd < 10; ++rnd) {
        const uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += W0; k1 += W1;
    }
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_gaussian(uint64_t seed, typename Vec2<R>::T *__restrict__ out,
                                                       uint64_t dim) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim; i += stride) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0u, 0u};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const uint64_t a = ((uint64_t)c[0] << 21) ^ (uint64_t)(c[1] >> 11);
        const uint64_t b = ((uint64_t)c[2] << 21) ^ (uint64_t)(c[3] >> 11);
        const double u1 = ((double)(a & ((1ull << 53) - 1)) + 0.5) * 0x1p-53;
        const double u2 = ((double)(b & ((1ull << 53) - 1)) + 0.5) * 0x1p-53;
        const double rad = sqrt(-log(u1));
        const double th = 6.283185307179586476925286766559 * u2;
        typename Vec2<R>::T o;
        o.x = (R)(rad * cos(th));
        o.y = (R)(rad * sin(th));
        out[i] = o;
    }
}

template <typename RS, typename RD>
__global__ __launch_bounds__(kThreads) void k_convert(const typename Vec2<RS>::T *__restrict__ src,
                                                      typename Vec2<RD>::T *__restrict__ dst, uint64_t dim) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim; i += stride) {
        typename Vec2<RD>::T o;
        o.x = (RD)src[i].x;
        o.y = (RD)src[i].y;
        dst[i] = o;
    }
}

// ---------------------------------------------------------------------------
// Batched Lanczos steps over many independent systems (one flattened grid):
// block b belongs to the instance i with lanes[i].blk0 <= b < blk0 + nblk and
// walks that instance's rows with a stride of nblk * kThreads.  The step index
// k is read from device memory (one counter for the batch), so the same
// captured step can be replayed; alpha / beta / coefficients are per-instance
// rows of device arrays.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int batch_instance(const BatchLane *__restrict__ L, int n, uint32_t b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (L[mid].blk0 <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_b_dot(const BatchLane *__restrict__ L, int n, double *__restrict__ partials) {
    using V2 = typename Vec2<R>::T;
    const BatchLane &l = L[batch_instance(L, n, blockIdx.x)];
    if (l.ndot > 0) return;   // the whole block: the H.psi kernel formed this lane's dot
    const V2 *x = (const V2 *)l.v, *y = (const V2 *)l.w;
    double v[2] = {0.0, 0.0};
    const uint64_t stride = (uint64_t)l.nblk * kThreads;
    for (uint64_t i = (uint64_t)(blockIdx.x - l.blk0) * kThreads + threadIdx.x; i < l.dim; i += stride) {
        const double xr = x[i].x, xi = x[i].y, yr = y[i].x, yi = y[i].y;
        v[0] += xr * yr + xi * yi;
        v[1] += xr * yi - xi * yr;
    }
    block_reduce_store<2>(v, partials + 2 * blockIdx.x);
}

// The batched recursion keeps unnormalised vectors u_k, v_k = c_k u_k with
// c_0 = 1 (unit start vectors) and c_k = 1/beta_k (0 once the Krylov space is
// exhausted), so no pass rescales w after its norm is known.
__device__ __forceinline__ double batch_scale(const double *__restrict__ beta_row, uint32_t k) {
    if (k == 0) return 1.0;
    const double b = beta_row[k];
    return b > 0.0 ? 1.0 / b : 0.0;
}

// one block per instance: alpha[i][k] = c_k^2 <u_k|H u_k>, the sum of its
// partials (fixed order)
__global__ __launch_bounds__(kThreads) void k_b_alpha(const BatchLane *__restrict__ L, const double *__restrict__ partials,
                                                      double *__restrict__ alpha, int lda,
                                                      const double *__restrict__ beta, int ldb,
                                                      const uint32_t *__restrict__ step) {
    const BatchLane &l = L[blockIdx.x];
    __shared__ double tot[2];
    double v[2] = {0.0, 0.0};
    if (l.ndot > 0) {
        for (uint32_t p = threadIdx.x; p < l.ndot; p += kThreads) {
            v[0] += l.dotp[2 * (size_t)p];
            v[1] += l.dotp[2 * (size_t)p + 1];
        }
    } else {
        for (uint32_t p = threadIdx.x; p < l.nblk; p += kThreads) {
            v[0] += partials[2 * (size_t)(l.blk0 + p)];
            v[1] += partials[2 * (size_t)(l.blk0 + p) + 1];
        }
    }
    block_reduce_store<2>(v, tot);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t k = *step;
        const double c = batch_scale(beta + (size_t)blockIdx.x * ldb, k);
        double *a = alpha + 2 * ((size_t)blockIdx.x * lda + k);
        a[0] = c * c * tot[0];
        a[1] = c * c * tot[1];
    }
}

// w = c_k w - alpha_k c_k v - beta_k c_{k-1} vp (w holds H u_k); partial |w|^2
// Streamed loads and stores (non-temporal hint) for the vector passes that
// touch each element once.
template <bool NT, typename T> __device__ __forceinline__ T ld_s(const T *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename T> __device__ __forceinline__ void st_s(T *p, T v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}
constexpr int kUnroll = 4;   // elements per lane in flight in the batched vector passes

template <typename R, bool NT>
__global__ __launch_bounds__(kThreads) void k_b_update(const BatchLane *__restrict__ L, int n,
                                                       const double *__restrict__ alpha, int lda,
                                                       const double *__restrict__ beta, int ldb,
                                                       const uint32_t *__restrict__ step, double *__restrict__ partials) {
    const int inst = batch_instance(L, n, blockIdx.x);
    const BatchLane &l = L[inst];
    const uint32_t k = *step;
    const double *brow = beta + (size_t)inst * ldb;
    const double c = batch_scale(brow, k), cp = k > 0 ? batch_scale(brow, k - 1) : 0.0;
    const double ar = alpha[2 * ((size_t)inst * lda + k)] * c, ai = alpha[2 * ((size_t)inst * lda + k) + 1] * c;
    const double b = brow[k] * cp;
    double acc[1] = {0.0};
    // kUnroll elements per lane per round, kThreads apart (coalesced)
    const uint64_t span = (uint64_t)kThreads * kUnroll;
    const uint64_t b0 = (uint64_t)(blockIdx.x - l.blk0) * span, bstride = (uint64_t)l.nblk * span;
    if constexpr (sizeof(R) == 4) {
        // complex64: two elements per lane access (16 bytes); dim is a power
        // of two >= 2
        typedef float F4 __attribute__((ext_vector_type(4)));
        const F4 *v = (const F4 *)l.v, *u = (const F4 *)l.vp;
        F4 *w = (F4 *)l.w;
        const uint64_t n4 = l.dim / 2;
        for (uint64_t base = b0; base < n4; base += bstride) {
            F4 x[kUnroll], y[kUnroll], z[kUnroll];
#pragma unroll
            for (int q = 0; q < kUnroll; ++q) {
                const uint64_t i = base + (uint64_t)q * kThreads + threadIdx.x;
                if (i < n4) {
                    x[q] = ld_s<NT>(v + i);
                    y[q] = ld_s<NT>(w + i);
                    z[q] = ld_s<NT>(u + i);
                }
            }
#pragma unroll
            for (int q = 0; q < kUnroll; ++q) {
                const uint64_t i = base + (uint64_t)q * kThreads + threadIdx.x;
                if (i >= n4) continue;
                const double r0 = c * (double)y[q].x - (ar * (double)x[q].x - ai * (double)x[q].y) - b * (double)z[q].x;
                const double i0 = c * (double)y[q].y - (ar * (double)x[q].y + ai * (double)x[q].x) - b * (double)z[q].y;
                const double r1 = c * (double)y[q].z - (ar * (double)x[q].z - ai * (double)x[q].w) - b * (double)z[q].z;
                const double i1 = c * (double)y[q].w - (ar * (double)x[q].w + ai * (double)x[q].z) - b * (double)z[q].w;
                F4 o;
                o.x = (float)r0;
                o.y = (float)i0;
                o.z = (float)r1;
                o.w = (float)i1;
                st_s<NT>(w + i, o);
                acc[0] += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
                acc[0] += (double)o.z * (double)o.z + (double)o.w * (double)o.w;
            }
        }
    } else {
        typedef double D2 __attribute__((ext_vector_type(2)));
        const D2 *v = (const D2 *)l.v, *u = (const D2 *)l.vp;
        D2 *w = (D2 *)l.w;
        for (uint64_t base = b0; base < l.dim; base += bstride) {
            D2 x[kUnroll], y[kUnroll], z[kUnroll];
#pragma unroll
            for (int q = 0; q < kUnroll; ++q) {
                const uint64_t i = base + (uint64_t)q * kThreads + threadIdx.x;
                if (i < l.dim) {
                    x[q] = ld_s<NT>(v + i);
                    y[q] = ld_s<NT>(w + i);
                    z[q] = ld_s<NT>(u + i);
                }
            }
#pragma unroll
            for (int q = 0; q < kUnroll; ++q) {
                const uint64_t i = base + (uint64_t)q * kThreads + threadIdx.x;
                if (i >= l.dim) continue;
                const double vr = x[q].x, vi = x[q].y;
                const double wr = c * y[q].x - (ar * vr - ai * vi) - b * z[q].x;
                const double wi = c * y[q].y - (ar * vi + ai * vr) - b * z[q].y;
                D2 o;
                o.x = wr;
                o.y = wi;
                st_s<NT>(w + i, o);
                acc[0] += wr * wr + wi * wi;
            }
        }
    }
    block_reduce_store<1>(acc, partials + blockIdx.x);
}

// one block per instance: beta[i][k+1] = |w|, or 0 when the Krylov space is
// exhausted (|w| at rounding level of |alpha_k|) or the value is not finite
__global__ __launch_bounds__(kThreads) void k_b_beta(const BatchLane *__restrict__ L, const double *__restrict__ partials,
                                                     const double *__restrict__ alpha, int lda,
                                                     double *__restrict__ beta, int ldb,
                                                     const uint32_t *__restrict__ step) {
    const BatchLane &l = L[blockIdx.x];
    __shared__ double tot[1];
    double v[1] = {0.0};
    for (uint32_t p = threadIdx.x; p < l.nblk; p += kThreads) v[0] += partials[l.blk0 + p];
    block_reduce_store<1>(v, tot);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t k = *step;
        const double a = fabs(alpha[2 * ((size_t)blockIdx.x * lda + k)]);
        double b = sqrt(tot[0]);
        if (!(b > 1e-13 * (a > 1.0 ? a : 1.0)) || !isfinite(b)) b = 0.0;
        beta[(size_t)blockIdx.x * ldb + k + 1] = b;
    }
}

// acc += coef[i][k] v_k = coef[i][k] c_k u_k  (replay of a recursion forming
// a Ritz vector)
template <typename R, bool NT>
__global__ __launch_bounds__(kThreads) void k_b_accum(const BatchLane *__restrict__ L, int n,
                                                      const double *__restrict__ coef, int ldc,
                                                      const double *__restrict__ beta, int ldb,
                                                      const uint32_t *__restrict__ step) {
    typedef R RV __attribute__((ext_vector_type(2)));
    const int inst = batch_instance(L, n, blockIdx.x);
    const BatchLane &l = L[inst];
    const double c = coef[(size_t)inst * ldc + *step] * batch_scale(beta + (size_t)inst * ldb, *step);
    const RV *v = (const RV *)l.v;
    RV *a = (RV *)l.acc;
    const uint64_t span = (uint64_t)kThreads * kUnroll;
    const uint64_t b0 = (uint64_t)(blockIdx.x - l.blk0) * span, bstride = (uint64_t)l.nblk * span;
    for (uint64_t base = b0; base < l.dim; base += bstride) {
        RV x[kUnroll], y[kUnroll];
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) {
            const uint64_t i = base + (uint64_t)q * kThreads + threadIdx.x;
            if (i < l.dim) {
                x[q] = ld_s<NT>(v + i);
                y[q] = ld_s<NT>(a + i);
            }
        }
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) {
            const uint64_t i = base + (uint64_t)q * kThreads + threadIdx.x;
            if (i >= l.dim) continue;
            RV o;
            o.x = (R)((double)y[q].x + c * (double)x[q].x);
            o.y = (R)((double)y[q].y + c * (double)x[q].y);
            st_s<NT>(a + i, o);
        }
    }
}

__global__ void k_b_next_step(uint32_t *step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1u;
}

inline uint32_t row_blocks(uint64_t dim) { return (uint32_t)((dim + kThreads - 1) / kThreads); }
inline uint32_t stream_blocks(uint64_t dim) {
    uint64_t b = (dim + kThreads - 1) / kThreads;
    return (uint32_t)(b < 8192 ? b : 8192);
}

// A launch covers at most kMaxGrid blocks; larger row counts (N = 32) are
// split into several launches that share one XCD mapping.
constexpr uint32_t kMaxGrid = 1u << 22;

}  // namespace

template <typename R> using V2p = typename Vec2<R>::T *;
template <typename R> using CV2p = const typename Vec2<R>::T *;

template <typename R>
static hipError_t launch_rows_hpsi(const HTables *t, const void *in, void *out, uint64_t dim, hipStream_t s) {
    const uint32_t nblk = row_blocks(dim);
    for (uint32_t base = 0; base < nblk; base += kMaxGrid) {
        const uint32_t n = (nblk - base) < kMaxGrid ? (nblk - base) : kMaxGrid;
        hipLaunchKernelGGL(k_hpsi<R>, dim3(n), dim3(kThreads), 0, s, t,
                           (const typename Vec2<R>::T *)in, (typename Vec2<R>::T *)out, dim, nblk, base);
    }
    return hipGetLastError();
}

template <typename R>
static hipError_t launch_rows_current(const CurrentTable *t, const void *in, void *out, uint64_t dim, hipStream_t s) {
    const uint32_t nblk = row_blocks(dim);
    for (uint32_t base = 0; base < nblk; base += kMaxGrid) {
        const uint32_t n = (nblk - base) < kMaxGrid ? (nblk - base) : kMaxGrid;
        hipLaunchKernelGGL(k_current<R>, dim3(n), dim3(kThreads), 0, s, t,
                           (const typename Vec2<R>::T *)in, (typename Vec2<R>::T *)out, dim, nblk, base);
    }
    return hipGetLastError();
}

template <typename R, int TB, int WPB>
static hipError_t launch_tile_tb(const GTables *g, const HTables *h, const void *in, void *out, double *dot_part,
                                 uint32_t ntiles, size_t lds, hipStream_t s) {
    // tiles above 64 KiB need the raised dynamic-LDS limit; raised once to
    // the full 160 KiB, outside any stream capture (the first launch of a
    // shape is an eager one)
    static bool raised = false;
    if (lds > 65536 && !raised) {
        hipError_t e = hipFuncSetAttribute((const void *)k_hpsi_tile<R, TB, WPB>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxTileLds);
        if (e != hipSuccess) return e;
        raised = true;
    }
    for (uint32_t base = 0; base < ntiles; base += kMaxGrid) {
        const uint32_t n = (ntiles - base) < kMaxGrid ? (ntiles - base) : kMaxGrid;
        hipLaunchKernelGGL((k_hpsi_tile<R, TB, WPB>), dim3(n), dim3(WPB * 64), lds, s, g, h, (CV2p<R>)in,
                           (V2p<R>)out, dot_part, ntiles, base);
    }
    return hipGetLastError();
}

size_t tile_lds_bytes(bool c64, int tile_bits, int n_uniform, int n_terms) {
    const size_t esz = c64 ? 8 : 16;
    return esz * (((size_t)1 << tile_bits) + (size_t)(n_terms - n_uniform) * 8) +
           (size_t)(n_terms - n_uniform) * 64 * sizeof(uint32_t);
}

// each wave holds 2^(TB-6)/WPB row groups in registers; beyond 8 groups the
// unrolled term loop spills
bool tile_shape_supported(bool c64, int tile_bits, int waves) {
    if (waves != 4 && waves != 8 && waves != 16) return false;
    if (tile_bits < 10 || tile_bits > 13) return false;
    const int groups = (1 << (tile_bits - 6)) / waves;
    return groups >= 2 && groups <= 8;
}

hipError_t launch_hpsi_tile(bool c64, const GTables *g, const HTables *h, int tile_bits, int waves, int n_uniform,
                            int n_terms, const void *in, void *out, double *dot_part, uint64_t dim, hipStream_t s) {
    const uint64_t ntiles64 = dim >> tile_bits;
    if (ntiles64 == 0 || ntiles64 > 0x7fffffffull || !tile_shape_supported(c64, tile_bits, waves))
        return hipErrorInvalidValue;
    const uint32_t ntiles = (uint32_t)ntiles64;
    const size_t lds = tile_lds_bytes(c64, tile_bits, n_uniform, n_terms);
    if (lds > kMaxTileLds) return hipErrorInvalidValue;
#define ESC_TILE_CASE(TBV, WV)                                                                          \
    if (tile_bits == TBV && waves == WV)                                                                \
        return c64 ? launch_tile_tb<float, TBV, WV>(g, h, in, out, dot_part, ntiles, lds, s)            \
                   : launch_tile_tb<double, TBV, WV>(g, h, in, out, dot_part, ntiles, lds, s);
    ESC_TILE_CASE(10, 4)
    ESC_TILE_CASE(11, 4)
    ESC_TILE_CASE(10, 8)
    ESC_TILE_CASE(11, 8)
    ESC_TILE_CASE(12, 8)
    ESC_TILE_CASE(11, 16)
    ESC_TILE_CASE(12, 16)
    ESC_TILE_CASE(13, 16)
#undef ESC_TILE_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_hpsi_c128(const HTables *t, const void *in, void *out, uint64_t dim, hipStream_t s) {
    return launch_rows_hpsi<double>(t, in, out, dim, s);
}
hipError_t launch_hpsi_c64(const HTables *t, const void *in, void *out, uint64_t dim, hipStream_t s) {
    return launch_rows_hpsi<float>(t, in, out, dim, s);
}
hipError_t launch_current_c128(const CurrentTable *t, const void *in, void *out, uint64_t dim, hipStream_t s) {
    return launch_rows_current<double>(t, in, out, dim, s);
}
hipError_t launch_current_c64(const CurrentTable *t, const void *in, void *out, uint64_t dim, hipStream_t s) {
    return launch_rows_current<float>(t, in, out, dim, s);
}

#define ESC_DISPATCH(c64, KERNEL, ...)                                  \
    do {                                                                \
        if (c64) hipLaunchKernelGGL(KERNEL<float>, __VA_ARGS__);        \
        else hipLaunchKernelGGL(KERNEL<double>, __VA_ARGS__);           \
    } while (0)


hipError_t launch_dot(bool c64, const void *x, const void *y, uint64_t dim, double *partials, double *result,
                      hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL(k_dot_partial<float>, dim3(kRedBlocks), dim3(kThreads), 0, s, (CV2p<float>)x,
                           (CV2p<float>)y, dim, partials);
    else
        hipLaunchKernelGGL(k_dot_partial<double>, dim3(kRedBlocks), dim3(kThreads), 0, s, (CV2p<double>)x,
                           (CV2p<double>)y, dim, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(2), dim3(kThreads), 0, s, partials, kRedBlocks, 2, result);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Reduction workspace that needs no allocation: one fixed device array
// (kRedBlocks value pairs, 16 KiB), held by each GPU the library runs on until
// the process ends.  Host code never passes the array's name to a kernel: its
// device address is resolved once per device (scratch_init, with that device
// current) and kept in g_scratch_ptr.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[2 * kRedBlocks];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

hipError_t scratch_init() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return hipSuccess;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
    return hipSuccess;
}

hipError_t launch_dot_noalloc(bool c64, const void *x, const void *y, uint64_t dim, double *result, hipStream_t s) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
    void *p;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        p = g_scratch_ptr[dev];
    }
    if (!p) {
        if ((e = scratch_init()) != hipSuccess) return e;
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        p = g_scratch_ptr[dev];
    }
    // every partial pair is written by k_dot_partial before k_reduce_final
    // reads it, so nothing needs clearing between calls
    return launch_dot(c64, x, y, dim, static_cast<double *>(p), result, s);
}

hipError_t launch_norm2(bool c64, const void *x, uint64_t dim, double *partials, double *result, hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL(k_norm2_partial<float>, dim3(kRedBlocks), dim3(kThreads), 0, s, (CV2p<float>)x, dim,
                           partials);
    else
        hipLaunchKernelGGL(k_norm2_partial<double>, dim3(kRedBlocks), dim3(kThreads), 0, s, (CV2p<double>)x, dim,
                           partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kThreads), 0, s, partials, kRedBlocks, 1, result);
    return hipGetLastError();
}

hipError_t launch_axpby(bool c64, double ar, double ai, const void *x, double br, double bi, void *y, uint64_t dim,
                        hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL(k_axpby<float>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, ar, ai, (CV2p<float>)x, br,
                           bi, (V2p<float>)y, dim);
    else
        hipLaunchKernelGGL(k_axpby<double>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, ar, ai, (CV2p<double>)x,
                           br, bi, (V2p<double>)y, dim);
    return hipGetLastError();
}

hipError_t launch_rec_init(double *rec, uint32_t *step, double c0, hipStream_t s) {
    hipLaunchKernelGGL(k_rec_init, dim3(1), dim3(1), 0, s, rec, step, c0);
    return hipGetLastError();
}

hipError_t launch_lanczos_update_rec(bool c64, void *w, const void *v, const void *u, const double *rec,
                                     const double *dres, const double *hist, const uint32_t *step, uint64_t dim,
                                     double *partials, hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL((k_lanczos_update_rec<float, false>), dim3(kRedBlocks), dim3(kThreads), 0, s,
                           (V2p<float>)w, (CV2p<float>)v, (CV2p<float>)u, rec, dres, hist, step, dim, partials);
    else
        hipLaunchKernelGGL((k_lanczos_update_rec<double, true>), dim3(kRedBlocks), dim3(kThreads), 0, s,
                           (V2p<double>)w, (CV2p<double>)v, (CV2p<double>)u, rec, dres, hist, step, dim, partials);
    return hipGetLastError();
}

hipError_t launch_rec_step(double *rec, double *dres, double *hist, uint32_t *step, uint32_t cap,
                           const double *partials, hipStream_t s) {
    hipLaunchKernelGGL(k_rec_step, dim3(1), dim3(kThreads), 0, s, rec, dres, hist, step, cap, partials, kRedBlocks);
    return hipGetLastError();
}

hipError_t launch_axpy_rec_step(bool c64, const double *coef, const uint32_t *step, const double *rec, const void *x,
                                void *y, uint64_t dim, hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL(k_axpy_rec_step<float>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, coef, step, rec,
                           (CV2p<float>)x, (V2p<float>)y, dim);
    else
        hipLaunchKernelGGL(k_axpy_rec_step<double>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, coef, step, rec,
                           (CV2p<double>)x, (V2p<double>)y, dim);
    return hipGetLastError();
}

hipError_t launch_rec_small(bool c64, const HTables *t, void *b0, void *b1, void *b2, double *rec, uint32_t *step,
                            double *hist, uint32_t cap, int nsteps, const double *coef, void *acc, uint64_t dim,
                            hipStream_t s) {
    if (dim > kSmallRecMaxDim || nsteps < 1 || (coef != nullptr && acc == nullptr)) return hipErrorInvalidValue;
    // rows per lane: 1 up to kSmallThreads rows, else dim / kSmallThreads
    // (dim is a power of two)
    const int nr = dim <= (uint64_t)kSmallThreads ? 1 : (int)(dim / kSmallThreads);
#define ESC_REC_SMALL(R, NRV)                                                                                     \
    hipLaunchKernelGGL((k_rec_small<R, NRV>), dim3(1), dim3(kSmallThreads), 0, s, t, (V2p<R>)b0, (V2p<R>)b1,      \
                       (V2p<R>)b2, rec, step, hist, cap, nsteps, coef, (V2p<R>)acc, (uint32_t)dim)
#define ESC_REC_SMALL_NR(R)                                   \
    switch (nr) {                                             \
        case 1: ESC_REC_SMALL(R, 1); break;                   \
        case 2: ESC_REC_SMALL(R, 2); break;                   \
        default: return hipErrorInvalidValue;                 \
    }
    if (c64) {
        ESC_REC_SMALL_NR(float)
    } else {
        ESC_REC_SMALL_NR(double)
    }
#undef ESC_REC_SMALL_NR
#undef ESC_REC_SMALL
    return hipGetLastError();
}

hipError_t launch_reduce(const double *partials, int npart, int nvals, double *result, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_final, dim3(nvals), dim3(kThreads), 0, s, partials, npart, nvals, result);
    return hipGetLastError();
}

hipError_t launch_lanczos_update(bool c64, void *w, const void *v, const void *u, const double *d_alpha, double beta,
                                 uint64_t dim, double *partials, double *result, hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL(k_lanczos_update<float>, dim3(kRedBlocks), dim3(kThreads), 0, s, (V2p<float>)w,
                           (CV2p<float>)v, (CV2p<float>)u, d_alpha, beta, dim, partials);
    else
        hipLaunchKernelGGL(k_lanczos_update<double>, dim3(kRedBlocks), dim3(kThreads), 0, s, (V2p<double>)w,
                           (CV2p<double>)v, (CV2p<double>)u, d_alpha, beta, dim, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kThreads), 0, s, partials, kRedBlocks, 1, result);
    return hipGetLastError();
}

hipError_t launch_normalize(bool c64, void *x, const double *d_norm2, uint64_t dim, hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL(k_normalize<float>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, (V2p<float>)x, d_norm2,
                           dim);
    else
        hipLaunchKernelGGL(k_normalize<double>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, (V2p<double>)x,
                           d_norm2, dim);
    return hipGetLastError();
}

hipError_t launch_scale(bool c64, void *x, double sc, uint64_t dim, hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL(k_scale<float>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, (V2p<float>)x, sc, dim);
    else
        hipLaunchKernelGGL(k_scale<double>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, (V2p<double>)x, sc, dim);
    return hipGetLastError();
}

hipError_t launch_multidot(bool c64, const void *base, int k, const void *w, uint64_t dim, double *partials,
                           double *result, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    dim3 grid(kMultiChunks, (k + kMultiVecs - 1) / kMultiVecs);
    if (c64)
        hipLaunchKernelGGL(k_multidot_partial<float>, grid, dim3(kThreads), 0, s, (CV2p<float>)base, k,
                           (CV2p<float>)w, dim, partials);
    else
        hipLaunchKernelGGL(k_multidot_partial<double>, grid, dim3(kThreads), 0, s, (CV2p<double>)base, k,
                           (CV2p<double>)w, dim, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(2 * k), dim3(kThreads), 0, s, partials, kMultiChunks, 2 * k, result);
    return hipGetLastError();
}

hipError_t launch_multi_axpy(bool c64, const void *base, int k, const double *d_coef, double sign, void *w,
                             uint64_t dim, bool overwrite, hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL(k_multi_axpy<float>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, (CV2p<float>)base, k,
                           d_coef, sign, (V2p<float>)w, dim, (int)overwrite);
    else
        hipLaunchKernelGGL(k_multi_axpy<double>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, (CV2p<double>)base,
                           k, d_coef, sign, (V2p<double>)w, dim, (int)overwrite);
    return hipGetLastError();
}

hipError_t launch_gaussian(bool c64, uint64_t seed, void *out, uint64_t dim, hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL(k_gaussian<float>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, seed, (V2p<float>)out,
                           dim);
    else
        hipLaunchKernelGGL(k_gaussian<double>, dim3(stream_blocks(dim)), dim3(kThreads), 0, s, seed,
                           (V2p<double>)out, dim);
    return hipGetLastError();
}

hipError_t launch_convert_from_c128(bool c64, const void *src, void *dst, uint64_t dim, hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(stream_blocks(dim)), dim3(kThreads), 0, s,
                           (CV2p<double>)src, (V2p<float>)dst, dim);
    else
        hipLaunchKernelGGL((k_convert<double, double>), dim3(stream_blocks(dim)), dim3(kThreads), 0, s,
                           (CV2p<double>)src, (V2p<double>)dst, dim);
    return hipGetLastError();
}

hipError_t launch_convert_to_c128(bool c64, const void *src, void *dst, uint64_t dim, hipStream_t s) {
    if (c64)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(stream_blocks(dim)), dim3(kThreads), 0, s,
                           (CV2p<float>)src, (V2p<double>)dst, dim);
    else
        hipLaunchKernelGGL((k_convert<double, double>), dim3(stream_blocks(dim)), dim3(kThreads), 0, s,
                           (CV2p<double>)src, (V2p<double>)dst, dim);
    return hipGetLastError();
}

// ---- batched Lanczos step pieces ----
hipError_t launch_batch_step(bool c64, int piece, const BatchLane *d_lanes, int n, uint32_t total_blocks,
                             double *partials, double *alpha, int lda, double *beta, int ldb, const double *coef,
                             int ldc, uint32_t *step, hipStream_t s) {
    const dim3 g(total_blocks), gi((uint32_t)n), t(kThreads);
    switch (piece) {
    case kBatchDot:
        if (c64) hipLaunchKernelGGL(k_b_dot<float>, g, t, 0, s, d_lanes, n, partials);
        else hipLaunchKernelGGL(k_b_dot<double>, g, t, 0, s, d_lanes, n, partials);
        hipLaunchKernelGGL(k_b_alpha, gi, t, 0, s, d_lanes, (const double *)partials, alpha, lda,
                           (const double *)beta, ldb, (const uint32_t *)step);
        break;
    case kBatchUpdate:
        // streamed accesses: each element is read and written once per step
        if (c64)
            hipLaunchKernelGGL((k_b_update<float, true>), g, t, 0, s, d_lanes, n, (const double *)alpha, lda,
                               (const double *)beta, ldb, (const uint32_t *)step, partials);
        else
            hipLaunchKernelGGL((k_b_update<double, true>), g, t, 0, s, d_lanes, n, (const double *)alpha, lda,
                               (const double *)beta, ldb, (const uint32_t *)step, partials);
        hipLaunchKernelGGL(k_b_beta, gi, t, 0, s, d_lanes, (const double *)partials, (const double *)alpha, lda, beta,
                           ldb, (const uint32_t *)step);
        hipLaunchKernelGGL(k_b_next_step, dim3(1), dim3(64), 0, s, step);
        break;
    case kBatchAccum:
        if (c64)
            hipLaunchKernelGGL((k_b_accum<float, true>), g, t, 0, s, d_lanes, n, coef, ldc, (const double *)beta, ldb,
                               (const uint32_t *)step);
        else
            hipLaunchKernelGGL((k_b_accum<double, true>), g, t, 0, s, d_lanes, n, coef, ldc, (const double *)beta,
                               ldb, (const uint32_t *)step);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---- the tridiagonal matrices and continued fractions of a batch ----
// One thread per instance: the recursion's length, the lowest eigenvalue of
// its T and the eigenvector (the replay coefficients, zero beyond the length).
__global__ __launch_bounds__(64) void k_batch_tridiag(int n, const double *__restrict__ alpha, int lda,
                                                      const double *__restrict__ beta, int ldb, int steps,
                                                      double *__restrict__ work, double *__restrict__ coef, int ldc,
                                                      double *__restrict__ e0, int *__restrict__ iters) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *bt = beta + (size_t)i * ldb;
    const int len = td_length(bt, steps);
    double *a = work + (size_t)i * 6 * steps;
    for (int k = 0; k < len; ++k) a[k] = alpha[2 * ((size_t)i * lda + k)];
    bool ok = false;
    const double lam = td_lowest(len, a, bt, &ok);
    double *x = coef + (size_t)i * ldc;
    if (ok) td_vector(len, a, bt, lam, a + steps, x);
    else
        for (int k = 0; k < len; ++k) x[k] = 0.0;
    for (int k = len; k < steps; ++k) x[k] = 0.0;
    e0[i] = ok ? lam : NAN;
    iters[i] = len;
}

// Continued fractions of every instance on every omega: grid (omega blocks,
// instances); f[i * nw + q] = n2_i <phi_i|(z - H_i)^-1|phi_i> / n2_i at
// z = omega_q + e0_i - i eta (no recursion when n2_i is not positive).
__global__ __launch_bounds__(256) void k_batch_cf(const double *__restrict__ alpha, int lda,
                                                  const double *__restrict__ beta, int ldb, int steps,
                                                  const double *__restrict__ n2, const double *__restrict__ e0,
                                                  double eta, const double *__restrict__ omega, int nw,
                                                  double2 *__restrict__ f) {
    const int i = blockIdx.y;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nw) return;
    const double *bt = beta + (size_t)i * ldb;
    const double nn = n2[i];
    const int it = nn > 0.0 ? td_length(bt, steps) : 0;
    double fr, fi;
    td_cf(it, alpha + 2 * (size_t)i * lda, 2, bt, nn, omega[q] + e0[i], -eta, &fr, &fi);
    f[(size_t)i * nw + q] = make_double2(fr, fi);
}

hipError_t launch_batch_tridiag(int n, const double *alpha, int lda, const double *beta, int ldb, int steps,
                                double *work, double *coef, int ldc, double *e0, int *iters, hipStream_t s) {
    if (n < 1 || steps < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_batch_tridiag, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, n, alpha, lda, beta, ldb,
                       steps, work, coef, ldc, e0, iters);
    return hipGetLastError();
}

hipError_t launch_batch_cf(int n, const double *alpha, int lda, const double *beta, int ldb, int steps,
                           const double *n2, const double *e0, double eta, const double *omega, int nw, void *f,
                           hipStream_t s) {
    if (n < 1 || nw < 1 || n > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_batch_cf, dim3((unsigned)((nw + 255) / 256), (unsigned)n), dim3(256), 0, s, alpha, lda, beta,
                       ldb, steps, n2, e0, eta, omega, nw, (double2 *)f);
    return hipGetLastError();
}

// ---- all eigenpairs of one tridiagonal matrix (n <= kMaxTridiag) ----
// One workgroup: every thread bisects for its eigenvalues, then takes the
// clusters that start at its indices and computes their vectors in order.
__global__ __launch_bounds__(256) void k_tridiag_eig(int n, const double *__restrict__ a,
                                                     const double *__restrict__ b, double *__restrict__ ev,
                                                     double *__restrict__ Z, double *__restrict__ work,
                                                     int *__restrict__ status) {
    __shared__ double s_ev[kMaxTridiag];
    __shared__ double s_lo, s_hi, s_piv, s_tn;
    __shared__ int s_ok;
    if (threadIdx.x == 0) {
        s_ok = td_range(n, a, b, &s_lo, &s_hi, &s_piv) ? 1 : 0;
        s_tn = td_eig_norm(n, a, b);
        status[0] = s_ok ? 0 : 1;
    }
    __syncthreads();
    if (!s_ok) return;
    for (int k = threadIdx.x; k < n; k += blockDim.x) s_ev[k] = n == 1 ? a[0] : td_kth(n, a, b, k, s_lo, s_hi, s_piv);
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) ev[k] = s_ev[k];
    for (int j = threadIdx.x; j < n; j += blockDim.x)
        if (td_eig_cluster_start(s_ev, j, s_tn))
            td_eig_vectors(n, a, b, s_ev, j, s_tn, work + (size_t)threadIdx.x * 5 * n, Z);
}

hipError_t launch_tridiag_eig(int n, const double *a, const double *b, double *ev, double *Z, double *work,
                              int *status, hipStream_t s) {
    if (n < 1 || n > kMaxTridiag) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_tridiag_eig, dim3(1), dim3(kTridiagThreads), 0, s, n, a, b, ev, Z, work, status);
    return hipGetLastError();
}

}  // namespace esc
