/*
 * esc.h -- C ABI of the MI355X Lanczos dynamical-correlation engine.
 *
 * Plain C types only: device buffers are passed as `void *` (caller-owned device
 * memory holding interleaved complex numbers, complex<double> for ESC_C128
 * contexts and complex<float> for ESC_C64), host arrays as typed pointers,
 * streams as `void *` (a hipStream_t; NULL = the default stream).
 *
 * Every call returns ESC_OK (0) or a negative status and then sets a
 * thread-local message readable with esc_last_error().  No C++ exception
 * crosses this boundary.  Calls on one context must not run concurrently;
 * distinct contexts are independent.  Entry points that evaluate continued
 * fractions on an omega grid (esc_dynamics, esc_rixs, esc_batch_dynamics)
 * spread that host loop over up to 16 threads that they create and join
 * before returning.
 *
 * Basis convention (as the reference): index r has bit k = site k and bit value
 * 0 = spin up (S^z = +1/2), S = sigma/2 (src/QuantumBasis.h:34-74).
 *
 * Reference interfaces replaced (paths relative to the fengshi96/EnergySpinCurrents
 * root) are named on each entry point.
 */
#ifndef ESC_H
#define ESC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ESC_OK               0
#define ESC_ERR_ARG         -1   /* invalid argument (std::invalid_argument in the reference) */
#define ESC_ERR_HIP         -2   /* HIP runtime failure */
#define ESC_ERR_ALLOC       -3   /* device or host allocation failed */
#define ESC_ERR_BREAKDOWN   -4   /* Lanczos beta == 0 (invariant subspace) */
#define ESC_ERR_NUMERIC     -5   /* non-finite input or eigensolver failure (runtime_error) */
#define ESC_ERR_UNSUPPORTED -6   /* model shape the kernels do not handle */

/* Vector storage precision of a context.  Scalars (alpha, beta, dots, norms,
 * tridiagonal solve, spectra) are always FP64. */
#define ESC_C128 0
#define ESC_C64  1

typedef struct esc_ctx esc_ctx;

/* Spin model on an explicit bond graph.
 * Replaces the data Hamiltonian::SetupConnectors/Connectors derives
 * (src/Hamiltonian.h:27-32,44-54,230-328) plus Lattice::N1neigh_
 * (src/Lattice.h:29).  Bond family f = 0 (x), 1 (y), 2 (z) carries the Kitaev
 * term K[f] S^f_i S^f_j.  Extensions with no reference counterpart: Heisenberg J
 * (all bonds) and Gamma (S^a_i S^b_j + S^b_i S^a_j on a f-bond, {a,b} the
 * other two axes).  Field term: +B . S on every site (src/Hamiltonian.h:167-198,
 * 363-375).  The diagonal is accumulated as the reference does: z bonds in
 * the given order, then (if J != 0) all bonds, then Bzz * Sz_total. */
typedef struct {
    int32_t n_sites;              /* N, 2 <= N <= 32 */
    int32_t n_bonds[3];           /* bonds per family (x, y, z) */
    const int32_t *bonds[3];      /* family f: n_bonds[f] pairs (i, j), i != j */
    const int32_t *neigh;         /* N x 3 row-major neighbour table, column 0 = z,
                                     1 = y, 2 = x, -1 = none; needed by the current
                                     operators only (may be NULL) */
    double K[3];                  /* Kxx, Kyy, Kzz */
    double J;                     /* Heisenberg coupling (extension) */
    double Gamma;                 /* Gamma coupling (extension) */
    double B[3];                  /* Bxx, Byy, Bzz */
} esc_model;

/* Operators for esc_apply_op (src/Observables.h:24-67 ApplyOp labels). */
#define ESC_OP_SX   0   /* S^x_site            src/Observables.h:506-521 */
#define ESC_OP_SY   1   /* S^y_site            src/Observables.h:488-503 */
#define ESC_OP_SZ   2   /* S^z_site            src/Observables.h:429-445 */
#define ESC_OP_JSX  3   /* spin current x      src/Observables.h:243-281 (site), :664-674 (total) */
#define ESC_OP_JSY  4   /* spin current y      src/Observables.h:202-240, :676-686 */
#define ESC_OP_JSZ  5   /* spin current z      src/Observables.h:161-199, :689-699 */
#define ESC_OP_JEX  6   /* energy current x    src/Observables.h:358-389 (A site), :448-458 (total over even sites) */
#define ESC_OP_JEY  7   /* energy current y    src/Observables.h:321-354 (B site), :460-470 (total over odd sites) */
#define ESC_SITE_TOTAL (-1)

/* Lanczos modes (src/DynLanczosSolver.h:38-239, LancType). */
#define ESC_LANC_ONLYEGS 0   /* no reorthogonalisation, no state formed */
#define ESC_LANC_GS      1   /* reorthogonalised; ground state formed */
#define ESC_LANC_ALL     2   /* reorthogonalised; basis kept for Ritz overlaps */
#define ESC_LANC_PLAIN   3   /* three-vector recursion, no basis kept (N=32 path) */

/* Start vectors (src/DynLanczosSolver.h:58-59,465-474). */
#define ESC_START_REF_UNIFORM  0  /* mt19937_64(seed), re/im = -1 + 2 U[0,1), drawn re then im (host, N <= 28) */
#define ESC_START_GAUSSIAN     1  /* counter-based (Philox4x32-10) N(0,1/2) per component */

/* ---------------------------------------------------------------- errors -- */
const char *esc_last_error(void);
int esc_version(void);

/* --------------------------------------------------------------- context -- */
/* Build device tables for a model.  Replaces Hamiltonian(...) + TightB_Ham()
 * + Diagonal() (src/Hamiltonian.h:20-25,94-206,331-377): no COO lists or
 * diagonal are stored; matrix elements are formed on the fly. */
int esc_ctx_create(const esc_model *model, int precision, esc_ctx **out);
int esc_ctx_destroy(esc_ctx *ctx);
/* Vector length (2^N) and bytes per element (16 or 8) of a context. */
int esc_ctx_info(const esc_ctx *ctx, int64_t *dim, int32_t *elem_bytes, int32_t *precision);

/* Kernel autotuner (no reference counterpart): times esc_hpsi with every
 * kernel configuration the context supports over `reps` launches on
 * d_in -> d_out, keeps the fastest for all later H.psi calls of the context
 * and reports it as tile_bits.  A configuration is 0 (one row per lane) or
 * 100000 * P + 10000 * V + 100 * TB + W: workgroups of 2^V waves sharing 2^TB
 * rows staged in LDS, walking windows of 2^W rows (one XCD per window) in P
 * planned passes (P = 1 when W = N), plus passes for terms no window held;
 * P = 5: windows chosen first to hold every term in as few passes as the
 * width W allows; P = 6: the same windows run in reverse order; P = 7 / 8:
 * the plans of 5 / 6 run by kernels generated for the lattice at run time
 * (hiprtc) in which each lane holds 2^k rows over k tile sites, with
 * 1000000 * k added to the configuration, and 10000000 * a layout digit:
 * bit 0 row pairs (complex64, two adjacent rows per 16-byte lane access),
 * bit 1 lanes on six contiguous sites (else only a 64-byte run is fixed).
 * d_out is overwritten. */
int esc_autotune(esc_ctx *ctx, const void *d_in, void *d_out, int reps, int *tile_bits, double *ms, void *stream);
/* Select the H.psi kernel explicitly: a configuration as reported by
 * esc_autotune or listed by esc_kernel_configs, or any generated-kernel
 * configuration (P = 7 / 8) that plans for the context's model. */
int esc_set_tile_bits(esc_ctx *ctx, int tile_bits);
/* List the kernel configurations of a context (*n = count; up to cap written). */
int esc_kernel_configs(const esc_ctx *ctx, int32_t *configs, int32_t cap, int32_t *n);
/* Generated-kernel configurations (P = 7 / 8 above) compile at run time on
 * first use.  This builds one into the code-object cache without a device
 * (ESC_JIT_CACHE, else jit_cache/ beside libesc.so), so that later contexts
 * of the same lattice and term structure load it instead of compiling; the
 * couplings are run-time arguments and do not enter the cache key.  source
 * (may be NULL) receives the generated HIP source, truncated to cap - 1
 * characters; *len its full length.  No reference counterpart. */
int esc_kernel_precompile(const esc_model *model, int precision, int config, char *source, size_t cap,
                          size_t *len);
/* The pass plan of an H.psi kernel configuration (> 0) for a model, without
 * a device (no reference counterpart; for tests and diagnostics).  Pass p <
 * *n_passes: passes[6p + 0] = window sites (bit mask), + 1 = tile sites,
 * + 2 = near terms (mask inside the tile), + 3 = far terms, + 4 = LDS bytes
 * of one workgroup's tile, + 5 = 1 if the pass adds the diagonal.  Term t <
 * *n_terms (the off-diagonal terms: bonds in family order, then flips):
 * term_pass[t] = its pass (-1: none), term_mask[t] = its flip mask.  Arrays
 * may be NULL; at most cap_passes / cap_terms entries are written. */
int esc_kernel_plan(const esc_model *model, int precision, int config, int64_t *passes, int cap_passes,
                    int *n_passes, int32_t *term_pass, uint32_t *term_mask, int cap_terms, int *n_terms);

/* ---------------------------------------------------------------- kernels -- */
/* d_out = H d_in.  Replaces Hamiltonian::MatrixVectorMult
 * (src/Hamiltonian.h:480-557).  d_in and d_out must not overlap. */
int esc_hpsi(esc_ctx *ctx, const void *d_in, void *d_out, void *stream);

/* d_out = Op(site) d_in, or the site sum for site == ESC_SITE_TOTAL.
 * Replaces Observables::ApplyOp (src/Observables.h:24-67) and
 * ApplyJ{sx,sy,sz,ex,ey}Total (:448-470,664-699).  Current operators need
 * model.neigh; a missing neighbour is ESC_ERR_ARG. */
int esc_apply_op(esc_ctx *ctx, int op, int site, const void *d_in, void *d_out, void *stream);

/* ------------------------------------------------------ vector utilities -- */
/* Start vector, normalised.  Replaces DynLanczosSolver::Randomize + normalize
 * (src/DynLanczosSolver.h:58-59,465-474). */
int esc_start_vector(esc_ctx *ctx, uint64_t seed, int kind, void *d_out, void *stream);
/* out[0] + i out[1] = <x|y> (conjugate-linear in x), FP64 accumulate,
 * deterministic.  Replaces Eigen's VectorXcd::dot as used at
 * src/DynLanczosSolver.h:76,95 and src/dynamics.h:250-262,370-381. */
int esc_dot(esc_ctx *ctx, const void *d_x, const void *d_y, double *out, void *stream);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  d_out[0..1] = <d_x|d_y> written by the kernels into
 * caller device memory (no host synchronisation, no allocation, the same bits
 * as esc_dot).  The partial sums live in one fixed 16 KiB device array of the
 * library per GPU (2 x 1024 doubles; the limit is that fixed size, which any
 * vector length fits), shared by every call on that GPU: calls that overlap
 * on one GPU race. */
int esc_dot_noalloc(esc_ctx *ctx, const void *d_x, const void *d_y, double *d_out, void *stream);
/* d_y = a d_x + b d_y, complex a = (a[0], a[1]), b likewise. */
int esc_axpby(esc_ctx *ctx, const double *a, const void *d_x, const double *b, void *d_y, void *stream);
/* Copy host complex<double> data into a context-precision device vector and back. */
int esc_upload(esc_ctx *ctx, const double *host, void *d_out, void *stream);
int esc_download(esc_ctx *ctx, const void *d_in, double *host, void *stream);

/* ---------------------------------------------------------------- Lanczos -- */
/* Lanczos from the normalised d_v0.  Replaces DynLanczosSolver::Lanczos_Nirav
 * (src/DynLanczosSolver.h:38-239).  alpha[0..iters-1], beta[0..iters]
 * (beta[0] = 0) are FP64 and recorded before reorthogonalisation, as the
 * reference does.  Exit at iteration steps-1 or when |dE0| < eps (:139-162).
 * evals (iters, ascending) and Z (iters x iters, column-major, column k =
 * eigenvector k) describe the final tridiagonal matrix.  For ESC_LANC_GS and
 * ESC_LANC_ALL the ground state sum_m Z(m,0) v_m is written to d_gs (may be
 * NULL otherwise).  ESC_LANC_ALL keeps the basis in the context for
 * esc_ritz_overlaps / esc_ritz_vector.  Capacity: steps <= 4096. */
int esc_lanczos(esc_ctx *ctx, const void *d_v0, int steps, double eps, int mode,
                double *alpha, double *beta, int *iters, double *evals, double *Z,
                void *d_gs, void *stream);

/* Solver=ED: every eigenpair of H.  The dense 2^N x 2^N matrix is formed on
 * the device from the context's own H.psi (column k = H e_k, so every element
 * is bit-exact) and diagonalised by rocSOLVER zheevd (divide and conquer,
 * FP64).  evals: 2^N doubles, ascending.  The eigenvectors become the kept
 * basis, exactly as after an ESC_LANC_ALL run with an identity rotation, so
 * esc_ritz_vector / esc_ritz_overlaps / ESC_SPEC_RITZ spectra read them as
 * PsiAll; d_gs (may be NULL) receives eigenvector 0.  complex128 contexts,
 * N <= 13.  Replaces ExactDiag + Diagonalize (src/Exactdiag.h:36-110) and
 * DynLanczosSolver::fillFromED (src/DynLanczosSolver.h:269-283). */
int esc_exact_diag(esc_ctx *ctx, double *evals, void *d_gs, void *stream);

/* Ground state with three resident Lanczos vectors plus the start and the
 * result (the memory plan for N = 32: 5 vectors of 2^32 complex64 = 172 GB).
 * Each cycle runs the recursion of esc_lanczos without reorthogonalisation
 * from the normalised start (same exit rule), then replays it -- the kernels
 * are deterministic, so the replay produces the same vectors -- accumulating
 * |gs> = sum_m Z(m,0) v_m; later cycles restart from the previous |gs>
 * (`restarts` extra cycles; each restart removes the drift that rounding in
 * complex64 storage leaves in the Ritz vector, and stops early once the
 * lowest Ritz pair's residual bound reaches four units of the storage
 * precision times |T|).  e0 = <gs|H|gs>, resid =
 * |H gs - e0 gs| (FP64 reductions), iters = H.psi applications in all.
 * Replaces the ground-state use of Lanczos_Nirav("GS") where the basis
 * cannot be stored (src/DynLanczosSolver.h:38-239). */
int esc_ground_state(esc_ctx *ctx, const void *d_v0, int steps, double eps, int restarts, double *e0,
                     double *resid, int *iters, void *d_gs, void *stream);

/* O[n] = <psi_n|d_vec>, n < n_states, for the Ritz states of the last
 * ESC_LANC_ALL run, computed as sum_m conj(Z(m,n)) <v_m|d_vec> without forming
 * the states.  Replaces the PsiAll construction plus Overlap loops
 * (src/DynLanczosSolver.h:208-234; src/dynamics.h:255-263,375-383).
 * out is interleaved complex (2 * n_states doubles). */
int esc_ritz_overlaps(esc_ctx *ctx, const void *d_vec, int n_states, double *out, void *stream);
/* d_out = PsiAll[k] = sum_m Z(m,k) v_m (src/DynLanczosSolver.h:213-223). */
int esc_ritz_vector(esc_ctx *ctx, int k, void *d_out, void *stream);

/* Static two-point correlations Corr(i, j) = <gs|A_i B_j|gs> of site
 * operators op_a, op_b (ESC_OP_SX..ESC_OP_JEY) over all site pairs, filled for
 * i >= j and mirrored as Observables::TwoPointCorr does
 * (src/Observables.h:132-155).  corr: N x N interleaved complex, row-major
 * (corr[2(iN + j)]).  Uses N + 1 device vectors. */
int esc_two_point(esc_ctx *ctx, int op_a, int op_b, const void *d_gs, double *corr, void *stream);

/* Plain three-vector Lanczos (no reorthogonalisation, no basis) from
 * d_phi / |d_phi| for the continued-fraction spectra; norm2 = |d_phi|^2.
 * No reference counterpart (the reference uses Ritz-Lehmann sums). */
int esc_lanczos_cf(esc_ctx *ctx, const void *d_phi, int steps, double *alpha, double *beta,
                   int *iters, double *norm2, void *stream);

/* ------------------------------------------------------------- dynamics -- */
#define ESC_DYN_JSJS 0   /* DynJsJs: 14 value columns (Re, Im) xx yy zz xy xz yz tt */
#define ESC_DYN_JEJE 1   /* DynJeJe: 5 value columns Im xx yy xy yx tt */
#define ESC_DYN_LOCALSZSZ 2  /* LocalDynSzSz: 8 value columns Im xx yy zz xy xz yz, S+e3 S+e3, S-e3 S-e3
                              * (src/dynamics.h:64-217); expect = sum over sites of <S^x>, <S^y>, <S^z> */
#define ESC_SPEC_RITZ 0  /* Lehmann sum over the Ritz states of the last ESC_LANC_ALL run */
#define ESC_SPEC_CF   1  /* continued fractions from J|gs> (cross terms by polarisation) */
#define ESC_SPEC_CF_REORTH 2  /* the same with full reorthogonalisation of each run's Lanczos
                               * vectors (cf_steps + 1 resident vectors): keeps complex64
                               * recursions as accurate as complex128 ones */

/* Full dynamical-correlation pipeline for one kind, replacing
 * StartDynamics -> DynJsJs / DynJeJe (src/dynamics.h:43-60,222-460) minus the
 * file writing: currents applied to d_gs, <J_a> and overlaps (RITZ) or
 * Lanczos continued fractions (CF, cf_steps each), evaluated on omega.
 * table is row-major nw x (1 + ncols): omega, then the columns in the
 * reference's file order; ncols = 14 (JSJS) or 5 (JEJE).  The JeJe yx column
 * carries no elastic subtraction (the reference subtracts an uninitialised
 * Jez there, src/dynamics.h:368-369,424).  e0 is used by CF only (RITZ uses the
 * lowest Ritz value, src/dynamics.h:234,357).  expect (may be NULL) receives
 * <J_a> for a = x, y, (z,) t as interleaved complex. */
int esc_dynamics(esc_ctx *ctx, int kind, int method, const void *d_gs, double e0, int cf_steps,
                 double eta, const double *omega, int nw, double *table, double *expect, void *stream);

/* ------------------------------------------------- lattice and parameters -- */
/* LocalDynRIXS (src/dynamics.h:465-562): RIXS at core-hole site `site`.
 * gs_ctx holds H (final states), int_ctx the intermediate Hamiltonian H_int
 * = H without the bonds of `site` (Hamiltonian::RIXS_killcenter,
 * src/Hamiltonian.h:57-91; build it with esc_ctx_create from bond lists that
 * leave those bonds out).  With inv_n = 1/(-E_n - i gamma) and
 * chi = sum_n |n><n|gs> inv_n over the intermediate states,
 *   I_X(w) = sum_{m >= 1} |<m|X chi>|^2 eta/((w - (E_m - Eg))^2 + eta^2),
 * X = 1, S^x_site, S^y_site, S^z_site.  ESC_SPEC_RITZ: n and m run over the
 * Ritz states of ESC_LANC_ALL runs kept on int_ctx and gs_ctx (the
 * reference's sums), Eg = the lowest Ritz value of gs_ctx, e0 and cf_steps
 * unused.  ESC_SPEC_CF: chi from cf_steps Lanczos steps of H_int started at
 * |gs>, I_X from the continued fraction of H at X chi minus the m = 0 term,
 * Eg = e0.  table: nw rows of 9 doubles (w, Re I_1, Re I_x, Re I_y, Re I_z,
 * Im I_1..Im I_z), the columns of RIXS_outSpectrum.real; the reference uses
 * gamma = 100, eta = 0.1, w = 0.1 k, k = 0..120. */
int esc_rixs(esc_ctx *gs_ctx, esc_ctx *int_ctx, int method, int site, const void *d_gs, double e0, int cf_steps,
             double gamma, double eta, const double *omega, int nw, double *table, void *stream);
/* DynSzSz (src/dynamics.h:566-611): for every pair of sites (i, j)
 *   sum_n Re(<S^z_j gs|n><n|S^z_i gs>) / (w - (E_n - Eg) - i eta)
 * over the Ritz states n of the ESC_LANC_ALL run kept on ctx (Eg = its lowest
 * Ritz value); table[w][i][j] = (real, imaginary part), nw x N x N x 2
 * doubles (SzSz_outSpectrum.real lists i, j, w, real, imaginary with w
 * outermost; the reference uses eta = 0.1, w = 0.1 k, k = 0..120). */
int esc_szsz(esc_ctx *ctx, const void *d_gs, double eta, const double *omega, int nw, double *table, void *stream);

/* Honeycomb torus neighbour table (N x 3 row-major, column 0 = z, 1 = y,
 * 2 = x, -1 = none) for Lx unit cells along a1 and Ly/2 along a2, site numbering
 * A = 2 (i Ly/2 + j), B = A + 1.  Replaces Lattice::Honeycomb1
 * (src/Lattice.h:434-576).  Ly must be even. */
int esc_honeycomb(int Lx, int Ly, int periodic_x, int periodic_y, int32_t *neigh);
/* Bond families from a neighbour table, pairs (i < j) in lexicographic order;
 * a family whose coupling K[f] is 0 is empty.  Replaces Hamiltonian::Connectors
 * (src/Hamiltonian.h:230-328).  Each pairs_f holds up to N pairs. */
int esc_connectors(int n_sites, const int32_t *neigh, const double *K, int32_t *pairs_x, int32_t *pairs_y,
                   int32_t *pairs_z, int32_t *counts);

/* Chain lattice for Model=HeisenbergChain, replacing Lattice::Chain
 * (src/Lattice.h:60-125): neigh is N x 2, column 0 = +x and column 1 = -x
 * neighbour (-1 when absent; wrapped when periodic). */
int esc_chain(int n_sites, int periodic, int32_t *neigh);
/* HeisenbergChain_Connectors (src/Hamiltonian.h:381-475): every family with a
 * non-zero coupling holds all chain bonds (i < j), in row-major order; with
 * these families the Kitaev-form terms K_f S^f_i S^f_j give the XYZ chain.
 * The J and Gamma extensions are honeycomb-only (leave them 0 on a chain). */
int esc_chain_connectors(int n_sites, const int32_t *neigh, const double *K, int32_t *pairs_x, int32_t *pairs_y,
                         int32_t *pairs_z, int32_t *counts);

/* Parameter file (key=value, first exact key match wins), replacing
 * ConstVariables::Initialize / matchstring / matchstring2
 * (src/ParametersEngine.h:53-229) including its defaults and checks.
 * Extension keys (optional): J, Gamma, Eta, OmegaMin, OmegaMax, OmegaPoints,
 * Precision (c128|c64), SpectralMode (Ritz|CF|CFReorth), StartVector
 * (Uniform|Gaussian), CFSteps, LocalSzSz (0|1), SzSz (0|1), LocalRIXS (0|1),
 * SpecialSite.  The reference declares SpecialSite but never reads it
 * (src/ParametersEngine.h:29,92-97); here it is a key, defaulting to the
 * value of the reference's commented-out line, NumberofSites / 2. */
typedef struct {
    int32_t Lx, Ly, NumberofSites, RandomSeed, LanczosSteps, Threads;
    int32_t IsPeriodicX, IsPeriodicY, LocalSkw, SpinCurrDynamics, EnergyCurrDynamics;
    double Kxx, Kyy, Kzz, Bxx, Byy, Bzz, Temperature, LanczosEps;
    char Model[64], Solver[64], LancType[64];
    /* extensions */
    double J, Gamma, Eta, OmegaMin, OmegaMax;
    int32_t OmegaPoints;      /* 0 = the reference grid omega = 0.002 k, k <= 1000 */
    int32_t Precision;        /* ESC_C128 / ESC_C64 */
    int32_t SpectralMode;     /* ESC_SPEC_RITZ / ESC_SPEC_CF / ESC_SPEC_CF_REORTH */
    int32_t StartVector;      /* ESC_START_* */
    int32_t CFSteps;
    int32_t LocalSzSz;        /* 1: LocalDynSzSz -> 2SSLocal_outSpectrum.real (src/dynamics.h:64-217) */
    int32_t SzSz;             /* 1: DynSzSz -> SzSz_outSpectrum.real (src/dynamics.h:566-611) */
    int32_t LocalRIXS;        /* 1: LocalDynRIXS -> RIXS_outSpectrum.real (src/dynamics.h:465-562) */
    int32_t SpecialSite;      /* core-hole site of LocalDynRIXS; default NumberofSites / 2 */
} esc_params;
int esc_params_load(const char *path, esc_params *out);

/* ------------------------------------------------ host FP64 small solves -- */
/* Symmetric tridiagonal eigensolve (implicit QL, Wilkinson shift), n <= 4096:
 * diag a[0..n-1], off-diagonal b[1..n-1] (b[0] ignored), as TriDiag builds T
 * (src/DynLanczosSolver.h:371-414).  Eigenvalues ascending; Z column-major
 * (may be NULL for eigenvalues only). */
int esc_tridiag_eig(int n, const double *a, const double *b, double *evals, double *Z);

/* Every eigenpair of T (a, b as above): Sturm-count bisection for the
 * eigenvalues, inverse iteration for the vectors, each iterate
 * orthogonalised against the earlier vectors of its cluster (consecutive
 * eigenvalues within 1e-3 |T|_1, LAPACK dstein's criterion); ascending
 * evals, Z column-major (column j = eigenvector j, first nonzero component
 * positive).  on_device = 0: the host; 1: one workgroup on the current
 * device -- the solver esc_lanczos and the RIXS fractions use -- with the
 * same bits.  n <= 4096.  TriDiag (src/DynLanczosSolver.h:371-414). */
int esc_tridiag_eigenpairs(int n, const double *a, const double *b, double *evals, double *Z, int on_device);

/* Lowest eigenpair of T (a, b as above, every b[k] != 0 for 1 <= k < n):
 * Sturm-count bisection for the eigenvalue, inverse iteration with a
 * partially pivoted LU of T - e0 I for the unit eigenvector z (may be NULL;
 * first nonzero component positive; inside a cluster of equal eigenvalues,
 * one unit vector of the cluster).  on_device = 0 runs the routine on the
 * host, 1 the kernel the batched ground states use on the current device;
 * both compute the same bits.  The ground-state half of TriDiag
 * (src/DynLanczosSolver.h:371-414): E0 and the coefficients of |gs>. */
int esc_tridiag_lowest(int n, const double *a, const double *b, double *e0, double *z, int on_device);

/* The raw continued fraction n2 <phi|(z - H)^-1|phi> / n2 of a recursion of
 * n levels (a[0..n-1], b[1..n-1], every b[k] != 0), z = w + e0 - i eta, on
 * nw frequencies; out interleaved complex, zero unless norm2 > 0.  on_device = 0: host threads;
 * 1: the kernel of the batched spectra on the current device; both compute
 * the same bits.  The fraction DynJsJs / DynJeJe's CF mode sums
 * (src/dynamics.h:270-336,389-455). */
int esc_cf_eval(int n, const double *a, const double *b, double norm2, double e0, double eta,
                const double *omega, int nw, double *out, int on_device);

/* Continued fraction: G(w) = norm2 / (z - a0 - b1^2/(z - a1 - ...)),
 * z = w + e0 - i eta, for nw frequencies.  Optional elastic subtraction of
 * |<J>|^2/(w - i eta) when elastic != NULL (elastic[0] + i elastic[1] =
 * conj(<J_a>) <J_b>).  out: interleaved complex.  No reference counterpart. */
int esc_cf_spectrum(int n, const double *a, const double *b, double norm2, double e0, double eta,
                    const double *omega, int nw, const double *elastic, double *out);

/* Lehmann sum over Ritz or exact states, as DynJsJs/DynJeJe evaluate it
 * (src/dynamics.h:270-316,389-425):
 * I(w) = sum_n conj(Oa[n]) Ob[n] / (w - (E_n - E_0) - i eta) - elastic/(w - i eta),
 * E_0 = evals[0]; Oa, Ob, elastic interleaved complex; out interleaved. */
int esc_lehmann_spectrum(int n, const double *evals, const double *Oa, const double *Ob,
                         const double *elastic, double eta, const double *omega, int nw,
                         double *out);

/* --------------------------------------------------------------- batches -- */
/* Many independent systems (any sizes and couplings, one precision) solved in
 * lockstep: each Lanczos step of all instances is replayed as one captured
 * HIP graph (the instances' H.psi launches and four batched kernels that
 * keep alpha / beta on the device).  No reference counterpart: the reference
 * solves one system per process (main.cpp:41-112); these entry points run
 * its ground-state + JsJs / JeJe pipeline for a whole parameter scan.
 * Each instance holds 6 vectors.  The calls are synchronous. */
typedef struct esc_batch esc_batch;
int esc_batch_create(int n, const esc_model *models, int precision, esc_batch **out);
int esc_batch_destroy(esc_batch *b);
/* Ground states: `steps` lockstep steps without reorthogonalisation from
 * Gaussian start vectors (seed + i, as esc_start_vector), then a replay
 * forming |gs_i> = sum_k Z(k,0) v_k.  e0[i], iters[i] (recursion length; it
 * ends early only on an exhausted Krylov space), alpha (n x steps) and beta
 * (n x (steps + 1)) may be NULL.  use_graph = 0 issues the same launches
 * without graphs (same results bit for bit). */
int esc_batch_ground_state(esc_batch *b, uint64_t seed, int steps, int use_graph, double *e0, int *iters,
                           double *alpha, double *beta);
/* Tune every instance's H.psi kernel (esc_autotune on a Gaussian vector);
 * configs (may be NULL) receives the n choices.  Clears the ground states. */
int esc_batch_autotune(esc_batch *b, int reps, int *configs);
/* Copy instance i's ground state into d_out (the instance's precision). */
int esc_batch_state(esc_batch *b, int i, void *d_out, void *stream);
/* JsJs / JeJe spectra of every instance by continued fractions (cf_steps per
 * run), rows as esc_dynamics: tables is n x nw x (1 + 14) for JsJs or
 * n x nw x (1 + 5) for JeJe; expect (may be NULL) n x 4 complex <J_a>. */
int esc_batch_dynamics(esc_batch *b, int kind, int cf_steps, double eta, const double *omega, int nw,
                       int use_graph, double *tables, double *expect);

/* ----------------------------------------------------------- gradients -- */
/* Loss L = sum_q (S(w_q) - target_q)^2 of the spectrum
 *   S(w) = Im <phi|(w + E0 - i eta - H)^-1|phi>,  phi = O|gs>,
 * evaluated as the continued fraction of cf_steps Lanczos steps from phi,
 * and its gradient with respect to theta = (K, J, Gamma, h): K moves all three
 * Kitaev couplings together (direction (1,1,1) in Kx, Ky, Kz), h is the field
 * strength along [111] (B = h (1,1,1)/sqrt 3).  O is a total current
 * (ESC_OP_JSX..ESC_OP_JEY, whose couplings follow K) or a single-site
 * operator (ESC_OP_SX..SZ at `site`).  The ground state is formed in FP64
 * (gs_steps-step replayed Lanczos with one restart, then Newton residual
 * corrections).  Gradient: adjoint of the continued fraction, of the Lanczos
 * recursion (checkpoint > 0: its vectors checkpointed every `checkpoint`
 * steps and recomputed per segment in the reverse sweep, memory 3 + 2
 * cf_steps / checkpoint vectors; checkpoint = 0: all cf_steps + 1 vectors kept
 * and reorthogonalised, which keeps long recursions on their exact-arithmetic
 * course so that the gradient matches finite differences of the loss),
 * of O's couplings, of E0 (Hellmann-Feynman) and of |gs> (one conjugate-
 * gradient solve with H - E0 on the complement of |gs>).  Outputs: loss,
 * spectrum (nw, may be NULL), grad[4], e0, de0[4] = <gs|dH/dtheta|gs> (may be
 * NULL).  complex128 throughout.  No reference counterpart. */
#define ESC_THETA_K     0
#define ESC_THETA_J     1
#define ESC_THETA_GAMMA 2
#define ESC_THETA_H     3
int esc_spectrum_grad(const esc_model *model, int op, int site, int gs_steps, int cf_steps, int checkpoint,
                      double eta, const double *omega, int nw, const double *target, double *loss,
                      double *spectrum, double *grad, double *e0, double *de0, void *stream);

/* The same gradient for a loss over a whole JsJs or JeJe table (kind
 * ESC_DYN_JSJS / ESC_DYN_JEJE): L = sum_q sum_c (T_c(w_q) - target[q ncols
 * + c])^2 over the table's columns c (JsJs: 14, the real and imaginary parts
 * of xx, yy, zz, xy, xz, yz, tt; JeJe: 5, Im of xx, yy, xy, yx, tt), T as
 * esc_dynamics writes it in CF mode: the continued fractions of O_r|gs> for
 * the runs O_r = J_a, J_a + J_b, J_a + i J_b, sum_a J_a, the cross channels by
 * polarisation, and the elastic terms conj(<J_a>) <J_b> / (w - i eta)
 * subtracted (src/dynamics.h:222-460).  Every run contributes through its
 * own recursion adjoint; the ground state's adjoint (one CG solve) collects
 * O_r^dagger phi_bar_r of all runs and the expectation values' 2 J_a|gs>
 * terms.  table: nw x (1 + ncols), first column omega (may be NULL).  Other
 * arguments and outputs as esc_spectrum_grad.  No reference counterpart. */
int esc_table_grad(const esc_model *model, int kind, int gs_steps, int cf_steps, int checkpoint, double eta,
                   const double *omega, int nw, const double *target, double *loss, double *table, double *grad,
                   double *e0, double *de0, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* ESC_H */
