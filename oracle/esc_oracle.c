/*
 * esc_oracle.c -- CPU restatement of the reference's Lanczos dynamical-correlation
 * path, used ONLY as test infrastructure (the checker).  Nothing in the product
 * library links, loads or calls this file; only tests/, the smoke check and the
 * CPU-baseline leg of bench.py do.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * fengshi96/EnergySpinCurrents root).  Complex vectors are interleaved (re, im)
 * doubles; basis index r has bit k = site k, bit value 0 = spin up
 * (src/QuantumBasis.h:34-35,57).
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   orc_honeycomb1        closed-form torus + hand-derived 12-site table
 *   orc_connectors        lexicographic bond order, known 12-site table
 *   orc_coo_*             dense Kronecker-product Hamiltonian (numpy, N<=10),
 *                         Tr H = 0, Tr H^2 closed form, Hermiticity
 *   orc_hpsi_gather       equality with orc_coo_* (reference model) and with the
 *                         Kronecker builder (J, Gamma extensions)
 *   orc_apply_site_op,
 *   orc_current_*         dense Kronecker operators built from Pauli matrices
 *   orc_mt19937_64        C++ standard's 10000th-output requirement
 *                         ([rand.predef]: 9981545732273789042 for seed 5489)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

/* ------------------------------------------------------------------------- */
/* Lattice: restatement of Lattice::Honeycomb1 (src/Lattice.h:434-576).        */
/* neigh is row-major N x 3 with column 0 = z, 1 = y, 2 = x neighbour          */
/* (src/Lattice.h:529-575 column use).  Returns 0 on success, -1 on odd Ly.    */
/* ------------------------------------------------------------------------- */
int orc_honeycomb1(int Lx, int Ly, int perX, int perY, int *neigh, int *indx, int *indy)
{
    int N = Lx * Ly;                       /* src/ParametersEngine.h:60 */
    int LLX = Lx, LLY = Ly / 2;            /* src/Lattice.h:437-438 */
    if (Ly % 2 != 0 || Lx < 1 || LLY < 1) return -1;
    double scalex = 2.0, scaley = 4.0 / sqrt(3.0);
    double t1x = 1.0 * scalex;
    double t2x = 0.5 * scalex, t2y = sqrt(3.0) / 2.0 * scaley;   /* shift = "right" */
    int ncx = LLX * 2 + LLY, ncy = LLY * 2;                       /* :458 */
    int *Nc = (int *)malloc(sizeof(int) * ncx * ncy);
    for (int k = 0; k < ncx * ncy; k++) Nc[k] = -1;
#define NC(x, y) Nc[(x) + (y) * ncx]
    for (int k = 0; k < N * 3; k++) neigh[k] = -1;               /* N1neigh_.fill(-1) */
    int xv = 0, counter = 0;
    for (int i = 0; i < LLX; i++) {                               /* :461-486 */
        if (i == 0) xv = 0; else xv = (int)(xv + t1x);
        int x0 = xv, y0 = 0, x1 = (int)(xv + 1.0), y1 = 1;
        for (int j = 0; j < LLY; j++) {
            int cxa = (int)(x0 + j * t2x), cxb = (int)(x1 + j * t2x);
            int cya = (int)(y0 + j * t2y), cyb = (int)(y1 + j * t2y);
            indx[counter] = cxa; indy[counter] = cya; NC(cxa, cya) = counter; counter++;
            indx[counter] = cxb; indy[counter] = cyb; NC(cxb, cyb) = counter; counter++;
        }
    }
    int xmax = 0, ymax = 0;
    for (int k = 0; k < N; k++) { if (indx[k] > xmax) xmax = indx[k]; if (indy[k] > ymax) ymax = indy[k]; }
#define NB(i, c) neigh[(i) * 3 + (c)]
    for (int i = 0; i < N; i++) {                                 /* :494-572 */
        int ix = indx[i], iy = indy[i], jx, jy;
        jx = ix + 1; jy = iy + 1;                                 /* x: column 2 */
        if (jx < xmax + 1 && jy < ymax + 1 && NC(jx, jy) != -1) { int j = NC(jx, jy); NB(i, 2) = j; NB(j, 2) = i; }
        jx = ix + 1; jy = iy - 1;                                 /* y: column 1 */
        if (jx < xmax + 1 && jy < ymax && jy >= 0 && NC(jx, jy) != -1) { int j = NC(jx, jy); NB(i, 1) = j; NB(j, 1) = i; }
        jx = ix; jy = iy + 1;                                     /* z: column 0 */
        if (jx < xmax && jy < ymax + 1 && NC(jx, jy) != -1) { int j = NC(jx, jy); NB(i, 0) = j; NB(j, 0) = i; }
        if (perY) {                                               /* :530-537 */
            jx = ix - LLY; jy = 0;
            if (jx < xmax && iy == ymax && NC(jx, jy) != -1) { int j = NC(jx, jy); NB(i, 0) = j; NB(j, 0) = i; }
        }
        if (perX && iy % 2 == 0 && NC(ix, iy) < LLY * 2) {        /* :548-555 */
            jx = ix + LLX * 2 - 1; jy = iy + 1;
            if (jx < xmax + 1 && iy < ymax + 1 && NC(jx, jy) != -1) { int j = NC(jx, jy); NB(i, 1) = j; NB(j, 1) = i; }
        }
    }
#undef NB
#undef NC
    free(Nc);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Bond families: restatement of Hamiltonian::Connectors                      */
/* (src/Hamiltonian.h:230-328).  Family order here: 0 = x, 1 = y, 2 = z.      */
/* pairs[f] receives counts[f] pairs (i<j) in lexicographic order.  A family  */
/* whose coupling is 0 has no bonds (numNonZeros, :269-271).                   */
/* ------------------------------------------------------------------------- */
int orc_connectors(int N, const int *neigh, const double *K /*Kxx,Kyy,Kzz*/,
                   int *pairs_x, int *pairs_y, int *pairs_z, int *counts)
{
    /* neighbour column per family: x -> 2, y -> 1, z -> 0 (:241-263) */
    const int col[3] = {2, 1, 0};
    int *out[3] = {pairs_x, pairs_y, pairs_z};
    char *sp = (char *)calloc((size_t)N * N, 1);
    for (int f = 0; f < 3; f++) {
        memset(sp, 0, (size_t)N * N);
        for (int i = 0; i < N; i++) {
            int j = neigh[i * 3 + col[f]];
            if (i < j && K[f] != 0.0) sp[i * N + j] = 1;
        }
        int c = 0;
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++)
                if (sp[i * N + j]) { out[f][2 * c] = i; out[f][2 * c + 1] = j; c++; }
        counts[f] = c;
    }
    free(sp);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* COO Hamiltonian: restatement of TightB_Ham (src/Hamiltonian.h:94-206),     */
/* Diagonal (:331-377) and MatrixVectorMult (:480-557).  A family with one     */
/* bond is skipped (:116,143,347 together with :283-285).                      */
/* ------------------------------------------------------------------------- */
typedef struct {
    int64_t nx, ny;           /* entries in the x and y lists */
    uint32_t *xi, *xf, *yi, *yf;
    double *xv, *yv;          /* interleaved complex */
    double *diag;             /* interleaved complex, dim entries */
    int64_t dim;
} orc_coo;

static inline int bitof(uint64_t k, int s) { return (int)((k >> s) & 1u); }

void *orc_coo_build(int N, const int *pairs_x, int nbx, const int *pairs_y, int nby,
                    const int *pairs_z, int nbz, const double *K, const double *B)
{
    orc_coo *h = (orc_coo *)calloc(1, sizeof(orc_coo));
    int64_t dim = (int64_t)1 << N;
    h->dim = dim;
    int use_x = (nbx != 1) && nbx > 0, use_y = (nby != 1) && nby > 0, use_z = (nbz != 1) && nbz > 0;
    h->nx = (use_x ? (int64_t)nbx * dim : 0) + dim * N;
    h->ny = (use_y ? (int64_t)nby * dim : 0) + dim * N;
    h->xi = (uint32_t *)malloc(sizeof(uint32_t) * h->nx); h->xf = (uint32_t *)malloc(sizeof(uint32_t) * h->nx);
    h->yi = (uint32_t *)malloc(sizeof(uint32_t) * h->ny); h->yf = (uint32_t *)malloc(sizeof(uint32_t) * h->ny);
    h->xv = (double *)malloc(sizeof(double) * 2 * h->nx); h->yv = (double *)malloc(sizeof(double) * 2 * h->ny);
    h->diag = (double *)calloc(2 * dim, sizeof(double));
    /* Entries sit in the reference's push_back order: bond b, ket at
     * b*dim + ket, then the field entries at base + ket*N + s.  Each entry's
     * index is fixed, so the loops over ket run in parallel. */
    const int64_t fx = use_x ? (int64_t)nbx * dim : 0, fy = use_y ? (int64_t)nby * dim : 0;
    if (use_x)                                                     /* :113-132 */
        for (int b = 0; b < nbx; b++) {
            int si = pairs_x[2 * b], sj = pairs_x[2 * b + 1];
#pragma omp parallel for schedule(static)
            for (int64_t ket = 0; ket < dim; ket++) {
                int64_t e = (int64_t)b * dim + ket;
                h->xi[e] = (uint32_t)ket;
                h->xf[e] = (uint32_t)(ket ^ ((uint64_t)1 << si) ^ ((uint64_t)1 << sj));
                h->xv[2 * e] = 0.25 * K[0]; h->xv[2 * e + 1] = 0.0;   /* Sx*Sx*Kxx */
            }
        }
#pragma omp parallel for schedule(static)
    for (int64_t ket = 0; ket < dim; ket++)                         /* :167-179 */
        for (int s = 0; s < N; s++) {
            int64_t e = fx + ket * N + s;
            h->xi[e] = (uint32_t)(ket ^ ((uint64_t)1 << s)); h->xf[e] = (uint32_t)ket;
            h->xv[2 * e] = 0.5 * B[0]; h->xv[2 * e + 1] = 0.0;
        }
    if (use_y)                                                     /* :140-158 */
        for (int b = 0; b < nby; b++) {
            int si = pairs_y[2 * b], sj = pairs_y[2 * b + 1];
#pragma omp parallel for schedule(static)
            for (int64_t ket = 0; ket < dim; ket++) {
                int64_t e = (int64_t)b * dim + ket;
                double ci = bitof(ket, si) == 0 ? 0.5 : -0.5;       /* ApplySy imag part */
                double cj = bitof(ket, sj) == 0 ? 0.5 : -0.5;
                h->yi[e] = (uint32_t)ket;
                h->yf[e] = (uint32_t)(ket ^ ((uint64_t)1 << si) ^ ((uint64_t)1 << sj));
                h->yv[2 * e] = -(ci * cj) * K[1]; h->yv[2 * e + 1] = 0.0;   /* (i ci)(i cj) Kyy */
            }
        }
#pragma omp parallel for schedule(static)
    for (int64_t ket = 0; ket < dim; ket++)                         /* :186-198 */
        for (int s = 0; s < N; s++) {
            int64_t e = fy + ket * N + s;
            double c = bitof(ket, s) == 0 ? 0.5 : -0.5;
            h->yi[e] = (uint32_t)(ket ^ ((uint64_t)1 << s)); h->yf[e] = (uint32_t)ket;
            h->yv[2 * e] = 0.0; h->yv[2 * e + 1] = B[1] * c;
        }
    if (use_z)                                                     /* :344-360 */
        for (int b = 0; b < nbz; b++) {
            int si = pairs_z[2 * b], sj = pairs_z[2 * b + 1];
#pragma omp parallel for
            for (int64_t ket = 0; ket < dim; ket++) {
                double zi = bitof(ket, si) == 0 ? 0.5 : -0.5, zj = bitof(ket, sj) == 0 ? 0.5 : -0.5;
                h->diag[2 * ket] += (K[2] * zi) * zj;
            }
        }
#pragma omp parallel for
    for (int64_t ket = 0; ket < dim; ket++) {                       /* :365-375 */
        double szt = 0.0;
        for (int s = 0; s < N; s++) szt += bitof(ket, s) == 0 ? 0.5 : -0.5;
        h->diag[2 * ket] += B[2] * szt;
    }
    return h;
}

void orc_coo_free(void *p)
{
    orc_coo *h = (orc_coo *)p;
    if (!h) return;
    free(h->xi); free(h->xf); free(h->yi); free(h->yf); free(h->xv); free(h->yv); free(h->diag); free(h);
}

int64_t orc_coo_nnz(void *p) { orc_coo *h = (orc_coo *)p; return h->nx + h->ny; }

/* Export the COO lists (for the dense ED builder): row, col, value. */
void orc_coo_export(void *p, uint32_t *rows, uint32_t *cols, double *vals, double *diag)
{
    orc_coo *h = (orc_coo *)p;
    memcpy(rows, h->xi, sizeof(uint32_t) * h->nx); memcpy(cols, h->xf, sizeof(uint32_t) * h->nx);
    memcpy(vals, h->xv, sizeof(double) * 2 * h->nx);
    memcpy(rows + h->nx, h->yi, sizeof(uint32_t) * h->ny); memcpy(cols + h->nx, h->yf, sizeof(uint32_t) * h->ny);
    memcpy(vals + 2 * h->nx, h->yv, sizeof(double) * 2 * h->ny);
    memcpy(diag, h->diag, sizeof(double) * 2 * h->dim);
}

/* MatrixVectorMult (src/Hamiltonian.h:480-557): per-thread private outputs,   */
/* static contiguous chunks, then a thread-ordered reduction.                  */
static void cmac(double *acc, const double *a, const double *x)
{
    double re = a[0] * x[0] - a[1] * x[1], im = a[0] * x[1] + a[1] * x[0];
    acc[0] += re; acc[1] += im;
}

int orc_coo_apply(void *p, const double *in, double *out, int nthreads)
{
    orc_coo *h = (orc_coo *)p;
    int64_t dim = h->dim;
    double **tv = (double **)malloc(sizeof(double *) * nthreads);
    for (int t = 0; t < nthreads; t++) tv[t] = (double *)calloc(2 * dim, sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
        int tid = omp_get_thread_num();
        double *o = tv[tid];
#pragma omp for schedule(static)
        for (int64_t i = 0; i < dim; i++) cmac(o + 2 * i, h->diag + 2 * i, in + 2 * i);
#pragma omp for schedule(static)
        for (int64_t e = 0; e < h->nx; e++) cmac(o + 2 * (int64_t)h->xi[e], h->xv + 2 * e, in + 2 * (int64_t)h->xf[e]);
#pragma omp for schedule(static)
        for (int64_t e = 0; e < h->ny; e++) cmac(o + 2 * (int64_t)h->yi[e], h->yv + 2 * e, in + 2 * (int64_t)h->yf[e]);
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < dim; i++) {
        double re = 0.0, im = 0.0;
        for (int t = 0; t < nthreads; t++) { re += tv[t][2 * i]; im += tv[t][2 * i + 1]; }
        out[2 * i] = re; out[2 * i + 1] = im;
    }
    for (int t = 0; t < nthreads; t++) free(tv[t]);
    free(tv);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Matrix-free gather form of the same operator, extended by the Heisenberg J */
/* and Gamma couplings (no reference counterpart; pinned by the dense         */
/* Kronecker builder).  Matrix elements per row r, s_k = +1 if bit k is 0:    */
/*   diagonal  sum_z (Kz/4) s_i s_j  [reference order, :344-360]              */
/*             + sum_all (J/4) s_i s_j + Bz * (sum_k s_k / 2)  [:365-375]      */
/*   x bond    Kx/4 + J/4 (1 - s_i s_j)                        [:113-132]      */
/*   y bond   -Ky/4 s_i s_j + J/4 (1 - s_i s_j)                [:140-158]      */
/*   z bond    J/4 (1 - s_i s_j) - i G/4 (s_i + s_j)                           */
/*   flip k    Bx/2 - i By/2 s_k                                [:167-198]     */
/*   Gamma     x bond (i,j): flips i and j, -i G/4 s_i s_j;                    */
/*             y bond (i,j): flip i gets G/4 s_j, flip j gets G/4 s_i.         */
/* Families with one bond are skipped as in the COO form when skip1 != 0.      */
/* ------------------------------------------------------------------------- */
/* Value of the input at column c: direct indexing, or (keys != NULL) the
 * entry of a sorted key list, for evaluating single rows of huge systems. */
static inline const double *colval(const double *in, const int64_t *keys, int64_t nkeys, int64_t c)
{
    if (!keys) return in + 2 * c;
    int64_t lo = 0, hi = nkeys - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (keys[mid] < c) lo = mid + 1; else hi = mid;
    }
    return in + 2 * lo;   /* the caller supplies every column a row needs */
}

static inline void hpsi_row(int N, const int *const *P, const int *nb, const double *K, double J, double G,
                            const double *B, const double *in, const int64_t *keys, int64_t nkeys,
                            int64_t r, double *out_re, double *out_im)
{
    double d = 0.0;
    for (int b = 0; b < nb[2]; b++) {
        int si = P[2][2 * b], sj = P[2][2 * b + 1];
        d += (bitof(r, si) == bitof(r, sj)) ? 0.25 * K[2] : -0.25 * K[2];
    }
    if (J != 0.0)
        for (int f = 0; f < 3; f++)
            for (int b = 0; b < nb[f]; b++) {
                int si = P[f][2 * b], sj = P[f][2 * b + 1];
                d += (bitof(r, si) == bitof(r, sj)) ? 0.25 * J : -0.25 * J;
            }
    double szt = 0.5 * (double)(N - 2 * __builtin_popcountll((unsigned long long)r));
    d += B[2] * szt;
    const double *x = colval(in, keys, nkeys, r);
    double re = d * x[0], im = d * x[1];
    for (int f = 0; f < 3; f++)
        for (int b = 0; b < nb[f]; b++) {
            int si = P[f][2 * b], sj = P[f][2 * b + 1];
            double s_i = bitof(r, si) ? -1.0 : 1.0, s_j = bitof(r, sj) ? -1.0 : 1.0;
            double cr = 0.25 * J * (1.0 - s_i * s_j), ci = 0.0;
            if (f == 0) cr += 0.25 * K[0];
            if (f == 1) cr += -0.25 * K[1] * s_i * s_j;
            if (f == 2) ci = -0.25 * G * (s_i + s_j);
            const double *y = colval(in, keys, nkeys, r ^ (((int64_t)1 << si) | ((int64_t)1 << sj)));
            re += cr * y[0] - ci * y[1];
            im += cr * y[1] + ci * y[0];
            if (G != 0.0 && f == 0) {            /* Gamma on x bond: single flips */
                double gi = -0.25 * G * s_i * s_j;
                const double *y1 = colval(in, keys, nkeys, r ^ ((int64_t)1 << si));
                const double *y2 = colval(in, keys, nkeys, r ^ ((int64_t)1 << sj));
                re += -gi * y1[1] - gi * y2[1];
                im += gi * y1[0] + gi * y2[0];
            }
            if (G != 0.0 && f == 1) {            /* Gamma on y bond: single flips */
                const double *y1 = colval(in, keys, nkeys, r ^ ((int64_t)1 << si));
                const double *y2 = colval(in, keys, nkeys, r ^ ((int64_t)1 << sj));
                re += 0.25 * G * s_j * y1[0] + 0.25 * G * s_i * y2[0];
                im += 0.25 * G * s_j * y1[1] + 0.25 * G * s_i * y2[1];
            }
        }
    for (int k = 0; k < N; k++) {
        double sk = bitof(r, k) ? -1.0 : 1.0;
        double cr = 0.5 * B[0], ci = -0.5 * B[1] * sk;
        const double *y = colval(in, keys, nkeys, r ^ ((int64_t)1 << k));
        re += cr * y[0] - ci * y[1];
        im += cr * y[1] + ci * y[0];
    }
    *out_re = re; *out_im = im;
}

int orc_hpsi_gather(int N, const int *pairs_x, int nbx, const int *pairs_y, int nby,
                    const int *pairs_z, int nbz, const double *K, double J, double G,
                    const double *B, int skip1, const double *in, double *out)
{
    int64_t dim = (int64_t)1 << N;
    const int *P[3] = {pairs_x, pairs_y, pairs_z};
    int nb[3] = {nbx, nby, nbz};
    if (skip1) for (int f = 0; f < 3; f++) if (nb[f] == 1) nb[f] = 0;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < dim; r++)
        hpsi_row(N, P, nb, K, J, G, B, in, NULL, 0, r, &out[2 * r], &out[2 * r + 1]);
    return 0;
}

/* The same rows for a list of row indices of a system too large to hold:
 * keys (sorted, ascending) lists every column those rows read and vals
 * (interleaved complex) their input values.  N may be up to 62. */
int orc_hpsi_rows(int N, const int *pairs_x, int nbx, const int *pairs_y, int nby,
                  const int *pairs_z, int nbz, const double *K, double J, double G,
                  const double *B, int skip1, int64_t nrows, const int64_t *rows,
                  int64_t nkeys, const int64_t *keys, const double *vals, double *out)
{
    const int *P[3] = {pairs_x, pairs_y, pairs_z};
    int nb[3] = {nbx, nby, nbz};
    if (skip1) for (int f = 0; f < 3; f++) if (nb[f] == 1) nb[f] = 0;
    for (int64_t q = 0; q < nrows; q++)
        hpsi_row(N, P, nb, K, J, G, B, vals, keys, nkeys, rows[q], &out[2 * q], &out[2 * q + 1]);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Single-site operators: restatement of Observables::ApplySx/Sy/Sz           */
/* (src/Observables.h:429-445,488-521) with QBasis coefficients               */
/* (src/QuantumBasis.h:48-74).  op: 0 = Sx, 1 = Sy, 2 = Sz.  Scatter form:    */
/* out(ketout) += coef * in(ket), out zero-initialised.                        */
/* ------------------------------------------------------------------------- */
void orc_apply_site_op(int N, int op, int site, const double *in, double *out)
{
    int64_t dim = (int64_t)1 << N;
    memset(out, 0, sizeof(double) * 2 * dim);
    for (int64_t ket = 0; ket < dim; ket++) {
        int bit = bitof(ket, site);
        int64_t ko = ket;
        double cr = 0.0, ci = 0.0;
        if (op == 0) { ko = ket ^ ((int64_t)1 << site); cr = 0.5; }
        else if (op == 1) { ko = ket ^ ((int64_t)1 << site); ci = bit == 0 ? 0.5 : -0.5; }
        else { cr = bit == 0 ? 0.5 : -0.5; }
        double a[2] = {cr, ci};
        cmac(out + 2 * ko, a, in + 2 * ket);
    }
}

static double *vnew(int64_t dim) { return (double *)calloc(2 * dim, sizeof(double)); }
static void vscale(int64_t dim, double s, double *v) { for (int64_t i = 0; i < 2 * dim; i++) v[i] = s * v[i]; }

/* Two-operator string: out = K * OpA_{sa} OpB_{sb} in  (OpB applied first). */
static double *str2(int N, const double *in, int opB, int sb, int opA, int sa, double K)
{
    int64_t dim = (int64_t)1 << N;
    double *t1 = vnew(dim), *t2 = vnew(dim);
    orc_apply_site_op(N, opB, sb, in, t1);
    orc_apply_site_op(N, opA, sa, t1, t2);
    vscale(dim, K, t2);
    free(t1);
    return t2;
}

static double *str3(int N, const double *in, int op1, int s1, int op2, int s2, int op3, int s3, double K)
{
    int64_t dim = (int64_t)1 << N;
    double *t1 = vnew(dim), *t2 = vnew(dim), *t3 = vnew(dim);
    orc_apply_site_op(N, op1, s1, in, t1);
    orc_apply_site_op(N, op2, s2, t1, t2);
    orc_apply_site_op(N, op3, s3, t2, t3);
    vscale(dim, K, t3);
    free(t1); free(t2);
    return t3;
}

/* Spin current at one site.  comp: 0 = x (src/Observables.h:243-281),        */
/* 1 = y (:202-240), 2 = z (:161-199).  The field parts are computed and then  */
/* dropped by the reference (return statements :198,:239,:280), so they are   */
/* not formed here.                                                            */
int orc_current_js_site(int N, const int *neigh, const double *K, int comp, int si,
                        const double *in, double *out)
{
    int64_t dim = (int64_t)1 << N;
    int xn = neigh[si * 3 + 2], yn = neigh[si * 3 + 1], zn = neigh[si * 3 + 0];
    if (xn < 0 || yn < 0 || zn < 0) return -1;
    double *L, *R;
    if (comp == 2) {        /* Kyy Sx_i Sy_y - Kxx Sy_i Sx_x */
        L = str2(N, in, 1, yn, 0, si, K[1]);
        R = str2(N, in, 0, xn, 1, si, K[0]);
    } else if (comp == 1) { /* Kxx Sz_i Sx_x - Kzz Sx_i Sz_z */
        L = str2(N, in, 0, xn, 2, si, K[0]);
        R = str2(N, in, 2, zn, 0, si, K[2]);
    } else {                /* Kzz Sy_i Sz_z - Kyy Sz_i Sy_y */
        L = str2(N, in, 2, zn, 1, si, K[2]);
        R = str2(N, in, 1, yn, 2, si, K[1]);
    }
    for (int64_t i = 0; i < 2 * dim; i++) out[i] = L[i] - R[i];
    free(L); free(R);
    return 0;
}

/* Energy current at one site: comp 0 = Jex (A site, src/Observables.h:358-389),*/
/* comp 1 = Jey (B site, :321-354).                                            */
int orc_current_je_site(int N, const int *neigh, const double *K, int comp, int si,
                        const double *in, double *out)
{
    int64_t dim = (int64_t)1 << N;
    double *L, *R;
    if (comp == 0) {
        int ix = neigh[si * 3 + 2];
        if (ix < 0) return -1;
        int ixpy = neigh[ix * 3 + 1], ixpz = neigh[ix * 3 + 0];
        if (ixpy < 0 || ixpz < 0) return -1;
        L = str3(N, in, 1, ixpy, 2, ix, 0, si, K[0] * K[1]);   /* Sx_i Sz_ix Sy_ixpy */
        R = str3(N, in, 2, ixpz, 1, ix, 0, si, K[0] * K[2]);   /* Sx_i Sy_ix Sz_ixpz */
        for (int64_t i = 0; i < 2 * dim; i++) out[i] = R[i] - L[i];
    } else {
        int iy = neigh[si * 3 + 1];
        if (iy < 0) return -1;
        int iypx = neigh[iy * 3 + 2], iypz = neigh[iy * 3 + 0];
        if (iypx < 0 || iypz < 0) return -1;
        L = str3(N, in, 0, iypx, 2, iy, 1, si, K[0] * K[1]);   /* Sy_i Sz_iy Sx_iypx */
        R = str3(N, in, 2, iypz, 0, iy, 1, si, K[1] * K[2]);   /* Sy_i Sx_iy Sz_iypz */
        for (int64_t i = 0; i < 2 * dim; i++) out[i] = L[i] - R[i];
    }
    free(L); free(R);
    return 0;
}

/* Totals: Js over all sites (src/Observables.h:664-699), Jex over even sites  */
/* (:448-458), Jey over odd sites (:460-470).  kind: 0..2 = Jsx..Jsz,          */
/* 3 = Jex, 4 = Jey.  out = 0; out += term(site) in site order.                */
int orc_current_total(int N, const int *neigh, const double *K, int kind,
                      const double *in, double *out)
{
    int64_t dim = (int64_t)1 << N;
    double *t = vnew(dim);
    memset(out, 0, sizeof(double) * 2 * dim);
    int start = (kind == 4) ? 1 : 0, step = (kind >= 3) ? 2 : 1;
    for (int s = start; s < N; s += step) {
        int rc = kind < 3 ? orc_current_js_site(N, neigh, K, kind, s, in, t)
                          : orc_current_je_site(N, neigh, K, kind - 3, s, in, t);
        if (rc) { free(t); return rc; }
        for (int64_t i = 0; i < 2 * dim; i++) out[i] += t[i];
    }
    free(t);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Start vector: restatement of DynLanczosSolver::Randomize                   */
/* (src/DynLanczosSolver.h:465-474) with ConstVariables::grnd                  */
/* (src/ParametersEngine.h:152-154): std::mt19937_64(RandomSeed) feeding      */
/* uniform_real_distribution<double>(0,1), i.e. double(g()) * 2^-64 with the  */
/* draw rounded to nearest (generate_canonical with one 64-bit draw).         */
/* mt19937_64 parameters are those of the C++ standard ([rand.predef]).       */
/* ------------------------------------------------------------------------- */
typedef struct { uint64_t mt[312]; int idx; } orc_mt64;

static void mt64_seed(orc_mt64 *g, uint64_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 312; i++)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = 312;
}

static uint64_t mt64_next(orc_mt64 *g)
{
    const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL, A = 0xB5026F5AA96619E9ULL;
    if (g->idx >= 312) {
        for (int i = 0; i < 312; i++) {
            uint64_t x = (g->mt[i] & UM) | (g->mt[(i + 1) % 312] & LM);
            uint64_t xa = x >> 1;
            if (x & 1ULL) xa ^= A;
            g->mt[i] = g->mt[(i + 156) % 312] ^ xa;
        }
        g->idx = 0;
    }
    uint64_t y = g->mt[g->idx++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return y;
}

/* Raw generator outputs, for the standard's known-answer check. */
void orc_mt19937_64(uint64_t seed, int64_t n, uint64_t *out)
{
    orc_mt64 g; mt64_seed(&g, seed);
    for (int64_t i = 0; i < n; i++) out[i] = mt64_next(&g);
}

/* v[i] = (-1 + 2u, -1 + 2u') drawn re first, then im; NOT normalised.        */
/* The reference was built against libc++ (macOS), whose generate_canonical   */
/* returns double(x) / 2^64 without clamping a draw that rounds up to 1.0;    */
/* that convention is followed here.                                          */
void orc_randomize(uint64_t seed, int64_t dim, double *v)
{
    orc_mt64 g; mt64_seed(&g, seed);
    const double two64 = 18446744073709551616.0;
    for (int64_t i = 0; i < dim; i++) {
        double u1 = (double)mt64_next(&g) / two64;
        double u2 = (double)mt64_next(&g) / two64;
        v[2 * i] = -1.0 + u1 * 2.0;
        v[2 * i + 1] = -1.0 + u2 * 2.0;
    }
}
