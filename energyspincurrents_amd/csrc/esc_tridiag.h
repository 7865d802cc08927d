// esc_tridiag.h -- FP64 routines on the Lanczos tridiagonal matrix T and its
// continued fraction, written once for the host and the device (the batched
// solves run them in kernels, the single-system paths on the host).  No
// multiply-add is contracted, so both sides compute the same bits.
//
// T is given as a[0..n-1] (diagonal) and b[1..n-1] (b[i] couples rows i-1
// and i; b[0] is not read), the layout of the alpha / beta rows the
// recursions record (src/DynLanczosSolver.h:94-101,371-414).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

namespace esc {

// length of a recorded recursion: up to and including the first step whose
// beta_{k+1} came out zero (exhausted Krylov space), else `steps`
__host__ __device__ inline int td_length(const double *beta, int steps) {
    for (int k = 0; k < steps; ++k)
        if (beta[k + 1] == 0.0) return k + 1;
    return steps;
}

// Gershgorin interval [lo, hi] of T and the Sturm-count pivot floor; false
// if T has a non-finite entry
__host__ __device__ inline bool td_bounds(int n, const double *a, const double *b, double *lo, double *hi,
                                          double *pivmin) {
#pragma clang fp contract(off)
    double l = INFINITY, h = -INFINITY, bmax = 0.0;
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? fabs(b[i]) : 0.0) + (i + 1 < n ? fabs(b[i + 1]) : 0.0);
        l = fmin(l, a[i] - r);
        h = fmax(h, a[i] + r);
        if (i > 0) bmax = fmax(bmax, b[i] * b[i]);
    }
    if (!(l > -INFINITY && h < INFINITY)) return false;
    *lo = l;
    *hi = h;
    *pivmin = DBL_MIN * fmax(1.0, bmax);
    return true;
}

// number of eigenvalues of T below x: the negative pivots of
// q_i = (a_i - x) - b_i^2 / q_{i-1}, pivots below pivmin in magnitude taken
// as -pivmin
__host__ __device__ inline int td_below(int n, const double *a, const double *b, double x, double pivmin) {
#pragma clang fp contract(off)
    int c = 0;
    double q = a[0] - x;
    for (int i = 0;;) {
        if (fabs(q) < pivmin) q = -pivmin;
        if (q < 0.0) ++c;
        if (++i == n) break;
        q = (a[i] - x) - (b[i] * b[i]) / q;
    }
    return c;
}

// Bisection bracket of the whole spectrum: the Gershgorin interval widened so
// that no eigenvalue lies below lo; false for a non-finite T
__host__ __device__ inline bool td_range(int n, const double *a, const double *b, double *lo, double *hi,
                                         double *pivmin) {
#pragma clang fp contract(off)
    if (!td_bounds(n, a, b, lo, hi, pivmin)) return false;
    const double margin = 2.0 * DBL_EPSILON * fmax(fabs(*lo), fabs(*hi)) + 2.0 * *pivmin;
    *lo -= margin;
    *hi += margin;
    return true;
}

// The k-th smallest eigenvalue (k = 0 first) by bisection on the Sturm count
// inside [lo, hi] (below(lo) <= k < below(hi)): O(n) per halving, within a
// few ulp of |lambda| (or 2^-128 of the Gershgorin width)
__host__ __device__ inline double td_kth(int n, const double *a, const double *b, int k, double lo, double hi,
                                         double pivmin) {
#pragma clang fp contract(off)
    for (int it = 0; it < 256; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (!(mid > lo && mid < hi)) break;   // interval at machine resolution
        if (hi - lo <= 4.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi))) break;
        if (td_below(n, a, b, mid, pivmin) >= k + 1) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

// Lowest eigenvalue; *ok = false for a non-finite T
__host__ __device__ inline double td_lowest(int n, const double *a, const double *b, bool *ok) {
#pragma clang fp contract(off)
    *ok = true;
    if (n == 1) {
        *ok = a[0] > -INFINITY && a[0] < INFINITY;
        return a[0];
    }
    double lo, hi, pivmin;
    if (!td_range(n, a, b, &lo, &hi, &pivmin)) {
        *ok = false;
        return NAN;
    }
    return td_kth(n, a, b, 0, lo, hi, pivmin);
}

// Eigenvector of T for the eigenvalue lam by inverse iteration: T - lam I is
// factored once by Gaussian elimination with partial pivoting (the row
// interchanges of LAPACK's dgttrf: L unit lower bidiagonal, U with two
// superdiagonals), pivots smaller than eps ||T|| replaced by eps ||T||, and
// three solves from a pseudo-random start (seed), each normalised.  x (n)
// gets the unit eigenvector, its first nonzero component positive.  work
// holds 5n doubles.  Inside a cluster of equal eigenvalues (the repeated
// copies a recursion without reorthogonalisation produces) the result is one
// unit vector of the cluster's invariant subspace; with prev (nprev unit
// vectors, stride ldp) each iterate is orthogonalised against them, as
// LAPACK's dstein does within a cluster.
__host__ __device__ inline void td_vector(int n, const double *a, const double *b, double lam, double *work,
                                          double *x, uint64_t seed = 0x9e3779b97f4a7c15ull,
                                          const double *prev = nullptr, int nprev = 0, int ldp = 0) {
#pragma clang fp contract(off)
    if (n == 1) {
        x[0] = 1.0;
        return;
    }
    double *dl = work, *d = work + n, *du = work + 2 * n, *du2 = work + 3 * n, *piv = work + 4 * n;
    double tnorm = 0.0;
    for (int i = 0; i < n; ++i) {
        d[i] = a[i] - lam;
        tnorm = fmax(tnorm, fabs(a[i]) + (i > 0 ? fabs(b[i]) : 0.0) + (i + 1 < n ? fabs(b[i + 1]) : 0.0));
        if (i + 1 < n) { dl[i] = b[i + 1]; du[i] = b[i + 1]; }
        du2[i] = 0.0;
        piv[i] = 0.0;
    }
    const double tiny = DBL_EPSILON * fmax(tnorm, DBL_MIN);
    for (int i = 0; i + 1 < n; ++i) {
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] == 0.0) d[i] = tiny;
            const double f = dl[i] / d[i];
            dl[i] = f;
            d[i + 1] -= f * du[i];
        } else {
            // interchange rows i and i+1
            const double f = d[i] / dl[i];
            d[i] = dl[i];
            dl[i] = f;
            const double t = du[i];
            du[i] = d[i + 1];
            d[i + 1] = t - f * d[i + 1];
            if (i + 2 < n) {
                du2[i] = du[i + 1];
                du[i + 1] = -f * du[i + 1];
            }
            piv[i] = 1.0;
        }
    }
    for (int i = 0; i < n; ++i)
        if (fabs(d[i]) < tiny) d[i] = d[i] < 0.0 ? -tiny : tiny;
    // start: uniform in (-1, 1) from a linear congruential sequence
    uint64_t s = seed;
    for (int i = 0; i < n; ++i) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        x[i] = (double)(s >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    for (int pass = 0; pass < 3; ++pass) {
        // x = L^-1 P x
        for (int i = 0; i + 1 < n; ++i) {
            if (piv[i] != 0.0) {
                const double t = x[i];
                x[i] = x[i + 1];
                x[i + 1] = t - dl[i] * x[i];
            } else {
                x[i + 1] -= dl[i] * x[i];
            }
        }
        // x = U^-1 x
        x[n - 1] /= d[n - 1];
        x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / d[n - 2];
        for (int i = n - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / d[i];
        // against the earlier vectors of the cluster (modified Gram-Schmidt),
        // so that equal or close eigenvalues get orthogonal vectors
        if (nprev > 0) {
            double m = 0.0;
            for (int i = 0; i < n; ++i) m = fmax(m, fabs(x[i]));
            if (m > 0.0 && m < INFINITY)
                for (int i = 0; i < n; ++i) x[i] /= m;
            for (int q = 0; q < nprev; ++q) {
                const double *p = prev + (size_t)q * ldp;
                double dp = 0.0;
                for (int i = 0; i < n; ++i) dp += p[i] * x[i];
                for (int i = 0; i < n; ++i) x[i] -= dp * p[i];
            }
        }
        // normalise, guarding against overflow of the sum of squares
        double m = 0.0;
        for (int i = 0; i < n; ++i) m = fmax(m, fabs(x[i]));
        if (!(m > 0.0) || !(m < INFINITY)) {
            for (int i = 0; i < n; ++i) x[i] = i == 0 ? 1.0 : 0.0;
            return;
        }
        double s2 = 0.0;
        for (int i = 0; i < n; ++i) {
            x[i] /= m;
            s2 += x[i] * x[i];
        }
        const double r = 1.0 / sqrt(s2);
        for (int i = 0; i < n; ++i) x[i] *= r;
    }
    int f = 0;
    while (f + 1 < n && x[f] == 0.0) ++f;
    if (x[f] < 0.0)
        for (int i = 0; i < n; ++i) x[i] = -x[i];
}

// All eigenpairs.  Eigenvalues are split into clusters where consecutive
// ones lie within 1e-3 |T|_1 (LAPACK dstein's criterion); a cluster's
// vectors are computed in order, each orthogonalised against the ones before
// it.  td_eig_norm: |T|_1; td_eig_cluster_start: whether eigenvalue j starts a
// cluster; td_eig_vectors: the vectors of the cluster starting at j0 into
// Z[j * n + i] (column j = eigenvector j).
__host__ __device__ inline double td_eig_norm(int n, const double *a, const double *b) {
#pragma clang fp contract(off)
    double t = 0.0;
    for (int i = 0; i < n; ++i)
        t = fmax(t, fabs(a[i]) + (i > 0 ? fabs(b[i]) : 0.0) + (i + 1 < n ? fabs(b[i + 1]) : 0.0));
    return t;
}

__host__ __device__ inline bool td_eig_cluster_start(const double *ev, int j, double tnorm) {
#pragma clang fp contract(off)
    return j == 0 || ev[j] - ev[j - 1] > 1e-3 * tnorm;
}

__host__ __device__ inline void td_eig_vectors(int n, const double *a, const double *b, const double *ev, int j0,
                                               double tnorm, double *work, double *Z) {
    int j1 = j0 + 1;
    while (j1 < n && !td_eig_cluster_start(ev, j1, tnorm)) ++j1;
    for (int j = j0; j < j1; ++j)
        td_vector(n, a, b, ev[j], work, Z + (size_t)j * n, 0x9e3779b97f4a7c15ull + 0x632be59bd9b4e019ull * (uint64_t)j,
                  Z + (size_t)j0 * n, j - j0, n);
}

// n2 <phi|(z - H)^-1|phi> / n2 as the continued fraction of the recursion
// (alpha[k * sa], beta[k], k < it) at z = zr + i zi, zi < 0 (src/dynamics.h:
// 270-336).  Each level is 1 / (x + i y) with y <= zi < 0 (the tail's
// imaginary part is never negative), so (x - i y) / (x^2 + y^2) stays in
// range.  Products are rounded before the sums on both sides.
__host__ __device__ inline void td_cf(int it, const double *alpha, int sa, const double *beta, double n2, double zr,
                                      double zi, double *fr, double *fi) {
#pragma clang fp contract(off)
    double gr = 0.0, gi = 0.0;
    for (int k = it - 1; k >= 0; --k) {
        double xr = zr - alpha[(int64_t)k * sa], xi = zi;
        if (k + 1 < it) {
            const double b2 = beta[k + 1] * beta[k + 1];
            xr -= b2 * gr;
            xi -= b2 * gi;
        }
        const double dd = xr * xr + xi * xi;
        gr = xr / dd;
        gi = -xi / dd;
    }
    *fr = n2 * gr;
    *fi = n2 * gi;
}

}  // namespace esc
