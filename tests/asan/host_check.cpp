// Exercises the host-only paths of the C ABI under AddressSanitizer and
// UndefinedBehaviorSanitizer: the parameter parser, lattices and bond
// families, the FP64 tridiagonal eigensolver and the continued-fraction and
// Lehmann spectra, and the pass planners of the tiled H.psi kernels (generic
// and generated) with the kernel generator.  No device is used.  Prints
// "host checks passed" and exits 0 when every call behaves.
#include "esc.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

static int failures = 0;
#define EXPECT(cond, what)                                              \
    do {                                                                \
        if (!(cond)) {                                                  \
            std::fprintf(stderr, "FAILED: %s (%s)\n", what, esc_last_error()); \
            ++failures;                                                 \
        }                                                               \
    } while (0)

struct Model {
    std::vector<int32_t> neigh, bonds[3];
    esc_model m{};
};

static Model honeycomb(int lx, int ly, double K, double J, double G, double h) {
    Model md;
    const int N = lx * ly;
    md.neigh.assign(3 * N, 0);
    EXPECT(esc_honeycomb(lx, ly, 1, 1, md.neigh.data()) == ESC_OK, "esc_honeycomb");
    for (auto &b : md.bonds) b.assign(2 * N, 0);
    int32_t cnt[3] = {0, 0, 0};
    const double Kv[3] = {1.0, 1.0, 1.0};
    EXPECT(esc_connectors(N, md.neigh.data(), Kv, md.bonds[0].data(), md.bonds[1].data(), md.bonds[2].data(), cnt) ==
               ESC_OK,
           "esc_connectors");
    md.m.n_sites = N;
    for (int f = 0; f < 3; ++f) {
        md.m.n_bonds[f] = cnt[f];
        md.m.bonds[f] = md.bonds[f].data();
        md.m.K[f] = K;
        md.m.B[f] = h / std::sqrt(3.0);
    }
    md.m.neigh = md.neigh.data();
    md.m.J = J;
    md.m.Gamma = G;
    return md;
}

int main(int argc, char **argv) {
    const std::string golden = argc > 1 ? argv[1] : "tests/golden/Midinput.inp";
    // parameters: the reference's file, then a missing file
    esc_params p{};
    EXPECT(esc_params_load(golden.c_str(), &p) == ESC_OK, "esc_params_load");
    EXPECT(esc_params_load("/nonexistent/input.inp", &p) != ESC_OK, "missing file is an error");
    // lattices and chains of many sizes, including invalid ones
    for (int lx = 1; lx <= 4; ++lx)
        for (int ly = 2; ly <= 8; ly += 2) {
            std::vector<int32_t> nb(3 * lx * ly);
            EXPECT(esc_honeycomb(lx, ly, lx % 2, ly % 4 == 0, nb.data()) == ESC_OK, "honeycomb sizes");
        }
    {
        std::vector<int32_t> nb(3 * 8);
        EXPECT(esc_honeycomb(2, 3, 1, 1, nb.data()) != ESC_OK, "odd Ly is refused");
        std::vector<int32_t> cn(2 * 12), px(24), py(24), pz(24);
        int32_t cnt[3];
        const double K[3] = {1.0, 0.5, 0.0};
        EXPECT(esc_chain(12, 1, cn.data()) == ESC_OK, "esc_chain");
        EXPECT(esc_chain_connectors(12, cn.data(), K, px.data(), py.data(), pz.data(), cnt) == ESC_OK,
               "esc_chain_connectors");
    }
    // tridiagonal eigensolver, continued fraction and Lehmann sums
    std::mt19937_64 rng(7);
    std::normal_distribution<double> nd;
    for (int n : {1, 2, 5, 40, 300}) {
        std::vector<double> a(n), b(n + 1, 0.0), ev(n), Z((size_t)n * n);
        for (int i = 0; i < n; ++i) a[i] = nd(rng);
        for (int i = 1; i <= n; ++i) b[i] = std::fabs(nd(rng)) + (i % 7 == 0 ? 0.0 : 0.1);
        EXPECT(esc_tridiag_eig(n, a.data(), b.data(), ev.data(), Z.data()) == ESC_OK, "esc_tridiag_eig");
        for (int i = 1; i < n; ++i) EXPECT(ev[i] >= ev[i - 1], "eigenvalues ascending");
        std::vector<double> w(101), out(2 * 101), Oa(2 * n), Ob(2 * n);
        for (int q = 0; q <= 100; ++q) w[q] = 0.06 * q;
        EXPECT(esc_cf_spectrum(n, a.data(), b.data(), 1.3, -2.0, 0.05, w.data(), 101, nullptr, out.data()) == ESC_OK,
               "esc_cf_spectrum");
        for (int i = 0; i < 2 * n; ++i) { Oa[i] = nd(rng); Ob[i] = nd(rng); }
        const double el[2] = {0.1, -0.2};
        EXPECT(esc_lehmann_spectrum(n, ev.data(), Oa.data(), Ob.data(), el, 0.05, w.data(), 101, out.data()) == ESC_OK,
               "esc_lehmann_spectrum");
        for (double x : out) EXPECT(std::isfinite(x), "finite spectrum");
    }
    // pass plans of many kernel configurations, and generated sources
    for (auto geo : {std::pair<int, int>(2, 8), {3, 6}, {2, 10}, {3, 8}, {4, 8}}) {
        Model md = honeycomb(geo.first, geo.second, -1.0, 0.1, 0.2, 0.2);
        const int N = md.m.n_sites;
        int planned = 0;
        for (int prec : {ESC_C128, ESC_C64})
            for (int P : {1, 5, 6, 7, 8})
                for (int tb : {11, 12, 13})
                    for (int w : {17, 18}) {
                        const int ww = w > N ? N : w;
                        if ((P == 1) != (ww == N)) continue;
                        // generated-kernel plans (ten local searches each for complex128):
                        // one shape per lattice and order
                        if (P >= 7 && (tb != 12 || w != 18)) continue;
                        const int k = P >= 7 ? 3 : 0;
                        const int cfg = 1000000 * k + 100000 * P + 30000 + 100 * tb + ww;
                        std::vector<int64_t> rows(6 * 16);
                        std::vector<int32_t> tp(256);
                        std::vector<uint32_t> tm(256);
                        int np = 0, nt = 0;
                        const int rc = esc_kernel_plan(&md.m, prec, cfg, rows.data(), 16, &np, tp.data(), tm.data(),
                                                       256, &nt);
                        EXPECT(rc == ESC_OK || rc == ESC_ERR_UNSUPPORTED, "esc_kernel_plan status");
                        if (rc != ESC_OK) continue;
                        ++planned;
                        for (int t = 0; t < nt; ++t) EXPECT(tp[t] >= 0 && tp[t] < np, "every term in a pass");
                    }
        EXPECT(planned >= 4, "configurations planned");
        size_t len = 0;
        const int rc = esc_kernel_precompile(&md.m, ESC_C128, 3731200 + (N > 18 ? 18 : N), nullptr, 0, &len);
        EXPECT(rc == ESC_OK && len > 1000, "generated source");
    }
    if (failures) {
        std::fprintf(stderr, "%d host checks failed\n", failures);
        return 1;
    }
    std::printf("host checks passed\n");
    return 0;
}
