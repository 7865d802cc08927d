// esc_host.cpp -- host-side geometry and parameter-file handling of the C ABI
// (esc_honeycomb, esc_connectors, esc_chain, esc_chain_connectors, esc_params_load).
#include "esc.h"
#include "esc_internal.h"

#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct HostError : std::runtime_error {
    int code;
    HostError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

template <typename F>
int host_guard(F &&f) {
    try {
        f();
        esc::clear_error();
        return ESC_OK;
    } catch (const HostError &e) {
        esc::set_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        esc::set_error(e.what());
        return ESC_ERR_ARG;
    }
}

// Value lookup with the reference's rule (src/ParametersEngine.h:157-229): the
// text before the first '=' must equal the key exactly, and the rest must
// start with a readable value (leading blanks skipped); the first such line
// wins.  Returns false if no line matches.
template <typename T>
bool find_key(const std::vector<std::string> &lines, const std::string &key, T &value) {
    for (const std::string &line : lines) {
        std::istringstream is(line);
        std::string head;
        if (!std::getline(is, head, '=')) continue;
        T v{};
        if (is >> v && head == key) {
            value = v;
            return true;
        }
    }
    return false;
}

template <typename T>
T need_key(const std::vector<std::string> &lines, const std::string &key) {
    T v{};
    if (!find_key(lines, key, v)) throw HostError(ESC_ERR_ARG, key + "= argument is missing in the input file!");
    return v;
}

void copy_str(char *dst, const std::string &s) {
    std::strncpy(dst, s.c_str(), 63);
    dst[63] = 0;
}

}  // namespace

extern "C" int esc_honeycomb(int Lx, int Ly, int periodic_x, int periodic_y, int32_t *neigh) {
    return host_guard([&] {
        if (!neigh) throw HostError(ESC_ERR_ARG, "null neigh");
        if (Lx < 1 || Ly < 2 || Ly % 2 != 0) throw HostError(ESC_ERR_ARG, "honeycomb needs Lx >= 1 and an even Ly >= 2");
        const int LLX = Lx, LLY = Ly / 2, N = Lx * Ly;
        for (int k = 0; k < 3 * N; ++k) neigh[k] = -1;
        auto A = [&](int i, int j) { return 2 * (i * LLY + j); };
        auto link = [&](int a, int b, int col) {
            neigh[a * 3 + col] = b;
            neigh[b * 3 + col] = a;
        };
        for (int i = 0; i < LLX; ++i)
            for (int j = 0; j < LLY; ++j) {
                const int a = A(i, j), b = a + 1;
                link(a, b, 2);                                     // x bond inside the cell
                if (i + 1 < LLX) link(b, A(i + 1, j), 1);          // y bond to the next cell along a1
                else if (periodic_x) link(A(0, j), b, 1);          // wrap along a1
                if (j + 1 < LLY) link(b, A(i, j + 1), 0);          // z bond to the next cell along a2
                else if (periodic_y) link(b, A(i, 0), 0);          // wrap along a2
            }
    });
}

extern "C" int esc_connectors(int n_sites, const int32_t *neigh, const double *K, int32_t *pairs_x, int32_t *pairs_y,
                              int32_t *pairs_z, int32_t *counts) {
    return host_guard([&] {
        if (!neigh || !K || !pairs_x || !pairs_y || !pairs_z || !counts) throw HostError(ESC_ERR_ARG, "null argument");
        if (n_sites < 1) throw HostError(ESC_ERR_ARG, "n_sites must be positive");
        int32_t *out[3] = {pairs_x, pairs_y, pairs_z};
        const int col[3] = {2, 1, 0};   // x, y, z families live in neighbour columns 2, 1, 0
        for (int f = 0; f < 3; ++f) {
            int c = 0;
            if (K[f] != 0.0)
                for (int i = 0; i < n_sites; ++i) {
                    // pairs come out sorted by (i, j): for a fixed i the partner
                    // in one family is unique
                    const int j = neigh[i * 3 + col[f]];
                    if (j > i && j < n_sites) {
                        out[f][2 * c] = i;
                        out[f][2 * c + 1] = j;
                        ++c;
                    }
                }
            counts[f] = c;
        }
    });
}

extern "C" int esc_chain(int n_sites, int periodic, int32_t *neigh) {
    return host_guard([&] {
        if (!neigh) throw HostError(ESC_ERR_ARG, "null argument");
        if (n_sites < 2 || n_sites > 64) throw HostError(ESC_ERR_ARG, "n_sites must be in [2, 64]");
        // Lattice::Chain (src/Lattice.h:60-125): column 0 = +x neighbour,
        // column 1 = -x neighbour, wrapped when periodic
        for (int i = 0; i < n_sites; ++i) {
            neigh[2 * i] = (i + 1 < n_sites) ? i + 1 : (periodic ? 0 : -1);
            neigh[2 * i + 1] = (i - 1 >= 0) ? i - 1 : (periodic ? n_sites - 1 : -1);
        }
    });
}

extern "C" int esc_chain_connectors(int n_sites, const int32_t *neigh, const double *K, int32_t *pairs_x,
                                    int32_t *pairs_y, int32_t *pairs_z, int32_t *counts) {
    return host_guard([&] {
        if (!neigh || !K || !pairs_x || !pairs_y || !pairs_z || !counts) throw HostError(ESC_ERR_ARG, "null argument");
        if (n_sites < 2) throw HostError(ESC_ERR_ARG, "n_sites must be >= 2");
        // HeisenbergChain_Connectors (src/Hamiltonian.h:381-475): every family
        // holds the chain bonds (i < j from both neighbour columns) of a
        // non-zero coupling, listed in row-major order of the coupling matrix
        std::vector<char> bond((size_t)n_sites * n_sites, 0);
        for (int i = 0; i < n_sites; ++i)
            for (int c = 0; c < 2; ++c) {
                const int j = neigh[2 * i + c];
                if (j > i && j < n_sites) bond[(size_t)i * n_sites + j] = 1;
            }
        int32_t *out[3] = {pairs_x, pairs_y, pairs_z};
        for (int f = 0; f < 3; ++f) {
            int n = 0;
            if (K[f] != 0.0)
                for (int i = 0; i < n_sites; ++i)
                    for (int j = 0; j < n_sites; ++j)
                        if (bond[(size_t)i * n_sites + j]) {
                            out[f][2 * n] = i;
                            out[f][2 * n + 1] = j;
                            ++n;
                        }
            counts[f] = n;
        }
    });
}

extern "C" int esc_params_load(const char *path, esc_params *p) {
    return host_guard([&] {
        if (!path || !p) throw HostError(ESC_ERR_ARG, "null argument");
        std::ifstream f(path);
        if (!f) throw HostError(ESC_ERR_ARG, std::string("cannot open input file ") + path);
        std::vector<std::string> lines;
        for (std::string ln; std::getline(f, ln);) lines.push_back(ln);
        std::memset(p, 0, sizeof(*p));
        p->Lx = (int32_t)need_key<double>(lines, "NumberofSitesLx");
        p->Ly = (int32_t)need_key<double>(lines, "NumberofSitesLy");
        p->NumberofSites = p->Lx * p->Ly;
        p->RandomSeed = (int32_t)need_key<double>(lines, "RandomSeed");
        copy_str(p->Model, need_key<std::string>(lines, "Model"));
        copy_str(p->Solver, need_key<std::string>(lines, "Solver"));
        p->Kxx = need_key<double>(lines, "Kxx");
        p->Kyy = need_key<double>(lines, "Kyy");
        p->Kzz = need_key<double>(lines, "Kzz");
        p->Bxx = need_key<double>(lines, "Bxx");
        p->Byy = need_key<double>(lines, "Byy");
        p->Bzz = need_key<double>(lines, "Bzz");
        p->Temperature = need_key<double>(lines, "Temperature");
        p->LanczosEps = need_key<double>(lines, "LanczosEps");
        p->IsPeriodicX = need_key<double>(lines, "IsPeriodicX") != 0.0;
        p->IsPeriodicY = need_key<double>(lines, "IsPeriodicY") != 0.0;
        copy_str(p->LancType, need_key<std::string>(lines, "LancType"));
        p->Threads = (int32_t)need_key<double>(lines, "Threads");
        double v;
        p->LanczosSteps = find_key(lines, "LanczosSteps", v) ? (int32_t)v : 200;
        // extensions (no reference counterpart); parsed before the checks that
        // depend on the spectral mode
        std::string s;
        p->J = find_key(lines, "J", v) ? v : 0.0;
        p->Gamma = find_key(lines, "Gamma", v) ? v : 0.0;
        p->Eta = find_key(lines, "Eta", v) ? v : 0.005;
        p->OmegaMin = find_key(lines, "OmegaMin", v) ? v : 0.0;
        p->OmegaMax = find_key(lines, "OmegaMax", v) ? v : 2.0;
        p->OmegaPoints = find_key(lines, "OmegaPoints", v) ? (int32_t)v : 0;
        p->Precision = ESC_C128;
        if (find_key(lines, "Precision", s)) {
            if (s == "c64") p->Precision = ESC_C64;
            else if (s != "c128") throw HostError(ESC_ERR_ARG, "Precision must be c128 or c64");
        }
        p->SpectralMode = ESC_SPEC_RITZ;
        if (find_key(lines, "SpectralMode", s)) {
            if (s == "CF") p->SpectralMode = ESC_SPEC_CF;
            else if (s == "CFReorth") p->SpectralMode = ESC_SPEC_CF_REORTH;
            else if (s != "Ritz") throw HostError(ESC_ERR_ARG, "SpectralMode must be Ritz, CF or CFReorth");
        }
        p->StartVector = ESC_START_REF_UNIFORM;
        if (find_key(lines, "StartVector", s)) {
            if (s == "Gaussian") p->StartVector = ESC_START_GAUSSIAN;
            else if (s != "Uniform") throw HostError(ESC_ERR_ARG, "StartVector must be Uniform or Gaussian");
        }
        p->CFSteps = find_key(lines, "CFSteps", v) ? (int32_t)v : p->LanczosSteps;
        // dynamics flags default to 0; the Ritz spectra need LancType=All
        // (src/ParametersEngine.h:113-144)
        const bool ritz = p->SpectralMode == ESC_SPEC_RITZ;
        const std::string lt = p->LancType;
        const char *dyn_keys[5] = {"LocalSkw", "SpinCurrDynamics", "EnergyCurrDynamics", "LocalSzSz", "LocalRIXS"};
        int32_t *dyn_vals[5] = {&p->LocalSkw, &p->SpinCurrDynamics, &p->EnergyCurrDynamics, &p->LocalSzSz,
                                &p->LocalRIXS};
        for (int k = 0; k < 5; ++k) {
            *dyn_vals[k] = find_key(lines, dyn_keys[k], v) ? (v != 0.0) : 0;
            if (*dyn_vals[k] && ritz && lt != "All")
                throw HostError(ESC_ERR_ARG, "You need to set LancType=All for Dynamical calculations --- you dummy!");
        }
        // DynSzSz runs its own Lanczos (src/dynamics.h:575-576) and needs no kept basis
        p->SzSz = find_key(lines, "SzSz", v) ? (v != 0.0) : 0;
        p->SpecialSite = find_key(lines, "SpecialSite", v) ? (int32_t)v : p->NumberofSites / 2;
        if (p->SpecialSite < 0 || p->SpecialSite >= p->NumberofSites)
            throw HostError(ESC_ERR_ARG, "SpecialSite must be a site index in [0, NumberofSites)");
    });
}
