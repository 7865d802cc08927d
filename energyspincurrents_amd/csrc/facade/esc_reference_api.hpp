// esc_reference_api.hpp -- the reference's class interfaces over the C ABI.
//
// A program written against the reference (main.cpp + src/*.h of
// fengshi96/EnergySpinCurrents) finds the same class names, constructors and
// methods here:
//   ConstVariables(string)                    src/ParametersEngine.h:21-44
//   Lattice(ConstVariables&)                  src/Lattice.h:21-58
//   QBasis(ConstVariables&)                   src/QuantumBasis.h:18-163
//   Hamiltonian(ConstVariables&, Lattice&, QBasis&), TightB_Ham, Diagonal,
//       MatrixVectorMult                      src/Hamiltonian.h:18-557
//   DynLanczosSolver(ConstVariables&, QBasis&, Hamiltonian&), Lanczos_Nirav,
//       TriDeval, PsiAll, Vorig               src/DynLanczosSolver.h:24-239
//   Observables(...), ApplyOp, ApplyJ*Total, measureLocal*   src/Observables.h
//   StartDynamics(string&, ...)               src/dynamics.h:43-60
// Vectors are device-resident (esc::VectorXcd replaces Eigen::VectorXcd); the
// H.psi, operators, Lanczos and spectra run in libesc.so on the GPU.
#pragma once

#include "esc.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace esc {

using dcomplex = std::complex<double>;

inline void check(int rc) {
    if (rc == ESC_OK) return;
    const std::string msg = esc_last_error();
    if (rc == ESC_ERR_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// --------------------------------------------------------------------------
// Device vector: owns 2^N complex entries of the context precision.
// --------------------------------------------------------------------------
class Context;

class VectorXcd {
public:
    VectorXcd() = default;
    explicit VectorXcd(std::shared_ptr<Context> ctx);
    VectorXcd(const VectorXcd &o);
    VectorXcd &operator=(const VectorXcd &o);
    VectorXcd(VectorXcd &&o) noexcept { swap(o); }
    VectorXcd &operator=(VectorXcd &&o) noexcept { swap(o); return *this; }
    ~VectorXcd() { release(); }

    int64_t size() const;
    void *data() const { return d_; }
    const std::shared_ptr<Context> &context() const { return ctx_; }

    // Eigen-style helpers used by the reference's code paths
    dcomplex dot(const VectorXcd &y) const;        // <this|y>, conjugate-linear in this
    double norm() const { return std::sqrt(dot(*this).real()); }
    void normalize();
    void setZero();
    std::vector<dcomplex> to_host() const;
    dcomplex coeff(int64_t i) const;               // one element (device read)
    dcomplex operator()(int64_t i) const { return coeff(i); }
    VectorXcd operator+(const VectorXcd &o) const { return axpby(1.0, o, 1.0); }
    VectorXcd operator-(const VectorXcd &o) const { return axpby(-1.0, o, 1.0); }
    VectorXcd &operator+=(const VectorXcd &o);
    VectorXcd scaled(dcomplex s) const { return axpby(0.0, *this, s); }

private:
    VectorXcd axpby(dcomplex a, const VectorXcd &x, dcomplex b) const;   // a x + b this
    void swap(VectorXcd &o) noexcept { std::swap(ctx_, o.ctx_); std::swap(d_, o.d_); }
    void release();
    std::shared_ptr<Context> ctx_;
    void *d_ = nullptr;
};

inline VectorXcd operator*(dcomplex s, const VectorXcd &v) { return v.scaled(s); }
inline VectorXcd operator*(double s, const VectorXcd &v) { return v.scaled(dcomplex(s, 0.0)); }

// --------------------------------------------------------------------------
// Context wrapper
// --------------------------------------------------------------------------
class Context {
public:
    Context(const esc_model &m, int precision) { check(esc_ctx_create(&m, precision, &h_)); info(); }
    ~Context() { esc_ctx_destroy(h_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    esc_ctx *get() const { return h_; }
    int64_t dim() const { return dim_; }
    int32_t elem_bytes() const { return esz_; }
    // one Lanczos basis is kept per context; each run gets a number so that
    // the Ritz states of an earlier run are not read from a later run's basis
    int lanczos_runs = 0;

private:
    void info() { check(esc_ctx_info(h_, &dim_, &esz_, nullptr)); }
    esc_ctx *h_ = nullptr;
    int64_t dim_ = 0;
    int32_t esz_ = 16;
};

inline VectorXcd::VectorXcd(std::shared_ptr<Context> ctx) : ctx_(std::move(ctx)) {
    if (hipMalloc(&d_, ctx_->dim() * ctx_->elem_bytes()) != hipSuccess)
        throw std::runtime_error("device allocation of a state vector failed");
}
inline VectorXcd::VectorXcd(const VectorXcd &o) : ctx_(o.ctx_) {
    if (!ctx_) return;
    if (hipMalloc(&d_, ctx_->dim() * ctx_->elem_bytes()) != hipSuccess)
        throw std::runtime_error("device allocation of a state vector failed");
    if (hipMemcpy(d_, o.d_, ctx_->dim() * ctx_->elem_bytes(), hipMemcpyDeviceToDevice) != hipSuccess)
        throw std::runtime_error("device copy failed");
}
inline VectorXcd &VectorXcd::operator=(const VectorXcd &o) {
    if (this != &o) { VectorXcd t(o); swap(t); }
    return *this;
}
inline void VectorXcd::release() {
    if (d_) (void)hipFree(d_);
    d_ = nullptr;
}
inline int64_t VectorXcd::size() const { return ctx_ ? ctx_->dim() : 0; }
inline dcomplex VectorXcd::dot(const VectorXcd &y) const {
    double r[2];
    check(esc_dot(ctx_->get(), d_, y.d_, r, nullptr));
    return {r[0], r[1]};
}
inline void VectorXcd::normalize() {
    const double n = norm();
    const double a[2] = {0.0, 0.0}, b[2] = {1.0 / n, 0.0};
    check(esc_axpby(ctx_->get(), a, d_, b, d_, nullptr));
}
inline void VectorXcd::setZero() {
    if (hipMemset(d_, 0, ctx_->dim() * ctx_->elem_bytes()) != hipSuccess) throw std::runtime_error("hipMemset failed");
}
inline std::vector<dcomplex> VectorXcd::to_host() const {
    std::vector<dcomplex> h((size_t)size());
    check(esc_download(ctx_->get(), d_, reinterpret_cast<double *>(h.data()), nullptr));
    return h;
}
inline dcomplex VectorXcd::coeff(int64_t i) const {
    if (ctx_->elem_bytes() == 16) {
        dcomplex v;
        if (hipMemcpy(&v, (char *)d_ + 16 * i, 16, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("read failed");
        return v;
    }
    std::complex<float> v;
    if (hipMemcpy(&v, (char *)d_ + 8 * i, 8, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("read failed");
    return {v.real(), v.imag()};
}
inline VectorXcd VectorXcd::axpby(dcomplex a, const VectorXcd &x, dcomplex b) const {
    VectorXcd out(*this);
    const double av[2] = {a.real(), a.imag()}, bv[2] = {b.real(), b.imag()};
    check(esc_axpby(ctx_->get(), av, x.d_, bv, out.d_, nullptr));
    return out;
}
inline VectorXcd &VectorXcd::operator+=(const VectorXcd &o) {
    const double av[2] = {1.0, 0.0}, bv[2] = {1.0, 0.0};
    check(esc_axpby(ctx_->get(), av, o.d_, bv, d_, nullptr));
    return *this;
}

// --------------------------------------------------------------------------
// ConstVariables (src/ParametersEngine.h:21-44): esc_params_load + the seeded
// generator.  Extension keys are documented in include/esc.h.
// --------------------------------------------------------------------------
struct ConstVariables {
    explicit ConstVariables(const std::string &inputfile) {
        esc_params p;
        check(esc_params_load(inputfile.c_str(), &p));
        Lx = p.Lx; Ly = p.Ly; NumberofSites = p.NumberofSites; RandomSeed = p.RandomSeed;
        LanczosSteps = p.LanczosSteps; Threads = p.Threads;
        IsPeriodicX = p.IsPeriodicX; IsPeriodicY = p.IsPeriodicY;
        LocalSkw = p.LocalSkw; SpinCurrDynamics = p.SpinCurrDynamics; EnergyCurrDynamics = p.EnergyCurrDynamics;
        Kxx = p.Kxx; Kyy = p.Kyy; Kzz = p.Kzz; Bxx = p.Bxx; Byy = p.Byy; Bzz = p.Bzz;
        Temperature = p.Temperature; LanczosEps = p.LanczosEps;
        Model = p.Model; Solver = p.Solver; LancType = p.LancType;
        J = p.J; Gamma = p.Gamma; Eta = p.Eta; OmegaMin = p.OmegaMin; OmegaMax = p.OmegaMax;
        OmegaPoints = p.OmegaPoints; Precision = p.Precision; SpectralMode = p.SpectralMode;
        StartVector = p.StartVector; CFSteps = p.CFSteps; LocalSzSz = p.LocalSzSz;
        SzSz = p.SzSz; LocalRIXS = p.LocalRIXS; SpecialSite = p.SpecialSite;
        generator_.seed((uint64_t)RandomSeed);
        if (LancType == "OnlyEgs") std::cout << "WARNING:: Reortho will be turned off for LancType=OnlyEgs (set by default) " << std::endl;
    }
    std::mt19937_64 generator_;
    int Lx, Ly, NumberofSites, SpecialSite, LanczosSteps, RandomSeed, Threads;
    double LanczosEps, Kxx, Kyy, Kzz, Bxx, Byy, Bzz, Temperature;
    bool IsPeriodicX, IsPeriodicY;
    std::string Model, Solver, LancType;
    bool LocalSkw, RIXSDynamics = false, SpinCurrDynamics, EnergyCurrDynamics;
    // extensions
    double J, Gamma, Eta, OmegaMin, OmegaMax;
    int OmegaPoints, Precision, SpectralMode, StartVector, CFSteps;
    bool LocalSzSz, SzSz, LocalRIXS;
    int lanczos_runs = 0;   // Gaussian start vectors: seed RandomSeed + run

    // uniform [0,1) draw: double(g()) * 2^-64 (src/ParametersEngine.h:152-154)
    double grnd() { return (double)generator_() * (1.0 / 18446744073709551616.0); }

    std::vector<double> omega_grid() const {
        std::vector<double> w;
        if (OmegaPoints <= 0) {
            for (int om = 0; om <= 1000; ++om) w.push_back(0.002 * om);   // src/dynamics.h:231-232,270-271
        } else {
            const double d = OmegaPoints > 1 ? (OmegaMax - OmegaMin) / (OmegaPoints - 1) : 0.0;
            for (int k = 0; k < OmegaPoints; ++k) w.push_back(OmegaMin + d * k);
        }
        return w;
    }
};

// --------------------------------------------------------------------------
// Lattice (src/Lattice.h): N1neigh_(i, c), c = 0 (z), 1 (y), 2 (x)
// --------------------------------------------------------------------------
class Lattice {
public:
    struct NeighTable {
        std::vector<int32_t> v;
        int cols = 3;
        int operator()(int i, int c) const { return v[(size_t)i * cols + c]; }
    };
    // src/Lattice.h:37-58: Kitaev -> Honeycomb1, HeisenbergChain -> Chain
    explicit Lattice(ConstVariables &p) : variables_(p) {
        if (p.Model == "Kitaev") {
            N1neigh_.v.resize((size_t)p.NumberofSites * 3);
            check(esc_honeycomb(p.Lx, p.Ly, p.IsPeriodicX, p.IsPeriodicY, N1neigh_.v.data()));
            std::cout << "creating Honeycomb lattice: right shifted " << std::endl;
        } else if (p.Model == "HeisenbergChain") {
            N1neigh_.cols = 2;
            N1neigh_.v.resize((size_t)p.NumberofSites * 2);
            check(esc_chain(p.NumberofSites, p.IsPeriodicX, N1neigh_.v.data()));
            std::cout << "creating Chain lattice " << std::endl;
        } else {
            std::cerr << "Model=" << p.Model << "\n";
            throw std::string("Unknown Model parameter \n");
        }
    }
    bool chain() const { return N1neigh_.cols == 2; }
    NeighTable N1neigh_;

private:
    ConstVariables &variables_;
};

// --------------------------------------------------------------------------
// QBasis (src/QuantumBasis.h:18-163): identity basis on 2^N states.
// --------------------------------------------------------------------------
class QBasis {
public:
    struct IdentityBasis {
        int64_t n;
        int64_t size() const { return n; }
        int64_t operator[](int64_t k) const { return k; }
    };
    explicit QBasis(ConstVariables &p) : basis{(int64_t)1 << p.NumberofSites}, nHil_(basis.n), Nsite_(p.NumberofSites) {
        std::cout << "number of ket basis: " << basis.size() << std::endl;
    }
    IdentityBasis basis;
    int64_t nHil_;
    static int findBit(int64_t i, int position) { return (int)((i >> position) & 1); }
    static int64_t flipBit(int64_t ketin, int position) { return ketin ^ ((int64_t)1 << position); }
    void ApplySz(int site, int64_t ketin, int64_t &ketout, dcomplex &coef) const {
        coef = findBit(ketin, site) == 0 ? dcomplex(0.5, 0.0) : dcomplex(-0.5, 0.0);
        ketout = ketin;
    }
    void ApplySx(int site, int64_t ketin, int64_t &ketout, dcomplex &coef) const {
        coef = dcomplex(0.5, 0.0);
        ketout = flipBit(ketin, site);
    }
    void ApplySy(int site, int64_t ketin, int64_t &ketout, dcomplex &coef) const {
        coef = findBit(ketin, site) == 0 ? dcomplex(0.0, 0.5) : dcomplex(0.0, -0.5);
        ketout = flipBit(ketin, site);
    }

private:
    int Nsite_;
};

// --------------------------------------------------------------------------
// Hamiltonian (src/Hamiltonian.h): bond families in the reference order and a
// device context.  TightB_Ham / Diagonal build nothing (matrix elements are
// formed on the fly) and are kept for call compatibility.
// --------------------------------------------------------------------------
class Hamiltonian {
public:
    Hamiltonian(ConstVariables &p, Lattice &lat, QBasis &basis) : variables_(p), Lat_(lat), Basis_(basis) {
        const int N = p.NumberofSites;
        const double K[3] = {p.Kxx, p.Kyy, p.Kzz};
        const double ones[3] = {1.0, 1.0, 1.0};
        for (auto &f : pairs_) f.assign((size_t)2 * N, 0);
        int32_t cnt[3];
        if (lat.chain()) {   // src/Hamiltonian.h:44-53: HeisenbergChain_Connectors
            if (p.J != 0.0 || p.Gamma != 0.0)
                throw std::invalid_argument("J and Gamma are honeycomb extensions; use Kxx, Kyy, Kzz on a chain");
            check(esc_chain_connectors(N, lat.N1neigh_.v.data(), K, pairs_[0].data(), pairs_[1].data(),
                                       pairs_[2].data(), cnt));
        } else {
            // families follow the lattice; a zero coupling adds exact zeros
            check(esc_connectors(N, lat.N1neigh_.v.data(), (p.J != 0.0 || p.Gamma != 0.0) ? ones : K,
                                 pairs_[0].data(), pairs_[1].data(), pairs_[2].data(), cnt));
        }
        for (int f = 0; f < 3; ++f) {
            // a family with exactly one bond is skipped (src/Hamiltonian.h:116,143,347)
            if (cnt[f] == 1) cnt[f] = 0;
            pairs_[f].resize((size_t)2 * cnt[f]);
        }
        make_context();
    }
    void TightB_Ham() { std::cout << "create Tight-Binding/Off-Diag Hamiltonian (matrix-free)" << std::endl; }
    void Diagonal() { std::cout << "create Diagonal matrix elements (on the fly)" << std::endl; }
    // the bond families are built once in the constructor (src/Hamiltonian.h:44-54)
    void SetupConnectors() {}

    // Intermediate Hamiltonian of LocalDynRIXS (src/Hamiltonian.h:57-91): every
    // bond touching `center` loses its coupling.  Like the reference this
    // changes this Hamiltonian for all later calls; vectors and Lanczos runs
    // made before keep the previous context alive and stay valid.
    void RIXS_killcenter(int &center) {
        std::cout << " ==== RIXS - Intermediate Hamiltonian ==== " << std::endl;
        const double K[3] = {variables_.Kxx, variables_.Kyy, variables_.Kzz};
        for (int f = 0; f < 3; ++f) {
            std::vector<int32_t> keep;
            for (size_t b = 0; b + 1 < pairs_[f].size(); b += 2) {
                const int i = pairs_[f][b], j = pairs_[f][b + 1];
                const bool killed = i == center || j == center;
                std::cout << i << " \t " << j << " \t " << (killed ? 0.0 : K[f]) << std::endl;
                if (!killed) { keep.push_back(i); keep.push_back(j); }
            }
            pairs_[f] = keep;
        }
        make_context();
    }

    // H|vin> (src/Hamiltonian.h:480-557)
    VectorXcd MatrixVectorMult(VectorXcd &Vin) {
        if (Vin.size() != ctx_->dim()) throw std::invalid_argument("LanczosSolver.h::Dot-Product size mismatch!");
        VectorXcd out(ctx_);
        check(esc_hpsi(ctx_->get(), Vin.data(), out.data(), nullptr));
        return out;
    }
    const std::shared_ptr<Context> &context() const { return ctx_; }
    const std::vector<int32_t> &pairs(int family) const { return pairs_[family]; }

    // a second device context of the same Hamiltonian (its own vectors)
    std::shared_ptr<Context> fresh_context() const { return build_context(); }

private:
    void make_context() { ctx_ = build_context(); }
    std::shared_ptr<Context> build_context() const {
        const ConstVariables &p = variables_;
        esc_model m{};
        m.n_sites = p.NumberofSites;
        for (int f = 0; f < 3; ++f) {
            m.n_bonds[f] = (int32_t)(pairs_[f].size() / 2);
            m.bonds[f] = pairs_[f].data();
        }
        m.neigh = Lat_.chain() ? nullptr : Lat_.N1neigh_.v.data();   // current operators are honeycomb-only
        m.K[0] = p.Kxx; m.K[1] = p.Kyy; m.K[2] = p.Kzz;
        m.B[0] = p.Bxx; m.B[1] = p.Byy; m.B[2] = p.Bzz;
        m.J = p.J;
        m.Gamma = p.Gamma;
        return std::make_shared<Context>(m, p.Precision);
    }
    ConstVariables &variables_;
    Lattice &Lat_;
    QBasis &Basis_;
    std::vector<int32_t> pairs_[3];
    std::shared_ptr<Context> ctx_;
};

// --------------------------------------------------------------------------
// Eigenvectors of Solver=ED (the reference's Eigen::MatrixXcd Evec): column k
// is eigenstate k, resident on the device as the context's kept basis.
// --------------------------------------------------------------------------
using VectorXd = std::vector<double>;

class MatrixXcd {
public:
    MatrixXcd() = default;
    MatrixXcd(std::shared_ptr<Context> ctx, int run) : ctx_(std::move(ctx)), run_(run) {}
    int64_t rows() const { return ctx_ ? ctx_->dim() : 0; }
    int64_t cols() const { return rows(); }
    VectorXcd col(int k) const {
        if (!ctx_ || ctx_->lanczos_runs != run_)
            throw std::runtime_error("Evec: the eigenvectors were replaced by a later solve on this Hamiltonian");
        VectorXcd v(ctx_);
        check(esc_ritz_vector(ctx_->get(), k, v.data(), nullptr));
        return v;
    }
    const std::shared_ptr<Context> &context() const { return ctx_; }
    int run() const { return run_; }

private:
    std::shared_ptr<Context> ctx_;
    int run_ = 0;
};

// --------------------------------------------------------------------------
// DynLanczosSolver (src/DynLanczosSolver.h:24-239)
// --------------------------------------------------------------------------
class DynLanczosSolver {
public:
    // own_basis: run on a fresh context of the same Hamiltonian, so that the
    // kept basis of another solver on it survives (the reference keeps one
    // PsiAll per solver object)
    DynLanczosSolver(ConstVariables &p, QBasis &basis, Hamiltonian &H, bool own_basis = false)
        : variables_(p), Basis_(basis), Hamil_(H), own_basis_(own_basis) {}

    // Ritz states of the last "All" run, formed on demand (PsiAll[k])
    struct RitzStates {
        DynLanczosSolver *s = nullptr;
        int n = 0;
        int size() const { return n; }
        VectorXcd operator[](int k) const {
            s->check_basis();
            VectorXcd v(s->ctx_);
            check(esc_ritz_vector(s->ctx_->get(), k, v.data(), nullptr));
            return v;
        }
    };
    RitzStates PsiAll;
    std::vector<double> TriDeval;
    VectorXcd Vorig;
    std::vector<double> alpha, beta;
    bool basis_kept = false;

    VectorXcd Lanczos_Nirav(const std::string &PsiOrAll) {
        std::cout << "Begin Lanczos! --- " << std::endl;
        int mode;
        if (PsiOrAll == "OnlyEgs") mode = ESC_LANC_ONLYEGS;
        else if (PsiOrAll == "GS") mode = ESC_LANC_GS;
        else if (PsiOrAll == "All") mode = ESC_LANC_ALL;
        else throw std::runtime_error("DynLanczosSolver: Use PsiOrAll==GS or All");
        auto ctx = own_basis_ ? Hamil_.fresh_context() : Hamil_.context();
        ctx_ = ctx;
        run_ = ++ctx->lanczos_runs;
        Vorig = VectorXcd(ctx);
        Randomize(Vorig);
        const int M = variables_.LanczosSteps;
        alpha.assign(M + 1, 0.0);
        beta.assign(M + 2, 0.0);
        TriDeval.assign(M + 1, 0.0);
        int iters = 0;
        VectorXcd gs(ctx);
        check(esc_lanczos(ctx->get(), Vorig.data(), M, variables_.LanczosEps, mode, alpha.data(), beta.data(), &iters,
                          TriDeval.data(), nullptr, mode == ESC_LANC_ONLYEGS ? nullptr : gs.data(), nullptr));
        alpha.resize(iters);
        beta.resize(iters + 1);
        TriDeval.resize(iters);
        std::cout << "  Lowest Eigen Value: " << std::setprecision(16) << TriDeval[0] << std::endl;
        std::cout << "  Alpha and beta of Tridiag Matrix : ";
        for (int m = 0; m < iters; ++m) std::cout << "(" << alpha[m] << "," << beta[m] << ") ";
        std::cout << std::endl << std::endl;
        basis_kept = mode == ESC_LANC_ALL;
        PsiAll = RitzStates{this, basis_kept ? iters : (mode == ESC_LANC_GS ? 1 : 0)};
        if (mode == ESC_LANC_ONLYEGS) gs.setZero();
        return gs;
    }

    // Solver=ED results as the solver's Ritz data: eigenvalues as TriDeval,
    // eigenvectors as PsiAll (src/DynLanczosSolver.h:269-283)
    void fillFromED(MatrixXcd Evec, VectorXd eval) {
        if (!Evec.context() || (int64_t)eval.size() != Evec.rows())
            throw std::invalid_argument("fillFromED: eigenvalues and eigenvectors do not match");
        ctx_ = Evec.context();
        run_ = Evec.run();
        TriDeval = eval;
        alpha.clear();
        beta.clear();
        basis_kept = true;
        PsiAll = RitzStates{this, (int)eval.size()};
    }

    // Start vector (src/DynLanczosSolver.h:58-59,465-474): re then im drawn
    // from the parameters' generator, which every run continues, then
    // normalised; the Gaussian extension uses seed RandomSeed + run.
    void Randomize(VectorXcd &V) {
        auto ctx = V.context();
        if (variables_.StartVector == ESC_START_REF_UNIFORM) {
            if (variables_.NumberofSites > 28)
                throw std::invalid_argument("the reference start vector is limited to N <= 28; set StartVector=Gaussian");
            std::vector<double> h(2 * (size_t)ctx->dim());
            for (double &x : h) x = -1.0 + variables_.grnd() * 2.0;
            double n2 = 0.0;
            for (double x : h) n2 += x * x;
            const double nrm = std::sqrt(n2);
            for (double &x : h) x /= nrm;
            check(esc_upload(ctx->get(), h.data(), V.data(), nullptr));
        } else {
            check(esc_start_vector(ctx->get(), (uint64_t)(variables_.RandomSeed + variables_.lanczos_runs),
                                   variables_.StartVector, V.data(), nullptr));
        }
        ++variables_.lanczos_runs;
    }

    // the context and the kept basis of this solver's last run
    const std::shared_ptr<Context> &context() const { return ctx_; }
    void check_basis() const {
        if (!basis_kept || !ctx_ || ctx_->lanczos_runs != run_)
            throw std::runtime_error("PsiAll: the Lanczos basis of this run is not kept (LancType=All is needed, "
                                     "and a later run on the same Hamiltonian replaces it)");
    }

private:
    std::shared_ptr<Context> ctx_;
    int run_ = 0;
    bool own_basis_ = false;
    ConstVariables &variables_;
    QBasis &Basis_;
    Hamiltonian &Hamil_;
};

// --------------------------------------------------------------------------
// Observables (src/Observables.h)
// --------------------------------------------------------------------------
class Observables {
public:
    Observables(ConstVariables &p, Lattice &lat, QBasis &basis, Hamiltonian &H)
        : variables_(p), Lat_(lat), Basis_(basis), Hamil_(H) {}

    VectorXcd ApplyOp(VectorXcd &inV, const std::string &OpA, const int &site) {
        static const char *names[] = {"Sx", "Sy", "Sz", "Jsx", "Jsy", "Jsz", "Jex", "Jey"};
        if (OpA == "I") return inV;
        for (int op = 0; op < 8; ++op)
            if (OpA == names[op]) return apply(op, site, inV);
        std::cerr << "Operator=" << OpA << " does not exist \n";
        throw std::string("Unknown Operator label \n");
    }
    VectorXcd ApplySx(VectorXcd &v, const int &i) { return apply(ESC_OP_SX, i, v); }
    VectorXcd ApplySy(VectorXcd &v, const int &i) { return apply(ESC_OP_SY, i, v); }
    VectorXcd ApplySz(VectorXcd &v, const int &i) { return apply(ESC_OP_SZ, i, v); }
    VectorXcd ApplyJsx(VectorXcd &v, const int &i) { return apply(ESC_OP_JSX, i, v); }
    VectorXcd ApplyJsy(VectorXcd &v, const int &i) { return apply(ESC_OP_JSY, i, v); }
    VectorXcd ApplyJsz(VectorXcd &v, const int &i) { return apply(ESC_OP_JSZ, i, v); }
    VectorXcd ApplyJex(VectorXcd &v, const int &i) { return apply(ESC_OP_JEX, i, v); }
    VectorXcd ApplyJey(VectorXcd &v, const int &i) { return apply(ESC_OP_JEY, i, v); }
    VectorXcd ApplyJsxTotal(VectorXcd &v) { return apply(ESC_OP_JSX, ESC_SITE_TOTAL, v); }
    VectorXcd ApplyJsyTotal(VectorXcd &v) { return apply(ESC_OP_JSY, ESC_SITE_TOTAL, v); }
    VectorXcd ApplyJszTotal(VectorXcd &v) { return apply(ESC_OP_JSZ, ESC_SITE_TOTAL, v); }
    VectorXcd ApplyJexTotal(VectorXcd &v) { return apply(ESC_OP_JEX, ESC_SITE_TOTAL, v); }
    VectorXcd ApplyJeyTotal(VectorXcd &v) { return apply(ESC_OP_JEY, ESC_SITE_TOTAL, v); }

    // src/Observables.h:107-128
    void measureLocalS(VectorXcd &GS) {
        std::cout << "\n-------- measuring: site, <sx>, <sy>, <sz>, <spin_total>, <sp>, <sm> --------" << std::endl;
        const dcomplex I(0.0, 1.0);
        for (int si = 0; si < variables_.NumberofSites; ++si) {
            VectorXcd tx = ApplySx(GS, si), ty = ApplySy(GS, si), tz = ApplySz(GS, si);
            const dcomplex sx = GS.dot(tx), sy = GS.dot(ty), sz = GS.dot(tz);
            std::cout << si << " \t " << sx << " \t " << sy << " \t " << sz << " \t " << sx + sy + sz << " \t "
                      << sx + I * sy << " \t " << sx - I * sy << std::endl;
        }
    }
    // src/Observables.h:88-104 (A sites carry Jex, B sites Jey)
    void measureLocalJe(VectorXcd &GS) {
        std::cout << "\n-------- measuring: site, <Je_xi>, <Je_yi>, <Je_zi>, <Je_total> --------" << std::endl;
        for (int si = 0; si < variables_.NumberofSites; ++si) {
            VectorXcd tx = ApplyJex(GS, si), ty = ApplyJey(GS, si);
            const dcomplex ox = GS.dot(tx), oy = GS.dot(ty);
            std::cout << si << " \t " << ox << " \t " << oy << " \t " << " \t " << ox + oy << std::endl;
        }
    }
    // src/Observables.h:70-85
    void measureLocalJs(VectorXcd &GS) {
        std::cout << "\n-------- measuring: site, <Js_xi>, <Js_yi>, <Js_zi>, <Js_total> --------" << std::endl;
        for (int si = 0; si < variables_.NumberofSites; ++si) {
            VectorXcd tx = ApplyJsx(GS, si), ty = ApplyJsy(GS, si), tz = ApplyJsz(GS, si);
            const dcomplex ox = GS.dot(tx), oy = GS.dot(ty), oz = GS.dot(tz);
            std::cout << si << " \t " << ox << " \t " << oy << " \t " << oz << " \t " << ox + oy + oz << std::endl;
        }
    }

    // src/Observables.h:132-155: Corr(si, sj) = <GS|A_si B_sj|GS> for si >= sj,
    // mirrored; returned row-major as N x N
    std::vector<dcomplex> TwoPointCorr(const std::string OpA, const std::string OpB, VectorXcd &GS_) {
        std::cout << "Calculating " << OpA << "." << OpB << " correlations " << std::endl;
        static const char *names[] = {"Sx", "Sy", "Sz", "Jsx", "Jsy", "Jsz", "Jex", "Jey"};
        int a = -1, b = -1;
        for (int op = 0; op < 8; ++op) {
            if (OpA == names[op]) a = op;
            if (OpB == names[op]) b = op;
        }
        if (a < 0 || b < 0) throw std::string("Unknown Operator label \n");
        const int N = variables_.NumberofSites;
        std::vector<double> raw(2 * (size_t)N * N);
        check(esc_two_point(Hamil_.context()->get(), a, b, GS_.data(), raw.data(), nullptr));
        std::vector<dcomplex> corr((size_t)N * N);
        for (size_t k = 0; k < corr.size(); ++k) corr[k] = dcomplex(raw[2 * k], raw[2 * k + 1]);
        return corr;
    }

private:
    VectorXcd apply(int op, int site, VectorXcd &in) {
        VectorXcd out(Hamil_.context());
        check(esc_apply_op(Hamil_.context()->get(), op, site, in.data(), out.data(), nullptr));
        return out;
    }
    ConstVariables &variables_;
    Lattice &Lat_;
    QBasis &Basis_;
    Hamiltonian &Hamil_;
};

// --------------------------------------------------------------------------
// StartDynamics (src/dynamics.h:43-60): spectrum table -> *_outSpectrum.real
// --------------------------------------------------------------------------
inline void write_spectrum_file(const std::string &path, const std::vector<double> &table, int ncols, int nw,
                                bool tab_after_last = false) {
    // byte format of src/dynamics.h:325-328,432-435: default ostream precision,
    // " \t " separators, every row terminated by " \n "; LocalDynSzSz also
    // writes " \t " after the last value (:187-192)
    std::ofstream f(path);
    for (int q = 0; q < nw; ++q) {
        const double *row = &table[(size_t)q * (1 + ncols)];
        f << row[0] << " \t ";
        for (int c = 1; c < ncols; ++c) f << row[c] << " \t ";
        f << row[ncols];
        if (tab_after_last) f << " \t ";
        f << " \n ";
    }
}

// the grid of DynSzSz / LocalDynRIXS: w = 0.1 k, k = 0..120 (src/dynamics.h:532-533,587-588),
// or the OmegaMin/OmegaMax/OmegaPoints extension
inline std::vector<double> tenth_grid(const ConstVariables &v) {
    if (v.OmegaPoints > 0) return v.omega_grid();
    std::vector<double> w;
    for (int om = 0; om <= 120; ++om) w.push_back(0.1 * om);
    return w;
}

// DynSzSz (src/dynamics.h:566-611): a second Lanczos run ("All") on the same
// Hamiltonian, then S^z_i S^z_j spectra over its Ritz states, written as
// "si \t sj \t w \t real \t imag" rows, w outermost, to SzSz_outSpectrum.real.
inline void DynSzSz(std::string &dynstr, ConstVariables &Parameters, Lattice &, QBasis &Basis, Hamiltonian &Hamil,
                    DynLanczosSolver &, VectorXcd &Psi) {
    std::cout << " ================================================ " << std::endl;
    std::cout << " ------- Startin Dynamics " << dynstr << " ------ " << std::endl;
    std::cout << " ================================================ " << std::endl;
    const int N = Parameters.NumberofSites;
    DynLanczosSolver DLanc(Parameters, Basis, Hamil, true);
    DLanc.Lanczos_Nirav("All");
    std::cout << DLanc.PsiAll.size() << " \t " << DLanc.PsiAll.size() << std::endl;
    const std::vector<double> w = tenth_grid(Parameters);
    const double eta = 0.1;
    std::vector<double> table(w.size() * N * N * 2);
    check(esc_szsz(DLanc.context()->get(), Psi.data(), eta, w.data(), (int)w.size(), table.data(), nullptr));
    std::ofstream f("SzSz_outSpectrum.real");
    for (size_t q = 0; q < w.size(); ++q)
        for (int si = 0; si < N; ++si)
            for (int sj = 0; sj < N; ++sj) {
                const double *v = &table[((q * N + si) * N + sj) * 2];
                f << si << " \t " << sj << " \t " << w[q] << " \t " << v[0] << " \t " << v[1] << std::endl;
            }
}

// LocalDynRIXS (src/dynamics.h:465-562): Eg from the ground-state run, the
// intermediate Hamiltonian without the bonds of SpecialSite, a Lanczos run
// ("All") on it, then the four channels (1, S^x, S^y, S^z at the site) with
// Gamma = 100 and eta = 0.1, written as "w \t Re.. \t Im.." rows to
// RIXS_outSpectrum.real.  SpectralMode=CF computes the same sums from
// Krylov spaces started at |gs> and at X chi (no second "All" run).
inline void LocalDynRIXS(std::string &dynstr, ConstVariables &Parameters, Lattice &, QBasis &Basis,
                         Hamiltonian &Hamiltonian, DynLanczosSolver &DLancGS, VectorXcd &Psi) {
    std::cout << " ================================================ " << std::endl;
    std::cout << " ------- Startin RIXS Local Dynamics " << dynstr << " ------ " << std::endl;
    std::cout << " ================================================ " << std::endl;
    int centeral_site = Parameters.SpecialSite;
    const bool ritz = Parameters.SpectralMode == ESC_SPEC_RITZ;
    if (ritz) DLancGS.check_basis();
    const double Eg = DLancGS.TriDeval.empty() ? 0.0 : DLancGS.TriDeval[0];
    auto gs_ctx = ritz ? DLancGS.context() : Psi.context();
    Hamiltonian.SetupConnectors();
    Hamiltonian.RIXS_killcenter(centeral_site);
    Hamiltonian.TightB_Ham();
    Hamiltonian.Diagonal();
    DynLanczosSolver DLancInter(Parameters, Basis, Hamiltonian);
    if (ritz) {
        DLancInter.Lanczos_Nirav("All");
        std::cout << DLancInter.PsiAll.size() << " \t " << DLancGS.PsiAll.size() << std::endl;
    }
    const std::vector<double> w = tenth_grid(Parameters);
    const double eta = 0.1, Gamma = 100;
    std::vector<double> table(w.size() * 9);
    check(esc_rixs(gs_ctx->get(), Hamiltonian.context()->get(), ritz ? ESC_SPEC_RITZ : ESC_SPEC_CF, centeral_site,
                   Psi.data(), Eg, Parameters.CFSteps, Gamma, eta, w.data(), (int)w.size(), table.data(), nullptr));
    std::ofstream f("RIXS_outSpectrum.real");
    for (size_t q = 0; q < w.size(); ++q) {
        const double *r = &table[q * 9];
        for (int c = 0; c < 9; ++c) {
            std::cout << r[c] << (c < 8 ? " \t " : "");
            f << r[c] << (c < 8 ? " \t " : "");
        }
        std::cout << std::endl;
        f << std::endl;
    }
}

// Solver=ED (src/Exactdiag.h:36-81): every eigenpair of the dense H on the
// device; D receives the eigenvalues (ascending), the returned matrix the
// eigenvectors
inline MatrixXcd ExactDiag(ConstVariables &variables_, QBasis &Basis_, Hamiltonian &Hamil_, Observables &Observ_,
                           VectorXd &D) {
    (void)variables_; (void)Basis_; (void)Observ_;
    std::cout << "Begin Exact Diagonalization! --- " << std::endl;
    auto ctx = Hamil_.context();
    D.assign((size_t)ctx->dim(), 0.0);
    check(esc_exact_diag(ctx->get(), D.data(), nullptr, nullptr));
    const int run = ++ctx->lanczos_runs;
    for (double d : D) std::cout << d << ", ";
    std::cout << " Ground-state energy:= " << std::setprecision(16) << D[0] << std::endl;
    return MatrixXcd(ctx, run);
}

inline void StartDynamics(std::string &dynstr, ConstVariables &variables, Lattice &Lat, QBasis &Basis,
                          Hamiltonian &Hamil, DynLanczosSolver &DLancGS, VectorXcd &Psi) {
    if (dynstr == "SzSz") return DynSzSz(dynstr, variables, Lat, Basis, Hamil, DLancGS, Psi);
    if (dynstr == "LocalRIXS") return LocalDynRIXS(dynstr, variables, Lat, Basis, Hamil, DLancGS, Psi);
    int kind;
    std::string file;
    if (dynstr == "Spin Current") { kind = ESC_DYN_JSJS; file = "JsJs_outSpectrum.real"; }
    else if (dynstr == "Energy Current") { kind = ESC_DYN_JEJE; file = "JeJe_outSpectrum.real"; }
    else if (dynstr == "LocalSzSz") { kind = ESC_DYN_LOCALSZSZ; file = "2SSLocal_outSpectrum.real"; }
    else {
        std::cerr << "dynstr=" << dynstr << "\n";
        throw std::string("Unknown dynstr parameter \n");
    }
    std::cout << " --> Startin Dynamics: " << dynstr << std::endl;
    const std::vector<double> w = variables.omega_grid();
    const int ncols = kind == ESC_DYN_JSJS ? 14 : (kind == ESC_DYN_JEJE ? 5 : 8);
    std::vector<double> table(w.size() * (1 + ncols));
    const int method = variables.SpectralMode;
    if (method == ESC_SPEC_RITZ && !DLancGS.basis_kept)
        throw std::invalid_argument("You need to set LancType=All for Dynamical calculations --- you dummy!");
    if (method == ESC_SPEC_RITZ) DLancGS.check_basis();
    const double e0 = DLancGS.TriDeval.empty() ? 0.0 : DLancGS.TriDeval[0];
    auto ctx = method == ESC_SPEC_RITZ ? DLancGS.context() : Hamil.context();
    check(esc_dynamics(ctx->get(), kind, method, Psi.data(), e0, variables.CFSteps, variables.Eta, w.data(),
                       (int)w.size(), table.data(), nullptr, nullptr));
    write_spectrum_file(file, table, ncols, (int)w.size(), kind == ESC_DYN_LOCALSZSZ);
    std::cout << std::setprecision(8);
    for (size_t q = 0; q < w.size(); q += 200) {
        std::cout << table[q * (1 + ncols)];
        for (int c = 1; c <= ncols; ++c) std::cout << " \t " << table[q * (1 + ncols) + c];
        std::cout << " \n ";
    }
    std::cout << " " << std::endl;
}

}  // namespace esc
