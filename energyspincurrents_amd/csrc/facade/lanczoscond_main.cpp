// LanczosCond -- drop-in command-line driver (usage: LanczosCond <inputfile>).
// Same call order as the reference driver (main.cpp:41-112): parameters,
// lattice, basis, Hamiltonian, Lanczos (or ED), local measurements, then the
// spin-current (LocalSkw=1) and energy-current (EnergyCurrDynamics=1) spectra,
// written to JsJs_outSpectrum.real / JeJe_outSpectrum.real in the working
// directory; with the extension keys LocalSzSz=1, SzSz=1 and LocalRIXS=1 the
// spectra of LocalDynSzSz (2SSLocal_outSpectrum.real), DynSzSz
// (SzSz_outSpectrum.real) and LocalDynRIXS at SpecialSite
// (RIXS_outSpectrum.real).  Everything below the class API runs on the GPU
// through libesc.so.
#include "esc_reference_api.hpp"

#include <iostream>
#include <stdexcept>
#include <string>

int main(int argc, char *argv[]) {
    using namespace esc;
    std::cout << "LanczosCond (MI355X engine)" << std::endl;
    if (argc != 2) throw std::invalid_argument("USE:: executable inputfile");
    std::string inputfile = argv[1];
    std::cout << argv[1] << std::endl;
    try {
        ConstVariables Parameters(inputfile);
        Lattice Lat(Parameters);
        QBasis Basis(Parameters);
        Hamiltonian Ham(Parameters, Lat, Basis);
        Ham.TightB_Ham();
        Ham.Diagonal();
        Observables Observ(Parameters, Lat, Basis, Ham);
        DynLanczosSolver Lanczos(Parameters, Basis, Ham);
        VectorXcd Psi;
        const bool ed = Parameters.Solver == "ED";
        if (ed) {
            VectorXd eval;
            MatrixXcd Evec = ExactDiag(Parameters, Basis, Ham, Observ, eval);
            Psi = Evec.col(0);
            Lanczos.fillFromED(Evec, eval);
        } else {
            Psi = Lanczos.Lanczos_Nirav(Parameters.LancType);
            std::cout << std::setprecision(6) << std::fixed;
            std::cout << " Ritz values: ";
            for (double e : Lanczos.TriDeval) std::cout << e << " ";
            std::cout << std::endl;
        }
        std::cout << std::setprecision(8) << std::fixed;
        if (ed || Parameters.LancType != "OnlyEgs") {
            Observ.measureLocalS(Psi);
            if (!Lat.chain()) Observ.measureLocalJe(Psi);   // currents live on the honeycomb only
        }
        if (Lat.chain() && (Parameters.LocalSkw || Parameters.EnergyCurrDynamics))
            throw std::invalid_argument("current spectra are defined on the honeycomb only (Model=Kitaev)");
        std::cout.unsetf(std::ios_base::floatfield);
        if (Parameters.LocalSkw) {
            std::cout << " ==================================================================== " << std::endl;
            std::cout << "                   Spin Dynamical Correlations                         " << std::endl;
            std::cout << " ==================================================================== " << std::endl;
            std::string str = "Spin Current";
            StartDynamics(str, Parameters, Lat, Basis, Ham, Lanczos, Psi);
        }
        if (Parameters.EnergyCurrDynamics) {
            std::cout << " ==================================================================== " << std::endl;
            std::cout << "              Energy Current Dynamical Correlations                    " << std::endl;
            std::cout << " ==================================================================== " << std::endl;
            std::string str1 = "Energy Current";
            StartDynamics(str1, Parameters, Lat, Basis, Ham, Lanczos, Psi);
        }
        if (Parameters.LocalSzSz) {   // extension key: src/dynamics.h:64-217, unreachable from main.cpp
            std::cout << " ==================================================================== " << std::endl;
            std::cout << "                  Local SzSz Dynamical Correlations                    " << std::endl;
            std::cout << " ==================================================================== " << std::endl;
            std::string str2 = "LocalSzSz";
            StartDynamics(str2, Parameters, Lat, Basis, Ham, Lanczos, Psi);
        }
        if (Parameters.SzSz) {        // extension key: src/dynamics.h:566-611
            std::string str3 = "SzSz";
            StartDynamics(str3, Parameters, Lat, Basis, Ham, Lanczos, Psi);
        }
        if (Parameters.LocalRIXS) {   // extension key: src/dynamics.h:465-562 (last: it changes the Hamiltonian)
            std::string str4 = "LocalRIXS";
            StartDynamics(str4, Parameters, Lat, Basis, Ham, Lanczos, Psi);
        }
    } catch (const std::exception &e) {
        std::cerr << "LanczosCond: " << e.what() << std::endl;
        return 1;
    } catch (const std::string &e) {
        std::cerr << "LanczosCond: " << e << std::endl;
        return 1;
    }
    return 0;
}
