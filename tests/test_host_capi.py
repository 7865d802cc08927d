"""CPU tests of the C ABI: the library loads and exports every entry point of
include/esc.h; the host-side functions (lattice, bond families, parameter file,
tridiagonal solver, spectra) match the oracle.  No GPU call is made."""
import os
import re

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, ROOT

esc = pytest.importorskip("energyspincurrents_amd")
from energyspincurrents_amd import _capi  # noqa: E402


def test_library_exports_every_declared_entry_point():
    hdr = open(os.path.join(ROOT, "include", "esc.h")).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(esc_\w+)\s*\(", hdr, re.M))
    assert declared == set(_capi.EXPORTS)
    lib = _capi.lib()
    for name in declared:
        assert hasattr(lib, name), name


@pytest.mark.parametrize("Lx", [1, 2, 3, 4])
@pytest.mark.parametrize("Ly", [2, 4, 6, 8])
@pytest.mark.parametrize("px,py", [(1, 1), (0, 1), (1, 0), (0, 0)])
def test_honeycomb_and_bonds_match_reference_restatement(Lx, Ly, px, py):
    lat = esc.Honeycomb(Lx, Ly, px, py)
    ref, _, _ = O.honeycomb1(Lx, Ly, px, py)
    assert np.array_equal(lat.neigh, ref)
    for K in [(1.0, 1.0, 1.0), (0.0, -1.0, 0.5)]:
        got = lat.bonds(K)
        want = O.connectors(ref, K)
        for f in range(3):
            assert np.array_equal(got[f], want[f])


def test_honeycomb_rejects_odd_ly():
    with pytest.raises(esc.EscError):
        esc.Honeycomb(3, 5)


def test_tridiag_eig_matches_lapack():
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 50, 300):
        a = rng.standard_normal(n)
        b = np.concatenate([[0.0], np.abs(rng.standard_normal(n - 1))])
        ev, Z = esc.tridiag_eig(a, b)
        T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
        assert np.abs(ev - np.linalg.eigvalsh(T)).max() < 1e-12 * max(1, np.abs(ev).max())
        assert np.abs(T @ Z - Z * ev).max() < 1e-11
        assert np.abs(Z.T @ Z - np.eye(n)).max() < 1e-12


def test_tridiag_eig_degenerate_and_zero_offdiagonal():
    a = np.array([1.0, 1.0, 2.0, 2.0])
    b = np.array([0.0, 0.0, 0.0, 0.0])
    ev, Z = esc.tridiag_eig(a, b)
    assert np.array_equal(ev, [1.0, 1.0, 2.0, 2.0])


def test_cf_and_lehmann_spectra_match_oracle():
    rng = np.random.default_rng(1)
    n = 40
    a = rng.standard_normal(n)
    b = np.concatenate([[0.0], np.abs(rng.standard_normal(n))])
    w = np.linspace(0, 6, 300)
    got = esc.cf_spectrum(a, b, 1.7, -2.0, 0.05, w)
    ref = O.cf_resolvent(a, b[:n], 1.7, w - 2.0 - 0.05j)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    ev = np.sort(rng.standard_normal(n))
    Oa = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Ob = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    el = np.conj(Oa[0]) * Ob[0]
    got = esc.lehmann_spectrum(ev, Oa, Ob, 0.005, w, elastic=el)
    ref = O.lehmann(ev, np.stack([Oa, Ob], 1), np.array([Oa[0], Ob[0]]), w, 0.005, [(0, 1)])[(0, 1)]
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_params_midinput():
    p = esc.load_params(os.path.join(GOLDEN, "Midinput.inp"))
    assert (p["Lx"], p["Ly"], p["NumberofSites"], p["RandomSeed"]) == (3, 4, 12, 3)
    assert (p["Model"], p["Solver"], p["LancType"]) == ("Kitaev", "none", "All")
    assert (p["Kxx"], p["Kyy"], p["Kzz"], p["Bxx"], p["Byy"], p["Bzz"]) == (1.0, 1.0, 1.0, 0.3, 0.3, 0.3)
    assert (p["LanczosSteps"], p["Threads"], p["LanczosEps"]) == (100, 8, 0.0)
    assert (p["LocalSkw"], p["SpinCurrDynamics"], p["EnergyCurrDynamics"]) == (0, 0, 1)
    assert (p["IsPeriodicX"], p["IsPeriodicY"]) == (1, 1)
    # extension defaults reproduce the reference behaviour
    assert (p["J"], p["Gamma"], p["Eta"], p["OmegaPoints"], p["Precision"], p["SpectralMode"]) == \
        (0.0, 0.0, 0.005, 0, 0, 0)


def _write(tmp_path, text):
    f = tmp_path / "in.inp"
    f.write_text(text)
    return str(f)


BASE = open(os.path.join(GOLDEN, "Midinput.inp")).read()


def test_params_missing_key_message(tmp_path):
    text = "\n".join(ln for ln in BASE.splitlines() if not ln.startswith("Kyy="))
    with pytest.raises(esc.EscError, match="Kyy= argument is missing in the input file!"):
        esc.load_params(_write(tmp_path, text))


def test_params_defaults_and_first_match(tmp_path):
    text = BASE.replace("LanczosSteps=100", "LanczosSteps=abc\nLanczosSteps=55\nLanczosSteps=77")
    p = esc.load_params(_write(tmp_path, text))
    assert p["LanczosSteps"] == 55               # first readable exact match wins
    text = "\n".join(ln for ln in BASE.splitlines() if not ln.startswith("LanczosSteps"))
    assert esc.load_params(_write(tmp_path, text))["LanczosSteps"] == 200
    text = BASE.replace("Kxx=1.0", " Kxx=9.0\nKxx =8.0\nKxx=  2.5")
    assert esc.load_params(_write(tmp_path, text))["Kxx"] == 2.5   # key text must match exactly


def test_params_dynamics_need_lanctype_all(tmp_path):
    text = BASE.replace("LancType=All", "LancType=GS")
    with pytest.raises(esc.EscError, match="LancType=All"):
        esc.load_params(_write(tmp_path, text))
    # the continued-fraction mode does not need the Ritz states
    p = esc.load_params(_write(tmp_path, text + "\nSpectralMode=CF\n"))
    assert p["SpectralMode"] == 1


def test_params_extensions(tmp_path):
    text = BASE + "\nJ=0.1\nGamma=0.2\nEta=0.05\nOmegaMax=6\nOmegaPoints=2000\nPrecision=c64\nStartVector=Gaussian\n"
    p = esc.load_params(_write(tmp_path, text))
    assert (p["J"], p["Gamma"], p["Eta"], p["OmegaMax"], p["OmegaPoints"]) == (0.1, 0.2, 0.05, 6.0, 2000)
    assert (p["Precision"], p["StartVector"]) == (1, 1)
    with pytest.raises(esc.EscError):
        esc.load_params(_write(tmp_path, BASE + "\nPrecision=fp16\n"))


def test_context_rejects_bad_models_before_touching_the_device():
    import ctypes
    m = _capi.EscModel()
    m.n_sites = 1
    h = ctypes.c_void_p()
    rc = _capi.lib().esc_ctx_create(ctypes.byref(m), 0, ctypes.byref(h))
    assert rc == -1 and b"n_sites" in _capi.lib().esc_last_error()
    m.n_sites = 4
    m.n_bonds[0] = 1
    bonds = (ctypes.c_int32 * 2)(0, 0)
    m.bonds[0] = ctypes.cast(bonds, ctypes.c_void_p)
    rc = _capi.lib().esc_ctx_create(ctypes.byref(m), 0, ctypes.byref(h))
    assert rc == -1 and b"invalid bond" in _capi.lib().esc_last_error()
    m.n_bonds[0] = 0
    m.K[0] = float("nan")
    rc = _capi.lib().esc_ctx_create(ctypes.byref(m), 0, ctypes.byref(h))
    assert rc == -5


@pytest.mark.parametrize("mode,code", [("Ritz", 0), ("CF", 1), ("CFReorth", 2)])
def test_params_spectral_modes(tmp_path, mode, code):
    p = esc.load_params(_write(tmp_path, BASE + f"\nSpectralMode={mode}\n"))
    assert p["SpectralMode"] == code


def test_params_spectral_mode_rejects_unknown(tmp_path):
    with pytest.raises(esc.EscError, match="SpectralMode"):
        esc.load_params(_write(tmp_path, BASE + "\nSpectralMode=Pade\n"))


def test_params_local_szsz_key(tmp_path):
    p = esc.load_params(_write(tmp_path, BASE + "\nLocalSzSz=1\n"))
    assert p["LocalSzSz"] == 1
    assert esc.load_params(_write(tmp_path, BASE))["LocalSzSz"] == 0


def test_params_szsz_rixs_and_special_site(tmp_path):
    p = esc.load_params(_write(tmp_path, BASE))
    assert p["SzSz"] == 0 and p["LocalRIXS"] == 0
    assert p["SpecialSite"] == p["NumberofSites"] // 2      # the reference's commented-out default
    p = esc.load_params(_write(tmp_path, BASE + "\nSzSz=1\nLocalRIXS=1\nSpecialSite=5\n"))
    assert p["SzSz"] == 1 and p["LocalRIXS"] == 1 and p["SpecialSite"] == 5
    with pytest.raises(esc.EscError, match="SpecialSite"):
        esc.load_params(_write(tmp_path, BASE + "\nSpecialSite=99\n"))
    # the RIXS sums run over the ground-state run's Ritz states: LancType=All
    gs_only = BASE.replace("LancType=All", "LancType=GS").replace("EnergyCurrDynamics=1", "EnergyCurrDynamics=0")
    esc.load_params(_write(tmp_path, gs_only))
    with pytest.raises(esc.EscError, match="LancType=All"):
        esc.load_params(_write(tmp_path, gs_only + "\nLocalRIXS=1\n"))
    esc.load_params(_write(tmp_path, gs_only + "\nSzSz=1\n"))      # DynSzSz runs its own "All" run
