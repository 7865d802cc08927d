"""Model=HeisenbergChain on the device (src/Lattice.h:60-125,
src/Hamiltonian.h:381-475): H.psi against the oracle (matrix elements bit for
bit against its COO restatement, random vectors within 1e-14 / 1e-6 for c128 /
c64 in every kernel configuration), ground energies against the closed forms
(XX ring as free fermions, Heisenberg rings), the local S^a S^a spectra against
the exact Lehmann sum, and the CLI with Model=HeisenbergChain."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

esc = pytest.importorskip("energyspincurrents_amd")

from test_oracle_chain import HEISENBERG_RING, xx_ring_e0  # noqa: E402


def engine_for(m: O.ChainModel, precision="c128"):
    return esc.Engine(esc.SpinModel(esc.Chain(m.N, m.periodic), K=m.K, B=m.B), precision)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("kw", [dict(N=10, K=(1.0, 1.0, 1.0)),
                                dict(N=11, K=(0.7, -1.3, 0.4), B=(0.2, -0.5, 0.3), periodic=False)])
def test_chain_matrix_elements_bit_exact(kw):
    m = O.ChainModel(**kw)
    print(f"[chain bit-exact] {kw}", flush=True)
    H = O.CooHamiltonian(m).dense()
    eng = engine_for(m)
    rng = np.random.default_rng(3)
    for k in sorted(set([0, 1, m.dim - 1] + list(rng.integers(0, m.dim, 40)))):
        e = np.zeros(m.dim, np.complex128)
        e[k] = 1.0
        assert np.array_equal(eng.download(eng.hpsi(eng.upload(e))), H[:, k]), f"column {k}"


@pytest.mark.parametrize("prec,tol", [("c128", 1e-14), ("c64", 1e-6)])
@pytest.mark.parametrize("N,periodic", [(16, True), (18, False)])
def test_chain_random_vectors_every_config(prec, tol, N, periodic):
    m = O.ChainModel(N, K=(1.0, 0.8, 1.2), B=(0.1, 0.0, -0.2), periodic=periodic)
    rng = np.random.default_rng(N)
    v = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    if prec == "c64":
        v = v.astype(np.complex64).astype(np.complex128)
    ref = O.hpsi_gather(m, v)
    eng = engine_for(m, prec)
    x = eng.upload(v)
    for tb in eng.kernel_configs():
        eng.set_tile_bits(tb)
        print(f"[chain config {tb}] {prec} N={N}", flush=True)
        assert rel(eng.download(eng.hpsi(x)), ref) < tol, f"config {tb}"


@pytest.mark.parametrize("N", [16, 20])
def test_chain_ground_energies_closed_forms(N):
    print(f"[chain ground states N={N}]", flush=True)
    eng = engine_for(O.ChainModel(N, K=(1.0, 1.0, 0.0)))
    _, e0, resid, _ = eng.ground_state(eng.start_vector(7, "Gaussian"), steps=300, eps=1e-13, restarts=1)
    assert abs(e0 - xx_ring_e0(N)) < 1e-10 * abs(e0)
    assert resid < 1e-5


def test_chain_heisenberg_ring_and_exact_diagonalisation():
    eng = engine_for(O.ChainModel(10))
    _, e0, _, _ = eng.ground_state(eng.start_vector(7, "Gaussian"), steps=200, eps=1e-13, restarts=1)
    assert abs(e0 - HEISENBERG_RING[10]) < 1e-10
    m = O.ChainModel(14, K=(0.9, 1.1, 0.6), B=(0.1, 0.2, 0.05))
    w0, _ = O.ground_pair(m)
    eng = engine_for(m)
    _, e0, _, _ = eng.ground_state(eng.start_vector(7, "Gaussian"), steps=300, eps=1e-13, restarts=1)
    assert abs(e0 - w0) < 1e-10 * abs(w0)


def test_chain_local_szsz_equals_exact_lehmann():
    m = O.ChainModel(8, K=(1.0, 0.8, 1.2), B=(0.1, 0.0, -0.2))
    H = O.hamiltonian_dense_kron(m)
    e, U = np.linalg.eigh(H)
    gs = U[:, 0]
    w = np.linspace(0.0, 4.0, 300)
    exact = O.local_szsz_ref(m, e, U.T.copy(), gs, omega=w, eta=0.1)
    eng = engine_for(m)
    tab, _ = eng.dynamics("LocalSzSz", "CF", eng.upload(gs), w, 0.1, e0=e[0], cf_steps=150)
    assert np.abs(tab[:, 1:] - exact[:, 1:]).max() < 1e-6 * np.abs(exact[:, 1:]).max()
    C = eng.two_point("Sz", "Sz", eng.upload(gs))
    ref = O.two_point_ref(m, "Sz", "Sz", gs)
    assert np.abs(C - ref).max() < 1e-12
    with pytest.raises(esc.EscError):         # current operators are honeycomb-only
        eng.dynamics("JeJe", "CF", eng.upload(gs), w, 0.1, e0=e[0], cf_steps=20)


def test_cli_heisenberg_chain(tmp_path):
    cli = os.path.join(ROOT, "energyspincurrents_amd", "LanczosCond")
    text = open(os.path.join(GOLDEN, "Midinput.inp")).read()
    text = (text.replace("Model=Kitaev", "Model=HeisenbergChain").replace("NumberofSitesLx=3", "NumberofSitesLx=12")
            .replace("NumberofSitesLy=4", "NumberofSitesLy=1"))
    for k in ("Bxx", "Byy", "Bzz"):
        text = text.replace(f"{k}=0.3", f"{k}=0.0")
    text = text.replace("EnergyCurrDynamics=1", "EnergyCurrDynamics=0") + "\nLocalSzSz=1\nSpectralMode=CF\n"
    assert "Model=HeisenbergChain" in text and "NumberofSitesLx=12" in text
    (tmp_path / "Midinput.inp").write_text(text)
    r = subprocess.run([cli, str(tmp_path / "Midinput.inp")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "creating Chain lattice" in r.stdout
    tab = O.read_real(str(tmp_path / "2SSLocal_outSpectrum.real"))
    assert tab.shape[1] == 9 and np.isfinite(tab).all()
    # SU(2) symmetry of the Heisenberg ring: the xx, yy and zz channels agree
    # (continued fractions: one recursion per operator, so a degenerate
    # multiplet counts in full, which Ritz states of one Krylov space do not)
    scale = np.abs(tab[:, 3]).max()
    assert scale > 0
    assert np.abs(tab[:, 1] - tab[:, 3]).max() < 1e-3 * scale
    assert np.abs(tab[:, 2] - tab[:, 3]).max() < 1e-3 * scale
