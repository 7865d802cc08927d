"""LocalDynRIXS (src/dynamics.h:465-562) and DynSzSz (src/dynamics.h:566-611)
on the device (esc_rixs, esc_szsz) against the oracle.

* Lanczos runs as long as the space (N=8, 256 steps, full
  reorthogonalisation): the Ritz sums equal the dense resolvent formulas of
  tests/test_oracle_rixs.py, relative 1e-7; the CF mode of esc_rixs too.
* Shorter runs (40 steps), the reference's situation: equal to the oracle's
  literal restatement fed with the oracle's own Lanczos runs from the same
  mt19937_64 start vectors (the generator continues from one run to the next,
  as the reference's shared generator does), relative 1e-8.
* The LanczosCond CLI with SzSz=1 and LocalRIXS=1 writes both files in the
  reference's row layout with these values."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT
from test_oracle_rixs import _dense_rixs

pytestmark = pytest.mark.gpu

esc = pytest.importorskip("energyspincurrents_amd")

M8 = O.Model(2, 4, K=(-1.0, -0.8, -1.2), B=(0.1, 0.2, 0.3))
SITE = 3


def engines(m, site, precision="c128"):
    lat = esc.Honeycomb(m.Lx, m.Ly, m.perX, m.perY)
    model = esc.SpinModel(lat, K=m.K, J=m.J, Gamma=m.G, B=m.B)
    return esc.Engine(model, precision), esc.Engine(model.intermediate(site), precision)


def stream_vectors(seed, dim, count):
    """Consecutive reference start vectors of one mt19937_64 generator."""
    u = O.mt19937_64(seed, 2 * dim * count) / 2.0 ** 64
    out = []
    for k in range(count):
        h = -1.0 + 2.0 * u[2 * dim * k: 2 * dim * (k + 1)]
        v = h[0::2] + 1j * h[1::2]
        out.append(v / np.linalg.norm(v))
    return out


def test_rixs_full_space_ritz_and_cf_equal_dense_resolvent():
    m = M8
    print("[rixs full space N=8]", flush=True)
    eng, inter = engines(m, SITE)
    v1, v2 = stream_vectors(3, m.dim, 2)
    gs = eng.empty()
    res = eng.lanczos(eng.upload(v1), m.dim, 0.0, "All", gs)
    inter.lanczos(inter.upload(v2), m.dim, 0.0, "All", inter.empty())
    w = O.TENTH_OMEGA[::4]
    g = eng.download(gs)
    e = np.linalg.eigvalsh(O.hamiltonian_dense_kron(m))
    assert abs(res.evals[0] - e[0]) < 1e-11
    ref = _dense_rixs(m, SITE, g, e[0], w)
    scale = np.abs(ref).max()
    tab = eng.rixs(inter, SITE, gs, "Ritz", omega=w)
    assert np.abs(tab[:, 1:5] - ref).max() < 1e-7 * scale
    assert np.all(tab[:, 5:] == 0.0) and np.array_equal(tab[:, 0], w)
    tab_cf = eng.rixs(inter, SITE, gs, "CF", e0=e[0], cf_steps=120, omega=w)
    assert np.abs(tab_cf[:, 1:5] - ref).max() < 1e-6 * scale


def test_rixs_and_szsz_short_runs_match_oracle_restatement():
    m = M8
    print("[rixs / szsz 40 steps N=8]", flush=True)
    eng, inter = engines(m, SITE)
    v1, v2, v3 = stream_vectors(5, m.dim, 3)
    gs = eng.empty()
    eng.lanczos(eng.upload(v1), 40, 0.0, "All", gs)
    inter.lanczos(inter.upload(v2), 40, 0.0, "All", inter.empty())
    tab = eng.rixs(inter, SITE, gs, "Ritz")
    g_ref = O.lanczos_ref(lambda v: O.hpsi_gather(m, v), v1, 40)
    k = O.kill_site(m, SITE)
    i_ref = O.lanczos_ref(lambda v: O.hpsi_gather(k, v), v2, 40)
    ref = O.rixs_ref(m.N, SITE, g_ref.evals, O.ritz_vectors(g_ref), i_ref.evals, O.ritz_vectors(i_ref), g_ref.gs)
    assert np.abs(tab[:, 1:] - ref[:, 1:]).max() < 1e-8 * np.abs(ref[:, 1:]).max()
    # DynSzSz: a second "All" run on H (the third start vector), Psi from the first
    eng2, _ = engines(m, SITE)
    eng2.lanczos(eng2.upload(v3), 40, 0.0, "All", eng2.empty())
    S = eng2.szsz(gs)
    s_ref = O.lanczos_ref(lambda v: O.hpsi_gather(m, v), v3, 40)
    S_ref = O.szsz_ref(m.N, s_ref.evals, O.ritz_vectors(s_ref), g_ref.gs)
    assert np.abs(S - S_ref).max() < 1e-9 * np.abs(S_ref).max()


def test_szsz_full_space_equals_dense_resolvent():
    m = M8
    eng, _ = engines(m, SITE)
    v1, = stream_vectors(7, m.dim, 1)
    gs = eng.empty()
    eng.lanczos(eng.upload(v1), m.dim, 0.0, "All", gs)
    g = eng.download(gs)
    H = O.hamiltonian_dense_kron(m)
    e = np.linalg.eigvalsh(H)
    w = O.TENTH_OMEGA[::12]
    S = eng.szsz(gs, omega=w)
    Z = np.array([O.apply_site_op(m.N, "Sz", i, g) for i in range(m.N)])
    for q, wq in enumerate(w):
        G = Z.conj() @ np.linalg.solve((wq + e[0] - 0.1j) * np.eye(m.dim) - H, Z.T)
        ref = 0.5 * (G + G.T)
        assert np.abs(S[q] - ref).max() < 1e-8 * max(1.0, np.abs(ref).max())


def test_rixs_errors():
    eng, inter = engines(M8, SITE)
    gs = eng.start_vector(1, "Gaussian")
    with pytest.raises(esc.EscError):          # no kept bases
        eng.rixs(inter, SITE, gs, "Ritz")
    with pytest.raises(esc.EscError):
        eng.rixs(inter, 99, gs, "CF")
    with pytest.raises(esc.EscError):
        eng.szsz(gs)


def _read_rows(path):
    return np.array([[float(x) for x in ln.split(" \t ")] for ln in open(path).read().splitlines()])


def test_cli_szsz_and_rixs_files(tmp_path):
    """Midinput.inp keys for the 8-site cluster above, 40 Lanczos steps, with
    SzSz=1 and LocalRIXS=1 at SpecialSite=3: the files hold the oracle's
    values (six significant digits) in the reference's rows."""
    m = M8
    cli = os.path.join(ROOT, "energyspincurrents_amd", "LanczosCond")
    text = "\n".join([
        "NumberofSitesLx=2", "NumberofSitesLy=4", "IsPeriodicX=1", "IsPeriodicY=1", "RandomSeed=3",
        "Model=Kitaev", "Solver=none", "Temperature=1", "LanczosEps=0", "LanczosSteps=40", "Threads=1",
        "LancType=All", "LocalSkw=0", "SpinCurrDynamics=0", "EnergyCurrDynamics=0",
        "Kxx=-1.0", "Kyy=-0.8", "Kzz=-1.2", "Bxx=0.1", "Byy=0.2", "Bzz=0.3",
        "SzSz=1", "LocalRIXS=1", "SpecialSite=3", ""])
    (tmp_path / "Midinput.inp").write_text(text)
    r = subprocess.run([cli, str(tmp_path / "Midinput.inp")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    # the generator continues: ground state, the DynSzSz run, the RIXS run
    v1, v2, v3 = stream_vectors(3, m.dim, 3)
    g = O.lanczos_ref(lambda v: O.hpsi_gather(m, v), v1, 40)
    s = O.lanczos_ref(lambda v: O.hpsi_gather(m, v), v2, 40)
    k = O.kill_site(m, SITE)
    i = O.lanczos_ref(lambda v: O.hpsi_gather(k, v), v3, 40)
    raw = open(tmp_path / "SzSz_outSpectrum.real").read()
    assert raw.startswith("0 \t 0 \t 0 \t ") and raw.endswith("\n")
    rows = _read_rows(tmp_path / "SzSz_outSpectrum.real")
    assert rows.shape == (121 * 64, 5)
    S_ref = O.szsz_ref(m.N, s.evals, O.ritz_vectors(s), g.gs)
    got = (rows[:, 3] + 1j * rows[:, 4]).reshape(121, 8, 8)
    assert np.array_equal(rows[:, 0].reshape(121, 8, 8)[0], np.repeat(np.arange(8), 8).reshape(8, 8))
    assert np.abs(got - S_ref).max() < 1e-5 * np.abs(S_ref).max()
    rows = _read_rows(tmp_path / "RIXS_outSpectrum.real")
    assert rows.shape == (121, 9)
    ref = O.rixs_ref(m.N, SITE, g.evals, O.ritz_vectors(g), i.evals, O.ritz_vectors(i), g.gs)
    assert np.abs(rows[:, 1:] - ref[:, 1:]).max() < 1e-5 * np.abs(ref[:, 1:]).max()
