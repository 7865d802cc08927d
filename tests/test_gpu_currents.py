"""Single-site operators and current operators on the device (esc_apply_op)
against the oracle's restatement of src/Observables.h.

Bar: bit-exact for dyadic couplings (every factor is +-1/2 or +-i/2 and the
weights are exact) up to 16 sites, where k_current applies each row's strings
in the reference's order; 1e-14 relative otherwise and from 17 sites, where
the windowed tiled kernel applies the strings as gather terms (merged by
mask, a different summation order); c64 within 1e-6.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

esc = pytest.importorskip("energyspincurrents_amd")


def engine_for(m: O.Model, precision="c128"):
    lat = esc.Honeycomb(m.Lx, m.Ly, m.perX, m.perY)
    return esc.Engine(esc.SpinModel(lat, K=m.K, J=m.J, Gamma=m.G, B=m.B), precision)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def setup12():
    m = O.Model(3, 4, K=(1.0, 1.0, 1.0), B=(0.3, 0.3, 0.3))
    rng = np.random.default_rng(12)
    v = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    eng = engine_for(m)
    return m, v, eng, eng.upload(v)


@pytest.mark.parametrize("op", ["Sx", "Sy", "Sz"])
def test_site_ops_bit_exact(setup12, op):
    m, v, eng, dv = setup12
    for s in range(m.N):
        got = eng.download(eng.apply_op(op, s, dv))
        assert np.array_equal(got, O.apply_site_op(m.N, op, s, v)), f"{op} site {s}"


@pytest.mark.parametrize("name", ["Jsx", "Jsy", "Jsz", "Jex", "Jey"])
def test_current_sites_and_totals_bit_exact(setup12, name):
    m, v, eng, dv = setup12
    print(f"[current {name}]", flush=True)
    sites = range(0, m.N, 2) if name == "Jex" else range(1, m.N, 2) if name == "Jey" else range(m.N)
    for s in sites:
        got = eng.download(eng.apply_op(name, s, dv))
        assert np.array_equal(got, O.apply_current_site(m, name, s, v)), f"{name} site {s}"
    got = eng.download(eng.apply_total(name, dv))
    assert np.array_equal(got, O.apply_current_total(m, name, v))


@pytest.mark.parametrize("name", ["Jsx", "Jsy", "Jsz", "Jex", "Jey"])
def test_current_totals_non_dyadic(name):
    m = O.Model(2, 8, K=(-1.0, 0.7, 0.3))
    rng = np.random.default_rng(16)
    v = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    eng = engine_for(m)
    got = eng.download(eng.apply_total(name, eng.upload(v)))
    assert rel(got, O.apply_current_total(m, name, v)) < 1e-14
    eng64 = engine_for(m, "c64")
    v32 = v.astype(np.complex64).astype(np.complex128)
    got = eng64.download(eng64.apply_total(name, eng64.upload(v32)))
    assert rel(got, O.apply_current_total(m, name, v32)) < 1e-6


@pytest.mark.parametrize("lx,ly", [(3, 6), (2, 10), (3, 8)])
def test_tiled_current_operators_from_17_sites(lx, ly, capfd, monkeypatch):
    """From 17 sites totals and per-site currents run on the windowed tiled
    kernel (the strings as gather terms): against the oracle's op-by-op
    restatement within 1e-14 (c128) and 1e-6 (c64), K-J-Gamma-h couplings."""
    monkeypatch.setenv("ESC_VERBOSE", "1")
    m = O.Model(lx, ly, K=(-1.0, 0.7, 0.3), J=0.1, G=0.2, B=O.field111(0.2))
    rng = np.random.default_rng(m.N)
    v = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    eng = engine_for(m)
    dv = eng.upload(v)
    eng64 = engine_for(m, "c64")
    v32 = v.astype(np.complex64).astype(np.complex128)
    dv32 = eng64.upload(v32)
    for name in ("Jsx", "Jsy", "Jsz", "Jex", "Jey"):
        print(f"[tiled current] N={m.N} {name}", flush=True)
        ref = O.apply_current_total(m, name, v)
        assert rel(eng.download(eng.apply_total(name, dv)), ref) < 1e-14, name
        ref32 = O.apply_current_total(m, name, v32)
        assert rel(eng64.download(eng64.apply_total(name, dv32)), ref32) < 1e-6, name
    site = 2
    for name in ("Jsx", "Jex"):
        ref = O.apply_current_site(m, name, site, v)
        assert rel(eng.download(eng.apply_op(name, site, dv)), ref) < 1e-14, name
    err = capfd.readouterr().err
    assert "tiled kernel" in err, err[-2000:]
