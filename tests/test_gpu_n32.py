"""The 32-site system (2^32 rows, complex64: 34.4 GB per vector) on one GPU.

H.psi: sampled rows against the oracle's row evaluator (orc_hpsi_rows, the
same matrix elements as the full gather) within 1e-6 relative (complex64
storage), Hermiticity <x|Hy> = conj <y|Hx> within 1e-6.  Ground state with
five resident vectors and a short JeJe continued-fraction spectrum: finite,
variational (E0 below the start vector's energy), sum rule of the spectrum.
Needs about 210 GB of device memory.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

esc = pytest.importorskip("energyspincurrents_amd")
torch = pytest.importorskip("torch")

B32 = O.field111(0.3)


def model32():
    return O.Model(4, 8, K=(-1.0, -1.0, -1.0), B=B32)


def big_enough():
    if not torch.cuda.is_available():
        return False
    return torch.cuda.get_device_properties(0).total_memory >= 250 * 2**30


pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not big_enough(), reason="needs >= 250 GiB of device memory")]


def engine32():
    m = model32()
    lat = esc.Honeycomb(m.Lx, m.Ly)
    return m, esc.Engine(esc.SpinModel(lat, K=m.K, B=m.B), "c64")


def test_n32_hpsi_sampled_rows_and_hermiticity():
    m, eng = engine32()
    assert eng.dim == 1 << 32
    print("[N=32 hpsi] start vectors", flush=True)
    x = eng.start_vector(1, "Gaussian")
    y = eng.start_vector(2, "Gaussian")
    hx = eng.hpsi(x)
    torch.cuda.synchronize()
    print("[N=32 hpsi] sampled rows", flush=True)
    rng = np.random.default_rng(32)
    rows = np.concatenate([[0, m.dim - 1], rng.integers(0, m.dim, 62, dtype=np.int64)])
    cols = O.row_columns(m, rows)
    idx = torch.from_numpy(cols).to("cuda")
    vals = x[idx].to(torch.complex128).cpu().numpy()
    ref = O.hpsi_rows(m, rows, cols, vals)
    got = hx[torch.from_numpy(rows).to("cuda")].to(torch.complex128).cpu().numpy()
    assert np.linalg.norm(got - ref) < 1e-6 * np.linalg.norm(ref)
    print("[N=32 hpsi] hermiticity", flush=True)
    hy = eng.hpsi(y)
    a = eng.dot(x, hy)
    b = eng.dot(y, hx)
    assert abs(a - b.conjugate()) < 1e-6 * max(abs(a), 1e-3)
    del x, y, hx, hy, idx
    torch.cuda.empty_cache()


def test_n32_ground_state_and_jeje_cf():
    m, eng = engine32()
    print("[N=32 gs] start", flush=True)
    v0 = eng.start_vector(7, "Gaussian")
    e_start = eng.dot(v0, eng.hpsi(v0)).real
    torch.cuda.empty_cache()
    gs, e0, resid, it = eng.ground_state(v0, steps=40, eps=0.0, restarts=0)
    del v0
    torch.cuda.empty_cache()
    print(f"[N=32 gs] E0 {e0} resid {resid} H.psi {it}", flush=True)
    assert np.isfinite(e0) and np.isfinite(resid)
    assert e0 < e_start - 1.0
    assert it == 40 + 39 + 1
    w = np.linspace(0.0, 6.0, 200)
    print("[N=32 JeJe CF]", flush=True)
    tab, expect = eng.dynamics("JeJe", "CF", gs, w, 0.05, e0=e0, cf_steps=20)
    assert tab.shape == (200, 6) and np.all(np.isfinite(tab))
    # -Im I_aa(w) >= 0 for the diagonal channels up to the elastic subtraction
    assert tab[:, 1].min() > -1e-3 * np.abs(tab[:, 1]).max()
