"""esc_dot_noalloc: the dot product that allocates nothing (its partial sums
in the library's fixed per-GPU array) against esc_dot, bit for bit, on two
calls in a row, the second on a smaller vector (a 12-site context after a
20-site one), and its result in device memory."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

esc = pytest.importorskip("energyspincurrents_amd")


def engine_for(m: O.Model, precision):
    lat = esc.Honeycomb(m.Lx, m.Ly, m.perX, m.perY)
    return esc.Engine(esc.SpinModel(lat, K=m.K, J=m.J, Gamma=m.G, B=m.B), precision)


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_dot_noalloc_equals_dot_on_two_calls(precision):
    import torch
    print(f"[no-alloc dot] {precision}", flush=True)
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    for (lx, ly), seed in (((2, 10), 1), ((2, 6), 2)):        # 2^20 rows, then 2^12
        m = O.Model(lx, ly, K=(-1.0, -1.0, -1.0), J=0.1, G=0.2, B=O.field111(0.2))
        eng = engine_for(m, precision)
        g = torch.Generator(device="cuda").manual_seed(seed)
        dt = torch.complex64 if precision == "c64" else torch.complex128
        x = torch.randn(m.dim, dtype=dt, device="cuda", generator=g)
        y = torch.randn(m.dim, dtype=dt, device="cuda", generator=g)
        ref = eng.dot(x, y)
        eng.dot_noalloc(x, y, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert got[0] == ref.real and got[1] == ref.imag, (m.N, got, ref)
        # and against an independent host sum
        xh = x.cpu().numpy().astype(np.complex128)
        yh = y.cpu().numpy().astype(np.complex128)
        exact = np.vdot(xh, yh)
        assert abs(complex(got[0], got[1]) - exact) <= 1e-12 * np.linalg.norm(xh) * np.linalg.norm(yh)
