"""Gradients of a spectrum loss with respect to (K, J, Gamma, h)
(esc_spectrum_grad): against central finite differences in FP64 (steps 1e-4
and 5e-5, Richardson-combined), Hellmann-Feynman dE0/dtheta against finite
differences of E0 (relative 1e-6), checkpoint intervals that must give the
same gradient bit for bit, and the torch.autograd wrapper.

The loss is the output of a truncated Lanczos recursion; its late
coefficients amplify rounding (the recursion loses orthogonality), so the
loss computed in floating point carries noise that finite differences
divide by the step.  Bars: 1e-6 relative at N=8 (20 steps), 1e-5 at N=20
with 40 steps, 1e-2 / 2e-2 at N=20 with 120 steps.  The table losses (every column
of the JsJs / JeJe tables) are checked the same way at N=8."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

esc = pytest.importorskip("energyspincurrents_amd")
torch = pytest.importorskip("torch")

THETA = np.array([-1.0, 0.15, 0.12, 0.25])


def setup(lx=2, ly=4, op="Jex", site=0, cf_steps=20, nw=200, checkpoint=10):
    w = np.linspace(0.0, 6.0, nw)
    spec = esc.SpectrumLoss(esc.Honeycomb(lx, ly), w, eta=0.1, op=op, site=site, gs_steps=300, cf_steps=cf_steps,
                            checkpoint=checkpoint)
    target = np.random.default_rng(1).uniform(0.0, 1.0, nw)
    return spec, target


def central(spec, target, theta, p, step, key="loss"):
    tp, tm = theta.copy(), theta.copy()
    tp[p] += step
    tm[p] -= step
    return (spec.evaluate(tp, target)[key] - spec.evaluate(tm, target)[key]) / (2 * step)


def fd(spec, target, theta, p, h, key="loss", spread=False):
    """Central differences with steps h and h/2, Richardson-combined (error
    O(h^4)): the loss is sharply curved in K, which scales every energy.
    With spread, also |D(h) - D(h/2)| (the differences' own noise)."""
    d1 = central(spec, target, theta, p, h, key)
    d2 = central(spec, target, theta, p, h / 2, key)
    d = (4.0 * d2 - d1) / 3.0
    return (d, abs(d1 - d2)) if spread else d


@pytest.mark.parametrize("op,site", [("Jex", 0), ("Jsx", 0), ("Sz", 3)])
def test_gradient_matches_finite_differences(op, site):
    spec, target = setup(op=op, site=site)
    print(f"[grad N=8] {op}", flush=True)
    r = spec.evaluate(THETA, target)
    for p in range(4):
        d = fd(spec, target, THETA, p, 1e-4)
        assert abs(r["grad"][p] - d) < 1e-6 * max(1.0, abs(d)), f"theta[{p}]: adjoint {r['grad'][p]} fd {d}"


def test_hellmann_feynman_matches_finite_differences():
    spec, target = setup()
    r = spec.evaluate(THETA, target)
    for p in range(4):
        d = fd(spec, target, THETA, p, 1e-5, key="e0")
        assert abs(r["de0"][p] - d) < 1e-6 * max(1.0, abs(d))


def test_checkpoint_interval_does_not_change_the_gradient():
    g = []
    for c in (1, 7, 10, 30):
        spec, target = setup(checkpoint=c)
        g.append(spec.evaluate(THETA, target)["grad"])
    for x in g[1:]:
        assert np.array_equal(x, g[0])


def test_torch_autograd_function():
    spec, target = setup()
    theta = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    tgt = torch.tensor(target, dtype=torch.float64, requires_grad=True)
    loss = esc.spectrum_loss(theta, tgt, spec)
    loss.backward()
    r = spec.evaluate(THETA, target)
    assert abs(loss.item() - r["loss"]) < 1e-12 * r["loss"]
    assert np.allclose(theta.grad.numpy(), r["grad"], rtol=1e-12, atol=0)
    assert np.allclose(tgt.grad.numpy(), -2.0 * (r["spectrum"] - target), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("cf_steps,checkpoint,bar", [(40, 10, 1e-6), (120, 0, 1e-4), (120, 10, 1e-3)])
def test_n20_study_gradient(cf_steps, checkpoint, bar):
    """The 20-site inverse problem: K = -1 + d, J, Gamma, h random, 1000
    frequencies, eta = 0.1; finite differences of the loss in FP64 with steps
    1e-4 and 5e-5, Richardson-combined.  Bar on |grad - fd| / |fd| over the
    four components: 1e-6 for 40 steps (measured 7.5e-8), 1e-4 for 120
    steps reorthogonalised (measured 9.3e-6, about the finite differences'
    own spread of 2e-4 to 5e-4 per component), 1e-3 for 120 checkpointed
    steps (measured 1.4e-4; the plain recursion's loss is noisier, spreads of
    1e-3 to 1.2e-2 per component).  The adjoint's ground-state solve runs to
    the rounding level: its right-hand side is about 1e9 times the gradient
    at 120 steps."""
    rng = np.random.default_rng(20)
    theta = np.array([-1.0 + rng.uniform(0, 0.2), rng.uniform(0, 0.2), rng.uniform(0, 0.2), rng.uniform(0, 0.4)])
    w = np.linspace(0.0, 6.0, 1000)
    spec = esc.SpectrumLoss(esc.Honeycomb(2, 10), w, eta=0.1, op="Jex", gs_steps=300, cf_steps=cf_steps,
                            checkpoint=checkpoint)
    target = spec.evaluate(theta + np.array([0.05, -0.03, 0.02, 0.05]), np.zeros_like(w))["spectrum"]
    print(f"[grad N=20] {cf_steps} steps, checkpoint {checkpoint}", flush=True)
    r = spec.evaluate(theta, target)
    assert r["loss"] > 0
    d = np.zeros(4)
    for p in range(4):
        d[p], spread = fd(spec, target, theta, p, 1e-4, spread=True)
        print(f"  theta[{p}] adjoint {r['grad'][p]:.10g} fd {d[p]:.10g} spread {spread:.3g}", flush=True)
    err = np.linalg.norm(r["grad"] - d) / np.linalg.norm(d)
    print(f"  relative error {err:.3g}", flush=True)
    assert err < bar


@pytest.mark.parametrize("kind", ["JeJe", "JsJs"])
def test_table_loss_gradient_matches_finite_differences(kind):
    """Loss over the whole table (esc_table_grad): every column, the cross
    channels by polarisation and the elastic terms, N=8, 20 steps."""
    w = np.linspace(0.0, 6.0, 150)
    spec = esc.TableLoss(esc.Honeycomb(2, 4), w, eta=0.1, kind=kind, gs_steps=300, cf_steps=20, checkpoint=10)
    target = np.random.default_rng(2).uniform(0.0, 0.5, (len(w), spec.ncols))
    print(f"[table grad N=8] {kind}", flush=True)
    r = spec.evaluate(THETA, target)
    assert r["table"].shape == (len(w), 1 + spec.ncols) and np.array_equal(r["table"][:, 0], w)
    assert abs(r["loss"] - ((r["table"][:, 1:] - target) ** 2).sum()) < 1e-12 * r["loss"]
    for p in range(4):
        d = fd(spec, target, THETA, p, 1e-4)
        print(f"  theta[{p}] adjoint {r['grad'][p]:.12g} fd {d:.12g}", flush=True)
        assert abs(r["grad"][p] - d) < 1e-6 * max(1.0, abs(d)), f"theta[{p}]: adjoint {r['grad'][p]} fd {d}"


def test_table_loss_columns_match_the_spectra_entry_point():
    """The table the gradient call evaluates is the CF-mode JeJe table of
    esc_dynamics from the same ground state: same E0, columns within 1e-8."""
    w = np.linspace(0.0, 6.0, 150)
    spec = esc.TableLoss(esc.Honeycomb(2, 4), w, eta=0.1, kind="JeJe", gs_steps=300, cf_steps=20, checkpoint=10)
    r = spec.evaluate(THETA, np.zeros((len(w), 5)))
    eng = esc.Engine(spec.model(THETA), "c128")
    gs, e0, _, _ = eng.ground_state(eng.start_vector(3, "Gaussian"), steps=300, restarts=1)
    assert abs(e0 - r["e0"]) < 1e-10 * abs(e0)
    tab, _ = eng.dynamics("JeJe", "CF", gs, w, 0.1, e0=r["e0"], cf_steps=20)
    assert np.abs(tab - r["table"]).max() < 1e-8 * np.abs(tab[:, 1:]).max()


def test_hellmann_feynman_at_20_sites():
    """dE0/dtheta = <gs|dH/dtheta|gs> against finite differences of E0 on the
    20-site cluster of the inverse-problem study."""
    rng = np.random.default_rng(20)
    theta = np.array([-1.0 + rng.uniform(0, 0.2), rng.uniform(0, 0.2), rng.uniform(0, 0.2), rng.uniform(0, 0.4)])
    w = np.linspace(0.0, 6.0, 50)
    spec = esc.SpectrumLoss(esc.Honeycomb(2, 10), w, eta=0.1, op="Jex", gs_steps=300, cf_steps=4, checkpoint=2)
    r = spec.evaluate(theta, np.zeros_like(w))
    for p in range(4):
        d = fd(spec, np.zeros_like(w), theta, p, 1e-4, key="e0")
        print(f"[HF N=20] theta[{p}] {r['de0'][p]:.12g} fd {d:.12g}", flush=True)
        assert abs(r["de0"][p] - d) < 1e-7 * max(1.0, abs(d))
