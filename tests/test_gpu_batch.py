"""Batches of independent systems solved in lockstep (esc_batch_*): each
Lanczos step of all instances is one replayed HIP graph.

Bars: graph replays equal the eager launches bit for bit, on every run;
E0 of every instance against exact diagonalisation (N <= 12) within 1e-9 and
against the single-system ground state (N = 16) within 1e-9 relative; CF
spectra of every instance against the exact Lehmann sum within 1e-6 of the
peak; complex64 batches within 1e-6 relative of complex128 in E0.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

esc = pytest.importorskip("energyspincurrents_amd")


def spin_model(m: O.Model):
    lat = esc.Honeycomb(m.Lx, m.Ly, m.perX, m.perY)
    return esc.SpinModel(lat, K=m.K, J=m.J, Gamma=m.G, B=m.B)


def kjh(phi, h):
    """K = cos(phi), J = sin(phi) on every bond, field h along [111]."""
    return dict(K=(np.cos(phi),) * 3, J=float(np.sin(phi)), B=O.field111(h))


MIXED = [O.Model(2, 4, **kjh(0.3, 0.2)), O.Model(3, 4, K=(1.0, 1.0, 1.0), B=(0.3, 0.3, 0.3)),
         O.Model(2, 6, K=(-1.0, -1.0, -1.0), J=0.1, G=0.2, B=O.field111(0.2)), O.Model(2, 8, **kjh(2.0, 0.4)),
         O.Model(2, 4, **kjh(4.0, 0.1)), O.Model(3, 6, **kjh(1.0, 0.3))]


def test_batch_ground_states_graph_equals_eager_and_exact():
    print("[batch gs] mixed N=8,12,12,16,8,18", flush=True)
    b = esc.Batch([spin_model(m) for m in MIXED], "c128")
    e_eager, it_e, al_e, be_e = b.ground_state(seed=21, steps=200, use_graph=False)
    gs_eager = [b.state(i).cpu().numpy() for i in range(b.n)]
    for rep in range(2):                      # first run captures, the second replays
        e_g, it_g, al_g, be_g = b.ground_state(seed=21, steps=200, use_graph=True)
        assert np.array_equal(e_g, e_eager) and np.array_equal(it_g, it_e)
        assert np.array_equal(al_g, al_e) and np.array_equal(be_g, be_e)
        for i in range(b.n):
            assert np.array_equal(b.state(i).cpu().numpy(), gs_eager[i]), f"replay {rep} instance {i}"
    for i, m in enumerate(MIXED):
        if m.N <= 12:
            w0, _ = O.ground_pair(m)
            assert abs(e_eager[i] - w0) < 1e-9, f"instance {i}"
            g = gs_eager[i]
            assert abs(np.linalg.norm(g) - 1) < 1e-12
            assert np.linalg.norm(O.hpsi_gather(m, g) - w0 * g) < 1e-5
        else:
            eng = esc.Engine(spin_model(m), "c128")
            _, e1, _, _ = eng.ground_state(eng.start_vector(21 + i, "Gaussian"), steps=200, restarts=1)
            assert abs(e_eager[i] - e1) < 1e-9 * abs(e1)


@pytest.mark.parametrize("kind", ["JeJe", "JsJs"])
def test_batch_cf_spectra_equal_exact_lehmann(kind):
    ms = [O.Model(2, 4, K=(-1.0, -1.0, -1.0), J=0.1, G=0.2, B=O.field111(0.2)), O.Model(2, 4, **kjh(1.0, 0.3))]
    print(f"[batch {kind}]", flush=True)
    b = esc.Batch([spin_model(m) for m in ms], "c128")
    b.ground_state(seed=4, steps=200)
    w = np.linspace(0.0, 6.0, 300)
    eta = 0.1
    tab, ex = b.dynamics(kind, w, eta, cf_steps=150, use_graph=True)
    tab_e, ex_e = b.dynamics(kind, w, eta, cf_steps=150, use_graph=False)
    assert np.array_equal(tab, tab_e) and np.array_equal(ex, ex_e)
    for i, m in enumerate(ms):
        H = O.hamiltonian_dense_kron(m)
        e, U = np.linalg.eigh(H)
        exact = (O.jeje_ref if kind == "JeJe" else O.jsjs_ref)(m, e, U.T.copy(), U[:, 0], omega=w, eta=eta)
        scale = np.abs(exact[:, 1:]).max()
        assert np.abs(tab[i][:, 1:] - exact[:, 1:]).max() < 1e-6 * scale, f"instance {i}"


def test_batch_c64_energies_match_c128():
    """A small scan in the shape of the 64-instance study: sizes 16 and 18,
    K = cos(phi), J = sin(phi), h along [111]."""
    rng = np.random.default_rng(3)
    ms = [O.Model(*geom, **kjh(rng.uniform(0, 2 * np.pi), rng.uniform(0, 0.5)))
          for geom in [(2, 8), (3, 6), (2, 8), (3, 6)]]
    print("[batch c64 vs c128]", flush=True)
    e128, _, _, _ = esc.Batch([spin_model(m) for m in ms], "c128").ground_state(seed=9, steps=200)
    b64 = esc.Batch([spin_model(m) for m in ms], "c64")
    cfg = b64.autotune(reps=2)
    assert len(cfg) == len(ms)
    e64, _, _, _ = b64.ground_state(seed=9, steps=200)
    assert np.all(np.abs(e64 - e128) < 1e-6 * np.abs(e128))


def test_device_tridiagonal_and_fraction_routines_equal_the_host():
    """The kernels the batch runs for its ground states (lowest eigenpair of
    each T) and spectra (continued fractions) compute the host routines'
    bits: random matrices and the tridiagonal matrix of a plain 8-site run
    that repeats E0."""
    rng = np.random.default_rng(33)
    m = O.Model(2, 4, K=(-1.0, -1.0, -1.0), B=O.field111(0.3))
    al, be, n2 = O.lanczos_plain(lambda x: O.hpsi_gather(m, x), rng.standard_normal(m.dim) + 0j, 300)
    cases = [(al, be[:len(al)], n2)]
    for n in (1, 2, 7, 200):
        cases.append((rng.standard_normal(n), np.r_[0.0, 0.1 + np.abs(rng.standard_normal(n - 1))], 1.3))
    om = np.linspace(-2.0, 8.0, 1000)
    for a, b, nn in cases:
        print(f"[device tridiagonal] n={len(a)}", flush=True)
        e_h, z_h = esc.tridiag_lowest(a, b)
        e_d, z_d = esc.tridiag_lowest(a, b, device=True)
        assert e_h == e_d and np.array_equal(z_h, z_d)
        f_h = esc.cf_eval(a, b, nn, -e_h, 0.05, om)
        f_d = esc.cf_eval(a, b, nn, -e_h, 0.05, om, device=True)
        assert np.array_equal(f_h, f_d)


def test_batch_ground_state_energies_are_the_lowest_eigenvalues_of_the_returned_recursions():
    """E0 and the replay coefficients come from the device routine on the
    device's own alpha / beta rows: equal, bit for bit, to the host routine
    on the rows the call returns, and within rounding of LAPACK."""
    ms = MIXED[:4]
    b = esc.Batch([spin_model(m) for m in ms], "c128")
    e0, it, al, be = b.ground_state(seed=5, steps=120)
    for i in range(len(ms)):
        n = int(it[i])
        e_h, _ = esc.tridiag_lowest(al[i, :n], be[i, :n])
        assert e0[i] == e_h, (i, e0[i], e_h)
        T = np.diag(al[i, :n]) + np.diag(be[i, 1:n], 1) + np.diag(be[i, 1:n], -1)
        assert abs(e0[i] - np.linalg.eigvalsh(T)[0]) < 1e-12 * max(1.0, abs(e0[i]))


def test_device_all_eigenpairs_equal_the_host():
    """The tridiagonal eigensolver esc_lanczos and the RIXS fractions run on
    the device (bisection + clustered inverse iteration) computes the host
    routine's bits, on random matrices, Wilkinson's W21+ (close pairs) and a
    plain recursion's matrix with repeated eigenvalues."""
    rng = np.random.default_rng(41)
    m = O.Model(2, 4, K=(-1.0, -1.0, -1.0), B=O.field111(0.3))
    al, be, _ = O.lanczos_plain(lambda x: O.hpsi_gather(m, x), rng.standard_normal(m.dim) + 0j, 300)
    cases = [(al, be[:len(al)]), (np.abs(10.0 - np.arange(21.0)), np.r_[0.0, np.ones(20)])]
    for n in (1, 3, 64, 400):
        cases.append((rng.standard_normal(n), np.r_[0.0, 0.05 + np.abs(rng.standard_normal(n - 1))]))
    for a, b in cases:
        print(f"[device eigenpairs] n={len(a)}", flush=True)
        ev_h, Z_h = esc.tridiag_eigenpairs(a, b)
        ev_d, Z_d = esc.tridiag_eigenpairs(a, b, device=True)
        assert np.array_equal(ev_h, ev_d) and np.array_equal(Z_h, Z_d)


def test_shared_program_instances_equal_single_instance_batches():
    """Instances of one lattice and term structure (20 sites, K = cos(phi),
    J = sin(phi), [111] field) run their generated H.psi passes as one launch
    per pass (grid y = instance); an instance with Gamma has another program
    and runs alone.  Each instance's recursion and energy equal, bit for bit,
    those of a batch holding that instance only."""
    ms = [O.Model(2, 10, **kjh(0.4, 0.2)), O.Model(2, 10, **kjh(2.2, 0.35)),
          O.Model(2, 10, K=(-1.0, -1.0, -1.0), J=0.1, G=0.2, B=O.field111(0.2)), O.Model(2, 10, **kjh(4.1, 0.1))]
    print("[batch shared programs] 4 x N=20 c128", flush=True)
    e, it, al, be = esc.Batch([spin_model(m) for m in ms], "c128").ground_state(seed=5, steps=60)
    for i, m in enumerate(ms):
        e1, it1, al1, be1 = esc.Batch([spin_model(m)], "c128").ground_state(seed=5 + i, steps=60)
        assert e1[0] == e[i] and it1[0] == it[i], f"instance {i}"
        assert np.array_equal(al1[0], al[i]) and np.array_equal(be1[0], be[i]), f"instance {i}"
