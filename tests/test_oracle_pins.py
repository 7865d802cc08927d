"""Pins of the CPU oracle (oracle/) against the reference's committed outputs,
closed forms, an independent dense Kronecker formulation, full diagonalisation
and planted mistakes.  CPU only."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

# -- lattice -----------------------------------------------------------------

# 12-site table for Lx=3, Ly=4 derived by hand from src/Lattice.h:434-576
HAND_12 = {
    "x": [(0, 1), (2, 3), (4, 5), (6, 7), (8, 9), (10, 11)],
    "y": [(0, 9), (1, 4), (2, 11), (3, 6), (5, 8), (7, 10)],
    "z": [(0, 3), (1, 2), (4, 7), (5, 6), (8, 11), (9, 10)],
}


def test_honeycomb_12_site_known_answer():
    m = O.Model(3, 4)
    for f, name in enumerate("xyz"):
        assert [tuple(p) for p in m.bonds[f]] == HAND_12[name]
    # every site has exactly one neighbour of each kind
    assert (m.neigh >= 0).all()


def closed_form_neigh(Lx, Ly, px, py):
    LLY = Ly // 2
    N = Lx * Ly
    nb = -np.ones((N, 3), int)
    A = lambda i, j: 2 * (i * LLY + j)
    for i in range(Lx):
        for j in range(LLY):
            a, b = A(i, j), A(i, j) + 1
            nb[a, 2], nb[b, 2] = b, a
            if i + 1 < Lx:
                nb[b, 1], nb[A(i + 1, j), 1] = A(i + 1, j), b
            elif px:
                nb[A(0, j), 1], nb[b, 1] = b, A(0, j)
            if j + 1 < LLY:
                nb[b, 0], nb[A(i, j + 1), 0] = A(i, j + 1), b
            elif py:
                nb[b, 0], nb[A(i, 0), 0] = A(i, 0), b
    return nb


@pytest.mark.parametrize("Lx", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("Ly", [2, 4, 6, 8, 10])
@pytest.mark.parametrize("px,py", [(1, 1), (0, 1), (1, 0), (0, 0)])
def test_honeycomb_grid_restatement_equals_torus_closed_form(Lx, Ly, px, py):
    nb, _, _ = O.honeycomb1(Lx, Ly, px, py)
    assert np.array_equal(nb, closed_form_neigh(Lx, Ly, px, py))


# -- generator ---------------------------------------------------------------

def test_mt19937_64_standard_requirement():
    # C++ [rand.predef]: the 10000th consecutive invocation of a default-
    # constructed mt19937_64 produces 9981545732273789042
    assert int(O.mt19937_64(5489, 10000)[-1]) == 9981545732273789042


def test_start_vector_range_and_order():
    v = O.start_vector_ref(3, 16, normalize=False)
    raw = O.mt19937_64(3, 32).astype(np.float64) / 2.0**64
    assert np.array_equal(v.real, -1 + 2 * raw[0::2])
    assert np.array_equal(v.imag, -1 + 2 * raw[1::2])


# -- Hamiltonian ---------------------------------------------------------------

MODELS_8 = [
    dict(Lx=2, Ly=4, K=(1.0, 1.0, 1.0), B=(0.3, 0.3, 0.3)),
    dict(Lx=2, Ly=4, K=(-1.0, 0.7, 0.3), B=(0.1, -0.2, 0.3)),
    dict(Lx=1, Ly=8, K=(-1.0, -1.0, -1.0), B=(0.0, 0.0, 0.5)),
    dict(Lx=4, Ly=2, K=(0.4, -1.2, 0.9), B=(0.0, 0.25, 0.0)),
]


@pytest.mark.parametrize("kw", MODELS_8)
def test_coo_and_gather_equal_kronecker(kw):
    m = O.Model(**kw)
    Hk = O.hamiltonian_dense_kron(m)
    coo = O.CooHamiltonian(m)
    assert np.abs(coo.dense() - Hk).max() < 1e-15
    rng = np.random.default_rng(7)
    v = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    ref = Hk @ v
    assert np.abs(O.hpsi_gather(m, v) - ref).max() < 1e-13
    assert np.abs(coo.apply(v, 3) - ref).max() < 1e-13


@pytest.mark.parametrize("J,G", [(0.1, 0.0), (0.0, 0.2), (0.13, -0.21)])
def test_gather_extensions_equal_kronecker(J, G):
    m = O.Model(2, 4, K=(-1.0, 0.8, 0.6), J=J, G=G, B=(0.11, 0.2, -0.3))
    Hk = O.hamiltonian_dense_kron(m)
    assert np.abs(Hk - Hk.conj().T).max() == 0.0
    v = np.random.default_rng(3).standard_normal(m.dim) + 0.5j
    assert np.abs(O.hpsi_gather(m, v) - Hk @ v).max() < 1e-13


def test_trace_identities():
    m = O.Model(3, 4, K=(-1.0, 0.7, 0.3), B=(0.1, -0.2, 0.3))
    H = O.CooHamiltonian(m).dense()
    assert abs(np.trace(H)) < 1e-12
    nb = sum(len(b) for b in m.bonds)
    assert nb == 18
    expect = sum(len(m.bonds[f]) * m.K[f] ** 2 / 16 for f in range(3)) + m.N * sum(b * b for b in m.B) / 4
    assert abs(np.vdot(H, H).real / m.dim - expect) < 1e-12


def _hexagons(m):
    """6-cycles of the bond graph with their outward bond type per site."""
    adj = {i: {} for i in range(m.N)}
    for f, pairs in enumerate(m.bonds):
        for i, j in pairs:
            adj[i][j] = f
            adj[j][i] = f
    cycles = set()
    for s in range(m.N):
        stack = [(s, [s])]
        while stack:
            u, path = stack.pop()
            for w in adj[u]:
                if len(path) == 6 and w == s:
                    cycles.add(tuple(sorted(path)))
                elif w not in path and len(path) < 6:
                    stack.append((w, path + [w]))
    hexes = []
    for cyc in cycles:
        cs = set(cyc)
        out = {}
        for i in cyc:
            inner = [f for j, f in adj[i].items() if j in cs]
            outer = [f for j, f in adj[i].items() if j not in cs]
            if len(inner) != 2 or len(outer) != 1:
                break
            out[i] = outer[0]
        else:
            hexes.append(out)
    return hexes


def test_plaquette_operators_commute_only_without_field():
    names = ("Sx", "Sy", "Sz")
    m0 = O.Model(3, 4, K=(1.0, 1.0, 1.0))
    m1 = O.Model(3, 4, K=(1.0, 1.0, 1.0), B=(0.3, 0.3, 0.3))
    H0 = O.CooHamiltonian(m0).dense()
    H1 = O.CooHamiltonian(m1).dense()
    # N/2 hexagonal faces plus non-contractible 6-loops of the small torus; every
    # closed-loop operator of the Kitaev model commutes with H at zero field
    hexes = _hexagons(m0)
    assert len(hexes) >= m0.N // 2
    for out in hexes[::5]:
        W = np.eye(m0.dim, dtype=complex)
        for i, f in out.items():
            W = W @ (2 * O.site_op_dense(m0.N, names[f], i))
        assert np.abs(H0 @ W - W @ H0).max() < 1e-12
        assert np.abs(H1 @ W - W @ H1).max() > 1e-3


# -- currents ------------------------------------------------------------------

def _dense_current(m, name, i):
    S = lambda a, k: O.site_op_dense(m.N, a, k)
    nz, ny, nx = m.neigh[i]
    Kx, Ky, Kz = m.K
    if name == "Jsx":
        return Kz * S("Sy", i) @ S("Sz", nz) - Ky * S("Sz", i) @ S("Sy", ny)
    if name == "Jsy":
        return Kx * S("Sz", i) @ S("Sx", nx) - Kz * S("Sx", i) @ S("Sz", nz)
    if name == "Jsz":
        return Ky * S("Sx", i) @ S("Sy", ny) - Kx * S("Sy", i) @ S("Sx", nx)
    if name == "Jex":
        ix = nx
        ixpz, ixpy = m.neigh[ix][0], m.neigh[ix][1]
        return (Kx * Kz * S("Sx", i) @ S("Sy", ix) @ S("Sz", ixpz)
                - Kx * Ky * S("Sx", i) @ S("Sz", ix) @ S("Sy", ixpy))
    iy = ny
    iypz, iypx = m.neigh[iy][0], m.neigh[iy][2]
    return (Kx * Ky * S("Sy", i) @ S("Sz", iy) @ S("Sx", iypx)
            - Ky * Kz * S("Sy", i) @ S("Sx", iy) @ S("Sz", iypz))


@pytest.mark.parametrize("name", ["Jsx", "Jsy", "Jsz", "Jex", "Jey"])
def test_currents_equal_kronecker(name):
    m = O.Model(2, 4, K=(-1.0, 0.7, 0.3))
    v = np.random.default_rng(11).standard_normal(m.dim) + 1j * np.random.default_rng(12).standard_normal(m.dim)
    sites = range(0, m.N, 2) if name == "Jex" else range(1, m.N, 2) if name == "Jey" else range(m.N)
    tot = np.zeros((m.dim, m.dim), complex)
    for i in sites:
        D = _dense_current(m, name, i)
        tot += D
        assert np.abs(O.apply_current_site(m, name, i, v) - D @ v).max() < 1e-13
    assert np.abs(O.apply_current_total(m, name, v) - tot @ v).max() < 1e-12


@pytest.mark.parametrize("op", ["Sx", "Sy", "Sz"])
def test_site_ops_equal_kronecker(op):
    N = 6
    v = np.random.default_rng(5).standard_normal(1 << N) + 0.3j
    for s in range(N):
        assert np.abs(O.apply_site_op(N, op, s, v) - O.site_op_dense(N, op, s) @ v).max() < 1e-15


# -- Lanczos and spectra -------------------------------------------------------

def test_lanczos_ground_energy_equals_full_diagonalisation():
    m = O.Model(2, 4, K=(-1.0, -1.0, -1.0), B=O.field111(0.2))
    H = O.hamiltonian_dense_kron(m)
    e = np.linalg.eigvalsh(H)
    coo = O.CooHamiltonian(m)
    res = O.lanczos_ref(coo.apply, O.start_vector_ref(3, m.dim), 60)
    assert abs(res.evals[0] - e[0]) < 1e-12
    # reorthogonalised basis is orthonormal
    Vb = res.basis[: res.iters]
    assert np.abs(Vb.conj() @ Vb.T - np.eye(res.iters)).max() < 1e-12


def test_cf_equals_exact_lehmann_sum():
    m = O.Model(2, 4, K=(-1.0, -1.0, -1.0), B=O.field111(0.3))
    H = O.hamiltonian_dense_kron(m)
    e, U = np.linalg.eigh(H)
    gs = U[:, 0]
    phi = O.apply_current_total(m, "Jex", gs)
    w = np.linspace(0, 6, 400)
    exact = ((np.abs(U.conj().T @ phi) ** 2)[None, :] / (w[:, None] + e[0] - e[None, :] - 0.1j)).sum(1)
    a, b, n2 = O.lanczos_plain(lambda v: H @ v, phi, 120)
    cf = O.cf_resolvent(a, b, n2, w + e[0] - 0.1j)
    assert np.abs(cf - exact).max() < 1e-8 * np.abs(exact).max()


@pytest.mark.parametrize("kind", ["JeJe", "JsJs"])
def test_cf_dynamics_table_equals_exact_lehmann_tables(kind):
    """The oracle's continued-fraction tables (polarised cross channels,
    elastic terms) equal the Lehmann tables of full diagonalisation, and a
    planted sign error in the polarisation is caught."""
    m = O.Model(2, 4, K=(-1.0, -1.0, -1.0), J=0.1, G=0.2, B=O.field111(0.2))
    H = O.hamiltonian_dense_kron(m)
    e, U = np.linalg.eigh(H)
    gs = U[:, 0]
    w = np.linspace(0, 6, 300)
    exact = (O.jeje_ref if kind == "JeJe" else O.jsjs_ref)(m, e, U.T.copy(), gs, omega=w, eta=0.1)
    got = O.cf_dynamics_table(m, gs, e[0], w, 0.1, 150, kind, apply_h=lambda v: H @ v)
    scale = np.abs(exact[:, 1:]).max()
    assert np.abs(got - exact).max() < 1e-8 * scale
    # a transposed cross channel (xy <-> yx) differs at the percent level
    bad = got.copy()
    if kind == "JeJe":
        bad[:, [3, 4]] = bad[:, [4, 3]]
    else:
        bad[:, 8] = -bad[:, 8]          # Im xy with the wrong sign of i in the polarisation
    assert np.abs(bad - exact).max() > 1e-3 * scale


def _ref_run(B, steps=100, K=1.0):
    m = O.Model(3, 4, K=(K, K, K), B=B)
    coo = O.CooHamiltonian(m)
    res = O.lanczos_ref(coo.apply, O.start_vector_ref(3, m.dim), steps)
    return m, res, O.ritz_vectors(res)


def _l1(tab, ref):
    return np.abs(tab[:, 1:] - ref[:, 1:]).sum(0) / np.abs(ref[:, 1:]).sum(0)


def test_golden_jsjs_reproduced():
    """JsJs_outSpectrum.real = Midinput.inp with B = 0 (2e-4 L1 per column)."""
    ref = O.read_real(os.path.join(GOLDEN, "JsJs_outSpectrum.real"))
    m, res, st = _ref_run((0.0, 0.0, 0.0))
    tab = O.jsjs_ref(m, res.evals, st, res.gs)
    # the grid omega = 0.002 * k is printed with 6 significant digits
    assert [O.format_g6(w) for w in tab[:, 0]] == [O.format_g6(w) for w in ref[:, 0]]
    assert _l1(tab, ref).max() < 5e-3


def test_golden_jeje_reproduced():
    """JeJe_outSpectrum.real = Midinput.inp as committed (B = 0.3); ~2 % L1 is the
    rounding-chaos floor of a 100-step reorthogonalised run."""
    ref = O.read_real(os.path.join(GOLDEN, "JeJe_outSpectrum.real"))
    m, res, st = _ref_run((0.3, 0.3, 0.3))
    tab = O.jeje_ref(m, res.evals, st, res.gs)
    assert _l1(tab, ref).max() < 0.06
    w = np.abs(tab[:, 1:].sum(0) / ref[:, 1:].sum(0) - 1)
    assert w.max() < 0.02


@pytest.mark.parametrize("mistake", ["steps", "ky_sign", "seed"])
def test_golden_rejects_planted_mistakes(mistake):
    """The golden comparison is sensitive: a wrong step count, a flipped Kyy sign
    or another seed must fail it."""
    ref = O.read_real(os.path.join(GOLDEN, "JsJs_outSpectrum.real"))
    if mistake == "steps":
        m, res, st = _ref_run((0.0, 0.0, 0.0), steps=101)
    else:
        m = O.Model(3, 4, K=(1.0, -1.0 if mistake == "ky_sign" else 1.0, 1.0))
        coo = O.CooHamiltonian(m)
        res = O.lanczos_ref(coo.apply, O.start_vector_ref(4 if mistake == "seed" else 3, m.dim), 100)
        st = O.ritz_vectors(res)
    tab = O.jsjs_ref(m, res.evals, st, res.gs)
    assert _l1(tab, ref).max() > 0.05


def test_real_file_format_roundtrip(tmp_path):
    ref_path = os.path.join(GOLDEN, "JeJe_outSpectrum.real")
    tab = O.read_real(ref_path)
    out = tmp_path / "JeJe_outSpectrum.real"
    O.write_real(str(out), tab)
    assert out.read_bytes() == open(ref_path, "rb").read()


@pytest.mark.parametrize("kw", [dict(Lx=2, Ly=6, K=(-1.0, 0.5, 0.7), J=0.1, G=0.2, B=(0.1, 0.2, 0.3)),
                                dict(Lx=2, Ly=8, K=(1.0, 1.0, 1.0), B=(0.3, 0.3, 0.3))])
def test_sampled_rows_equal_full_gather(kw):
    """orc_hpsi_rows (rows of H.v from a sparse column list, used to check
    systems too large to hold) reproduces orc_hpsi_gather row for row,
    including repeated rows and rows 0 and dim-1."""
    m = O.Model(**kw)
    rng = np.random.default_rng(m.N)
    v = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    full = O.hpsi_gather(m, v)
    rows = np.concatenate([[0, m.dim - 1], rng.integers(0, m.dim, 64), [5, 5]])
    cols = O.row_columns(m, rows)
    assert np.array_equal(O.hpsi_rows(m, rows, cols, v[cols]), full[rows])


def test_sparse_ground_pair_equals_dense():
    """The eigsh-based ground state (used to check N = 12 runs) equals the
    dense eigensolve of the Kronecker-built H at N = 8 and 10."""
    for m in (O.Model(2, 4, K=(-1.0, -1.0, -1.0), J=0.1, G=0.2, B=O.field111(0.2)),
              O.Model(1, 10, K=(1.0, 0.8, 1.2), B=(0.3, 0.1, 0.2))):
        H = O.hamiltonian_dense_kron(m)
        w = np.linalg.eigvalsh(H)
        e0, g = O.ground_pair(m)
        assert abs(e0 - w[0]) < 1e-11
        assert np.linalg.norm(H @ g - e0 * g) < 1e-8


def test_golden_local_szsz_reproduced():
    """2SSLocal_outSpectrum.real (LocalDynSzSz, src/dynamics.h:64-217) =
    Midinput.inp with B = 0, the same run as JsJs (measured 3e-4 L1 per
    column); the committed field B = 0.3 does not reproduce it."""
    ref = O.read_real(os.path.join(GOLDEN, "2SSLocal_outSpectrum.real"))
    m, res, st = _ref_run((0.0, 0.0, 0.0))
    tab = O.local_szsz_ref(m, res.evals, st, res.gs)
    assert [O.format_g6(w) for w in tab[:, 0]] == [O.format_g6(w) for w in ref[:, 0]]
    assert _l1(tab, ref).max() < 2e-3
    m3, res3, st3 = _ref_run((0.3, 0.3, 0.3))
    assert _l1(O.local_szsz_ref(m3, res3.evals, st3, res3.gs), ref).max() > 0.5


def test_local_szsz_lehmann_equals_resolvent():
    """With all eigenstates, the Lehmann sums of local_szsz_ref equal the
    resolvent forms sum_i <A_i gs|(w + E0 - i eta - H)^-1|B_i gs> minus the
    elastic terms (dense solves at N = 8), for xx, xy and the rotated channel."""
    m = O.Model(2, 4, K=(-1.0, -1.0, -1.0), J=0.1, G=0.2, B=O.field111(0.2))
    H = O.hamiltonian_dense_kron(m)
    e, U = np.linalg.eigh(H)
    gs = U[:, 0]
    w = np.linspace(0.0, 3.0, 7)
    eta = 0.1
    tab = O.local_szsz_ref(m, e, U.T.copy(), gs, omega=w, eta=eta)
    xx, xy, pp = np.zeros(len(w), complex), np.zeros(len(w), complex), np.zeros(len(w), complex)
    for i in range(m.N):
        sx, sy, sz = (O.apply_site_op(m.N, n, i, gs) for n in ("Sx", "Sy", "Sz"))
        sp = O.SPE3[0] * sx + O.SPE3[1] * sy + O.SPE3[2] * sz
        sm = O.SME3[0] * sx + O.SME3[1] * sy + O.SME3[2] * sz
        for q, om in enumerate(w):
            R = np.linalg.inv((om + e[0] - 1j * eta) * np.eye(m.dim) - H)
            dn1 = om - 1j * eta
            xx[q] += np.vdot(sx, R @ sx) - abs(np.vdot(gs, sx)) ** 2 / dn1
            xy[q] += np.vdot(sx, R @ sy) - np.conj(np.vdot(gs, sx)) * np.vdot(gs, sy) / dn1
            pp[q] += np.vdot(sp, R @ sp) - np.vdot(gs, sm) * np.vdot(gs, sp) / dn1
    assert np.abs(tab[:, 1] - xx.imag).max() < 1e-10
    assert np.abs(tab[:, 4] - xy.imag).max() < 1e-10
    assert np.abs(tab[:, 7] - pp.imag).max() < 1e-10


def test_two_point_ref_against_dense_operators():
    """TwoPointCorr restatement: <gs|S^a_i S^b_j|gs> from dense Kronecker site
    operators (i >= j, mirrored), and Sz.Sz on the diagonal = 1/4."""
    m = O.Model(2, 4, K=(-1.0, -1.0, -1.0), J=0.1, G=0.2, B=O.field111(0.2))
    H = O.hamiltonian_dense_kron(m)
    gs = np.linalg.eigh(H)[1][:, 0]
    C = O.two_point_ref(m, "Sx", "Sz", gs)
    for i in range(m.N):
        for j in range(i + 1):
            d = np.vdot(gs, O.site_op_dense(m.N, "Sx", i) @ (O.site_op_dense(m.N, "Sz", j) @ gs))
            assert abs(C[i, j] - d) < 1e-13 and C[j, i] == C[i, j]
    Czz = O.two_point_ref(m, "Sz", "Sz", gs)
    assert np.allclose(np.diag(Czz), 0.25, atol=1e-13)
