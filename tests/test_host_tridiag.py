"""The FP64 routines on a recursion's tridiagonal matrix and its continued
fraction that the batched solves run on the device (esc_tridiag.h), here on
the host through esc_tridiag_lowest / esc_cf_eval.

Pins: the lowest eigenpair against LAPACK (numpy eigh) on random, graded,
split-free and tiny matrices; on the tridiagonal matrix of a plain 8-site
Lanczos run long enough to repeat its lowest eigenvalue (the copies that a
recursion without reorthogonalisation produces), the Ritz vector built from
the returned eigenvector against the exact ground state; the continued
fraction against the oracle's complex-division restatement and against the
exact resolvent <phi|(z - H)^-1|phi> of a dense matrix.
"""
import numpy as np
import pytest

import oracle as O

esc = pytest.importorskip("energyspincurrents_amd")


def tmat(a, b):
    n = len(a)
    return np.diag(a) + np.diag(b[1:n], 1) + np.diag(b[1:n], -1)


@pytest.mark.parametrize("n,scale", [(1, 1.0), (2, 1.0), (3, 1e-3), (10, 1.0), (200, 1.0), (1000, 1e3)])
def test_lowest_pair_matches_lapack(n, scale):
    rng = np.random.default_rng(n)
    a = scale * rng.standard_normal(n)
    b = np.r_[0.0, scale * (0.1 + np.abs(rng.standard_normal(n - 1)))]
    T = tmat(a, b)
    w, V = np.linalg.eigh(T)
    e0, z = esc.tridiag_lowest(a, b)
    tn = np.abs(T).sum(axis=1).max()
    assert abs(e0 - w[0]) <= 8 * np.finfo(float).eps * tn
    assert abs(np.linalg.norm(z) - 1.0) < 1e-14
    assert np.linalg.norm(T @ z - e0 * z) <= 1e-13 * tn
    if n > 1 and w[1] - w[0] > 1e-6 * tn:
        assert abs(abs(z @ V[:, 0]) - 1.0) < 1e-12
    nz = np.flatnonzero(z)
    assert z[nz[0]] > 0.0


def test_lowest_pair_rejects_a_zero_coupling_and_non_finite_entries():
    with pytest.raises(esc.EscError):
        esc.tridiag_lowest([1.0, 2.0, 3.0], [0.0, 1.0, 0.0])
    with pytest.raises(esc.EscError):
        esc.tridiag_lowest([1.0, np.inf], [0.0, 1.0])


def test_planted_sign_mistake_in_the_off_diagonal_is_invisible_but_a_wrong_index_is_not():
    """T is similar to the matrix with |b|, so the routine must give the same
    e0 for -b; a shifted coupling (b[k] placed at k + 1) changes e0."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal(30)
    b = np.r_[0.0, 0.2 + np.abs(rng.standard_normal(29))]
    e0, _ = esc.tridiag_lowest(a, b)
    assert e0 == esc.tridiag_lowest(a, np.r_[0.0, -b[1:]])[0]
    shifted = np.r_[0.0, 0.5, b[1:-1]]
    assert abs(esc.tridiag_lowest(a, shifted)[0] - e0) > 1e-6


def test_ritz_vector_of_a_plain_recursion_with_repeated_lowest_eigenvalue():
    """8-site Kitaev-field cluster, 300 plain Lanczos steps: T holds several
    copies of E0.  The returned eigenvector, used as replay coefficients,
    gives a Ritz vector as close to the exact ground state as the one from
    LAPACK's eigenvector."""
    m = O.Model(2, 4, K=(-1.0, -1.0, -1.0), B=O.field111(0.3))
    H = O.hamiltonian_dense_kron(m)
    ev, U = np.linalg.eigh(H)
    gs = U[:, 0]
    rng = np.random.default_rng(12)
    v0 = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    al, be, _ = O.lanczos_plain(lambda x: H @ x, v0, 300)
    n = len(al)
    a, b = al, be[:n]
    w = np.linalg.eigh(tmat(a, b))[0]
    copies = int(np.sum(np.abs(w - ev[0]) < 1e-8))
    assert copies >= 2, "the run should repeat E0"
    e0, z = esc.tridiag_lowest(a, b)
    assert abs(e0 - ev[0]) < 1e-12 * abs(ev[0])
    # replay the recursion, accumulating sum_k z_k v_k
    v = v0 / np.linalg.norm(v0)
    vp = np.zeros_like(v)
    y = np.zeros_like(v)
    for k in range(n):
        y += z[k] * v
        if k + 1 == n:
            break
        wv = H @ v - b[k] * vp - a[k] * v
        vp, v = v, wv / be[k + 1]
    y /= np.linalg.norm(y)
    assert abs(abs(np.vdot(gs, y)) - 1.0) < 1e-9
    assert np.linalg.norm(H @ y - e0 * y) < 1e-5


@pytest.mark.parametrize("n", [1, 5, 120])
def test_continued_fraction_matches_oracle_restatement(n):
    rng = np.random.default_rng(n + 40)
    a = rng.standard_normal(n)
    b = np.r_[0.0, 0.1 + np.abs(rng.standard_normal(n - 1))] if n > 1 else np.zeros(1)
    om = np.linspace(-1.0, 6.0, 301)
    got = esc.cf_eval(a, b, 1.7, 0.3, 0.05, om)
    ref = O.cf_resolvent(a, b, 1.7, om + 0.3 - 0.05j)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.all(esc.cf_eval(a, b, 0.0, 0.3, 0.05, om) == 0)


def test_continued_fraction_equals_dense_resolvent():
    """A full-length recursion of a 60 x 60 Hermitian matrix from phi
    reproduces <phi|(z - H)^-1|phi> exactly (up to rounding)."""
    rng = np.random.default_rng(60)
    X = rng.standard_normal((60, 60)) + 1j * rng.standard_normal((60, 60))
    H = (X + X.conj().T) / 2
    phi = rng.standard_normal(60) + 1j * rng.standard_normal(60)
    n2 = np.vdot(phi, phi).real
    r = O.lanczos_ref(lambda x: H @ x, phi / np.sqrt(n2), 60, reortho=True)
    al, be = np.real(r.alpha), np.abs(r.beta)
    om = np.linspace(-5.0, 5.0, 41)
    eta = 0.3
    got = esc.cf_eval(al, be[:len(al)], n2, 0.0, eta, om)
    ref = np.array([np.vdot(phi, np.linalg.solve((w - 1j * eta) * np.eye(60) - H, phi)) for w in om])
    assert np.abs(got - ref).max() < 1e-8 * np.abs(ref).max()


def check_eigenpairs(a, b, ev, Z, tol_res=1e-13, tol_orth=1e-12):
    T = tmat(a, b)
    w = np.linalg.eigvalsh(T)
    tn = np.abs(T).sum(axis=1).max()
    assert np.abs(ev - w).max() <= 16 * np.finfo(float).eps * tn
    assert np.linalg.norm(T @ Z - Z * ev, axis=0).max() <= tol_res * tn
    assert np.abs(Z.T @ Z - np.eye(len(a))).max() <= tol_orth


@pytest.mark.parametrize("n,scale", [(1, 1.0), (2, 1.0), (50, 1.0), (300, 1e-2)])
def test_all_eigenpairs_match_lapack(n, scale):
    rng = np.random.default_rng(100 + n)
    a = scale * rng.standard_normal(n)
    b = np.r_[0.0, scale * (0.05 + np.abs(rng.standard_normal(n - 1)))]
    ev, Z = esc.tridiag_eigenpairs(a, b)
    check_eigenpairs(a, b, ev, Z)
    assert np.all(Z[np.argmax(Z != 0, axis=0), np.arange(n)] > 0)


def test_all_eigenpairs_of_wilkinson_matrix_with_close_pairs():
    """W21+: diagonal |10 - k|, off-diagonal 1; its largest eigenvalues come in
    pairs that agree to 1e-14, so the clusters' vectors must be
    orthogonalised against each other."""
    a = np.abs(10.0 - np.arange(21.0))
    b = np.r_[0.0, np.ones(20)]
    ev, Z = esc.tridiag_eigenpairs(a, b)
    assert ev[-1] - ev[-2] < 1e-12
    check_eigenpairs(a, b, ev, Z)


def test_all_eigenpairs_of_a_plain_recursion_with_repeated_eigenvalues():
    m = O.Model(2, 4, K=(-1.0, -1.0, -1.0), B=O.field111(0.3))
    H = O.hamiltonian_dense_kron(m)
    rng = np.random.default_rng(7)
    al, be, _ = O.lanczos_plain(lambda x: H @ x, rng.standard_normal(m.dim) + 0j, 300)
    n = len(al)
    ev, Z = esc.tridiag_eigenpairs(al, be[:n])
    check_eigenpairs(al, be[:n], ev, Z)
