"""Oracle restatements of LocalDynRIXS (src/dynamics.h:465-562) and DynSzSz
(src/dynamics.h:566-611), pinned on the CPU: fed with the exact eigenpairs of
H and of the intermediate Hamiltonian (src/Hamiltonian.h:57-91), the
reference's sums over states equal dense resolvent formulas built from linear
solves, an independent route:
  RIXS   I_X(w) = Im <phi|(w + E0 - i eta - H)^-1|phi> - |<gs|phi>|^2 eta/(w^2 + eta^2),
         phi = X (-H_int - i Gamma)^-1 |gs>
  SzSz   real + i imag = (G_ij + G_ji)/2, G_ij = <Sz_i gs|(w + E0 - i eta - H)^-1|Sz_j gs>."""
import numpy as np
import pytest

import oracle as O

MODELS = [O.Model(3, 2, K=(-1.0, -0.8, -1.2), B=(0.1, 0.2, 0.3)),
          O.Model(2, 4, K=(-1.0, -1.0, -1.0), J=0.1, G=0.2, B=O.field111(0.2))]


def _dense_rixs(m, site, gs, E0, omega, eta=0.1, gamma=100.0):
    H = O.hamiltonian_dense_kron(m)
    Hi = O.hamiltonian_dense_kron(O.kill_site(m, site))
    chi = np.linalg.solve(-Hi - 1j * gamma * np.eye(m.dim), gs)
    cols = []
    for X in (None, "Sx", "Sy", "Sz"):
        phi = chi if X is None else O.apply_site_op(m.N, X, site, chi)
        el = abs(np.vdot(gs, phi)) ** 2
        I = []
        for w in omega:
            r = np.linalg.solve((w + E0 - 1j * eta) * np.eye(m.dim) - H, phi)
            I.append(np.vdot(phi, r).imag - el * eta / (w * w + eta * eta))
        cols.append(np.array(I))
    return np.column_stack(cols)


@pytest.mark.parametrize("mi", [0, 1])
@pytest.mark.parametrize("site", [0, 3])
def test_rixs_ref_equals_dense_resolvent(mi, site):
    m = MODELS[mi]
    e, U = np.linalg.eigh(O.hamiltonian_dense_kron(m))
    ei, Ui = np.linalg.eigh(O.hamiltonian_dense_kron(O.kill_site(m, site)))
    gs = U[:, 0]
    w = O.TENTH_OMEGA[::6]
    tab = O.rixs_ref(m.N, site, e, U.T.copy(), ei, Ui.T.copy(), gs, omega=w)
    ref = _dense_rixs(m, site, gs, e[0], w)
    assert np.abs(tab[:, 1:5] - ref).max() < 1e-10 * np.abs(ref).max()
    assert np.all(tab[:, 5:] == 0.0)
    # the bonds of the site are gone, the others stay
    k = O.kill_site(m, site)
    assert all(site not in b for fam in k.bonds for b in fam.tolist())
    assert sum(len(f) for f in k.bonds) == sum(len(f) for f in m.bonds) - 3


def test_rixs_elastic_term_is_not_negligible():
    # both sides leave out the final state m = 0; the comparison above pins
    # that only if its term is large next to what remains
    m = MODELS[0]
    e, U = np.linalg.eigh(O.hamiltonian_dense_kron(m))
    ei, Ui = np.linalg.eigh(O.hamiltonian_dense_kron(O.kill_site(m, 2)))
    gs = U[:, 0]
    tab = O.rixs_ref(m.N, 2, e, U.T.copy(), ei, Ui.T.copy(), gs, omega=[0.0])
    chi = Ui @ ((Ui.conj().T @ gs) * O.inv_complex(0.0 - ei, 100.0))
    # identity channel: <gs|chi> is nearly |chi| (Gamma = 100 makes chi ~ i gs / 100)
    assert abs(np.vdot(gs, chi)) ** 2 / 0.1 > 10 * tab[0, 1]


@pytest.mark.parametrize("mi", [0, 1])
def test_szsz_ref_equals_dense_resolvent(mi):
    m = MODELS[mi]
    H = O.hamiltonian_dense_kron(m)
    e, U = np.linalg.eigh(H)
    gs = U[:, 0]
    w = O.TENTH_OMEGA[::10]
    S = O.szsz_ref(m.N, e, U.T.copy(), gs, omega=w)
    Z = np.array([O.apply_site_op(m.N, "Sz", i, gs) for i in range(m.N)])
    for q, wq in enumerate(w):
        R = np.linalg.solve((wq + e[0] - 0.1j) * np.eye(m.dim) - H, Z.T)   # columns (z - H)^-1 Sz_j gs
        G = Z.conj() @ R                                                   # G[i, j]
        ref = 0.5 * (G + G.T)
        assert np.abs(S[q] - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())
