"""CPU oracle for the Lanczos dynamical-correlation path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  The product
library never calls it.

It wraps ``libesc_oracle.so`` (plain C restatement, ``esc_oracle.c``) and adds the
numpy parts of the reference algorithm.  Each function names the reference
file:line it restates (paths relative to the fengshi96/EnergySpinCurrents root).

Parity pins (checked in ``tests/test_oracle_*.py``):
  honeycomb / bonds       closed form + hand-derived 12-site table
  chain / chain bonds     hand-derived 4-site tables; XX chain (free fermions)
                          and transverse-field Ising closed forms, ring E0s
  hamiltonian (COO, gather) dense Kronecker product of Pauli matrices, Tr H,
                          Tr H^2 closed form, Hermiticity, W_p commutation
  currents                dense Kronecker operators
  lanczos_ref             E0 against full diagonalisation; Krylov orthonormality;
                          committed JeJe_outSpectrum.real of the reference
  tridiag_eig             numpy/LAPACK eigh of the dense tridiagonal
  lehmann / cf spectra    exact Lehmann sum from full ED; sum rules
  mt19937_64              the C++ standard's 10000th-output requirement
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    """Load the compiled C restatement (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libesc_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i32, f64, i64, u64 = ctypes.c_int, ctypes.c_double, ctypes.c_int64, ctypes.c_uint64
        L.orc_honeycomb1.argtypes = [i32, i32, i32, i32, P, P, P]
        L.orc_connectors.argtypes = [i32, P, P, P, P, P, P]
        L.orc_coo_build.argtypes = [i32, P, i32, P, i32, P, i32, P, P]
        L.orc_coo_build.restype = P
        L.orc_coo_free.argtypes = [P]
        L.orc_coo_nnz.argtypes = [P]
        L.orc_coo_nnz.restype = i64
        L.orc_coo_export.argtypes = [P, P, P, P, P]
        L.orc_coo_apply.argtypes = [P, P, P, i32]
        L.orc_hpsi_gather.argtypes = [i32, P, i32, P, i32, P, i32, P, f64, f64, P, i32, P, P]
        L.orc_hpsi_rows.argtypes = [i32, P, i32, P, i32, P, i32, P, f64, f64, P, i32, i64, P, i64, P, P, P]
        L.orc_apply_site_op.argtypes = [i32, i32, i32, P, P]
        L.orc_current_js_site.argtypes = [i32, P, P, i32, i32, P, P]
        L.orc_current_je_site.argtypes = [i32, P, P, i32, i32, P, P]
        L.orc_current_total.argtypes = [i32, P, P, i32, P, P]
        L.orc_mt19937_64.argtypes = [u64, i64, P]
        L.orc_randomize.argtypes = [u64, i64, P]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# Geometry and model
# --------------------------------------------------------------------------
def honeycomb1(Lx, Ly, perX=True, perY=True):
    """Neighbour table (N x 3; col 0 = z, 1 = y, 2 = x) -- src/Lattice.h:434-576."""
    N = Lx * Ly
    neigh = np.empty((N, 3), dtype=np.int32)
    ix = np.zeros(N, dtype=np.int32)
    iy = np.zeros(N, dtype=np.int32)
    rc = lib().orc_honeycomb1(Lx, Ly, int(perX), int(perY), _p(neigh), _p(ix), _p(iy))
    if rc:
        raise ValueError("Honeycomb1 needs an even Ly >= 2 and Lx >= 1")
    return neigh, ix, iy


def connectors(neigh, K):
    """Bond families x, y, z as lists of (i<j) pairs -- src/Hamiltonian.h:230-328."""
    N = neigh.shape[0]
    neigh = np.ascontiguousarray(neigh, dtype=np.int32)
    Kd = np.asarray(K, dtype=np.float64)
    px = np.zeros((N, 2), np.int32)
    py = np.zeros((N, 2), np.int32)
    pz = np.zeros((N, 2), np.int32)
    cnt = np.zeros(3, np.int32)
    lib().orc_connectors(N, _p(neigh), _p(Kd), _p(px), _p(py), _p(pz), _p(cnt))
    return [px[: cnt[0]].copy(), py[: cnt[1]].copy(), pz[: cnt[2]].copy()]


@dataclass
class Model:
    """Honeycomb model: Kitaev K (per family), Heisenberg J, Gamma G, field B."""

    Lx: int
    Ly: int
    K: tuple = (1.0, 1.0, 1.0)
    J: float = 0.0
    G: float = 0.0
    B: tuple = (0.0, 0.0, 0.0)
    perX: bool = True
    perY: bool = True
    neigh: np.ndarray = field(init=False, repr=False)
    bonds: list = field(init=False, repr=False)

    def __post_init__(self):
        self.neigh, _, _ = honeycomb1(self.Lx, self.Ly, self.perX, self.perY)
        # Families exist whenever the lattice has the bond; the coupling only
        # decides their weight (a zero-K family adds exact zeros).
        self.bonds = connectors(self.neigh, (1.0, 1.0, 1.0))
        self.K = tuple(float(k) for k in self.K)
        self.B = tuple(float(b) for b in self.B)

    def families(self, K):
        return connectors(self.neigh, K)

    @property
    def N(self):
        return self.Lx * self.Ly

    @property
    def dim(self):
        return 1 << self.N


def chain(N, periodic=True):
    """Neighbour table (N x 2; col 0 = +x, col 1 = -x, -1 if absent) --
    src/Lattice.h:60-125."""
    neigh = np.full((N, 2), -1, np.int32)
    for i in range(N):
        if i + 1 < N:
            neigh[i, 0] = i + 1
        elif periodic:
            neigh[i, 0] = 0
        if i - 1 >= 0:
            neigh[i, 1] = i - 1
        elif periodic:
            neigh[i, 1] = N - 1
    return neigh


def chain_connectors(neigh, K):
    """Bond families x, y, z of the chain: each family of a non-zero coupling
    holds every bond i < j, in row-major order of the coupling matrix --
    src/Hamiltonian.h:381-475."""
    N = neigh.shape[0]
    coupled = np.zeros((N, N), bool)
    for i in range(N):
        for c in (0, 1):
            j = int(neigh[i, c])
            if i < j:
                coupled[i, j] = True
    pairs = np.argwhere(coupled).astype(np.int32)      # row-major (i, j) order
    return [pairs.copy() if K[f] != 0.0 else np.zeros((0, 2), np.int32) for f in range(3)]


@dataclass
class ChainModel:
    """HeisenbergChain model: H = sum_bonds sum_f K_f S^f_i S^f_j + B.S."""

    N: int
    K: tuple = (1.0, 1.0, 1.0)
    B: tuple = (0.0, 0.0, 0.0)
    periodic: bool = True
    J: float = 0.0
    G: float = 0.0
    neigh: np.ndarray = field(init=False, repr=False)
    bonds: list = field(init=False, repr=False)

    def __post_init__(self):
        self.K = tuple(float(k) for k in self.K)
        self.B = tuple(float(b) for b in self.B)
        self.neigh = chain(self.N, self.periodic)
        self.bonds = chain_connectors(self.neigh, self.K)

    def families(self, K):
        return chain_connectors(self.neigh, K)

    @property
    def dim(self):
        return 1 << self.N


def field111(h):
    """[111] field of magnitude h as (Bxx, Byy, Bzz)."""
    b = h / np.sqrt(3.0)
    return (b, b, b)


# --------------------------------------------------------------------------
# Hamiltonian application
# --------------------------------------------------------------------------
def hpsi_gather(m: Model, v):
    """Matrix-free H.v in FP64 (gather form of src/Hamiltonian.h:94-377,480-557,
    extended by J and Gamma)."""
    v = np.ascontiguousarray(v, dtype=np.complex128)
    out = np.empty_like(v)
    px, py, pz = (np.ascontiguousarray(b, np.int32) for b in m.bonds)
    K = np.asarray(m.K, np.float64)
    B = np.asarray(m.B, np.float64)
    lib().orc_hpsi_gather(m.N, _p(px), len(px), _p(py), len(py), _p(pz), len(pz), _p(K),
                          float(m.J), float(m.G), _p(B), 1, _p(v), _p(out))
    return out


def row_columns(m: Model, rows):
    """Sorted columns that rows of H.v read (the rows, their bond and single-
    flip partners): the values orc_hpsi_rows needs."""
    rows = np.asarray(rows, np.int64)
    masks = [0] + [(1 << int(i)) | (1 << int(j)) for fam in m.bonds for i, j in fam] + [1 << k for k in range(m.N)]
    cols = rows[:, None] ^ np.asarray(masks, np.int64)[None, :]
    return np.unique(cols.ravel())


def hpsi_rows(m: Model, rows, cols, vals):
    """Rows of H.v (same elements as hpsi_gather) for a system too large to
    hold: cols = row_columns(m, rows), vals = v[cols]."""
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int64)
    vals = np.ascontiguousarray(vals, np.complex128)
    out = np.empty(len(rows), np.complex128)
    px, py, pz = (np.ascontiguousarray(b, np.int32) for b in m.bonds)
    K = np.asarray(m.K, np.float64)
    B = np.asarray(m.B, np.float64)
    lib().orc_hpsi_rows(m.N, _p(px), len(px), _p(py), len(py), _p(pz), len(pz), _p(K), float(m.J), float(m.G),
                        _p(B), 1, len(rows), _p(rows), len(cols), _p(cols), _p(vals), _p(out))
    return out


def ground_pair(m: Model, tol=1e-13):
    """Lowest eigenpair of the gather-form H by implicitly restarted Lanczos
    (ARPACK through scipy.sparse.linalg.eigsh) on a LinearOperator over
    hpsi_gather: the exact-diagonalisation check for N where a dense
    eigensolve is slow (N = 12: dim 4096)."""
    from scipy.sparse.linalg import LinearOperator, eigsh
    op = LinearOperator((m.dim, m.dim), matvec=lambda v: hpsi_gather(m, v.ravel()), dtype=np.complex128)
    rng = np.random.default_rng(12345)          # a generic start: no symmetry sector is missed
    v0 = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    w, U = eigsh(op, k=1, which="SA", tol=tol, v0=v0)
    return float(w[0]), U[:, 0]


class CooHamiltonian:
    """Faithful COO restatement: TightB_Ham + Diagonal + MatrixVectorMult
    (src/Hamiltonian.h:94-206,331-377,480-557).  Reference model only (no J, G)."""

    def __init__(self, m: Model):
        assert m.J == 0.0 and m.G == 0.0
        self.m = m
        # the reference builds families only for non-zero couplings (:269-285)
        self.bonds = m.families(m.K)
        px, py, pz = (np.ascontiguousarray(b, np.int32) for b in self.bonds)
        self._keep = (px, py, pz)
        K = np.asarray(m.K, np.float64)
        B = np.asarray(m.B, np.float64)
        self.h = lib().orc_coo_build(m.N, _p(px), len(px), _p(py), len(py), _p(pz), len(pz), _p(K), _p(B))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_coo_free(self.h)
            self.h = None

    def apply(self, v, nthreads=4):
        v = np.ascontiguousarray(v, dtype=np.complex128)
        out = np.empty_like(v)
        lib().orc_coo_apply(self.h, _p(v), _p(out), nthreads)
        return out

    def dense(self):
        """Dense H as the ED path assembles it (src/Exactdiag.h:45-69), with the
        matrix zero-initialised (the reference leaves it uninitialised, :42)."""
        nnz = lib().orc_coo_nnz(self.h)
        dim = self.m.dim
        rows = np.empty(nnz, np.uint32)
        cols = np.empty(nnz, np.uint32)
        vals = np.empty(nnz, np.complex128)
        diag = np.empty(dim, np.complex128)
        lib().orc_coo_export(self.h, _p(rows), _p(cols), _p(vals), _p(diag))
        H = np.zeros((dim, dim), np.complex128)
        H[np.arange(dim), np.arange(dim)] += diag
        np.add.at(H, (rows.astype(np.int64), cols.astype(np.int64)), vals)
        return H


# --------------------------------------------------------------------------
# Dense Kronecker builder (independent formulation, small N)
# --------------------------------------------------------------------------
_SX = np.array([[0, 1], [1, 0]], np.complex128) / 2
_SY = np.array([[0, -1j], [1j, 0]], np.complex128) / 2
_SZ = np.array([[1, 0], [0, -1]], np.complex128) / 2
SPIN = {"Sx": _SX, "Sy": _SY, "Sz": _SZ}


def site_op_dense(N, name, site):
    """S^a_site as a dense 2^N matrix; basis index bit k = site k, bit 0 = up."""
    # bit k of r is the k-th least-significant bit -> site k is the (N-1-k)-th
    # tensor factor in numpy's kron ordering.
    op = np.array([[1.0]], np.complex128)
    for k in range(N - 1, -1, -1):
        op = np.kron(op, SPIN[name] if k == site else np.eye(2))
    return op


def hamiltonian_dense_kron(m: Model):
    """H = sum_bonds [K S^g S^g + J S.S + G (S^a S^b + S^b S^a)] + B.S, dense."""
    N = m.N
    S = {a: [site_op_dense(N, a, k) for k in range(N)] for a in ("Sx", "Sy", "Sz")}
    axes = ("Sx", "Sy", "Sz")
    H = np.zeros((m.dim, m.dim), np.complex128)
    for f, pairs in enumerate(m.bonds):
        g = axes[f]
        a, b = [ax for ax in axes if ax != g]
        for i, j in pairs:
            H += m.K[f] * S[g][i] @ S[g][j]
            if m.J:
                H += m.J * sum(S[ax][i] @ S[ax][j] for ax in axes)
            if m.G:
                H += m.G * (S[a][i] @ S[b][j] + S[b][i] @ S[a][j])
    for k in range(N):
        for c, ax in enumerate(axes):
            H += m.B[c] * S[ax][k]
    return H


# --------------------------------------------------------------------------
# Site operators and currents (C restatement of src/Observables.h)
# --------------------------------------------------------------------------
_OPS = {"Sx": 0, "Sy": 1, "Sz": 2}
CURRENTS = {"Jsx": 0, "Jsy": 1, "Jsz": 2, "Jex": 3, "Jey": 4}


def apply_site_op(N, name, site, v):
    v = np.ascontiguousarray(v, np.complex128)
    out = np.empty_like(v)
    lib().orc_apply_site_op(N, _OPS[name], site, _p(v), _p(out))
    return out


def apply_current_site(m: Model, name, site, v):
    v = np.ascontiguousarray(v, np.complex128)
    out = np.empty_like(v)
    K = np.asarray(m.K, np.float64)
    nb = np.ascontiguousarray(m.neigh, np.int32)
    if name.startswith("Js"):
        rc = lib().orc_current_js_site(m.N, _p(nb), _p(K), CURRENTS[name], site, _p(v), _p(out))
    else:
        rc = lib().orc_current_je_site(m.N, _p(nb), _p(K), CURRENTS[name] - 3, site, _p(v), _p(out))
    if rc:
        raise ValueError("missing neighbour for current operator")
    return out


def apply_current_total(m: Model, name, v):
    """ApplyJ{sx,sy,sz}Total / ApplyJ{ex,ey}Total (src/Observables.h:448-470,664-699)."""
    v = np.ascontiguousarray(v, np.complex128)
    out = np.empty_like(v)
    K = np.asarray(m.K, np.float64)
    nb = np.ascontiguousarray(m.neigh, np.int32)
    rc = lib().orc_current_total(m.N, _p(nb), _p(K), CURRENTS[name], _p(v), _p(out))
    if rc:
        raise ValueError("missing neighbour for current operator")
    return out


# --------------------------------------------------------------------------
# Start vector (src/DynLanczosSolver.h:58-59,465-474)
# --------------------------------------------------------------------------
def mt19937_64(seed, n):
    out = np.empty(n, np.uint64)
    lib().orc_mt19937_64(seed, n, _p(out))
    return out


def start_vector_ref(seed, dim, normalize=True):
    v = np.empty(dim, np.complex128)
    lib().orc_randomize(seed, dim, _p(v))
    if normalize:
        v = v / np.linalg.norm(v)   # Eigen normalize(): v / ||v||_2
    return v


# --------------------------------------------------------------------------
# Lanczos: restatement of DynLanczosSolver::Lanczos_Nirav
# (src/DynLanczosSolver.h:38-239) and TriDiag (:371-414)
# --------------------------------------------------------------------------
def tridiag_eig(alpha, beta, vectors=True):
    """Eigen-decomposition of the Hermitian tridiagonal T with T[i,i] = alpha[i],
    T[i,i+1] = beta[i+1] (beta[0] unused), as zheev sees it (upper triangle,
    diagonal imaginary parts ignored).  Returns ascending eigenvalues and the
    eigenvector matrix Z (columns)."""
    a = np.real(np.asarray(alpha, np.complex128))
    b = np.asarray(beta, np.complex128)
    n = len(a)
    T = np.diag(a.astype(np.complex128))
    for i in range(n - 1):
        T[i, i + 1] = b[i + 1]
        T[i + 1, i] = np.conj(b[i + 1])
    if vectors:
        w, Z = np.linalg.eigh(T, UPLO="U")
        return w, Z
    return np.linalg.eigvalsh(T, UPLO="U"), None


@dataclass
class LanczosResult:
    alpha: np.ndarray
    beta: np.ndarray
    evals: np.ndarray       # TriDeval (ascending)
    Z: np.ndarray           # tridiagonal eigenvectors, column k = state k
    basis: np.ndarray       # (EXITiter+1, dim) Lanczos vectors
    gs: np.ndarray          # PsiAll[0]
    iters: int              # EXITiter
    energies: list          # D0 per iteration (stdout trace)


def lanczos_ref(apply_h, v0, steps, eps=0.0, reortho=True):
    """Lanczos_Nirav: v0 normalised start, alpha/beta recorded before the
    single-pass modified Gram-Schmidt reorthogonalisation (:105-119), exit at
    iter == steps-1 or |dE0| < eps (:139-162)."""
    V = [np.asarray(v0, np.complex128)]
    alpha, beta = [], [0.0 + 0.0j]
    w = apply_h(V[0])
    a = np.vdot(V[0], w)
    alpha.append(a)
    w = w - a * V[0]
    b = np.linalg.norm(w)
    beta.append(b + 0j)
    V.append(w / b)
    E0, it, energies = 1.0, 0, []
    done = False
    while not done:
        it += 1
        v1, v0_ = V[it], V[it - 1]
        w = apply_h(v1)
        a = np.vdot(v1, w)
        alpha.append(a)
        w = w - a * v1 - beta[it] * v0_
        b = np.linalg.norm(w)
        beta.append(b + 0j)
        w = w / b
        if reortho:
            for i in range(it + 1):
                w = w - V[i] * np.vdot(V[i], w)
            w = w / np.linalg.norm(w)
        V.append(w)
        D0 = tridiag_eig(alpha, beta, vectors=False)[0][0]
        diff = abs(E0 - D0)
        energies.append(D0)
        if diff < eps:
            done = True
        E0 = D0
        if it == steps - 1:
            done = True
    n = it + 1
    evals, Z = tridiag_eig(alpha, beta, vectors=True)
    basis = np.array(V[: n + 1])
    gs = np.tensordot(Z[:, 0], basis[:n], axes=(0, 0))
    return LanczosResult(np.array(alpha), np.array(beta), evals, Z, basis, gs, n, energies)


def ritz_vectors(res: LanczosResult, k=None):
    """PsiAll[k] = sum_m V_m Z(m,k) (:208-234)."""
    n = res.iters
    k = n if k is None else k
    return np.tensordot(res.Z[:, :k].T, res.basis[:n], axes=(1, 0))


# --------------------------------------------------------------------------
# Spectra: restatement of DynJsJs / DynJeJe (src/dynamics.h:222-460)
# --------------------------------------------------------------------------
REF_OMEGA = 0.002 * np.arange(1001)   # src/dynamics.h:231-232,354-355 (omega = domega*om)
REF_ETA = 0.005                       # src/dynamics.h:233,356


def lehmann(evals, overlaps, expect, omega, eta, pairs):
    """I_ab(w) = sum_n conj(O_na) O_nb / (w - (E_n - E_0) - i eta)
                 - conj(<J_a>) <J_b> / (w - i eta)   (src/dynamics.h:270-316,389-425)."""
    E0 = evals[0]
    w = np.asarray(omega, np.float64)[:, None]
    den = (w - (np.asarray(evals)[None, :] - E0)) - 1j * eta
    dn1 = np.asarray(omega, np.float64) - 1j * eta
    out = {}
    for (a, b) in pairs:
        num = np.conj(overlaps[:, a]) * overlaps[:, b]
        out[(a, b)] = (num[None, :] / den).sum(axis=1) - np.conj(expect[a]) * expect[b] / dn1
    return out


def jeje_ref(m: Model, evals, states, gs, omega=REF_OMEGA, eta=REF_ETA):
    """DynJeJe columns: omega, Im xx, Im yy, Im xy, Im yx, Im tt
    (src/dynamics.h:365-435).  The yx elastic term uses Jez, which the reference
    never initialises (:368-369,424); it is taken as 0 here."""
    jx = apply_current_total(m, "Jex", gs)
    jy = apply_current_total(m, "Jey", gs)
    jt = jx + jy
    ex = np.array([np.vdot(gs, jx), np.vdot(gs, jy), np.vdot(gs, jt), 0.0])
    O = np.stack([states.conj() @ jx, states.conj() @ jy, states.conj() @ jt], axis=1)
    # overlaps column 3: same as column 1 for the yx numerator, elastic uses 0
    O4 = np.concatenate([O, O[:, 1:2]], axis=1)
    I = lehmann(evals, O4, ex, omega, eta, [(0, 0), (1, 1), (0, 1), (1, 0), (2, 2)])
    # (1,0) elastic term: conj(<Jey>) * Jez with Jez = 0 -> recompute without it
    E0 = evals[0]
    wv = np.asarray(omega)[:, None]
    den = (wv - (np.asarray(evals)[None, :] - E0)) - 1j * eta
    I[(1, 0)] = ((np.conj(O[:, 1]) * O[:, 0])[None, :] / den).sum(axis=1)
    cols = [I[(0, 0)].imag, I[(1, 1)].imag, I[(0, 1)].imag, I[(1, 0)].imag, I[(2, 2)].imag]
    return np.column_stack([np.asarray(omega)] + cols)


def jsjs_ref(m: Model, evals, states, gs, omega=REF_OMEGA, eta=REF_ETA):
    """DynJsJs columns: omega, (Re, Im) of xx, yy, zz, xy, xz, yz, then Re tt, Im tt
    (src/dynamics.h:244-328)."""
    j = [apply_current_total(m, n, gs) for n in ("Jsx", "Jsy", "Jsz")]
    j.append(j[0] + j[1] + j[2])
    ex = np.array([np.vdot(gs, x) for x in j])
    O = np.stack([states.conj() @ x for x in j], axis=1)
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2), (3, 3)]
    I = lehmann(evals, O, ex, omega, eta, pairs)
    cols = []
    for p in pairs:
        cols += [I[p].real, I[p].imag]
    return np.column_stack([np.asarray(omega)] + cols)


def apply_op(m: Model, name, site, v):
    """Observables::ApplyOp for one site (src/Observables.h:24-67)."""
    if name in ("Sx", "Sy", "Sz"):
        return apply_site_op(m.N, name, site, v)
    return apply_current_site(m, name, site, v)


def two_point_ref(m: Model, op_a, op_b, gs):
    """Observables::TwoPointCorr (src/Observables.h:132-155): Corr(si, sj) =
    <gs|A_si B_sj|gs> for si >= sj, copied to (sj, si)."""
    N = m.N
    C = np.zeros((N, N), complex)
    for sj in range(N):
        out1 = apply_op(m, op_b, sj, gs)
        for si in range(sj, N):
            C[si, sj] = np.vdot(gs, apply_op(m, op_a, si, out1))
            C[sj, si] = C[si, sj]
    return C


# rotated spin-flip operators of src/Observables.h:539-589, with the
# reference's own six-digit coefficients
SPE3 = (complex(-0.408248, 0.707107), complex(-0.408248, -0.707107), 0.816497)
SME3 = (complex(-0.408248, -0.707107), complex(-0.408248, 0.707107), 0.816497)


def local_szsz_ref(m: Model, evals, states, gs, omega=REF_OMEGA, eta=REF_ETA):
    """LocalDynSzSz columns: omega, Im of I_xx, I_yy, I_zz, I_xy, I_xz, I_yz
    summed over sites, then Im of the S+e3 S+e3 and S-e3 S-e3 channels
    (src/dynamics.h:64-217).  Per site: overlaps <n|S^a_i|gs> with the elastic
    term conj(<S^a_i>) <S^b_i> / (w - i eta) subtracted; the two rotated
    channels subtract <S-e3_i><S+e3_i> (as written, without a conjugate)."""
    N = m.N
    w = np.asarray(omega, np.float64)
    E0 = evals[0]
    den = (w[:, None] - (np.asarray(evals)[None, :] - E0)) - 1j * eta
    dn1 = w - 1j * eta
    acc = {k: np.zeros(len(w), complex) for k in ("xx", "yy", "zz", "xy", "xz", "yz", "pp", "mm")}
    for si in range(N):
        sx, sy, sz = (apply_site_op(N, n, si, gs) for n in ("Sx", "Sy", "Sz"))
        sp = SPE3[0] * sx + SPE3[1] * sy + SPE3[2] * sz
        sm = SME3[0] * sx + SME3[1] * sy + SME3[2] * sz
        vec = {"x": sx, "y": sy, "z": sz, "p": sp, "m": sm}
        ex = {a: np.vdot(gs, v) for a, v in vec.items()}
        O = {a: states.conj() @ v for a, v in vec.items()}
        for a, b in (("x", "x"), ("y", "y"), ("z", "z"), ("x", "y"), ("x", "z"), ("y", "z")):
            acc[a + b] += (np.conj(O[a]) * O[b])[None, :].__truediv__(den).sum(axis=1) - np.conj(ex[a]) * ex[b] / dn1
        acc["pp"] += (np.abs(O["p"]) ** 2)[None, :].__truediv__(den).sum(axis=1) - ex["m"] * ex["p"] / dn1
        acc["mm"] += (np.abs(O["m"]) ** 2)[None, :].__truediv__(den).sum(axis=1) - ex["p"] * ex["m"] / dn1
    cols = [acc[k].imag for k in ("xx", "yy", "zz", "xy", "xz", "yz", "pp", "mm")]
    return np.column_stack([w] + cols)


def kill_site(m, site):
    """Hamiltonian::RIXS_killcenter (src/Hamiltonian.h:57-91): the model with
    every bond of `site` removed (its couplings set to zero)."""
    import copy
    k = copy.copy(m)
    k.bonds = [b[(b[:, 0] != site) & (b[:, 1] != site)] if len(b) else b for b in m.bonds]
    return k


def inv_complex(x, y):
    """invComplex (src/dynamics.h:34-39): (x + i y) / (x^2 + y^2)."""
    d = x * x + y * y
    return x / d + 1j * y / d


TENTH_OMEGA = 0.1 * np.arange(121)    # src/dynamics.h:532-533,587-588


def rixs_ref(N, site, evals_gs, states_gs, evals_int, states_int, gs, omega=TENTH_OMEGA, eta=0.1, gamma=100.0):
    """LocalDynRIXS (src/dynamics.h:497-560) as written: A_X(m, n) = <n|gs>
    <m|X|n> invComplex(0 - E_n, Gamma) for X = 1, S^x, S^y, S^z at `site`, then
    I_X(w) = sum_{m >= 1} |sum_n A_X(m, n)|^2 eta / ((w - (E_m - E_0))^2 + eta^2).
    states_* are rows (Ritz or eigen vectors).  Returns the 9-column table."""
    n0 = states_int.conj() @ gs                            # nstate.dot(Psi)
    inv = inv_complex(0.0 - np.asarray(evals_int), gamma)
    Eg = evals_gs[0]
    w = np.asarray(omega, np.float64)
    xm = w[:, None] - (np.asarray(evals_gs)[None, 1:] - Eg)
    lor = eta / (xm * xm + eta * eta)
    cols = []
    for X in (None, "Sx", "Sy", "Sz"):
        Xn = states_int if X is None else np.array([apply_site_op(N, X, site, v) for v in states_int])
        mXn = states_gs.conj() @ Xn.T                      # mstate.dot(X nstate)
        S = (n0[None, :] * mXn * inv[None, :]).sum(axis=1)
        cols.append(((S * S.conj()).real[None, 1:] * lor).sum(axis=1))
    return np.column_stack([w] + cols + [np.zeros_like(w)] * 4)


def szsz_ref(N, evals, states, gs, omega=TENTH_OMEGA, eta=0.1):
    """DynSzSz (src/dynamics.h:566-611) as written: for every (si, sj, w) the
    sums over states n of Re(<Szj|n><n|Szi>) x/(x^2+eta^2) and
    ... eta/(x^2+eta^2), x = w - (E_n - E_0).  Returns S[w, si, sj] = real + i imag."""
    Z = np.array([apply_site_op(N, "Sz", i, gs) for i in range(N)])
    O = Z.conj() @ states.T                                # O[i, n] = <Szi|n>
    w = np.asarray(omega, np.float64)
    x = w[:, None] - (np.asarray(evals)[None, :] - evals[0])
    kre = x / (x * x + eta * eta)
    kim = eta / (x * x + eta * eta)
    out = np.zeros((len(w), N, N), complex)
    for si in range(N):
        for sj in range(N):
            tmp = (O[sj] * O[si].conj()).real              # Szj.dot(n) * n.dot(Szi)
            out[:, si, sj] = kre @ tmp + 1j * (kim @ tmp)
    return out


def format_g6(x):
    """C++ ostream default formatting (precision 6, %g-like)."""
    s = "%g" % x
    return s


def write_real(path, table):
    """Byte format of *_outSpectrum.real (src/dynamics.h:325-328,432-435): each
    row is 'w \\t v1 \\t ... \\t vK \\n ' (values with default ostream
    precision), so every row after the first starts with a space and the file
    ends with a lone space."""
    with open(path, "w") as f:
        for row in table:
            f.write(" \t ".join(format_g6(x) for x in row))
            f.write(" \n ")


def read_real(path):
    with open(path) as f:
        rows = [ln.split() for ln in f.read().split("\n") if ln.strip()]
    return np.array([[float(x) for x in r] for r in rows])


# --------------------------------------------------------------------------
# Continued fraction (north-star spectral mode; no reference counterpart)
# --------------------------------------------------------------------------
def cf_resolvent(alpha, beta, norm2, z):
    """<phi|(z - H)^-1|phi> from the Lanczos run started at phi/|phi|:
    norm2 / (z - a0 - b1^2 / (z - a1 - b2^2 / ...)); beta[k] couples k-1,k."""
    a = np.real(np.asarray(alpha))
    b = np.abs(np.asarray(beta))
    z = np.asarray(z, np.complex128)
    g = np.zeros_like(z)
    for k in range(len(a) - 1, -1, -1):
        nxt = (b[k + 1] ** 2 * g) if k + 1 < len(a) else 0.0
        g = 1.0 / (z - a[k] - nxt)
    return norm2 * g


def lanczos_plain(apply_h, phi, steps):
    """Three-vector Lanczos without reorthogonalisation from phi/|phi|."""
    nrm = np.linalg.norm(phi)
    v = phi / nrm
    vp = np.zeros_like(v)
    alpha, beta = [], [0.0]
    for k in range(steps):
        w = apply_h(v) - beta[-1] * vp
        a = np.vdot(v, w).real
        w = w - a * v
        alpha.append(a)
        b = np.linalg.norm(w)
        beta.append(b)
        if b == 0.0:
            break
        vp, v = v, w / b
    return np.array(alpha), np.array(beta), nrm ** 2


def cf_dynamics_table(m: Model, gs, e0, omega, eta, steps, kind="JeJe", apply_h=None):
    """JsJs / JeJe table by continued fractions from J|gs> (north-star spectral
    mode): the columns, channel set and elastic subtraction of DynJsJs / DynJeJe
    (src/dynamics.h:244-328, 365-435), each I_ab(w) = <J_a gs|(w + e0 - i eta -
    H)^-1|J_b gs>.  Diagonal channels are one plain Lanczos run each; a cross
    channel (a, b) comes from the runs started at J_a + J_b and J_a + i J_b by
    polarisation.  The JeJe yx column has no elastic term (Jez is never set in
    the reference, src/dynamics.h:368-369,424).  Pinned by the exact Lehmann
    sum from full diagonalisation (tests/test_oracle_pins.py)."""
    if apply_h is None:
        apply_h = lambda v: hpsi_gather(m, v)  # noqa: E731
    names = ("Jex", "Jey") if kind == "JeJe" else ("Jsx", "Jsy", "Jsz")
    J = [apply_current_total(m, n, gs) for n in names]
    nc = len(J)
    J.append(sum(J))
    ex = [np.vdot(gs, x) for x in J]
    z = np.asarray(omega, np.float64) + e0 - 1j * eta

    def res(phi):
        a, b, n2 = lanczos_plain(apply_h, phi, steps)
        return cf_resolvent(a, b, n2, z)

    R = {(a, a): res(J[a]) for a in range(nc + 1)}
    if kind == "JeJe":
        pairs = [(0, 0), (1, 1), (0, 1), (1, 0), (nc, nc)]
    else:
        pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2), (nc, nc)]
    for (a, b) in pairs:
        if a == b or (a, b) in R:
            continue
        s1 = res(J[a] + J[b])
        si = res(J[a] + 1j * J[b])
        u = s1 - R[(a, a)] - R[(b, b)]
        v = si - R[(a, a)] - R[(b, b)]
        R[(a, b)] = 0.5 * (u - 1j * v)
        R[(b, a)] = 0.5 * (u + 1j * v)
    dn1 = np.asarray(omega, np.float64) - 1j * eta
    cols = [np.asarray(omega, np.float64)]
    for (a, b) in pairs:
        val = R[(a, b)]
        if not (kind == "JeJe" and (a, b) == (1, 0)):
            val = val - np.conj(ex[a]) * ex[b] / dn1
        if kind == "JeJe":
            cols.append(val.imag)
        else:
            cols += [val.real, val.imag]
    return np.column_stack(cols)
