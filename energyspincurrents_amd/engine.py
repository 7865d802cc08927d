"""Python mirror of the reference's operator interfaces over the C ABI.

Classes follow the reference roles (src/Lattice.h, src/Hamiltonian.h,
src/DynLanczosSolver.h, src/Observables.h, src/dynamics.h):

  Honeycomb   Lattice::Honeycomb1 + Hamiltonian::Connectors  (esc_honeycomb / esc_connectors)
  Chain       Lattice::Chain + HeisenbergChain_Connectors     (esc_chain / esc_chain_connectors)
  SpinModel   the coupling set (Kxx..Kzz, Bxx..Bzz, and the J / Gamma extensions)
  Engine      one device context: H.psi (MatrixVectorMult), ApplyOp / ApplyJ*Total,
              Lanczos_Nirav, overlaps, StartDynamics (esc_dynamics)

Device vectors are torch tensors (complex128 for "c128", complex64 for "c64")
on the current HIP device; torch supplies memory and the stream only, every
operation runs in libesc.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._capi import (DYN_KINDS, ESC_C64, ESC_C128, LANC_MODES, OPS, SITE_TOTAL, SPEC_METHODS, START_KINDS,
                    EscModel, EscParams, check, lib)


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dev_ptr(x, dtype, n, name, optional=False):
    """Device pointer of a vector argument, checked before anything launches.

    The kernels index n elements of the context's precision from the pointer,
    so a wrong dtype, a short tensor, a strided view or host memory would be an
    out-of-bounds device access: each raises ValueError/TypeError here."""
    if x is None:
        if optional:
            return None
        raise ValueError(f"{name}: a device vector is required")
    import torch
    if not isinstance(x, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor on the GPU, got {type(x).__name__}")
    if x.device.type != "cuda":
        raise ValueError(f"{name}: tensor is on {x.device}, the engine needs device memory")
    if x.dtype != dtype:
        raise ValueError(f"{name}: dtype {x.dtype} does not match the engine precision ({dtype})")
    if x.numel() != n:
        raise ValueError(f"{name}: {x.numel()} elements, the basis has {n}")
    if not x.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    return ctypes.c_void_p(x.data_ptr())


def _stream(stream):
    if stream is not None:
        return ctypes.c_void_p(stream if isinstance(stream, int) else stream.cuda_stream)
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Honeycomb:
    """Honeycomb torus with Lx unit cells along a1 and Ly/2 along a2 (src/Lattice.h:434-576)."""

    def __init__(self, Lx, Ly, periodic_x=True, periodic_y=True):
        self.Lx, self.Ly = int(Lx), int(Ly)
        self.periodic_x, self.periodic_y = bool(periodic_x), bool(periodic_y)
        self.N = self.Lx * self.Ly
        self.neigh = np.empty((self.N, 3), np.int32)
        check(lib().esc_honeycomb(self.Lx, self.Ly, int(periodic_x), int(periodic_y), _np_ptr(self.neigh)))

    def bonds(self, K=(1.0, 1.0, 1.0)):
        """Bond families x, y, z as (n, 2) int32 arrays (src/Hamiltonian.h:230-328)."""
        Kd = np.asarray(K, np.float64)
        px = np.zeros((self.N, 2), np.int32)
        py = np.zeros((self.N, 2), np.int32)
        pz = np.zeros((self.N, 2), np.int32)
        cnt = np.zeros(3, np.int32)
        check(lib().esc_connectors(self.N, _np_ptr(self.neigh), _np_ptr(Kd), _np_ptr(px), _np_ptr(py), _np_ptr(pz),
                                   _np_ptr(cnt)))
        return [px[: cnt[0]].copy(), py[: cnt[1]].copy(), pz[: cnt[2]].copy()]


class Chain:
    """Chain of N sites for Model=HeisenbergChain (src/Lattice.h:60-125): column
    0 of the neighbour table is the +x neighbour, column 1 the -x neighbour."""

    def __init__(self, N, periodic=True):
        self.N = int(N)
        self.periodic = bool(periodic)
        self.neigh = np.empty((self.N, 2), np.int32)
        check(lib().esc_chain(self.N, int(periodic), _np_ptr(self.neigh)))

    def bonds(self, K=(1.0, 1.0, 1.0)):
        """Bond families x, y, z: every family of a non-zero coupling holds all
        chain bonds (src/Hamiltonian.h:381-475)."""
        Kd = np.asarray(K, np.float64)
        px, py, pz = (np.zeros((self.N, 2), np.int32) for _ in range(3))
        cnt = np.zeros(3, np.int32)
        check(lib().esc_chain_connectors(self.N, _np_ptr(self.neigh), _np_ptr(Kd), _np_ptr(px), _np_ptr(py),
                                         _np_ptr(pz), _np_ptr(cnt)))
        return [px[: cnt[0]].copy(), py[: cnt[1]].copy(), pz[: cnt[2]].copy()]


def field111(h):
    b = h / np.sqrt(3.0)
    return (b, b, b)


@dataclass
class SpinModel:
    lattice: "Honeycomb | Chain"
    K: tuple = (1.0, 1.0, 1.0)
    J: float = 0.0
    Gamma: float = 0.0
    B: tuple = (0.0, 0.0, 0.0)
    skip_single_bond_family: bool = False   # reference rule, src/Hamiltonian.h:116,143,347
    core_hole: int | None = None            # bonds of this site removed (src/Hamiltonian.h:57-91)

    def intermediate(self, site):
        """The intermediate Hamiltonian of LocalDynRIXS: this model with every
        bond of `site` removed (Hamiltonian::RIXS_killcenter)."""
        import dataclasses
        return dataclasses.replace(self, core_hole=int(site))

    def struct(self):
        """esc_model for this model; keeps the arrays it points into alive on self."""
        chain = isinstance(self.lattice, Chain)
        if chain and (self.J != 0.0 or self.Gamma != 0.0):
            raise ValueError("J and Gamma are honeycomb extensions; set K on a chain")
        # honeycomb families exist for every lattice bond (a zero coupling only
        # zeroes the weight); chain families exist for non-zero couplings
        self._bonds = self.lattice.bonds(self.K if chain else (1.0, 1.0, 1.0))
        if self.skip_single_bond_family:
            self._bonds = [b if len(b) != 1 else b[:0] for b in self._bonds]
        if self.core_hole is not None:
            c = self.core_hole
            self._bonds = [b[(b[:, 0] != c) & (b[:, 1] != c)] for b in self._bonds]
        m = EscModel()
        m.n_sites = self.lattice.N
        for f in range(3):
            self._bonds[f] = np.ascontiguousarray(self._bonds[f], np.int32)
            m.n_bonds[f] = len(self._bonds[f])
            m.bonds[f] = self._bonds[f].ctypes.data if len(self._bonds[f]) else None
        m.neigh = None if chain else self.lattice.neigh.ctypes.data   # current operators are honeycomb-only
        for f in range(3):
            m.K[f] = float(self.K[f])
            m.B[f] = float(self.B[f])
        m.J = float(self.J)
        m.Gamma = float(self.Gamma)
        return m


@dataclass
class LanczosResult:
    alpha: np.ndarray
    beta: np.ndarray
    iters: int
    evals: np.ndarray
    Z: np.ndarray          # (iters, iters), column k = eigenvector k


class Engine:
    """One esc_ctx: the device-side operator set of one model."""

    def __init__(self, model: SpinModel, precision="c128"):
        self.model = model
        self.precision = precision
        self._m = model.struct()
        h = ctypes.c_void_p()
        check(lib().esc_ctx_create(ctypes.byref(self._m), ESC_C64 if precision == "c64" else ESC_C128,
                                   ctypes.byref(h)))
        self._h = h
        d, e, p = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().esc_ctx_info(h, ctypes.byref(d), ctypes.byref(e), ctypes.byref(p)))
        self.dim, self.elem_bytes = d.value, e.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib().esc_ctx_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- device vectors ---------------------------------------------------
    @property
    def torch_dtype(self):
        import torch
        return torch.complex64 if self.precision == "c64" else torch.complex128

    def empty(self):
        import torch
        return torch.empty(self.dim, dtype=self.torch_dtype, device="cuda")

    def _v(self, x, name, optional=False):
        return _dev_ptr(x, self.torch_dtype, self.dim, name, optional)

    def upload(self, host: np.ndarray, out=None):
        out = self.empty() if out is None else out
        h = np.ascontiguousarray(host, np.complex128).reshape(-1)
        if h.size != self.dim:
            raise ValueError(f"host vector has {h.size} elements, the basis has {self.dim}")
        check(lib().esc_upload(self._h, _np_ptr(h), self._v(out, "out"), _stream(None)))
        return out

    def download(self, x) -> np.ndarray:
        h = np.empty(self.dim, np.complex128)
        check(lib().esc_download(self._h, self._v(x, "x"), _np_ptr(h), _stream(None)))
        return h

    # -- operators --------------------------------------------------------
    def hpsi(self, x, out=None, stream=None):
        """out = H x  (Hamiltonian::MatrixVectorMult, src/Hamiltonian.h:480-557)."""
        out = self.empty() if out is None else out
        check(lib().esc_hpsi(self._h, self._v(x, "x"), self._v(out, "out"), _stream(stream)))
        return out

    def autotune(self, x, out=None, reps=5, stream=None):
        """Time every H.psi tile size on x and keep the fastest (esc_autotune)."""
        out = self.empty() if out is None else out
        tb, ms = ctypes.c_int(), ctypes.c_double()
        check(lib().esc_autotune(self._h, self._v(x, "x"), self._v(out, "out"), int(reps), ctypes.byref(tb),
                                 ctypes.byref(ms), _stream(stream)))
        return tb.value, ms.value

    def set_tile_bits(self, tile_bits: int):
        check(lib().esc_set_tile_bits(self._h, int(tile_bits)))

    def kernel_configs(self):
        """H.psi kernel configurations: 0 (one row per lane) or
        100000 * planned passes (5: cover-first planner, 6: its windows in
        reverse order; 7, 8: the generated sub-cube kernels over those plans)
        + 10000 * log2(waves per workgroup) + 100 * tile bits + window bits,
        plus, for the generated kernels, 1000000 * log2(rows per lane) and
        10000000 * layout: bit 0 over row pairs (complex64, 16-byte
        accesses), bit 1 lanes on six contiguous sites (offered from 25
        sites without row pairs; otherwise only a 64-byte run of lane sites
        is fixed)."""
        n = ctypes.c_int32()
        check(lib().esc_kernel_configs(self._h, None, 0, ctypes.byref(n)))
        buf = (ctypes.c_int32 * max(n.value, 1))()
        check(lib().esc_kernel_configs(self._h, buf, n.value, ctypes.byref(n)))
        return [buf[i] for i in range(n.value)]

    def apply_op(self, op: str, site: int, x, out=None, stream=None):
        """Observables::ApplyOp / ApplyJ*Total (site = -1) (src/Observables.h:24-67,448-470,664-699)."""
        out = self.empty() if out is None else out
        check(lib().esc_apply_op(self._h, OPS[op], int(site), self._v(x, "x"), self._v(out, "out"), _stream(stream)))
        return out

    def apply_total(self, op: str, x, out=None, stream=None):
        return self.apply_op(op, SITE_TOTAL, x, out, stream)

    def start_vector(self, seed: int, kind="Uniform", out=None, stream=None):
        out = self.empty() if out is None else out
        check(lib().esc_start_vector(self._h, ctypes.c_uint64(seed), START_KINDS[kind], self._v(out, "out"),
                                     _stream(stream)))
        return out

    def dot(self, x, y, stream=None) -> complex:
        r = np.zeros(2)
        check(lib().esc_dot(self._h, self._v(x, "x"), self._v(y, "y"), _np_ptr(r), _stream(stream)))
        return complex(r[0], r[1])

    def dot_noalloc(self, x, y, out, stream=None):
        """out[0], out[1] = Re, Im <x|y> in device memory (float64 tensor of 2),
        computed without any allocation (esc_dot_noalloc)."""
        import torch
        if not (isinstance(out, torch.Tensor) and out.dtype == torch.float64 and out.numel() >= 2
                and out.device.type == "cuda" and out.is_contiguous()):
            raise ValueError("out: contiguous float64 device tensor of at least 2 elements")
        check(lib().esc_dot_noalloc(self._h, self._v(x, "x"), self._v(y, "y"), ctypes.c_void_p(out.data_ptr()),
                                    _stream(stream)))
        return out

    def axpby(self, a: complex, x, b: complex, y, stream=None):
        av = np.array([a.real, a.imag]) if isinstance(a, complex) else np.array([float(a), 0.0])
        bv = np.array([b.real, b.imag]) if isinstance(b, complex) else np.array([float(b), 0.0])
        check(lib().esc_axpby(self._h, _np_ptr(av), self._v(x, "x"), _np_ptr(bv), self._v(y, "y"), _stream(stream)))
        return y

    # -- Lanczos ----------------------------------------------------------
    def lanczos(self, v0, steps: int, eps=0.0, mode="All", gs=None, stream=None) -> LanczosResult:
        """DynLanczosSolver::Lanczos_Nirav (src/DynLanczosSolver.h:38-239)."""
        a = np.zeros(steps + 1)
        b = np.zeros(steps + 2)
        ev = np.zeros(steps + 1)
        Z = np.zeros((steps + 1) * (steps + 1))
        it = ctypes.c_int()
        check(lib().esc_lanczos(self._h, self._v(v0, "v0"), int(steps), float(eps), LANC_MODES[mode], _np_ptr(a),
                                _np_ptr(b), ctypes.byref(it), _np_ptr(ev), _np_ptr(Z), self._v(gs, "gs", optional=True),
                                _stream(stream)))
        n = it.value
        return LanczosResult(a[:n].copy(), b[: n + 1].copy(), n, ev[:n].copy(),
                             Z[: n * n].reshape(n, n).T.copy())

    def exact_diag(self, gs=None, stream=None):
        """Solver=ED: every eigenvalue (ascending) of the dense H, diagonalised
        on the device (ExactDiag + fillFromED, src/Exactdiag.h:36-110,
        src/DynLanczosSolver.h:269-283).  The eigenvectors are kept as the
        Ritz states for Ritz spectra; gs (optional) receives the lowest."""
        ev = np.zeros(self.dim)
        check(lib().esc_exact_diag(self._h, _np_ptr(ev), self._v(gs, "gs", optional=True), _stream(stream)))
        return ev

    def ground_state(self, v0, steps=200, eps=0.0, restarts=0, out=None, stream=None):
        """Ground state with 5 resident vectors (esc_ground_state): two-pass
        Lanczos without reorthogonalisation, `restarts` restarts from the
        previous state.  Returns (out, e0, residual norm, H.psi count)."""
        out = self.empty() if out is None else out
        e0, res, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(lib().esc_ground_state(self._h, self._v(v0, "v0"), int(steps), float(eps), int(restarts), ctypes.byref(e0),
                                     ctypes.byref(res), ctypes.byref(it), self._v(out, "out"), _stream(stream)))
        return out, e0.value, res.value, it.value

    def ritz_overlaps(self, vec, n_states, stream=None) -> np.ndarray:
        out = np.zeros(2 * n_states)
        check(lib().esc_ritz_overlaps(self._h, self._v(vec, "vec"), int(n_states), _np_ptr(out), _stream(stream)))
        return out[0::2] + 1j * out[1::2]

    def ritz_vector(self, k, out=None, stream=None):
        out = self.empty() if out is None else out
        check(lib().esc_ritz_vector(self._h, int(k), self._v(out, "out"), _stream(stream)))
        return out

    def two_point(self, op_a: str, op_b: str, gs, stream=None):
        """Observables::TwoPointCorr (src/Observables.h:132-155): N x N matrix
        <gs|A_i B_j|gs>, filled for i >= j and mirrored."""
        N = self.model.lattice.N
        out = np.zeros(2 * N * N)
        check(lib().esc_two_point(self._h, OPS[op_a], OPS[op_b], self._v(gs, "gs"), _np_ptr(out), _stream(stream)))
        return (out[0::2] + 1j * out[1::2]).reshape(N, N)

    def lanczos_cf(self, phi, steps, stream=None):
        a = np.zeros(steps + 1)
        b = np.zeros(steps + 2)
        it = ctypes.c_int()
        n2 = ctypes.c_double()
        check(lib().esc_lanczos_cf(self._h, self._v(phi, "phi"), int(steps), _np_ptr(a), _np_ptr(b), ctypes.byref(it),
                                   ctypes.byref(n2), _stream(stream)))
        n = it.value
        return a[:n].copy(), b[: n + 1].copy(), n2.value

    def dynamics(self, kind: str, method: str, gs, omega, eta, e0=0.0, cf_steps=200, stream=None):
        """StartDynamics -> DynJsJs / DynJeJe without the file (src/dynamics.h:43-60,222-460).
        Returns (table, expectations)."""
        omega = np.ascontiguousarray(omega, np.float64)
        ncols = {"JsJs": 14, "JeJe": 5, "LocalSzSz": 8}[kind]
        table = np.zeros((len(omega), 1 + ncols))
        expect = np.zeros(2 * 4)
        check(lib().esc_dynamics(self._h, DYN_KINDS[kind], SPEC_METHODS[method], self._v(gs, "gs"), float(e0),
                                 int(cf_steps), float(eta), _np_ptr(omega), len(omega), _np_ptr(table),
                                 _np_ptr(expect), _stream(stream)))
        nexp = 4 if kind == "JsJs" else 3
        return table, (expect[0::2] + 1j * expect[1::2])[:nexp]


    def rixs(self, inter: "Engine", site, gs, method="Ritz", e0=0.0, cf_steps=100, gamma=100.0, eta=0.1,
             omega=None, stream=None):
        """LocalDynRIXS (src/dynamics.h:465-562) at core-hole `site`; `inter` is the
        engine of model.intermediate(site).  Ritz mode needs "All" Lanczos runs
        kept on both engines.  Returns the (nw, 9) table of RIXS_outSpectrum.real."""
        omega = np.ascontiguousarray(tenth_omega() if omega is None else omega, np.float64)
        table = np.zeros((len(omega), 9))
        check(lib().esc_rixs(self._h, inter._h, SPEC_METHODS[method], int(site), self._v(gs, "gs"), float(e0),
                             int(cf_steps), float(gamma), float(eta), _np_ptr(omega), len(omega), _np_ptr(table),
                             _stream(stream)))
        return table

    def szsz(self, gs, eta=0.1, omega=None, stream=None):
        """DynSzSz (src/dynamics.h:566-611) over the Ritz states of the kept "All"
        run.  Returns S[w, i, j] = real + i imag of the reference's sums."""
        omega = np.ascontiguousarray(tenth_omega() if omega is None else omega, np.float64)
        N = self.model.lattice.N
        table = np.zeros((len(omega), N, N, 2))
        check(lib().esc_szsz(self._h, self._v(gs, "gs"), float(eta), _np_ptr(omega), len(omega), _np_ptr(table),
                             _stream(stream)))
        return table[..., 0] + 1j * table[..., 1]


# -- host FP64 small solves -------------------------------------------------
class Batch:
    """Many models solved in lockstep (esc_batch_*): each Lanczos step of all
    instances is one replayed HIP graph.  Instances may differ in size and
    couplings; all share one precision."""

    def __init__(self, models, precision="c64"):
        import copy
        self.models = list(models)
        self.precision = precision
        # one copy per entry: struct() keeps the bond arrays it points into on
        # the model object, so a model listed twice must not share them
        self._copies = [copy.copy(m) for m in self.models]
        self._structs = [m.struct() for m in self._copies]
        arr = (EscModel * len(self._structs))(*self._structs)
        self._arr = arr
        h = ctypes.c_void_p()
        check(lib().esc_batch_create(len(self._structs), arr, ESC_C64 if precision == "c64" else ESC_C128,
                                     ctypes.byref(h)))
        self._h = h
        self.n = len(self._structs)
        self.e0 = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib().esc_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, reps=3):
        """Pick every instance's H.psi kernel (esc_batch_autotune); returns the configurations."""
        cfg = np.zeros(self.n, np.int32)
        check(lib().esc_batch_autotune(self._h, int(reps), _np_ptr(cfg)))
        return cfg

    def ground_state(self, seed=0, steps=200, use_graph=True):
        """Returns (e0[n], iters[n], alpha (n, steps), beta (n, steps + 1))."""
        e0 = np.zeros(self.n)
        it = np.zeros(self.n, np.int32)
        al = np.zeros((self.n, steps))
        be = np.zeros((self.n, steps + 1))
        check(lib().esc_batch_ground_state(self._h, int(seed), int(steps), int(bool(use_graph)), _np_ptr(e0),
                                           _np_ptr(it), _np_ptr(al), _np_ptr(be)))
        self.e0 = e0
        return e0, it, al, be

    def state(self, i, out=None, stream=None):
        """Instance i's ground state as a device tensor."""
        import torch
        if not 0 <= int(i) < self.n:
            raise ValueError(f"instance {i} out of range [0, {self.n})")
        m = self.models[i]
        dt = torch.complex64 if self.precision == "c64" else torch.complex128
        if out is None:
            out = torch.empty(1 << m.lattice.N, dtype=dt, device="cuda")
        check(lib().esc_batch_state(self._h, int(i), _dev_ptr(out, dt, 1 << m.lattice.N, "out"), _stream(stream)))
        if stream is None:
            torch.cuda.current_stream().synchronize()
        return out

    def dynamics(self, kind, omega, eta, cf_steps=150, use_graph=True):
        """JsJs / JeJe spectra of every instance (continued fractions).
        Returns (tables (n, nw, 1 + ncols), expectations (n, ncomp))."""
        omega = np.ascontiguousarray(omega, np.float64)
        ncols = 14 if kind == "JsJs" else 5
        tables = np.zeros((self.n, len(omega), 1 + ncols))
        expect = np.zeros((self.n, 8))
        check(lib().esc_batch_dynamics(self._h, DYN_KINDS[kind], int(cf_steps), float(eta), _np_ptr(omega),
                                       len(omega), int(bool(use_graph)), _np_ptr(tables), _np_ptr(expect)))
        nexp = 4 if kind == "JsJs" else 3
        return tables, (expect[:, 0::2] + 1j * expect[:, 1::2])[:, :nexp]


def default_kernel_config(n_sites, precision="c128"):
    """The generated-kernel configuration a context of n_sites sites takes by
    default from 19 sites (default_tile_bits in esc_capi.cpp); also used for
    smaller lattices, where the window holds every site."""
    w = n_sites if n_sites <= 18 else (17 if n_sites <= 22 else 18)
    # layout digit (x0000000): bit 0 over row pairs (16-byte accesses), bit 1
    # lanes on six contiguous sites
    if precision == "c64":
        return (13731200 if n_sites <= 22 else (14821200 if n_sites <= 26 else 13831300)) + w
    return 3731200 + w


def precompile_kernel(model: SpinModel, precision="c128", config=None):
    """Build a generated H.psi kernel configuration for this lattice into the
    code-object cache without a device (esc_kernel_precompile) and return its
    HIP source.  config None: the configuration a context of this size uses
    by default (complex128: 2^3 rows per lane, 8 waves, 2^12-row tiles;
    complex64 over row pairs: 2^3 rows per lane, 8 waves, 2^12-row tiles up
    to 22 sites; 2^4 rows per lane, 4 waves, 2^12-row tiles, windows in
    reverse order up to 26 sites; 2^3 rows per lane, 8 waves, 2^13-row
    tiles, windows in reverse order beyond; at most 17 window sites up to 22
    sites, 18 beyond)."""
    N = model.lattice.N
    if config is None:
        config = default_kernel_config(N, precision)
    st = model.struct()
    n = ctypes.c_size_t()
    check(lib().esc_kernel_precompile(ctypes.byref(st), ESC_C64 if precision == "c64" else ESC_C128, int(config),
                                      None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    check(lib().esc_kernel_precompile(ctypes.byref(st), ESC_C64 if precision == "c64" else ESC_C128, int(config),
                                      buf, len(buf), ctypes.byref(n)))
    return buf.value.decode()


def kernel_plan(model: SpinModel, precision="c128", config=None):
    """The pass plan of an H.psi kernel configuration without a device
    (esc_kernel_plan): a list of per-pass dicts (window, tile, near, far,
    lds_bytes, diag) and, per off-diagonal term, (mask, pass)."""
    if config is None:
        config = default_kernel_config(model.lattice.N, precision)
    st = model.struct()
    prec = ESC_C64 if precision == "c64" else ESC_C128
    npass, nterm = ctypes.c_int(), ctypes.c_int()
    check(lib().esc_kernel_plan(ctypes.byref(st), prec, int(config), None, 0, ctypes.byref(npass), None, None, 0,
                                ctypes.byref(nterm)))
    rows = np.zeros(6 * max(npass.value, 1), np.int64)
    tp = np.zeros(max(nterm.value, 1), np.int32)
    tm = np.zeros(max(nterm.value, 1), np.uint32)
    check(lib().esc_kernel_plan(ctypes.byref(st), prec, int(config), _np_ptr(rows), npass.value,
                                ctypes.byref(npass), _np_ptr(tp), _np_ptr(tm), nterm.value, ctypes.byref(nterm)))
    keys = ("window", "tile", "near", "far", "lds_bytes", "diag")
    passes = [dict(zip(keys, (int(x) for x in rows[6 * p:6 * p + 6]))) for p in range(npass.value)]
    return passes, [(int(tm[t]), int(tp[t])) for t in range(nterm.value)]


def tridiag_eig(alpha, beta, vectors=True):
    a = np.ascontiguousarray(np.real(alpha), np.float64)
    b = np.ascontiguousarray(np.abs(beta), np.float64)
    n = len(a)
    ev = np.zeros(n)
    Z = np.zeros(n * n) if vectors else None
    check(lib().esc_tridiag_eig(n, _np_ptr(a), _np_ptr(b), _np_ptr(ev), _np_ptr(Z) if vectors else None))
    return ev, (Z.reshape(n, n).T.copy() if vectors else None)


def tridiag_lowest(alpha, beta, device=False):
    """Lowest eigenpair (e0, z) of the tridiagonal matrix of a recursion
    (esc_tridiag_lowest): alpha[0..n-1], beta[1..n-1] (beta[0] ignored); on
    the host, or with device=True by the batched solves' kernel."""
    a = np.ascontiguousarray(np.real(alpha), np.float64)
    b = np.ascontiguousarray(np.abs(beta), np.float64)
    n = len(a)
    if len(b) < n:
        raise ValueError("beta needs n entries (beta[0] ignored)")
    e0 = ctypes.c_double()
    z = np.zeros(n)
    check(lib().esc_tridiag_lowest(n, _np_ptr(a), _np_ptr(b), ctypes.byref(e0), _np_ptr(z), int(bool(device))))
    return e0.value, z


def tridiag_eigenpairs(alpha, beta, device=False):
    """Every eigenpair (evals, Z) of the tridiagonal matrix of a recursion
    (esc_tridiag_eigenpairs): bisection and inverse iteration, on the host or
    with device=True on the current device; Z[:, j] is eigenvector j."""
    a = np.ascontiguousarray(np.real(alpha), np.float64)
    b = np.ascontiguousarray(np.abs(beta), np.float64)
    n = len(a)
    if len(b) < n:
        raise ValueError("beta needs n entries (beta[0] ignored)")
    ev = np.zeros(n)
    Z = np.zeros(n * n)
    check(lib().esc_tridiag_eigenpairs(n, _np_ptr(a), _np_ptr(b), _np_ptr(ev), _np_ptr(Z), int(bool(device))))
    return ev, Z.reshape(n, n).T.copy()


def cf_eval(alpha, beta, norm2, e0, eta, omega, device=False):
    """The raw continued fraction of a recursion on every omega (esc_cf_eval),
    on the host threads or, with device=True, by the batched spectra's kernel."""
    a = np.ascontiguousarray(np.real(alpha), np.float64)
    b = np.ascontiguousarray(beta, np.float64)
    if len(b) < len(a):
        raise ValueError("beta needs n entries (beta[0] ignored)")
    om = np.ascontiguousarray(omega, np.float64)
    out = np.zeros(2 * len(om))
    check(lib().esc_cf_eval(len(a), _np_ptr(a), _np_ptr(b), float(norm2), float(e0), float(eta), _np_ptr(om),
                            len(om), _np_ptr(out), int(bool(device))))
    return out[0::2] + 1j * out[1::2]


def cf_spectrum(alpha, beta, norm2, e0, eta, omega, elastic=None):
    a = np.ascontiguousarray(alpha, np.float64)
    b = np.ascontiguousarray(beta, np.float64)
    om = np.ascontiguousarray(omega, np.float64)
    out = np.zeros(2 * len(om))
    el = None if elastic is None else np.array([complex(elastic).real, complex(elastic).imag])
    check(lib().esc_cf_spectrum(len(a), _np_ptr(a), _np_ptr(b), float(norm2), float(e0), float(eta), _np_ptr(om),
                                len(om), _np_ptr(el) if el is not None else None, _np_ptr(out)))
    return out[0::2] + 1j * out[1::2]


def lehmann_spectrum(evals, Oa, Ob, eta, omega, elastic=None):
    ev = np.ascontiguousarray(evals, np.float64)
    oa = np.ascontiguousarray(np.column_stack([np.real(Oa), np.imag(Oa)]).ravel())
    ob = np.ascontiguousarray(np.column_stack([np.real(Ob), np.imag(Ob)]).ravel())
    om = np.ascontiguousarray(omega, np.float64)
    out = np.zeros(2 * len(om))
    el = None if elastic is None else np.array([complex(elastic).real, complex(elastic).imag])
    check(lib().esc_lehmann_spectrum(len(ev), _np_ptr(ev), _np_ptr(oa), _np_ptr(ob),
                                     _np_ptr(el) if el is not None else None, float(eta), _np_ptr(om), len(om),
                                     _np_ptr(out)))
    return out[0::2] + 1j * out[1::2]


def load_params(path) -> dict:
    """ConstVariables(inputfile) (src/ParametersEngine.h:53-229) as a dict."""
    p = EscParams()
    check(lib().esc_params_load(str(path).encode(), ctypes.byref(p)))
    out = {}
    for name, _ in EscParams._fields_:
        v = getattr(p, name)
        out[name] = v.decode() if isinstance(v, bytes) else v
    return out


def tenth_omega():
    """The grid of DynSzSz and LocalDynRIXS: w = 0.1 k, k = 0..120 (src/dynamics.h:532-533,587-588)."""
    return 0.1 * np.arange(121)


def reference_omega():
    """omega = domega * om, om = 0..1000, domega = 0.002 (src/dynamics.h:231-232,270-271)."""
    return np.array([0.002 * k for k in range(1001)])
