"""ctypes binding of the C ABI in include/esc.h (libesc.so, built in-tree).

The library is the only compute path: if it is missing, importing anything that
needs it raises immediately; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libesc.so")

ESC_OK = 0
ESC_C128, ESC_C64 = 0, 1
OPS = {"Sx": 0, "Sy": 1, "Sz": 2, "Jsx": 3, "Jsy": 4, "Jsz": 5, "Jex": 6, "Jey": 7}
SITE_TOTAL = -1
LANC_MODES = {"OnlyEgs": 0, "GS": 1, "All": 2, "Plain": 3}
START_KINDS = {"Uniform": 0, "Gaussian": 1}
DYN_KINDS = {"JsJs": 0, "JeJe": 1, "LocalSzSz": 2}
SPEC_METHODS = {"Ritz": 0, "CF": 1, "CFReorth": 2}

# every entry point declared in include/esc.h
EXPORTS = [
    "esc_last_error", "esc_version", "esc_ctx_create", "esc_ctx_destroy", "esc_ctx_info", "esc_hpsi",
    "esc_apply_op", "esc_start_vector", "esc_dot", "esc_axpby", "esc_upload", "esc_download", "esc_lanczos",
    "esc_ritz_overlaps", "esc_ritz_vector", "esc_lanczos_cf", "esc_dynamics", "esc_honeycomb",
    "esc_connectors", "esc_params_load", "esc_tridiag_eig", "esc_cf_spectrum", "esc_lehmann_spectrum",
    "esc_autotune", "esc_set_tile_bits", "esc_kernel_configs", "esc_ground_state",
    "esc_batch_create", "esc_batch_destroy", "esc_batch_ground_state", "esc_batch_state", "esc_batch_dynamics",
    "esc_batch_autotune", "esc_spectrum_grad", "esc_two_point", "esc_chain", "esc_chain_connectors",
    "esc_rixs", "esc_szsz", "esc_exact_diag", "esc_kernel_precompile", "esc_dot_noalloc", "esc_kernel_plan",
    "esc_tridiag_lowest", "esc_cf_eval", "esc_table_grad", "esc_tridiag_eigenpairs",
]


class EscError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"esc error {code}: {msg}")
        self.code = code


class EscModel(ctypes.Structure):
    _fields_ = [
        ("n_sites", ctypes.c_int32),
        ("n_bonds", ctypes.c_int32 * 3),
        ("bonds", ctypes.c_void_p * 3),
        ("neigh", ctypes.c_void_p),
        ("K", ctypes.c_double * 3),
        ("J", ctypes.c_double),
        ("Gamma", ctypes.c_double),
        ("B", ctypes.c_double * 3),
    ]


class EscParams(ctypes.Structure):
    _fields_ = [
        ("Lx", ctypes.c_int32), ("Ly", ctypes.c_int32), ("NumberofSites", ctypes.c_int32),
        ("RandomSeed", ctypes.c_int32), ("LanczosSteps", ctypes.c_int32), ("Threads", ctypes.c_int32),
        ("IsPeriodicX", ctypes.c_int32), ("IsPeriodicY", ctypes.c_int32), ("LocalSkw", ctypes.c_int32),
        ("SpinCurrDynamics", ctypes.c_int32), ("EnergyCurrDynamics", ctypes.c_int32),
        ("Kxx", ctypes.c_double), ("Kyy", ctypes.c_double), ("Kzz", ctypes.c_double),
        ("Bxx", ctypes.c_double), ("Byy", ctypes.c_double), ("Bzz", ctypes.c_double),
        ("Temperature", ctypes.c_double), ("LanczosEps", ctypes.c_double),
        ("Model", ctypes.c_char * 64), ("Solver", ctypes.c_char * 64), ("LancType", ctypes.c_char * 64),
        ("J", ctypes.c_double), ("Gamma", ctypes.c_double), ("Eta", ctypes.c_double),
        ("OmegaMin", ctypes.c_double), ("OmegaMax", ctypes.c_double),
        ("OmegaPoints", ctypes.c_int32), ("Precision", ctypes.c_int32), ("SpectralMode", ctypes.c_int32),
        ("StartVector", ctypes.c_int32), ("CFSteps", ctypes.c_int32), ("LocalSzSz", ctypes.c_int32),
        ("SzSz", ctypes.c_int32), ("LocalRIXS", ctypes.c_int32), ("SpecialSite", ctypes.c_int32),
    ]


_LIB = None


def lib():
    """Load libesc.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m energyspincurrents_amd.build`")
    # torch ships its own HIP runtime (libamdhip64.so, soname libamdhip64.so.7);
    # loading it before libesc.so makes libesc bind to that same runtime, so
    # device pointers and streams from torch are valid in libesc.  Loading
    # libesc first would put two HIP runtimes in one process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P, I, D, I64, U64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64, ctypes.c_uint64
    sig = {
        "esc_last_error": ([], ctypes.c_char_p),
        "esc_version": ([], I),
        "esc_ctx_create": ([P, I, P], I),
        "esc_ctx_destroy": ([P], I),
        "esc_ctx_info": ([P, P, P, P], I),
        "esc_hpsi": ([P, P, P, P], I),
        "esc_apply_op": ([P, I, I, P, P, P], I),
        "esc_start_vector": ([P, U64, I, P, P], I),
        "esc_dot": ([P, P, P, P, P], I),
        "esc_axpby": ([P, P, P, P, P, P], I),
        "esc_upload": ([P, P, P, P], I),
        "esc_download": ([P, P, P, P], I),
        "esc_lanczos": ([P, P, I, D, I, P, P, P, P, P, P, P], I),
        "esc_exact_diag": ([P, P, P, P], I),
        "esc_dot_noalloc": ([P, P, P, P, P], I),
        "esc_kernel_plan": ([P, I, I, P, I, P, P, P, I, P], I),
        "esc_kernel_precompile": ([P, I, I, P, ctypes.c_size_t, P], I),
        "esc_ritz_overlaps": ([P, P, I, P, P], I),
        "esc_ritz_vector": ([P, I, P, P], I),
        "esc_lanczos_cf": ([P, P, I, P, P, P, P, P], I),
        "esc_ground_state": ([P, P, I, D, I, P, P, P, P, P], I),
        "esc_batch_create": ([I, P, I, P], I),
        "esc_batch_destroy": ([P], I),
        "esc_batch_ground_state": ([P, U64, I, I, P, P, P, P], I),
        "esc_batch_state": ([P, I, P, P], I),
        "esc_batch_autotune": ([P, I, P], I),
        "esc_spectrum_grad": ([P, I, I, I, I, I, D, P, I, P, P, P, P, P, P, P], I),
        "esc_table_grad": ([P, I, I, I, I, D, P, I, P, P, P, P, P, P, P], I),
        "esc_two_point": ([P, I, I, P, P, P], I),
        "esc_chain": ([I, I, P], I),
        "esc_rixs": ([P, P, I, I, P, D, I, D, D, P, I, P, P], I),
        "esc_szsz": ([P, P, D, P, I, P, P], I),
        "esc_chain_connectors": ([I, P, P, P, P, P, P], I),
        "esc_batch_dynamics": ([P, I, I, D, P, I, I, P, P], I),
        "esc_dynamics": ([P, I, I, P, D, I, D, P, I, P, P, P], I),
        "esc_honeycomb": ([I, I, I, I, P], I),
        "esc_connectors": ([I, P, P, P, P, P, P], I),
        "esc_params_load": ([ctypes.c_char_p, P], I),
        "esc_tridiag_eig": ([I, P, P, P, P], I),
        "esc_cf_spectrum": ([I, P, P, D, D, D, P, I, P, P], I),
        "esc_tridiag_lowest": ([I, P, P, P, P, I], I),
        "esc_tridiag_eigenpairs": ([I, P, P, P, P, I], I),
        "esc_cf_eval": ([I, P, P, D, D, D, P, I, P, I], I),
        "esc_lehmann_spectrum": ([I, P, P, P, P, D, P, I, P], I),
        "esc_autotune": ([P, P, P, I, P, P, P], I),
        "esc_set_tile_bits": ([P, I], I),
        "esc_kernel_configs": ([P, P, I, P], I),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _LIB = L
    return L


def check(rc):
    if rc != ESC_OK:
        raise EscError(rc, lib().esc_last_error().decode(errors="replace"))
    return rc
