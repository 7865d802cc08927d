"""Build recipe for the native parts of energyspincurrents_amd.

Produces, in-tree:
  energyspincurrents_amd/libesc.so   HIP kernels + C ABI (include/esc.h), gfx950
  energyspincurrents_amd/LanczosCond drop-in CLI of the reference driver
  oracle/libesc_oracle.so            CPU oracle (test infrastructure, built only
                                     so that tests and the baseline can load it)

Usage: python -m energyspincurrents_amd.build  [--no-cli] [--no-oracle]
"""
from __future__ import annotations

import math
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libesc.so")
CLI = os.path.join(PKG, "LanczosCond")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for every device pass.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", f"-I{INC}", f"-I{CSRC}",
          "-Wno-unused-command-line-argument", "-pthread"]

LIB_SOURCES = ["esc_kernels.hip", "esc_capi.cpp", "esc_plan.cpp", "esc_spectra.cpp", "esc_batch.cpp", "esc_grad.cpp",
               "esc_host.cpp",
               "esc_cube.cpp"]
CLI_SOURCES = ["facade/lanczoscond_main.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib():
    objs = []
    for src in LIB_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(CSRC, os.path.splitext(os.path.basename(src))[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *COMMON, *DEVICE_FLAGS, *lang, "-c", path, "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, *COMMON, "-shared", *objs, "-o", tmp, "-lrocsolver", "-lrocblas", "-lhiprtc", "-ldl"])
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_cli():
    srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return None
    tmp = CLI + ".tmp"
    _run([HIPCC, "-O2", "-std=c++17", f"-I{INC}", f"-I{CSRC}", f"-I{os.path.join(CSRC, 'facade')}",
          *srcs, "-o", tmp, f"-L{PKG}", "-lesc", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI)
    return CLI


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def cube_candidates(n_sites, c64=True):
    """The generated-kernel configurations a context of n_sites may try (the
    list of cube_candidates in esc_capi.cpp): 2^12- and 2^13-row tiles, 2^3
    or 2^4 rows per lane, 4 or 8 waves, both window orders; complex64 also
    over row pairs (layout bit 0: 10000000 + ...); from 25 sites also with
    lanes on six contiguous sites, without row pairs (layout bit 1:
    20000000 + ...)."""
    w = n_sites if n_sites <= 18 else (17 if n_sites <= 22 else 18)
    out = []
    for tb in (12, 13):
        if tb > w:
            continue
        for k in (3, 4):
            for v in (2, 3):
                for P in (7, 8):
                    for paired in ((0, 1) if c64 else (0,)):
                        for contiguous in ((0, 2) if n_sites >= 25 and not paired else (0,)):
                            if w == n_sites and P == 8:
                                continue
                            out.append(10000000 * (paired | contiguous) + 1000000 * k + 100000 * P + 10000 * v
                                       + 100 * tb + w)
    return out


def warm_jit_cache():
    """Compile the generated H.psi kernels of the benchmark lattices into
    jit_cache/ beside libesc.so, so that contexts of these models load them
    instead of compiling (the couplings are not part of the cache key): the
    default configurations of the single-system workloads, the lane layouts
    the 24-site GPU tests select, and every
    candidate of the lattices that bench.py autotunes (the batched scan's
    16-, 18-, 20- and 24-site clusters, the 32-site cluster)."""
    sys.path.insert(0, ROOT)
    from concurrent.futures import ProcessPoolExecutor
    import energyspincurrents_amd as esc
    kjgh = dict(K=(-1.0, -1.0, -1.0), J=0.1, Gamma=0.2, B=esc.field111(0.2))
    # the batched scan's term structure: K = cos(phi), J = sin(phi), field
    # along [111] (no Gamma; the diagonal summed bond by bond)
    scan = dict(K=(math.cos(1.0),) * 3, J=math.sin(1.0), B=esc.field111(0.3))
    n32 = dict(K=(-1.0, -1.0, -1.0), B=esc.field111(0.3))
    jobs = [((3, 8), kjgh, "c128", None), ((3, 8), kjgh, "c64", None), ((2, 10), kjgh, "c128", None),
            ((2, 10), kjgh, "c64", None)]
    # the lane layouts the 24-site GPU tests set explicitly
    jobs += [((3, 8), kjgh, "c128", cfg) for cfg in (23731218, 23831318)]
    jobs += [((3, 8), kjgh, "c64", cfg) for cfg in (24831318, 34831318, 4831318)]
    for (lx, ly), kw in (((2, 8), scan), ((3, 6), scan), ((2, 10), scan), ((3, 8), scan), ((4, 8), n32)):
        jobs += [((lx, ly), kw, "c64", cfg) for cfg in cube_candidates(lx * ly)]

    # the run-time compiler serialises within a process: one process per job
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        done = list(ex.map(_precompile_job, jobs))
    print(f"+ generated H.psi kernels: {sum(done)} configurations of the benchmark lattices", flush=True)


def _precompile_job(job):
    """One generated-kernel configuration into the disk cache; 1 if it applies
    to the lattice, 0 if not (ESC_ERR_UNSUPPORTED)."""
    sys.path.insert(0, ROOT)
    import energyspincurrents_amd as esc
    (lx, ly), kw, prec, cfg = job
    try:
        esc.precompile_kernel(esc.SpinModel(esc.Honeycomb(lx, ly), **kw), prec, cfg)
    except esc.EscError as e:
        if e.code != -6:   # -6: the configuration does not apply to this lattice
            raise
        return 0
    return 1


def build(cli=True, oracle=True):
    build_lib()
    if cli:
        build_cli()
    if oracle:
        build_oracle()
    warm_jit_cache()


if __name__ == "__main__":
    build(cli="--no-cli" not in sys.argv, oracle="--no-oracle" not in sys.argv)
