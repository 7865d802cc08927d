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

LIB_SOURCES = ["esc_kernels.hip", "esc_capi.cpp", "esc_batch.cpp", "esc_grad.cpp", "esc_host.cpp", "esc_cube.cpp"]
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


def warm_jit_cache():
    """Compile the generated H.psi kernels of the benchmark lattices into
    jit_cache/ beside libesc.so, so that contexts of these models load them
    instead of compiling (the couplings are not part of the cache key)."""
    sys.path.insert(0, ROOT)
    import energyspincurrents_amd as esc
    kjgh = dict(K=(-1.0, -1.0, -1.0), J=0.1, Gamma=0.2, B=esc.field111(0.2))
    # the batched scan's term structure: K = cos(phi), J = sin(phi), field
    # along [111] (no Gamma; the diagonal summed bond by bond)
    scan = dict(K=(math.cos(1.0),) * 3, J=math.sin(1.0), B=esc.field111(0.3))
    jobs = [((3, 8), kjgh, "c128"), ((3, 8), kjgh, "c64"), ((2, 10), kjgh, "c128"), ((2, 10), kjgh, "c64"),
            ((4, 8), dict(K=(-1.0, -1.0, -1.0), B=esc.field111(0.3)), "c64"), ((2, 10), scan, "c64"),
            ((3, 8), scan, "c64")]
    for (lx, ly), kw, prec in jobs:
        esc.precompile_kernel(esc.SpinModel(esc.Honeycomb(lx, ly), **kw), prec)
        print(f"+ generated H.psi kernels: {lx * ly} sites, {prec}", flush=True)


def build(cli=True, oracle=True):
    build_lib()
    if cli:
        build_cli()
    if oracle:
        build_oracle()
    warm_jit_cache()


if __name__ == "__main__":
    build(cli="--no-cli" not in sys.argv, oracle="--no-oracle" not in sys.argv)
