"""Generated H.psi kernels (esc_cube.cpp): source generation and run-time
compilation for gfx950 on the host (no device), through the C ABI.

Checks: every default configuration compiles for the benchmark lattices in
both precisions; a generated pass applies each term exactly once, and the
passes together apply every bond and flip of the model once (counted from
the source's own term comments); models the generated kernels do not cover
(the chain's XYZ pairs) are refused with ESC_ERR_UNSUPPORTED; a diagonal whose
closed form would round differently is summed bond by bond; no packed-FP32
instruction appears in the code object.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O

esc = pytest.importorskip("energyspincurrents_amd")


@pytest.fixture(autouse=True)
def jit_cache(tmp_path, monkeypatch):
    monkeypatch.setenv("ESC_JIT_CACHE", str(tmp_path / "jit"))
    return tmp_path / "jit"


def model(lx, ly, **kw):
    return esc.SpinModel(esc.Honeycomb(lx, ly), **kw)


KJGH = dict(K=(-1.0, -1.0, -1.0), J=0.1, Gamma=0.2, B=esc.field111(0.2))


def terms_in(src):
    """(kind, sites) of every term comment of the generated passes."""
    out = []
    for m in re.finditer(r"// (bond (\d+)-(\d+) family (\d)|flip (\d+))\n", src):
        out.append(("b", int(m.group(2)), int(m.group(3))) if m.group(2) else ("f", int(m.group(5))))
    return out


@pytest.mark.parametrize("lx,ly,prec", [(3, 8, "c128"), (3, 8, "c64"), (2, 10, "c128"), (2, 10, "c64"),
                                        (4, 8, "c64")])
def test_default_configuration_compiles_and_covers_every_term(lx, ly, prec, jit_cache):
    m = model(lx, ly, **KJGH)
    src = esc.precompile_kernel(m, prec)
    assert os.listdir(jit_cache), "code object not cached"
    terms = terms_in(src)
    bonds = [tuple(sorted(b)) for fam in m.lattice.bonds() for b in fam.tolist()]
    want = sorted([("b",) + b for b in bonds] + [("f", k) for k in range(m.lattice.N)])
    assert sorted(terms) == want
    assert src.count('extern "C" __global__') >= 2


def test_pure_kitaev_has_no_flip_terms_and_compiles():
    src = esc.precompile_kernel(model(3, 8, K=(-1.0, -1.0, -1.0)), "c128")
    assert all(t[0] == "b" for t in terms_in(src))


def test_unsupported_models_are_refused():
    with pytest.raises(esc.EscError) as e:
        esc.precompile_kernel(esc.SpinModel(esc.Chain(20), K=(1.0, 0.5, 0.25)), "c128")
    assert e.value.code == -6


def test_inexact_closed_form_diagonal_is_summed_bond_by_bond():
    """Kz = 0.3: the partial sums of +-Kz/4 are not exact, so the generated
    diagonal adds the z bonds, then the Heisenberg bonds, one by one in the
    reference's order (10 + 30 at 20 sites); Kz = -1 keeps the closed form."""
    src = esc.precompile_kernel(model(2, 10, K=(-1.0, 0.7, 0.3), J=0.1), "c128")
    body = src[src.index("esc_diag("):src.index("}", src.index("esc_diag("))]
    assert body.count("? -kz4 : kz4") == 10 and body.count("? -j4 : j4") == 30
    assert "esc_diag_c" not in src
    src = esc.precompile_kernel(model(2, 10, K=(-1.0, -1.0, -1.0), J=0.1), "c128")
    assert "? -kz4 : kz4" not in src and "esc_diag(" not in src
    # closed form: per-lane words of (R0 ^ (R0 >> d)) & m, one popcount per
    # word and row of the lane's sub-cube
    assert "esc_diag_c(" in src and "const unsigned zw0 = 0u | ((R0 ^ (R0 >>" in src


def test_closed_form_diagonal_counts_match_bond_by_bond():
    """The generated closed form counts anti-aligned bonds per row from words
    formed once per lane, (R0 ^ (R0 >> d)) & m, xor a constant of the slot;
    evaluate those emitted expressions on the host for random rows and compare
    with the bond-by-bond counts of the lattice."""
    mdl = model(3, 8, K=(-1.0, -1.0, -1.0), J=0.1, B=esc.field111(0.2))
    src = esc.precompile_kernel(mdl, "c128", 3731218)
    last = src[src.rindex("// pass "):]
    words = dict(re.findall(r"const unsigned ((?:zw|jw)\d+) = (0u(?: \| \(\(R0 \^ \(R0 >> \d+\)\) & 0x[0-9a-f]+u\))+);", last))
    calls = re.findall(r"esc_diag_c\(([^,]+), ([^,]+), pc0 \+ (\d+)u,", last)
    slot_rows = [int(x, 16) for x in re.findall(r"out \+ \(base \| grw \| 0x([0-9a-f]+)u \| lrow\)\)\); ar\[\d+\] = v\.x", last)]
    assert len(calls) == len(slot_rows) == 8 and words
    om = O.Model(3, 8, K=(-1.0, -1.0, -1.0), J=0.1, B=O.field111(0.2))
    px, py, pz = om.bonds            # families x, y, z (src/Hamiltonian.h:230-328)
    bonds_z = [tuple(b) for b in pz]
    bonds_all = [tuple(b) for fam in (px, py, pz) for b in fam]
    assert len(bonds_z) == 12 and len(bonds_all) == 36
    rng = np.random.default_rng(5)
    free = 0xffffff
    for X in slot_rows:
        free &= ~X
    def ev(expr, R0):
        return eval(expr.replace("u", "").replace("R0", str(R0)))
    for _ in range(200):
        R0 = int(rng.integers(0, 1 << 24)) & free
        vals = {k: ev(e, R0) for k, e in words.items()}
        for (cz_e, cj_e, pcd), X in zip(calls, slot_rows):
            r = R0 | X
            def cnt(e):
                tot = 0
                for w, c in re.findall(r"__popc\(((?:zw|jw)\d+) \^ 0x([0-9a-f]+)u\)", e):
                    tot += bin(vals[w] ^ int(c, 16)).count("1")
                return tot
            assert cnt(cz_e) == sum(((r >> i) ^ (r >> j)) & 1 for i, j in bonds_z)
            assert cnt(cj_e) == sum(((r >> i) ^ (r >> j)) & 1 for i, j in bonds_all)
            assert bin(R0).count("1") + int(pcd) == bin(r).count("1")


def test_code_object_has_no_packed_fp32(jit_cache):
    # a term structure no other test compiles, so the code object is built
    # here and lands in this test's cache directory
    esc.precompile_kernel(model(3, 6, K=(-1.0, -1.0, -1.0), J=0.1, B=(0.1, 0.0, 0.2)), "c64")
    objdump = "/opt/rocm/llvm/bin/llvm-objdump"
    if not os.path.exists(objdump):
        pytest.skip("llvm-objdump not available")
    for f in os.listdir(jit_cache):
        asm = subprocess.run([objdump, "-d", "--mcpu=gfx950", str(jit_cache / f)], capture_output=True,
                             text=True).stdout
        assert "v_fma_f32" in asm or "v_fmac_f32" in asm
        assert "v_pk_fma_f32" not in asm and "v_pk_add_f32" not in asm and "v_pk_mul_f32" not in asm
