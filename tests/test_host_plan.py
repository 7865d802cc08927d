"""Pass planners of the tiled H.psi kernels (generic build_passes and the
generated sub-cube plans), checked on the host through esc_kernel_plan for
many lattices and configurations.

Invariants: every off-diagonal term is applied by exactly one pass; its mask
lies inside that pass's window; a pass's near terms are exactly its terms
inside its tile; tiles hold tile_bits sites, among them the six lane sites 0..5
(generic) or the sites below and of the lanes' run (generated: 0..2 or 0..3
with a 64-byte run, 0..5 or 0..6 with contiguous lanes), and lie in their
windows; windows hold w sites (generic) or at most w (generated); the
LDS of a workgroup fits the CU; the first pass, and only it, adds the
diagonal.
"""
import numpy as np
import pytest

esc = pytest.importorskip("energyspincurrents_amd")

KJGH = dict(K=(-1.0, -1.0, -1.0), J=0.1, Gamma=0.2, B=esc.field111(0.2))


def models():
    out = []
    for lx, ly in ((2, 8), (3, 6), (2, 10), (3, 8), (4, 8)):
        out.append((f"honeycomb {lx}x{ly} KJGh", esc.SpinModel(esc.Honeycomb(lx, ly), **KJGH)))
    out.append(("honeycomb 3x8 pure Kitaev", esc.SpinModel(esc.Honeycomb(3, 8), K=(-1.0, -1.0, -1.0))))
    out.append(("honeycomb 2x10 core hole at 3", esc.SpinModel(esc.Honeycomb(2, 10), core_hole=3, **KJGH)))
    out.append(("chain 20 XYZ", esc.SpinModel(esc.Chain(20), K=(1.0, 0.5, 0.25), B=(0.1, 0.0, 0.2))))
    return out


def configs(N, c64):
    out = []
    for P in (1, 2, 3, 5, 6):
        for v, waves in ((2, 4), (3, 8), (4, 16)):
            for tb in (10, 11, 12, 13):
                for w in (N, 16, 17, 18, 19):
                    ww = min(w, N)
                    if (ww == N) != (P == 1):
                        continue
                    out.append(100000 * P + 10000 * v + 100 * tb + ww)
    # generated-kernel plans: the full grid up to 26 sites, k = 3 with 8 waves
    # beyond (each plan there runs ten local searches)
    for P in (7, 8):
        for k in ((3, 4) if N <= 26 else (3,)):
            for v in ((2, 3) if N <= 26 else (3,)):
                for tb in (12, 13):
                    for w in (17, 18):
                        # layout digit: bit 0 over row pairs, bit 1 contiguous lanes
                        for lay in ((0, 1, 2, 3) if c64 else (0, 2)):
                            out.append(10000000 * lay + 1000000 * k + 100000 * P + 10000 * v + 100 * tb + min(w, N))
    return sorted(set(out))


@pytest.mark.parametrize("name,model", models(), ids=[m[0] for m in models()])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_plan_invariants(name, model, prec):
    N = model.lattice.N
    esz = 8 if prec == "c64" else 16
    ok = 0
    for cfg in configs(N, prec == "c64"):
        try:
            passes, terms = esc.kernel_plan(model, prec, cfg)
        except esc.EscError as e:
            assert e.code == -6, (cfg, str(e))       # shape not available for this model: refused cleanly
            continue
        ok += 1
        cube = (cfg // 100000) % 10 in (7, 8)
        tb = (cfg // 100) % 100
        w = cfg % 100
        assert len(passes) >= 1
        # the generic tiled kernel adds the diagonal in its first pass, the
        # generated sub-cube kernels in their last (with the output re-read)
        want = [0] * (len(passes) - 1) + [1] if cube else [1] + [0] * (len(passes) - 1)
        assert [p["diag"] for p in passes] == want, cfg
        # generic plans hold the six lane sites 0..5; generated plans the
        # sites below and of the lanes' run: six contiguous lane sites (0..5,
        # 0..6 with row pairs) or a 64-byte run and the next site (0..2 with
        # 16-byte rows, 0..3 with 8-byte rows or row pairs)
        lay = cfg // 10000000
        if not cube:
            fixed = 63
        elif lay & 2:
            fixed = 127 if lay & 1 else 63
        else:
            fixed = 7 if (prec == "c128" and not lay & 1) else 15
        for p in passes:
            W, T = p["window"], p["tile"]
            assert T & fixed == fixed and bin(T).count("1") == tb, (cfg, p)
            assert T & ~W == 0, (cfg, p)
            assert (bin(W).count("1") <= w) if cube else (bin(W).count("1") == w), (cfg, p)
            assert p["lds_bytes"] <= (131072 if cube else 160 * 1024 - 1024), (cfg, p)
            if cube:
                assert p["lds_bytes"] == esz << tb
        for mask, q in terms:
            assert 0 <= q < len(passes), (cfg, hex(mask))
            assert mask & ~passes[q]["window"] == 0, (cfg, hex(mask))
        for qi, p in enumerate(passes):
            mine = [m for m, q in terms if q == qi]
            assert p["near"] + p["far"] == len(mine), cfg
            assert sum(1 for m in mine if m & ~p["tile"] == 0) == p["near"], cfg
    assert ok >= 3, f"{name}: only {ok} configurations planned"
