"""bench.py's command-line contract and the agreement of its constants with
the committed profile summaries (CPU only: nothing here touches a GPU)."""
import importlib.util
import os
import subprocess
import sys

from conftest import ROOT


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_before_any_gpu_use():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_documented_flags_parse():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--no-cpu", "--no-spectrum", "--no-n32", "--no-batch",
                 "--kernel-config", "--autotune", "--pipelines", "--no-cpu-coo", "--jeje-e2e"):
        assert flag in r.stdout, flag


def test_profiled_configurations_and_traffic_match_the_profiles():
    """The configurations bench.py times are the ones the committed kernel
    trace and traffic summaries cover, and roofline.traffic is (FETCH KiB x 2.0
    + WRITE KiB) x 1024 of the named summary row, both passes."""
    b = load_bench()
    cfgs = b.PROFILED_CONFIGS
    assert set(b.MEASURED_TRAFFIC) == {(p, c) for p in cfgs for c in cfgs[p]}
    trace = open(os.path.join(ROOT, "profiles", "hpsi_cube_n24_kernel_trace.md")).read()
    for prec in cfgs:
        for c in cfgs[prec]:
            assert f"| {prec} | {c} |" in trace
    for (prec, cfg), (total, source) in b.MEASURED_TRAFFIC.items():
        path, row = source.split(": ", 1)
        lines = open(os.path.join(ROOT, path)).read().splitlines()
        i = next(n for n, ln in enumerate(lines) if ln.startswith(f"| {row} |"))
        assert prec in row and str(cfg) in row
        cells = [[c.strip() for c in ln.strip().strip("|").split("|")] for ln in lines[i:i + 2]]
        assert [c[1] for c in cells] == ["1", "2"]
        fetch = sum(int(c[2]) for c in cells)
        write = sum(int(c[4]) for c in cells)
        assert total == (2 * fetch + write) * 1024


def test_host_jeje_pipeline_equals_exact_lehmann_sum():
    """The host leg of the end-to-end JeJe timing (plain Lanczos ground state
    with replay, four continued fractions, cross terms by polarisation) against
    DynJeJe over the exact eigenstates of a dense 8-site Hamiltonian."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    b = load_bench()
    m = O.Model(2, 4, K=(-1.0, -1.0, -1.0), B=O.field111(0.2))
    e, U = np.linalg.eigh(O.hamiltonian_dense_kron(m))
    w = np.linspace(0.0, 6.0, 400)
    exact = O.jeje_ref(m, e, U.T.copy(), U[:, 0], omega=w, eta=0.1)
    rng = np.random.default_rng(3)
    v0 = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    e0, tab = b.cpu_jeje_pipeline(m, v0, 60, 150, w, 0.1)
    assert abs(e0 - e[0]) < 1e-12
    assert np.abs(tab[:, 1:] - exact[:, 1:]).max() < 1e-10 * np.abs(exact[:, 1:]).max()
