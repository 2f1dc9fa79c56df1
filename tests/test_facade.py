"""The naga:: C++ facade over the C-ABI: builds a small C++ program against
naga_amd/naga.hpp and libnaga_amd.so and runs it."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "naga_amd")


@pytest.fixture(scope="module")
def facade_exe(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libnaga_amd.so")):
        pytest.fail("libnaga_amd.so is not built")
    exe = str(tmp_path_factory.mktemp("facade") / "facade_main")
    cmd = ["g++", "-std=c++17", "-O1", "-o", exe,
           os.path.join(ROOT, "tests", "cpp", "facade_main.cpp"),
           "-I", os.path.join(ROOT, "include"), "-I", LIBDIR,
           "-L", LIBDIR, "-lnaga_amd", f"-Wl,-rpath,{LIBDIR}"]
    subprocess.run(cmd, check=True)
    return exe


def test_facade_error_mapping(facade_exe):
    r = subprocess.run([facade_exe, "errors"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "errors ok" in r.stdout


@pytest.mark.gpu
def test_facade_matches_python_mirror(facade_exe):
    r = subprocess.run([facade_exe, "run", "10"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = dict(zip(r.stdout.split()[0::2], r.stdout.split()[1::2]))
    import naga_amd
    h = 0.02
    nodes = naga_amd.UniformGridProvider([0, 0], [1, 1], h, 0.1, 0.5, jitter=0.25).get()
    p = naga_amd.problem_parameters(9, 0.6, 0.4 * h)
    eng = naga_amd.SimulationEngine(lattice=9, k=32, rbf=naga_amd.RBF_POLYHARMONIC3,
                                    rbf_boundary=naga_amd.RBF_MULTIQUADRIC)
    eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
    eng.init_domain(nodes)
    eng.set_initial_pulse([0.5, 0.5], 1e-3, 0.05)
    eng.step_forward(10)
    m = eng.macroscopic(fresh=False)
    assert int(vals["nodes"]) == eng.num_nodes
    ref = float(np.sum(m["rho"].astype(np.float64) - 1.0))
    assert abs(float(vals["sum_drho"]) - ref) <= 1e-6 * max(1.0, abs(ref))
    assert abs(float(vals["max_u"]) - float(np.abs(m["u"]).max())) <= 1e-7


@pytest.mark.gpu
def test_facade_state_source_observer(facade_exe):
    r = subprocess.run([facade_exe, "extras"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "extras ok" in r.stdout


@pytest.mark.gpu
def test_facade_advection_and_csr(facade_exe):
    r = subprocess.run([facade_exe, "operators"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "operators ok" in r.stdout


@pytest.mark.gpu
def test_facade_reference_named_entry_points(facade_exe, tmp_path):
    r = subprocess.run([facade_exe, "dropin", str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "dropin ok" in r.stdout
    assert sorted(p.name for p in tmp_path.iterdir()) == ["result.0.vtp", "result.1.vtp"]
