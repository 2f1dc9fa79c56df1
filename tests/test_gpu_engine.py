"""GPU parity of the full nonlocal-LBM step against the FP64 oracle.

Tolerance (from the north-star accuracy bar): relative L2 of the density
deviation rho - rho0 and of the velocity <= 1e-5 after the run, FP32 GPU
state versus the FP64 oracle of the same method on the same nodes.
"""
import numpy as np
import pytest

from tests.scenes import gpu_engine, grid_nodes, oracle_engine, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def cfg1(oracle64):
    """Config 1: D2Q9, unit square, h = 0.01, jitter 0.25h, k = 16, r^3 RBF,
    PML 10h with sigma_max 0.5, pulse 1e-3 / 0.05 at the centre."""
    h = 0.01
    nodes = grid_nodes(oracle64, 2, h, layer=10 * h)
    gpu = gpu_engine(nodes, 9, 16, h)
    ora = oracle_engine(oracle64, nodes, 9, 16, h)
    gpu.set_initial_pulse([0.5, 0.5], 1e-3, 0.05)
    ora.set_pulse([0.5, 0.5], 1e-3, 0.05)
    return dict(gpu=gpu, ora=ora, nodes=nodes, h=h)


def test_cfg1_setup_products(cfg1):
    g = cfg1["gpu"].setup_products()
    o = cfg1["ora"].setup()
    assert np.array_equal(g["idx"], o["idx"])
    assert np.abs(g["W"] - o["W"]).max() < 1e-6 * np.abs(o["W"]).max()
    assert np.abs(g["filt"] - o["filt"]).max() < 1e-6
    assert np.array_equal(g["bidx"], o["bidx"])
    assert np.abs(g["bw"] - o["bw"]).max() < 1e-6 * max(1.0, np.abs(o["bw"]).max())


def test_cfg1_initial_state(cfg1):
    g = cfg1["gpu"].macroscopic(fresh=True)
    o = cfg1["ora"].get(fresh=True)
    assert rel_l2(g["drho"], o["drho"]) < 1e-7


@pytest.mark.parametrize("nsteps", [1, 20, 100])
def test_cfg1_steps(cfg1, nsteps):
    gpu, ora = cfg1["gpu"], cfg1["ora"]
    done = int(round(ora.time / 0.004))
    gpu.step_forward(nsteps - done)
    ora.step(nsteps - done)
    g = gpu.macroscopic(fresh=True)
    o = ora.get(fresh=True)
    err_rho = rel_l2(g["drho"], o["drho"])
    err_u = rel_l2(g["u"], o["u"])
    print(f"steps={nsteps} rel_l2 drho={err_rho:.3e} u={err_u:.3e}")
    assert err_rho < TOL and err_u < TOL


def test_cfg1_pml_state(cfg1):
    gpu, ora = cfg1["gpu"], cfg1["ora"]
    e_g, q_g = gpu.pml_state()
    q1, q2 = ora.pml()
    n = cfg1["nodes"]
    lay = slice(n["n_bulk"], n["n_bulk"] + n["n_layer"])
    w = np.array([4 / 9] + [1 / 36, 1 / 9] * 4)
    e_o = q1 - w[:, None]
    assert rel_l2(e_g[:, lay], e_o[:, lay]) < 1e-4
    assert rel_l2(q_g[:, lay], q2[:, lay]) < 1e-4


def test_3d_small_steps(oracle64):
    """D3Q27 on a jittered 21^3 unit-cube grid (h = 0.05), k = 32, PML 3h,
    off-centre pulse; 10 steps against the FP64 oracle."""
    h = 0.05
    nodes = grid_nodes(oracle64, 3, h, layer=3 * h)
    gpu = gpu_engine(nodes, 27, 32, h)
    ora = oracle_engine(oracle64, nodes, 27, 32, h)
    gpu.set_initial_pulse([0.4, 0.5, 0.5], 1e-3, 0.15)
    ora.set_pulse([0.4, 0.5, 0.5], 1e-3, 0.15)
    s = gpu.setup_products()
    so = ora.setup()
    assert np.array_equal(s["idx"], so["idx"])
    assert np.array_equal(s["bidx"], so["bidx"])
    gpu.step_forward(10)
    ora.step(10)
    g = gpu.macroscopic(fresh=True)
    o = ora.get(fresh=True)
    err_rho = rel_l2(g["drho"], o["drho"])
    err_u = rel_l2(g["u"], o["u"])
    print(f"3d rel_l2 drho={err_rho:.3e} u={err_u:.3e}")
    assert err_rho < TOL and err_u < TOL
