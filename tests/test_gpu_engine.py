"""GPU parity of the full nonlocal-LBM step against the FP64 oracle.

Tolerance: relative L2 of the density
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
    """The 2D benchmark scene: D2Q9, unit square, h = 0.01, jitter 0.25h, k = 16, r^3 RBF,
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


@pytest.mark.parametrize("nsteps", [1, 20, 100, 1000])
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


def test_3d_200_steps(oracle64):
    """The 3D configuration's physics and run length (D3Q27, k = 32, r^3
    RBF, PML 8h, tau 0.6, dt 0.4h, pulse 1e-3 / 0.05 at (0.3, 0.5, 0.5),
    200 steps) on the h = 0.02 cube (132,651 nodes) against the FP64
    oracle."""
    import os
    h = 0.02
    nodes = grid_nodes(oracle64, 3, h, layer=8 * h)
    gpu = gpu_engine(nodes, 27, 32, h)
    ora = oracle64.engine(nodes, Q=27, k=32, k_boundary=32, rbf=1, rbf_boundary=1, h=h,
                          threads=min(16, os.cpu_count() or 1))
    gpu.set_initial_pulse([0.3, 0.5, 0.5], 1e-3, 0.05)
    ora.set_pulse([0.3, 0.5, 0.5], 1e-3, 0.05)
    gpu.step_forward(200)
    ora.step(200)
    g, o = gpu.macroscopic(fresh=True), ora.get(fresh=True)
    err_rho, err_u = rel_l2(g["drho"], o["drho"]), rel_l2(g["u"], o["u"])
    print(f"3d 200 steps rel_l2 drho={err_rho:.3e} u={err_u:.3e}")
    assert err_rho < TOL and err_u < TOL


def test_density_source_parity(oracle64):
    """Density sources (two overlapping sinusoidal sources and a one-shot
    one) match the oracle over 40 steps."""
    h = 0.02
    nodes = grid_nodes(oracle64, 2, h, layer=5 * h)
    gpu = gpu_engine(nodes, 9, 16, h)
    ora = oracle_engine(oracle64, nodes, 9, 16, h)
    gpu.reset()
    ora.set_pulse([0.5, 0.5], 0.0, 1.0)
    import naga_amd
    dt = 0.4 * h
    srcs = [([0.45, 0.5], 0.12, 2e-3, 4.0, 0.3, 0.0, float("inf")),
            ([0.55, 0.52], 0.08, -1e-3, 7.0, 0.0, 3 * dt, 30 * dt),
            ([0.3, 0.3], 0.1, 1e-3, 0.0, 0.0, 0.0, dt)]
    for c, R, A, f, ph, t0, t1 in srcs:
        gpu.register_density_source(naga_amd.DensitySource(c, R, A, f, ph, t0, t1))
        ora.add_density_source(c, R, A, f, ph, t0, t1)
    gpu.step_forward(40)
    ora.step(40)
    g = gpu.macroscopic(fresh=True)
    o = ora.get(fresh=True)
    err_rho, err_u = rel_l2(g["drho"], o["drho"]), rel_l2(g["u"], o["u"])
    print(f"rel_l2 drho={err_rho:.3e} u={err_u:.3e}")
    assert err_rho < TOL and err_u < TOL
    gh, oh = gpu.macroscopic(fresh=False), ora.get(fresh=False)
    assert np.abs(gh["rho"] - oh["rho"]).max() < 1e-6


def test_3d_density_source(oracle64):
    h = 0.05
    nodes = grid_nodes(oracle64, 3, h, layer=2 * h)
    gpu = gpu_engine(nodes, 27, 32, h)
    ora = oracle_engine(oracle64, nodes, 27, 32, h)
    gpu.reset()
    ora.set_pulse([0.5, 0.5, 0.5], 0.0, 1.0)
    import naga_amd
    gpu.register_density_source(naga_amd.DensitySource([0.5, 0.45, 0.5], 0.2, 1e-3, 3.0))
    ora.add_density_source([0.5, 0.45, 0.5], 0.2, 1e-3, 3.0)
    gpu.step_forward(10)
    ora.step(10)
    g, o = gpu.macroscopic(fresh=True), ora.get(fresh=True)
    assert rel_l2(g["drho"], o["drho"]) < TOL and rel_l2(g["u"], o["u"]) < TOL


@pytest.mark.parametrize("Q", [9, 27])
def test_geometric_pml_indexing(oracle64, Q):
    """pml_ref_index = 0 (the layer field reads Q at j) matches the oracle."""
    D = 2 if Q == 9 else 3
    h = 0.02 if D == 2 else 0.05
    k = 16 if D == 2 else 32
    nodes = grid_nodes(oracle64, D, h, layer=4 * h)
    gpu = gpu_engine(nodes, Q, k, h, pml_ref_index=False)
    ora = oracle_engine(oracle64, nodes, Q, k, h, pml_ref_index=False)
    c = [0.4] + [0.5] * (D - 1)
    gpu.set_initial_pulse(c, 1e-3, 0.1)
    ora.set_pulse(c, 1e-3, 0.1)
    gpu.step_forward(15)
    ora.step(15)
    g, o = gpu.macroscopic(fresh=True), ora.get(fresh=True)
    assert rel_l2(g["drho"], o["drho"]) < TOL and rel_l2(g["u"], o["u"]) < TOL


def _mirror_check(nodes, mirror, Q, k, h, steps):
    gpu = gpu_engine(nodes, Q, k, h, pml_ref_index=False)
    D = 2 if Q == 9 else 3
    lo = nodes["points"].min(0).astype(np.float64)
    hi = nodes["points"].max(0).astype(np.float64)
    c = 0.5 * (lo + hi)
    c[0] = 0.0
    gpu.set_initial_pulse(c, 1e-3, 0.1 * (hi[1] - lo[1]))
    gpu.step_forward(steps)
    m = gpu.macroscopic(fresh=True)
    flip = np.ones((D, 1))
    flip[0] = -1
    drho, u = m["drho"], m["u"].astype(np.float64)
    e_rho = np.abs(drho - drho[mirror]).max() / np.abs(drho).max()
    e_u = np.abs(u - flip * u[:, mirror]).max() / np.abs(u).max()
    return e_rho, e_u


def test_mirror_symmetry_small(oracle64):
    from tests.scenes import mirror_nodes
    nodes, mirror = mirror_nodes(3, h=0.05, m=6, n=11)
    e_rho, e_u = _mirror_check(nodes, mirror, 27, 32, 0.05, 10)
    print(f"mirror drho={e_rho:.2e} u={e_u:.2e}")
    assert e_rho < 1e-5 and e_u < 1e-5


def test_mirror_symmetry_1m_nodes():
    """Full-size check (about one million D3Q27 nodes, the benchmark's
    scale): a pulse on the mirror plane of an exactly x-symmetric node set
    stays symmetric to FP32 rounding after 20 steps, so every pass indexes
    its stencils and rows consistently at this size."""
    from tests.scenes import mirror_nodes
    nodes, mirror = mirror_nodes(3, h=0.01, m=51, n=101, layer=0.08)
    assert nodes["points"].shape[0] > 1_000_000
    e_rho, e_u = _mirror_check(nodes, mirror, 27, 32, 0.01, 20)
    print(f"mirror 1M drho={e_rho:.2e} u={e_u:.2e}")
    assert e_rho < 1e-4 and e_u < 1e-4


@pytest.mark.parametrize("Q", [9, 27])
def test_density_probes(oracle64, Q):
    """Probes interpolate rho - rho0 with RPIM weights over the 32 nearest
    nodes (density_observer.cuh:62-113); checked against the oracle's kNN
    and RPIM weights applied to the engine's own density."""
    import naga_amd
    D = 2 if Q == 9 else 3
    h = 0.02 if D == 2 else 0.1
    k = 16 if D == 2 else 32
    nodes = grid_nodes(oracle64, D, h, layer=3 * h)
    gpu = gpu_engine(nodes, Q, k, h)
    gpu.set_initial_pulse([0.45] + [0.5] * (D - 1), 1e-3, 0.1)
    gpu.step_forward(8)
    rng = np.random.default_rng(4)
    probes = rng.uniform(0.1, 0.9, (7, D)).astype(np.float32)
    sid = gpu.add_probes(probes, naga_amd.RBF_MULTIQUADRIC)
    pts = nodes["points"]
    _, idx = oracle64.knn(pts, probes, 32)
    w = oracle64.rpim_weights(pts, idx, probes, 0, h)
    fresh = gpu.macroscopic(fresh=True)["drho"]
    want = 1.0 + (w * fresh[idx]).sum(1)
    got = gpu.sample_probes(sid, fresh=True)
    scale = np.abs(fresh).max()
    assert np.abs(got - want).max() < 1e-5 * scale
    held = gpu.macroscopic(fresh=False)["rho"].astype(np.float64)
    want_h = 1.0 + (w * (held[idx] - 1.0)).sum(1)
    assert np.abs(gpu.sample_probes(sid) - want_h).max() < 1e-5 * scale + 1e-7
    obs = naga_amd.DensityObserver(gpu, probes[0])
    obs.advance(3)
    assert len(obs.density_values) == 3 and obs.time_values[1] > obs.time_values[0]


@pytest.mark.parametrize("Q,kmax,kmin", [(27, 48, 24), (9, 32, 16)])
def test_variable_stencils_parity(oracle64, Q, kmax, kmin):
    """Box minus a sphere/circle with a refined shell and per-node stencil
    widths by local density (k in [kmin, kmax]): the step matches the oracle
    given the same widths."""
    import naga_amd
    D = 2 if Q == 9 else 3
    h = 0.04 if D == 2 else 0.1
    centre = [0.5] * D
    nodes = naga_amd.ObstacleGridProvider([0] * D, [1] * D, h, [(centre, 0.2)], 3 * h, 0.5,
                                          jitter=0.25, refine_width=0.1).get()
    eng = naga_amd.SimulationEngine(lattice=Q, k=kmax, k_boundary=32 if D == 3 else 16,
                                    rbf=naga_amd.RBF_POLYHARMONIC3,
                                    rbf_boundary=naga_amd.RBF_POLYHARMONIC3, k_min=kmin)
    p = naga_amd.problem_parameters(Q, 0.6, 0.4 * h)
    eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
    eng.init_domain(nodes)
    kv = eng.stencil_sizes()
    assert kv.min() >= kmin and kv.max() <= kmax and len(np.unique(kv)) > 3
    nd = nodes.as_dict()
    ora = oracle64.engine(nd, Q=Q, k=kmax, k_boundary=32 if D == 3 else 16, rbf=1,
                          rbf_boundary=1, h=h, kvar=kv)
    pc = [0.2] + [0.5] * (D - 1)
    eng.set_initial_pulse(pc, 1e-3, 0.1)
    ora.set_pulse(pc, 1e-3, 0.1)
    eng.step_forward(12)
    ora.step(12)
    g, o = eng.macroscopic(fresh=True), ora.get(fresh=True)
    err_rho, err_u = rel_l2(g["drho"], o["drho"]), rel_l2(g["u"], o["u"])
    print(f"variable k Q={Q}: rel_l2 drho={err_rho:.3e} u={err_u:.3e}")
    assert err_rho < TOL and err_u < TOL


def test_ragged_records_bit_identical(monkeypatch):
    """Per-node stencil widths with tile unions: each tile keeps only the
    groups of 8 records its widest node needs (the records past a node's
    width have zero weights), so the state matches full-width records bit
    for bit; NAGA_RAGGED=0 keeps full-width rows."""
    import naga_amd
    h = 0.05
    nodes = naga_amd.ObstacleGridProvider([0] * 3, [1] * 3, h, [([0.5] * 3, 0.2)], 3 * h, 0.5,
                                          jitter=0.25, refine_width=0.1).get()
    runs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("NAGA_RAGGED", flag)
        eng = naga_amd.SimulationEngine(lattice=27, k=48, k_boundary=32,
                                        rbf=naga_amd.RBF_POLYHARMONIC3,
                                        rbf_boundary=naga_amd.RBF_POLYHARMONIC3, k_min=24)
        p = naga_amd.problem_parameters(27, 0.6, 0.2 * h)
        eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
        eng.init_domain(nodes)
        eng.set_initial_pulse([0.3, 0.5, 0.5], 1e-3, 0.1)
        eng.step_forward(15)
        runs.append(eng.distributions()["f"])
        eng.close()
    assert np.array_equal(runs[0], runs[1])


def test_ragged_scenes_batched_like_alone():
    """Two scenes with per-node stencil widths (ragged record rows) and
    many boundary tiles past their last fluid tile, stepped in one batch,
    advance exactly as each does alone: the stage-1 and filter passes run
    over those boundary-only tiles without touching the per-tile record
    tables (one entry per fluid tile)."""
    import naga_amd

    def scene(h):
        nodes = naga_amd.ObstacleGridProvider([0] * 3, [1] * 3, h, [([0.5] * 3, 0.25)], 3 * h,
                                              0.5, jitter=0.25, refine_width=0.1).get()
        eng = naga_amd.SimulationEngine(lattice=27, k=48, k_boundary=32,
                                        rbf=naga_amd.RBF_POLYHARMONIC3,
                                        rbf_boundary=naga_amd.RBF_POLYHARMONIC3, k_min=24)
        p = naga_amd.problem_parameters(27, 0.6, 0.2 * h)
        eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
        eng.init_domain(nodes)
        assert nodes.num_boundary_points > 48 * 32  # boundary-only tiles exist
        eng.set_initial_pulse([0.3, 0.5, 0.5], 1e-3, 0.1)
        return eng

    hs = (0.04, 0.05)
    alone = [scene(h) for h in hs]
    for e in alone:
        e.step_forward(12)
    together = [scene(h) for h in hs]
    batch = naga_amd.SceneBatch(together)
    batch.step_forward(12)
    for a, b in zip(alone, together):
        assert np.array_equal(a.macroscopic(fresh=True)["drho"], b.macroscopic(fresh=True)["drho"])
    batch.close()
    for e in alone + together:
        e.close()


def test_tile_unions_bit_identical(oracle64, monkeypatch):
    """The gather kernels read neighbour rows from the tile unions staged in
    LDS by default and straight from HBM with NAGA_UNIONS=0: the same values
    in the same order, so the states agree bit for bit."""
    h = 0.04
    nodes = grid_nodes(oracle64, 3, h, layer=3 * h)
    runs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("NAGA_UNIONS", flag)
        eng = gpu_engine(nodes, 27, 32, h)
        eng.set_initial_pulse([0.4, 0.5, 0.5], 1e-3, 0.15)
        eng.step_forward(20)
        runs.append(eng.distributions()["f"])
    assert np.array_equal(runs[0], runs[1])


def test_3d_k16_unions(oracle64, monkeypatch):
    """D3Q27 with k = 16 (64-node union tiles, packed rows of 15 units):
    10 steps match the FP64 oracle, and the tile-union and direct-gather
    paths agree bit for bit."""
    h = 0.05
    nodes = grid_nodes(oracle64, 3, h, layer=3 * h)
    runs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("NAGA_UNIONS", flag)
        eng = gpu_engine(nodes, 27, 16, h)
        eng.set_initial_pulse([0.4, 0.5, 0.5], 1e-3, 0.15)
        eng.step_forward(10)
        runs.append((eng.distributions()["f"], eng.macroscopic(fresh=True)))
    assert np.array_equal(runs[0][0], runs[1][0])
    ora = oracle_engine(oracle64, nodes, 27, 16, h)
    ora.set_pulse([0.4, 0.5, 0.5], 1e-3, 0.15)
    ora.step(10)
    g, o = runs[0][1], ora.get(fresh=True)
    assert rel_l2(g["drho"], o["drho"]) < TOL and rel_l2(g["u"], o["u"]) < TOL


def test_split_tiles_bit_identical(oracle64, monkeypatch):
    """With a union cap below most 64-node tiles' unions, tiles whose two
    half-tile unions fit are split (a second block per tile serves the
    second half); the others read their rows through the list.  Split,
    unsplit over-cap and direct-gather runs agree bit for bit."""
    h = 0.04
    nodes = grid_nodes(oracle64, 3, h, layer=3 * h)
    nf = nodes["n_bulk"] + nodes["n_layer"]
    # the setup's split decision, from the oracle's kNN (bit-exact with the GPU's)
    pts = np.ascontiguousarray(nodes["points"], dtype=np.float32)
    _, idx = oracle64.knn(pts, pts[:nf], 32)
    cap, n_split = 300, 0
    for t0 in range(0, nf - 32, 64):
        tile = idx[t0:t0 + 64]
        if len(np.unique(tile)) > cap and max(len(np.unique(tile[:32])),
                                              len(np.unique(tile[32:]))) <= cap:
            n_split += 1
    assert n_split > 10
    runs = []
    for env in ({"NAGA_UCAP": str(cap)}, {"NAGA_UCAP": str(cap), "NAGA_SPLIT": "0"},
                {"NAGA_UNIONS": "0"}):
        for key in ("NAGA_UCAP", "NAGA_SPLIT", "NAGA_UNIONS"):
            monkeypatch.delenv(key, raising=False)
        for key, val in env.items():
            monkeypatch.setenv(key, val)
        eng = gpu_engine(nodes, 27, 32, h)
        eng.set_initial_pulse([0.4, 0.5, 0.5], 1e-3, 0.15)
        eng.step_forward(20)
        runs.append(eng.distributions()["f"])
    assert np.array_equal(runs[0], runs[2])
    assert np.array_equal(runs[1], runs[2])


def test_irregular_domain_2m_nodes(monkeypatch):
    """The irregular-domain scene at full size: the unit cube minus a
    central sphere of radius 0.15 at h = 0.008 with 2x refinement within
    0.05 of it (about 2M nodes) and per-node stencil widths 24-48, time
    step 0.4 of the refined spacing.  300 steps stay finite and bounded;
    the tile-union and direct-gather paths agree bit for bit (k = 48 unions,
    some over the LDS cap)."""
    import naga_amd
    h = 0.008
    nodes = naga_amd.ObstacleGridProvider([0, 0, 0], [1, 1, 1], h, [([0.5, 0.5, 0.5], 0.15)],
                                          8 * h, 0.5, jitter=0.25, refine_width=0.05).get()
    assert 1.9e6 < nodes.size < 2.3e6

    def run(steps):
        eng = naga_amd.SimulationEngine(lattice=27, k=48, k_boundary=32,
                                        rbf=naga_amd.RBF_POLYHARMONIC3,
                                        rbf_boundary=naga_amd.RBF_POLYHARMONIC3, k_min=24)
        p = naga_amd.problem_parameters(27, 0.6, 0.4 * (h / 2))  # 0.4 of the refined spacing
        eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
        eng.init_domain(nodes)
        eng.set_initial_pulse([0.3, 0.5, 0.5], 1e-3, 0.05)
        ms = eng.step_timed(steps)
        print(f"irregular domain: {nodes.size} nodes, {ms / steps:.3f} ms per step")
        out = eng.distributions()["f"], eng.macroscopic(fresh=True)
        eng.close()
        return out

    monkeypatch.setenv("NAGA_UNIONS", "1")
    f1, m = run(20)
    monkeypatch.setenv("NAGA_UNIONS", "0")
    f0, _ = run(20)
    assert np.array_equal(f1, f0)
    monkeypatch.setenv("NAGA_UNIONS", "1")
    _, m = run(300)
    assert np.all(np.isfinite(m["drho"])) and np.all(np.isfinite(m["u"]))
    assert np.abs(m["drho"]).max() < 1.5e-3


@pytest.mark.parametrize("Q", [9, 27])
def test_step_graph_replays_match_direct_launches(oracle64, Q):
    """step_forward(n >= 2) replays one captured step; step_forward(1)
    launches directly.  Replays interleaved with direct steps and host
    reads give the direct-launch state bit for bit after every call."""
    D = 2 if Q == 9 else 3
    h = 0.02 if D == 2 else 0.1
    nodes = grid_nodes(oracle64, D, h, layer=3 * h)
    a = gpu_engine(nodes, Q, 16 if D == 2 else 32, h)
    b = gpu_engine(nodes, Q, 16 if D == 2 else 32, h)
    for e in (a, b):
        e.set_initial_pulse([0.4] + [0.5] * (D - 1), 1e-3, 0.1)
    for n in (3, 1, 5, 1, 2, 20):
        b.step_forward(n)
        for _ in range(n):
            a.step_forward(1)
        assert np.array_equal(a.distributions()["f"], b.distributions()["f"]), n
        assert abs(a.time() - b.time()) < 1e-12
