"""Parity of the benchmark configurations against the FP64 oracle:

* the 1,030,301-node D3Q27 cube at full size: kNN lists bit-exact,
  divergence weights to 1e-6, 20 steps at rel-L2 <= 1e-5;
* the batched form: 32 scenes drawn by the batched workload's generator
  (sizes capped at 8k nodes), stepped as one autotuned batch, every scene
  against the oracle after 60 steps;
* the served form: a Poisson trace drawn by the served workload's
  generator (12 requests, 4k-15k nodes, 30-90 steps) served with continuous batching,
  every result against the oracle.
The parity bar is rel-L2 <= 1e-5 of rho - rho0 and u (FP32 state against
the FP64 oracle of the same method)."""
import time

import numpy as np
import pytest

import naga_amd
from naga_amd import workloads
from tests.scenes import kvar_like_engine, rel_l2

pytestmark = pytest.mark.gpu

BAR = 1e-5


def _say(msg):
    print(msg, flush=True)


def test_1m_node_cube_full_size_parity(oracle64):
    """The D3Q27 benchmark cube at its size: h = 0.01 (101^3 nodes),
    jitter 0.25h, k = 32, r^3 RBF, PML 8h sigma_max 0.5, tau 0.6, dt 0.4h,
    pulse 1e-3 / 0.05 at (0.3, 0.5, 0.5)."""
    h = 0.01
    nodes = oracle64.uniform_grid(3, [0] * 3, [1] * 3, h, jitter=0.25, layer=8 * h,
                                  sigma_max=0.5)
    t0 = time.perf_counter()
    ref = oracle64.engine(nodes, Q=27, k=32, h=h, tau=0.6, threads=16)
    _say(f"1M-node cube: oracle setup {time.perf_counter() - t0:.1f} s, {ref.N} nodes")
    sim = naga_amd.SimulationNodes(nodes["points"], nodes["normals"], nodes["sigma"],
                                   nodes["n_bulk"], nodes["n_layer"], nodes["n_boundary"], h)
    eng = naga_amd.SimulationEngine(lattice=27, k=32, rbf=naga_amd.RBF_POLYHARMONIC3)
    p = naga_amd.problem_parameters(27, 0.6, 0.4 * h)
    eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
    eng.init_domain(sim)
    g, o = eng.setup_products(), ref.setup()
    assert np.array_equal(g["idx"], o["idx"])
    assert np.array_equal(g["bidx"], o["bidx"])
    assert np.abs(g["W"] - o["W"]).max() <= 1e-6 * np.abs(o["W"]).max()
    assert np.abs(g["bw"] - o["bw"]).max() <= 1e-6
    _say("1M-node cube: kNN bit-exact, weights within 1e-6")
    eng.set_initial_pulse([0.3, 0.5, 0.5], 1e-3, 0.05)
    ref.set_pulse([0.3, 0.5, 0.5], 1e-3, 0.05)
    eng.step_forward(20)
    t0 = time.perf_counter()
    ref.step(20)
    _say(f"1M-node cube: oracle 20 steps {time.perf_counter() - t0:.1f} s")
    a, b = eng.macroscopic(fresh=True), ref.get()
    e_rho, e_u = rel_l2(a["drho"], b["drho"]), rel_l2(a["u"], b["u"])
    _say(f"1M-node cube: 20 steps rel-L2 drho {e_rho:.2e} u {e_u:.2e}")
    assert e_rho <= BAR and e_u <= BAR
    eng.close()


def test_batched_scenes_vs_oracle(oracle64):
    scenes = workloads.batched_scenes(32, n_lo=2e3, n_hi=8e3)
    engines, nodes = [], []
    for s in scenes:
        e, nd = workloads.batched_engine(s)
        engines.append(e)
        nodes.append(nd)
    batch = naga_amd.SceneBatch(engines)
    gb = batch.autotune(3)
    steps = 60
    batch.step_forward(steps)
    _say(f"batched scenes: {len(scenes)} scenes, "
         f"{sum(e.num_nodes for e in engines)} nodes, gather batch {gb}")
    worst = 0.0
    for s, e, nd in zip(scenes, engines, nodes):
        ref = oracle64.engine(nd.as_dict(), Q=9, k=s.k, h=s.h, tau=0.6, threads=16)
        ref.set_pulse(s.pulse, 1e-3, 0.05)
        ref.step(steps)
        a, b = e.macroscopic(fresh=True), ref.get()
        err = max(rel_l2(a["drho"], b["drho"]), rel_l2(a["u"], b["u"]))
        worst = max(worst, err)
        assert err <= BAR, (s.h, s.k, len(s.obstacles), err)
    _say(f"batched scenes: worst rel-L2 {worst:.2e}")
    batch.close()


def _oracle_solve(oracle, req):
    D = 2 if req.lattice == 9 else 3
    k = req.k or (16 if D == 2 else 32)
    nd = naga_amd.request_nodes(req).as_dict()
    kv = kvar_like_engine(oracle, nd["points"], k, req.k_min) if req.k_min > 0 else None
    hmin = 0.5 * req.h if req.refine_width > 0 else req.h
    ref = oracle.engine(nd, Q=req.lattice, k=k, h=req.h, tau=req.tau or 0.6,
                        dt=0.4 * hmin, threads=16, kvar=kv)
    ref.set_pulse([req.pulse_center[d] for d in range(D)], req.pulse_amplitude,
                  req.pulse_width)
    ref.step(req.steps)
    return ref.get()


def test_served_trace_vs_oracle(oracle64):
    trace = workloads.served_requests(12, n_lo=4e3, n_hi=1.5e4, s_lo=30, s_hi=90)
    srv = naga_amd.SolveServer(hbm_budget_bytes=16 << 30, max_active=64, chunk_steps=25)
    ids = []
    for r, gap in zip(trace.requests, trace.gaps):
        time.sleep(float(gap))
        ids.append(srv.submit(r))
    srv.drain()
    st = srv.stats()
    assert st["completed"] == len(ids) and st["failed"] == 0
    _say(f"served trace: {len(ids)} requests served, p50 {st['p50_ms']:.0f} ms "
         f"p99 {st['p99_ms']:.0f} ms")
    worst = 0.0
    for r, i in zip(trace.requests, ids):
        got = srv.result(i)
        ref = _oracle_solve(oracle64, r)
        err = max(rel_l2(got["drho"], ref["drho"]), rel_l2(got["u"], ref["u"]))
        worst = max(worst, err)
        assert err <= BAR, (r.lattice, r.h, r.steps, err)
    _say(f"served trace: worst rel-L2 {worst:.2e}")
    srv.close()
