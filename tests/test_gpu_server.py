"""Serving loop: requests admitted by HBM budget and stepped together with
continuous batching return the same result as the same solve alone."""
import numpy as np
import pytest

import naga_amd
from tests.scenes import rel_l2

pytestmark = pytest.mark.gpu


def _trace(seed=3, n=12):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        lat = 9 if i % 3 else 27
        D = 2 if lat == 9 else 3
        h = float(rng.choice([0.02, 0.025, 0.03125])) if D == 2 else float(rng.choice([0.1, 0.125]))
        nob = int(rng.integers(0, 3))
        obst = [([float(v) for v in rng.uniform(0.3, 0.7, D)], float(rng.uniform(0.05, 0.1)))
                for _ in range(nob)]
        steps = int(rng.integers(5, 60))
        centre = [float(v) for v in rng.uniform(0.2, 0.8, D)]
        reqs.append(naga_amd.make_request(lat, h, steps, obst, seed=int(rng.integers(1, 1 << 30)),
                                          pulse_center=centre, pulse_width=0.1))
    return reqs


def test_served_results_equal_alone():
    reqs = _trace()
    srv = naga_amd.SolveServer(hbm_budget_bytes=8 << 30, max_active=64, chunk_steps=7)
    ids = [srv.submit(r) for r in reqs]
    srv.drain()
    st = srv.stats()
    got_all = [srv.result(i, release=False) for i in ids]  # raises with a failed request's error
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["p99_ms"] >= st["p50_ms"] > 0 and st["node_updates_per_s"] > 0
    for r, i, got in zip(reqs, ids, got_all):
        ref = naga_amd.solve_alone(r)
        assert got["nodes"] == ref["nodes"]
        assert rel_l2(got["drho"], ref["drho"]) < 1e-6, (r.lattice, r.h, r.steps)
        assert rel_l2(got["u"], ref["u"]) < 1e-6
    srv.close()


def test_small_budget_admits_one_at_a_time():
    """A budget below two scenes' footprint serialises admission; results
    are unchanged."""
    reqs = _trace(seed=5, n=4)
    srv = naga_amd.SolveServer(hbm_budget_bytes=1, max_active=64, chunk_steps=10)
    ids = [srv.submit(r) for r in reqs]
    srv.drain()
    assert srv.stats()["completed"] == 4
    for r, i in zip(reqs, ids):
        assert rel_l2(srv.result(i)["drho"], naga_amd.solve_alone(r)["drho"]) < 1e-6
    srv.close()


def test_served_refined_variable_stencils_equal_alone():
    """Requests laid out as the irregular-domain workload (box minus a
    central sphere or circle, 2x refinement near it, per-node stencil widths
    by density) served together equal the same solves alone."""
    reqs = []
    for i, (lat, h, steps) in enumerate([(27, 0.08, 20), (9, 0.02, 35), (27, 0.1, 12),
                                         (9, 0.025, 25)]):
        D = 2 if lat == 9 else 3
        k, kmin = (48, 24) if D == 3 else (24, 12)
        reqs.append(naga_amd.make_request(
            lat, h, steps, [([0.5] * D, 0.15)], k=k, k_min=kmin, refine_width=1.5 * h,
            seed=100 + i, pulse_center=[0.3] + [0.5] * (D - 1), pulse_width=0.1))
    srv = naga_amd.SolveServer(hbm_budget_bytes=8 << 30, max_active=8, chunk_steps=9)
    ids = [srv.submit(r) for r in reqs]
    srv.drain()
    assert srv.stats()["completed"] == len(reqs)
    for r, i in zip(reqs, ids):
        got, ref = srv.result(i), naga_amd.solve_alone(r)
        assert got["nodes"] == ref["nodes"]
        assert np.all(np.isfinite(got["drho"])) and np.all(np.isfinite(got["u"]))
        assert rel_l2(got["drho"], ref["drho"]) < 1e-6, (r.lattice, r.h)
        assert rel_l2(got["u"], ref["u"]) < 1e-6
    srv.close()
