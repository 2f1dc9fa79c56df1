"""Synthetic workloads of the batched and served forms, shared by bench.py
and the parity tests: scene and request descriptions drawn from one seeded
stream, so a test can draw the same kind of workload at capped sizes.

* batched_scenes: unit squares with 0-5 circular obstacles of radius
  0.03-0.1, node counts log-uniform in [n_lo, n_hi] (2k-50k in the
  256-scene workload), k in [12, 24], a random pulse position; D2Q9, PML 8h
  at the faces and 3h around the obstacles, tau 0.6, dt 0.4h.
* served_requests: Poisson arrivals, D2Q9/D3Q27 50/50, node counts
  log-uniform in [n_lo, n_hi] (5k-2M), steps uniform in [s_lo, s_hi]
  (100-2000); D3Q27 requests are the unit cube minus a central sphere of
  radius 0.15 refined 2x within 0.05 of it (for h <= 0.05, jitter 0.1h;
  coarser requests unrefined and unjittered), k 24-48 by local density;
  D2Q9 requests the unit square with 0-5
  circles, k 16, jitter 0.25h, an absorbing layer of 3h around the circles.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class BatchedScene:
    h: float
    k: int
    obstacles: list
    seed: int
    pulse: list
    layer_cells: int = 8
    immersed_cells: int = 3

    def provider(self):
        import naga_amd
        return naga_amd.ObstacleGridProvider([0, 0], [1, 1], self.h, self.obstacles,
                                             self.layer_cells * self.h, 0.5, jitter=0.25,
                                             seed=self.seed,
                                             immersed_layer_thickness=self.immersed_cells * self.h,
                                             immersed_absorption_rate=0.5)


def batched_scenes(n_scenes: int = 256, n_lo: float = 2e3, n_hi: float = 5e4,
                   seed: int = 0x4E414741) -> list:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_scenes):
        n = float(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        h = 1.0 / (np.sqrt(n) - 1.0)
        obst = [([float(v) for v in rng.uniform(0.15, 0.85, 2)], float(rng.uniform(0.03, 0.1)))
                for _ in range(int(rng.integers(0, 6)))]
        k = int(rng.integers(12, 25))
        s = int(rng.integers(1, 1 << 62))
        pulse = [float(v) for v in rng.uniform(0.2, 0.8, 2)]
        out.append(BatchedScene(h=h, k=k, obstacles=obst, seed=s, pulse=pulse))
    return out


def batched_engine(scene: BatchedScene):
    """The GPU engine of one batched scene, pulse set (amplitude 1e-3, width
    0.05)."""
    import naga_amd
    nodes = scene.provider().get()
    eng = naga_amd.SimulationEngine(lattice=9, k=scene.k, k_boundary=scene.k,
                                    rbf=naga_amd.RBF_POLYHARMONIC3,
                                    rbf_boundary=naga_amd.RBF_POLYHARMONIC3)
    p = naga_amd.problem_parameters(9, 0.6, 0.4 * scene.h)
    eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
    eng.init_domain(nodes)
    eng.set_initial_pulse(scene.pulse, 1e-3, 0.05)
    return eng, nodes


@dataclass
class ServedTrace:
    requests: list
    gaps: np.ndarray = field(default_factory=lambda: np.zeros(0))


def served_requests(n_requests: int = 32, rate: float = 50.0, n_lo: float = 5e3,
                    n_hi: float = 2e6, s_lo: int = 100, s_hi: int = 2000,
                    seed: int = 0x4E414742) -> ServedTrace:
    import naga_amd
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n_requests):
        lat = 9 if rng.random() < 0.5 else 27
        D = 2 if lat == 9 else 3
        n = float(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        h = 1.0 / (n ** (1.0 / D) - 1.0)
        pc = [float(v) for v in rng.uniform(0.2, 0.3, 1)] + \
             [float(v) for v in rng.uniform(0.35, 0.65, D - 1)]
        if D == 2:
            obst = [([float(v) for v in rng.uniform(0.15, 0.85, 2)], float(rng.uniform(0.03, 0.1)))
                    for _ in range(int(rng.integers(0, 6)))]
            kw = dict(k=16)
        else:
            obst = [([0.5, 0.5, 0.5], 0.15)]
            # The refinement shell only where the sphere spans at least 3
            # spacings (h <= 0.05), with jitter 0.1h; coarser requests on the
            # unjittered grid.  Over the 200-request trace (up to 2000 steps)
            # these choices keep every request bounded, while a mode grows
            # next to the sphere with jitter 0.25h (some refined requests),
            # 0.15h (one refined request of 1223 steps, and two coarse ones
            # of about 1500 steps whose sphere spans 2.9 spacings) and on a
            # refined sphere of 2.4 spacings at any jitter.
            fine = h <= 0.05
            kw = dict(k=48, k_min=24, refine_width=0.05 if fine else 0.0,
                      jitter=0.1 if fine else 0.0)
        reqs.append(naga_amd.make_request(
            lat, h, int(rng.integers(s_lo, s_hi + 1)), obst, seed=int(rng.integers(1, 1 << 62)),
            pulse_center=pc, pulse_width=0.05, **kw))
    gaps = rng.exponential(1.0 / rate, n_requests)
    return ServedTrace(requests=reqs, gaps=gaps)
