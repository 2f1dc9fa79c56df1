"""Benchmark: D3Q27 nonlocal-LBM node-updates/s on one MI355X.

Workload (the headline configuration): unit cube, h = 0.01 (101^3 =
1,030,301 nodes), interior nodes jittered by +-0.25h (SplitMix64 seed
0x4E414741), k = 32 stencils with polyharmonic r^3 RBF + linear terms,
FP64 (36 x 36) solves, central-moment collision tau = 0.6, PML 8h with
sigma_max 0.5, Gaussian pulse 1e-3 / 0.05 at (0.3, 0.5, 0.5), RK4 with
dt = 0.4h.  Synthetic nodes, state resident in HBM before timing.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  value = nodes * K / wall time of the K timed steps.
roofline: the dominant kernel (an RK4 stage) timed with HIP events on the
engine's stream, against its algorithmic bytes per launch.
cpu_baseline: the oracle's FP32 OpenMP port of the same method on a bounded
sample (same workload at h = 0.02), steps only.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E
# The RK4 stage kernels of the 1M-node workload (tile unions, 64-node tiles);
# the dominant one, stage 4, is the roofline line.
STAGE_KERNELS = {s: f"k_stage_u<{s}, 32, 8, true, 64, false>" for s in (1, 2, 3, 4)}
ROOFLINE_STAGE = 4


def traffic_profile(stage: int) -> str:
    """HBM bytes per launch of a stage kernel on this workload, from two
    rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) summarised by
    naga_amd.profsum; the file records the commands."""
    return os.path.join(ROOT, "profiles", f"k_stage{stage}_d3q27_h0.01_traffic.md")


def measured_traffic(h: float, stage: int):
    path = traffic_profile(stage)
    if h != 0.01 or not os.path.exists(path):
        return None
    text = open(path).read()
    if f"`{STAGE_KERNELS[stage]}`" not in text.splitlines()[0]:
        return None  # the profile is of another kernel build
    for line in text.splitlines():
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1])
    return None


def copy_bandwidth(nbytes: int = 2 << 30, reps: int = 10) -> float:
    """Measured device-to-device copy rate (bytes read + written per second)
    of the runtime's copy over nbytes, the practical HBM ceiling beside the
    8 TB/s nominal peak."""
    from naga_amd import devmem
    return devmem.copy_rate(nbytes, reps)


def copy_calibration(nbytes: int) -> int:
    """One device-to-device copy of nbytes (the runtime's copy kernel): the
    known-size reference for scaling HBM byte counters."""
    from naga_amd import devmem
    devmem.copy_once(nbytes)
    print(json.dumps({"copied_bytes": nbytes}))
    return 0


def stage_bytes_per_node(Q: int, D: int, k: int, stage: int) -> int:
    """Algorithmic bytes of one RK4 stage per fluid node: the stencil slab
    (k records of 4 + 4D bytes) plus FP32 rows of Q values: the gathered
    input read once, the state row (the same array as the input in stage
    1), the stage output, and in stage 4 the two earlier stage outputs."""
    rows = {1: 2, 2: 3, 3: 3, 4: 5}[stage]
    return k * (4 + 4 * D) + rows * 4 * Q


def step_model_bytes(nodes, Q, D, k):
    """Whole-step algorithmic bytes per node-update (bulk/layer/boundary
    model): bulk 76Q + 8k + 16k(1+D), layer +24Q + 4k(1+D), boundary
    2(8k + 8Q) + 8Q."""
    nb, nl, nbd = nodes.num_bulk_points, nodes.num_layer_points, nodes.num_boundary_points
    bulk = 76 * Q + 8 * k + 16 * k * (1 + D)
    layer = bulk + 24 * Q + 4 * k * (1 + D)
    bnd = 2 * (8 * k + 8 * Q) + 8 * Q
    return (nb * bulk + nl * layer + nbd * bnd) / (nb + nl + nbd)


def cpu_baseline(threads: int) -> dict:
    import naga_amd
    from oracle import Oracle

    h = 0.02
    o = Oracle(32)
    nodes = o.uniform_grid(3, [0, 0, 0], [1, 1, 1], h, jitter=0.25, layer=8 * h,
                           sigma_max=0.5)
    t0 = time.perf_counter()
    eng = o.engine(nodes, Q=27, k=32, h=h, tau=0.6, threads=threads)
    cpu_setup = time.perf_counter() - t0
    # the GPU setup of the same sample, for the setup comparison
    g = naga_amd.SimulationEngine(lattice=27, k=32, rbf=naga_amd.RBF_POLYHARMONIC3)
    p = naga_amd.problem_parameters(27, 0.6, 0.4 * h)
    g.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
    sim = naga_amd.SimulationNodes(nodes["points"], nodes["normals"], nodes["sigma"],
                                   nodes["n_bulk"], nodes["n_layer"], nodes["n_boundary"], h)
    t0 = time.perf_counter()
    g.init_domain(sim)
    gpu_setup = time.perf_counter() - t0
    g.close()
    eng.set_pulse([0.3, 0.5, 0.5], 1e-3, 0.05)
    eng.step(1)  # warm-up
    n = 80  # about 10 s of CPU work on 16 threads
    t0 = time.perf_counter()
    eng.step(n)
    dt = time.perf_counter() - t0
    N = eng.N
    return {"value": N * n / dt, "unit": "node-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"D3Q27 unit cube h=0.02 ({N} nodes, k=32, PML 8h), {n} steps "
                      f"after 1 warm-up step, FP32 oracle port, OpenMP {threads} threads",
            "setup_s": cpu_setup, "gpu_setup_s_same_sample": gpu_setup}


def batched_workload(steps: int, n_scenes: int = 256) -> dict:
    """Many small D2Q9 scenes (naga_amd.workloads.batched_scenes: node counts
    log-uniform in [2k, 50k], 0-5 circular obstacles of radius 0.03-0.1, k in
    [12, 24], random pulse).  The same scenes are stepped one scene at a time
    (each engine's own launches, scene after scene) and together in one
    batch (one launch per pass over all scenes, graph replay, autotuned)."""
    import naga_amd
    from naga_amd import workloads
    engines, total = [], 0
    t0 = time.perf_counter()
    for scene in workloads.batched_scenes(n_scenes):
        eng, _ = workloads.batched_engine(scene)
        engines.append(eng)
        total += eng.num_nodes
    setup_s = time.perf_counter() - t0
    print(f"batched: {n_scenes} scenes, {total} nodes, setup {setup_s:.1f} s", file=sys.stderr,
          flush=True)
    for e in engines:
        e.step_forward(2)
    t0 = time.perf_counter()
    for e in engines:
        e.step_forward(steps)
    plain_s = time.perf_counter() - t0
    batch = naga_amd.SceneBatch(engines)
    gb = batch.autotune(3)
    tune = batch.tuning()
    batch.step_forward(2)
    t0 = time.perf_counter()
    batch.step_forward(steps)
    batched_s = time.perf_counter() - t0
    plain, batched = total * steps / plain_s, total * steps / batched_s
    # every scene's state after its 2 + steps + 2 + steps steps: finite and
    # bounded by the initial pulse amplitude (a diverging scene would show)
    peak = max(float(np.abs(e.macroscopic(fresh=True)["drho"]).max()) for e in engines)
    batch.close()
    return {"scenes": n_scenes, "nodes": total, "steps": steps, "setup_s": setup_s,
            "max_abs_drho_after": peak,
            "plain_node_updates_per_s": plain, "batched_node_updates_per_s": batched,
            "speedup": batched / plain, "gather_batch": gb, "q9_lanes": tune["q9_lanes"],
            "q9_unions": tune["q9_unions"], "graph": tune["graph"],
            "workload": "D2Q9, 256 scenes, N log-uniform 2k-50k, 0-5 obstacles r 0.03-0.1, "
                        "k 12-24, PML 8h (3h around obstacles), tau 0.6, dt 0.4h"}


SIZE_CLASSES = ((0, 5e4, "<50k"), (5e4, 5e5, "50k-500k"), (5e5, 1e12, ">=500k"))


def served_workload(n_requests: int = 32, rate: float = 50.0) -> dict:
    """Poisson arrivals at `rate` requests/s (naga_amd.workloads.served_requests:
    D2Q9/D3Q27 50/50, node counts log-uniform in [5k, 2M], steps uniform in
    [100, 2000], random pulse position; D3Q27 requests the unit cube minus a
    central sphere of radius 0.15, 2x node density within 0.05 of it, k from
    24 to 48 by local density; D2Q9 requests the unit square with 0-5 circles
    of radius 0.03-0.1, k = 16).  Served with continuous batching, then the
    same requests solved one at a time with their own launches (setup
    included in both).  Every served result is checked to stay bounded, and
    the plain loop's time is broken down by request size."""
    import naga_amd
    from naga_amd import workloads
    trace = workloads.served_requests(n_requests, rate)
    reqs = trace.requests
    srv = naga_amd.SolveServer(hbm_budget_bytes=96 << 30, max_active=64, chunk_steps=50)
    t0 = time.perf_counter()
    ids = []
    for r, g in zip(reqs, trace.gaps):
        time.sleep(float(g))
        ids.append(srv.submit(r))
    srv.drain()
    st = srv.stats()
    peaks = [float(np.abs(srv.result(i)["drho"]).max()) for i in ids]
    srv.close()
    print(f"served: {n_requests} requests in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    work, classes = 0.0, {name: {"requests": 0, "node_updates": 0.0, "plain_s": 0.0}
                          for _, _, name in SIZE_CLASSES}
    t1 = time.perf_counter()
    for r in reqs:
        ta = time.perf_counter()
        out = naga_amd.solve_alone(r)
        dt = time.perf_counter() - ta
        nu = float(out["nodes"]) * r.steps
        work += nu
        for lo, hi, name in SIZE_CLASSES:
            if lo <= out["nodes"] < hi:
                c = classes[name]
                c["requests"] += 1
                c["node_updates"] += nu
                c["plain_s"] += dt
    plain_s = time.perf_counter() - t1
    for c in classes.values():
        c["work_share"] = c["node_updates"] / work
        c["plain_node_updates_per_s"] = c["node_updates"] / c["plain_s"] if c["plain_s"] else None
    print(f"served: plain loop {plain_s:.1f} s", file=sys.stderr, flush=True)
    return {"requests": n_requests, "rate_per_s": rate, "p50_ms": st["p50_ms"],
            "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
            "served_node_updates_per_s": st["node_updates_per_s"],
            "plain_node_updates_per_s": work / plain_s,
            "speedup": st["node_updates_per_s"] / (work / plain_s),
            "trace_s": time.perf_counter() - t0 - plain_s,
            "results_bounded": int(sum(bool(np.isfinite(p) and p < 1e-3) for p in peaks)),
            "max_abs_drho_after": max((p for p in peaks if np.isfinite(p)), default=None),
            "unbounded": [{"lattice": r.lattice, "h": r.h, "steps": r.steps,
                           "max_abs_drho_after": p}
                          for r, p in zip(reqs, peaks) if not (np.isfinite(p) and p < 1e-3)],
            "plain_by_size": classes,
            "workload": "Poisson 50 req/s, D2Q9/D3Q27 50/50, N log-uniform 5k-2M, "
                        "steps 100-2000; 3D: cube minus sphere r 0.15, 2x refinement "
                        "within 0.05 and jitter 0.1h (h <= 0.05; else unjittered), "
                        "k 24-48; 2D: 0-5 circles with a 3h absorbing layer, k 16"}


def profile_workload(name: str, steps: int) -> dict:
    """One workload alone for a profiler run: the 256-scene batched D2Q9
    workload (autotuned, then `steps` batched steps), the irregular 3D
    scene (`steps` steps), or the set-up of the 1M-node D3Q27 scene (done
    twice, the second timed by phase); returns its time per step or its
    set-up phases."""
    import naga_amd
    from naga_amd import workloads
    if name == "setup":
        h = 0.01
        nodes = naga_amd.UniformGridProvider([0, 0, 0], [1, 1, 1], h, 8 * h, 0.5,
                                             jitter=0.25).get()
        p = naga_amd.problem_parameters(27, 0.6, 0.4 * h)
        out = {}
        for rep in range(2):
            eng = naga_amd.SimulationEngine(lattice=27, k=32, rbf=naga_amd.RBF_POLYHARMONIC3)
            eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
            t0 = time.perf_counter()
            eng.init_domain(nodes)
            out = {"workload": "setup", "nodes": eng.num_nodes,
                   "setup_s": time.perf_counter() - t0, "phases_ms": eng.setup_phases()}
            eng.close()
        return out
    if name == "batched":
        engines = [workloads.batched_engine(s)[0] for s in workloads.batched_scenes(256)]
        batch = naga_amd.SceneBatch(engines)
        batch.autotune(3)
        batch.step_forward(5)
        t0 = time.perf_counter()
        batch.step_forward(steps)
        wall = time.perf_counter() - t0
        nodes = sum(e.num_nodes for e in engines)
        batch.close()
    else:
        h = 0.008
        nd = naga_amd.ObstacleGridProvider([0, 0, 0], [1, 1, 1], h, [([0.5, 0.5, 0.5], 0.15)],
                                           8 * h, 0.5, jitter=0.25, refine_width=0.05).get()
        eng = naga_amd.SimulationEngine(lattice=27, k=48, k_boundary=32,
                                        rbf=naga_amd.RBF_POLYHARMONIC3,
                                        rbf_boundary=naga_amd.RBF_POLYHARMONIC3, k_min=24)
        p = naga_amd.problem_parameters(27, 0.6, 0.4 * (h / 2))
        eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
        eng.init_domain(nd)
        eng.set_initial_pulse([0.3, 0.5, 0.5], 1e-3, 0.05)
        eng.step_forward(5)
        t0 = time.perf_counter()
        eng.step_forward(steps)
        wall = time.perf_counter() - t0
        nodes = eng.num_nodes
        eng.close()
    return {"workload": name, "nodes": nodes, "steps": steps, "ms_per_step": wall * 1e3 / steps,
            "node_updates_per_s": nodes * steps / wall}


def other_configs(threads: int) -> dict:
    """Throughput of the 2D plain scene (D2Q9, 10,201 nodes, k = 16, PML
    10h, 1000 steps) and the irregular 3D scene (cube minus a refined
    sphere, 2.24M nodes, k 24-48, 300 steps) on the GPU, and of the 2D
    scene on the oracle's FP32 OpenMP port; setup times beside them."""
    import naga_amd
    from oracle import Oracle
    out = {}
    h = 0.01
    o = Oracle(32)
    nd = o.uniform_grid(2, [0, 0], [1, 1], h, jitter=0.25, layer=10 * h, sigma_max=0.5)
    nodes = naga_amd.SimulationNodes(nd["points"], nd["normals"], nd["sigma"], nd["n_bulk"],
                                     nd["n_layer"], nd["n_boundary"], h)
    eng = naga_amd.SimulationEngine(lattice=9, k=16, rbf=naga_amd.RBF_POLYHARMONIC3)
    p = naga_amd.problem_parameters(9, 0.6, 0.4 * h)
    eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
    t0 = time.perf_counter()
    eng.init_domain(nodes)
    setup = time.perf_counter() - t0
    eng.set_initial_pulse([0.5, 0.5], 1e-3, 0.05)
    eng.step_forward(10)
    t0 = time.perf_counter()
    eng.step_forward(1000)
    wall = time.perf_counter() - t0
    out["2d_plain"] = {"nodes": eng.num_nodes, "steps": 1000, "setup_s": setup,
                       "node_updates_per_s": eng.num_nodes * 1000 / wall}
    eng.close()
    t0 = time.perf_counter()
    ora = o.engine(nd, Q=9, k=16, h=h, tau=0.6, threads=threads)
    cpu_setup = time.perf_counter() - t0
    ora.set_pulse([0.5, 0.5], 1e-3, 0.05)
    ora.step(1)
    t0 = time.perf_counter()
    ora.step(200)
    out["2d_plain"]["cpu_node_updates_per_s"] = ora.N * 200 / (time.perf_counter() - t0)
    out["2d_plain"]["cpu_setup_s"] = cpu_setup
    out["2d_plain"]["cpu_cores"] = threads
    h = 0.008
    nodes = naga_amd.ObstacleGridProvider([0, 0, 0], [1, 1, 1], h, [([0.5, 0.5, 0.5], 0.15)],
                                          8 * h, 0.5, jitter=0.25, refine_width=0.05).get()
    eng = naga_amd.SimulationEngine(lattice=27, k=48, k_boundary=32,
                                    rbf=naga_amd.RBF_POLYHARMONIC3,
                                    rbf_boundary=naga_amd.RBF_POLYHARMONIC3, k_min=24)
    p = naga_amd.problem_parameters(27, 0.6, 0.4 * (h / 2))
    eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
    t0 = time.perf_counter()
    eng.init_domain(nodes)
    setup = time.perf_counter() - t0
    eng.set_initial_pulse([0.3, 0.5, 0.5], 1e-3, 0.05)
    eng.step_forward(5)
    t0 = time.perf_counter()
    eng.step_forward(300)
    wall = time.perf_counter() - t0
    out["3d_irregular"] = {"nodes": eng.num_nodes, "steps": 300, "setup_s": setup,
                           "node_updates_per_s": eng.num_nodes * 300 / wall,
                           "max_abs_drho_after": float(np.abs(
                               eng.macroscopic(fresh=True)["drho"]).max())}
    eng.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--h", type=float, default=0.01)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=20)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device copy of BYTES (counter calibration)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched and served workloads")
    ap.add_argument("--batched-steps", type=int, default=500)
    ap.add_argument("--served-requests", type=int, default=32)
    ap.add_argument("--no-other-configs", action="store_true",
                    help="skip the 2D plain and irregular 3D scenes")
    ap.add_argument("--profile-workload", choices=("batched", "irregular", "setup"), default=None,
                    help="only run that workload for --steps steps (for a profiler)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.profile_workload:
        print(json.dumps(profile_workload(args.profile_workload, args.steps)))
        return 0
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; multi-GPU sharding is "
              "designed (halo exchange per stencil pass) but not built", file=sys.stderr)
        return 2

    import naga_amd

    Q, D, k, h = 27, 3, 32, args.h
    t0 = time.perf_counter()
    nodes = naga_amd.UniformGridProvider([0, 0, 0], [1, 1, 1], h, 8 * h, 0.5,
                                         jitter=0.25).get()
    t_nodes = time.perf_counter() - t0
    p = naga_amd.problem_parameters(Q, 0.6, 0.4 * h)
    eng = naga_amd.SimulationEngine(lattice=Q, k=k, rbf=naga_amd.RBF_POLYHARMONIC3)
    eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
    t0 = time.perf_counter()
    eng.init_domain(nodes)
    t_setup = time.perf_counter() - t0
    eng.set_initial_pulse([0.3, 0.5, 0.5], 1e-3, 0.05)
    N = eng.num_nodes

    eng.step_forward(args.warmup)
    t0 = time.perf_counter()
    gpu_ms = eng.step_timed(args.steps)
    wall = time.perf_counter() - t0
    value = N * args.steps / wall

    # per-kernel times (HIP events on the engine stream), averaged
    prof = {}
    for _ in range(args.profile_reps):
        for kname, ms in eng.profile_step().items():
            prof[kname] = prof.get(kname, 0.0) + ms / args.profile_reps
    nfluid = nodes.num_bulk_points + nodes.num_layer_points
    stages = {}
    for s in (1, 2, 3, 4):
        ms = prof[f"s{s}_rk4_stage"]
        b = nfluid * stage_bytes_per_node(Q, D, k, s)
        stages[f"s{s}"] = {"kernel": STAGE_KERNELS[s], "bytes_per_launch": b, "avg_ms": ms,
                           "frac": b / (ms * 1e-3) / HBM_PEAK,
                           "traffic": measured_traffic(h, s)}
    top = stages[f"s{ROOFLINE_STAGE}"]
    achieved = top["bytes_per_launch"] / (top["avg_ms"] * 1e-3)
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": top["traffic"],
                "kernel": top["kernel"] + f" (RK4 stage {ROOFLINE_STAGE}, the longest kernel)",
                "bytes_per_launch": top["bytes_per_launch"], "avg_ms": top["avg_ms"],
                "stages": stages}
    model = step_model_bytes(nodes, Q, D, k)
    out = {
        "metric": "D3Q27 nonlocal-LBM node-updates/s @1M nodes",
        "value": value, "unit": "node-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "dtype_note": "FP32 state and stencils; FP64 collision and setup solves",
        "data": "synthetic",
        "config": {"workload": f"D3Q27 unit cube h={h} jittered 0.25h, k=32 r^3, "
                               f"PML 8h sigma_max 0.5, tau 0.6, dt 0.4h",
                   "nodes": N, "bulk": nodes.num_bulk_points, "layer": nodes.num_layer_points,
                   "boundary": nodes.num_boundary_points},
        "gpu_event_ms_per_step": gpu_ms / args.steps,
        "setup_s": t_setup, "nodes_s": t_nodes, "setup_phases_ms": eng.setup_phases(),
        "step_model_bytes_per_node_update": model,
        "step_model_hbm_frac": value * model / HBM_PEAK,
        "kernel_ms": prof,
        "roofline": roofline,
        "copy_gbs": copy_bandwidth() / 1e9,
    }
    if not args.no_batched:
        eng.close()
        for name, fn in (("batched", lambda: batched_workload(args.batched_steps)),
                         ("served", lambda: served_workload(args.served_requests))):
            try:
                out[name] = fn()
            except Exception as exc:  # reported, not fatal to the headline line
                out[name] = {"error": f"{type(exc).__name__}: {exc}"}
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if not args.no_other_configs and not args.no_batched:
        try:
            out["other_configs"] = other_configs(threads)
        except Exception as exc:  # reported, not fatal to the headline line
            out["other_configs"] = {"error": f"{type(exc).__name__}: {exc}"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(out, default=float))
    return 0


if __name__ == "__main__":
    sys.exit(main())
