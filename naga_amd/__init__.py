"""MI355X-native nonlocal lattice-Boltzmann acoustics (NAGA's nonlocal LBM
step and setup) -- Python host mirror of the reference interface.

The compute path is libnaga_amd.so (HIP kernels for gfx950 behind the C-ABI
declared in include/naga_amd.h).  This module only marshals host arrays:
it loads the library or raises, it never computes a fallback.

Names follow the reference's namespaces:

* ``naga::fluids::nonlocal_lbm``: :class:`SimulationEngine`,
  :class:`SimulationNodes`, :class:`UniformGridProvider`, :func:`lattice_tables`;
* ``naga::segmentation``: :func:`batched_nearest_neighbors`;
* ``naga::linalg``: :func:`batched_lu_solve` (replaces batched_matrix_inverse);
* ``naga::interpolation``: :func:`rpim_weights` (radial_point_method);
* ``naga::nonlocal_calculus``: :class:`DivergenceOperator`.

Errors raise ``ValueError`` where the reference throws std::invalid_argument
and ``RuntimeError`` where it throws std::runtime_error, with the
reference's "naga::..." prefixes.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnaga_amd.so")

CS_LAT = 0.57735026918962576450914878050196
RBF_MULTIQUADRIC = 0
RBF_POLYHARMONIC3 = 1

_ENGINE_PREFIX = "naga::fluids::nonlocal_lbm::simulation_engine::"


class SceneConfig(C.Structure):
    _fields_ = [
        ("lattice", C.c_int32), ("k", C.c_int32), ("k_boundary", C.c_int32),
        ("rbf", C.c_int32), ("rbf_boundary", C.c_int32),
        ("pml_ref_index", C.c_int32), ("k_min", C.c_int32),
        ("nu", C.c_double), ("rho0", C.c_double), ("dt", C.c_double),
        ("c", C.c_double), ("h", C.c_double), ("filter_strength", C.c_double),
        ("solve_precision", C.c_int32),
    ]


MAX_OBSTACLES = 8


class Request(C.Structure):
    """naga_request: one acoustic-pulse solve for the serving loop."""
    _fields_ = [
        ("lattice", C.c_int32), ("k", C.c_int32), ("steps", C.c_int32),
        ("n_obstacles", C.c_int32), ("h", C.c_double), ("jitter", C.c_double),
        ("seed", C.c_uint64), ("obstacles", C.c_double * (MAX_OBSTACLES * 4)),
        ("layer_thickness", C.c_double), ("peak_absorption", C.c_double),
        ("tau", C.c_double), ("pulse_center", C.c_double * 3),
        ("pulse_amplitude", C.c_double), ("pulse_width", C.c_double),
        ("k_min", C.c_int32), ("refine_width", C.c_double),
        ("immersed_layer_thickness", C.c_double), ("immersed_absorption", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
        ("resident", C.c_int64), ("queued", C.c_int64), ("bytes_in_use", C.c_int64),
        ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
        ("node_updates_per_s", C.c_double),
    ]


_lib = None
_BATCH_PREFIX = "naga::batch: "


def load() -> C.CDLL:
    """Load libnaga_amd.so; raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m naga_amd.build` "
            "(the HIP library is the only compute path)")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i32, i64, f64, u64 = C.c_int32, C.c_int64, C.c_double, C.c_uint64
    sig = {
        "naga_last_error": (C.c_char_p, []),
        "naga_lattice_tables": (C.c_int, [i32, vp, vp, vp, vp]),
        "naga_collide": (C.c_int, [i32, vp, f64, vp, f64, f64, vp]),
        "naga_engine_create": (C.c_int, [C.POINTER(SceneConfig), C.POINTER(vp)]),
        "naga_engine_destroy": (C.c_int, [vp]),
        "naga_engine_init_domain": (C.c_int, [vp, vp, i64, i64, i64, vp, vp]),
        "naga_engine_set_pulse": (C.c_int, [vp, vp, f64, f64]),
        "naga_engine_set_distributions": (C.c_int, [vp, vp]),
        "naga_engine_step": (C.c_int, [vp, i32]),
        "naga_engine_step_timed": (C.c_int, [vp, i32, vp]),
        "naga_engine_profile_step": (C.c_int, [vp, vp]),
        "naga_engine_get_macroscopic": (C.c_int, [vp, vp, vp, vp, i32]),
        "naga_engine_get_distributions": (C.c_int, [vp, vp, vp]),
        "naga_engine_get_setup": (C.c_int, [vp, vp, vp, vp, vp, vp]),
        "naga_engine_get_pml": (C.c_int, [vp, vp, vp]),
        "naga_engine_info": (C.c_int, [vp, vp, vp, vp, vp]),
        "naga_engine_setup_phases": (C.c_int, [vp, vp]),
        "naga_engine_write_vtp": (C.c_int, [vp, C.c_char_p, i32]),
        "naga_batched_matrix_inverse_f64": (C.c_int, [vp, i32, i64, vp, vp]),
        "naga_rpim_scratchpad_size": (i64, [i64, i32, i32, i32, i32]),
        "naga_rpim_create": (C.c_int, [vp, i64, vp, i32, vp, i64, i32, i32, i32, f64,
                                       C.POINTER(vp)]),
        "naga_rpim_destroy": (C.c_int, [vp]),
        "naga_rpim_interpolate": (C.c_int, [vp, vp, i64, i32, vp, C.c_float]),
        "naga_rpim_get_weights": (C.c_int, [vp, vp]),
        "naga_segmentation_create": (C.c_int, [vp, i64, i32, i32, f64, C.POINTER(vp)]),
        "naga_segmentation_destroy": (C.c_int, [vp]),
        "naga_segmentation_shape": (C.c_int, [vp, vp, vp, vp, vp]),
        "naga_segmentation_partitions": (C.c_int, [vp, vp, vp, vp]),
        "naga_segmentation_knn": (C.c_int, [vp, vp, i64, i32, vp, vp]),
        "naga_rk4_stage": (C.c_int, [i32, i64, i64, i32, vp, vp, f64, i32, vp, vp, vp, vp, vp]),
        "naga_uniform_grid_size": (i64, [i32, vp, vp, f64]),
        "naga_uniform_grid": (C.c_int, [i32, vp, vp, f64, f64, u64, f64, f64, vp, vp, vp, vp]),
        "naga_knn": (C.c_int, [vp, i64, vp, i64, i32, i32, vp, vp]),
        "naga_knn_noalloc": (C.c_int, [vp, i64, vp, i64, i32, i32, vp, vp, vp]),
        "naga_batched_lu_solve_f64": (C.c_int, [vp, i32, i64, vp, i32, vp, vp]),
        "naga_rpim_weights": (C.c_int, [vp, i64, vp, i32, vp, i64, i32, i32, f64, vp]),
        "naga_divergence_weights": (C.c_int, [vp, i64, vp, i64, i32, i32, i32, vp, vp]),
        "naga_divergence_apply": (C.c_int, [vp, vp, i64, i32, i32, vp, i64, C.c_float, vp]),
        "naga_engine_add_density_source": (C.c_int, [vp, vp, f64, f64, f64, f64, f64, f64]),
        "naga_engine_clear_density_sources": (C.c_int, [vp]),
        "naga_batch_create": (C.c_int, [C.POINTER(vp)]),
        "naga_batch_destroy": (C.c_int, [vp]),
        "naga_batch_add": (C.c_int, [vp, vp]),
        "naga_batch_remove": (C.c_int, [vp, vp]),
        "naga_batch_set_graph": (C.c_int, [vp, i32]),
        "naga_batch_step": (C.c_int, [vp, i32]),
        "naga_batch_step_timed": (C.c_int, [vp, i32, vp]),
        "naga_obstacle_grid": (C.c_int, [i32, vp, vp, f64, f64, u64, i32, vp, f64, f64, f64,
                                         f64, f64, i64, vp, vp, vp, vp]),
        "naga_server_create": (C.c_int, [i64, i32, i32, i32, i32, C.POINTER(vp)]),
        "naga_server_destroy": (C.c_int, [vp]),
        "naga_server_submit": (C.c_int, [vp, C.POINTER(Request), C.POINTER(i64)]),
        "naga_server_status": (C.c_int, [vp, i64, C.POINTER(i32)]),
        "naga_server_result": (C.c_int, [vp, i64, C.POINTER(i64), vp, vp, vp, C.POINTER(f64)]),
        "naga_server_release": (C.c_int, [vp, i64]),
        "naga_server_drain": (C.c_int, [vp]),
        "naga_server_get_stats": (C.c_int, [vp, C.POINTER(ServerStats)]),
        "naga_batch_autotune": (C.c_int, [vp, i32, C.POINTER(i32)]),
        "naga_batch_set_q9_kernels": (C.c_int, [vp, i32, i32]),
        "naga_batch_tuning": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                        C.POINTER(i32)]),
        "naga_engine_state_size": (C.c_int, [vp, C.POINTER(i64)]),
        "naga_engine_save_state": (C.c_int, [vp, vp, i64]),
        "naga_engine_load_state": (C.c_int, [vp, vp, i64]),
        "naga_engine_add_probes": (C.c_int, [vp, vp, i64, i32, C.POINTER(i32)]),
        "naga_engine_sample_probes": (C.c_int, [vp, i32, i32, vp]),
        "naga_engine_get_stencil_sizes": (C.c_int, [vp, vp]),
        "naga_advection_create": (C.c_int, [vp, i64, i32, i32, i32, C.POINTER(vp)]),
        "naga_advection_from_weights": (C.c_int, [vp, vp, i64, i32, i32, C.POINTER(vp)]),
        "naga_advection_destroy": (C.c_int, [vp]),
        "naga_advection_step": (C.c_int, [vp, vp, i32, vp, f64, f64, i64, i32]),
        "naga_advection_weights": (C.c_int, [vp, vp, vp]),
        "naga_csr_create": (C.c_int, [i64, i64, vp, vp, vp, C.POINTER(vp)]),
        "naga_csr_from_stencil": (C.c_int, [i64, i64, i32, vp, vp, C.POINTER(vp)]),
        "naga_csr_destroy": (C.c_int, [vp]),
        "naga_csr_mult": (C.c_int, [vp, C.c_float, C.c_float, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


EXPORTED = [
    "naga_last_error", "naga_lattice_tables", "naga_collide",
    "naga_engine_create", "naga_engine_destroy", "naga_engine_init_domain",
    "naga_engine_set_pulse", "naga_engine_set_distributions", "naga_engine_step",
    "naga_engine_step_timed", "naga_engine_profile_step",
    "naga_engine_get_macroscopic", "naga_engine_get_distributions",
    "naga_engine_get_setup", "naga_engine_get_pml", "naga_engine_info",
    "naga_engine_setup_phases", "naga_engine_write_vtp", "naga_batched_matrix_inverse_f64",
    "naga_rpim_scratchpad_size", "naga_rpim_create", "naga_rpim_destroy",
    "naga_rpim_interpolate", "naga_rpim_get_weights", "naga_segmentation_create",
    "naga_segmentation_destroy", "naga_segmentation_shape", "naga_segmentation_partitions",
    "naga_segmentation_knn", "naga_rk4_stage",
    "naga_uniform_grid_size", "naga_uniform_grid", "naga_knn", "naga_knn_noalloc",
    "naga_batched_lu_solve_f64", "naga_rpim_weights", "naga_divergence_weights",
    "naga_divergence_apply", "naga_engine_add_density_source",
    "naga_engine_clear_density_sources", "naga_batch_create", "naga_batch_destroy",
    "naga_batch_add", "naga_batch_remove", "naga_batch_set_graph", "naga_batch_step",
    "naga_batch_step_timed", "naga_obstacle_grid", "naga_server_create",
    "naga_server_destroy", "naga_server_submit", "naga_server_status", "naga_server_result",
    "naga_server_release", "naga_server_drain", "naga_server_get_stats",
    "naga_batch_autotune", "naga_batch_set_q9_kernels", "naga_batch_tuning",
    "naga_engine_state_size", "naga_engine_save_state",
    "naga_engine_load_state", "naga_engine_add_probes", "naga_engine_sample_probes",
    "naga_engine_get_stencil_sizes", "naga_advection_create", "naga_advection_from_weights",
    "naga_advection_destroy", "naga_advection_step", "naga_advection_weights",
    "naga_csr_create", "naga_csr_from_stencil", "naga_csr_destroy", "naga_csr_mult",
]

STEP_KERNELS = ["p1_collide", "pl_pml_local", "pm_layer_means", "pf_pml_field",
                "p2_pml_divergence", "p3_boundary_interp", "p4_filter_bounce",
                "s1_rk4_stage", "s2_rk4_stage", "s3_rk4_stage", "s4_rk4_stage",
                "p5_boundary_interp"]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(status: int, prefix: str = "naga::") -> None:
    if status == 0:
        return
    msg = prefix + load().naga_last_error().decode(errors="replace")
    if status == 1:
        raise ValueError(msg)
    if status == 2:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- lattices
def lattice_tables(Q: int):
    """Velocities c[Q, D], weights w[Q], bounce-back bb[Q], collision
    matrix K[Q, Q] of D2Q9 (Q = 9) or D3Q27 (Q = 27)."""
    lib = load()
    D = 2 if Q == 9 else 3
    c = np.zeros(Q * D)
    w = np.zeros(Q)
    bb = np.zeros(Q, np.int32)
    K = np.zeros(Q * Q)
    _check(lib.naga_lattice_tables(Q, _ptr(c), _ptr(w), _ptr(bb), _ptr(K)))
    return c.reshape(Q, D), w, bb, K.reshape(Q, Q)


def collide(Q, f, rho, u, nu_lat, dt_lat):
    """Host (FP64) evaluation of the device collision formula for one node."""
    lib = load()
    f = _f64(f)
    u = _f64(u)
    df = np.zeros(Q)
    _check(lib.naga_collide(Q, _ptr(f), float(rho), _ptr(u), float(nu_lat),
                            float(dt_lat), _ptr(df)))
    return df


# ------------------------------------------------------------------- nodes
@dataclasses.dataclass
class SimulationNodes:
    """simulation_nodes<T> (fluids/nonlocal_lattice_boltzmann/
    simulation_nodes.cuh:41-66), points stored [N, D] (one row per node)."""
    points: np.ndarray
    boundary_normals: np.ndarray
    layer_absorption: np.ndarray
    num_bulk_points: int
    num_layer_points: int
    num_boundary_points: int
    nodal_spacing: float
    num_ghost_nodes: int = 0

    @property
    def size(self) -> int:
        return self.num_bulk_points + self.num_layer_points + self.num_boundary_points

    def as_dict(self) -> dict:
        return dict(points=self.points, normals=self.boundary_normals,
                    sigma=self.layer_absorption, n_bulk=self.num_bulk_points,
                    n_layer=self.num_layer_points, n_boundary=self.num_boundary_points)


class UniformGridProvider:
    """uniform_grid_provider<Lattice> (node_providers/uniform_grid_provider.cuh:
    45-237) with seeded jitter of the interior nodes (SplitMix64)."""

    def __init__(self, min_point, max_point, grid_spacing,
                 absorption_layer_thickness=0.0, peak_absorption_rate=0.0,
                 jitter=0.0, seed=0x4E414741):
        self.lo = _f64(min_point)
        self.hi = _f64(max_point)
        self.h = float(grid_spacing)
        self.layer = float(absorption_layer_thickness)
        self.peak = float(peak_absorption_rate)
        self.jitter = float(jitter)
        self.seed = int(seed)

    def get(self) -> SimulationNodes:
        lib = load()
        D = self.lo.size
        N = lib.naga_uniform_grid_size(D, _ptr(self.lo), _ptr(self.hi), self.h)
        if N <= 0:
            raise ValueError("naga::ranges::uniform_grid:: invalid grid")
        pts = np.zeros(N * D, np.float32)
        nrm = np.zeros(N * D, np.float32)
        sig = np.zeros(N, np.float32)
        counts = np.zeros(3, np.int64)
        _check(lib.naga_uniform_grid(D, _ptr(self.lo), _ptr(self.hi), self.h,
                                     self.jitter, self.seed, self.layer, self.peak,
                                     _ptr(pts), _ptr(nrm), _ptr(sig), _ptr(counts)),
               "naga::fluids::nonlocal_lbm::uniform_grid_provider::")
        nb, nl, nbd = (int(x) for x in counts)
        return SimulationNodes(pts.reshape(N, D), nrm[: nbd * D].reshape(nbd, D).copy(),
                               sig[:nl].copy(), nb, nl, nbd, self.h)


# ------------------------------------------------------------------ engine
def problem_parameters(lattice: int, tau: float, dt: float, c: float | None = None):
    """(nu, dt, c) giving relaxation time tau: D3Q27 relaxes shear at
    1/(3 nu_lat + 1/2)/dt_lat (detail/d3q27_lattice.cuh:177), D2Q9 at
    1/(3 nu_lat + dt_lat/2) (detail/d2q9_lattice.cuh:136)."""
    c = CS_LAT if c is None else c
    ratio = CS_LAT / c
    dt_lat = dt / ratio
    nu_lat = (tau - 0.5) / 3.0 if lattice == 27 else (tau - 0.5) * dt_lat / 3.0
    return dict(nu=nu_lat * ratio, dt=dt, c=c)


class DensitySource:
    """A density source (the declarative form of density_source<Lattice>,
    density_source.cuh:41-82): at the start of each step at time t, nodes
    within `radius` of `center` receive
    s = amplitude * cos(2 pi frequency t + phase) * P(r / radius),
    P(x) = (1 - (2x)^2 / 4) exp(-(2x)^2), while t_start <= t < t_end."""

    def __init__(self, center, radius, amplitude, frequency=0.0, phase=0.0,
                 t_start=0.0, t_end=float("inf")):
        self.center = list(center)
        self.radius = float(radius)
        self.amplitude = float(amplitude)
        self.frequency = float(frequency)
        self.phase = float(phase)
        self.t_start = float(t_start)
        self.t_end = float(t_end)


class ObstacleGridProvider:
    """Box minus circles/spheres (naga_obstacle_grid): obstacles is a list of
    (centre, radius); layer absorption at the box faces and, with
    immersed_layer_thickness and immersed_absorption_rate > 0, around the
    obstacles (the immersed boundary layers of
    conforming_point_cloud_provider.cuh:250-280)."""

    def __init__(self, min_point, max_point, grid_spacing, obstacles=(),
                 absorption_layer_thickness=0.0, peak_absorption_rate=0.0,
                 jitter=0.0, seed=0x4E414741, refine_width=0.0,
                 immersed_layer_thickness=0.0, immersed_absorption_rate=0.0):
        self.refine_width = float(refine_width)
        self.immersed_layer = float(immersed_layer_thickness)
        self.immersed_rate = float(immersed_absorption_rate)
        self.lo = _f64(min_point)
        self.hi = _f64(max_point)
        self.D = self.lo.size
        self.h = float(grid_spacing)
        self.obst = np.array([list(c) + [r] for c, r in obstacles], np.float64).reshape(-1, self.D + 1)
        self.layer = float(absorption_layer_thickness)
        self.peak = float(peak_absorption_rate)
        self.jitter = float(jitter)
        self.seed = int(seed)

    def get(self) -> "SimulationNodes":
        lib = load()
        counts = np.zeros(3, np.int64)
        ob = np.ascontiguousarray(self.obst.reshape(-1))
        args = (self.D, _ptr(self.lo), _ptr(self.hi), self.h, self.jitter, C.c_uint64(self.seed),
                len(self.obst), _ptr(ob) if ob.size else None, self.refine_width, self.layer,
                self.peak, self.immersed_layer, self.immersed_rate)
        _check(lib.naga_obstacle_grid(*args, 0, None, None, None, _ptr(counts)),
               "naga::nonlocal_lbm::obstacle_grid_provider::get: ")
        n = int(counts.sum())
        pts = np.zeros((n, self.D), np.float32)
        nrm = np.zeros((max(int(counts[2]), 1), self.D), np.float32)
        sig = np.zeros(max(int(counts[1]), 1), np.float32)
        _check(lib.naga_obstacle_grid(*args, n, _ptr(pts), _ptr(nrm), _ptr(sig), _ptr(counts)),
               "naga::nonlocal_lbm::obstacle_grid_provider::get: ")
        return SimulationNodes(pts, nrm[: counts[2]], sig[: counts[1]], int(counts[0]),
                               int(counts[1]), int(counts[2]), self.h)


class SimulationEngine:
    """simulation_engine<Lattice> (fluids/nonlocal_lattice_boltzmann/
    simulation_engine.cuh:39-121) backed by one GPU-resident scene."""

    def __init__(self, lattice: int = 27, k: int | None = None,
                 k_boundary: int | None = None, rbf: int = RBF_POLYHARMONIC3,
                 rbf_boundary: int | None = None, pml_ref_index: bool = True,
                 k_min: int = 0, filter_strength: float = 0.0, solve_precision: int = 64):
        """filter_strength: selective filter strength per step; 0 gives the
        reference's 0.75 (detail/simulation_engine.cuh:1242), a negative
        value no filter.  solve_precision: 64 (FP64 RBF weight solves) or 32
        (FP32, the reference's; with k = 32 and multiquadric RBFs the
        reference-parity mode)."""
        if lattice not in (9, 27):
            raise ValueError(_ENGINE_PREFIX + "lattice must be D2Q9 or D3Q27")
        self.k_min = int(k_min)
        self.filter_strength = float(filter_strength)
        self.solve_precision = int(solve_precision)
        self.lattice = lattice
        self.dimensions = 2 if lattice == 9 else 3
        self.k = 32 if k is None else int(k)
        self.k_boundary = self.k if k_boundary is None else int(k_boundary)
        self.rbf = int(rbf)
        self.rbf_boundary = self.rbf if rbf_boundary is None else int(rbf_boundary)
        self.pml_ref_index = bool(pml_ref_index)
        self._params = None
        self._h = None
        self._nodes = None

    # set_problem_parameters (detail/simulation_engine.cuh:94-124)
    def set_problem_parameters(self, fluid_viscosity, nominal_density, time_step,
                               speed_of_sound, characteristic_frequency=0.0):
        self._params = dict(nu=float(fluid_viscosity), rho0=float(nominal_density),
                            dt=float(time_step), c=float(speed_of_sound))

    def init_domain(self, nodes):
        """init_domain(simulation_nodes | node_provider)."""
        if self._params is None:
            raise ValueError(_ENGINE_PREFIX + "set_problem_parameters must be called first")
        if hasattr(nodes, "get") and not isinstance(nodes, SimulationNodes):
            nodes = nodes.get()
        D = self.dimensions
        pts = _f32(nodes.points).reshape(-1)
        if np.asarray(nodes.points).ndim != 2 or np.asarray(nodes.points).shape[1] != D:
            raise ValueError(_ENGINE_PREFIX + "The number of dimensions in the simulation "
                             "domain does not match the number of dimensions in the lattice.")
        if nodes.num_ghost_nodes:
            raise ValueError(_ENGINE_PREFIX + "ghost nodes are not supported")
        if pts.size != nodes.size * D:
            raise ValueError(_ENGINE_PREFIX + "The number of points in the simulation domain "
                             "does not match the sum of the number of bulk, layer, and "
                             "boundary points.")
        nrm = _f32(nodes.boundary_normals).reshape(-1)
        if nrm.size != nodes.num_boundary_points * D:
            raise ValueError(_ENGINE_PREFIX + "The number of boundary normals does not match "
                             "the number of boundary points.")
        sig = _f32(nodes.layer_absorption).reshape(-1)
        if sig.size != nodes.num_layer_points:
            raise ValueError(_ENGINE_PREFIX + "The number of layer absorption coefficients "
                             "does not match the number of layer points.")
        lib = load()
        self.close()
        cfg = SceneConfig(lattice=self.lattice, k=self.k, k_boundary=self.k_boundary,
                          rbf=self.rbf, rbf_boundary=self.rbf_boundary,
                          pml_ref_index=int(self.pml_ref_index), k_min=self.k_min,
                          nu=self._params["nu"], rho0=self._params["rho0"],
                          dt=self._params["dt"], c=self._params["c"],
                          h=float(nodes.nodal_spacing), filter_strength=self.filter_strength,
                          solve_precision=self.solve_precision)
        h = C.c_void_p()
        _check(lib.naga_engine_create(C.byref(cfg), C.byref(h)), _ENGINE_PREFIX)
        self._h = h
        self._cfg = cfg
        _check(lib.naga_engine_init_domain(h, _ptr(pts), nodes.num_bulk_points,
                                           nodes.num_layer_points,
                                           nodes.num_boundary_points,
                                           _ptr(nrm) if nrm.size else None,
                                           _ptr(sig) if sig.size else None), _ENGINE_PREFIX)
        self._nodes = nodes

    def close(self):
        if self._h is not None:
            load().naga_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _live(self):
        if self._h is None:
            raise ValueError(_ENGINE_PREFIX + "init_domain must be called first")
        return self._h

    @property
    def num_nodes(self) -> int:
        return self._nodes.size

    def set_initial_pulse(self, center, amplitude, width):
        _check(load().naga_engine_set_pulse(self._live(), _ptr(_f64(center)),
                                            float(amplitude), float(width)), _ENGINE_PREFIX)

    def reset(self):
        """reset() (detail/simulation_engine.cuh:1377-1386): f = w, time 0."""
        self.set_initial_pulse(np.zeros(self.dimensions), 0.0, 1.0)

    def add_probes(self, points, rbf: int = RBF_MULTIQUADRIC) -> int:
        """Density probe points (density_observer, density_observer.cuh:
        62-113); returns the probe set's id."""
        pts = _f32(points).reshape(-1, self.dimensions)
        sid = C.c_int32()
        _check(load().naga_engine_add_probes(self._live(), _ptr(pts), pts.shape[0], int(rbf),
                                             C.byref(sid)), _ENGINE_PREFIX)
        self._probe_sizes = getattr(self, "_probe_sizes", {})
        self._probe_sizes[sid.value] = pts.shape[0]
        return sid.value

    def sample_probes(self, probe_set: int, fresh: bool = False) -> np.ndarray:
        out = np.zeros(self._probe_sizes[probe_set])
        _check(load().naga_engine_sample_probes(self._live(), int(probe_set), int(fresh),
                                                _ptr(out)), _ENGINE_PREFIX)
        return out

    def stencil_sizes(self) -> np.ndarray:
        """k_i per node (all k unless k_min was given)."""
        out = np.zeros(self.num_nodes, np.uint8)
        _check(load().naga_engine_get_stencil_sizes(self._live(), _ptr(out)), _ENGINE_PREFIX)
        return out

    def save_state(self) -> bytes:
        """Snapshot of the state carried between steps (save_state,
        simulation_engine.cuh:108-116) as bytes."""
        lib = load()
        n = C.c_int64()
        _check(lib.naga_engine_state_size(self._live(), C.byref(n)), _ENGINE_PREFIX)
        buf = C.create_string_buffer(n.value)
        _check(lib.naga_engine_save_state(self._live(), buf, n.value), _ENGINE_PREFIX)
        return buf.raw

    def load_state(self, state: bytes):
        _check(load().naga_engine_load_state(self._live(), state, len(state)), _ENGINE_PREFIX)

    def register_density_source(self, source: "DensitySource"):
        """register_density_source (simulation_engine.cuh:79-81) for a
        declarative source; see DensitySource."""
        c = np.zeros(3)
        c[: self.dimensions] = np.asarray(source.center, np.float64)[: self.dimensions]
        _check(load().naga_engine_add_density_source(
            self._live(), _ptr(c), float(source.radius), float(source.amplitude),
            float(source.frequency), float(source.phase), float(source.t_start),
            float(source.t_end)), _ENGINE_PREFIX)

    def clear_density_sources(self):
        _check(load().naga_engine_clear_density_sources(self._live()), _ENGINE_PREFIX)

    def set_distributions(self, f):
        _check(load().naga_engine_set_distributions(self._live(), _ptr(_f64(f))),
               _ENGINE_PREFIX)

    def step_forward(self, nsteps: int = 1):
        """step_forward() x nsteps (detail/simulation_engine.cuh:1325-1375).
        With observers attached each step first hands them the state at its
        start (update_observers, :1332), then advances one step."""
        obs = getattr(self, "_observers", None)
        if not obs:
            _check(load().naga_engine_step(self._live(), int(nsteps)), _ENGINE_PREFIX)
            return
        for _ in range(int(nsteps)):
            sol = self.solution(fresh=True)
            t = self.time()
            for o in list(obs):
                o.update(t, self._nodes, self._params, sol)
            _check(load().naga_engine_step(self._live(), 1), _ENGINE_PREFIX)

    def attach_observer(self, observer):
        """attach_observer (simulation_engine.cuh:90-92): an object with
        update(time, domain, parameters, solution), called at the start of
        every step with the macroscopic values of that moment."""
        self._observers = getattr(self, "_observers", [])
        self._observers.append(observer)

    def detach_observer(self, observer):
        obs = getattr(self, "_observers", [])
        self._observers = [o for o in obs if o is not observer]

    def write_vtp(self, path, fresh: bool = False):
        """vtk_observer's output format for this scene (see naga_engine_write_vtp)."""
        _check(load().naga_engine_write_vtp(self._live(), os.fsencode(str(path)), int(fresh)),
               _ENGINE_PREFIX)

    def step_timed(self, nsteps: int) -> float:
        ms = np.zeros(1)
        _check(load().naga_engine_step_timed(self._live(), int(nsteps), _ptr(ms)),
               _ENGINE_PREFIX)
        return float(ms[0])

    def profile_step(self) -> dict:
        ms = np.zeros(len(STEP_KERNELS))
        _check(load().naga_engine_profile_step(self._live(), _ptr(ms)), _ENGINE_PREFIX)
        return dict(zip(STEP_KERNELS, ms.tolist()))

    SETUP_PHASES = ("knn", "divergence_weights", "filter_slabs_unions", "boundary_interpolator",
                    "pml_and_state")

    def setup_phases(self) -> dict:
        """Wall time (ms) of the last init_domain's phases on the engine's
        stream."""
        ms = np.zeros(len(self.SETUP_PHASES))
        _check(load().naga_engine_setup_phases(self._live(), _ptr(ms)), _ENGINE_PREFIX)
        return dict(zip(self.SETUP_PHASES, ms.tolist()))

    def info(self) -> dict:
        n = np.zeros(1, np.int64)
        t = np.zeros(1)
        fr = np.zeros(1, np.int64)
        sm = np.zeros(1)
        _check(load().naga_engine_info(self._live(), _ptr(n), _ptr(t), _ptr(fr), _ptr(sm)))
        return dict(nodes=int(n[0]), time=float(t[0]), frame=int(fr[0]), setup_ms=float(sm[0]))

    def time(self) -> float:
        return self.info()["time"]

    def macroscopic(self, fresh: bool = True) -> dict:
        N, D = self.num_nodes, self.dimensions
        rho = np.zeros(N, np.float32)
        u = np.zeros(D * N, np.float32)
        drho = np.zeros(N)
        _check(load().naga_engine_get_macroscopic(self._live(), _ptr(rho), _ptr(u),
                                                  _ptr(drho), int(fresh)), _ENGINE_PREFIX)
        return dict(rho=rho, u=u.reshape(D, N), drho=drho)

    def distributions(self) -> dict:
        N, Q = self.num_nodes, self.lattice
        f = np.zeros(Q * N)
        g = np.zeros(Q * N)
        _check(load().naga_engine_get_distributions(self._live(), _ptr(f), _ptr(g)),
               _ENGINE_PREFIX)
        return dict(f=f.reshape(Q, N), g=g.reshape(Q, N))

    def solution(self, fresh: bool = False) -> dict:
        """state_variables: distributions plus macroscopic values as held by
        the engine (reference semantics, computed at the start of the last
        step), or fresh ones of the current state."""
        out = self.distributions()
        out.update(self.macroscopic(fresh=fresh))
        return out

    def setup_products(self) -> dict:
        N, k, kb, D = self.num_nodes, self.k, self.k_boundary, self.dimensions
        nb = self._nodes.num_boundary_points
        idx = np.zeros(N * k, np.uint32)
        W = np.zeros(N * k * D, np.float32)
        filt = np.zeros(N * k, np.float32)
        bidx = np.zeros(max(nb, 1) * kb, np.uint32)
        bw = np.zeros(max(nb, 1) * kb, np.float32)
        _check(load().naga_engine_get_setup(self._live(), _ptr(idx), _ptr(W), _ptr(filt),
                                            _ptr(bidx), _ptr(bw)), _ENGINE_PREFIX)
        return dict(idx=idx.reshape(N, k), W=W.reshape(N, k, D), filt=filt.reshape(N, k),
                    bidx=bidx[: nb * kb].reshape(nb, kb), bw=bw[: nb * kb].reshape(nb, kb))

    def pml_state(self):
        N, Q = self.num_nodes, self.lattice
        e = np.zeros(Q * N)
        q = np.zeros(Q * N)
        _check(load().naga_engine_get_pml(self._live(), _ptr(e), _ptr(q)), _ENGINE_PREFIX)
        return e.reshape(Q, N), q.reshape(Q, N)


# ------------------------------------------------------- setup primitives
def batched_nearest_neighbors(k: int, query, ref):
    """segmentation::batched_nearest_neighbors (segmentation/nearest_neighbors.cuh:
    454-477), exact: returns (distances_squared [nq, k], indices [nq, k])."""
    ref = _f32(ref)
    query = _f32(query)
    D = ref.shape[1]
    nq = query.shape[0]
    d2 = np.zeros(nq * k, np.float32)
    idx = np.zeros(nq * k, np.uint32)
    _check(load().naga_knn(_ptr(ref), ref.shape[0], _ptr(query), nq, D, int(k),
                           _ptr(d2), _ptr(idx)), "naga::segmentation::")
    return d2.reshape(nq, k), idx.reshape(nq, k)


def batched_lu_solve(A, B):
    """Solve A_b X_b = B_b for a batch of n x n systems in FP64 (replaces
    linalg::batched_matrix_inverse, linalg/batched_matrix_inverse.cuh:325-397).
    A: [batch, n, n], B: [batch, n] or [batch, n, nrhs].  Returns (X, info)."""
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    batch, n, _ = A.shape
    vec = B.ndim == 2
    B3 = B[:, :, None] if vec else B
    nrhs = B3.shape[2]
    Ac = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))  # column-major per item
    Bc = np.ascontiguousarray(np.transpose(B3, (0, 2, 1)))
    X = np.zeros_like(Bc)
    info = np.zeros(batch, np.int32)
    _check(load().naga_batched_lu_solve_f64(_ptr(Ac), n, batch, _ptr(Bc), nrhs, _ptr(X),
                                            _ptr(info)), "naga::linalg::")
    X = np.transpose(X, (0, 2, 1))
    return (X[:, :, 0] if vec else X), info


def rpim_weights(src, idx, query, rbf=RBF_POLYHARMONIC3, spacing=1.0):
    """interpolation::radial_point_method::compute_weights, group size 1:
    weights [nq, k] for interpolating from src[idx[q]] to query[q]."""
    src = _f32(src)
    query = _f32(query)
    idx = np.ascontiguousarray(idx, np.uint32)
    nq, k = idx.shape
    w = np.zeros(nq * k)
    _check(load().naga_rpim_weights(_ptr(src), src.shape[0], _ptr(idx), k, _ptr(query), nq,
                                    src.shape[1], int(rbf), float(spacing), _ptr(w)),
           "naga::interpolation::")
    return w.reshape(nq, k)


def batched_matrix_inverse(A, return_info: bool = False):
    """linalg::batched_matrix_inverse(A) (linalg/batched_matrix_inverse.cuh:
    325-397): A [batch, n, n] -> A^-1 per item (FP64).  The reference leaves
    singular items unchecked; return_info gives the per-item pivot info."""
    A = np.asarray(A, np.float64)
    if A.ndim != 3 or A.shape[1] != A.shape[2]:
        raise ValueError("naga::linalg::Matrices must be square")
    b, n, _ = A.shape
    Ac = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))  # column-major items
    out = np.zeros_like(Ac)
    info = np.zeros(b, np.int32)
    _check(load().naga_batched_matrix_inverse_f64(_ptr(Ac), n, b, _ptr(out), _ptr(info)),
           "naga::linalg::")
    inv = np.transpose(out, (0, 2, 1))
    return (inv, info) if return_info else inv


class RadialPointMethod:
    """interpolation::radial_point_method<T> (interpolation/
    radial_point_method.cuh:92-280): weights of each query over its group's
    stencil (group_size consecutive queries share one stencil and one RBF
    system), and interpolation with them."""

    def __init__(self, handle, n_src, n_q, k):
        self._h, self.n_src, self.n_q, self.k = handle, n_src, n_q, k

    @staticmethod
    def get_scratchpad_size(query_size, support_size, dimensions, group_size=1,
                            dtype=np.float32) -> int:
        v = load().naga_rpim_scratchpad_size(int(query_size), int(support_size),
                                             int(dimensions), int(group_size),
                                             int(np.dtype(dtype).itemsize))
        if v < 0:
            raise ValueError("naga::interpolation::radial_point_method::query_size must be a "
                             "multiple of group_size")
        return int(v)

    @classmethod
    def create_interpolator(cls, source_points, interpolating_indices, query_points,
                            approx_particle_spacing, group_size=1, rbf=RBF_MULTIQUADRIC):
        src = _f32(source_points)
        q = _f32(query_points)
        idx = np.ascontiguousarray(interpolating_indices, np.uint32)
        k = idx.shape[1]
        h = C.c_void_p()
        _check(load().naga_rpim_create(_ptr(src), src.shape[0], _ptr(idx), k, _ptr(q),
                                       q.shape[0], src.shape[1], int(group_size), int(rbf),
                                       float(approx_particle_spacing), C.byref(h)),
               "naga::interpolation::radial_point_method::")
        return cls(h, src.shape[0], q.shape[0], k)

    @classmethod
    def compute_weights(cls, source_points, interpolating_indices, query_points,
                        approx_particle_spacing, group_size=1, rbf=RBF_MULTIQUADRIC):
        r = cls.create_interpolator(source_points, interpolating_indices, query_points,
                                    approx_particle_spacing, group_size, rbf)
        w = r.weights()
        r.close()
        return w

    def weights(self):
        w = np.zeros(self.n_q * self.k, np.float32)
        _check(load().naga_rpim_get_weights(self._h, _ptr(w)),
               "naga::interpolation::radial_point_method::")
        return w.reshape(self.n_q, self.k)

    def interpolate(self, field, centering_offset=0.0):
        """field [n_src] or [n_src, m] -> [n_q] or [n_q, m]."""
        f = _f32(field)
        m = 1 if f.ndim == 1 else f.shape[1]
        out = np.zeros((self.n_q, m), np.float32)
        _check(load().naga_rpim_interpolate(self._h, _ptr(f), f.shape[0], m, _ptr(out),
                                            float(centering_offset)),
               "naga::interpolation::radial_point_method::")
        return out[:, 0] if f.ndim == 1 else out

    def close(self):
        if self._h is not None:
            load().naga_rpim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NdCubicSegmentation:
    """segmentation::nd_cubic_segmentation<T, D>(points, approx_partition_size
    | partition_width) (segmentation/nd_cubic_segmentation.cuh:219-461),
    binned on the device."""

    def __init__(self, points, approx_partition_size: int = 32, partition_width: float = 0.0):
        pts = _f32(points)
        if pts.ndim != 2:
            raise ValueError("naga::The number of dimensions in the points array does not "
                             "match the number of dimensions in the segmentation")
        self.points = pts
        h = C.c_void_p()
        approx = 0 if partition_width > 0 else int(approx_partition_size)
        _check(load().naga_segmentation_create(_ptr(pts), pts.shape[0], pts.shape[1], approx,
                                               float(partition_width), C.byref(h)),
               "naga::segmentation::")
        self._h = h
        shape = np.zeros(3, np.int64)
        lo = np.zeros(3)
        w = C.c_double()
        cnt = C.c_int64()
        _check(load().naga_segmentation_shape(h, _ptr(shape), _ptr(lo), C.byref(w), C.byref(cnt)),
               "naga::segmentation::")
        D = pts.shape[1]
        self.shape = tuple(int(v) for v in shape[:D])
        self.lower_bounds = lo[:D].copy()
        self.partition_width = w.value
        self.upper_bounds = self.lower_bounds + np.array(self.shape) * w.value
        self.partition_count = int(cnt.value)

    def partitions(self):
        """(sizes, index offsets, indices), cells with x fastest."""
        sizes = np.zeros(self.partition_count, np.uint32)
        offs = np.zeros(self.partition_count, np.uint32)
        idx = np.zeros(self.points.shape[0], np.uint32)
        _check(load().naga_segmentation_partitions(self._h, _ptr(sizes), _ptr(offs), _ptr(idx)),
               "naga::segmentation::")
        return sizes, offs, idx

    def batched_nearest_neighbors(self, k: int, query):
        q = _f32(query)
        d2 = np.zeros(q.shape[0] * k, np.float32)
        idx = np.zeros(q.shape[0] * k, np.uint32)
        _check(load().naga_segmentation_knn(self._h, _ptr(q), q.shape[0], int(k), _ptr(d2),
                                            _ptr(idx)), "naga::segmentation::")
        return d2.reshape(-1, k), idx.reshape(-1, k)

    def close(self):
        if self._h is not None:
            load().naga_segmentation_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rk4_stage(stage, y, stencil_idx, weights, dt, n_fluid=None, stage_input=None, y1=None,
              y2=None):
    """One RK4 stage of the streaming equation (naga_rk4_stage): y and the
    stage input are [n, Q] deviation rows; stencil_idx [n_fluid, k],
    weights [n_fluid, k, D]."""
    y = _f32(y)
    n, q = y.shape
    idx = np.ascontiguousarray(stencil_idx, np.uint32)
    W = _f32(weights)
    nf = idx.shape[0] if n_fluid is None else int(n_fluid)
    k = idx.shape[1]
    out = np.zeros_like(y)
    ins = None if stage_input is None else _f32(stage_input)
    a1 = None if y1 is None else _f32(y1)
    a2 = None if y2 is None else _f32(y2)
    _check(load().naga_rk4_stage(q, n, nf, k, _ptr(idx), _ptr(W), float(dt), int(stage),
                                 _ptr(y), None if ins is None else _ptr(ins),
                                 None if a1 is None else _ptr(a1),
                                 None if a2 is None else _ptr(a2), _ptr(out)),
           "naga::nonlocal_lbm::rk4_solver::")
    return out


class DivergenceOperator:
    """nonlocal_calculus::divergence_operator<T, D> (nonlocal_calculus/
    divergence.cuh:44-612)."""

    def __init__(self, weights, support_indices):
        self._W = np.ascontiguousarray(weights, np.float64)
        self._idx = np.ascontiguousarray(support_indices, np.uint32)

    @classmethod
    def create(cls, domain, query=None, k: int = 32, rbf: int = RBF_POLYHARMONIC3,
               override_query_size: int = 0):
        """create(domain[, query_points, override_query_size])
        (nonlocal_calculus/divergence.cuh:53-166): stencils over domain for
        each query point; override_query_size > 0 sizes the operator for that
        many queries, the rows past the query count zero."""
        domain = _f32(domain)
        query = domain if query is None else _f32(query)
        D = domain.shape[1]
        nq = query.shape[0]
        size = nq if int(override_query_size) == 0 else int(override_query_size)
        if size < nq:
            raise ValueError("naga::nonlocal_calculus::divergence_operator::override_query_size "
                             "must be at least as large as the number of query points.")
        idx = np.zeros(size * k, np.uint32)
        W = np.zeros(size * k * D)
        _check(load().naga_divergence_weights(_ptr(domain), domain.shape[0], _ptr(query), nq,
                                              D, int(k), int(rbf), _ptr(idx), _ptr(W)),
               "naga::nonlocal_calculus::divergence_operator::")
        return cls(W.reshape(size, k, D), idx.reshape(size, k))

    def weights(self):
        return self._W

    def support_indices(self):
        return self._idx

    def apply(self, field, centering_offset=0.0):
        """out[i] = sum_s sum_d W(d, s, i) (F_d(idx(s, i)) - offset), FP32."""
        F = _f32(field)
        n, k, D = self._W.shape
        if F.ndim != 2 or F.shape[1] != D:
            raise ValueError("naga::nonlocal_calculus::divergence_operator::"
                             "Field map has incorrect dimensions.")
        out = np.zeros(n, np.float32)
        W32 = _f32(self._W)
        _check(load().naga_divergence_apply(_ptr(self._idx), _ptr(W32), n, k, D, _ptr(F),
                                            F.shape[0], float(centering_offset), _ptr(out)),
               "naga::nonlocal_calculus::divergence_operator::")
        return out


class ConstantVelocityField:
    """constant_velocity_field (nonlocal_calculus/advection.cuh:58-85): the
    same velocity at every node."""

    def __init__(self, velocity):
        self.velocity = _f32(velocity).reshape(-1)


class AdvectionOperator:
    """nonlocal_calculus::advection_operator<T, D> (nonlocal_calculus/
    advection.cuh:88-524): RK4 steps of df/dt = div(-v (f - offset)) over a
    divergence operator; stencils and weights stay on the device."""

    _PREFIX = "naga::nonlocal_calculus::advection_operator::"

    def __init__(self, handle, n, dims, k):
        self._h = handle
        self._n = int(n)
        self._D = int(dims)
        self._k = int(k)

    @classmethod
    def create(cls, domain, k: int = 32, rbf: int = RBF_POLYHARMONIC3):
        """advection_operator::create(domain) (advection.cuh:143-151)."""
        domain = _f32(domain)
        n, D = domain.shape
        h = C.c_void_p()
        _check(load().naga_advection_create(_ptr(domain), n, D, int(k), int(rbf), C.byref(h)),
               cls._PREFIX)
        return cls(h, n, D, k)

    @classmethod
    def from_divergence(cls, div_op: "DivergenceOperator"):
        """advection_operator::create(divergence_op) (advection.cuh:153-159)."""
        W = _f32(div_op.weights())
        idx = np.ascontiguousarray(div_op.support_indices(), np.uint32)
        n, k, D = W.shape
        h = C.c_void_p()
        _check(load().naga_advection_from_weights(_ptr(idx), _ptr(W), n, D, k, C.byref(h)),
               cls._PREFIX)
        return cls(h, n, D, k)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.naga_advection_destroy(h)
            self._h = None

    def divergence_op(self) -> "DivergenceOperator":
        """The operator's stencils and FP32 weights (advection.cuh:494-496)."""
        n, k, D = self._n, self._k, self._D
        W = np.zeros(n * k * D, np.float32)
        idx = np.zeros(n * k, np.uint32)
        _check(load().naga_advection_weights(self._h, _ptr(idx), _ptr(W)), self._PREFIX)
        return DivergenceOperator(W.reshape(n, k, D), idx.reshape(n, k))

    def step_forward(self, velocity_field, f0, dt, centering_offset=0.0,
                     boundary_idx_start=0, steps=1):
        """step_forward (advection.cuh:460-492): `steps` advection_task RK4
        steps (advection.cuh:207-276) from f0; returns f.  velocity_field:
        ConstantVelocityField or an (n, D) per-node array.  Nodes at or past
        boundary_idx_start (> 0) keep their f0 value."""
        f = _f32(f0).copy()
        if f.shape != (self._n,):
            raise ValueError(self._PREFIX + "f0 must have the same number of elements as the "
                             "domain for which this operator was created.")
        if isinstance(velocity_field, ConstantVelocityField):
            v, per = velocity_field.velocity, 0
            if v.shape != (self._D,):
                raise ValueError(self._PREFIX + "velocity has incorrect dimensions.")
        else:
            v, per = _f32(velocity_field), 1
            if v.shape != (self._n, self._D):
                raise ValueError(self._PREFIX + "velocity field has incorrect shape.")
        _check(load().naga_advection_step(self._h, _ptr(v), per, _ptr(f), float(dt),
                                          float(centering_offset), int(boundary_idx_start),
                                          int(steps)), self._PREFIX)
        return f


class CsrMatrix:
    """linalg::matrix<T, sparse_csr> (linalg/matrix.cuh:231-437), device
    resident, FP32 values."""

    _PREFIX = "naga::linalg::matrix::"

    def __init__(self, handle, rows, cols):
        self._h = handle
        self.rows = int(rows)
        self.columns = int(cols)

    @classmethod
    def create(cls, rows, cols, row_offsets, column_indices, values):
        rp = np.ascontiguousarray(row_offsets, np.int64)
        ci = np.ascontiguousarray(column_indices, np.uint32)
        v = _f32(values)
        if rp.shape != (rows + 1,) or ci.shape != v.shape or ci.size != rp[-1]:
            raise ValueError(cls._PREFIX + "inconsistent CSR arrays")
        h = C.c_void_p()
        _check(load().naga_csr_create(rows, cols, _ptr(rp), _ptr(ci), _ptr(v), C.byref(h)),
               cls._PREFIX)
        return cls(h, rows, cols)

    @classmethod
    def create_from_index_stencil(cls, index_stencil, values, columns=None):
        """create_from_index_stencil (matrix.cuh:300-380): row r holds the
        stencil entries index_stencil[r] with values[r], in stencil order;
        square unless `columns` is given."""
        idx = np.ascontiguousarray(index_stencil, np.uint32)
        v = _f32(values)
        if idx.ndim != 2 or v.shape != idx.shape:
            raise ValueError(cls._PREFIX + "index stencil and values must have the same shape")
        rows, k = idx.shape
        cols = rows if columns is None else int(columns)
        h = C.c_void_p()
        _check(load().naga_csr_from_stencil(rows, cols, k, _ptr(idx), _ptr(v), C.byref(h)),
               cls._PREFIX)
        return cls(h, rows, cols)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.naga_csr_destroy(h)
            self._h = None

    def mult(self, x, y=None, alpha=1.0, beta=0.0):
        """matrix_mult (matrix.cuh:438-629): alpha A x + beta y."""
        x = _f32(x)
        if x.shape != (self.columns,):
            raise ValueError(self._PREFIX + "x has incorrect size")
        out = np.zeros(self.rows, np.float32) if y is None else _f32(y).copy()
        if out.shape != (self.rows,):
            raise ValueError(self._PREFIX + "y has incorrect size")
        _check(load().naga_csr_mult(self._h, float(alpha), float(beta), _ptr(x), _ptr(out)),
               self._PREFIX)
        return out


class SceneBatch:
    """Many engines stepped together (one launch per pass over all scenes,
    replayed from a captured hipGraph).  The engines stay usable on their
    own for reading results; destroying an engine removes it."""

    def __init__(self, engines=(), graph: bool = True):
        lib = load()
        h = C.c_void_p()
        _check(lib.naga_batch_create(C.byref(h)), _BATCH_PREFIX)
        self._h = h
        self._engines = []
        self.set_graph(graph)
        for e in engines:
            self.add(e)

    def add(self, engine):
        _check(load().naga_batch_add(self._h, engine._live()), _BATCH_PREFIX)
        self._engines.append(engine)

    def remove(self, engine):
        _check(load().naga_batch_remove(self._h, engine._live()), _BATCH_PREFIX)
        self._engines.remove(engine)

    def set_graph(self, on: bool):
        _check(load().naga_batch_set_graph(self._h, int(bool(on))), _BATCH_PREFIX)

    def step_forward(self, nsteps: int = 1):
        _check(load().naga_batch_step(self._h, int(nsteps)), _BATCH_PREFIX)

    def step_timed(self, nsteps: int) -> float:
        ms = C.c_double()
        _check(load().naga_batch_step_timed(self._h, int(nsteps), C.byref(ms)), _BATCH_PREFIX)
        return ms.value

    def autotune(self, trial_steps: int = 5) -> int:
        gb = C.c_int32()
        _check(load().naga_batch_autotune(self._h, int(trial_steps), C.byref(gb)), _BATCH_PREFIX)
        return gb.value

    def set_q9_kernels(self, lanes: int, unions: bool | None = None):
        """D2Q9 kernel form of the stencil passes: 3 or 9 lanes per node;
        with 3, whether scenes with tile unions read them from LDS (default:
        yes for 3 lanes, no for 9)."""
        if unions is None:
            unions = int(lanes) == 3
        _check(load().naga_batch_set_q9_kernels(self._h, int(lanes), int(bool(unions))),
               _BATCH_PREFIX)

    def tuning(self) -> dict:
        """The current gather batch, D2Q9 kernel form and graph replay."""
        gb, lanes, un, graph = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _check(load().naga_batch_tuning(self._h, C.byref(gb), C.byref(lanes), C.byref(un),
                                        C.byref(graph)), _BATCH_PREFIX)
        return {"gather_batch": gb.value, "q9_lanes": lanes.value, "q9_unions": bool(un.value),
                "graph": bool(graph.value)}

    def __len__(self):
        return len(self._engines)

    def close(self):
        if self._h is not None:
            load().naga_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_SERVER_PREFIX = "naga::server: "


def make_request(lattice, h, steps, obstacles=(), k=0, jitter=0.25, seed=0x4E414741,
                 layer_thickness=None, peak_absorption=0.5, tau=0.6, pulse_center=None,
                 pulse_amplitude=1e-3, pulse_width=0.05, k_min=0,
                 refine_width=0.0, immersed_layer_thickness=None,
                 immersed_absorption=0.5) -> Request:
    """A solve of the unit box (square or cube) minus obstacles
    [(centre, radius), ...].  The absorbing layer at the box faces defaults
    to 8 spacings, at most 0.1 (wider layers would cover most of a coarse
    scene); in 2D the layer around the obstacles to 3 spacings with peak
    absorption 0.5 (without it, scattered nodes next to curved surfaces can
    carry a growing mode over hundreds of steps; around a refined sphere in
    3D the same layer itself grows a mode, so 3D requests default to none).
    refine_width > 0 doubles
    the node density within that distance of the obstacles (the time step
    is then 0.4 of the refined spacing h/2); k_min > 0 gives per-node
    stencil widths in [k_min, k]."""
    D = 2 if lattice == 9 else 3
    r = Request()
    r.lattice, r.k, r.steps = int(lattice), int(k), int(steps)
    r.n_obstacles = len(obstacles)
    if r.n_obstacles > MAX_OBSTACLES:
        raise ValueError(_SERVER_PREFIX + "too many obstacles")
    for j, (c, rad) in enumerate(obstacles):
        for d in range(D):
            r.obstacles[j * (D + 1) + d] = float(c[d])
        r.obstacles[j * (D + 1) + D] = float(rad)
    r.h, r.jitter, r.seed = float(h), float(jitter), int(seed)
    r.layer_thickness = min(8 * h, 0.1) if layer_thickness is None else float(layer_thickness)
    r.immersed_layer_thickness = ((3 * h if D == 2 else 0.0) if immersed_layer_thickness is None
                                  else float(immersed_layer_thickness))
    r.immersed_absorption = float(immersed_absorption)
    r.peak_absorption, r.tau = float(peak_absorption), float(tau)
    pc = [0.5] * D if pulse_center is None else list(pulse_center)
    for d in range(D):
        r.pulse_center[d] = float(pc[d])
    r.pulse_amplitude, r.pulse_width = float(pulse_amplitude), float(pulse_width)
    r.k_min, r.refine_width = int(k_min), float(refine_width)
    return r


def request_nodes(req: Request) -> "SimulationNodes":
    """The scene a request describes (what the server builds for it)."""
    D = 2 if req.lattice == 9 else 3
    obst = [([req.obstacles[j * (D + 1) + d] for d in range(D)], req.obstacles[j * (D + 1) + D])
            for j in range(req.n_obstacles)]
    return ObstacleGridProvider([0] * D, [1] * D, req.h, obst, req.layer_thickness,
                                req.peak_absorption, jitter=req.jitter, seed=req.seed,
                                refine_width=req.refine_width,
                                immersed_layer_thickness=req.immersed_layer_thickness,
                                immersed_absorption_rate=req.immersed_absorption).get()


def solve_alone(req: Request) -> dict:
    """The same solve on one engine with its own launches (reference for
    the served result)."""
    k = req.k or (16 if req.lattice == 9 else 32)
    eng = SimulationEngine(lattice=req.lattice, k=k, k_boundary=k, rbf=RBF_POLYHARMONIC3,
                           rbf_boundary=RBF_POLYHARMONIC3, pml_ref_index=True,
                           k_min=req.k_min)
    hmin = 0.5 * req.h if req.refine_width > 0 else req.h  # finest spacing
    p = problem_parameters(req.lattice, req.tau or 0.6, 0.4 * hmin)
    eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
    eng.init_domain(request_nodes(req))
    D = 2 if req.lattice == 9 else 3
    eng.set_initial_pulse([req.pulse_center[d] for d in range(D)], req.pulse_amplitude,
                          req.pulse_width)
    eng.step_forward(req.steps)
    out = eng.macroscopic(fresh=True)
    out["nodes"] = eng.num_nodes
    eng.close()
    return out


class SolveServer:
    """Serving loop with continuous batching (naga_server_*)."""

    def __init__(self, hbm_budget_bytes: int, max_active: int = 64, chunk_steps: int = 50,
                 graph: bool = True, setup_threads: int = 4):
        h = C.c_void_p()
        _check(load().naga_server_create(int(hbm_budget_bytes), int(max_active),
                                         int(chunk_steps), int(bool(graph)), int(setup_threads),
                                         C.byref(h)),
               _SERVER_PREFIX)
        self._h = h
        self._dims = {}

    def submit(self, req: Request) -> int:
        i = C.c_int64()
        _check(load().naga_server_submit(self._h, C.byref(req), C.byref(i)), _SERVER_PREFIX)
        self._dims[i.value] = 2 if req.lattice == 9 else 3
        return i.value

    def done(self, rid: int) -> bool:
        st = C.c_int32()
        _check(load().naga_server_status(self._h, int(rid), C.byref(st)), _SERVER_PREFIX)
        return st.value == 1

    def result(self, rid: int, release: bool = True) -> dict:
        lib = load()
        n = C.c_int64()
        lat = C.c_double()
        _check(lib.naga_server_result(self._h, int(rid), C.byref(n), None, None, None,
                                      C.byref(lat)), _SERVER_PREFIX)
        N, D = n.value, self._dims[rid]
        rho = np.zeros(N, np.float32)
        u = np.zeros(D * N, np.float32)
        drho = np.zeros(N)
        _check(lib.naga_server_result(self._h, int(rid), C.byref(n), _ptr(rho), _ptr(u),
                                      _ptr(drho), C.byref(lat)), _SERVER_PREFIX)
        if release:
            lib.naga_server_release(self._h, int(rid))
        return dict(rho=rho, u=u.reshape(D, N), drho=drho, latency_ms=lat.value, nodes=N)

    def drain(self):
        _check(load().naga_server_drain(self._h), _SERVER_PREFIX)

    def stats(self) -> dict:
        st = ServerStats()
        _check(load().naga_server_get_stats(self._h, C.byref(st)), _SERVER_PREFIX)
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}

    def close(self):
        if self._h is not None:
            load().naga_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DensityObserver:
    """density_observer<Lattice> (simulation_observers/density_observer.cuh:
    47-113): records the density at one location at every step, as the
    reference's update() does at the start of each step."""

    def __init__(self, engine: "SimulationEngine", location, time_multiplier: float = 1.0,
                 rbf: int = RBF_MULTIQUADRIC):
        self.engine = engine
        self.set = engine.add_probes(np.asarray(location, np.float64).reshape(1, -1), rbf)
        self.time_multiplier = float(time_multiplier)
        self.density_values, self.time_values = [], []

    def advance(self, nsteps: int = 1):
        """Step the engine, recording the observed value of each step."""
        for _ in range(int(nsteps)):
            t = self.engine.time()
            self.engine.step_forward(1)
            self.density_values.append(float(self.engine.sample_probes(self.set)[0]))
            self.time_values.append(t * self.time_multiplier)

    def update(self, time, domain=None, parameters=None, solution=None):
        """simulation_observer::update (simulation_observer.cuh:41-55), as an
        attached observer: the density of the state at the step's start."""
        self.density_values.append(float(self.engine.sample_probes(self.set, fresh=True)[0]))
        self.time_values.append(float(time) * self.time_multiplier)


class VtkObserver:
    """vtk_observer<Lattice> (simulation_observers/vtk_observer.cuh:45-431):
    as an attached observer, writes output_directory/result.<frame>.vtp with
    the state at each step's start.  frame = floor(time_multiplier * t *
    frame_rate) when frame_rate != 0 (a frame already written is skipped),
    else ceil(t / dt)."""

    def __init__(self, engine: "SimulationEngine", output_directory, frame_rate: float = 0.0,
                 time_multiplier: float = 1.0):
        self.engine = engine
        self.output_directory = str(output_directory)
        self.frame_rate = float(frame_rate)
        self.time_multiplier = float(time_multiplier)
        self.current_frame = 0
        self.files = []
        os.makedirs(self.output_directory, exist_ok=True)

    def update(self, time, domain=None, parameters=None, solution=None):
        if self.frame_rate != 0.0:
            frame = int(self.time_multiplier * time * self.frame_rate)
            if frame < self.current_frame:
                return
        else:
            dt = self.engine._params["dt"]
            frame = int(np.ceil(time / dt - 1e-9))
        self.current_frame = frame + 1
        path = os.path.join(self.output_directory, f"result.{frame}.vtp")
        self.engine.write_vtp(path, fresh=True)
        self.files.append(path)


def read_vtp(path) -> dict:
    """Point arrays of a .vtp file written by write_vtp (inline binary,
    UInt32 headers): name -> array [n] or [n, components]."""
    import base64
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    out = {}
    types = {"Float32": np.float32, "Int32": np.int32}
    for da in root.iter("DataArray"):
        raw = base64.b64decode("".join(da.text.split()))
        nbytes = int(np.frombuffer(raw[:4], np.uint32)[0])
        v = np.frombuffer(raw[4:4 + nbytes], types[da.get("type")])
        c = int(da.get("NumberOfComponents", "1"))
        out[da.get("Name")] = v.reshape(-1, c) if c > 1 else v
    return out
