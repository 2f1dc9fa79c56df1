"""ctypes wrapper of the CPU oracle (test infrastructure only).

The oracle is a C restatement of NAGA's nonlocal LBM step and setup, built
from ``naga_oracle.c`` once in FP64 (the checker) and once in FP32 (the CPU
baseline).  Only the tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


class Config(C.Structure):
    _fields_ = [
        ("Q", C.c_int), ("k", C.c_int), ("k_boundary", C.c_int),
        ("rbf", C.c_int), ("rbf_boundary", C.c_int),
        ("pml_ref_index", C.c_int), ("threads", C.c_int),
        ("nu", C.c_double), ("rho0", C.c_double), ("dt", C.c_double),
        ("c", C.c_double), ("h", C.c_double), ("filter", C.c_double),
        ("kvar", C.c_void_p), ("solve_bits", C.c_int),
    ]


def build(force: bool = False) -> None:
    """Compile both precisions with the Makefile beside this file."""
    libs = [os.path.join(HERE, f"libnaga_oracle_f{b}.so") for b in (32, 64)]
    if force or not all(os.path.exists(p) for p in libs):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []),
                       check=True)


class Oracle:
    """One precision of the oracle (``bits`` = 64 or 32)."""

    def __init__(self, bits: int = 64):
        build()
        self.bits = bits
        self.real = np.float64 if bits == 64 else np.float32
        self.lib = C.CDLL(os.path.join(HERE, f"libnaga_oracle_f{bits}.so"))
        p = f"o{bits}_"
        L = self.lib
        rp = np.ctypeslib.ndpointer(self.real, flags="C_CONTIGUOUS")
        self._rp = rp

        def fn(name, res, args):
            f = getattr(L, p + name)
            f.restype = res
            f.argtypes = args
            return f

        self._lattice = fn("lattice_tables", C.c_int,
                           [C.c_int, _f64p, _f64p, np.ctypeslib.ndpointer(np.int32), _f64p])
        self._collide = fn("collide", C.c_int,
                           [C.c_int, rp, self._ct(), rp, self._ct(), self._ct(), rp])
        self._feq = fn("feq", self._ct(), [C.c_int, self._ct(), rp, C.c_int])
        self._rk4 = fn("rk4_scalar", C.c_double,
                       [C.c_double, C.c_double, C.c_double, C.c_int])
        self._grid_size = fn("grid_size", C.c_int64,
                             [C.c_int, _f64p, _f64p, C.c_double, _i64p])
        self._grid = fn("uniform_grid", C.c_int,
                        [C.c_int, _f64p, _f64p, C.c_double, C.c_double,
                         C.c_uint64, C.c_double, C.c_double, _f32p, _f32p,
                         _f32p, _i64p, _i64p])
        knn_args = [C.c_int, _f32p, C.c_int64, _f32p, C.c_int64, C.c_int,
                    _f32p, _u32p]
        self._knn = fn("knn", C.c_int, knn_args)
        self._knn_brute = fn("knn_brute", C.c_int, knn_args)
        self._lu = fn("lu_solve", C.c_int, [C.c_int, C.c_int, _f64p, _f64p])
        self._rpim = fn("rpim_weights_ex", C.c_int,
                        [C.c_int, _f32p, _u32p, C.c_int, _f32p, C.c_int64,
                         C.c_int, C.c_double, C.c_int, _f64p])
        self._divw = fn("divergence_weights_ex", C.c_int,
                        [C.c_int, _f32p, C.c_int64, _f32p, C.c_int64, _u32p,
                         _f32p, C.c_int, C.c_int, C.c_void_p, C.c_int, _f64p, _f64p])
        self._diva = fn("divergence_apply", C.c_int,
                        [C.c_int, C.c_int64, C.c_int, _u32p, _f64p, _f64p,
                         C.c_double, _f64p])
        self._create = fn("engine_create", C.c_void_p,
                          [C.POINTER(Config), _f32p, C.c_int64, C.c_int64,
                           C.c_int64, _f32p, _f32p])
        self._destroy = fn("engine_destroy", None, [C.c_void_p])
        self._pulse = fn("engine_set_pulse", C.c_int,
                         [C.c_void_p, _f64p, C.c_double, C.c_double])
        self._setf = fn("engine_set_distributions", C.c_int, [C.c_void_p, _f64p])
        self._step = fn("engine_step", C.c_int, [C.c_void_p, C.c_int])
        self._add_src = fn("engine_add_density_source", C.c_int,
                           [C.c_void_p, _f64p] + [C.c_double] * 6)
        self._clear_src = fn("engine_clear_density_sources", C.c_int, [C.c_void_p])
        opt64 = C.c_void_p
        self._get = fn("engine_get", C.c_int,
                       [C.c_void_p, opt64, opt64, opt64, opt64, C.c_int])
        self._get_setup = fn("engine_get_setup", C.c_int,
                             [C.c_void_p, opt64, opt64, opt64, opt64, opt64])
        self._get_pml = fn("engine_get_pml", C.c_int, [C.c_void_p, opt64, opt64])
        self._set_pml = fn("engine_set_pml", C.c_int, [C.c_void_p, opt64, opt64])
        self._pml_stage = fn("engine_pml_stage", C.c_int, [C.c_void_p, _f64p, _f64p])
        self._time = fn("engine_time", C.c_double, [C.c_void_p])

    def _ct(self):
        return C.c_double if self.bits == 64 else C.c_float

    # --- lattice and collision ------------------------------------------
    def lattice(self, Q: int):
        D = 2 if Q == 9 else 3
        c = np.zeros(Q * D)
        w = np.zeros(Q)
        bb = np.zeros(Q, np.int32)
        K = np.zeros(Q * Q)
        if self._lattice(Q, c, w, bb, K):
            raise ValueError("bad lattice")
        return c.reshape(Q, D), w, bb, K.reshape(Q, Q)

    def collide(self, Q, f, rho, u, nu_lat, dt_lat):
        f = np.ascontiguousarray(f, self.real)
        u = np.ascontiguousarray(u, self.real)
        df = np.zeros(Q, self.real)
        self._collide(Q, f, rho, u, nu_lat, dt_lat, df)
        return df

    def feq(self, Q, rho, u, alpha):
        return self._feq(Q, rho, np.ascontiguousarray(u, self.real), alpha)

    def rk4_scalar(self, y0, lam, dt, n):
        return self._rk4(y0, lam, dt, n)

    # --- nodes ----------------------------------------------------------
    def uniform_grid(self, D, lo, hi, h, jitter=0.25, seed=0x4E414741,
                     layer=0.0, sigma_max=0.0):
        lo = np.asarray(lo, np.float64)
        hi = np.asarray(hi, np.float64)
        dims = np.zeros(3, np.int64)
        N = int(self._grid_size(D, lo, hi, h, dims))
        pts = np.zeros(N * D, np.float32)
        nrm = np.zeros(N * D, np.float32)
        sig = np.zeros(N, np.float32)
        counts = np.zeros(3, np.int64)
        gidx = np.zeros(N, np.int64)
        self._grid(D, lo, hi, h, jitter, seed, layer, sigma_max, pts, nrm,
                   sig, counts, gidx)
        nb, nl, nbd = (int(x) for x in counts)
        return dict(points=pts.reshape(N, D), normals=nrm[: nbd * D].reshape(nbd, D),
                    sigma=sig[:nl].copy(), n_bulk=nb, n_layer=nl,
                    n_boundary=nbd, grid_index=gidx, dims=dims[:D].copy())

    # --- kNN, LU, RPIM, divergence ----------------------------------------
    def knn(self, ref, query, k, brute=False):
        ref = np.ascontiguousarray(ref, np.float32)
        query = np.ascontiguousarray(query, np.float32)
        D = ref.shape[1]
        nq = query.shape[0]
        d2 = np.zeros(nq * k, np.float32)
        idx = np.zeros(nq * k, np.uint32)
        f = self._knn_brute if brute else self._knn
        if f(D, ref, ref.shape[0], query, nq, k, d2, idx):
            raise ValueError("knn failed")
        return d2.reshape(nq, k), idx.reshape(nq, k)

    def lu_solve(self, A, B):
        A = np.asfortranarray(A, np.float64)
        n = A.shape[0]
        B2 = np.asarray(B, np.float64)
        nrhs = 1 if B2.ndim == 1 else B2.shape[1]
        Ac = np.ascontiguousarray(A.T.reshape(-1))  # column-major storage
        Bc = np.ascontiguousarray(B2.reshape(n, nrhs).T.reshape(-1))
        info = self._lu(n, nrhs, Ac, Bc)
        X = Bc.reshape(nrhs, n).T
        return (X[:, 0] if B2.ndim == 1 else X), info

    def rpim_weights(self, src, idx, query, rbf, spacing, solve_bits=64):
        src = np.ascontiguousarray(src, np.float32)
        query = np.ascontiguousarray(query, np.float32)
        idx = np.ascontiguousarray(idx, np.uint32)
        nq, k = idx.shape
        w = np.zeros(nq * k)
        if self._rpim(src.shape[1], src, idx, k, query, nq, rbf, spacing, solve_bits, w):
            raise ValueError("singular RPIM system")
        return w.reshape(nq, k)

    def divergence_weights(self, src, query, idx, d2, rbf, kvar=None, solve_bits=64):
        src = np.ascontiguousarray(src, np.float32)
        query = np.ascontiguousarray(query, np.float32)
        idx = np.ascontiguousarray(idx, np.uint32)
        d2 = np.ascontiguousarray(d2, np.float32)
        nq, k = idx.shape
        D = src.shape[1]
        W = np.zeros(nq * k * D)
        delta = np.zeros(nq)
        kv = None if kvar is None else np.ascontiguousarray(kvar, np.uint8)
        if self._divw(D, src, src.shape[0], query, nq, idx, d2, k, rbf,
                      None if kv is None else kv.ctypes.data, solve_bits, W, delta):
            raise ValueError("singular divergence system")
        return W.reshape(nq, k, D), delta

    def divergence_apply(self, idx, W, F, off=0.0):
        idx = np.ascontiguousarray(idx, np.uint32)
        W = np.ascontiguousarray(W, np.float64)
        F = np.ascontiguousarray(F, np.float64)
        n, k, D = W.shape
        out = np.zeros(n)
        self._diva(D, n, k, idx, W, F, off, out)
        return out

    def advection_rate(self, idx, W, v, f, off):
        """rate = div(F), F_j = v_j (-(f_j - off)): the divergence field map of
        nonlocal_calculus/advection.cuh:102-141 applied with
        divergence_operator::apply."""
        v = np.asarray(v, np.float64)
        n = idx.shape[0]
        if v.ndim == 1:
            v = np.broadcast_to(v, (n, v.shape[0]))
        F = v * (-(np.asarray(f, np.float64) - off))[:, None]
        return self.divergence_apply(idx, W, F)

    def advection_step(self, idx, W, v, f0, dt, off=0.0, boundary_start=0, steps=1):
        """advection_task::impl (nonlocal_calculus/advection.cuh:207-276) in
        FP64: k1 = rate(f0); f = f0 + dt a_i k_i on [0, end) with
        a = (1/2, 1/2, 1) (:231-237, :568), k_{i+1} = rate(f) (:246);
        f = f0 + dt sum b_i k_i on [0, end), b = (1, 2, 2, 1)/6 (:245-276,
        :569-570); end = boundary_start when > 0, else every node.  Nodes at
        and past end keep f0 (assign_array(f0, f), :223)."""
        a = (0.5, 0.5, 1.0)
        b = (1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0)
        f0 = np.asarray(f0, np.float64).copy()
        n = f0.shape[0]
        end = boundary_start if boundary_start > 0 else n
        for _ in range(steps):
            ks = [self.advection_rate(idx, W, v, f0, off)]
            f = f0.copy()
            for i in range(3):
                f[:end] = f0[:end] + dt * a[i] * ks[i][:end]
                ks.append(self.advection_rate(idx, W, v, f, off))
            total = sum(bi * ki for bi, ki in zip(b, ks))
            f[:end] = f0[:end] + dt * total[:end]
            f0 = f
        return f0

    @staticmethod
    def csr_from_stencil(idx, values):
        """create_from_index_stencil (linalg/matrix.cuh:300-368): row offsets
        k r, row r's entries in stencil order (the per-row sort writes entry i
        back to position i, :337-351)."""
        n, k = idx.shape
        rp = np.arange(n + 1, dtype=np.int64) * k
        return rp, np.asarray(idx, np.int64).reshape(-1), np.asarray(values, np.float64).reshape(-1)

    @staticmethod
    def csr_mult(rp, ci, vals, x, alpha=1.0, beta=0.0, y=None):
        """y = alpha A x + beta y (matrix_mult, linalg/matrix.cuh:438-629)."""
        rows = rp.shape[0] - 1
        out = np.zeros(rows)
        for r in range(rows):
            e0, e1 = rp[r], rp[r + 1]
            out[r] = np.dot(vals[e0:e1], np.asarray(x, np.float64)[ci[e0:e1]])
        out *= alpha
        if beta != 0.0:
            out += beta * np.asarray(y, np.float64)
        return out

    # --- engine ---------------------------------------------------------
    def engine(self, nodes, **kw):
        return OracleEngine(self, nodes, **kw)


def physics_params(Q: int, tau: float, dt: float, c: float | None = None):
    """nu for the requested relaxation time tau.

    D3Q27 relaxes shear at 1/(3 nu_lat + 1/2)/dt_lat
    (detail/d3q27_lattice.cuh:177); D2Q9 at 1/(3 nu_lat + dt_lat/2)
    (detail/d2q9_lattice.cuh:136).  With c = c_s the lattice and physical
    units coincide.
    """
    cs = 0.57735026918962576450914878050196
    c = cs if c is None else c
    ratio = cs / c
    dt_lat = dt / ratio
    nu_lat = (tau - 0.5) / 3.0 if Q == 27 else (tau - 0.5) * dt_lat / 3.0
    return dict(nu=nu_lat * ratio, c=c, dt=dt)


class OracleEngine:
    def __init__(self, oracle: Oracle, nodes: dict, Q: int, k: int,
                 k_boundary: int | None = None, rbf: int = 1,
                 rbf_boundary: int | None = None, tau: float = 0.6,
                 dt: float | None = None, h: float = 0.01, rho0: float = 1.0,
                 pml_ref_index: int = 1, threads: int = 0, c=None,
                 filter_strength: float = 0.75, kvar=None, solve_bits: int = 64):
        self.o = oracle
        self.Q = Q
        self.D = 2 if Q == 9 else 3
        self.N = nodes["n_bulk"] + nodes["n_layer"] + nodes["n_boundary"]
        self.nodes = nodes
        self.k = k
        self.kb = k if k_boundary is None else k_boundary
        dt = 0.4 * h if dt is None else dt
        ph = physics_params(Q, tau, dt, c)
        cfg = Config(Q=Q, k=k, k_boundary=self.kb, rbf=rbf,
                     rbf_boundary=rbf if rbf_boundary is None else rbf_boundary,
                     pml_ref_index=pml_ref_index, threads=threads,
                     nu=ph["nu"], rho0=rho0, dt=ph["dt"], c=ph["c"], h=h,
                     filter=filter_strength, solve_bits=solve_bits)
        # per-node stencil widths; kept alive for the engine's setup
        self._kvar = None if kvar is None else np.ascontiguousarray(kvar, np.uint8)
        cfg.kvar = None if self._kvar is None else self._kvar.ctypes.data
        self.cfg = cfg
        pts = np.ascontiguousarray(nodes["points"], np.float32).reshape(-1)
        nrm = np.ascontiguousarray(nodes["normals"], np.float32).reshape(-1)
        if nrm.size == 0:
            nrm = np.zeros(1, np.float32)
        sig = np.ascontiguousarray(nodes["sigma"], np.float32)
        if sig.size == 0:
            sig = np.zeros(1, np.float32)
        self.h = oracle._create(C.byref(cfg), pts, nodes["n_bulk"],
                                nodes["n_layer"], nodes["n_boundary"], nrm, sig)
        if not self.h:
            raise ValueError("oracle engine_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.o._destroy(self.h)
            self.h = None

    def set_pulse(self, center, amp, width):
        self.o._pulse(self.h, np.asarray(center, np.float64), amp, width)

    def set_distributions(self, f):
        self.o._setf(self.h, np.ascontiguousarray(f, np.float64).reshape(-1))

    def add_density_source(self, center, radius, amplitude, frequency=0.0, phase=0.0,
                           t_start=0.0, t_end=float("inf")):
        c = np.zeros(3)
        c[: self.D] = np.asarray(center, np.float64)[: self.D]
        if self.o._add_src(self.h, c, radius, amplitude, frequency, phase, t_start, t_end):
            raise ValueError("density source rejected")

    def clear_density_sources(self):
        self.o._clear_src(self.h)

    def step(self, n=1):
        self.o._step(self.h, n)

    def get(self, fresh=True):
        N, Q, D = self.N, self.Q, self.D
        f = np.zeros(Q * N)
        rho = np.zeros(N)
        u = np.zeros(D * N)
        drho = np.zeros(N)
        self.o._get(self.h, f.ctypes.data, rho.ctypes.data, u.ctypes.data,
                    drho.ctypes.data, int(fresh))
        return dict(f=f.reshape(Q, N), rho=rho, u=u.reshape(D, N), drho=drho)

    def setup(self):
        N, k, D = self.N, self.k, self.D
        nb = self.nodes["n_boundary"]
        idx = np.zeros(N * k, np.uint32)
        W = np.zeros(N * k * D)
        filt = np.zeros(N * k)
        bidx = np.zeros(max(nb, 1) * self.kb, np.uint32)
        bw = np.zeros(max(nb, 1) * self.kb)
        self.o._get_setup(self.h, idx.ctypes.data, W.ctypes.data,
                          filt.ctypes.data, bidx.ctypes.data, bw.ctypes.data)
        return dict(idx=idx.reshape(N, k), W=W.reshape(N, k, D),
                    filt=filt.reshape(N, k),
                    bidx=bidx[: nb * self.kb].reshape(nb, self.kb),
                    bw=bw[: nb * self.kb].reshape(nb, self.kb))

    def pml(self):
        N, Q = self.N, self.Q
        q1 = np.zeros(Q * N)
        q2 = np.zeros(Q * N)
        self.o._get_pml(self.h, q1.ctypes.data, q2.ctypes.data)
        return q1.reshape(Q, N), q2.reshape(Q, N)

    def set_pml(self, q1=None, q2=None):
        """Overwrite the stored PML history (Q1: previous equilibrium, Q2:
        previous trapezoid value), each Q x N."""
        a = None if q1 is None else np.ascontiguousarray(q1, np.float64).reshape(-1)
        b = None if q2 is None else np.ascontiguousarray(q2, np.float64).reshape(-1)
        self.o._set_pml(self.h, None if a is None else a.ctypes.data,
                        None if b is None else b.ctypes.data)

    def pml_stage(self, s1):
        """The PML terms of one step applied to the stage input s1 (Q x N),
        with the macroscopic values of the current distributions."""
        s1 = np.ascontiguousarray(s1, np.float64).reshape(-1)
        out = np.zeros_like(s1)
        self.o._pml_stage(self.h, s1, out)
        return out.reshape(self.Q, self.N)

    @property
    def time(self):
        return self.o._time(self.h)
