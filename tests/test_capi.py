"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/naga_amd.h declares, and its host-evaluated lattice tables and FP64
collision agree with the oracle (no GPU needed)."""
import os
import re

import numpy as np
import pytest

import naga_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "naga_amd.h")).read()
    return sorted(set(re.findall(r"\b(naga_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols(lib):
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(naga_amd.EXPORTED) == syms


@pytest.mark.parametrize("Q", [9, 27])
def test_lattice_tables_match_oracle(oracle64, Q):
    c, w, bb, K = naga_amd.lattice_tables(Q)
    co, wo, bbo, Ko = oracle64.lattice(Q)
    assert np.array_equal(c, co)
    assert np.allclose(w, wo, rtol=0, atol=1e-16)
    assert np.array_equal(bb, bbo)
    assert np.array_equal(K, Ko)


@pytest.mark.parametrize("Q", [9, 27])
def test_host_collision_matches_oracle(oracle64, Q):
    c, w, _, _ = naga_amd.lattice_tables(Q)
    rng = np.random.default_rng(Q)
    for amp in (1e-4, 1e-2, 5e-2):
        f = w * (1 + amp * rng.standard_normal(Q))
        rho = f.sum()
        u = (f[:, None] * c).sum(0) / rho
        for nu, dt in ((0.0333, 0.004), (0.05, 0.3)):
            a = naga_amd.collide(Q, f, rho, u, nu, dt)
            b = oracle64.collide(Q, f, rho, u, nu, dt)
            # increments are O(1e-5) differences of O(1) moments: compare at
            # FP64 resolution of the populations
            assert np.abs(a - b).max() <= 1e-13 + 1e-10 * np.abs(b).max()
            # a density source makes rho differ from sum f (u is kept)
            rs = rho * (1 + 3e-4)
            a = naga_amd.collide(Q, f, rs, u, nu, dt)
            b = oracle64.collide(Q, f, rs, u, nu, dt)
            assert np.abs(a - b).max() <= 1e-13 + 1e-10 * np.abs(b).max()


def test_engine_rejects_bad_config(lib):
    eng = naga_amd.SimulationEngine(lattice=27, k=70)
    eng.set_problem_parameters(0.01, 1.0, 0.004, naga_amd.CS_LAT)
    nodes = naga_amd.SimulationNodes(np.zeros((100, 3), np.float32), np.zeros((0, 3)),
                                     np.zeros(0), 100, 0, 0, 0.01)
    with pytest.raises(ValueError):
        eng.init_domain(nodes)


def test_engine_shape_errors():
    eng = naga_amd.SimulationEngine(lattice=9)
    eng.set_problem_parameters(0.01, 1.0, 0.004, naga_amd.CS_LAT)
    bad = naga_amd.SimulationNodes(np.zeros((10, 3), np.float32), np.zeros((0, 2)),
                                   np.zeros(0), 10, 0, 0, 0.01)
    with pytest.raises(ValueError, match="number of dimensions"):
        eng.init_domain(bad)
    bad2 = naga_amd.SimulationNodes(np.zeros((10, 2), np.float32), np.zeros((3, 2)),
                                    np.zeros(0), 8, 0, 2, 0.01)
    with pytest.raises(ValueError, match="boundary normals"):
        eng.init_domain(bad2)


def test_csr_and_advection_argument_errors():
    """Shape and index errors are reported before any device work."""
    import ctypes as C
    lib = naga_amd.load()
    idx = np.array([[0, 5]], np.uint32)
    vals = np.ones((1, 2), np.float32)
    h = C.c_void_p()
    st = lib.naga_csr_from_stencil(1, 5, 2, idx.ctypes.data_as(C.c_void_p),
                                   vals.ctypes.data_as(C.c_void_p), C.byref(h))
    assert st == 1 and b"column index" in lib.naga_last_error()
    assert not h.value
    W = np.ones((1, 2, 2), np.float32)
    st = lib.naga_advection_from_weights(idx.ctypes.data_as(C.c_void_p),
                                         W.ctypes.data_as(C.c_void_p), 1, 2, 2, C.byref(h))
    assert st == 1 and b"stencil index" in lib.naga_last_error()
    with pytest.raises(ValueError):
        naga_amd.CsrMatrix.create_from_index_stencil(idx, np.ones((2, 2)))
