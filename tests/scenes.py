"""Shared scene builders for the parity tests (nodes from the oracle's
provider feed both the GPU engine and the oracle engine, so both see the
same inputs bit for bit)."""
import numpy as np

import naga_amd

SEED = 0x4E414741


def grid_nodes(oracle, D, h, layer, sigma_max=0.5, jitter=0.25, lo=0.0, hi=1.0):
    return oracle.uniform_grid(D, [lo] * D, [hi] * D, h, jitter=jitter, seed=SEED,
                               layer=layer, sigma_max=sigma_max)


def to_sim_nodes(nodes, h):
    return naga_amd.SimulationNodes(nodes["points"], nodes["normals"], nodes["sigma"],
                                    nodes["n_bulk"], nodes["n_layer"], nodes["n_boundary"], h)


def gpu_engine(nodes, Q, k, h, tau=0.6, dt=None, rbf=1, kb=None, rbf_b=None,
               pml_ref_index=True):
    dt = 0.4 * h if dt is None else dt
    eng = naga_amd.SimulationEngine(lattice=Q, k=k, k_boundary=kb, rbf=rbf,
                                    rbf_boundary=rbf_b, pml_ref_index=pml_ref_index)
    p = naga_amd.problem_parameters(Q, tau, dt)
    eng.set_problem_parameters(p["nu"], 1.0, p["dt"], p["c"])
    eng.init_domain(to_sim_nodes(nodes, h))
    return eng


def oracle_engine(oracle, nodes, Q, k, h, tau=0.6, dt=None, rbf=1, kb=None, rbf_b=None,
                  pml_ref_index=True):
    return oracle.engine(nodes, Q=Q, k=k, k_boundary=kb, rbf=rbf, rbf_boundary=rbf_b,
                         tau=tau, dt=dt, h=h, pml_ref_index=int(pml_ref_index))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
