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
               pml_ref_index=True, filter_strength=0.0):
    dt = 0.4 * h if dt is None else dt
    eng = naga_amd.SimulationEngine(lattice=Q, k=k, k_boundary=kb, rbf=rbf,
                                    rbf_boundary=rbf_b, pml_ref_index=pml_ref_index,
                                    filter_strength=filter_strength)
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


def mirror_nodes(D, h=0.05, m=8, n=17, layer=0.15, sigma_max=0.5, seed=3):
    """A node set that is exactly symmetric under x -> -x.

    A jittered half grid (x = (i + 1/2) h, i < m; other axes j h, j < n) is
    reflected by negating x, which is exact in floating point, so every
    FP32 distance, stencil and weight of a node equals that of its mirror
    image with x flipped.  Boundary nodes close the box (i = m-1 or any other
    axis at 0 or n-1); layer and absorption follow the provider's rules.
    Returns the node dict and, per node, the index of its mirror image."""
    rng = np.random.default_rng(seed)
    grids = np.stack(np.meshgrid(*([np.arange(m)] + [np.arange(n)] * (D - 1)),
                                 indexing="ij"), -1).reshape(-1, D)
    bnd = (grids[:, 0] == m - 1) | np.any((grids[:, 1:] == 0) | (grids[:, 1:] == n - 1), axis=1)
    x = grids.astype(np.float64) * h
    x[:, 0] += 0.5 * h
    x[~bnd] += (2 * rng.random((int((~bnd).sum()), D)) - 1) * 0.25 * h
    nrm = np.zeros((len(x), D))
    nrm[:, 0] = np.where(grids[:, 0] == m - 1, -1.0, 0.0)
    nrm[:, 1:] = np.where(grids[:, 1:] == 0, 1.0, np.where(grids[:, 1:] == n - 1, -1.0, 0.0))
    half = x.astype(np.float32)
    flip = np.ones(D, np.float32)
    flip[0] = -1
    P = np.concatenate([half, half * flip])
    bnd = np.concatenate([bnd, bnd])
    nrm = np.concatenate([nrm, nrm * flip])
    from scipy.spatial import cKDTree
    # distance to the nearest boundary node (negating x leaves it unchanged)
    dist = cKDTree(P[bnd].astype(np.float64)).query(P.astype(np.float64))[0]
    inl = ~bnd & (dist <= layer)
    blk = ~bnd & ~inl
    order = np.concatenate([np.where(blk)[0], np.where(inl)[0], np.where(bnd)[0]])
    dl = dist[order][blk.sum():blk.sum() + inl.sum()].astype(np.float32)
    r = (np.float32(layer) - dl) / np.float32(layer)
    nb = nrm[order][-int(bnd.sum()):]
    nodes = dict(points=np.ascontiguousarray(P[order]), n_bulk=int(blk.sum()),
                 n_layer=int(inl.sum()), n_boundary=int(bnd.sum()),
                 normals=(nb / np.linalg.norm(nb, axis=1, keepdims=True)).astype(np.float32),
                 sigma=(sigma_max * r * r).astype(np.float32))
    pos = {tuple(p): i for i, p in enumerate(nodes["points"].tolist())}
    mirror = np.array([pos[(-p[0],) + tuple(p[1:])] for p in nodes["points"].tolist()])
    return nodes, mirror


def kvar_like_engine(oracle, points, k, k_min):
    """Per-node stencil widths by local density, as the engine forms them
    (k_min > 0): d_i the mean distance to the 4 nearest other nodes, d_ref
    its 90th percentile (element floor(0.9 (N - 1)) in ascending order),
    k_i = round(k_min d_ref / d_i) clamped to [k_min, k] (halves away from
    zero); computed here from the oracle's kNN."""
    pts = np.ascontiguousarray(points, np.float32)
    d2, _ = oracle.knn(pts, pts, k)
    r = np.sqrt(d2[:, 1:5].astype(np.float64))
    dm = (((r[:, 0] + r[:, 1]) + r[:, 2]) + r[:, 3]) / 4.0
    N = dm.size
    dref = np.partition(dm, int(0.9 * (N - 1)))[int(0.9 * (N - 1))]
    want = float(k_min) * dref / np.maximum(dm, 1e-30)
    return np.clip(np.floor(want + 0.5), k_min, k).astype(np.uint8)
