"""GPU parity of the setup path (kNN, FP64 LU, RPIM, divergence weights,
node provider) against the CPU oracle, through the C-ABI."""
import json
import os

import numpy as np
import pytest

import naga_amd

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def _rng(seed=1234):
    return np.random.default_rng(seed)


@pytest.mark.parametrize("D,n,k", [(2, 3000, 16), (3, 4000, 32), (3, 2000, 1), (2, 500, 64)])
def test_knn_bit_exact_random(oracle64, D, n, k):
    rng = _rng(D * 100 + k)
    ref = rng.random((n, D)).astype(np.float32)
    q = (rng.random((n // 2, D)) * 1.2 - 0.1).astype(np.float32)
    d2g, ig = naga_amd.batched_nearest_neighbors(k, q, ref)
    d2o, io = oracle64.knn(ref, q, k)
    assert np.array_equal(ig, io)
    assert np.array_equal(d2g.view(np.uint32), d2o.view(np.uint32))


def test_knn_regular_grid_ties(oracle64):
    """examples/knn_validation_3d.cu:37-138: 20^3 grid of spacing 0.1, k=64,
    queries in [0, 1.9]^3; on a lattice many distances tie, broken by index."""
    g = np.arange(20) * 0.1
    ref = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    rng = _rng(7)
    q = (rng.random((30, 3)) * 1.9).astype(np.float32)
    q = np.concatenate([q, ref[::97]])  # queries exactly on grid nodes
    d2g, ig = naga_amd.batched_nearest_neighbors(64, q, ref)
    d2o, io = oracle64.knn(ref, q, 64)
    bo, bi = oracle64.knn(ref, q, 64, brute=True)
    assert np.array_equal(io, bi) and np.array_equal(d2o, bo)
    assert np.array_equal(ig, io)
    assert np.array_equal(d2g, d2o)


def test_knn_planted_cluster():
    """examples/batched_knn_search.cu:66-157: a 4^3 cluster of spacing 0.25
    planted at the origin of a coarse grid; its 64 points are the 64
    nearest neighbours of the cluster point at the origin."""
    g = np.linspace(2.0, 10.0, 30)
    coarse = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    c = (np.arange(4) - 1.5) * 0.25
    cluster = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3)
    ref = np.concatenate([cluster, coarse]).astype(np.float32)
    _, idx = naga_amd.batched_nearest_neighbors(64, ref[:1], ref)
    assert set(idx[0].tolist()) == set(range(64))


def test_batched_lu_known_answer():
    """examples/batched_matrix_inversion.cu:45-69: the inverse of
    [[1,2,3],[4,5,6],[7,8,1]] to 1e-6 (golden values from the reference)."""
    kat = json.load(open(os.path.join(GOLDEN, "inverse3x3.json")))
    A = np.array(kat["A"])
    Ainv = np.array(kat["A_inv"])
    batch = 1000
    X, info = naga_amd.batched_lu_solve(np.repeat(A[None], batch, 0),
                                        np.repeat(np.eye(3)[None], batch, 0))
    assert not info.any()
    assert np.abs(X - Ainv).max() < 1e-6


@pytest.mark.parametrize("n,nrhs", [(19, 1), (36, 3), (52, 5), (64, 2)])
def test_batched_lu_vs_numpy(n, nrhs):
    rng = _rng(n)
    A = rng.standard_normal((50, n, n)) + n * np.eye(n)
    B = rng.standard_normal((50, n, nrhs))
    X, info = naga_amd.batched_lu_solve(A, B)
    assert not info.any()
    ref = np.linalg.solve(A, B)
    assert np.abs(X - ref).max() / np.abs(ref).max() < 1e-12


def test_batched_lu_pivoting_and_singular():
    # zero leading entry forces a row exchange; a zero matrix is flagged
    A = np.array([[[0.0, 1.0], [1.0, 0.0]], [[0.0, 0.0], [0.0, 0.0]]])
    B = np.array([[2.0, 3.0], [1.0, 1.0]])
    X, info = naga_amd.batched_lu_solve(A, B)
    assert info[0] == 0 and np.allclose(X[0], [3.0, 2.0])
    assert info[1] != 0


@pytest.mark.parametrize("D,rbf", [(2, 1), (3, 1), (2, 0), (3, 0)])
def test_rpim_weights_vs_oracle(oracle64, D, rbf):
    rng = _rng(10 + D + rbf)
    src = (rng.random((800, D)) * 0.3).astype(np.float32)
    q = (rng.random((200, D)) * 0.2 + 0.05).astype(np.float32)
    k = 16 if D == 2 else 32
    _, idx = oracle64.knn(src, q, k)
    wg = naga_amd.rpim_weights(src, idx, q, rbf=rbf, spacing=0.01)
    wo = oracle64.rpim_weights(src, idx, q, rbf, 0.01)
    assert np.abs(wg - wo).max() < 1e-6 * max(1.0, np.abs(wo).max())
    # linear reproduction: sum w = 1, sum w x = x_q
    assert np.abs(wg.sum(1) - 1).max() < 1e-8
    xs = src.astype(np.float64)[idx]
    assert np.abs((wg[:, :, None] * xs).sum(1) - q).max() < 1e-8


@pytest.mark.parametrize("D,k,rbf", [(2, 16, 1), (3, 32, 1), (2, 32, 0)])
def test_divergence_weights_vs_oracle(oracle64, D, k, rbf):
    o = oracle64
    h = 0.05 if D == 2 else 0.1
    nodes = o.uniform_grid(D, [0] * D, [1] * D, h, jitter=0.25)
    pts = nodes["points"]
    op = naga_amd.DivergenceOperator.create(pts, k=k, rbf=rbf)
    d2, idx = o.knn(pts, pts, k)
    assert np.array_equal(op.support_indices(), idx)
    W, _ = o.divergence_weights(pts, pts, idx, d2, rbf)
    scale = np.abs(W).max()
    assert np.abs(op.weights() - W).max() < 1e-7 * scale
    # consistency: a constant field has zero divergence; a linear field
    # F = (x, 0, ..) gives the same quadrature constant at every interior
    # node (RPIM reproduces linear fields exactly), as in the oracle
    F = np.zeros_like(pts)
    F[:, 0] = pts[:, 0]
    out = op.apply(F)
    ref = o.divergence_apply(idx, W, F.astype(np.float64))
    interior = np.all((pts > 0.3) & (pts < 0.7), axis=1)
    assert np.abs(out[interior] - ref[interior]).max() < 1e-4
    assert np.ptp(out[interior]) < 1e-4
    const = op.apply(np.ones_like(pts))
    assert np.abs(const[interior]).max() < 1e-3


def test_divergence_apply_vs_oracle(oracle64):
    rng = _rng(3)
    n, k, D = 500, 16, 2
    idx = rng.integers(0, n, (n, k)).astype(np.uint32)
    W = rng.standard_normal((n, k, D)).astype(np.float32)
    F = rng.standard_normal((n, D)).astype(np.float32)
    out = naga_amd.DivergenceOperator(W, idx).apply(F, 0.25)
    ref = oracle64.divergence_apply(idx, W.astype(np.float64), F.astype(np.float64), 0.25)
    assert np.abs(out - ref).max() < 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("D,h,layer", [(2, 0.01, 0.1), (3, 0.05, 0.16)])
def test_uniform_grid_provider_vs_oracle(oracle64, D, h, layer):
    prov = naga_amd.UniformGridProvider([0] * D, [1] * D, h, layer, 0.5, jitter=0.25)
    got = prov.get()
    ref = oracle64.uniform_grid(D, [0] * D, [1] * D, h, jitter=0.25, layer=layer,
                                sigma_max=0.5)
    assert (got.num_bulk_points, got.num_layer_points, got.num_boundary_points) == (
        ref["n_bulk"], ref["n_layer"], ref["n_boundary"])
    assert np.array_equal(got.points, ref["points"])
    assert np.array_equal(got.boundary_normals, ref["normals"])
    assert np.array_equal(got.layer_absorption, ref["sigma"])
