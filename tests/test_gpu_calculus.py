"""Standalone advection operator and CSR matrices on the GPU against the
oracle (nonlocal_calculus/advection.cuh:143-524, linalg/matrix.cuh:300-629)."""
import numpy as np
import pytest

import naga_amd

pytestmark = pytest.mark.gpu


def _domain(o, D, h):
    nd = o.uniform_grid(D, [0] * D, [1] * D, h, jitter=0.25)
    return nd["points"]


@pytest.mark.parametrize("D,h,k", [(2, 0.02, 16), (3, 0.1, 32)])
def test_advection_matches_oracle(oracle64, D, h, k):
    o = oracle64
    pts = _domain(o, D, h)
    n = len(pts)
    op = naga_amd.AdvectionOperator.create(pts, k=k)
    div = op.divergence_op()
    d2, idx = o.knn(pts, pts, k)
    assert np.array_equal(div.support_indices(), idx)
    Wo, _ = o.divergence_weights(pts, pts, idx, d2, 1)
    Wg = div.weights().astype(np.float64)
    assert np.linalg.norm(Wg - Wo) / np.linalg.norm(Wo) < 1e-5
    rng = np.random.default_rng(3)
    x = pts.astype(np.float64)
    f0 = np.exp(-np.sum((x - 0.4) ** 2, 1) / 0.1 ** 2) + 0.3
    v = np.array([0.4, -0.3, 0.2][:D])
    dt = 0.25 * h
    # constant velocity, all nodes, centering offset
    got = op.step_forward(naga_amd.ConstantVelocityField(v), f0, dt, centering_offset=0.3,
                          steps=10)
    want = o.advection_step(idx, Wg, v, f0, dt, 0.3, steps=10)
    assert np.linalg.norm(got - want) / np.linalg.norm(want - 0.3) < 1e-5
    # per-node velocity and a boundary slice
    vn = rng.uniform(-0.5, 0.5, (n, D))
    end = n - n // 5
    got = op.step_forward(vn, f0, dt, boundary_idx_start=end, steps=5)
    want = o.advection_step(idx, Wg, vn, f0, dt, 0.0, boundary_start=end, steps=5)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5
    assert np.array_equal(got[end:], f0[end:].astype(np.float32))


def test_advection_from_divergence_operator(oracle64):
    """create(divergence_op) over the same stencils and FP32 weights runs the
    same kernels: the results are identical."""
    pts = _domain(oracle64, 2, 0.03)
    a = naga_amd.AdvectionOperator.create(pts, k=16)
    b = naga_amd.AdvectionOperator.from_divergence(a.divergence_op())
    f0 = np.cos(3 * pts[:, 0]) * np.sin(2 * pts[:, 1])
    vel = naga_amd.ConstantVelocityField([0.2, 0.1])
    fa = a.step_forward(vel, f0, 0.005, steps=4)
    fb = b.step_forward(vel, f0, 0.005, steps=4)
    assert np.array_equal(fa, fb)
    with pytest.raises(ValueError):
        a.step_forward(vel, f0[:-1], 0.005)


def test_advection_translates_a_pulse(oracle64):
    pts = _domain(oracle64, 2, 0.02)
    x = pts.astype(np.float64)
    op = naga_amd.AdvectionOperator.create(pts, k=16)
    v = np.array([0.5, 0.25])
    c0 = np.array([0.35, 0.4])
    f0 = np.exp(-np.sum((x - c0) ** 2, 1) / 0.08 ** 2)
    f = op.step_forward(naga_amd.ConstantVelocityField(v), f0, 0.01, steps=40)
    c1 = c0 + 0.4 * v
    inside = np.all((x > 0.15) & (x < 0.85), axis=1)
    com = (f[inside, None] * x[inside]).sum(0) / f[inside].sum()
    assert np.abs(com - c1).max() < 0.01


def test_csr_from_stencil_matches_oracle(oracle64):
    rng = np.random.default_rng(11)
    n, k, cols = 5000, 27, 4000
    idx = rng.integers(0, cols, (n, k)).astype(np.uint32)
    vals = rng.standard_normal((n, k)).astype(np.float32)
    A = naga_amd.CsrMatrix.create_from_index_stencil(idx, vals, columns=cols)
    rp, ci, cv = oracle64.csr_from_stencil(idx, vals)
    x = rng.standard_normal(cols).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    for alpha, beta in [(1.0, 0.0), (0.5, -2.0), (2.0, 1.0)]:
        got = A.mult(x, y, alpha=alpha, beta=beta)
        want = oracle64.csr_mult(rp, ci, cv, x, alpha, beta, y)
        scale = np.abs(vals).sum(1) * np.abs(x).max() * abs(alpha) + abs(beta) * np.abs(y)
        assert np.all(np.abs(got - want) <= 1e-6 * scale)


def test_csr_general_rows_and_errors(oracle64):
    rng = np.random.default_rng(12)
    rows, cols = 300, 250
    lens = rng.integers(0, 9, rows)
    lens[:5] = 0  # empty rows
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ci = rng.integers(0, cols, rp[-1]).astype(np.uint32)
    cv = rng.standard_normal(rp[-1]).astype(np.float32)
    A = naga_amd.CsrMatrix.create(rows, cols, rp, ci, cv)
    x = rng.standard_normal(cols).astype(np.float32)
    got = A.mult(x)
    want = oracle64.csr_mult(rp, ci.astype(np.int64), cv.astype(np.float64), x)
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    assert np.all(got[:5] == 0.0)
    bad = ci.copy()
    bad[7] = cols
    with pytest.raises(ValueError):
        naga_amd.CsrMatrix.create(rows, cols, rp, bad, cv)
    square = naga_amd.CsrMatrix.create_from_index_stencil(
        np.array([[0, 1], [1, 2], [2, 0]]), np.ones((3, 2)))
    assert np.array_equal(square.mult(np.array([1.0, 2.0, 4.0])), [3.0, 6.0, 5.0])
