"""Pins of the CPU oracle (the checker) against known answers, independent
computations and properties the mathematics fixes.  CPU only."""
import itertools
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CS2 = 1.0 / 3.0


# ---------------------------------------------------------------- lattices
def q(t):
    return 3.0 * t * t - 2.0


# The cascade basis as lattice polynomials (an independent representation of
# the collision-matrix columns).
BASIS27 = [
    lambda x, y, z: 1.0, lambda x, y, z: x, lambda x, y, z: y, lambda x, y, z: z,
    lambda x, y, z: x * y, lambda x, y, z: x * z, lambda x, y, z: y * z,
    lambda x, y, z: x * x - y * y, lambda x, y, z: x * x + y * y - 2 * z * z,
    lambda x, y, z: x * x + y * y + z * z - 2,
    lambda x, y, z: x * (3 * (y * y + z * z) - 4), lambda x, y, z: y * (3 * (x * x + z * z) - 4),
    lambda x, y, z: z * (3 * (x * x + y * y) - 4),
    lambda x, y, z: x * (y * y - z * z), lambda x, y, z: y * (x * x - z * z),
    lambda x, y, z: z * (x * x - y * y), lambda x, y, z: x * y * z,
    lambda x, y, z: (q(x) * q(y) + q(x) * q(z) + q(y) * q(z)) / 3,
    lambda x, y, z: (q(x) * q(y) + q(x) * q(z) - 2 * q(y) * q(z)) / 3,
    lambda x, y, z: q(x) * (y * y - z * z), lambda x, y, z: y * z * q(x),
    lambda x, y, z: x * z * q(y), lambda x, y, z: x * y * q(z),
    lambda x, y, z: x * q(y) * q(z), lambda x, y, z: y * q(x) * q(z),
    lambda x, y, z: z * q(x) * q(y), lambda x, y, z: q(x) * q(y) * q(z),
]
BASIS9 = [
    lambda x, y: 1.0, lambda x, y: x, lambda x, y: y, lambda x, y: q(x) + q(y),
    lambda x, y: x * x - y * y, lambda x, y: -x * y, lambda x, y: -y * q(x),
    lambda x, y: -x * q(y), lambda x, y: q(x) * q(y),
]


@pytest.mark.parametrize("Q", [9, 27])
def test_lattice_invariants(oracle64, Q):
    c, w, bb, K = oracle64.lattice(Q)
    D = c.shape[1]
    assert abs(w.sum() - 1) < 1e-15
    assert np.abs(w @ c).max() < 1e-16
    assert np.allclose(np.einsum("a,ai,aj->ij", w, c, c), CS2 * np.eye(D), atol=1e-15)
    # fourth-order isotropy: sum w c_x^4 = 3 c_s^4, sum w c_x^2 c_y^2 = c_s^4
    assert abs((w * c[:, 0] ** 4).sum() - 3 * CS2 * CS2) < 1e-15
    assert abs((w * c[:, 0] ** 2 * c[:, 1] ** 2).sum() - CS2 * CS2) < 1e-15
    # product weights w = prod_d w1(c_d), w1 = (1/6, 2/3, 1/6)
    w1 = {-1: 1 / 6, 0: 2 / 3, 1: 1 / 6}
    assert np.allclose(w, [np.prod([w1[int(v)] for v in row]) for row in c], atol=1e-16)
    assert np.array_equal(bb[bb], np.arange(Q))
    assert np.array_equal(c[bb], -c)
    # collision matrix columns are the cascade's lattice polynomials
    basis = BASIS27 if Q == 27 else BASIS9
    Kp = np.array([[P(*row) for P in basis] for row in c])
    assert np.array_equal(K, Kp)


@pytest.mark.parametrize("Q", [9, 27])
def test_feq_moments(oracle64, Q):
    """The reference's equilibrium (with its -u.u term,
    equilibrium_distribution.cuh:61-63) has sum feq = rho (1 + |u|^2/2) and
    sum feq c = rho u."""
    c, w, _, _ = oracle64.lattice(Q)
    rng = np.random.default_rng(Q)
    for _ in range(5):
        rho = 1 + 0.1 * rng.standard_normal()
        u = 0.05 * rng.standard_normal(c.shape[1])
        feq = np.array([oracle64.feq(Q, rho, u, a) for a in range(Q)])
        assert abs(feq.sum() - rho * (1 + 0.5 * u @ u)) < 1e-14
        assert np.abs(feq @ c - rho * u).max() < 1e-14
        assert np.allclose(np.array([oracle64.feq(Q, rho, 0 * u, a) for a in range(Q)]), w * rho)


def _cascade_k(Q, c, K, f, rho, u, rates, reference_form=False):
    """First-principles cascade: k_0..k_D = 0 and k = Omega B(u)^-1 (rho e - C)
    with C_j = sum_a f_a P_j(c_a - u), B_jl = sum_a P_l(c_a) P_j(c_a - u), e_j
    the unit-density Maxwellian (c_s^2 = 1/3) central moments of P_j.
    reference_form: C_j expanded over the central moments with the zeroth
    taken as rho and the first as zero (the reference's closed forms contain
    neither; it matters when rho != sum f or u != j / sum f)."""
    D = c.shape[1]
    basis = BASIS27 if Q == 27 else BASIS9
    x = c - u
    Pshift = np.array([[P(*row) for P in basis] for row in x])  # [a, j]
    B = Pshift.T @ K
    mons = list(itertools.product(range(3), repeat=D))
    V = np.array([[np.prod([row[d] ** m[d] for d in range(D)]) for m in mons] for row in c])
    coef = np.linalg.solve(V, K)  # P_j = sum_mon coef[mon, j] mon
    if reference_form:
        M = np.array([f @ np.prod([x[:, d] ** m[d] for d in range(D)], axis=0) for m in mons])
        for i, m in enumerate(mons):
            if sum(m) == 0:
                M[i] = rho
            elif sum(m) == 1:
                M[i] = 0.0
        C = M @ coef
    else:
        C = f @ Pshift
    # Maxwellian central moments <x^a y^b z^c> via exact Gaussian moments
    gm = {0: 1.0, 1: 0.0, 2: CS2}
    e = np.array([sum(coef[i, j] * np.prod([gm[m[d]] for d in range(D)])
                      for i, m in enumerate(mons)) for j in range(Q)])
    r = list(range(D + 1, Q))
    k = np.zeros(Q)
    k[r] = np.asarray(rates)[r] * np.linalg.solve(B[np.ix_(r, r)], (rho * e - C)[r])
    return k


def test_collision_d3q27_is_the_cascade(oracle64):
    """detail/d3q27_lattice.cuh:125-378 equals the first-principles cascade
    for random f and u != 0 (to the FP32 rounding of the unit rates)."""
    c, w, _, K = oracle64.lattice(27)
    rng = np.random.default_rng(5)
    for amp in (1e-3, 2e-2, 5e-2):
        f = w * (1 + amp * rng.standard_normal(27))
        rho = f.sum()
        u = f @ c / rho
        nu, dt = 0.0333, 0.004
        df = oracle64.collide(27, f, rho, u, nu, dt)
        k = np.linalg.solve(K, df / dt)

        def rate(unit):
            return float(np.float32(unit)) / dt

        shear = 1 / (3 * nu + 0.5) / dt
        rates = [0] * 4 + [shear] * 6 + [rate(1.5)] * 3 + [rate(1.83)] * 3 + [rate(1.4)] \
            + [rate(1.61)] + [rate(1.98)] * 5 + [rate(1.74)] * 4
        kfp = _cascade_k(27, c, K, f, rho, u, rates)
        assert np.abs(k - kfp).max() <= 1e-9 * np.abs(kfp).max()
        # conservation: mass and momentum increments vanish
        assert abs(df.sum()) < 1e-15 and np.abs(df @ c).max() < 1e-15
        # rho != sum f (a density source) and u != j / sum f
        rs, us = rho * (1 + 1e-3), u * 0.9 + 1e-4
        k = np.linalg.solve(K, oracle64.collide(27, f, rs, us, nu, dt) / dt)
        kfp = _cascade_k(27, c, K, f, rs, us, rates, reference_form=True)
        assert np.abs(k - kfp).max() <= 1e-9 * np.abs(kfp).max()


def test_collision_d2q9(oracle64):
    """D2Q9 (detail/d2q9_lattice.cuh:120-299, completed): exact cascade at
    u = 0; for u != 0, k3 and k4 are the cascade and k5 is the cascade with
    the sign of its rho u_x u_y term flipped, as the reference writes it
    (detail/d2q9_lattice.cuh:203); mass and momentum are conserved."""
    c, w, bb, K = oracle64.lattice(9)
    rng = np.random.default_rng(9)
    nu, dt = 0.01, 0.3
    om_ab = 1 / (3 * nu + 0.5 * dt)
    rates = [0, 0, 0, 1 / dt, om_ab, om_ab, 1 / dt, 1 / dt, 1 / dt]
    f = w * (1 + 0.05 * rng.standard_normal(9))
    f = 0.5 * (f + f[bb])  # zero momentum
    rho = f.sum()
    u = f @ c / rho
    k = np.linalg.solve(K, oracle64.collide(9, f, rho, u, nu, dt) / dt)
    kfp = _cascade_k(9, c, K, f, rho, u, rates)
    assert np.abs(k - kfp).max() < 1e-12 * np.abs(kfp).max()
    for _ in range(3):
        f = w * (1 + 0.05 * rng.standard_normal(9))
        rho = f.sum()
        u = f @ c / rho
        df = oracle64.collide(9, f, rho, u, nu, dt)
        k = np.linalg.solve(K, df / dt)
        kfp = _cascade_k(9, c, K, f, rho, u, rates)
        assert np.allclose(k[3:5], kfp[3:5], rtol=1e-10, atol=1e-14)
        assert abs(k[5] - (kfp[5] + om_ab * rho * u[0] * u[1] / 2)) < 1e-12
        assert abs(df.sum()) < 1e-15 and np.abs(df @ c).max() < 1e-15


@pytest.mark.parametrize("Q", [9, 27])
def test_collision_rest_equilibrium_is_fixed_point(oracle64, oracle32, Q):
    c, w, _, _ = oracle64.lattice(Q)
    for o in (oracle64, oracle32):
        for rho in (1.0, 1.001):
            df = o.collide(Q, w * rho, rho, np.zeros(c.shape[1]), 0.0333, 0.004)
            assert np.abs(df).max() < (1e-14 if o.bits == 64 else 2e-5)


def test_rk4_scalar_closed_form(oracle64, oracle32):
    """rk4_solver.cuh:48-171 on y' = lambda y: one step multiplies by
    1 + z + z^2/2 + z^3/6 + z^4/24."""
    for lam, dt, n in ((-1.0, 0.1, 10), (-3.0, 0.05, 7), (0.5, 0.2, 5)):
        z = lam * dt
        R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
        assert abs(oracle64.rk4_scalar(1.0, lam, dt, n) - R ** n) < 1e-14
        assert abs(oracle32.rk4_scalar(1.0, lam, dt, n) - R ** n) < 1e-6


# ------------------------------------------------------------- node provider
def splitmix64_stream(seed):
    s = seed
    while True:
        s = (s + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        yield z ^ (z >> 31)


def test_splitmix64_published_vector():
    g = splitmix64_stream(0)
    assert [next(g) for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                            0x06C45D188009454F]


def hilbert(g, D):
    """Hilbert key of grid point g (Skilling's transposed-index algorithm)."""
    bits = 32 if D == 2 else 21
    x = [int(v) for v in g] + [0] * (3 - D)
    top = 1 << (bits - 1)
    q = top
    while q > 1:
        p = q - 1
        for i in range(D):
            if x[i] & q:
                x[0] ^= p
            else:
                t = (x[0] ^ x[i]) & p
                x[0] ^= t
                x[i] ^= t
        q >>= 1
    for i in range(1, D):
        x[i] ^= x[i - 1]
    flip, q = 0, top
    while q > 1:
        if x[D - 1] & q:
            flip ^= q - 1
        q >>= 1
    key = 0
    for b in range(bits - 1, -1, -1):
        for i in range(D):
            key = (key << 1) | (((x[i] ^ flip) >> b) & 1)
    return key


@pytest.mark.parametrize("D,n", [(2, 16), (3, 8)])
def test_hilbert_key_is_a_curve(D, n):
    """What fixes the curve: on a 2^m grid the keys are distinct and
    consecutive keys are grid neighbours (L1 distance 1), with the walk
    starting at the origin."""
    pts = [tuple(int(v) for v in np.unravel_index(t, (n,) * D)) for t in range(n ** D)]
    order = sorted(pts, key=lambda g: hilbert(g, D))
    assert len({hilbert(g, D) for g in pts}) == n ** D
    assert order[0] == (0,) * D
    steps = [sum(abs(a - b) for a, b in zip(p, q)) for p, q in zip(order, order[1:])]
    assert all(s == 1 for s in steps)


@pytest.mark.parametrize("D,h,L", [(2, 0.02, 0.1), (3, 0.1, 0.25)])
def test_uniform_grid(oracle64, D, h, L):
    o = oracle64
    seed = 12345
    nd = o.uniform_grid(D, [0] * D, [1] * D, h, jitter=0.25, seed=seed, layer=L, sigma_max=0.5)
    n1 = int(round(1 / h)) + 1
    N = n1 ** D
    nb, nl, nbd = nd["n_bulk"], nd["n_layer"], nd["n_boundary"]
    assert nb + nl + nbd == N
    assert nbd == N - (n1 - 2) ** D
    pts = nd["points"].astype(np.float64)
    gidx = nd["grid_index"]
    gi = np.stack([(gidx // n1 ** d) % n1 for d in range(D)], 1)
    # exact jitter from the published SplitMix64 stream, D draws per node
    stream = splitmix64_stream(seed)
    u01 = np.array([next(stream) >> 11 for _ in range(N * D)], np.float64) * 2.0 ** -53
    u01 = u01.reshape(N, D)
    sp = 1.0 / (n1 - 1)
    expect = gi * sp + np.where(
        np.any((gi == 0) | (gi == n1 - 1), axis=1, keepdims=True), 0.0,
        (2 * u01[gidx] - 1) * 0.25 * h)
    assert np.array_equal(pts.astype(np.float32), expect.astype(np.float32))
    bnd = slice(nb + nl, N)
    assert np.all(np.any((gi[bnd] == 0) | (gi[bnd] == n1 - 1), axis=1))
    # layer: brute-force FP32 distance to the boundary set <= L
    bp = nd["points"][bnd]
    ip = nd["points"][: nb + nl]
    d2 = ((ip[:, None, :] - bp[None, :, :]) ** 2).sum(-1).min(1)
    inlayer = d2 <= np.float32(L) ** 2
    assert not inlayer[:nb].any() and inlayer[nb:].all()
    r = (np.float32(L) - np.sqrt(d2[nb:].astype(np.float32))) / np.float32(L)
    assert np.allclose(nd["sigma"], 0.5 * r * r, rtol=1e-6, atol=1e-7)
    # inward unit normals
    nrm = nd["normals"]
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-6)
    side = np.where(gi[bnd] == 0, 1.0, np.where(gi[bnd] == n1 - 1, -1.0, 0.0))
    assert np.allclose(nrm, side / np.linalg.norm(side, axis=1, keepdims=True), atol=1e-7)
    # Hilbert order within each class
    for sl in (slice(0, nb), slice(nb, nb + nl), bnd):
        keys = [hilbert(g, D) for g in gi[sl]]
        assert keys == sorted(keys)


# --------------------------------------------------------------------- kNN
def knn_numpy(ref, q, k):
    ref = ref.astype(np.float32)
    q = q.astype(np.float32)
    D = ref.shape[1]
    out_d, out_i = [], []
    for x in q:
        diff = x[None, :] - ref
        s = diff[:, 0] * diff[:, 0]
        for d in range(1, D):
            s = (diff[:, d] * diff[:, d]) + s
        order = np.lexsort((np.arange(len(ref)), s))[:k]
        out_d.append(s[order])
        out_i.append(order)
    return np.array(out_d, np.float32), np.array(out_i, np.uint32)


@pytest.mark.parametrize("D,k", [(2, 16), (3, 32), (3, 1)])
def test_knn_vs_numpy_and_brute(oracle64, D, k):
    rng = np.random.default_rng(D + k)
    ref = rng.random((700, D)).astype(np.float32)
    qs = (rng.random((150, D)) * 1.3 - 0.15).astype(np.float32)
    d2, idx = oracle64.knn(ref, qs, k)
    bd, bi = oracle64.knn(ref, qs, k, brute=True)
    nd, ni = knn_numpy(ref, qs, k)
    assert np.array_equal(idx, bi) and np.array_equal(idx, ni)
    assert np.array_equal(d2, bd) and np.array_equal(d2, nd)


def test_knn_reference_grid_kat(oracle64):
    """examples/knn_validation_3d.cu:37-138 protocol: 20^3 grid of spacing
    0.1, k = 64, 30 queries in [0, 1.9]^3; neighbour sets equal brute force
    (here with exact ties resolved by index)."""
    g = np.arange(20) * 0.1
    ref = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    qs = (np.random.default_rng(0).random((30, 3)) * 1.9).astype(np.float32)
    _, idx = oracle64.knn(ref, qs, 64)
    _, ni = knn_numpy(ref, qs, 64)
    assert np.array_equal(idx, ni)


# ---------------------------------------------------------------------- LU
def test_lu_known_answer(oracle64):
    kat = json.load(open(os.path.join(GOLDEN, "inverse3x3.json")))
    X, info = oracle64.lu_solve(np.array(kat["A"]), np.eye(3))
    assert info == 0
    assert np.abs(X - np.array(kat["A_inv"])).max() < 1e-6


def test_lu_vs_numpy(oracle64):
    rng = np.random.default_rng(1)
    for n in (5, 19, 36, 52):
        A = rng.standard_normal((n, n))
        B = rng.standard_normal((n, 3))
        X, info = oracle64.lu_solve(A, B)
        assert info == 0
        assert np.abs(X - np.linalg.solve(A, B)).max() < 1e-10 * np.abs(X).max()
    X, info = oracle64.lu_solve(np.zeros((3, 3)), np.ones(3))
    assert info == 1


# -------------------------------------------------------------------- RPIM
@pytest.mark.parametrize("D,rbf", [(2, 1), (3, 1), (2, 0), (3, 0)])
def test_rpim_vs_scipy(oracle64, D, rbf):
    """Cardinal weights of the augmented RBF interpolant equal SciPy's
    RBFInterpolator with the same kernel and a linear polynomial tail
    (multiquadric: SciPy's -sqrt(1 + (r/c)^2) is a constant multiple of
    sqrt(r^2 + c^2), which leaves the cardinal functions unchanged)."""
    from scipy.interpolate import RBFInterpolator
    rng = np.random.default_rng(D * 7 + rbf)
    k = 16 if D == 2 else 32
    src = (rng.random((k, D)) * 0.05).astype(np.float32)
    qs = (rng.random((5, D)) * 0.04 + 0.005).astype(np.float32)
    idx = np.tile(np.arange(k, dtype=np.uint32), (len(qs), 1))
    spacing = 0.01
    w = oracle64.rpim_weights(src, idx, qs, rbf, spacing)
    c = 0.1 * spacing
    kw = dict(kernel="cubic") if rbf == 1 else dict(kernel="multiquadric", epsilon=1 / c)
    ref = np.zeros((len(qs), k))
    for s in range(k):
        e = np.zeros(k)
        e[s] = 1.0
        ref[:, s] = RBFInterpolator(src.astype(np.float64), e, degree=1, **kw)(qs.astype(np.float64))
    assert np.abs(w - ref).max() < 1e-6 * max(1, np.abs(ref).max())
    # linear reproduction and the interpolating property
    assert np.abs(w.sum(1) - 1).max() < 1e-9
    assert np.abs(w @ src.astype(np.float64) - qs).max() < 1e-9
    wi = oracle64.rpim_weights(src, idx[:3], src[:3], rbf, spacing)
    assert np.abs(wi - np.eye(k)[:3]).max() < 1e-8


# -------------------------------------------------------------- divergence
@pytest.mark.parametrize("D,h,k", [(2, 0.02, 16), (3, 0.1, 32)])
def test_divergence_consistency(oracle64, D, h, k):
    """nonlocal divergence (detail/divergence.inl): zero on constants, one
    common quadrature constant on linear fields at every interior node
    (RPIM reproduces them exactly), and close to 1 (the kernel's continuum
    normalisation); the reference demo field (examples/nonlocal_divergence.cu:
    40-57) converges."""
    o = oracle64
    nd = o.uniform_grid(D, [0] * D, [1] * D, h, jitter=0.25)
    pts = nd["points"]
    d2, idx = o.knn(pts, pts, k)
    W, delta = o.divergence_weights(pts, pts, idx, d2, 1)
    assert np.allclose(delta, 0.05 * np.sqrt(d2[:, 1].astype(np.float64)))
    interior = np.all((pts > 0.3) & (pts < 0.7), axis=1)
    x = pts.astype(np.float64)
    const = o.divergence_apply(idx, W, np.ones_like(x))
    assert np.abs(const[interior]).max() < 1e-8 * np.abs(W).max()
    for d in range(D):
        F = np.zeros_like(x)
        F[:, d] = x[:, d]
        lin = o.divergence_apply(idx, W, F)[interior]
        assert np.ptp(lin) < 1e-8
        assert 0.9 < lin.mean() < 1.1
        F = np.zeros_like(x)
        F[:, (d + 1) % D] = x[:, d]  # divergence-free linear field
        assert np.abs(o.divergence_apply(idx, W, F)[interior]).max() < 1e-8
    if D == 2:
        F = np.stack([np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1]),
                      np.zeros(len(x))], 1)
        got = o.divergence_apply(idx, W, F)[interior] / lin.mean()
        want = 2 * np.pi * np.cos(2 * np.pi * x[interior, 0]) * np.cos(2 * np.pi * x[interior, 1])
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 0.02


# Linear-field constants of the reference's quadrature, in closed form
# (nonlocal_calculus/detail/quadrature_points.cuh:40-152,
# detail/divergence.inl:53-227).  The kernel alpha(y - x) = K(r) n is
# normalised so that the continuum divergence of a linear field is exact
# (2D: (240/7)/(pi delta^5) k(r), int r^2 k = 7/240 delta^5; 3D:
# 48/(pi delta^6) k(r), int r^3 k = delta^6/64; k(r) = r(2 delta - 3r) below
# delta/2, (delta - r)^2 below delta).  RPIM reproduces linear fields, so the
# discrete divergence of F = x_d e_d is the quadrature of that normalised
# integral: the 4-point Gauss-Legendre rule (the reference's 6-digit table)
# on the radial moment times the angular rule.  The theta rule (0, 90, 180,
# 270 degrees, weight pi/2) integrates cos^2 exactly; the 3D phi rule is the
# 4-point midpoint rule on [0, pi] (phi = 22.5 + 45 j degrees), which
# integrates sin^3 (x and y) and sin cos^2 (z) with different errors: the
# 3D operator is anisotropic.
GL_X = np.array([0.339981, -0.339981, 0.861136, -0.861136])
GL_W = np.array([0.652145, 0.652145, 0.347855, 0.347855])


def _kernel_profile(r):
    return np.where(2 * r < 1, r * (2 - 3 * r), np.where(r < 1, (1 - r) ** 2, 0.0))


def quadrature_linear_constants(D):
    """(c_x, c_y[, c_z]): discrete divergence of F = x_d e_d."""
    from scipy.integrate import quad
    p = D  # radial moment r^D k(r) (r^(D-1) from the volume element)
    r = (GL_X + 1) / 2
    gl = np.sum(GL_W / 2 * r ** p * _kernel_profile(r))
    exact = (quad(lambda x: x ** p * _kernel_profile(x), 0, 0.5)[0]
             + quad(lambda x: x ** p * _kernel_profile(x), 0.5, 1)[0])
    radial = gl / exact
    th = np.pi / 2 * np.arange(4)
    theta = (np.pi / 2) * np.sum(np.cos(th) ** 2) / np.pi  # = 1
    if D == 2:
        return np.array([radial * theta] * 2)
    phi = np.pi / 4 * np.arange(4) + np.pi / 8
    sx = (np.pi / 4) * np.sum(np.sin(phi) ** 3) / (4.0 / 3.0)
    sz = (np.pi / 4) * np.sum(np.sin(phi) * np.cos(phi) ** 2) / (2.0 / 3.0)
    return np.array([radial * theta * sx, radial * theta * sx, radial * sz])


@pytest.mark.parametrize("D,h,k", [(2, 0.01, 16), (2, 0.05, 16), (3, 0.05, 32), (3, 0.1, 32)])
def test_divergence_linear_constants(oracle64, D, h, k):
    """The oracle's divergence of x_d e_d at interior nodes equals the
    closed-form quadrature constant of axis d at every spacing: 2D
    0.980808 (radial Gauss-Legendre error alone); 3D 0.972372 along x and
    y, 1.063616 along z."""
    o = oracle64
    nd = o.uniform_grid(D, [0] * D, [1] * D, h, jitter=0.25)
    pts = nd["points"]
    d2, idx = o.knn(pts, pts, k)
    W, _ = o.divergence_weights(pts, pts, idx, d2, 1)
    interior = np.all((pts > 0.3) & (pts < 0.7), axis=1)
    x = pts.astype(np.float64)
    want = quadrature_linear_constants(D)
    got = []
    for d in range(D):
        F = np.zeros_like(x)
        F[:, d] = x[:, d]
        got.append(o.divergence_apply(idx, W, F)[interior].mean())
    print(f"D={D} h={h}: linear-field constants {np.round(got, 7)} (closed form {np.round(want, 7)})")
    assert np.abs(np.array(got) - want).max() < 1e-8
    if D == 3:
        assert abs(want[0] - 0.9723721) < 1e-6 and abs(want[2] - 1.0636160) < 1e-6
    else:
        assert abs(want[0] - 0.9808084) < 1e-6


def directional_peak_times(step, drho, points, center, radii, h, steps, dt, cone_deg=25.0):
    """Peak times of the mean density deviation on each sphere |x - c| = R
    (within 0.6 h), restricted to the cones of half-angle cone_deg around
    the +-x, +-y and +-z axes, and over the whole sphere.  step() advances
    one step, drho() returns rho - rho0.  Parabolic interpolation between
    steps.  Returns {0: t_x, 1: t_y, 2: t_z, "shell": t}."""
    p = points.astype(np.float64) - np.asarray(center)
    r = np.linalg.norm(p, axis=1)
    cos = np.cos(np.radians(cone_deg))
    masks = {ax: [(np.abs(p[:, ax]) > cos * r) & (np.abs(r - R) < 0.6 * h) for R in radii]
             for ax in range(3)}
    masks["shell"] = [np.abs(r - R) < 0.6 * h for R in radii]
    series = {key: [] for key in masks}
    for _ in range(steps):
        step()
        d = drho()
        for key, ms in masks.items():
            series[key].append([d[m].mean() for m in ms])
    out = {}
    for key, sr in series.items():
        sr = np.array(sr)
        t = []
        for j in range(len(radii)):
            i = int(np.argmax(sr[:, j]))
            assert 0 < i < len(sr) - 1, "no interior peak"
            a, b, c = sr[i - 1:i + 2, j]
            t.append((i + 1) * dt + 0.5 * dt * (a - c) / (a - 2 * b + c))
        out[key] = np.array(t)
    return out


def test_engine_pulse_speed_3d(oracle64):
    """3D acoustics of the reference method, filter off: a central Gaussian
    pulse (width 0.06) on the unit cube at h = 0.02 travels along each axis
    at c_s times that axis's quadrature constant within 1% (0.972 c_s along
    x and y, 1.064 c_s along z), so the 3D front is not spherical; the
    whole-shell average lies between the two."""
    h, width = 0.02, 0.06
    radii = np.array([0.15, 0.2, 0.25, 0.3])
    nd = oracle64.uniform_grid(3, [0] * 3, [1] * 3, h, jitter=0.25, layer=0.1, sigma_max=0.5)
    e = oracle64.engine(nd, Q=27, k=32, h=h, threads=8, filter_strength=0.0)
    e.set_pulse([0.5, 0.5, 0.5], 1e-3, width)
    dt = 0.4 * h
    cs = np.sqrt(CS2)
    steps = int(1.25 * radii.max() / (0.97 * cs) / dt)
    t = directional_peak_times(lambda: e.step(1), lambda: e.get()["drho"], nd["points"],
                               [0.5] * 3, radii, h, steps, dt)
    c = quadrature_linear_constants(3)
    speeds = {key: np.polyfit(t[key], radii, 1)[0] / cs for key in t}
    print("3D directional speed / c_s:", {k_: round(v, 4) for k_, v in speeds.items()})
    for ax in range(3):
        assert abs(speeds[ax] / c[ax] - 1) < 0.01, (ax, speeds[ax], c[ax])
    assert c[0] < speeds["shell"] < c[2]


def test_divergence_apply_vs_numpy(oracle64):
    rng = np.random.default_rng(2)
    n, k, D = 300, 16, 3
    idx = rng.integers(0, n, (n, k)).astype(np.uint32)
    W = rng.standard_normal((n, k, D))
    F = rng.standard_normal((n, D))
    out = oracle64.divergence_apply(idx, W, F, 0.5)
    ref = np.einsum("isd,isd->i", W, F[idx] - 0.5)
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()


# ------------------------------------------------------------------ engine
def test_engine_rest_state_is_steady(oracle64):
    nd = oracle64.uniform_grid(2, [0, 0], [1, 1], 0.05, jitter=0.25, layer=0.1, sigma_max=0.5)
    e = oracle64.engine(nd, Q=9, k=16, h=0.05)
    e.step(5)
    s = e.get()
    assert np.abs(s["drho"]).max() < 1e-14 and np.abs(s["u"]).max() < 1e-14


def test_engine_fp32_tracks_fp64(oracle64, oracle32):
    nd = oracle64.uniform_grid(2, [0, 0], [1, 1], 0.02, jitter=0.25, layer=0.1, sigma_max=0.5)
    a = oracle64.engine(nd, Q=9, k=16, h=0.02)
    b = oracle32.engine(nd, Q=9, k=16, h=0.02)
    for e in (a, b):
        e.set_pulse([0.5, 0.5], 1e-3, 0.08)
        e.step(5)
    ra, rb = a.get(), b.get()
    assert np.linalg.norm(ra["drho"] - rb["drho"]) / np.linalg.norm(ra["drho"]) < 1e-2


@pytest.mark.parametrize("Q", [9, 27])
def test_engine_mirror_symmetry(oracle64, Q):
    """A pulse centred on the mirror plane of an exactly x-symmetric node set
    stays symmetric: rho and u_y even, u_x odd (a sign, axis or index slip in
    any pass breaks this).  With the reference's PML indexing
    (pml_absorption.inl:445-448 reads Q at j - num_bulk, which ties the
    field to storage order rather than geometry) the symmetry is lost, which
    the last check confirms."""
    from tests.scenes import mirror_nodes
    D = 2 if Q == 9 else 3
    nd, mirror = mirror_nodes(D, m=8 if D == 2 else 5, n=17 if D == 2 else 9)
    k = 16 if D == 2 else 32
    lo = nd["points"].min(0).astype(np.float64)
    hi = nd["points"].max(0).astype(np.float64)
    centre = 0.5 * (lo + hi)
    centre[0] = 0.0
    res = {}
    for ref in (0, 1):
        e = oracle64.engine(nd, Q=Q, k=k, h=0.05, pml_ref_index=ref)
        e.set_pulse(centre, 1e-3, 0.1)
        e.step(10)
        s = e.get()
        flip = np.ones((D, 1))
        flip[0] = -1
        res[ref] = max(np.abs(s["drho"] - s["drho"][mirror]).max() / np.abs(s["drho"]).max(),
                       np.abs(s["u"] - flip * s["u"][:, mirror]).max() / np.abs(s["u"]).max())
    assert res[0] < 1e-10
    assert res[1] > 1e-3


def acoustic_peak_time(r, width, t_lo, t_hi):
    """Peak time at radius r of the exact 2D linear acoustic solution for an
    initial Gaussian density bump exp(-|x|^2 / (2 width^2)) at rest:
    rho'(r, t) = int_0^inf width^2 exp(-k^2 width^2 / 2) cos(c_s k t) J0(k r) k dk."""
    from scipy.special import j0
    cs = np.sqrt(CS2)
    kk = np.linspace(0, 12 / width, 6000)
    amp = width * width * np.exp(-0.5 * (kk * width) ** 2) * kk
    ts = np.linspace(t_lo, t_hi, 2001)
    vals = [np.trapezoid(amp * np.cos(cs * kk * t) * j0(kk * r), kk) for t in ts]
    return ts[int(np.argmax(vals))]


def _peak_times(oracle, filter_strength, radii, width, h=0.01, steps=260):
    nd = oracle.uniform_grid(2, [0, 0], [1, 1], h, jitter=0.25, layer=0.1, sigma_max=0.5)
    e = oracle.engine(nd, Q=9, k=16, h=h, threads=8, filter_strength=filter_strength)
    e.set_pulse([0.5, 0.5], 1e-3, width)
    p = nd["points"].astype(np.float64)
    r = np.hypot(p[:, 0] - 0.5, p[:, 1] - 0.5)
    rings = [np.abs(r - R) < 0.5 * h for R in radii]
    dt = 0.4 * h
    series = []
    for _ in range(steps):
        e.step(1)
        d = e.get()["drho"]
        series.append([d[m].mean() for m in rings])
    series = np.array(series)
    out = []
    for j in range(len(radii)):
        i = int(np.argmax(series[:, j]))
        a, b, c = series[i - 1:i + 2, j]
        out.append((i + 1) * dt + 0.5 * dt * (a - c) / (a - 2 * b + c))
    return np.array(out)


def test_engine_pulse_speed(oracle64):
    """Acoustics: the front of a Gaussian pulse crosses probe rings at the
    times of the exact linear solution (speed c_s = 1/sqrt(3)) within 1% of
    the time of flight, with the selective filter off.  The reference's filter
    (strength 0.75 per step, simulation_engine.cuh:1242, applied per step and
    so independent of dt) smooths the front and delays it; with it on the
    front is 2-6% slow at dt = 0.4 h, which the second check records."""
    radii = np.array([0.15, 0.2, 0.25, 0.3, 0.35])
    width = 0.04
    cs = np.sqrt(CS2)
    exact = np.array([acoustic_peak_time(R, width, 0.8 * R / cs - 0.1, 1.2 * R / cs)
                      for R in radii])
    t0 = _peak_times(oracle64, 0.0, radii, width)
    assert np.abs(t0 - exact).max() < 0.01 * radii.min() / cs, (t0, exact)
    speed0 = np.polyfit(t0, radii, 1)[0] / cs
    assert abs(speed0 - 1) < 0.01
    t1 = _peak_times(oracle64, 0.75, radii, width)
    speed1 = np.polyfit(t1, radii, 1)[0] / cs
    assert 0.94 < speed1 < 0.98, speed1


def test_density_source(oracle64):
    """Declarative density source (detail/simulation_engine.cuh:347-364,
    1057-1107): from rest, one step with a source active during [0, dt)
    sets the step's density output to rho0 + s/2 at every node (the
    reference's rho <- (rho + (rho + s)) / 2, once per node) with
    s = A P(r / R), P(x) = (1 - (2x)^2/4) exp(-(2x)^2), and adds
    sum(s) / rho0 of mass (to the scheme's non-conservation, < 0.5%); a
    source that starts later leaves the rest state untouched until then."""
    h, rho0, A, R = 0.02, 1.2, 1e-3, 0.1
    nd = oracle64.uniform_grid(2, [0, 0], [1, 1], h, jitter=0.25, layer=0.1, sigma_max=0.5)
    dt = 0.4 * h
    e = oracle64.engine(nd, Q=9, k=16, h=h, rho0=rho0)
    e.set_pulse([0.5, 0.5], 0.0, 1.0)
    e.add_density_source([0.5, 0.5], R, A, t_end=dt)
    m0 = e.get()["f"].sum()
    e.step(1)
    p = nd["points"].astype(np.float64)
    x = 2 * np.hypot(p[:, 0] - 0.5, p[:, 1] - 0.5) / R
    s = np.where(x < 2, A * (1 - x * x / 4) * np.exp(-x * x), 0.0)
    assert s.max() == pytest.approx(A, rel=2e-2)
    assert np.abs(e.get(fresh=False)["rho"] - (rho0 + s / 2)).max() < 1e-14
    gain = e.get()["f"].sum() - m0
    assert abs(gain / (s.sum() / rho0) - 1) < 5e-3
    late = oracle64.engine(nd, Q=9, k=16, h=h, rho0=rho0)
    late.set_pulse([0.5, 0.5], 0.0, 1.0)
    late.add_density_source([0.5, 0.5], R, A, frequency=3.0, t_start=5.5 * dt)
    late.step(6)
    assert np.abs(late.get()["drho"]).max() < 1e-13
    late.step(1)
    assert np.abs(late.get()["drho"]).max() > 1e-6


def test_variable_stencil_weights(oracle64):
    """Per-node stencil widths: node i's weights are those of a fixed-width
    operator over its first k_i neighbours, zero after them."""
    nd = oracle64.uniform_grid(3, [0] * 3, [1] * 3, 0.125, jitter=0.25)
    pts = nd["points"]
    d2, idx = oracle64.knn(pts, pts, 32)
    rng = np.random.default_rng(8)
    kv = rng.integers(20, 33, len(pts)).astype(np.uint8)
    W, _ = oracle64.divergence_weights(pts, pts, idx, d2, 1, kvar=kv)
    for i in rng.choice(len(pts), 6, replace=False):
        ki = int(kv[i])
        Wi, _ = oracle64.divergence_weights(pts, pts[i:i + 1], idx[i:i + 1, :ki],
                                            d2[i:i + 1, :ki], 1)
        assert np.allclose(W[i, :ki], Wi[0], rtol=1e-12, atol=1e-12)
        assert np.all(W[i, ki:] == 0)


# -------------------------------------------------------------- advection
def _advection_matrix(idx, W, v):
    """Dense L with rate(f) = L (f - off), assembled entry by entry."""
    n, k, D = W.shape
    L = np.zeros((n, n))
    for i in range(n):
        for s in range(k):
            L[i, idx[i, s]] -= np.dot(W[i, s], v)
    return L


@pytest.mark.parametrize("end", [0, 70])
def test_advection_rk4_closed_form(oracle64, end):
    """For a constant velocity the advection task (advection.cuh:207-276) is
    RK4 on the linear system dg/dt = P L g, g = f - off, P the projection
    onto [0, end): one step is the degree-4 Taylor polynomial of exp(dt P L),
    and nodes past end keep f0."""
    o = oracle64
    nd = o.uniform_grid(2, [0, 0], [1, 1], 0.1, jitter=0.25)
    pts = nd["points"]
    d2, idx = o.knn(pts, pts, 12)
    W, _ = o.divergence_weights(pts, pts, idx, d2, 1)
    n = len(pts)
    v = np.array([0.3, -0.2])
    rng = np.random.default_rng(5)
    f0 = rng.standard_normal(n)
    dt, off = 0.02, 0.25
    P = np.eye(n)
    if end:
        P[end:, end:] = 0.0
    M = dt * P @ _advection_matrix(idx, W, v)
    g = f0 - off
    want, term = g.copy(), g.copy()
    for m in range(1, 5):
        term = M @ term / m
        want += term
    want += off
    got = o.advection_step(idx, W, v, f0, dt, off, boundary_start=end)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    if end:
        assert np.array_equal(got[end:], f0[end:])
    # two steps are two applications of the same polynomial
    got2 = o.advection_step(idx, W, v, f0, dt, off, boundary_start=end, steps=2)
    one = o.advection_step(idx, W, v, got, dt, off, boundary_start=end)
    assert np.abs(got2 - one).max() < 1e-12 * np.abs(one).max()


def test_advection_translates_a_pulse(oracle64):
    """df/dt = -v . grad f carries a smooth pulse with velocity v."""
    o = oracle64
    nd = o.uniform_grid(2, [0, 0], [1, 1], 0.02, jitter=0.25)
    pts = nd["points"].astype(np.float64)
    d2, idx = o.knn(nd["points"], nd["points"], 16)
    W, _ = o.divergence_weights(nd["points"], nd["points"], idx, d2, 1)
    v = np.array([0.5, 0.25])
    width, c0 = 0.08, np.array([0.35, 0.4])
    f0 = np.exp(-np.sum((pts - c0) ** 2, 1) / width ** 2)
    T, dt = 0.4, 0.01
    f = o.advection_step(idx, W, v, f0, dt, 0.0, steps=int(round(T / dt)))
    c1 = c0 + v * T
    want = np.exp(-np.sum((pts - c1) ** 2, 1) / width ** 2)
    inside = np.all((pts > 0.15) & (pts < 0.85), axis=1)
    err = np.linalg.norm((f - want)[inside]) / np.linalg.norm(want[inside])
    assert err < 0.06
    com = (f[inside, None] * pts[inside]).sum(0) / f[inside].sum()
    assert np.abs(com - c1).max() < 0.01


# -------------------------------------------------------------------- CSR
def test_csr_from_stencil_vs_scipy(oracle64):
    """create_from_index_stencil keeps stencil order (duplicate columns
    included) and its product equals SciPy's CSR product."""
    import scipy.sparse as sp
    rng = np.random.default_rng(9)
    n, k, cols = 200, 7, 150
    idx = rng.integers(0, cols, (n, k)).astype(np.uint32)
    idx[3, :3] = 11  # duplicate columns in one row
    vals = rng.standard_normal((n, k))
    rp, ci, cv = oracle64.csr_from_stencil(idx, vals)
    assert np.array_equal(rp, np.arange(n + 1) * k)
    assert np.array_equal(ci, idx.reshape(-1).astype(np.int64))
    A = sp.csr_matrix((vals.reshape(-1), idx.reshape(-1).astype(np.int64), rp),
                      shape=(n, cols))
    x = rng.standard_normal(cols)
    y = rng.standard_normal(n)
    got = oracle64.csr_mult(rp, ci, cv, x, alpha=0.5, beta=-2.0, y=y)
    assert np.allclose(got, 0.5 * (A @ x) - 2.0 * y, rtol=1e-12, atol=1e-12)
