"""Pins of the oracle's PML terms and of the D2Q9 collision's third- and
fourth-order moments for u != 0 -- by closed forms, exact responses and
invariants, each chosen so that a plausible slip (a coefficient 2 sigma Q1
instead of 3 sigma Q1, Q2 formed from the new Q1 instead of the stored one,
layer means over all nodes, a field read at j instead of j - num_bulk, a
dropped or sign-flipped substitution term) fails one of them.  CPU only."""
import itertools

import numpy as np
import pytest

from tests.test_oracle import BASIS9, CS2

CS = np.sqrt(CS2)


def _all_layer_nodes(D, h, sigma, seed=7):
    """A jittered grid whose every node is an absorbing-layer node with the
    same sigma, and no boundary nodes: a spatially uniform state stays
    uniform (divergences of uniform fields vanish to rounding), so one step
    reduces to the per-node recurrences."""
    n = int(round(1.0 / h)) + 1
    g = np.array(list(itertools.product(range(n), repeat=D)), np.float64) * h
    rng = np.random.default_rng(seed)
    pts = (g + rng.uniform(-0.25 * h, 0.25 * h, g.shape)).astype(np.float32)
    N = len(pts)
    return dict(points=pts, normals=np.zeros((0, D), np.float32),
                sigma=np.full(N, sigma, np.float32), n_bulk=0, n_layer=N, n_boundary=0)


@pytest.mark.parametrize("Q", [9, 27])
def test_pml_uniform_layer_recurrences(oracle64, Q):
    """compute_Qvalues / apply / post_collision (pml_absorption.inl:551-631,
    734-926) on a uniform state inside a uniform layer: with f~eq = feq - w,
    Q1 = (f~eq + (Q1s - w)) dt/2, Q2 = (Q1 + Q2s) dt/2 (one-step trapezoids,
    Q2 from the STORED previous Q1), s1 = f - sigma (3 f~eq + 3 sigma Q1 +
    sigma^2 Q2) (no dt factor), then Q2s <- Q1, Q1s <- feq; the collision
    (pinned separately) adds its increment from the pre-PML f.  Three steps
    from a non-equilibrium state with u != 0."""
    D = 2 if Q == 9 else 3
    h = 0.1 if D == 2 else 0.25
    k = 16
    sig = float(np.float32(0.37))  # the layer absorption is FP32 input
    tau = 0.6
    dt = 0.4 * h
    nd = _all_layer_nodes(D, h, sig)
    N = nd["n_layer"]
    e = oracle64.engine(nd, Q=Q, k=k, h=h, tau=tau)
    c, w, _, _ = oracle64.lattice(Q)
    rng = np.random.default_rng(11 + Q)
    f = w * (1.0 + 2e-2 * rng.standard_normal(Q))
    e.set_distributions(np.repeat(f[:, None], N, axis=1))
    nu_lat = (tau - 0.5) / 3.0 if Q == 27 else (tau - 0.5) * dt / 3.0
    Q1s, Q2s = w.copy(), np.zeros(Q)  # the operator's reset state
    for _ in range(3):
        rho = f.sum()
        u = f @ c / rho
        fte = np.array([oracle64.feq(Q, rho, u, a) for a in range(Q)]) - w
        q1 = (fte + (Q1s - w)) * dt / 2
        q2 = (q1 + Q2s) * dt / 2
        s1 = f - sig * (3 * fte + 3 * sig * q1 + sig * sig * q2)
        Q2s, Q1s = q1, fte + w
        f = s1 + oracle64.collide(Q, f, rho, u, nu_lat, dt)
        e.step(1)
        got = e.get()["f"]
        scale = np.abs(f - w).max()
        assert np.abs(got - f[:, None]).max() < 1e-11 * scale
        g1, g2 = e.pml()
        assert np.abs(g1 - Q1s[:, None]).max() < 1e-13
        assert np.abs(g2 - Q2s[:, None]).max() < 1e-11 * np.abs(Q2s).max()
    # a state that the damping actually moved (sigma terms are not negligible)
    assert np.abs(sig * 3 * fte).max() > 1e-3 * np.abs(f - w).max()


def _layer_scene(oracle, ref, Q=9):
    D = 2 if Q == 9 else 3
    h = 0.05 if D == 2 else 0.125
    nd = oracle.uniform_grid(D, [0] * D, [1] * D, h, jitter=0.25, layer=2 * h, sigma_max=0.5)
    e = oracle.engine(nd, Q=Q, k=16, h=h, pml_ref_index=ref)
    c, w, _, _ = oracle.lattice(Q)
    N = e.N
    e.set_distributions(np.repeat(w[:, None], N, axis=1))  # rest: feq - w = 0
    return nd, e, c, w, 0.4 * h


@pytest.mark.parametrize("Q", [9, 27])
@pytest.mark.parametrize("ref", [1, 0])
def test_pml_field_delta_response(oracle64, ref, Q):
    """The auxiliary fields (pml_div_Q1_field_map / pml_div_Q2_field_map,
    pml_absorption.inl:411-529) from one non-zero history entry at a bulk
    node p0: Q1 = dQ dt/2 there, Q2 = Q1 dt/2, every layer mean zero.  With
    the reference's indexing the layer node j0 = p0 + num_bulk reads that
    entry (:445-448), with sigma[j0 - num_bulk], so the stage output is the
    divergence of a single source,
      s1[a0, i] = sum_{s: idx(s,i) = j0} sum_d W(d,s,i) c_{a0,d}
                  (-2 sigma Q1 + sigma^2 Q2),
    and exactly zero for every other (alpha, node).  Read at j instead, the
    bulk entry is never read and the output is exactly zero."""
    nd, e, c, w, dt = _layer_scene(oracle64, ref, Q)
    N, nb = e.N, nd["n_bulk"]
    p0, a0, dq = 3, Q - 2, 0.25
    assert p0 < min(nb, nd["n_layer"])
    q1s = np.repeat(w[:, None], N, axis=1)
    q1s[a0, p0] += dq
    e.set_pml(q1s, np.zeros((Q, N)))
    out = e.pml_stage(np.zeros((Q, N)))
    st = e.setup()
    idx, W = st["idx"], st["W"]
    j0 = p0 + nb
    sig = float(nd["sigma"][p0])
    q1 = dq * dt / 2
    q2 = q1 * dt / 2
    F = c[a0] * (-2.0 * sig * q1 + sig * sig * q2)
    hit = idx == j0
    expect = np.einsum("isd,d->is", W, F)
    expect = np.where(hit, expect, 0.0).sum(axis=1)
    assert hit.any(axis=1).sum() >= 8
    # "zero" is the rest state's rounding: sum w and u are exact to 1e-16
    tiny = 1e-13 * np.abs(expect).max()
    if ref == 0:
        assert np.abs(out).max() < tiny
        return
    mask = np.ones((Q, N), bool)
    mask[a0] = False
    assert np.abs(out[mask]).max() < tiny
    assert np.abs(out[a0][~hit.any(axis=1)]).max() < tiny
    assert np.abs(out[a0] - expect).max() < 1e-13 * np.abs(expect).max()


@pytest.mark.parametrize("Q", [9, 27])
def test_pml_layer_means_cancel_a_layer_constant(oracle64, Q):
    """History constant over the layer (per alpha), other constants outside
    it, field read at j: the layer means (pml_absorption.inl:775-800, over the layer
    only) equal the layer values, so both auxiliary fields vanish and the
    stage output is the local damping alone,
      s1 = -sigma (3 sigma Q1 + sigma^2 Q2) on layer nodes, 0 elsewhere."""
    nd, e, c, w, dt = _layer_scene(oracle64, 0, Q)
    N, nb, nl = e.N, nd["n_bulk"], nd["n_layer"]
    C = 0.05 * (1 + np.arange(Q)) / Q
    q1s = np.repeat(w[:, None], N, axis=1)
    q1s[:, nb:nb + nl] += C[:, None]
    q1s[:, :nb] -= 2 * C[:, None]       # bulk and boundary history: read by no
    q1s[:, nb + nl:] += 3 * C[:, None]  # term here, but by means over all nodes
    e.set_pml(q1s, np.zeros((Q, N)))
    out = e.pml_stage(np.zeros((Q, N)))
    sig = nd["sigma"].astype(np.float64)
    q1 = C * dt / 2
    q2 = q1 * dt / 2
    expect = np.zeros((Q, N))
    expect[:, nb:nb + nl] = -sig[None, :] * (3 * sig[None, :] * q1[:, None]
                                             + sig[None, :] ** 2 * q2[:, None])
    assert np.abs(out - expect).max() < 1e-10 * np.abs(expect).max()
    # means taken over all nodes would leave a field of size sigma |W| Q1:
    # far above the tolerance
    assert np.abs(expect).max() > 1e-6


def test_pml_zero_sigma_is_no_layer(oracle64):
    """sigma = 0 in the layer is bit for bit the same run as no layer at all
    (the same nodes counted as bulk): every PML term carries a factor
    sigma."""
    h = 0.05
    nd = oracle64.uniform_grid(2, [0, 0], [1, 1], h, jitter=0.25, layer=2 * h, sigma_max=0.5)
    z = dict(nd, sigma=np.zeros_like(nd["sigma"]))
    flat = dict(nd, sigma=np.zeros(0, np.float32), n_bulk=nd["n_bulk"] + nd["n_layer"],
                n_layer=0)
    runs = []
    for nodes in (z, flat):
        e = oracle64.engine(nodes, Q=9, k=16, h=h)
        e.set_pulse([0.35, 0.5], 1e-3, 0.08)
        e.step(15)
        runs.append(e.get()["f"])
    assert np.array_equal(runs[0], runs[1])


def test_pml_absorbs_an_outgoing_pulse(oracle64):
    """A pulse leaving the interior is absorbed by the layer: the acoustic
    energy left inside after the front has reached the walls and come back
    falls as sigma_max grows (with sigma = 0 the bounce-back walls reflect
    it).  A wrong sign of a damping or field term amplifies instead."""
    h = 0.025
    energy = []
    for smax in (0.0, 0.25, 0.5):
        nd = oracle64.uniform_grid(2, [0, 0], [1, 1], h, jitter=0.25, layer=0.2,
                                   sigma_max=smax)
        e = oracle64.engine(nd, Q=9, k=16, h=h, threads=8)
        e.set_pulse([0.5, 0.5], 1e-3, 0.05)
        p = nd["points"].astype(np.float64)
        inner = np.abs(p - 0.5).max(axis=1) < 0.3
        e.step(int(round(1.0 / CS / (0.4 * h))))  # front out to the walls and back
        s = e.get()
        energy.append(float((s["drho"][inner] ** 2).sum() + (s["u"][:, inner] ** 2).sum()))
    assert energy[1] < 0.6 * energy[0] and energy[2] < energy[1], energy


def _d2q9_forward_substitution(c, K, f, rho, u, k, dt):
    """Third- and fourth-order moments of the D2Q9 cascade with full
    relaxation (rate 1/dt): the post-collision central moment j of
    f + dt K k equals the equilibrium one, solved by forward substitution on
    B(u)_{jl} = sum_a P_j(c_a - u) K_al with the given lower k."""
    x = c - u
    Ps = np.array([[P(*row) for P in BASIS9] for row in x])
    B = Ps.T @ K
    C = f @ Ps
    mons = list(itertools.product(range(3), repeat=2))
    V = np.array([[np.prod([row[d] ** m[d] for d in range(2)]) for m in mons] for row in c])
    coef = np.linalg.solve(V, K)
    gm = {0: 1.0, 1: 0.0, 2: CS2}
    e = np.array([sum(coef[i, j] * np.prod([gm[m[d]] for d in range(2)])
                      for i, m in enumerate(mons)) for j in range(9)])
    assert np.abs(np.triu(B, 1)).max() < 1e-14  # B(u) is lower triangular
    out = {}
    for j in (6, 7, 8):
        out[j] = ((rho * e[j] - C[j]) / dt - sum(B[j, l] * k[l] for l in range(j))) / B[j, j]
    return out


def test_collision_d2q9_higher_moments(oracle64):
    """D2Q9 k6-k8 for u != 0 (detail/d2q9_lattice.cuh:207-299).  Each is the
    cascade's full relaxation solved by forward substitution with the
    reference's own lower k (its k5 sign quirk included), plus the terms by
    which the reference's closed forms differ from that substitution:
      k6: + (3 u_y k3 + u_y k4 - 4 u_x k5)   (term3 of compute_k6, :227-228,
          enters with the opposite sign),
      k7: + (3 u_x k3 - u_x k4 - 4 u_y k5)   (the same in compute_k7,
          :252-253),
      k8: + u_y^2 k4                          (compute_k8 multiplies k4/2 by
          u_x^2 + u_y^2, :292-295, where the substitution has u_x^2 - u_y^2).
    At u = 0 every extra term vanishes and the projection is the exact
    cascade (test_collision_d2q9)."""
    c, w, _, K = oracle64.lattice(9)
    rng = np.random.default_rng(21)
    for nu, dt in ((0.01, 0.3), (0.002, 0.004)):
        for _ in range(6):
            f = w * (1 + 0.05 * rng.standard_normal(9))
            rho = f.sum()
            u = f @ c / rho
            k = np.linalg.solve(K, oracle64.collide(9, f, rho, u, nu, dt) / dt)
            fs = _d2q9_forward_substitution(c, K, f, rho, u, k, dt)
            ux, uy = u
            extra = {6: 3 * uy * k[3] + uy * k[4] - 4 * ux * k[5],
                     7: 3 * ux * k[3] - ux * k[4] - 4 * uy * k[5],
                     8: uy * uy * k[4]}
            for j in (6, 7, 8):
                assert abs(k[j] - (fs[j] + extra[j])) < 1e-11 * np.abs(k).max(), j
            # the extra terms are not negligible: the substitution alone fails
            assert max(abs(k[j] - fs[j]) for j in (6, 7)) > 1e-4 * np.abs(k[6:8]).max()
