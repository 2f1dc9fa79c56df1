"""Derive and emit the D3Q27 cascaded central-moment collision.

The collision of detail/d3q27_lattice.cuh:125-378 is characterised here from
first principles instead of being transcribed:

* the increment is  df = dt * K k,  where column j of K is a lattice
  polynomial P_j evaluated at the 27 velocities (listed in ``BASIS``; the
  reference's integer table is exactly these values);
* the cascaded moments of f about the local velocity u are
  C_j = sum_a f_a P_j(c_a - u), with the zeroth central moment taken as rho
  and the first central moments as zero (the reference's formulas contain
  neither);
* k_0..k_3 = 0 (mass and momentum are conserved) and the remaining k solve
  B(u) k' = rho e - C with k_j = omega_j k'_j, where
  B(u)_{jl} = sum_a P_l(c_a) P_j(c_a - u) is lower triangular with a
  constant diagonal, and e_j is P_j projected on the central moments of the
  unit-density Maxwellian (c_s^2 = 1/3).

The emitted header computes k' by forward substitution on B(u) (polynomial
entries, common subexpressions shared) from the 27 raw central moments of f
about u, which the caller forms with a dimension-by-dimension transform.
All arithmetic is FP64.

Run:  python naga_amd/csrc/gen_collision.py   (rewrites cascade_d3q27.inc)
"""
from __future__ import annotations

import itertools
import os

import sympy as sp
from sympy.printing.c import C99CodePrinter

HERE = os.path.dirname(os.path.abspath(__file__))

# Velocity ordering of the D3Q27 lattice; it is part of the interface
# (distributions are exchanged per alpha), detail/d3q27_lattice.cuh:45-55.
VEL = [
    (0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1),
    (0, 0, -1), (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1),
    (-1, 0, 1), (1, 0, -1), (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1),
    (0, -1, -1), (1, 1, 1), (-1, 1, 1), (1, -1, 1), (-1, -1, 1), (1, 1, -1),
    (-1, 1, -1), (1, -1, -1), (-1, -1, -1),
]

X, Y, Z = sp.symbols("x y z")


def q(t):
    """One-dimensional second-order lattice polynomial 3t^2 - 2."""
    return 3 * t**2 - 2


# Cascade basis, in k-index order: conserved (0-3), second order (4-9),
# third order (10-16), fourth (17-22), fifth (23-25), sixth (26).
BASIS = [
    sp.Integer(1), X, Y, Z,
    X * Y, X * Z, Y * Z,
    X**2 - Y**2,
    X**2 + Y**2 - 2 * Z**2,
    X**2 + Y**2 + Z**2 - 2,
    X * (3 * (Y**2 + Z**2) - 4),
    Y * (3 * (X**2 + Z**2) - 4),
    Z * (3 * (X**2 + Y**2) - 4),
    X * (Y**2 - Z**2),
    Y * (X**2 - Z**2),
    Z * (X**2 - Y**2),
    X * Y * Z,
    (q(X) * q(Y) + q(X) * q(Z) + q(Y) * q(Z)) / 3,
    (q(X) * q(Y) + q(X) * q(Z) - 2 * q(Y) * q(Z)) / 3,
    q(X) * (Y**2 - Z**2),
    Y * Z * q(X),
    X * Z * q(Y),
    X * Y * q(Z),
    X * q(Y) * q(Z),
    Y * q(X) * q(Z),
    Z * q(X) * q(Y),
    q(X) * q(Y) * q(Z),
]

MONS = list(itertools.product(range(3), repeat=3))


def collision_matrix():
    """K[a][j] = P_j(c_a) as integers."""
    K = []
    for (cx, cy, cz) in VEL:
        row = []
        for P in BASIS:
            v = sp.nsimplify(P.subs({X: cx, Y: cy, Z: cz}))
            assert v == int(v)
            row.append(int(v))
        K.append(row)
    return K


def poly_coeffs(P):
    """Coefficients of P over the monomials x^a y^b z^c, a,b,c <= 2."""
    poly = sp.Poly(sp.expand(P), X, Y, Z)
    out = {}
    for (a, b, c), v in poly.terms():
        assert a <= 2 and b <= 2 and c <= 2
        out[(a, b, c)] = sp.Rational(v)
    return out


def derive():
    K = collision_matrix()
    Q = len(VEL)
    ux, uy, uz = sp.symbols("ux uy uz")
    # B(u)_{jl} = sum_a P_l(c_a) P_j(c_a - u)
    shifted = [[sp.expand(P.subs({X: cx - ux, Y: cy - uy, Z: cz - uz},
                                 simultaneous=True)) for (cx, cy, cz) in VEL]
               for P in BASIS]
    B = [[sp.expand(sum(K[a][l] * shifted[j][a] for a in range(Q)))
          for l in range(Q)] for j in range(Q)]
    R = range(4, Q)
    for j in R:
        assert B[j][j].free_symbols == set(), "diagonal must be constant"
        for l in R:
            if l > j:
                assert B[j][l] == 0, "cascade must be lower triangular"
    # unit-density Maxwellian central moments: <1>=1, <t>=0, <t^2>=1/3
    meq = {0: sp.Integer(1), 1: sp.Integer(0), 2: sp.Rational(1, 3)}
    coeffs = [poly_coeffs(P) for P in BASIS]
    e = [sum(v * meq[a] * meq[b] * meq[c] for (a, b, c), v in cf.items())
         for cf in coeffs]
    return K, B, e, coeffs, (ux, uy, uz)


class _Printer(C99CodePrinter):
    """Integer powers as products (no pow() calls in device code)."""

    def _print_Pow(self, expr):
        b, ex = expr.as_base_exp()
        if ex.is_Integer and 1 < int(ex) <= 8:
            base = self.parenthesize(b, 100)
            return "*".join([base] * int(ex))
        return super()._print_Pow(expr)


def ccode(ex):
    return _Printer().doprint(ex)


def emit(path):
    K, B, e, coeffs, (ux, uy, uz) = derive()
    Q = len(VEL)
    R = list(range(4, Q))
    m = {mon: sp.Symbol(f"m[{mon[0]}][{mon[1]}][{mon[2]}]") for mon in MONS}
    rho = sp.Symbol("rho")
    kp = {j: sp.Symbol(f"kp{j}") for j in R}
    # off-diagonal polynomial entries, shared subexpressions extracted once
    offd = [(j, l, B[j][l]) for j in R for l in R if l < j and B[j][l] != 0]
    repl, red = sp.cse([t for (_, _, t) in offd], optimizations="basic")
    lines = []
    w = lines.append
    w("// Generated by gen_collision.py -- do not edit by hand.")
    w("// D3Q27 cascaded central-moment collision in FP64: raw central")
    w("// moments m[a][b][c] of f about u in, increment df = dt * K * k out.")
    w("// omega[j] are the relaxation rates of cascade index j (4..26).")
    w("NAGA_HD inline void cascade_d3q27(const double (&m)[3][3][3], double rho,")
    w("                                  double ux, double uy, double uz,")
    w("                                  const double* omega, double dt,")
    w("                                  double* df) {")
    for (s, ex) in repl:
        w(f"    const double {s} = {ccode(ex)};")
    offv = {}
    for (j, l, _), ex in zip(offd, red):
        name = f"b{j}_{l}"
        offv[(j, l)] = sp.Symbol(name)
        w(f"    const double {name} = {ccode(ex)};")
    # The reference's closed forms contain no zeroth or first central moment:
    # they are the cascade with m000 = rho and m100 = m010 = m001 = 0, which
    # differs from the general solve when rho != sum f or u != j / sum f
    # (a density source changes rho without changing u).
    subs = {m[(0, 0, 0)]: rho, m[(1, 0, 0)]: 0, m[(0, 1, 0)]: 0, m[(0, 0, 1)]: 0}
    for j in R:
        rhs = (rho * e[j] - sum(v * m[mon] for mon, v in coeffs[j].items())).subs(subs)
        acc = sp.expand(rhs) - sum(offv[(j, l)] * kp[l]
                                   for l in R if (j, l) in offv)
        # multiply by the reciprocal (a compile-time constant): FP64 division
        # is a long instruction sequence on the GPU
        w(f"    const double kp{j} = ({ccode(acc)}) * (1.0 / {ccode(B[j][j])}.0);")
    w("    double k[27];")
    w("    k[0] = k[1] = k[2] = k[3] = 0.0;")
    for j in R:
        w(f"    k[{j}] = omega[{j}] * kp{j};")
    for a in range(Q):
        terms = [(K[a][j], j) for j in R if K[a][j] != 0]
        expr = " + ".join(f"({c}.0) * k[{j}]" for c, j in terms) or "0.0"
        w(f"    df[{a}] = dt * ({expr});")
    w("}")
    w("")
    w("// K[a][j] = P_j(c_a): the collision matrix of the cascade basis.")
    w("static constexpr int kCascadeK27[27][27] = {")
    for row in K:
        w("    {" + ", ".join(str(v) for v in row) + "},")
    w("};")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    emit(os.path.join(HERE, "cascade_d3q27.inc"))
    print("wrote", os.path.join(HERE, "cascade_d3q27.inc"))
