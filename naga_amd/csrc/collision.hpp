// Cascaded central-moment collision, evaluated in FP64 per node.
//
// D3Q27: the cascade of detail/d3q27_lattice.cuh:125-378, derived from its
// defining properties by gen_collision.py (cascade_d3q27.inc).
// D2Q9: the projection of detail/d2q9_lattice.cuh:120-299 as the reference
// writes it (the engine never reached it at HEAD: the call at
// detail/simulation_engine.cuh:509-516 passes six arguments to a
// seven-argument function; the unused characteristic frequency is dropped).
//
// Both take the full distribution f (= w + g) in FP64: the stored state is
// the FP32 deviation g, and the moments are formed in FP64 so that the
// O(1) equilibrium part cancels exactly instead of at FP32 resolution.
#pragma once

#include "lattice.hpp"

namespace naga_amd {

#include "cascade_d3q27.inc"

struct CollisionParams {
    double omega[27];  // D3Q27 cascade rates (1/time), index = k index
    double om_ab;      // D2Q9 shear rate 1/(3 nu + dt/2)
    double inv_dt;     // D2Q9 higher-order rate 1/dt
    double dt;         // lattice time step
};

// unit-time rate -> rate at dt, detail/d3q27_lattice.cuh:200-204
inline double cascade_rate(double unit_rate, double dt) {
    double unit_nu = (1.0 / unit_rate - 0.5) / 3.0;
    return 1.0 / (3.0 * unit_nu + 0.5) / dt;
}

// Relaxation rates: shear/bulk from the viscosity; higher orders from the
// fixed unit-time rates of the reference (FP32 literals there, so their FP32
// values are used: 1.5f, 1.83f, 1.4f, 1.61f, 1.98f, 1.74f).
inline CollisionParams make_collision_params(int Q, double nu_lat, double dt_lat) {
    CollisionParams p{};
    p.dt = dt_lat;
    if (Q == 27) {
        const double shear = 1.0 / (3.0 * nu_lat + 0.5) / dt_lat;
        for (int j = 4; j <= 9; ++j) p.omega[j] = shear;
        const struct {
            int lo, hi;
            float unit;
        } groups[] = {{10, 12, 1.5f},  {13, 15, 1.83f}, {16, 16, 1.4f},
                      {17, 17, 1.61f}, {18, 22, 1.98f}, {23, 26, 1.74f}};
        for (const auto& g : groups)
            for (int j = g.lo; j <= g.hi; ++j)
                p.omega[j] = cascade_rate(double(g.unit), dt_lat);
    } else {
        p.om_ab = 1.0 / (3.0 * nu_lat + 0.5 * dt_lat);
        p.inv_dt = 1.0 / dt_lat;
    }
    return p;
}

// Raw central moments m[a][b][c] = sum_alpha f (cx-ux)^a (cy-uy)^b (cz-uz)^c
// by one sweep per dimension over the 3x3x3 velocity cube.
NAGA_HD inline void central_moments27(const double* f, double ux, double uy,
                                      double uz, double (&m)[3][3][3]) {
    using L = Lattice<27>;
    double cube[3][3][3];
#pragma unroll
    for (int a = 0; a < 27; ++a)
        cube[L::c(a, 0) + 1][L::c(a, 1) + 1][L::c(a, 2) + 1] = f[a];
    // x sweep: t[p][iy][iz]
    double t[3][3][3];
#pragma unroll
    for (int iy = 0; iy < 3; ++iy)
#pragma unroll
        for (int iz = 0; iz < 3; ++iz) {
            double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
            for (int ix = 0; ix < 3; ++ix) {
                double d = double(ix - 1) - ux, v = cube[ix][iy][iz];
                s0 += v;
                s1 += v * d;
                s2 += v * d * d;
            }
            t[0][iy][iz] = s0;
            t[1][iy][iz] = s1;
            t[2][iy][iz] = s2;
        }
    double r[3][3][3];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int iz = 0; iz < 3; ++iz) {
            double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
            for (int iy = 0; iy < 3; ++iy) {
                double d = double(iy - 1) - uy, v = t[p][iy][iz];
                s0 += v;
                s1 += v * d;
                s2 += v * d * d;
            }
            r[p][0][iz] = s0;
            r[p][1][iz] = s1;
            r[p][2][iz] = s2;
        }
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
            for (int iz = 0; iz < 3; ++iz) {
                double d = double(iz - 1) - uz, v = r[p][q][iz];
                s0 += v;
                s1 += v * d;
                s2 += v * d * d;
            }
            m[p][q][0] = s0;
            m[p][q][1] = s1;
            m[p][q][2] = s2;
        }
}

// D2Q9 projection (detail/d2q9_lattice.cuh:165-299), written over named
// groups of populations.  Alpha order: rest, nw, w, sw, s, se, e, ne, n.
NAGA_HD inline void project_d2q9(const double* f, double rho, double ux, double uy,
                                 const CollisionParams& p, double* k) {
    const double r0 = f[0], nw = f[1], we = f[2], sw = f[3], so = f[4];
    const double se = f[5], ea = f[6], ne = f[7], no = f[8];
    const double uxx = ux * ux, uyy = uy * uy;
    const double axis = ea + no + so + we;
    const double diag = se + sw + ne + nw;
    k[0] = k[1] = k[2] = 0.0;
    k[3] = p.inv_dt * (rho * (uxx + uyy) - axis - 2.0 * (diag - rho / 3.0)) / 12.0;
    k[4] = p.om_ab * (no + so - ea - we + rho * (uxx - uyy)) / 4.0;
    k[5] = p.om_ab * (ne + sw - nw - se + rho * (ux * uy)) / 4.0;
    {
        const double a = (se + sw - ne - nw - 2.0 * uxx * uy * rho
                          + uy * (rho - no - so - r0)) / 4.0;
        const double b = ux / 2.0 * (ne - nw - se + sw);
        const double c = -uy / 2.0 * (-3.0 * k[3] - k[4]) - 2.0 * ux * k[5];
        k[6] = -p.inv_dt * (a + b) + c;
    }
    {
        const double a = (sw + nw - se - ne - 2.0 * ux * uyy * rho
                          + ux * (rho - we - ea - r0)) / 4.0;
        const double b = uy / 2.0 * (ne + sw - se - nw);
        const double c = -ux / 2.0 * (-3.0 * k[3] + k[4]) - 2.0 * uy * k[5];
        k[7] = -p.inv_dt * (a + b) + c;
    }
    {
        const double t1 = rho / 9.0 - diag;
        const double t2 = 2.0 * (ux * (ne - nw + se - sw) + uy * (ne + nw - se - sw));
        const double t3 = 4.0 * ux * uy * (nw - ne + se - sw);
        const double t4 = -uxx * (no + ne + nw + so + se + sw);
        const double t5 = uyy * (3.0 * uxx * rho - ea - ne - nw - se - sw - we);
        const double t6 = -2.0 * k[3] - 2.0 * ux * k[7] - 2.0 * uy * k[6]
                          + 4.0 * ux * uy * k[5]
                          - (1.5 * k[3] - k[4] / 2.0) * (uxx + uyy);
        k[8] = p.inv_dt * (0.25 * (t1 + t2 + t3 + t4 + t5)) + t6;
    }
}

// D2Q9 collision-matrix column j at velocity (x, y): the lattice
// polynomials 1, x, y, q(x)+q(y), x^2-y^2, -xy, -y q(x), -x q(y), q(x)q(y)
// with q(t) = 3t^2 - 2 (their lattice values form the reference table,
// detail/d2q9_lattice.cuh:72-82).
NAGA_HD constexpr int k9_column(int j, int x, int y) {
    return j == 0   ? 1
           : j == 1 ? x
           : j == 2 ? y
           : j == 3 ? (3 * x * x - 2) + (3 * y * y - 2)
           : j == 4 ? x * x - y * y
           : j == 5 ? -x * y
           : j == 6 ? -y * (3 * x * x - 2)
           : j == 7 ? -x * (3 * y * y - 2)
                    : (3 * x * x - 2) * (3 * y * y - 2);
}

// df = collision increment for one node (simulation_engine.cuh:469-531):
// f full distribution, rho = sum f, u unitless velocity.
template <int Q>
NAGA_HD inline void collide(const double* f, double rho, const double* u,
                            const CollisionParams& p, double* df) {
    if constexpr (Q == 27) {
        double m[3][3][3];
        central_moments27(f, u[0], u[1], u[2], m);
        cascade_d3q27(m, rho, u[0], u[1], u[2], p.omega, p.dt, df);
    } else {
        using L = Lattice<9>;
        double k[9];
        project_d2q9(f, rho, u[0], u[1], p, k);
#pragma unroll
        for (int a = 0; a < 9; ++a) {
            double acc = 0.0;
#pragma unroll
            for (int j = 3; j < 9; ++j)
                acc += double(k9_column(j, L::c(a, 0), L::c(a, 1))) * k[j];
            df[a] = acc * p.dt;
        }
    }
}

// Collision matrix of either lattice, K[a][j] (for the C-ABI table query).
inline int collision_matrix_entry(int Q, int a, int j) {
    if (Q == 27) return kCascadeK27[a][j];
    return k9_column(j, Lattice<9>::c(a, 0), Lattice<9>::c(a, 1));
}

}  // namespace naga_amd
