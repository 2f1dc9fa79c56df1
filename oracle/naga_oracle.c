/*
 * naga_oracle.c -- CPU restatement of NAGA's nonlocal lattice-Boltzmann
 * acoustics step and its setup.  TEST INFRASTRUCTURE: the checker and the
 * CPU baseline, never part of the product path.  See naga_oracle.h for the
 * build recipe (compiled once per precision) and the citation convention.
 *
 * Build flags: -O2 -ffp-contract=off -fopenmp (no fast-math) so every FP32
 * operation is one correctly rounded IEEE operation in program order.
 */
#include "naga_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* lattices.cuh:46-49 */
#define CS_LAT 0.57735026918962576450914878050196

/* ======================================================================
 * Lattice tables.
 * D2Q9: detail/d2q9_lattice.cuh:44-88 (velocities, weights, K),
 *       :151-163 (bounce-back).
 * D3Q27: detail/d3q27_lattice.cuh:45-98 (velocities, weights, K),
 *       :380-387 (bounce-back).
 * The alpha ordering is part of the interface (distributions are exchanged
 * per alpha), so it is kept as in the reference.
 * ====================================================================== */
static const int C9[9][2] = {{0, 0},  {-1, 1}, {-1, 0}, {-1, -1}, {0, -1},
                             {1, -1}, {1, 0},  {1, 1},  {0, 1}};
static const double W9[9] = {4.0 / 9.0,  1.0 / 36.0, 1.0 / 9.0,
                             1.0 / 36.0, 1.0 / 9.0,  1.0 / 36.0,
                             1.0 / 9.0,  1.0 / 36.0, 1.0 / 9.0};
static const int BB9[9] = {0, 5, 6, 7, 8, 1, 2, 3, 4};
static const int K9[9][9] = {{1, 0, 0, -4, 0, 0, 0, 0, 4},
                             {1, -1, 1, 2, 0, 1, -1, 1, 1},
                             {1, -1, 0, -1, 1, 0, 0, -2, -2},
                             {1, -1, -1, 2, 0, -1, 1, 1, 1},
                             {1, 0, -1, -1, -1, 0, -2, 0, -2},
                             {1, 1, -1, 2, 0, 1, 1, -1, 1},
                             {1, 1, 0, -1, 1, 0, 0, 2, -2},
                             {1, 1, 1, 2, 0, -1, -1, -1, 1},
                             {1, 0, 1, -1, -1, 0, 2, 0, -2}};

static const int C27[27][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
    {0, 0, 1},   {0, 0, -1},  {1, 1, 0},  {-1, 1, 0},  {1, -1, 0},
    {-1, -1, 0}, {1, 0, 1},   {-1, 0, 1}, {1, 0, -1},  {-1, 0, -1},
    {0, 1, 1},   {0, -1, 1},  {0, 1, -1}, {0, -1, -1}, {1, 1, 1},
    {-1, 1, 1},  {1, -1, 1},  {-1, -1, 1}, {1, 1, -1}, {-1, 1, -1},
    {1, -1, -1}, {-1, -1, -1}};
static const int BB27[27] = {0,  2,  1,  4,  3,  6,  5,  10, 9,
                             8,  7,  14, 13, 12, 11, 18, 17, 16,
                             15, 26, 25, 24, 23, 22, 21, 20, 19};
static const int K27[27][27] = {
    {1, 0, 0, 0, 0, 0, 0, 0, 0, -2, 0, 0, 0, 0, 0, 0, 0, 4, 0, 0, 0, 0, 0, 0, 0, 0, -8},
    {1, 1, 0, 0, 0, 0, 0, 1, 1, -1, -4, 0, 0, 0, 0, 0, 0, 0, -4, 0, 0, 0, 0, 4, 0, 0, 4},
    {1, -1, 0, 0, 0, 0, 0, 1, 1, -1, 4, 0, 0, 0, 0, 0, 0, 0, -4, 0, 0, 0, 0, -4, 0, 0, 4},
    {1, 0, 1, 0, 0, 0, 0, -1, 1, -1, 0, -4, 0, 0, 0, 0, 0, 0, 2, -2, 0, 0, 0, 0, 4, 0, 4},
    {1, 0, -1, 0, 0, 0, 0, -1, 1, -1, 0, 4, 0, 0, 0, 0, 0, 0, 2, -2, 0, 0, 0, 0, -4, 0, 4},
    {1, 0, 0, 1, 0, 0, 0, 0, -2, -1, 0, 0, -4, 0, 0, 0, 0, 0, 2, 2, 0, 0, 0, 0, 0, 4, 4},
    {1, 0, 0, -1, 0, 0, 0, 0, -2, -1, 0, 0, 4, 0, 0, 0, 0, 0, 2, 2, 0, 0, 0, 0, 0, -4, 4},
    {1, 1, 1, 0, 1, 0, 0, 0, 2, 0, -1, -1, 0, 1, 1, 0, 0, -1, 1, 1, 0, 0, -2, -2, -2, 0, -2},
    {1, -1, 1, 0, -1, 0, 0, 0, 2, 0, 1, -1, 0, -1, 1, 0, 0, -1, 1, 1, 0, 0, 2, 2, -2, 0, -2},
    {1, 1, -1, 0, -1, 0, 0, 0, 2, 0, -1, 1, 0, 1, -1, 0, 0, -1, 1, 1, 0, 0, 2, -2, 2, 0, -2},
    {1, -1, -1, 0, 1, 0, 0, 0, 2, 0, 1, 1, 0, -1, -1, 0, 0, -1, 1, 1, 0, 0, -2, 2, 2, 0, -2},
    {1, 1, 0, 1, 0, 1, 0, 1, -1, 0, -1, 0, -1, -1, 0, 1, 0, -1, 1, -1, 0, -2, 0, -2, 0, -2, -2},
    {1, -1, 0, 1, 0, -1, 0, 1, -1, 0, 1, 0, -1, 1, 0, 1, 0, -1, 1, -1, 0, 2, 0, 2, 0, -2, -2},
    {1, 1, 0, -1, 0, -1, 0, 1, -1, 0, -1, 0, 1, -1, 0, -1, 0, -1, 1, -1, 0, 2, 0, -2, 0, 2, -2},
    {1, -1, 0, -1, 0, 1, 0, 1, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 1, -1, 0, -2, 0, 2, 0, 2, -2},
    {1, 0, 1, 1, 0, 0, 1, -1, -1, 0, 0, -1, -1, 0, -1, -1, 0, -1, -2, 0, -2, 0, 0, 0, -2, -2, -2},
    {1, 0, -1, 1, 0, 0, -1, -1, -1, 0, 0, 1, -1, 0, 1, -1, 0, -1, -2, 0, 2, 0, 0, 0, 2, -2, -2},
    {1, 0, 1, -1, 0, 0, -1, -1, -1, 0, 0, -1, 1, 0, -1, 1, 0, -1, -2, 0, 2, 0, 0, 0, -2, 2, -2},
    {1, 0, -1, -1, 0, 0, 1, -1, -1, 0, 0, 1, 1, 0, 1, 1, 0, -1, -2, 0, -2, 0, 0, 0, 2, 2, -2},
    {1, 1, 1, 1, 1, 1, 1, 0, 0, 1, 2, 2, 2, 0, 0, 0, 1, 1, 0, 0, 1, 1, 1, 1, 1, 1, 1},
    {1, -1, 1, 1, -1, -1, 1, 0, 0, 1, -2, 2, 2, 0, 0, 0, -1, 1, 0, 0, 1, -1, -1, -1, 1, 1, 1},
    {1, 1, -1, 1, -1, 1, -1, 0, 0, 1, 2, -2, 2, 0, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, -1, 1, 1},
    {1, -1, -1, 1, 1, -1, -1, 0, 0, 1, -2, -2, 2, 0, 0, 0, 1, 1, 0, 0, -1, -1, 1, -1, -1, 1, 1},
    {1, 1, 1, -1, 1, -1, -1, 0, 0, 1, 2, 2, -2, 0, 0, 0, -1, 1, 0, 0, -1, -1, 1, 1, 1, -1, 1},
    {1, -1, 1, -1, -1, 1, -1, 0, 0, 1, -2, 2, -2, 0, 0, 0, 1, 1, 0, 0, -1, 1, -1, -1, 1, -1, 1},
    {1, 1, -1, -1, -1, -1, 1, 0, 0, 1, 2, -2, -2, 0, 0, 0, 1, 1, 0, 0, 1, -1, -1, 1, -1, -1, 1},
    {1, -1, -1, -1, 1, 1, 1, 0, 0, 1, -2, -2, -2, 0, 0, 0, -1, 1, 0, 0, 1, 1, 1, -1, -1, -1, 1},
};

static double w27(int a) {
    int n = abs(C27[a][0]) + abs(C27[a][1]) + abs(C27[a][2]);
    return n == 0 ? 8.0 / 27.0 : n == 1 ? 2.0 / 27.0 : n == 2 ? 1.0 / 54.0
                                                               : 1.0 / 216.0;
}

static int lat_dims(int Q) { return Q == 9 ? 2 : 3; }
static double lat_c(int Q, int a, int d) {
    return Q == 9 ? (double)C9[a][d] : (double)C27[a][d];
}
static double lat_w(int Q, int a) { return Q == 9 ? W9[a] : w27(a); }
static int lat_bb(int Q, int a) { return Q == 9 ? BB9[a] : BB27[a]; }
static int lat_K(int Q, int l, int j) { return Q == 9 ? K9[l][j] : K27[l][j]; }

int NO_FN(lattice_tables)(int Q, double* c, double* w, int* bb, double* K) {
    if (Q != 9 && Q != 27) return 1;
    int D = lat_dims(Q);
    for (int a = 0; a < Q; ++a) {
        for (int d = 0; d < D; ++d) c[a * D + d] = lat_c(Q, a, d);
        w[a] = lat_w(Q, a);
        bb[a] = lat_bb(Q, a);
        for (int j = 0; j < Q; ++j) K[a * Q + j] = lat_K(Q, a, j);
    }
    return 0;
}

/* ======================================================================
 * Equilibrium: detail/equilibrium_distribution.cuh:41-64.  Keeps the
 * reference's "- u.u" term (no 1/(2 c_s^2) factor).  dot<D> sums the last
 * component first (math.hpp:198-207).
 * ====================================================================== */
static no_real dotD(int D, const no_real* a, const no_real* b) {
    no_real s = a[0] * b[0];
    for (int d = 1; d < D; ++d) s = a[d] * b[d] + s;
    return s;
}

static no_real feq_impl(int Q, no_real rho, const no_real* u, int alpha) {
    int D = lat_dims(Q);
    no_real c[3];
    for (int d = 0; d < D; ++d) c[d] = (no_real)lat_c(Q, alpha, d);
    no_real w = (no_real)lat_w(Q, alpha);
    no_real udotu = dotD(D, u, u);
    no_real udotc = dotD(D, u, c);
    const no_real cs = (no_real)CS_LAT;
    no_real cs2 = cs * cs;
    no_real cs4 = cs2 * cs2;
    return w * rho
           * ((no_real)1 + udotc / cs2 + (udotc * udotc) / ((no_real)2 * cs4)
              - udotu);
}

no_real NO_FN(feq)(int Q, no_real rho, const no_real* u, int alpha) {
    return feq_impl(Q, rho, u, alpha);
}

/* ======================================================================
 * Cascaded central-moment collision.
 * ====================================================================== */

/* loopless::pow<N> (math.hpp:183-196) for N = 0, 1, 2 */
static no_real ipow(no_real x, int n) {
    return n == 0 ? (no_real)1 : n == 1 ? x : x * x;
}

/* detail/d3q27_lattice.cuh:389-412 */
static no_real cmom27(const no_real* f, const no_real* u, int m, int n, int p) {
    no_real s = 0;
    for (int a = 0; a < 27; ++a) {
        s += f[a] * ipow((no_real)C27[a][0] - u[0], m)
             * ipow((no_real)C27[a][1] - u[1], n)
             * ipow((no_real)C27[a][2] - u[2], p);
    }
    return s;
}

/* relaxation rate from a unit-time rate: detail/d3q27_lattice.cuh:200-204 */
static no_real rate27(no_real om, no_real dt) {
    no_real unit_nu = ((no_real)1 / om - (no_real)0.5) / (no_real)3;
    return (no_real)1 / ((no_real)3 * unit_nu + (no_real)0.5) / dt;
}

/* detail/d3q27_lattice.cuh:125-378 */
static void project27(no_real* k, const no_real* f, no_real rho,
                      const no_real* u, no_real nu, no_real dt) {
    const no_real v0 = u[0], v1 = u[1], v2 = u[2];
    const no_real v0sq = v0 * v0, v1sq = v1 * v1, v2sq = v2 * v2;
    k[0] = k[1] = k[2] = k[3] = 0;
    const no_real mxy = cmom27(f, u, 1, 1, 0), mxz = cmom27(f, u, 1, 0, 1);
    const no_real myz = cmom27(f, u, 0, 1, 1), mxx = cmom27(f, u, 2, 0, 0);
    const no_real myy = cmom27(f, u, 0, 2, 0), mzz = cmom27(f, u, 0, 0, 2);
    const no_real mxxy = cmom27(f, u, 2, 1, 0), mxxz = cmom27(f, u, 2, 0, 1);
    const no_real mxyy = cmom27(f, u, 1, 2, 0), myyz = cmom27(f, u, 0, 2, 1);
    const no_real mxzz = cmom27(f, u, 1, 0, 2), myzz = cmom27(f, u, 0, 1, 2);
    const no_real mxyz = cmom27(f, u, 1, 1, 1), mxyzz = cmom27(f, u, 1, 1, 2);
    const no_real mxxyz = cmom27(f, u, 2, 1, 1), mxyyz = cmom27(f, u, 1, 2, 1);
    const no_real mxxyy = cmom27(f, u, 2, 2, 0), mxxzz = cmom27(f, u, 2, 0, 2);
    const no_real myyzz = cmom27(f, u, 0, 2, 2), mxxyzz = cmom27(f, u, 2, 1, 2);
    const no_real mxyyzz = cmom27(f, u, 1, 2, 2), mxxyyz = cmom27(f, u, 2, 2, 1);
    const no_real mxxyyzz = cmom27(f, u, 2, 2, 2);
    no_real om = (no_real)1 / ((no_real)3 * nu + (no_real)0.5) / dt;

    k[4] = om * (-mxy / 12);
    k[5] = om * (-mxz / 12);
    k[6] = om * (-myz / 12);
    k[7] = om * (-(mxx - myy) / 12);
    k[8] = om * (-(mxx + myy - 2 * mzz) / 36);
    k[9] = om * (-(mxx + myy + mzz - rho) / 18);

    om = rate27((no_real)1.5f, dt);
    k[10] = om * (-(6 * mxy * v1 + 3 * mxyy + 6 * mxz * v2 + 3 * mxzz
                    + 3 * myy * v0 + 3 * mzz * v0 - 2 * rho * v0) / 72);
    k[11] = om * (-(3 * mxx * v1 + 3 * mxxy + 6 * mxy * v0 + 6 * myz * v2
                    + 3 * myzz + 3 * mzz * v1 - 2 * rho * v1) / 72);
    k[12] = om * (-(3 * mxx * v2 + 3 * mxxz + 6 * mxz * v0 + 3 * myy * v2
                    + 3 * myyz + 6 * myz * v1 - 2 * rho * v2) / 72);

    om = rate27((no_real)1.83f, dt);
    k[13] = om * (-(2 * mxy * v1 + mxyy - 2 * mxz * v2 - mxzz + myy * v0
                    - mzz * v0) / 8);
    k[14] = om * (-(mxx * v1 + mxxy + 2 * mxy * v0 - 2 * myz * v2 - myzz
                    - mzz * v1) / 8);
    k[15] = om * (-(mxx * v2 + mxxz + 2 * mxz * v0 - myy * v2 - myyz
                    - 2 * myz * v1) / 8);

    om = rate27((no_real)1.4f, dt);
    k[16] = om * (-(mxy * v2 + mxyz + mxz * v1 + myz * v0) / 8);

    om = rate27((no_real)1.61f, dt);
    k[17] = om * (-(3 * mxx * v1sq + 3 * mxx * v2sq - 4 * mxx + 6 * mxxy * v1
                    + 3 * mxxyy + 6 * mxxz * v2 + 3 * mxxzz + 12 * mxy * v0 * v1
                    + 6 * mxyy * v0 + 12 * mxz * v0 * v2 + 6 * mxzz * v0
                    + 3 * myy * v0sq + 3 * myy * v2sq - 4 * myy + 6 * myyz * v2
                    + 3 * myyzz + 12 * myz * v1 * v2 + 6 * myzz * v1
                    + 3 * mzz * v0sq + 3 * mzz * v1sq - 4 * mzz - 2 * rho * v0sq
                    - 2 * rho * v1sq - 2 * rho * v2sq + 3 * rho) / 36);

    om = rate27((no_real)1.98f, dt);
    k[18] = om * (-(3 * mxx * v1sq + 3 * mxx * v2sq - 4 * mxx + 6 * mxxy * v1
                    + 3 * mxxyy + 6 * mxxz * v2 + 3 * mxxzz + 12 * mxy * v0 * v1
                    + 6 * mxyy * v0 + 12 * mxz * v0 * v2 + 6 * mxzz * v0
                    + 3 * myy * v0sq - 6 * myy * v2sq + 2 * myy - 12 * myyz * v2
                    - 6 * myyzz - 24 * myz * v1 * v2 - 12 * myzz * v1
                    + 3 * mzz * v0sq - 6 * mzz * v1sq + 2 * mzz - 2 * rho * v0sq
                    + rho * v1sq + rho * v2sq) / 72);
    k[19] = om * (-(3 * mxx * v1sq - 3 * mxx * v2sq + 6 * mxxy * v1 + 3 * mxxyy
                    - 6 * mxxz * v2 - 3 * mxxzz + 12 * mxy * v0 * v1
                    + 6 * mxyy * v0 - 12 * mxz * v0 * v2 - 6 * mxzz * v0
                    + 3 * myy * v0sq - 2 * myy - 3 * mzz * v0sq + 2 * mzz
                    - rho * v1sq + rho * v2sq) / 24);

    om = rate27((no_real)1.98f, dt);
    k[20] = om * (-(3 * mxx * v1 * v2 + 3 * mxxy * v2 + 3 * mxxyz + 3 * mxxz * v1
                    + 6 * mxy * v0 * v2 + 6 * mxyz * v0 + 6 * mxz * v0 * v1
                    + 3 * myz * v0sq - 2 * myz - rho * v1 * v2) / 24);
    k[21] = om * (-(6 * mxy * v1 * v2 + 3 * mxyy * v2 + 3 * mxyyz + 6 * mxyz * v1
                    + 3 * mxz * v1sq - 2 * mxz + 3 * myy * v0 * v2
                    + 3 * myyz * v0 + 6 * myz * v0 * v1 - rho * v0 * v2) / 24);
    k[22] = om * (-(3 * mxy * v2sq - 2 * mxy + 6 * mxyz * v2 + 3 * mxyzz
                    + 6 * mxz * v1 * v2 + 3 * mxzz * v1 + 6 * myz * v0 * v2
                    + 3 * myzz * v0 + 3 * mzz * v0 * v1 - rho * v0 * v1) / 24);

    om = rate27((no_real)1.74f, dt);
    k[23] = om * (-(6 * mxy * v1 * v2sq - 4 * mxy * v1 + 3 * mxyy * v2sq
                    - 2 * mxyy + 6 * mxyyz * v2 + 3 * mxyyzz
                    + 12 * mxyz * v1 * v2 + 6 * mxyzz * v1 + 6 * mxz * v1sq * v2
                    - 4 * mxz * v2 + 3 * mxzz * v1sq - 2 * mxzz
                    + 3 * myy * v0 * v2sq - 2 * myy * v0 + 6 * myyz * v0 * v2
                    + 3 * myyzz * v0 + 12 * myz * v0 * v1 * v2
                    + 6 * myzz * v0 * v1 + 3 * mzz * v0 * v1sq - 2 * mzz * v0
                    - rho * v0 * v1sq - rho * v0 * v2sq + rho * v0) / 24);
    k[24] = om * (-(3 * mxx * v1 * v2sq - 2 * mxx * v1 + 3 * mxxy * v2sq
                    - 2 * mxxy + 6 * mxxyz * v2 + 3 * mxxyzz + 6 * mxxz * v1 * v2
                    + 3 * mxxzz * v1 + 6 * mxy * v0 * v2sq - 4 * mxy * v0
                    + 12 * mxyz * v0 * v2 + 6 * mxyzz * v0
                    + 12 * mxz * v0 * v1 * v2 + 6 * mxzz * v0 * v1
                    + 6 * myz * v0sq * v2 - 4 * myz * v2 + 3 * myzz * v0sq
                    - 2 * myzz + 3 * mzz * v0sq * v1 - 2 * mzz * v1
                    - rho * v0sq * v1 - rho * v1 * v2sq + rho * v1) / 24);
    k[25] = om * (-(3 * mxx * v1sq * v2 - 2 * mxx * v2 + 6 * mxxy * v1 * v2
                    + 3 * mxxyy * v2 + 3 * mxxyyz + 6 * mxxyz * v1
                    + 3 * mxxz * v1sq - 2 * mxxz + 12 * mxy * v0 * v1 * v2
                    + 6 * mxyy * v0 * v2 + 6 * mxyyz * v0 + 12 * mxyz * v0 * v1
                    + 6 * mxz * v0 * v1sq - 4 * mxz * v0 + 3 * myy * v0sq * v2
                    - 2 * myy * v2 + 3 * myyz * v0sq - 2 * myyz
                    + 6 * myz * v0sq * v1 - 4 * myz * v1 - rho * v0sq * v2
                    - rho * v1sq * v2 + rho * v2) / 24);
    k[26] = om * (-(27 * mxx * v1sq * v2sq - 18 * mxx * v1sq - 18 * mxx * v2sq
                    + 12 * mxx + 54 * mxxy * v1 * v2sq - 36 * mxxy * v1
                    + 27 * mxxyy * v2sq - 18 * mxxyy + 54 * mxxyyz * v2
                    + 27 * mxxyyzz + 108 * mxxyz * v1 * v2 + 54 * mxxyzz * v1
                    + 54 * mxxz * v1sq * v2 - 36 * mxxz * v2 + 27 * mxxzz * v1sq
                    - 18 * mxxzz + 108 * mxy * v0 * v1 * v2sq
                    - 72 * mxy * v0 * v1 + 54 * mxyy * v0 * v2sq - 36 * mxyy * v0
                    + 108 * mxyyz * v0 * v2 + 54 * mxyyzz * v0
                    + 216 * mxyz * v0 * v1 * v2 + 108 * mxyzz * v0 * v1
                    + 108 * mxz * v0 * v1sq * v2 - 72 * mxz * v0 * v2
                    + 54 * mxzz * v0 * v1sq - 36 * mxzz * v0
                    + 27 * myy * v0sq * v2sq - 18 * myy * v0sq - 18 * myy * v2sq
                    + 12 * myy + 54 * myyz * v0sq * v2 - 36 * myyz * v2
                    + 27 * myyzz * v0sq - 18 * myyzz + 108 * myz * v0sq * v1 * v2
                    - 72 * myz * v1 * v2 + 54 * myzz * v0sq * v1 - 36 * myzz * v1
                    + 27 * mzz * v0sq * v1sq - 18 * mzz * v0sq - 18 * mzz * v1sq
                    + 12 * mzz - 9 * rho * v0sq * v1sq - 9 * rho * v0sq * v2sq
                    + 9 * rho * v0sq - 9 * rho * v1sq * v2sq + 9 * rho * v1sq
                    + 9 * rho * v2sq - 7 * rho) / 216);
}

/* detail/d2q9_lattice.cuh:120-299 (indices r=0, nw=1, w=2, sw=3, s=4, se=5,
 * e=6, ne=7, n=8).  The engine's 6-argument call (simulation_engine.cuh:
 * 509-516) does not compile against this 7-argument projection at HEAD; the
 * unused characteristic-frequency argument is dropped here. */
static void project9(no_real* k, const no_real* f, no_real rho,
                     const no_real* u, no_real nu, no_real dt) {
    enum { R = 0, NW, W, SW, S, SE, E, NE, N };
    const no_real ux = u[0], uy = u[1];
    const no_real om_ab = (no_real)1 / ((no_real)3 * nu + (no_real)0.5 * dt);
    const no_real inv_dt = (no_real)1 / dt;
    k[0] = k[1] = k[2] = 0;
    /* compute_k3 (:166-180) */
    k[3] = inv_dt
           * (rho * (ux * ux + uy * uy) - f[E] - f[N] - f[S] - f[W]
              - 2 * (f[SE] + f[SW] + f[NE] + f[NW] - rho / (no_real)3))
           / (no_real)12;
    /* compute_k4 (:182-193) */
    k[4] = om_ab * (f[N] + f[S] - f[E] - f[W] + rho * (ux * ux - uy * uy))
           / (no_real)4;
    /* compute_k5 (:195-205) */
    k[5] = om_ab * (f[NE] + f[SW] - f[NW] - f[SE] + rho * (ux * uy))
           / (no_real)4;
    /* compute_k6 (:207-231) */
    {
        no_real t1 = (f[SE] + f[SW] - f[NE] - f[NW] - 2 * ux * ux * uy * rho
                      + uy * (rho - f[N] - f[S] - f[R]))
                     / (no_real)4;
        no_real t2 = ux / (no_real)2 * (f[NE] - f[NW] - f[SE] + f[SW]);
        no_real t3 = -uy / (no_real)2 * (-3 * k[3] - k[4]) - 2 * ux * k[5];
        k[6] = -inv_dt * (t1 + t2) + t3;
    }
    /* compute_k7 (:233-257) */
    {
        no_real t1 = (f[SW] + f[NW] - f[SE] - f[NE] - 2 * ux * uy * uy * rho
                      + ux * (rho - f[W] - f[E] - f[R]))
                     / (no_real)4;
        no_real t2 = uy / (no_real)2 * (f[NE] + f[SW] - f[SE] - f[NW]);
        no_real t3 = -ux / (no_real)2 * (-3 * k[3] + k[4]) - 2 * uy * k[5];
        k[7] = -inv_dt * (t1 + t2) + t3;
    }
    /* compute_k8 (:259-299) */
    {
        no_real t1 = rho / (no_real)9 - f[NE] - f[NW] - f[SE] - f[SW];
        no_real t2 = 2 * (ux * (f[NE] - f[NW] + f[SE] - f[SW])
                          + uy * (f[NE] + f[NW] - f[SE] - f[SW]));
        no_real t3 = 4 * ux * uy * (f[NW] - f[NE] + f[SE] - f[SW]);
        no_real t4 = -ux * ux * (f[N] + f[NE] + f[NW] + f[S] + f[SE] + f[SW]);
        no_real t5 = uy * uy
                     * (3 * ux * ux * rho - f[E] - f[NE] - f[NW] - f[SE] - f[SW]
                        - f[W]);
        no_real t6 = -2 * k[3] - 2 * ux * k[7] - 2 * uy * k[6]
                     + 4 * ux * uy * k[5]
                     - ((no_real)3 / (no_real)2 * k[3] - k[4] / (no_real)2)
                           * (ux * ux + uy * uy);
        k[8] = inv_dt * ((no_real)1 / (no_real)4 * (t1 + t2 + t3 + t4 + t5))
               + t6;
    }
}

static void collide_impl(int Q, const no_real* f, no_real rho,
                         const no_real* u, no_real nu, no_real dt,
                         no_real* df) {
    no_real k[27];
    if (Q == 27)
        project27(k, f, rho, u, nu, dt);
    else
        project9(k, f, rho, u, nu, dt);
    /* simulation_engine.cuh:518-527 */
    for (int l = 0; l < Q; ++l) {
        no_real acc = 0;
        for (int j = 0; j < Q; ++j) acc += (no_real)lat_K(Q, l, j) * k[j];
        df[l] = acc * dt;
    }
}

int NO_FN(collide)(int Q, const no_real* f, no_real rho, const no_real* u,
                   no_real nu_lat, no_real dt_lat, no_real* df) {
    if (Q != 9 && Q != 27) return 1;
    collide_impl(Q, f, rho, u, nu_lat, dt_lat, df);
    return 0;
}

/* ======================================================================
 * RK4 combination (detail/rk4_solver.cuh:48-171) on y' = lambda y.
 * ====================================================================== */
double NO_FN(rk4_scalar)(double y0, double lambda, double dt, int nsteps) {
    const no_real a[3] = {(no_real)(1. / 2.), (no_real)(1. / 2.), (no_real)1.};
    const no_real b[4] = {(no_real)(1. / 6.), (no_real)(2. / 6.),
                          (no_real)(2. / 6.), (no_real)(1. / 6.)};
    no_real y = (no_real)y0, lam = (no_real)lambda, h = (no_real)dt;
    for (int n = 0; n < nsteps; ++n) {
        no_real y0s = y, acc, yk, kk;
        kk = lam * y0s;                 /* k1 */
        yk = y0s + (h * a[0]) * kk;
        kk = kk * b[0];
        acc = kk;
        kk = lam * yk;                  /* k2 */
        yk = y0s + (h * a[1]) * kk;
        kk = kk * b[1];
        acc = acc + kk;
        kk = lam * yk;                  /* k3 */
        yk = y0s + (h * a[2]) * kk;
        kk = kk * b[2];
        acc = acc + kk;
        kk = lam * yk;                  /* k4 */
        kk = kk * b[3];
        acc = acc + kk;
        y = y0s + h * acc;
    }
    return (double)y;
}

/* ======================================================================
 * Synthetic uniform-grid node provider.
 * ====================================================================== */
static uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Order of the nodes within each class: the Hilbert curve through the
 * grid points (J. Skilling, "Programming the Hilbert curve", AIP Conf.
 * Proc. 707, 2004): the axes are turned into the transposed Hilbert index
 * by the inverse undo and a Gray code, whose bits are then read most
 * significant first, one bit of each axis per level. */
static uint64_t hilbert_key(int D, const int64_t* g) {
    const int bits = D == 2 ? 32 : 21;
    uint32_t x[3];
    x[0] = (uint32_t)g[0];
    x[1] = (uint32_t)g[1];
    x[2] = D == 3 ? (uint32_t)g[2] : 0u;
    const uint32_t top = 1u << (bits - 1);
    for (uint32_t q = top; q > 1; q >>= 1) {
        const uint32_t p = q - 1;
        for (int i = 0; i < D; ++i) {
            if (x[i] & q) {
                x[0] ^= p;
            } else {
                const uint32_t t = (x[0] ^ x[i]) & p;
                x[0] ^= t;
                x[i] ^= t;
            }
        }
    }
    for (int i = 1; i < D; ++i) x[i] ^= x[i - 1];
    uint32_t flip = 0;
    for (uint32_t q = top; q > 1; q >>= 1)
        if (x[D - 1] & q) flip ^= q - 1;
    for (int i = 0; i < D; ++i) x[i] ^= flip;
    uint64_t key = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int i = 0; i < D; ++i) key = (key << 1) | ((x[i] >> b) & 1u);
    return key;
}

/* ranges/uniform_grid.cuh:55-65: size ceil(extent/h)+1 in T = float */
int64_t NO_FN(grid_size)(int D, const double* lo, const double* hi, double h,
                         int64_t* dims) {
    int64_t n = 1;
    for (int d = 0; d < D; ++d) {
        float ext = (float)hi[d] - (float)lo[d];
        dims[d] = (int64_t)ceilf(ext / (float)h) + 1;
        n *= dims[d];
    }
    return n;
}

typedef struct {
    uint64_t key;
    int64_t g;
} keyed_t;

static int cmp_keyed(const void* a, const void* b) {
    const keyed_t* x = (const keyed_t*)a;
    const keyed_t* y = (const keyed_t*)b;
    return x->key < y->key ? -1 : x->key > y->key ? 1 : 0;
}

int NO_FN(uniform_grid)(int D, const double* lo, const double* hi, double h,
                        double jitter_frac, uint64_t seed, double layer_L,
                        double sigma_max, float* pts, float* normals,
                        float* sigma, int64_t* counts, int64_t* grid_index) {
    int64_t dims[3];
    int64_t N = NO_FN(grid_size)(D, lo, hi, h, dims);
    double sp[3];
    for (int d = 0; d < D; ++d) sp[d] = (hi[d] - lo[d]) / (double)(dims[d] - 1);

    float* xyz = (float*)malloc(sizeof(float) * N * D);
    unsigned char* isb = (unsigned char*)malloc(N);
    uint64_t st = seed;
    int64_t nb = 0;
    for (int64_t g = 0; g < N; ++g) {
        int64_t gi[3], rem = g;
        int bnd = 0;
        for (int d = 0; d < D; ++d) {
            gi[d] = rem % dims[d];
            rem /= dims[d];
            if (gi[d] == 0 || gi[d] == dims[d] - 1) bnd = 1; /* uniform_grid.cuh:242-253 */
        }
        isb[g] = (unsigned char)bnd;
        nb += bnd;
        for (int d = 0; d < D; ++d) {
            double u01 = (double)(splitmix64(&st) >> 11) * 0x1.0p-53;
            double x = lo[d] + (double)gi[d] * sp[d];
            if (!bnd) x += (2.0 * u01 - 1.0) * jitter_frac * h;
            xyz[g * D + d] = (float)x;
        }
    }
    /* layer: 1-NN squared distance from interior nodes to boundary nodes
     * (uniform_grid_provider.cuh:104-127, detail/uniform_grid_provider.inl:
     * 111-171); layer iff d2 <= L^2 with L^2 formed in float. */
    int64_t ni = N - nb;
    float* bpts = (float*)malloc(sizeof(float) * (nb ? nb : 1) * D);
    float* ipts = (float*)malloc(sizeof(float) * (ni ? ni : 1) * D);
    int64_t* bg = (int64_t*)malloc(sizeof(int64_t) * (nb ? nb : 1));
    int64_t* ig = (int64_t*)malloc(sizeof(int64_t) * (ni ? ni : 1));
    int64_t cb = 0, ci = 0;
    for (int64_t g = 0; g < N; ++g) {
        if (isb[g]) {
            memcpy(bpts + cb * D, xyz + g * D, sizeof(float) * D);
            bg[cb++] = g;
        } else {
            memcpy(ipts + ci * D, xyz + g * D, sizeof(float) * D);
            ig[ci++] = g;
        }
    }
    float* d2 = (float*)malloc(sizeof(float) * (ni ? ni : 1));
    uint32_t* nn = (uint32_t*)malloc(sizeof(uint32_t) * (ni ? ni : 1));
    if (layer_L > 0 && ni > 0 && nb > 0)
        NO_FN(knn)(D, bpts, nb, ipts, ni, 1, d2, nn);
    float Lf = (float)layer_L;
    float L2 = Lf * Lf;

    keyed_t* kb = (keyed_t*)malloc(sizeof(keyed_t) * (ni ? ni : 1));
    keyed_t* kl = (keyed_t*)malloc(sizeof(keyed_t) * (ni ? ni : 1));
    keyed_t* kbd = (keyed_t*)malloc(sizeof(keyed_t) * (nb ? nb : 1));
    int64_t nbulk = 0, nlay = 0;
    for (int64_t t = 0; t < ni; ++t) {
        int64_t g = ig[t], gi[3], rem = g;
        for (int d = 0; d < D; ++d) {
            gi[d] = rem % dims[d];
            rem /= dims[d];
        }
        keyed_t e = {hilbert_key(D, gi), t};
        if (layer_L > 0 && d2[t] <= L2)
            kl[nlay++] = e;
        else
            kb[nbulk++] = e;
    }
    for (int64_t t = 0; t < nb; ++t) {
        int64_t g = bg[t], gi[3], rem = g;
        for (int d = 0; d < D; ++d) {
            gi[d] = rem % dims[d];
            rem /= dims[d];
        }
        keyed_t e = {hilbert_key(D, gi), t};
        kbd[t] = e;
    }
    qsort(kb, nbulk, sizeof(keyed_t), cmp_keyed);
    qsort(kl, nlay, sizeof(keyed_t), cmp_keyed);
    qsort(kbd, nb, sizeof(keyed_t), cmp_keyed);

    int64_t o = 0;
    for (int64_t t = 0; t < nbulk; ++t, ++o) {
        memcpy(pts + o * D, ipts + kb[t].g * D, sizeof(float) * D);
        if (grid_index) grid_index[o] = ig[kb[t].g];
    }
    for (int64_t t = 0; t < nlay; ++t, ++o) {
        memcpy(pts + o * D, ipts + kl[t].g * D, sizeof(float) * D);
        if (grid_index) grid_index[o] = ig[kl[t].g];
        /* detail/uniform_grid_provider.inl:148-159 */
        float dist = sqrtf(d2[kl[t].g]);
        float r = (Lf - dist) / Lf;
        sigma[t] = (float)sigma_max * (r * r);
    }
    for (int64_t t = 0; t < nb; ++t, ++o) {
        int64_t g = bg[kbd[t].g], gi[3], rem = g;
        memcpy(pts + o * D, bpts + kbd[t].g * D, sizeof(float) * D);
        if (grid_index) grid_index[o] = g;
        for (int d = 0; d < D; ++d) {
            gi[d] = rem % dims[d];
            rem /= dims[d];
        }
        /* uniform_grid.cuh:255-293, then negated to point inward
         * (uniform_grid_provider.cuh:89-94) */
        float nv[3] = {0.f, 0.f, 0.f};
        for (int d = 0; d < D; ++d)
            nv[d] = gi[d] == 0 ? -1.f : gi[d] == dims[d] - 1 ? 1.f : 0.f;
        float n2 = nv[0] * nv[0];
        for (int d = 1; d < D; ++d) n2 = nv[d] * nv[d] + n2;
        float nrm = sqrtf(n2);
        for (int d = 0; d < D; ++d)
            normals[t * D + d] = -(nv[d] / (nrm == 0.f ? 1.f : nrm));
    }
    counts[0] = nbulk;
    counts[1] = nlay;
    counts[2] = nb;
    free(xyz); free(isb); free(bpts); free(ipts); free(bg); free(ig);
    free(d2); free(nn); free(kb); free(kl); free(kbd);
    return 0;
}

/* ======================================================================
 * Exact kNN.
 * ====================================================================== */
static float d2_f32(int D, const float* a, const float* b) {
    /* distance_functions.hpp:151-160 with the recursion unrolled: the
     * highest dimension's term is added last. */
    float s = (a[0] - b[0]) * (a[0] - b[0]);
    for (int d = 1; d < D; ++d) {
        float t = (a[d] - b[d]) * (a[d] - b[d]);
        s = t + s;
    }
    return s;
}

static int key_less(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

static void topk_insert(int k, int* cnt, float* kd, uint32_t* ki, float d,
                        uint32_t id) {
    int n = *cnt;
    if (n == k) {
        if (!key_less(d, id, kd[k - 1], ki[k - 1])) return;
        n = k - 1;
    }
    int p = n;
    while (p > 0 && key_less(d, id, kd[p - 1], ki[p - 1])) {
        kd[p] = kd[p - 1];
        ki[p] = ki[p - 1];
        --p;
    }
    kd[p] = d;
    ki[p] = id;
    if (*cnt < k) ++*cnt;
}

int NO_FN(knn_brute)(int D, const float* ref, int64_t n_ref,
                     const float* query, int64_t n_q, int k, float* d2_out,
                     uint32_t* idx_out) {
    if (k <= 0 || n_ref < k) return 1;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n_q; ++q) {
        int cnt = 0;
        float* kd = d2_out + q * k;
        uint32_t* ki = idx_out + q * k;
        for (int64_t j = 0; j < n_ref; ++j)
            topk_insert(k, &cnt, kd, ki, d2_f32(D, query + q * D, ref + j * D),
                        (uint32_t)j);
    }
    return 0;
}

int NO_FN(knn)(int D, const float* ref, int64_t n_ref, const float* query,
               int64_t n_q, int k, float* d2_out, uint32_t* idx_out) {
    if (k <= 0 || n_ref < k) return 1;
    double lo[3], hi[3], vol = 1.0;
    for (int d = 0; d < D; ++d) {
        lo[d] = DBL_MAX;
        hi[d] = -DBL_MAX;
    }
    for (int64_t j = 0; j < n_ref; ++j)
        for (int d = 0; d < D; ++d) {
            double x = ref[j * D + d];
            if (x < lo[d]) lo[d] = x;
            if (x > hi[d]) hi[d] = x;
        }
    for (int d = 0; d < D; ++d) {
        if (hi[d] - lo[d] < 1e-6) hi[d] = lo[d] + 1e-6;
        vol *= hi[d] - lo[d];
    }
    double w = pow(4.0 * vol / (double)n_ref, 1.0 / D);
    int64_t dims[3] = {1, 1, 1}, ncell = 1;
    for (int d = 0; d < D; ++d) {
        dims[d] = (int64_t)ceil((hi[d] - lo[d]) / w);
        if (dims[d] < 1) dims[d] = 1;
        if (dims[d] > 4096) dims[d] = 4096;
        ncell *= dims[d];
    }
    int64_t* start = (int64_t*)calloc(ncell + 1, sizeof(int64_t));
    int64_t* cell_of = (int64_t*)malloc(sizeof(int64_t) * n_ref);
    uint32_t* sorted = (uint32_t*)malloc(sizeof(uint32_t) * n_ref);
    for (int64_t j = 0; j < n_ref; ++j) {
        int64_t c = 0, mul = 1;
        for (int d = 0; d < D; ++d) {
            int64_t ci = (int64_t)floor((ref[j * D + d] - lo[d]) / w);
            if (ci < 0) ci = 0;
            if (ci >= dims[d]) ci = dims[d] - 1;
            c += ci * mul;
            mul *= dims[d];
        }
        cell_of[j] = c;
        start[c + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    {
        int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (ncell + 1));
        memcpy(fill, start, sizeof(int64_t) * (ncell + 1));
        for (int64_t j = 0; j < n_ref; ++j) sorted[fill[cell_of[j]]++] = (uint32_t)j;
        free(fill);
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t q = 0; q < n_q; ++q) {
        const float* x = query + q * D;
        float* kd = d2_out + q * k;
        uint32_t* ki = idx_out + q * k;
        int cnt = 0;
        int64_t qc[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) {
            int64_t ci = (int64_t)floor((x[d] - lo[d]) / w);
            if (ci < 0) ci = 0;
            if (ci >= dims[d]) ci = dims[d] - 1;
            qc[d] = ci;
        }
        for (int64_t r = 0;; ++r) {
            int64_t a[3], b[3];
            for (int d = 0; d < 3; ++d) {
                a[d] = d < D ? qc[d] - r : 0;
                b[d] = d < D ? qc[d] + r : 0;
                if (d < D) {
                    if (a[d] < 0) a[d] = 0;
                    if (b[d] >= dims[d]) b[d] = dims[d] - 1;
                }
            }
            for (int64_t cz = a[2]; cz <= b[2]; ++cz)
                for (int64_t cy = a[1]; cy <= b[1]; ++cy)
                    for (int64_t cx = a[0]; cx <= b[0]; ++cx) {
                        int64_t cc[3] = {cx, cy, cz};
                        int64_t cheb = 0;
                        for (int d = 0; d < D; ++d) {
                            int64_t t = cc[d] > qc[d] ? cc[d] - qc[d] : qc[d] - cc[d];
                            if (t > cheb) cheb = t;
                        }
                        if (cheb != r) continue;
                        int64_t c = cx + dims[0] * (cy + dims[1] * cz);
                        for (int64_t t = start[c]; t < start[c + 1]; ++t) {
                            uint32_t j = sorted[t];
                            topk_insert(k, &cnt, kd, ki,
                                        d2_f32(D, x, ref + (int64_t)j * D), j);
                        }
                    }
            /* every unseen point lies outside the block of radius r */
            double bound = DBL_MAX;
            int all = 1;
            for (int d = 0; d < D; ++d) {
                if (qc[d] - r > 0) {
                    double t = x[d] - (lo[d] + (double)(qc[d] - r) * w);
                    if (t < bound) bound = t;
                    all = 0;
                }
                if (qc[d] + r < dims[d] - 1) {
                    double t = (lo[d] + (double)(qc[d] + r + 1) * w) - x[d];
                    if (t < bound) bound = t;
                    all = 0;
                }
            }
            if (all) break;
            if (cnt == k && bound > 0
                && (double)kd[k - 1] < bound * bound * (1.0 - 1e-5))
                break;
        }
    }
    free(start);
    free(cell_of);
    free(sorted);
    return 0;
}

/* ======================================================================
 * Dense LU with partial pivoting (Golub & Van Loan, Alg. 3.4.1) + solve.
 * ====================================================================== */
int NO_FN(lu_solve)(int n, int nrhs, double* A, double* B) {
    for (int j = 0; j < n; ++j) {
        int p = j;
        double best = fabs(A[j + (size_t)j * n]);
        for (int i = j + 1; i < n; ++i) {
            double v = fabs(A[i + (size_t)j * n]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        if (best == 0.0) return j + 1;
        if (p != j) {
            for (int m = 0; m < n; ++m) {
                double t = A[j + (size_t)m * n];
                A[j + (size_t)m * n] = A[p + (size_t)m * n];
                A[p + (size_t)m * n] = t;
            }
            for (int r = 0; r < nrhs; ++r) {
                double t = B[j + (size_t)r * n];
                B[j + (size_t)r * n] = B[p + (size_t)r * n];
                B[p + (size_t)r * n] = t;
            }
        }
        double piv = A[j + (size_t)j * n];
        for (int i = j + 1; i < n; ++i) {
            double l = A[i + (size_t)j * n] / piv;
            A[i + (size_t)j * n] = l;
            for (int m = j + 1; m < n; ++m)
                A[i + (size_t)m * n] -= l * A[j + (size_t)m * n];
            for (int r = 0; r < nrhs; ++r)
                B[i + (size_t)r * n] -= l * B[j + (size_t)r * n];
        }
    }
    for (int r = 0; r < nrhs; ++r) {
        double* b = B + (size_t)r * n;
        for (int j = n - 1; j >= 0; --j) {
            double s = b[j];
            for (int m = j + 1; m < n; ++m) s -= A[j + (size_t)m * n] * b[m];
            b[j] = s / A[j + (size_t)j * n];
        }
    }
    return 0;
}

/* The same elimination in FP32: the solve precision of the reference,
 * whose batched_matrix_inverse runs cuBLAS getrf/getri on T = float
 * (linalg/batched_matrix_inverse.cuh:229-250).  A and B are rounded to
 * FP32, solved, and the FP32 solution is returned in B. */
static int lu_solve_f32(int n, int nrhs, double* A, double* B) {
    float* a = (float*)malloc(sizeof(float) * (size_t)n * n);
    float* b = (float*)malloc(sizeof(float) * (size_t)n * nrhs);
    for (size_t t = 0; t < (size_t)n * n; ++t) a[t] = (float)A[t];
    for (size_t t = 0; t < (size_t)n * nrhs; ++t) b[t] = (float)B[t];
    int info = 0;
    for (int j = 0; j < n && !info; ++j) {
        int p = j;
        float best = fabsf(a[j + (size_t)j * n]);
        for (int i = j + 1; i < n; ++i) {
            float v = fabsf(a[i + (size_t)j * n]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        if (best == 0.0f) {
            info = j + 1;
            break;
        }
        if (p != j) {
            for (int m = 0; m < n; ++m) {
                float t = a[j + (size_t)m * n];
                a[j + (size_t)m * n] = a[p + (size_t)m * n];
                a[p + (size_t)m * n] = t;
            }
            for (int r = 0; r < nrhs; ++r) {
                float t = b[j + (size_t)r * n];
                b[j + (size_t)r * n] = b[p + (size_t)r * n];
                b[p + (size_t)r * n] = t;
            }
        }
        float piv = a[j + (size_t)j * n];
        for (int i = j + 1; i < n; ++i) {
            float l = a[i + (size_t)j * n] / piv;
            a[i + (size_t)j * n] = l;
            for (int m = j + 1; m < n; ++m) a[i + (size_t)m * n] -= l * a[j + (size_t)m * n];
            for (int r = 0; r < nrhs; ++r) b[i + (size_t)r * n] -= l * b[j + (size_t)r * n];
        }
    }
    if (!info)
        for (int r = 0; r < nrhs; ++r) {
            float* x = b + (size_t)r * n;
            for (int j = n - 1; j >= 0; --j) {
                float sum = x[j];
                for (int m = j + 1; m < n; ++m) sum -= a[j + (size_t)m * n] * x[m];
                x[j] = sum / a[j + (size_t)j * n];
            }
        }
    for (size_t t = 0; t < (size_t)n * nrhs; ++t) B[t] = (double)b[t];
    free(a);
    free(b);
    return info;
}

static int lu_solve_bits(int bits, int n, int nrhs, double* A, double* B) {
    return bits == 32 ? lu_solve_f32(n, nrhs, A, B) : NO_FN(lu_solve)(n, nrhs, A, B);
}

/* ======================================================================
 * RPIM (augmented RBF) weights.
 * ====================================================================== */
static double d2_f64(int D, const double* a, const double* b) {
    double s = (a[0] - b[0]) * (a[0] - b[0]);
    for (int d = 1; d < D; ++d) s = (a[d] - b[d]) * (a[d] - b[d]) + s;
    return s;
}

/* interpolation/radial_point_method.cuh:44-66 (multiquadric, q = 1/2,
 * alpha_c = 0.1) and the polyharmonic r^3 kernel. */
static double rbf_eval(int rbf, double r2, double spacing) {
    if (rbf == 1) return r2 * sqrt(r2);
    double c = spacing * 0.1;
    return pow(r2 + c * c, 0.5);
}

/* G0 = [[Phi, P], [P^T, 0]], P = [1, x(, y(, z))] in raw coordinates,
 * interpolation/detail/radial_point_method.cuh:349-398 */
static void assemble_G0(int D, const double* xs, int k, int rbf,
                        double spacing, double* G0) {
    int n = k + D + 1;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double v;
            if (i >= k && j >= k)
                v = 0.0;
            else if (i >= k)
                v = (i - k) == 0 ? 1.0 : xs[j * D + (i - k - 1)];
            else if (j >= k)
                v = (j - k) == 0 ? 1.0 : xs[i * D + (j - k - 1)];
            else
                v = rbf_eval(rbf, d2_f64(D, xs + i * D, xs + j * D), spacing);
            G0[i + (size_t)j * n] = v;
        }
}

/* G_x(q) = [phi(|x_s - x_q|^2); 1; x_q] (:400-433) */
static void assemble_Gx(int D, const double* xs, int k, const double* xq,
                        int rbf, double spacing, double* g) {
    for (int s = 0; s < k; ++s)
        g[s] = rbf_eval(rbf, d2_f64(D, xs + s * D, xq), spacing);
    g[k] = 1.0;
    for (int d = 0; d < D; ++d) g[k + 1 + d] = xq[d];
}

int NO_FN(rpim_weights)(int D, const float* src, const uint32_t* idx, int k,
                        const float* query, int64_t n_q, int rbf,
                        double spacing, double* w_out) {
    return NO_FN(rpim_weights_ex)(D, src, idx, k, query, n_q, rbf, spacing, 64, w_out);
}

int NO_FN(rpim_weights_ex)(int D, const float* src, const uint32_t* idx, int k,
                           const float* query, int64_t n_q, int rbf,
                           double spacing, int solve_bits, double* w_out) {
    int n = k + D + 1;
    int err = 0;
#pragma omp parallel
    {
        double* G0 = (double*)malloc(sizeof(double) * n * n);
        double* g = (double*)malloc(sizeof(double) * n);
        double xs[64 * 3], xq[3];
#pragma omp for schedule(dynamic, 64)
        for (int64_t q = 0; q < n_q; ++q) {
            for (int s = 0; s < k; ++s)
                for (int d = 0; d < D; ++d)
                    xs[s * D + d] = src[(int64_t)idx[q * k + s] * D + d];
            for (int d = 0; d < D; ++d) xq[d] = query[q * D + d];
            assemble_G0(D, xs, k, rbf, spacing, G0);
            assemble_Gx(D, xs, k, xq, rbf, spacing, g);
            /* w(s) = sum_i G0inv(i, s) g(i): solve G0^T y = g (:439-453) */
            for (int j = 0; j < n; ++j)
                for (int i = j + 1; i < n; ++i) {
                    double t = G0[i + (size_t)j * n];
                    G0[i + (size_t)j * n] = G0[j + (size_t)i * n];
                    G0[j + (size_t)i * n] = t;
                }
            if (lu_solve_bits(solve_bits, n, 1, G0, g)) {
#pragma omp atomic write
                err = 2;
            }
            for (int s = 0; s < k; ++s) w_out[q * k + s] = g[s];
        }
        free(G0);
        free(g);
    }
    return err;
}

/* ======================================================================
 * Quadrature points and nonlocal divergence weights.
 * ====================================================================== */
/* nonlocal_calculus/detail/quadrature_points.cuh:40-62 (6-digit constants
 * kept as published) */
static const double QR_W[4] = {0.652145, 0.652145, 0.347855, 0.347855};
static const double QR_X[4] = {0.339981, -0.339981, 0.861136, -0.861136};

/* :74-91 and :127-152 */
static void unscaled_quad_point(int D, int q, double* x) {
    int r_idx = q % 4;
    double r = (QR_X[r_idx] + 1.0) / 2.0;
    double theta_scale = 3.0 * M_PI / 2.0 / 3.0;
    if (D == 2) {
        int t_idx = q / 4;
        double th = t_idx * theta_scale;
        x[0] = r * cos(th);
        x[1] = r * sin(th);
    } else {
        int t_idx = (q / 4) % 4;
        int p_idx = q / 16;
        double th = t_idx * theta_scale;
        double phi_scale = 3.0 * M_PI / 2.0 / 3.0 / 2.0;
        double ph = p_idx * phi_scale + M_PI / 4.0 / 2.0;
        x[0] = r * sin(ph) * cos(th);
        x[1] = r * sin(ph) * sin(th);
        x[2] = r * cos(ph);
    }
}

/* antisymmetric w4 B-spline kernel, detail/divergence.inl:55-133 */
static void div_kernel(int D, double delta, const double* x, const double* y,
                       double* a) {
    double xy[3];
    for (int d = 0; d < D; ++d) xy[d] = y[d] - x[d];
    if (D == 2) {
        double r = sqrt(xy[1] * xy[1] + xy[0] * xy[0]);
        double ct = xy[0] / r, st = xy[1] / r;
        double C = (240.0 / 7.0) / (M_PI * pow(delta, 5));
        double g;
        if (2.0 * r / delta < 1)
            g = r * (2.0 * delta - 3.0 * r);
        else if (2.0 * r / delta < 2)
            g = (delta - r) * (delta - r);
        else
            g = 0.0;
        a[0] = C * g * ct;
        a[1] = C * g * st;
    } else {
        double rxy = sqrt(xy[1] * xy[1] + xy[0] * xy[0]);
        double r = sqrt(xy[2] * xy[2] + (xy[1] * xy[1] + xy[0] * xy[0]));
        double sp = rxy / r, cp = xy[2] / r;
        double ct = rxy != 0 ? xy[0] / rxy : 0, st = rxy != 0 ? xy[1] / rxy : 0;
        double C = 48.0 / (M_PI * pow(delta, 6));
        double g;
        if (2.0 * r / delta < 1)
            g = r * (2.0 * delta - 3.0 * r);
        else if (2.0 * r / delta < 2)
            g = (delta - r) * (delta - r);
        else
            g = 0.0;
        a[0] = C * g * sp * ct;
        a[1] = C * g * sp * st;
        a[2] = C * g * cp;
    }
}

int NO_FN(divergence_weights)(int D, const float* src, int64_t n_src,
                              const float* query, int64_t n_q,
                              const uint32_t* idx, const float* d2, int k,
                              int rbf, const unsigned char* kvar, double* W_out,
                              double* delta_out) {
    return NO_FN(divergence_weights_ex)(D, src, n_src, query, n_q, idx, d2, k, rbf, kvar, 64,
                                        W_out, delta_out);
}

int NO_FN(divergence_weights_ex)(int D, const float* src, int64_t n_src,
                                 const float* query, int64_t n_q,
                                 const uint32_t* idx, const float* d2, int k,
                                 int rbf, const unsigned char* kvar, int solve_bits,
                                 double* W_out, double* delta_out) {
    (void)n_src;
    const int kmax = k;
    const int nmax = kmax + D + 1;
    const int Nq = D == 2 ? 16 : 64;
    /* operators.cuh:84-117: delta = 0.05 * nearest-neighbour distance; the
     * multiquadric shape parameter uses the mean nearest distance over each
     * divergence_operator::create batch of 16384 queries
     * (nonlocal_calculus/divergence.cuh:118-160). */
    const int64_t batch = 16384;
    int64_t nbatch = (n_q + batch - 1) / batch;
    double* hbar = (double*)malloc(sizeof(double) * (nbatch ? nbatch : 1));
    for (int64_t b = 0; b < nbatch; ++b) {
        int64_t e = (b + 1) * batch < n_q ? (b + 1) * batch : n_q;
        double s = 0;
        for (int64_t i = b * batch; i < e; ++i) s += sqrt((double)d2[i * kmax + 1]);
        hbar[b] = s / (double)(e - b * batch);
    }
    int err = 0;
#pragma omp parallel
    {
        double* G0 = (double*)malloc(sizeof(double) * nmax * nmax);
        double* B = (double*)malloc(sizeof(double) * nmax * Nq);
        double xs[64 * 3];
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < n_q; ++i) {
            /* per-query stencil width (variable stencils): the first k of
             * the kmax neighbours, zero weights after them */
            const int k = kvar ? (int)kvar[i] : kmax;
            const int n = k + D + 1;
            double xi[3], xq[64][3];
            for (int d = 0; d < D; ++d) xi[d] = query[i * D + d];
            for (int s = 0; s < k; ++s)
                for (int d = 0; d < D; ++d)
                    xs[s * D + d] = src[(int64_t)idx[i * kmax + s] * D + d];
            double delta = 0.05 * sqrt((double)d2[i * kmax + 1]);
            double spacing = hbar[i / batch];
            if (delta_out) delta_out[i] = delta;
            /* quadrature_point_map (quadrature_points.cuh:238-263) */
            for (int q = 0; q < Nq; ++q) {
                unscaled_quad_point(D, q, xq[q]);
                for (int d = 0; d < D; ++d) xq[q][d] = xq[q][d] * delta + xi[d];
            }
            assemble_G0(D, xs, k, rbf, spacing, G0);
            for (int q = 0; q < Nq; ++q)
                assemble_Gx(D, xs, k, xq[q], rbf, spacing, B + (size_t)q * n);
            for (int j = 0; j < n; ++j)
                for (int r = j + 1; r < n; ++r) {
                    double t = G0[r + (size_t)j * n];
                    G0[r + (size_t)j * n] = G0[j + (size_t)r * n];
                    G0[j + (size_t)r * n] = t;
                }
            if (lu_solve_bits(solve_bits, n, Nq, G0, B)) {
#pragma omp atomic write
                err = 2;
            }
            /* compute_divergence_weights (divergence.inl:161-222) */
            double* Wi = W_out + i * (int64_t)kmax * D;
            for (int t = 0; t < kmax * D; ++t) Wi[t] = 0.0;
            for (int q = 0; q < Nq; ++q) {
                double a[3];
                div_kernel(D, delta, xi, xq[q], a);
                double qw = QR_W[q % 4] * delta / 2.0;
                qw *= 2.0 * M_PI / 4.0;
                if (D == 3) qw *= M_PI / 4.0;
                double dx[3];
                for (int d = 0; d < D; ++d) dx[d] = xq[q][d] - xi[d];
                if (D == 2) {
                    qw *= sqrt(dx[1] * dx[1] + dx[0] * dx[0]);
                } else {
                    double r = sqrt(dx[2] * dx[2] + (dx[1] * dx[1] + dx[0] * dx[0]));
                    double rxy = sqrt(dx[1] * dx[1] + dx[0] * dx[0]);
                    qw *= r * rxy;
                }
                for (int t = 0; t < k * D; ++t) {
                    int d = t % D, s = t / D;
                    Wi[s * D + d] += a[d] * qw
                                     * (B[(size_t)q * n + s] + (s == 0 ? 1.0 : 0.0));
                }
            }
        }
        free(G0);
        free(B);
    }
    free(hbar);
    return err;
}

/* tree reduction over the stencil (divergence.cuh:468-478) */
static no_real tree_sum(no_real* v, int k) {
    for (int s = 1; s < k; s *= 2)
        for (int t = 0; 2 * s * t + s < k; ++t) v[2 * s * t] += v[2 * s * t + s];
    return v[0];
}

int NO_FN(divergence_apply)(int D, int64_t n, int k, const uint32_t* idx,
                            const double* W, const double* F, double off,
                            double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        no_real v[128];
        for (int s = 0; s < k; ++s) {
            const double* Fj = F + (int64_t)idx[i * k + s] * D;
            no_real acc = 0;
            for (int d = 0; d < D; ++d)
                acc += (no_real)W[(i * k + s) * D + d]
                       * ((no_real)Fj[d] - (no_real)off);
            v[s] = acc;
        }
        out[i] = (double)tree_sum(v, k);
    }
    return 0;
}

/* ======================================================================
 * The engine (detail/simulation_engine.cuh:85-1486).
 * ====================================================================== */
struct no_engine {
    no_config cfg;
    int Q, D, k, kb;
    int64_t N, nbulk, nlayer, nbnd;
    float* pts;
    float* normals;
    no_real* sigma;       /* per layer node */
    uint32_t* idx;        /* N*k */
    no_real* W;           /* N*k*D */
    double* W64;          /* as computed */
    no_real* filt;        /* N*k normalised filter weights */
    uint32_t* bidx;       /* nbnd*kb into [0, bulk+layer) */
    no_real* bw;          /* nbnd*kb */
    no_real *f, *s1, *s2, *f0, *df, *rates, *Q1, *Q2;
    no_real *rho, *u;
    no_real c[27][3], w[27];
    int bb[27];
    no_real rho0, vel_scale, dt_lat, nu_lat, dt;
    double time;
    int nsrc;
    double src[NO_MAX_SOURCES][9]; /* centre[3], radius, amplitude, frequency,
                                      phase, t_start, t_end */
};

static void eng_macroscopic(no_engine* e, const no_real* f) {
    /* simulation_engine.cuh:53-82, 366-433 */
    const int Q = e->Q, D = e->D;
    const int64_t N = e->N;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        no_real rho = 0, u[3] = {0, 0, 0};
        for (int a = 0; a < Q; ++a) {
            no_real fa = f[a * N + i];
            rho += fa;
            for (int d = 0; d < D; ++d) u[d] += fa * e->c[a][d];
        }
        for (int d = 0; d < D; ++d) u[d] /= rho;
        e->rho[i] = rho * e->rho0;
        for (int d = 0; d < D; ++d)
            e->u[d * N + i] = u[d] * e->vel_scale;
    }
}

/* Density sources: the deterministic restatement of the density_source
 * plugin path (simulation_engine.cuh:347-364 computes the per-node source
 * terms at the step's start time, :1057-1107 applies them).  Source j adds
 *   s_j(x, t) = A_j cos(2 pi f_j t + phi_j) P(|x - c_j| / R_j)
 * for |x - c_j| < R_j and t_start <= t < t_end, with the radial profile of
 * the reference example's source, P(r) = (1 - (2r)^2 / 4) exp(-(2r)^2)
 * (examples/nonlocal_lattice_boltzmann.cu:97-106).  Where the summed source
 * s is non-zero: f_a += w_a s / rho0 and rho <- (rho + (rho + s)) / 2,
 * once per node (the reference repeats the density update once per
 * distribution without synchronisation); u is unchanged. */
static double src_profile(double r, double R) {
    const double x = 2.0 * r / R;
    return (1.0 - x * x / 4.0) * exp(-x * x);
}

static void eng_sources(no_engine* e) {
    if (!e->nsrc) return;
    const int Q = e->Q, D = e->D;
    const int64_t N = e->N;
    const double t = e->time;
    double amp[NO_MAX_SOURCES];
    for (int j = 0; j < e->nsrc; ++j) {
        const double* S = e->src[j];
        amp[j] = (t >= S[7] && t < S[8]) ? S[4] * cos(2.0 * M_PI * S[5] * t + S[6]) : 0.0;
    }
    for (int64_t i = 0; i < N; ++i) {
        double s = 0.0;
        int hit = 0;
        for (int j = 0; j < e->nsrc; ++j) {
            const double* S = e->src[j];
            double r2 = 0.0;
            for (int d = 0; d < D; ++d) {
                const double dx = (double)e->pts[i * D + d] - S[d];
                r2 += dx * dx;
            }
            const double r = sqrt(r2);
            if (r < S[3]) {
                s += src_profile(r, S[3]) * amp[j];
                hit = 1;
            }
        }
        if (!hit || s == 0.0) continue;
        for (int a = 0; a < Q; ++a)
            e->f[a * N + i] += (no_real)(lat_w(Q, a) * s / (double)e->rho0);
        e->rho[i] = (no_real)(((double)e->rho[i] + ((double)e->rho[i] + s)) / 2.0);
    }
}

int NO_FN(engine_add_density_source)(no_engine* e, const double* center, double radius,
                                     double amplitude, double frequency, double phase,
                                     double t_start, double t_end) {
    if (e->nsrc >= NO_MAX_SOURCES || !(radius > 0)) return 1;
    double* S = e->src[e->nsrc++];
    for (int d = 0; d < 3; ++d) S[d] = d < e->D ? center[d] : 0.0;
    S[3] = radius;
    S[4] = amplitude;
    S[5] = frequency;
    S[6] = phase;
    S[7] = t_start;
    S[8] = t_end;
    return 0;
}

int NO_FN(engine_clear_density_sources)(no_engine* e) {
    e->nsrc = 0;
    return 0;
}

/* divergence of F_d(j) = c_d * (-(f(j) - w)) at node i (pde_right_hand_side
 * field map, nonlocal_calculus/advection.cuh:117-126) */
static no_real eng_div_stream(const no_engine* e, int a, const no_real* fa,
                              int64_t i) {
    const int k = e->k, D = e->D;
    no_real v[128];
    for (int s = 0; s < k; ++s) {
        int64_t j = e->idx[i * k + s];
        no_real sc = -(fa[j] - e->w[a]);
        no_real acc = 0;
        for (int d = 0; d < D; ++d) {
            no_real Fd = e->c[a][d] * sc;
            acc += e->W[(i * k + s) * D + d] * (Fd - (no_real)0);
        }
        v[s] = acc;
    }
    return tree_sum(v, k);
}

/* PML layer field (pml_absorption.inl:440-453 and :500-513) */
static no_real eng_div_pml(const no_engine* e, int a, const no_real* Qa,
                           no_real mean, int second, int64_t i) {
    const int k = e->k, D = e->D;
    const int64_t lb = e->nbulk, le = e->nbulk + e->nlayer;
    no_real v[128];
    for (int s = 0; s < k; ++s) {
        int64_t j = e->idx[i * k + s];
        no_real acc = 0;
        if (j >= lb && j < le) {
            int64_t p = j - lb;
            no_real sig = e->sigma[p];
            no_real qv = e->cfg.pml_ref_index ? Qa[p] : Qa[j];
            for (int d = 0; d < D; ++d) {
                no_real Fd = second ? e->c[a][d] * sig * sig * (qv - mean)
                                    : (no_real)-2 * e->c[a][d] * sig * (qv - mean);
                acc += e->W[(i * k + s) * D + d] * (Fd - (no_real)0);
            }
        }
        v[s] = acc;
    }
    return tree_sum(v, k);
}

static void eng_pml(no_engine* e) {
    const int Q = e->Q, D = e->D;
    const int64_t N = e->N, lb = e->nbulk, le = e->nbulk + e->nlayer;
    /* compute_Qvalues (pml_absorption.inl:551-631) */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        no_real rh = e->rho[i] / e->rho0, uh[3];
        for (int d = 0; d < D; ++d) uh[d] = e->u[d * N + i] / e->vel_scale;
        for (int a = 0; a < Q; ++a) {
            no_real fte = feq_impl(Q, rh, uh, a) - e->w[a];
            no_real prev = e->Q1[a * N + i] - e->w[a];
            no_real q1 = (fte + prev) * e->dt_lat / (no_real)2;
            e->Q1[a * N + i] = q1;
            no_real q2 = (q1 + e->Q2[a * N + i]) * e->dt_lat / (no_real)2;
            e->Q2[a * N + i] = q2;
            if (i < lb || i >= le) continue;
            no_real sig = e->sigma[i - lb];
            e->s1[a * N + i] -= sig * ((no_real)3 * fte + (no_real)3 * sig * q1
                                       + (sig * sig) * q2);
        }
    }
    /* layer means and divergence of the two auxiliary fields (:770-855) */
    for (int a = 0; a < Q; ++a) {
        for (int second = 0; second < 2; ++second) {
            const no_real* Qa = (second ? e->Q2 : e->Q1) + a * N;
            no_real sum = 0;
            for (int64_t i = lb; i < le; ++i) sum += Qa[i];
            no_real mean = e->nlayer ? sum / (no_real)e->nlayer : (no_real)0;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < N; ++i) e->s2[a * N + i] = eng_div_pml(e, a, Qa, mean, second, i);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < N; ++i)
                e->s1[a * N + i] = e->s1[a * N + i] + (no_real)1 * e->s2[a * N + i];
        }
    }
    /* post_collision (:875-896): Q2 <- Q1, Q1 <- feq */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        no_real rh = e->rho[i] / e->rho0, uh[3];
        for (int d = 0; d < D; ++d) uh[d] = e->u[d * N + i] / e->vel_scale;
        for (int a = 0; a < Q; ++a) {
            e->Q2[a * N + i] = e->Q1[a * N + i];
            e->Q1[a * N + i] = feq_impl(Q, rh, uh, a);
        }
    }
}

/* simulation_engine.cuh:434-467 -> radial_point_method.cuh:230-250 */
static void eng_interp_boundary(no_engine* e, no_real* f) {
    const int Q = e->Q, kb = e->kb;
    const int64_t N = e->N, b0 = e->nbulk + e->nlayer;
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < e->nbnd; ++b)
        for (int a = 0; a < Q; ++a) {
            no_real sum = 0;
            for (int s = 0; s < kb; ++s)
                sum += e->bw[b * kb + s]
                       * (f[a * N + e->bidx[b * kb + s]] - e->w[a]);
            f[a * N + b0 + b] = sum + e->w[a];
        }
}

/* simulation_engine.cuh:469-531 */
static void eng_collision(no_engine* e) {
    const int Q = e->Q, D = e->D;
    const int64_t N = e->N;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        no_real fi[27], df[27], uh[3];
        for (int a = 0; a < Q; ++a) fi[a] = e->f[a * N + i];
        no_real rh = e->rho[i] / e->rho0;
        for (int d = 0; d < D; ++d)
            uh[d] = e->u[d * N + i] * ((no_real)CS_LAT / (no_real)e->cfg.c);
        collide_impl(Q, fi, rh, uh, e->nu_lat, e->dt_lat, df);
        for (int a = 0; a < Q; ++a) e->s1[a * N + i] += df[a];
    }
}

/* simulation_engine.cuh:1233-1323 with the weights normalised once */
static void eng_filter(no_engine* e) {
    const int Q = e->Q, k = e->k;
    const int64_t N = e->N, be = e->nbulk + e->nlayer;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i)
        for (int a = 0; a < Q; ++a) {
            const no_real* s1 = e->s1 + a * N;
            if (i >= be) {
                e->s2[a * N + i] = s1[i];
                continue;
            }
            no_real fi = s1[i], sum = 0;
            for (int s = 0; s < k; ++s)
                sum += e->filt[i * k + s] * (s1[e->idx[i * k + s]] - fi);
            e->s2[a * N + i] = fi + (no_real)e->cfg.filter * sum;
        }
}

/* simulation_engine.cuh:789-944 */
static void eng_bounce_back(no_engine* e) {
    const int Q = e->Q, D = e->D;
    const int64_t N = e->N, b0 = e->nbulk + e->nlayer;
    for (int64_t b = 0; b < e->nbnd; ++b) {
        int64_t i = b0 + b;
        for (int d = 0; d < D; ++d) e->u[d * N + i] = 0;
        for (int a = 0; a < Q; ++a) {
            no_real cn = 0;
            for (int d = 0; d < D; ++d)
                cn += e->c[a][d] * (no_real)e->normals[b * D + d];
            if (cn <= 0) continue;
            e->f[a * N + i] = e->f[e->bb[a] * N + i];
        }
    }
}

/* pde_right_hand_side<...>::operator() (pde_right_hand_side.cuh:82-176),
 * completed for D2Q9 by analogy */
static void eng_rhs(no_engine* e, const no_real* f) {
    const int Q = e->Q;
    const int64_t N = e->N, be = e->nbulk + e->nlayer;
    for (int a = 0; a < Q; ++a) {
        const no_real* fa = f + a * N;
        no_real* ra = e->rates + a * N;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) ra[i] = i < be ? eng_div_stream(e, a, fa, i) : (no_real)0;
    }
}

/* rk4_solver::step_forward (rk4_solver.cuh:48-171) */
static void eng_rk4(no_engine* e) {
    const int Q = e->Q;
    const int64_t QN = (int64_t)Q * e->N;
    const no_real dt = e->dt_lat;
    const no_real a[3] = {(no_real)(1. / 2.), (no_real)(1. / 2.), (no_real)1.};
    const no_real b[4] = {(no_real)(1. / 6.), (no_real)(2. / 6.),
                          (no_real)(2. / 6.), (no_real)(1. / 6.)};
    memcpy(e->f0, e->f, sizeof(no_real) * QN);
    for (int st = 0; st < 4; ++st) {
        eng_rhs(e, e->f);
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < QN; ++t) {
            no_real kk = e->rates[t];
            if (st < 3) e->f[t] = e->f0[t] + (dt * a[st]) * kk;
            kk = kk * b[st];
            e->df[t] = st == 0 ? kk : e->df[t] + kk;
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < QN; ++t) e->f[t] = e->f0[t] + dt * e->df[t];
}

no_engine* NO_FN(engine_create)(const no_config* cfg, const float* pts,
                                int64_t n_bulk, int64_t n_layer,
                                int64_t n_boundary, const float* normals,
                                const float* sigma) {
    if (!cfg || (cfg->Q != 9 && cfg->Q != 27)) return NULL;
    int D = lat_dims(cfg->Q);
    int64_t N = n_bulk + n_layer + n_boundary;
    int k = cfg->k, kb = cfg->k_boundary;
    if (k < 2 || k > 128 || kb < 1 || kb > 64 || N < k
        || n_bulk + n_layer < kb)
        return NULL;
#ifdef _OPENMP
    if (cfg->threads > 0) omp_set_num_threads(cfg->threads);
#endif
    no_engine* e = (no_engine*)calloc(1, sizeof(no_engine));
    e->cfg = *cfg;
    e->Q = cfg->Q;
    e->D = D;
    e->k = k;
    e->kb = kb;
    e->N = N;
    e->nbulk = n_bulk;
    e->nlayer = n_layer;
    e->nbnd = n_boundary;
    e->pts = (float*)malloc(sizeof(float) * N * D);
    memcpy(e->pts, pts, sizeof(float) * N * D);
    e->normals = (float*)malloc(sizeof(float) * (n_boundary ? n_boundary : 1) * D);
    if (n_boundary) memcpy(e->normals, normals, sizeof(float) * n_boundary * D);
    e->sigma = (no_real*)malloc(sizeof(no_real) * (n_layer ? n_layer : 1));
    for (int64_t t = 0; t < n_layer; ++t) e->sigma[t] = (no_real)sigma[t];
    for (int a = 0; a < e->Q; ++a) {
        for (int d = 0; d < D; ++d) e->c[a][d] = (no_real)lat_c(e->Q, a, d);
        e->w[a] = (no_real)lat_w(e->Q, a);
        e->bb[a] = lat_bb(e->Q, a);
    }
    /* set_problem_parameters (simulation_engine.cuh:94-124) */
    no_real ratio = (no_real)CS_LAT / (no_real)cfg->c;
    e->rho0 = (no_real)cfg->rho0;
    e->dt = (no_real)cfg->dt;
    e->dt_lat = (no_real)cfg->dt / ratio;
    e->nu_lat = (no_real)cfg->nu / ratio;
    e->vel_scale = (no_real)cfg->c / (no_real)CS_LAT;

    /* divergence operator over all nodes (init_domain, :169-181) */
    e->idx = (uint32_t*)malloc(sizeof(uint32_t) * N * k);
    float* d2 = (float*)malloc(sizeof(float) * N * k);
    NO_FN(knn)(D, pts, N, pts, N, k, d2, e->idx);
    e->W64 = (double*)malloc(sizeof(double) * N * k * D);
    const int bits = cfg->solve_bits == 32 ? 32 : 64;
    int err = NO_FN(divergence_weights_ex)(D, pts, N, pts, N, e->idx, d2, k,
                                           cfg->rbf, cfg->kvar, bits, e->W64, NULL);
    e->W = (no_real*)malloc(sizeof(no_real) * N * k * D);
    for (int64_t t = 0; t < N * k * D; ++t) e->W[t] = (no_real)e->W64[t];
    /* filter weights: Gaussian of sigma_g = h/2.5 over the divergence
     * stencil (:1237-1315), normalised once */
    e->filt = (no_real*)malloc(sizeof(no_real) * N * k);
    {
        double sg = cfg->h / 2.5;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            double gsum = 0, g[128], xi[3], xj[3];
            const int ki = cfg->kvar ? (int)cfg->kvar[i] : k;
            for (int d = 0; d < D; ++d) xi[d] = pts[i * D + d];
            for (int s = 0; s < ki; ++s) {
                for (int d = 0; d < D; ++d)
                    xj[d] = pts[(int64_t)e->idx[i * k + s] * D + d];
                double r = sqrt(d2_f64(D, xi, xj));
                g[s] = exp(-0.5 * (r / sg) * (r / sg));
                gsum += g[s];
            }
            for (int s = 0; s < k; ++s)
                e->filt[i * k + s] = s < ki ? (no_real)(g[s] / gsum) : (no_real)0;
        }
    }
    free(d2);
    /* boundary interpolator (:223-258) */
    int64_t nsrc = n_bulk + n_layer;
    e->bidx = (uint32_t*)malloc(sizeof(uint32_t) * (n_boundary ? n_boundary : 1) * kb);
    e->bw = (no_real*)malloc(sizeof(no_real) * (n_boundary ? n_boundary : 1) * kb);
    if (n_boundary) {
        float* bd2 = (float*)malloc(sizeof(float) * n_boundary * kb);
        double* bw64 = (double*)malloc(sizeof(double) * n_boundary * kb);
        NO_FN(knn)(D, pts, nsrc, pts + nsrc * D, n_boundary, kb, bd2, e->bidx);
        err |= NO_FN(rpim_weights_ex)(D, pts, e->bidx, kb, pts + nsrc * D, n_boundary,
                                      cfg->rbf_boundary, cfg->h, bits, bw64);
        for (int64_t t = 0; t < n_boundary * kb; ++t) e->bw[t] = (no_real)bw64[t];
        free(bd2);
        free(bw64);
    }
    int64_t QN = (int64_t)e->Q * N;
    no_real** bufs[] = {&e->f, &e->s1, &e->s2, &e->f0, &e->df, &e->rates, &e->Q1, &e->Q2};
    for (size_t t = 0; t < sizeof(bufs) / sizeof(bufs[0]); ++t)
        *bufs[t] = (no_real*)calloc(QN, sizeof(no_real));
    e->rho = (no_real*)calloc(N, sizeof(no_real));
    e->u = (no_real*)calloc(N * D, sizeof(no_real));
    /* pml_absorption_operator constructor (:744-768) and reset (:1377-1431) */
    for (int a = 0; a < e->Q; ++a)
        for (int64_t i = 0; i < N; ++i) {
            e->Q1[a * N + i] = e->w[a];
            e->f[a * N + i] = e->w[a];
        }
    eng_macroscopic(e, e->f);
    e->time = 0;
    if (err) {
        NO_FN(engine_destroy)(e);
        return NULL;
    }
    return e;
}

void NO_FN(engine_destroy)(no_engine* e) {
    if (!e) return;
    void* ptrs[] = {e->pts, e->normals, e->sigma, e->idx, e->W, e->W64,
                    e->filt, e->bidx, e->bw, e->f, e->s1, e->s2, e->f0,
                    e->df, e->rates, e->Q1, e->Q2, e->rho, e->u};
    for (size_t t = 0; t < sizeof(ptrs) / sizeof(ptrs[0]); ++t) free(ptrs[t]);
    free(e);
}

int NO_FN(engine_set_pulse)(no_engine* e, const double* center, double amp,
                            double width) {
    const int64_t N = e->N;
    for (int64_t i = 0; i < N; ++i) {
        double r2 = 0;
        for (int d = 0; d < e->D; ++d) {
            double t = (double)e->pts[i * e->D + d] - center[d];
            r2 += t * t;
        }
        double p = amp * exp(-r2 / (2.0 * width * width));
        for (int a = 0; a < e->Q; ++a)
            e->f[a * N + i] = (no_real)(lat_w(e->Q, a) * (1.0 + p));
    }
    eng_macroscopic(e, e->f);
    e->time = 0.0;
    return 0;
}

int NO_FN(engine_set_distributions)(no_engine* e, const double* f) {
    for (int64_t t = 0; t < (int64_t)e->Q * e->N; ++t) e->f[t] = (no_real)f[t];
    eng_macroscopic(e, e->f);
    return 0;
}

/* step_forward (simulation_engine.cuh:1325-1375) */
int NO_FN(engine_step)(no_engine* e, int nsteps) {
    const int64_t QN = (int64_t)e->Q * e->N;
    for (int n = 0; n < nsteps; ++n) {
        eng_macroscopic(e, e->f);
        eng_sources(e);
        memcpy(e->s1, e->f, sizeof(no_real) * QN);
        eng_pml(e);
        eng_interp_boundary(e, e->s1);
        eng_collision(e);
        eng_filter(e);
        no_real* t = e->f;
        e->f = e->s2;
        e->s2 = t;
        eng_bounce_back(e);
        eng_rk4(e);
        eng_interp_boundary(e, e->f);
        e->time += (double)e->dt;
    }
    return 0;
}

int NO_FN(engine_get)(no_engine* e, double* f, double* rho, double* u,
                      double* drho, int fresh) {
    const int Q = e->Q, D = e->D;
    const int64_t N = e->N;
    if (f)
        for (int64_t t = 0; t < (int64_t)Q * N; ++t) f[t] = (double)e->f[t];
    for (int64_t i = 0; i < N; ++i) {
        if (fresh) {
            no_real r = 0, uu[3] = {0, 0, 0};
            double dr = 0;
            for (int a = 0; a < Q; ++a) {
                no_real fa = e->f[a * N + i];
                r += fa;
                dr += (double)fa - lat_w(Q, a);
                for (int d = 0; d < D; ++d) uu[d] += fa * e->c[a][d];
            }
            if (rho) rho[i] = (double)(r * e->rho0);
            if (drho) drho[i] = dr * (double)e->rho0;
            for (int d = 0; d < D; ++d)
                if (u) u[d * N + i] = (double)(uu[d] / r * e->vel_scale);
        } else {
            if (rho) rho[i] = (double)e->rho[i];
            if (drho) drho[i] = (double)e->rho[i] - (double)e->rho0;
            for (int d = 0; d < D; ++d)
                if (u) u[d * N + i] = (double)e->u[d * N + i];
        }
    }
    return 0;
}

int NO_FN(engine_get_setup)(no_engine* e, uint32_t* idx, double* W,
                            double* filt, uint32_t* bidx, double* bw) {
    const int64_t N = e->N;
    if (idx) memcpy(idx, e->idx, sizeof(uint32_t) * N * e->k);
    if (W) memcpy(W, e->W64, sizeof(double) * N * e->k * e->D);
    if (filt)
        for (int64_t t = 0; t < N * e->k; ++t) filt[t] = (double)e->filt[t];
    if (bidx) memcpy(bidx, e->bidx, sizeof(uint32_t) * e->nbnd * e->kb);
    if (bw)
        for (int64_t t = 0; t < e->nbnd * e->kb; ++t) bw[t] = (double)e->bw[t];
    return 0;
}

int NO_FN(engine_get_pml)(no_engine* e, double* Q1, double* Q2) {
    for (int64_t t = 0; t < (int64_t)e->Q * e->N; ++t) {
        if (Q1) Q1[t] = (double)e->Q1[t];
        if (Q2) Q2[t] = (double)e->Q2[t];
    }
    return 0;
}

int NO_FN(engine_set_pml)(no_engine* e, const double* Q1, const double* Q2) {
    for (int64_t t = 0; t < (int64_t)e->Q * e->N; ++t) {
        if (Q1) e->Q1[t] = (no_real)Q1[t];
        if (Q2) e->Q2[t] = (no_real)Q2[t];
    }
    return 0;
}

int NO_FN(engine_pml_stage)(no_engine* e, const double* s1_in, double* s1_out) {
    const int64_t QN = (int64_t)e->Q * e->N;
    eng_macroscopic(e, e->f);
    for (int64_t t = 0; t < QN; ++t) e->s1[t] = (no_real)s1_in[t];
    eng_pml(e);
    for (int64_t t = 0; t < QN; ++t) s1_out[t] = (double)e->s1[t];
    return 0;
}

double NO_FN(engine_time)(no_engine* e) { return e->time; }
