/*
 * naga_oracle.h -- CPU restatement of NAGA's nonlocal lattice-Boltzmann
 * acoustics step and its setup.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against, and the CPU baseline that bench.py times.
 * It is never linked into, loaded by, or called from the product path.
 *
 * One source file is compiled twice:
 *   -DNO_REAL=double -DNO_PFX=o64_   -> libnaga_oracle_f64.so (the FP64 oracle)
 *   -DNO_REAL=float  -DNO_PFX=o32_   -> libnaga_oracle_f32.so (FP32 CPU port)
 * Setup arithmetic (RBF assembly, LU solves, quadrature, weights) is FP64
 * (the RBF solves optionally FP32, as the reference's); the stepping state
 * and arithmetic use NO_REAL.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the NAGA source tree).
 */
#ifndef NAGA_ORACLE_H
#define NAGA_ORACLE_H

#include <stdint.h>

#ifndef NO_REAL
#define NO_REAL double
#endif
#ifndef NO_PFX
#define NO_PFX o64_
#endif
#define NO_CAT2(a, b) a##b
#define NO_CAT(a, b) NO_CAT2(a, b)
#define NO_FN(name) NO_CAT(NO_PFX, name)

typedef NO_REAL no_real;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- lattices ---------------------------------------------------------- */
/* velocities c[Q*D] (alpha-major), weights w[Q], bounce-back map bb[Q],
 * collision matrix K[Q*Q] (row l, column j).  Q = 9 or 27. */
int NO_FN(lattice_tables)(int Q, double* c, double* w, int* bb, double* K);

/* Cascaded central-moment collision increment for one node:
 * df[l] = dt * sum_j K[l][j] k[j],  k = moment projection of f.
 * rho, u are unitless (rho = sum f, u = sum f c / sum f).
 * detail/d3q27_lattice.cuh:125-378, detail/d2q9_lattice.cuh:120-299,
 * detail/simulation_engine.cuh:469-531. */
int NO_FN(collide)(int Q, const no_real* f, no_real rho, const no_real* u,
                   no_real nu_lat, no_real dt_lat, no_real* df);

/* Equilibrium distribution for one velocity (the reference's form,
 * detail/equilibrium_distribution.cuh:41-64). */
no_real NO_FN(feq)(int Q, no_real rho, const no_real* u, int alpha);

/* ---- RK4 on a scalar linear ODE y' = lambda*y (pins the RK4 combination
 * used by the engine; detail/rk4_solver.cuh:48-171) --------------------- */
double NO_FN(rk4_scalar)(double y0, double lambda, double dt, int nsteps);

/* ---- synthetic node provider -------------------------------------------
 * Uniform grid on [lo, hi]^D (node_providers/uniform_grid_provider.cuh:65-161,
 * ranges/uniform_grid.cuh:55-97,238-294) with seeded SplitMix64 jitter of the
 * interior nodes, PML layer by 1-NN distance to the boundary nodes and the
 * build's deterministic [bulk | layer | boundary] order (Hilbert key of the
 * grid index within each class).
 * counts[0..2] = bulk, layer, boundary.  Output arrays sized by the caller
 * after naga grid size query (grid_size). */
int64_t NO_FN(grid_size)(int D, const double* lo, const double* hi, double h,
                         int64_t* dims_out);
int NO_FN(uniform_grid)(int D, const double* lo, const double* hi, double h,
                        double jitter_frac, uint64_t seed, double layer_L,
                        double sigma_max, float* pts, float* normals,
                        float* sigma, int64_t* counts, int64_t* grid_index);

/* ---- exact k-nearest neighbours ----------------------------------------
 * Key (d2_fp32, index) ascending, d2 = dz^2 + (dy^2 + dx^2) in FP32 with no
 * contraction (distance_functions.hpp:151-160); segmentation/
 * nearest_neighbors.cuh:95-477 visits cells in random order and stops one
 * shell after the fill: the build fixes it to the exact k smallest keys. */
int NO_FN(knn)(int D, const float* ref, int64_t n_ref, const float* query,
               int64_t n_q, int k, float* d2_out, uint32_t* idx_out);
int NO_FN(knn_brute)(int D, const float* ref, int64_t n_ref,
                     const float* query, int64_t n_q, int k, float* d2_out,
                     uint32_t* idx_out);

/* ---- dense LU with partial pivoting, FP64 (replaces cuBLAS getrf/getri,
 * linalg/batched_matrix_inverse.cuh:229-250).  A column-major n x n,
 * B column-major n x nrhs, overwritten by the solution.  Returns 0 or the
 * 1-based index of the first zero pivot. */
int NO_FN(lu_solve)(int n, int nrhs, double* A, double* B);

/* ---- RPIM weights -------------------------------------------------------
 * rbf: 0 = multiquadric sqrt(r^2 + (0.1*spacing)^2), 1 = polyharmonic r^3.
 * interpolation/detail/radial_point_method.cuh:92-472,
 * interpolation/radial_point_method.cuh:44-66.
 * Weights w[q*k + s] for query q (group size 1). */
int NO_FN(rpim_weights)(int D, const float* src, const uint32_t* idx, int k,
                        const float* query, int64_t n_q, int rbf,
                        double spacing, double* w_out);
/* The same with the systems solved in FP32 (solve_bits = 32, the reference's
 * precision) or FP64 (64). */
int NO_FN(rpim_weights_ex)(int D, const float* src, const uint32_t* idx, int k,
                           const float* query, int64_t n_q, int rbf,
                           double spacing, int solve_bits, double* w_out);

/* ---- nonlocal divergence weights ---------------------------------------
 * W[(i*k + s)*D + d] for query i, restating operator_builder + RPIM at the
 * quadrature points + compute_divergence_weights
 * (nonlocal_calculus/operators.cuh:77-136, detail/quadrature_points.cuh:
 * 74-152,188-282, detail/divergence.inl:55-227).  idx/d2 are the kNN of the
 * queries over src (self first).  delta_out[i] = 0.05*sqrt(d2[i][1]). */
int NO_FN(divergence_weights)(int D, const float* src, int64_t n_src,
                              const float* query, int64_t n_q,
                              const uint32_t* idx, const float* d2, int k,
                              int rbf, const unsigned char* kvar, double* W_out,
                              double* delta_out);
int NO_FN(divergence_weights_ex)(int D, const float* src, int64_t n_src,
                                 const float* query, int64_t n_q,
                                 const uint32_t* idx, const float* d2, int k,
                                 int rbf, const unsigned char* kvar, int solve_bits,
                                 double* W_out, double* delta_out);

/* Apply a divergence operator to a vector field F[j*D + d]:
 * out[i] = tree-sum over s of sum_d W(d,s,i) (F_d(idx(s,i)) - off),
 * nonlocal_calculus/divergence.cuh:377-488. */
int NO_FN(divergence_apply)(int D, int64_t n, int k, const uint32_t* idx,
                            const double* W, const double* F, double off,
                            double* out);

/* ---- the engine ---------------------------------------------------------
 * detail/simulation_engine.cuh:85-1486.  Nodes ordered [bulk|layer|boundary]
 * (no ghost nodes).  Distributions are returned alpha-major f[alpha*N + i]. */
typedef struct no_engine no_engine;

typedef struct no_config {
    int Q;              /* 9 or 27 */
    int k;              /* divergence stencil width */
    int k_boundary;     /* boundary interpolation stencil width */
    int rbf;            /* 0 = multiquadric, 1 = polyharmonic r^3 */
    int rbf_boundary;   /* RBF for the boundary interpolator */
    int pml_ref_index;  /* 1: layer field reads Q at (j - num_bulk) like the
                           reference; 0: reads Q at j */
    int threads;        /* OpenMP threads for stepping (0 = default) */
    double nu;          /* physical viscosity */
    double rho0;        /* nominal density */
    double dt;          /* physical time step */
    double c;           /* physical speed of sound */
    double h;           /* nodal spacing */
    double filter;      /* selective filter strength: 0.75 in the reference
                           (simulation_engine.cuh:1242); other values exist
                           only for tests of the scheme without it */
    const unsigned char* kvar; /* optional per-node stencil widths <= k */
    int solve_bits;     /* 64 (default, also for 0) or 32: the RBF systems of the
                           divergence and boundary weights solved in FP32, as
                           the reference does (batched_matrix_inverse.cuh:
                           229-250) */
} no_config;

no_engine* NO_FN(engine_create)(const no_config* cfg, const float* pts,
                                int64_t n_bulk, int64_t n_layer,
                                int64_t n_boundary, const float* normals,
                                const float* sigma);
void NO_FN(engine_destroy)(no_engine* e);
/* f_alpha = w_alpha * (1 + amp*exp(-|x-x0|^2/(2 width^2))) */
int NO_FN(engine_set_pulse)(no_engine* e, const double* center, double amp,
                            double width);
int NO_FN(engine_set_distributions)(no_engine* e, const double* f);
int NO_FN(engine_step)(no_engine* e, int nsteps);
/* f: Q*N alpha-major; rho: N; u: D*N (d-major); drho: N = rho - rho0
 * computed without cancellation.  fresh = 1 recomputes rho, u from the
 * current f; fresh = 0 returns the engine's macroscopic state. */
int NO_FN(engine_get)(no_engine* e, double* f, double* rho, double* u,
                      double* drho, int fresh);
/* setup products: idx N*k, W N*k*D, filter N*k, bidx Nb*kb, bw Nb*kb,
 * pml Q1/Q2 Q*N */
int NO_FN(engine_get_setup)(no_engine* e, uint32_t* idx, double* W,
                            double* filt, uint32_t* bidx, double* bw);
int NO_FN(engine_get_pml)(no_engine* e, double* Q1, double* Q2);
/* Test hooks of the PML stage alone: engine_set_pml overwrites the stored
 * PML history (Q1: the previous equilibrium, Q2: the previous trapezoid
 * value, both Q*N alpha-major); engine_pml_stage takes the macroscopic
 * values of the current f, sets the stage input s1 = s1_in (Q*N), runs the
 * PML terms of one step (compute_Qvalues, the layer's local damping, the
 * divergences of both auxiliary fields, post_collision) and returns the
 * stage output in s1_out. */
int NO_FN(engine_set_pml)(no_engine* e, const double* Q1, const double* Q2);
int NO_FN(engine_pml_stage)(no_engine* e, const double* s1_in, double* s1_out);
double NO_FN(engine_time)(no_engine* e);
/* density sources (see eng_sources in naga_oracle.c); at most
 * NO_MAX_SOURCES; returns non-zero when full or radius <= 0 */
#define NO_MAX_SOURCES 16
int NO_FN(engine_add_density_source)(no_engine* e, const double* center, double radius,
                                     double amplitude, double frequency, double phase,
                                     double t_start, double t_end);
int NO_FN(engine_clear_density_sources)(no_engine* e);

#ifdef __cplusplus
}
#endif
#endif
