/*
 * naga_amd.h -- C-ABI of the MI355X-native nonlocal lattice-Boltzmann
 * acoustics step and its setup (one GPU, HIP/CDNA4).
 *
 * Plain pointers and sizes only.  Arrays are host memory, copied in and out;
 * device state is owned by the handles.  Every function returns a status
 * (NAGA_OK = 0) and sets a thread-local message readable with
 * naga_last_error().  Reference interfaces are cited as path:line relative
 * to the NAGA source tree; the C++ facade naga_amd/naga.hpp maps these
 * entry points back to the reference's naga:: classes.
 */
#ifndef NAGA_AMD_H
#define NAGA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NAGA_OK 0
#define NAGA_INVALID_ARGUMENT 1
#define NAGA_OUT_OF_MEMORY 2
#define NAGA_DEVICE_ERROR 3
#define NAGA_NUMERICAL_ERROR 4
#define NAGA_NOT_SUPPORTED 5

#define NAGA_RBF_MULTIQUADRIC 0
#define NAGA_RBF_POLYHARMONIC3 1

typedef struct naga_engine naga_engine;
typedef struct naga_batch naga_batch;

/* Scene configuration.  Physics parameters follow
 * simulation_engine::set_problem_parameters(nu, rho0, dt, c)
 * (detail/simulation_engine.cuh:94-124); the stencil constants that the
 * reference hard-codes are fields here. */
typedef struct naga_scene_config {
    int32_t lattice;        /* 9 (D2Q9) or 27 (D3Q27) */
    int32_t k;              /* divergence stencil width (reference: 32) */
    int32_t k_boundary;     /* boundary RPIM stencil width (reference: 32) */
    int32_t rbf;            /* RBF of the divergence weights */
    int32_t rbf_boundary;   /* RBF of the boundary interpolator */
    int32_t pml_ref_index;  /* 1: PML layer field reads Q at (j - num_bulk)
                               as pml_absorption.inl:445-448 does; 0: at j */
    int32_t k_min;          /* 0: every node uses k neighbours; else node i
                               uses k_i in [k_min, k] by local density,
                               k_i = k_min d_ref / d_i (d_i: mean distance to
                               the 4 nearest nodes, d_ref its 90th
                               percentile); stencil entries past k_i carry
                               zero weights */
    double nu;              /* fluid viscosity */
    double rho0;            /* nominal density */
    double dt;              /* time step */
    double c;               /* speed of sound */
    double h;               /* nodal spacing (simulation_nodes::nodal_spacing) */
    double filter_strength; /* selective filter strength per step: 0 gives the
                               reference's 0.75 (detail/simulation_engine.cuh:
                               1242), > 0 that strength, < 0 no filter */
    int32_t solve_precision; /* bits of the (k + D + 1)^2 RBF solves of the
                               divergence and boundary weights: 64 (or 0) FP64;
                               32 FP32, the reference's precision
                               (linalg/batched_matrix_inverse.cuh:229-250) --
                               with k = 32 and multiquadric RBFs the
                               reference-parity mode */
} naga_scene_config;

const char* naga_last_error(void);

/* Lattice tables: c[Q*D], w[Q], bb[Q], K[Q*Q] (row alpha, column j).
 * detail/d2q9_lattice.cuh:44-88,151-163; detail/d3q27_lattice.cuh:45-105,
 * 380-387. */
int naga_lattice_tables(int32_t q, double* c, double* w, int32_t* bb, double* K);

/* Collision increment of one node on the host (FP64), for tests of the
 * device formula: f full distribution, rho = sum f, u unitless velocity.
 * lattice_interface<...>::compute_moment_projection + collision matrix,
 * detail/simulation_engine.cuh:486-527. */
int naga_collide(int32_t q, const double* f, double rho, const double* u,
                 double nu_lat, double dt_lat, double* df);

/* ---- engine: simulation_engine<Lattice>
 *      (fluids/nonlocal_lattice_boltzmann/simulation_engine.cuh:39-121) ---- */
int naga_engine_create(const naga_scene_config* cfg, naga_engine** out);
int naga_engine_destroy(naga_engine* e);

/* init_domain(simulation_nodes) (detail/simulation_engine.cuh:126-339):
 * points[N*D] ordered [bulk | layer | boundary] (no ghost nodes),
 * normals[n_boundary*D] inward, layer_absorption[n_layer]. */
int naga_engine_init_domain(naga_engine* e, const float* points, int64_t n_bulk,
                            int64_t n_layer, int64_t n_boundary,
                            const float* normals, const float* layer_absorption);

/* Initial condition f_alpha = w_alpha (1 + amp exp(-|x-c|^2 / (2 width^2))),
 * the one-shot density source of examples/nonlocal_lattice_boltzmann.cu:
 * 79-125 applied as an initial state; resets time to 0. */
int naga_engine_set_pulse(naga_engine* e, const double* center, double amp,
                          double width);
/* Density source, the declarative replacement of a density_source plugin
 * (density_source.cuh:41-82; applied as detail/simulation_engine.cuh:
 * 347-364,1057-1107 do, with the density update made once per node).  At
 * the start of each step at time t, every node within `radius` of `center`
 * gets s = amplitude * cos(2 pi frequency t + phase) * P(r / radius),
 * P(x) = (1 - (2x)^2 / 4) exp(-(2x)^2) (the profile of the reference
 * example's source), while t_start <= t < t_end; sources add.  Where s != 0:
 * f_alpha += w_alpha s / rho0 and rho <- rho + s / 2.  At most 16. */
int naga_engine_add_density_source(naga_engine* e, const double* center, double radius,
                                   double amplitude, double frequency, double phase,
                                   double t_start, double t_end);
int naga_engine_clear_density_sources(naga_engine* e);
/* Density probes, the density_observer of
 * simulation_observers/density_observer.cuh:62-113: each probe point is
 * interpolated from its 32 nearest nodes with RPIM weights at the nodal
 * spacing (rbf as for naga_rpim_weights; the reference uses the
 * multiquadric): value = rho0 + sum_s w_s (rho_s - rho0).  fresh = 0 uses
 * the density held from the start of the last step (what the observer sees
 * in update()), fresh = 1 the current state. */
int naga_engine_add_probes(naga_engine* e, const float* points, int64_t n, int32_t rbf,
                           int32_t* probe_set);
int naga_engine_sample_probes(naga_engine* e, int32_t probe_set, int32_t fresh,
                              double* values);
/* Snapshot/restore of the scene state that carries between steps
 * (distributions, PML history, held macroscopic values, clock) to a host
 * buffer; replaces simulation_engine::save_state/load_state
 * (fluids/nonlocal_lattice_boltzmann/simulation_engine.cuh:108-116) with a
 * binary format: a header (magic "NAGAST02", lattice, sizes, time, frame)
 * then FP32 arrays (distributions, the previous step's (rho - 1, u) for the
 * PML terms, PML Q1, density, velocity).  Setup products are rebuilt by
 * init_domain and are not part of the state. */
int naga_engine_state_size(naga_engine* e, int64_t* bytes);
int naga_engine_save_state(naga_engine* e, void* buf, int64_t capacity);
int naga_engine_load_state(naga_engine* e, const void* buf, int64_t bytes);
/* f[Q*N] alpha-major full distributions (FP64 in, stored as FP32 deviations). */
int naga_engine_set_distributions(naga_engine* e, const double* f);

/* step_forward() x nsteps (detail/simulation_engine.cuh:1325-1375). */
int naga_engine_step(naga_engine* e, int32_t nsteps);
/* The same, timed on the engine's stream with HIP events (ms). */
int naga_engine_step_timed(naga_engine* e, int32_t nsteps, double* ms);
/* Per-kernel timing of one step: ms[12] for p1 (collision), pl (PML local),
 * pm (layer means), pf (PML field), p2 (PML divergence), p3 (boundary
 * interpolation), p4 (filter + bounce-back), s1..s4 (RK4 stages), p5
 * (boundary interpolation). */
int naga_engine_profile_step(naga_engine* e, double* ms);

/* solution().macroscopic_values: rho[N] (FP32), u[D*N] (d-major, FP32) and
 * drho[N] = rho - rho0 (FP64, from the FP32 deviation state without
 * cancellation).  fresh = 1 recomputes them from the current state;
 * fresh = 0 returns the values held by the engine (computed at the start of
 * the last step, boundary velocities zeroed by bounce-back), as the
 * reference does. Any output may be NULL. */
int naga_engine_get_macroscopic(naga_engine* e, float* rho, float* u,
                                double* drho, int32_t fresh);
/* f[Q*N] alpha-major (FP64 = w + deviation) and/or the deviation g[Q*N]. */
int naga_engine_get_distributions(naga_engine* e, double* f, double* g);
/* k_i per node (uint8[N]; all k when k_min = 0). */
int naga_engine_get_stencil_sizes(naga_engine* e, uint8_t* k);
/* Setup products for parity tests: idx[N*k], W[N*k*D], filt[N*k],
 * bidx[Nb*kb], bw[Nb*kb]. Any may be NULL. */
int naga_engine_get_setup(naga_engine* e, uint32_t* idx, float* W, float* filt,
                          uint32_t* bidx, float* bw);
/* PML state: E = Q1 - w and Qp = Q2 (reference naming) as [Q*N] alpha-major;
 * only rows of nodes whose PML values are used are maintained. */
int naga_engine_get_pml(naga_engine* e, double* Q1_minus_w, double* Q2);
int naga_engine_info(naga_engine* e, int64_t* n_nodes, double* time,
                     int64_t* frame, double* setup_ms);
/* The output of vtk_observer::update_all (simulation_observers/
 * vtk_observer.cuh:86-247) for the engine's scene: one VTK XML PolyData
 * file (.vtp; inline binary, UInt32 headers, uncompressed) with the nodes as
 * vertices and the point arrays f (f_alpha / w_alpha, Q components),
 * density, velocity (3 components), absorption, normals (3) and type
 * (0 bulk, 1 layer, 2 boundary), all FP32 but type.  fresh = 0 writes the
 * macroscopic values held from the start of the last step (what an observer
 * sees in update()), fresh = 1 the current state.  The file name (the
 * reference's result.<frame>.vtp) is the caller's. */
int naga_engine_write_vtp(naga_engine* e, const char* path, int32_t fresh);
/* The last init_domain's phases, ms[5] of wall time on the engine's stream:
 * cell binning and kNN of the divergence stencils, divergence weights (RBF
 * solves), filter weights + stencil slabs + tile unions, boundary
 * interpolator (kNN and RPIM solves), PML lists and state.  No counterpart
 * in the reference (for measurement). */
int naga_engine_setup_phases(naga_engine* e, double* ms);

/* ---- batch: many scenes stepped together.  No counterpart in the
 *      reference (which steps one engine at a time): every pass of a step
 *      is one launch over all scenes of a lattice/kernel group, replayed
 *      from a captured hipGraph.  Engines keep ownership of their scenes;
 *      an engine is in at most one batch, and destroying it removes it. */
int naga_batch_create(naga_batch** out);
int naga_batch_destroy(naga_batch* b);
int naga_batch_add(naga_batch* b, naga_engine* e);     /* e needs a domain */
int naga_batch_remove(naga_batch* b, naga_engine* e);
int naga_batch_set_graph(naga_batch* b, int32_t on);  /* default on */
/* Autotuner: times trial_steps steps of the batch for each gather batch
 * (neighbours gathered per unrolled step: 8, 16, 32), with and without
 * graph replay and, when the batch holds D2Q9 scenes, for each D2Q9 kernel
 * form (3 lanes per node on 84-node tiles reading the tile unions staged in
 * LDS, the same gathering directly, 9 lanes per node on 28-node tiles);
 * keeps the fastest, and restores every scene's state afterwards. */
int naga_batch_autotune(naga_batch* b, int32_t trial_steps, int32_t* gather_batch);
/* D2Q9 kernel form of the batch's stencil passes: lanes per node 3
 * (default) or 9; unions 1 (default; lanes 3 only): scenes with tile
 * unions read them from LDS. */
int naga_batch_set_q9_kernels(naga_batch* b, int32_t lanes, int32_t unions);
/* The batch's current choices: gather batch (0: the kernels' default), D2Q9
 * lanes per node and unions, graph replay (1 or 0); null outputs are
 * skipped. */
int naga_batch_tuning(naga_batch* b, int32_t* gather_batch, int32_t* q9_lanes,
                      int32_t* q9_unions, int32_t* graph);
/* nsteps steps of every scene (each engine's time advances by its dt) */
int naga_batch_step(naga_batch* b, int32_t nsteps);
int naga_batch_step_timed(naga_batch* b, int32_t nsteps, double* ms);

/* ---- serving loop: acoustic-pulse solves with continuous batching.  No
 *      counterpart in the reference.  Requests are admitted while their
 *      estimated device footprint fits hbm_budget_bytes and fewer than
 *      max_active are resident; setup_threads threads build admitted scenes
 *      (host node generation, GPU setup on each scene's own stream) while a
 *      stepping thread advances all resident scenes together in chunks of
 *      at most chunk_steps steps; finished scenes leave at once and free
 *      their memory.  Scenes are the unit box minus obstacles
 *      (naga_obstacle_grid), dt = 0.4 of the finest spacing (h, or h/2
 *      with refinement), c = c_s, rho0 = 1, r^3 RBFs. */
typedef struct naga_server naga_server;
#define NAGA_MAX_OBSTACLES 8
typedef struct naga_request {
    int32_t lattice;           /* 9 (D2Q9) or 27 (D3Q27) */
    int32_t k;                 /* stencil width; 0: 16 (D2Q9) or 32 (D3Q27) */
    int32_t steps;
    int32_t n_obstacles;       /* <= NAGA_MAX_OBSTACLES */
    double h;                  /* grid spacing */
    double jitter;             /* interior jitter, fraction of h */
    uint64_t seed;
    double obstacles[NAGA_MAX_OBSTACLES * 4];  /* per obstacle: centre[D], radius */
    double layer_thickness;    /* absorbing layer at the box faces */
    double peak_absorption;
    double tau;                /* relaxation time (0: 0.6) */
    double pulse_center[3];
    double pulse_amplitude;
    double pulse_width;
    int32_t k_min;             /* 0: every node uses k; else per-node widths in
                                  [k_min, k] by local density (naga_scene_config) */
    double refine_width;       /* > 0: 2x node density within this distance of
                                  the obstacle surfaces (naga_obstacle_grid) */
    double immersed_layer_thickness;  /* absorbing layer around the obstacles */
    double immersed_absorption;       /* (naga_obstacle_grid; 0: none) */
} naga_request;

typedef struct naga_server_stats {
    int64_t submitted, completed, failed, resident, queued;
    int64_t bytes_in_use;
    double p50_ms, p99_ms, mean_ms;  /* latency, submit to result ready */
    double node_updates_per_s;       /* completed work / (last completion - first submit) */
} naga_server_stats;

int naga_server_create(int64_t hbm_budget_bytes, int32_t max_active, int32_t chunk_steps,
                       int32_t use_graph, int32_t setup_threads, naga_server** out);
int naga_server_destroy(naga_server* s);  /* waits for resident work */
int naga_server_submit(naga_server* s, const naga_request* req, int64_t* id);
/* *state: 0 queued or running, 1 done (see naga_server_result) */
int naga_server_status(naga_server* s, int64_t id, int32_t* state);
/* Result of a done request: *n_nodes always; rho[N], u[D*N], drho[N] when
 * not NULL (fresh values of the final state).  Fails with the request's
 * error if its setup failed. */
int naga_server_result(naga_server* s, int64_t id, int64_t* n_nodes, float* rho, float* u,
                       double* drho, double* latency_ms);
int naga_server_release(naga_server* s, int64_t id);
int naga_server_drain(naga_server* s);   /* block until all submitted are done */
int naga_server_get_stats(naga_server* s, naga_server_stats* out);

/* ---- node provider: uniform_grid_provider<Lattice>::get()
 *      (node_providers/uniform_grid_provider.cuh:65-161) with seeded
 *      SplitMix64 jitter of interior nodes; layer by exact 1-NN distance to
 *      the boundary nodes on the GPU.  Call naga_uniform_grid_size first. */
int64_t naga_uniform_grid_size(int32_t dims, const double* lo, const double* hi,
                               double h);
int naga_uniform_grid(int32_t dims, const double* lo, const double* hi, double h,
                      double jitter_frac, uint64_t seed, double layer_thickness,
                      double peak_absorption, float* points, float* normals,
                      float* sigma, int64_t* counts);

/* ---- procedural geometry: the box [lo, hi] minus n_obstacles balls
 *      (obstacles[j*(dims+1) ..] = centre, radius).  No single reference
 *      counterpart; it plays the role of the conforming point-cloud provider
 *      (node_providers/experimental/conforming_point_cloud_provider.cuh:
 *      147-525) for circles/spheres.  Grid nodes follow naga_uniform_grid
 *      (same SplitMix64 stream and jitter) and are dropped within half a
 *      spacing of an obstacle; obstacle surfaces get nodes at about the grid
 *      spacing (equal angles on circles, a Fibonacci lattice on spheres)
 *      with normals pointing into the fluid.  refine_width > 0 doubles the
 *      node density within that distance of the obstacle surfaces (the
 *      2^dims - 1 half-offset nodes of each grid cell there, jittered by
 *      jitter_frac * h/2 from a second stream; surface nodes at h/2).  The
 *      absorbing layer is the
 *      fluid within layer_thickness of the box faces (exact distance d),
 *      sigma = peak_absorption (1 - d/L)^2, and, as the provider's immersed
 *      boundary layers (conforming_point_cloud_provider.cuh:250-280), the
 *      fluid within immersed_layer_thickness Li of an obstacle surface and
 *      nearer to it than to the faces, sigma = immersed_absorption
 *      (1 - d/Li)^2 (0 for either: no such layer).  With points == NULL only
 *      counts[3] = {bulk, layer, boundary} are returned (size query). */
int naga_obstacle_grid(int32_t dims, const double* lo, const double* hi, double h,
                       double jitter_frac, uint64_t seed, int32_t n_obstacles,
                       const double* obstacles, double refine_width, double layer_thickness,
                       double peak_absorption, double immersed_layer_thickness,
                       double immersed_absorption, int64_t capacity, float* points,
                       float* normals, float* sigma, int64_t* counts);

/* ---- setup primitives ---------------------------------------------------- */
/* batched_nearest_neighbors (segmentation/nearest_neighbors.cuh:95-477),
 * exact: the k smallest (d2, index), d2 FP32 = dz^2 + (dy^2 + dx^2). */
int naga_knn(const float* ref, int64_t n_ref, const float* query, int64_t n_q,
             int32_t dims, int32_t k, float* d2_out, uint32_t* idx_out);
/* No counterpart in the reference.  naga_knn for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc): the same neighbours, bit for bit, with
 * every pointer a DEVICE pointer (ref[n_ref*dims], query[n_q*dims],
 * d2_out/idx_out[n_q*k]) and the cell binning in a static device scratch
 * array of the library (about 21 MB per GPU, held until the process ends),
 * so n_ref <= NAGA_KNN_NOALLOC_MAX_REF; larger n_ref is rejected with
 * NAGA_INVALID_ARGUMENT.  Work is queued on `stream` (a hipStream_t; NULL:
 * the null stream) and the call returns without waiting.  All calls on one
 * GPU share its scratch: calls that overlap on one GPU (other streams or
 * threads) race and must be ordered by the caller. */
#define NAGA_KNN_NOALLOC_MAX_REF (1 << 20)
int naga_knn_noalloc(const float* ref, int64_t n_ref, const float* query, int64_t n_q,
                     int32_t dims, int32_t k, float* d2_out, uint32_t* idx_out, void* stream);
/* batched_matrix_inverse(A, A_inv) (linalg/batched_matrix_inverse.cuh:
 * 325-397): A_inv[b] = A[b]^-1 for column-major n x n matrices (FP64 LU with
 * partial pivoting against the identity); info[b] = 0 or the first zero pivot
 * (1-based), which the reference leaves unchecked.  batch must be below
 * INT_MAX as in the reference. */
int naga_batched_matrix_inverse_f64(const double* A, int32_t n, int64_t batch, double* A_inv,
                                    int32_t* info);
/* The solve form of the same LU (linalg/batched_matrix_inverse.cuh:229-250
 * inverts and multiplies; this solves directly): A X = B per item in FP64,
 * A column-major n x n, B column-major n x nrhs; info[b] != 0 flags a zero
 * pivot. */
int naga_batched_lu_solve_f64(const double* A, int32_t n, int64_t batch,
                              const double* B, int32_t nrhs, double* X,
                              int32_t* info);

/* ---- interpolation::radial_point_method<T>
 *      (interpolation/radial_point_method.cuh:92-280) ---------------------- */
typedef struct naga_rpim naga_rpim;
/* get_scratchpad_size(query_size, support_size, dimensions, group_size)
 * (:97-118) for T of bytes_per_value (4 or 8) bytes; -1 (and the reference's
 * message in naga_last_error) when query_size is not a multiple of
 * group_size. */
int64_t naga_rpim_scratchpad_size(int64_t query_size, int32_t support_size, int32_t dims,
                                  int32_t group_size, int32_t bytes_per_value);
/* create_interpolator(source_points, interpolating_indices, query_points,
 * approx_particle_spacing, group_size, shape_function) (:143-169): src
 * [n_src*dims], idx [(n_q / group_size)*k] (the stencil of each group of
 * group_size consecutive queries), query [n_q*dims]; rbf 0 multiquadric
 * (the reference's default shape function) or 1 r^3.  Weights in FP64
 * solves, kept on the device as FP32. */
int naga_rpim_create(const float* src, int64_t n_src, const uint32_t* idx, int32_t k,
                     const float* query, int64_t n_q, int32_t dims, int32_t group_size,
                     int32_t rbf, double spacing, naga_rpim** out);
int naga_rpim_destroy(naga_rpim* r);
/* interpolate(field, destination, centering_offset) (:171-251): field
 * [n_field][m] (m components per source point; n_field must equal the
 * source count), dest [n_q][m] = sum_s w (field - offset) + offset. */
int naga_rpim_interpolate(naga_rpim* r, const float* field, int64_t n_field, int32_t m,
                          float* dest, float centering_offset);
/* the weights [n_q][k] (compute_weights' output) */
int naga_rpim_get_weights(naga_rpim* r, float* w);

/* ---- segmentation::nd_cubic_segmentation<T, D>
 *      (segmentation/nd_cubic_segmentation.cuh:219-461) ------------------- */
typedef struct naga_segmentation naga_segmentation;
/* nd_cubic_segmentation(points, approx_partition_size) (:227-303) when
 * approx_partition_size > 0, else nd_cubic_segmentation(points,
 * partition_width) (:305-379); points [n*dims] (dims 1..3), binned on the
 * device. */
int naga_segmentation_create(const float* points, int64_t n, int32_t dims,
                             int32_t approx_partition_size, double partition_width,
                             naga_segmentation** out);
int naga_segmentation_destroy(naga_segmentation* s);
/* shape()[3] (cells per axis, 1 past dims), lower bounds[3], partition width,
 * partition_count() = product of the shape.  Any output may be NULL. */
int naga_segmentation_shape(naga_segmentation* s, int64_t* shape, double* lower_bounds,
                            double* partition_width, int64_t* partition_count);
/* partition sizes [count], index offsets [count] and indices [n]: the points
 * of cell c (x fastest) are indices[offsets[c] .. offsets[c] + sizes[c]),
 * in no particular order within a cell (as the reference's atomics). */
int naga_segmentation_partitions(naga_segmentation* s, uint32_t* sizes, uint32_t* offsets,
                                 uint32_t* indices);
/* batched_nearest_neighbors(k, query_points, segmentation)
 * (segmentation/nearest_neighbors.cuh:454-477), exact as naga_knn. */
int naga_segmentation_knn(naga_segmentation* s, const float* query, int64_t n_q, int32_t k,
                          float* d2_out, uint32_t* idx_out);

/* One stage of the classic RK4 of the streaming equation
 * (detail/rk4_solver.cuh:48-171 with detail/pde_right_hand_side.cuh:82-176):
 * rate_alpha(i) = -sum_s sum_d W(d,s,i) c_{alpha,d} in_alpha(idx(s,i)) for
 * nodes i < n_fluid, zero past them.  Arrays are deviation rows g = f - w,
 * [n][q] FP32: y the state at the step start, in the stage input.
 *   stage 1: out = y + dt/2 rate(y)           (in unused)
 *   stage 2: out = y + dt/2 rate(in)          (in = stage 1's output)
 *   stage 3: out = y + dt rate(in)            (in = stage 2's output)
 *   stage 4: out = y + ((y1 - y) + 2 (y2 - y) + (in - y)) / 3 + dt/6 rate(in)
 *            (y1, y2 the outputs of stages 1 and 2, in that of stage 3),
 * which chains to y + dt/6 (k1 + 2 k2 + 2 k3 + k4).  idx [n_fluid*k]
 * (indices < n), W [n_fluid*k*D]; dt the lattice time step. */
int naga_rk4_stage(int32_t q, int64_t n, int64_t n_fluid, int32_t k, const uint32_t* idx,
                   const float* W, double dt, int32_t stage, const float* y, const float* in,
                   const float* y1, const float* y2, float* out);

/* radial_point_method::compute_weights, group size 1
 * (interpolation/radial_point_method.cuh:118-139): w[n_q*k]. */
int naga_rpim_weights(const float* src, int64_t n_src, const uint32_t* idx,
                      int32_t k, const float* query, int64_t n_q, int32_t dims,
                      int32_t rbf, double spacing, double* w);
/* divergence_operator::create (nonlocal_calculus/divergence.cuh:53-166):
 * kNN of the queries over src plus FP64 weights W[n_q*k*D] (FP64 out). */
int naga_divergence_weights(const float* src, int64_t n_src, const float* query,
                            int64_t n_q, int32_t dims, int32_t k, int32_t rbf,
                            uint32_t* idx_out, double* W_out);
/* divergence_operator::apply (nonlocal_calculus/divergence.cuh:183-212,
 * 377-488) in FP32: out[i] = sum_s sum_d W(d,s,i) (F_d(idx(s,i)) - off). */
int naga_divergence_apply(const uint32_t* idx, const float* W, int64_t n,
                          int32_t k, int32_t dims, const float* F, int64_t n_f,
                          float off, float* out);

/* ---- Standalone advection operator -------------------------------------
 * advection_operator::create(domain) (nonlocal_calculus/advection.cuh:
 * 143-151): exact kNN stencils of width k over the points (n x dims) and
 * their divergence weights (rbf as in naga_scene_config).  The operator's
 * state stays on the device. */
typedef struct naga_advection naga_advection;
int naga_advection_create(const float* points, int64_t n, int32_t dims, int32_t k,
                          int32_t rbf, naga_advection** out);
/* advection_operator::create(divergence_op) (advection.cuh:153-159) and the
 * load half of save/load (:498-523): an operator over existing stencils
 * idx (n x k, indices < n) and FP32 weights W (n x k x dims), e.g. those of
 * naga_advection_weights or naga_divergence_weights. */
int naga_advection_from_weights(const uint32_t* idx, const float* W, int64_t n,
                                int32_t dims, int32_t k, naga_advection** out);
int naga_advection_destroy(naga_advection* a);
/* nsteps of advection_task (advection.cuh:162-276): classic RK4 of
 * df/dt = div(-v (f - centering_offset)) with the divergence field map of
 * advection.cuh:102-141.  velocity: dims values (constant field) or n x dims
 * (per_node != 0).  f (n values) is read as f0 and overwritten by the result.
 * boundary_start > 0: only nodes [0, boundary_start) are updated, the rest
 * keep their value (the reference's end_slice); 0: every node. */
int naga_advection_step(naga_advection* a, const float* velocity, int32_t per_node,
                        float* f, double dt, double centering_offset,
                        int64_t boundary_start, int32_t nsteps);
/* stencil indices (n x k) and FP32 weights (n x k x dims); either may be
 * NULL (the save half of save/load, advection.cuh:498-514) */
int naga_advection_weights(naga_advection* a, uint32_t* idx, float* W);

/* ---- CSR matrices -------------------------------------------------------
 * linalg::matrix<T, sparse_csr> (linalg/matrix.cuh:231-437), device
 * resident, FP32 values.  naga_csr_from_stencil is
 * create_from_index_stencil(columns, index_stencil, values)
 * (matrix.cuh:300-368): row r holds the k entries idx[r*k + i] with values
 * values[r*k + i], in stencil order (the reference's per-row sort leaves the
 * order unchanged).  Column indices must be < cols; rows and cols < 2^31. */
typedef struct naga_csr naga_csr;
int naga_csr_create(int64_t rows, int64_t cols, const int64_t* row_offsets,
                    const uint32_t* column_indices, const float* values, naga_csr** out);
int naga_csr_from_stencil(int64_t rows, int64_t cols, int32_t k, const uint32_t* idx,
                          const float* values, naga_csr** out);
int naga_csr_destroy(naga_csr* m);
/* matrix_mult<csr, dense, dense> (matrix.cuh:438-629): y = alpha A x + beta y; x has cols
 * entries, y rows; y is only read when beta != 0. */
int naga_csr_mult(naga_csr* m, float alpha, float beta, const float* x, float* y);

#ifdef __cplusplus
}
#endif
#endif
