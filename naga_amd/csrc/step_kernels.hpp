// Stepping kernels of the nonlocal LBM on CDNA4 (gfx950).
//
// State layout in HBM: each node owns one row of RS floats holding the
// deviation g_alpha = f_alpha - w_alpha of its Q distributions (AoS).  Rows
// are ordered [bulk | layer | boundary].  The divergence stencil of fluid
// node i is a slab of k records {neighbour index, W_x, W_y(, W_z)}.
//
// Kernels that gather neighbours map one node to LPN consecutive lanes with
// lane = alpha ("lane over alpha"): the Q lanes of a node read one
// contiguous row of each neighbour, so every gather is one cache line, and
// the stencil sum over s runs sequentially inside the lane (no cross-lane
// reduction).  Kernels doing per-node algebra (macroscopic values, the FP64
// collision, the PML recurrences) map one node to one thread.
//
// One time step is twelve launches:
//   p1      macroscopic values, FP64 collision                (all nodes)
//   pl      PML local terms + partial layer sums              (nodes whose Q is used)
//   pm      layer means of the two PML auxiliary values       (2Q columns)
//   pf      PML auxiliary field rows                          (layer rows)
//   p2      PML divergence of the auxiliary fields            (nodes with layer neighbours)
//   p3      boundary interpolation of the post-PML state      (boundary;
//           the sources are read before their collision increment, which
//           p1 keeps aside for them, as the reference interpolates before
//           colliding)
//   p4      selective filter + bounce-back                    (all nodes)
//   s1..s4  the four RK4 stages of the streaming equation     (fluid nodes)
//   p5      boundary interpolation of the new state           (boundary)
#pragma once

#include "collision.hpp"
#include "naga_common.hpp"

namespace naga_amd {

struct StepGeom {
    int64_t N;        // all nodes
    int64_t nbulk;    // [0, nbulk) bulk
    int64_t nlayer;   // [nbulk, nfluid) absorbing layer
    int64_t nfluid;   // nbulk + nlayer; [nfluid, N) boundary
    int k;            // divergence stencil width
    int kb;           // boundary interpolation stencil width
    int pml_ref;      // layer field reads Q at (j - nbulk) as the reference
    float dt;         // lattice time step
    float rho0;       // nominal density
    float vel_scale;  // c / c_s
    float filter;     // selective filter strength (0.75)
};

struct StepBuffers {
    float* g;        // [N][RS] current deviation state
    float* A;        // [N][RS] stage ping
    float* B;        // [N][RS] stage pong
    float* acc;      // [N][RS] RK4 stage-3 output (the fourth stage buffer)
    float4* ru;      // [N] unitless (rho - 1, u) from p1
    float* Gf;       // [N][RS] PML auxiliary field (zero outside the layer)
    const int32_t* rq_list;  // nodes whose PML Q values are maintained
    int64_t n_rq;
    float* E;        // [N][4] (rho - 1, u) of the previous step's PML local
                     // terms, from which the previous equilibrium deviation
                     // (PML Q1 - w) is recomputed
    float* Qp;       // [N][RS] previous Q1 value (PML Q2 source)
    float* Q2v;      // [N][RS] current Q2 value
    const float* slab;    // [nfluid][k][1+D] stencil records
    const float* fslab;   // [nfluid][k][2] filter records {idx, weight}
    // D2Q9 tile unions: the same records with the index fields holding
    // positions in the node's tile union (null without unions)
    const float* uslab;
    const float* ufslab;
    const float* bslab;   // [nbnd][kb][2] boundary interpolation records
    const float* sigma;   // [nlayer] absorption
    const float* normals; // [nbnd][D] inward normals
    float* coll;     // [N][RS] collision increment of boundary-source nodes
    const uint8_t* bsrc;  // [nfluid] 1 if the node feeds the boundary interpolator
    const int32_t* pml_list;  // fluid nodes with a layer neighbour
    int64_t n_pml;
    const int32_t* ptiles;    // with tile unions: the tiles holding a pml_list node
    int64_t n_ptiles;
    float* rho;      // [N] density (reference semantics, FP32)
    float* u;        // [D][N] velocity
    double* partial; // [2Q][nparts] layer column sums (one column entry per part:
                     // a PML-local block, or a fused collision block)
    float* means;    // layer means: Q1 at [a], Q2 at [means_stride(Q) + a]
    int nparts;
    int part_lo;     // parts [part_lo, nparts) hold the layer sums
    const float* src;  // [N] density source of this step (physical units), or null
    // Tile unions (D3Q27 four-per-lane kernels): the fluid nodes form tiles
    // of 64 or 32; ulist[uoff[t] .. uoff[t+1]) are the distinct stencil neighbours
    // of tile t, and the index field of slab / fslab records holds the
    // neighbour's position in that list; slab and fslab rows are then
    // packed (packed_div_units, packed_filter_units).  Unions of at most
    // ucap rows are staged in LDS by the gather kernels.  A tile over the
    // cap with usplit[t] > 0 holds two half-tile unions (lengths usplit[t]
    // and the rest, each at most ucap): its block serves the first half,
    // and one more block per split tile, after the regular ones (tile
    // entries -1 - s in ptiles), serves the second half of split_tiles[s].
    const uint32_t* ulist;
    const int64_t* uoff;
    const int32_t* usplit;
    const int32_t* split_tiles;
    int64_t n_split;
    int ucap;
    int ucap_f;      // LDS union rows of the filter pass (at most ucap: its smaller
                     // records leave room for more blocks per CU)
    // Per-node stencil widths with tile unions: tile t's nodes keep
    // tile_ng[t] groups of 8 records (ragged rows of ragged_div_units /
    // ragged_filter_units(tile_ng[t]) units), its divergence and PML records
    // from unit rec_off[t] of slab / p2slab, its filter records from unit
    // frec_off[t] of fslab.  Null: every tile keeps k records per node.
    const int32_t* tile_ng;
    const int64_t* rec_off;
    const int64_t* frec_off;
    int self_first;  // every fluid node's first stencil entry is the node itself
                     // (its own row is then in its tile's union, record 0);
                     // NAGA_OWN_LDS=0 clears it
    // The same for the PML divergence, whose unions hold the layer
    // neighbours and one row outside the layer (Gf is zero there): its
    // packed records, lists, offsets and split tiles.
    const float* p2slab;
    const uint32_t* p2ulist;
    const int64_t* p2uoff;
    const int32_t* p2usplit;
    const int32_t* p2split_tiles;
    int64_t p2n_split;
    // Boundary interpolation tiles (32 boundary nodes): the distinct source
    // nodes of each tile, bulist[buoff[t] .. buoff[t+1]), and bslab index
    // fields holding positions in that list (null: no boundary unions).
    // Tiles of at most kBUnionCap sources are staged in LDS.
    const uint32_t* bulist;
    const int64_t* buoff;
};

// Batched launches: one launch of a pass covers the scenes of a batch; the
// blocks [bstart[s], bstart[s+1]) of the launch belong to scene s, which
// reads its geometry, buffers and collision parameters from scenes[s].
struct SceneDesc {
    StepGeom G;
    StepBuffers S;
    CollisionParams cp;
};
struct BatchArgs {
    const SceneDesc* scenes;  // null for a single-scene launch
    const int32_t* bstart;    // nscenes + 1 block offsets of this pass
    int nscenes;
};

// Density sources: per step, field[nodes[i]] = sum over entries e in
// [off[i], off[i+1]) of prof[e] * amp[ids[e]] (sources in a fixed order).
constexpr int kMaxSources = 16;
struct SourceAmps {
    float a[kMaxSources];
};
void launch_source_field(const int32_t* nodes, const int32_t* off, const int32_t* ids,
                         const float* prof, int64_t n, const SourceAmps& amps, float* field,
                         hipStream_t stream);

// Launch the kernels of one step on `stream`.  If `events` is not null,
// events[t] is recorded before kernel t and events[kStepKernels] after the
// last (kernels with no work still get their pair of marks).
constexpr int kStepKernels = 12;
// Parts of the layer sums: blocks of the PML-local pass, or, when the
// collision pass carries the PML local terms, its 64-node blocks up to the
// last layer node; parts below pml_part_lo hold no layer node.
int64_t pml_parts(int Q, int k, int64_t N, int64_t nfluid, int64_t n_rq);
int64_t pml_part_lo(int Q, int k, int64_t N, int64_t nbulk);
template <int Q>
void launch_step(const StepGeom& geo, const StepBuffers& buf,
                 const CollisionParams& cp, hipStream_t stream,
                 hipEvent_t* events = nullptr);

// One step of every scene of a batch: kStepKernels batched launches.  The
// host-side table describes the same scenes as the device table (block
// counts and kernel choices are derived from it); bstart holds
// kStepKernels consecutive arrays of nscenes + 1 block offsets (device).
struct BatchPlan {
    int Q = 9;
    int nscenes = 0;
    const SceneDesc* dev_scenes = nullptr;
    const int32_t* dev_bstart = nullptr;
    int64_t blocks[kStepKernels] = {};  // total blocks per pass
    int k_template = 0;                 // K template of the gather kernels (0: dynamic)
    bool unions = false;                // D3Q27 four-per-lane kernels
    int ucap = 0;                       // max LDS union rows over the scenes
    int ucap_f = 0;                     // the same for the filter pass
    int gather_batch = 0;               // neighbours gathered per unrolled step (0: default)
    bool tiles = false;                 // tile unions built for every scene
    int kmax = 0;                       // largest stencil width over the scenes
    bool ragged = false;                // per-tile record widths in every scene
    bool three = false;                 // D2Q9: three distributions per lane (else one)
};
// Block counts of each pass for one scene (the same counts launch_step
// uses).  three: D2Q9 lanes per node, 1 for three distributions per lane,
// 0 for one, -1 for the default (q9_three).
void step_block_counts(int Q, const StepGeom& G, const StepBuffers& S, int64_t* blocks,
                       int three = -1);
template <int Q>
void launch_step_batch(const BatchPlan& plan, hipStream_t stream);

// Whether the step runs the D3Q27 four-per-lane kernels, which may use
// the tile unions (StepBuffers::ulist/uoff/ucap) and slabs holding union
// positions: D3Q27, k in {16, 32, 48}, fewer than 2^25 nodes.
bool step_uses_unions(int Q, int k, int64_t N);
// Whether the D2Q9 stencil passes run three distributions per lane by
// default (NAGA_Q9_THREE=0: one), and whether they can (fewer than 2^26
// nodes: 32-bit buffer offsets of 36-byte rows).
bool q9_three(int Q, int64_t N);
bool q9_three_fits(int64_t N);
// Nodes per union tile (= per block of the union kernels) for stencil width
// k: 64 for k <= 32 (5.4 union rows per node against 7.0 with 32-node
// tiles, two blocks per CU; 1.40 against 1.44 ms per step on the 1M-node
// D3Q27 scene), else 32; NAGA_TILE=32 selects 32-node tiles for every k.
int union_tile_nodes(int k);
// Default LDS union rows per block: 400 for 64-node tiles (two blocks per
// CU), 296 for 32-node tiles and k <= 32 (three blocks per CU), 440 for
// k = 48 (two blocks per CU, the records taking 22 KB).  A tile with a
// larger union reads its rows from HBM through the union list.
int union_cap_default(int k);
// Stride of the two rows of layer means (D3Q27: 32, so each row's
// quarters load as 16-byte vectors; padding entries stay zero).
constexpr int means_stride(int Q) { return Q == 27 ? 32 : Q; }
constexpr int kBUnionCap = 256;   // LDS rows per boundary-interpolation tile
constexpr int kBTileNodes = 32;   // boundary nodes per boundary-interpolation tile
constexpr int kUnionRowsMax = 512;  // LDS union rows per block, upper bound
// D2Q9 tile unions: fluid nodes per tile (the block of the three-per-lane
// kernels: 21 nodes per wavefront, four wavefronts) and LDS union rows per
// tile, upper bound (36-byte rows).
constexpr int kQ9TileNodes = 84;
constexpr int kQ9UnionRowsMax = 512;

// Per-kernel entry points (used by the profiler and the standalone C-ABI).
template <int Q>
void launch_rk4_stage(int stage, const StepGeom& geo, const StepBuffers& buf,
                      hipStream_t stream);

// Density probes (density_observer.cuh:62-113): out[p] = rho0 +
// sum_s w[p][s] (rho(idx[p][s]) - rho0), with rho the held FP32 density
// (held != null) or computed fresh from the deviation rows g.
void launch_probe(int Q, int64_t n, int k, const uint32_t* idx, const float* w,
                  const float* held, const float* g, double rho0, double* out,
                  hipStream_t stream);

// Fresh macroscopic values from the current state: rho (FP32, reference
// units), drho = rho - rho0 in FP64 without cancellation, u (FP32).
template <int Q>
void launch_macroscopic(const StepGeom& geo, const float* g, float* rho,
                        double* drho, float* u, hipStream_t stream);

}  // namespace naga_amd
