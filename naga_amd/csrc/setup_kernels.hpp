// Setup on the GPU: exact kNN, FP64 batched LU solves, RPIM weights,
// nonlocal divergence weights, filter weights and stencil slabs.
#pragma once

#include <vector>

#include "device_buffer.hpp"
#include "naga_common.hpp"

namespace naga_amd {

// Uniform cell binning of points on the device, the nd_cubic_segmentation
// of segmentation/nd_cubic_segmentation.cuh:227-303: bounds by reduction
// (extents below 1e-3 widened to 1e-3), cells of width
// w = (per_cell * V / n)^(1/D) (V the bounding volume) or a given width,
// dims[d] = ceil(extent_d / w), cell of a point floor((x - lo) / w) clamped
// to the grid; points listed by cell with ascending indices within a cell.
// If the cells would exceed the capacity (2 n + 64, reached only by
// degenerate point sets) w grows by 5/4 until they fit.
struct CellGrid {
    double lo[3];
    double w;          // cell width
    int64_t dims[3];
    int64_t ncell;
    int D;
};

// Views of one binning in a scratch block of device memory: cell c holds
// sorted[start[c] .. start[c + 1]); counts[c] = start[c + 1] - start[c].
struct CellBins {
    CellGrid* grid;     // device
    uint32_t* start;    // cap + 1
    uint32_t* counts;   // cap + 1
    uint32_t* sorted;   // n
    int64_t cap;        // cell capacity
    int64_t n;
};
// Scratch layout of one binning of n points (offsets 256-byte aligned):
// the grid, per-block bounds, cell counts and starts, scan block sums, the
// sorted indices.
constexpr int kBoundBlocks = 256;
constexpr int kScanChunk = 2048;  // elements per scan block: 256 threads x 8
struct BinLayout {
    size_t grid, partial, counts, start, bsum, sorted, total;
    int64_t cap, nblocks;
};
constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr BinLayout bin_layout(int64_t n) {
    BinLayout L{};
    L.cap = 2 * n + 64;
    const int64_t M = L.cap + 1;
    L.nblocks = (M + kScanChunk - 1) / kScanChunk;
    size_t off = 0;
    L.grid = off;
    off = align256(off + sizeof(CellGrid));
    L.partial = off;
    off = align256(off + size_t(kBoundBlocks) * 6 * sizeof(double));
    L.counts = off;
    off = align256(off + size_t(M) * 4);
    L.start = off;
    off = align256(off + size_t(M) * 4);
    L.bsum = off;
    off = align256(off + size_t(L.nblocks) * 4);
    L.sorted = off;
    off = align256(off + size_t(n) * 4);
    L.total = off;
    return L;
}

// Scratch bytes for binning n points.
size_t cell_bins_bytes(int64_t n);
// Bins pts[n][D] (device) into scratch (device, at least cell_bins_bytes(n)
// bytes, 256-byte aligned) on `stream`: per_cell > 0 sets the width from the
// points per cell, else width > 0 is the cell width.  Allocates nothing.
CellBins launch_bin_points(int D, const float* pts, int64_t n, double per_cell, double width,
                           void* scratch, size_t bytes, hipStream_t stream);
// Binning with its own device memory.
struct DeviceCells {
    DeviceBuffer<uint8_t> mem;
    CellBins bins{};
};
DeviceCells bin_points_device(int D, const float* pts, int64_t n, hipStream_t stream,
                              double per_cell = 2.0, double width = 0.0);

// Exact k-nearest neighbours of each query among ref, key (d2, index)
// ascending, d2 in FP32 as dz^2 + (dy^2 + dx^2) with no contraction.
// All pointers are device pointers; bins are those of ref.
void launch_knn(const CellBins& bins, const float* ref, const float* query, int64_t n_q, int k,
                float* d2_out, uint32_t* idx_out, hipStream_t stream);

// Batched dense LU with partial pivoting + solve, FP64, one wavefront per
// system (n <= 64).  A: batch x n x n column-major; B: batch x n x nrhs
// column-major, overwritten by X.  info[b] = 0 or first zero pivot (1-based).
void launch_batched_lu_solve(int n, int nrhs, int64_t batch, double* A,
                             double* B, int* info, hipStream_t stream);

// Nonlocal divergence weights of every query (one wavefront each):
// W[(i*k + s)*D + d] (FP32 and, if W64 != nullptr, FP64).
// rbf: 0 multiquadric (shape 0.1*hbar), 1 polyharmonic r^3.
// hbar[b] is the mean nearest-neighbour distance of query batch b (16384).
// kvar (optional, device): per-query stencil width <= k; stencil entries
// past it get zero weights (idx/d2/W keep the row stride k).
void launch_divergence_weights(int D, const float* src, const float* query,
                               int64_t n_q, const uint32_t* idx,
                               const float* d2, int k, int rbf,
                               const double* hbar, float* W, double* W64,
                               int* info, hipStream_t stream,
                               const uint8_t* kvar = nullptr, int solve_bits = 64);

// Ragged copy of packed record rows (KP units per node, upg units per group
// of 8 records: 7 divergence, 3 filter): node i of tile t = i / T keeps
// upg * ng[t] units padded to ragged_div_units / ragged_filter_units(ng[t]),
// tile t's rows from unit off[t] of dst.
void launch_ragged_records(int64_t n, int T, int KP, int upg, const int32_t* ng,
                           const int64_t* off, const float* src, float* dst, hipStream_t stream);

// Whether idx[i * k] == i for every node i < n (the node is its own first
// neighbour); synchronises the stream.
bool stencils_self_first(int64_t n, int k, const uint32_t* idx, hipStream_t stream);

// Per-node stencil widths: the entries s >= kvar[i] of node i (zero weights)
// are pointed at i itself, so they add no rows to the tile unions and no
// node to the PML divergence list.
void launch_self_past_k(int64_t n, int k, const uint8_t* kvar, uint32_t* idx, hipStream_t stream);

// RPIM interpolation weights: w[q*k + s]; the queries of a group of `group`
// consecutive queries share the stencil idx[(q / group)*k ..]
// (radial_point_method::compute_weights' group_size).
// solve_bits (both weight kernels): 64 solves the (k + D + 1)^2 systems in
// FP64; 32 rounds them to FP32 and solves in FP32, the reference's solve
// precision (linalg/batched_matrix_inverse.cuh:229-250).
void launch_rpim_weights(int D, const float* src, const float* query,
                         int64_t n_q, const uint32_t* idx, int k, int rbf,
                         double spacing, float* w, double* w64, int* info,
                         hipStream_t stream, int solve_bits = 64, int group = 1);

// Normalised Gaussian filter weights over the stencil (sigma_g = h/2.5).
void launch_filter_weights(int D, const float* pts, int64_t n,
                           const uint32_t* idx, int k, double h, float* wf,
                           hipStream_t stream, const uint8_t* kvar = nullptr);

// Interleave indices and weights into slab records:
// rec[(i*k + s)*(1+ncomp) + 0] = idx bits, then ncomp weights.
void launch_make_slab(int64_t n, int k, int ncomp, const uint32_t* idx,
                      const float* w, float* slab, hipStream_t stream);

// Packs D3Q27 divergence records ({union position bits, wx, wy, wz} per
// stencil entry, k per node) into the packed_div_units layout.
void launch_pack_div_records(int64_t n, int k, const float* src, float* dst, hipStream_t st);
// Packs filter records ({union position bits, w} per stencil entry) into the
// packed_filter_units layout.
void launch_pack_filter_records(int64_t n, int k, const float* src, float* dst,
                                hipStream_t st);

// out[e] = idx[e] where it lies in [lo, hi), else zero_row (n entries).
void launch_layer_only(int64_t n, const uint32_t* idx, uint32_t lo, uint32_t hi,
                       uint32_t zero_row, uint32_t* out, hipStream_t stream);

// Tile unions: fluid nodes are grouped in tiles of tile_nodes consecutive
// nodes (k * tile_nodes <= 4096); a tile's union is the sorted list of the
// distinct neighbour indices of its nodes' k-stencils.  A tile whose union
// exceeds cap (> 0) while each half-tile's union fits is split: its list
// is the two half-tiles' unions one after the other and usplit[t] the
// first one's length (0: not split).  Counting pass (ulist == nullptr):
// counts[t] = list length and usplit[t].  Filling pass (uoff = exclusive
// scan of counts, ntiles + 1 entries; usplit from the counting pass):
// ulist[uoff[t] ..] = the list and, for each slab given (records of
// 1 + ncomp floats, field 0 the neighbour index), the index field is
// replaced by the neighbour's position in its (half-)tile's union.
void launch_tile_unions(int64_t nfluid, int k, int tile_nodes, const uint32_t* idx, int cap,
                        int32_t* counts, int32_t* usplit, const int64_t* uoff, uint32_t* ulist,
                        float* slab_a, int ncomp_a, float* slab_b, int ncomp_b,
                        hipStream_t stream);

// flag[i] = 1 if any stencil neighbour of fluid node i is in [lo, hi).
void launch_flag_neighbours(int64_t n, int k, const uint32_t* idx, int64_t lo,
                            int64_t hi, uint8_t* flag, hipStream_t stream);

// Divergence apply (divergence_operator::apply with the advection field
// map): out[i] = sum_s sum_d W(d,s,i) F_d(idx(s,i)), F given per node.
void launch_divergence_apply(int D, int64_t n, int k, const uint32_t* idx,
                             const float* W, const float* F, float off,
                             float* out, hipStream_t stream);

}  // namespace naga_amd
