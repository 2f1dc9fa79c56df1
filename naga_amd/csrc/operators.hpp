// Standalone setup operators under the reference's names: the cell
// segmentation with kNN over it, and the RPIM interpolator.
#pragma once

#include "device_buffer.hpp"
#include "setup_kernels.hpp"

namespace naga_amd {

// nd_cubic_segmentation (segmentation/nd_cubic_segmentation.cuh:219-461)
// of points held on the device, and batched_nearest_neighbors over it
// (segmentation/nearest_neighbors.cuh:454-477).
class Segmentation {
  public:
    // approx > 0: cells of about `approx` points; else cells of `width`
    Segmentation(const float* pts, int64_t n, int D, int approx, double width);
    ~Segmentation();
    Segmentation(const Segmentation&) = delete;
    Segmentation& operator=(const Segmentation&) = delete;

    const CellGrid& grid() const { return grid_; }
    int64_t points() const { return n_; }
    int dims() const { return D_; }
    // sizes[c], offsets[c] (exclusive scan of the sizes) for every cell and
    // indices[n]: the points of cell c at indices[offsets[c] ..)
    void partitions(uint32_t* sizes, uint32_t* offsets, uint32_t* indices);
    void knn(const float* query, int64_t n_q, int k, float* d2, uint32_t* idx);

  private:
    int D_;
    int64_t n_;
    hipStream_t st_ = nullptr;
    DeviceBuffer<float> pts_;
    DeviceCells cells_;
    CellGrid grid_{};
};

// radial_point_method<T>::create_interpolator / interpolate
// (interpolation/radial_point_method.cuh:143-251): weights of each query
// over its group's stencil, kept on the device with the stencils.
class Rpim {
  public:
    Rpim(const float* src, int64_t n_src, const uint32_t* idx, int k, const float* query,
         int64_t n_q, int D, int group, int rbf, double spacing, int solve_bits);
    ~Rpim();
    Rpim(const Rpim&) = delete;
    Rpim& operator=(const Rpim&) = delete;

    // dest[q][c] = sum_s w(s, q) (field[idx(s, q / group)][c] - off) + off for
    // m components c (field [n_src][m], dest [n_q][m])
    void interpolate(const float* field, int m, float* dest, float off);
    void weights(float* w) const;
    int64_t sources() const { return n_src_; }
    int64_t queries() const { return n_q_; }
    int support() const { return k_; }

  private:
    int64_t n_src_, n_q_;
    int k_, group_;
    hipStream_t st_ = nullptr;
    DeviceBuffer<uint32_t> idx_;
    DeviceBuffer<float> w_;
};

}  // namespace naga_amd
