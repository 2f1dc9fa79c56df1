// Standalone setup operators (see operators.hpp).
#include "operators.hpp"

#include <vector>

namespace naga_amd {

namespace {

// One thread per (query, component).
__global__ __launch_bounds__(256) void k_rpim_interpolate(int64_t n_q, int m, int k, int group,
                                                          const uint32_t* idx, const float* w,
                                                          const float* field, float off,
                                                          float* dest) {
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= n_q * m) return;
    const int64_t q = t / m;
    const int c = int(t - q * m);
    const uint32_t* row = idx + (q / group) * k;
    float sum = 0.0f;
    for (int s = 0; s < k; ++s) sum += w[q * k + s] * (field[int64_t(row[s]) * m + c] - off);
    dest[t] = sum + off;
}

}  // namespace

Segmentation::Segmentation(const float* pts, int64_t n, int D, int approx, double width)
    : D_(D), n_(n) {
    require(D >= 1 && D <= 3, "The number of dimensions must be 1, 2 or 3");
    require(pts != nullptr && n > 0, "no points to segment");
    require(approx > 0 || width > 0.0, "approx_partition_size or partition_width must be > 0");
    NAGA_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    try {
        pts_.allocate(size_t(n) * D);
        pts_.upload(pts, size_t(n) * D, st_);
        cells_ = bin_points_device(D, pts_.get(), n, st_, approx > 0 ? double(approx) : 0.0,
                                   approx > 0 ? 0.0 : width);
        NAGA_HIP(hipMemcpyAsync(&grid_, cells_.bins.grid, sizeof(CellGrid),
                                hipMemcpyDeviceToHost, st_));
        NAGA_HIP(hipStreamSynchronize(st_));
    } catch (...) {
        (void)hipStreamDestroy(st_);
        throw;
    }
}

Segmentation::~Segmentation() {
    if (st_) {
        (void)hipStreamSynchronize(st_);
        (void)hipStreamDestroy(st_);
    }
}

void Segmentation::partitions(uint32_t* sizes, uint32_t* offsets, uint32_t* indices) {
    const size_t nc = size_t(grid_.ncell);
    if (sizes)
        NAGA_HIP(hipMemcpyAsync(sizes, cells_.bins.counts, nc * 4, hipMemcpyDeviceToHost, st_));
    if (offsets)
        NAGA_HIP(hipMemcpyAsync(offsets, cells_.bins.start, nc * 4, hipMemcpyDeviceToHost, st_));
    if (indices)
        NAGA_HIP(hipMemcpyAsync(indices, cells_.bins.sorted, size_t(n_) * 4,
                                hipMemcpyDeviceToHost, st_));
    NAGA_HIP(hipStreamSynchronize(st_));
}

void Segmentation::knn(const float* query, int64_t n_q, int k, float* d2, uint32_t* idx) {
    require(D_ == 2 || D_ == 3, "the kNN takes 2 or 3 dimensions");
    require(k >= 1 && k <= 64, "k must be in [1, 64]");
    require(n_ >= k, "fewer reference points than neighbours requested");
    require(n_q >= 0 && (n_q == 0 || (query && d2 && idx)), "null argument");
    if (n_q == 0) return;
    DeviceBuffer<float> dq(size_t(n_q) * D_), dd2(size_t(n_q) * k);
    DeviceBuffer<uint32_t> didx(size_t(n_q) * k);
    dq.upload(query, dq.size(), st_);
    launch_knn(cells_.bins, pts_.get(), dq.get(), n_q, k, dd2.get(), didx.get(), st_);
    dd2.download(d2, dd2.size(), st_);
    didx.download(idx, didx.size(), st_);
    NAGA_HIP(hipStreamSynchronize(st_));
}

Rpim::Rpim(const float* src, int64_t n_src, const uint32_t* idx, int k, const float* query,
           int64_t n_q, int D, int group, int rbf, double spacing, int solve_bits)
    : n_src_(n_src), n_q_(n_q), k_(k), group_(group) {
    require(src && idx && query, "null argument");
    require(D == 2 || D == 3, "dims must be 2 or 3");
    require(k >= 1 && k + D + 1 <= 64, "support size + dims + 1 must be <= 64");
    require(group >= 1 && n_q % group == 0, "query_size must be a multiple of group_size");
    require(rbf == 0 || rbf == 1, "rbf must be 0 (multiquadric) or 1 (r^3)");
    require(spacing > 0.0, "approx_particle_spacing must be positive");
    const int64_t ng = n_q / group;
    for (int64_t t = 0; t < ng * k; ++t) require(idx[t] < n_src, "interpolating index out of range");
    NAGA_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    try {
        DeviceBuffer<float> ds(size_t(n_src) * D), dq(size_t(n_q) * D);
        DeviceBuffer<int> info(std::max<int64_t>(n_q, 1));
        ds.upload(src, ds.size(), st_);
        dq.upload(query, dq.size(), st_);
        idx_.allocate(size_t(ng) * k);
        idx_.upload(idx, idx_.size(), st_);
        w_.allocate(size_t(n_q) * k);
        launch_rpim_weights(D, ds.get(), dq.get(), n_q, idx_.get(), k, rbf, spacing, w_.get(),
                            nullptr, info.get(), st_, solve_bits, group);
        std::vector<int> hi(n_q);
        if (n_q) info.download(hi.data(), n_q, st_);
        NAGA_HIP(hipStreamSynchronize(st_));
        for (int v : hi) require(v == 0, "singular RPIM system");
    } catch (...) {
        (void)hipStreamDestroy(st_);
        st_ = nullptr;
        throw;
    }
}

Rpim::~Rpim() {
    if (st_) {
        (void)hipStreamSynchronize(st_);
        (void)hipStreamDestroy(st_);
    }
}

void Rpim::interpolate(const float* field, int m, float* dest, float off) {
    require(field && dest && m >= 1, "null argument or no components");
    DeviceBuffer<float> df(size_t(n_src_) * m), dd(size_t(n_q_) * m);
    df.upload(field, df.size(), st_);
    const int64_t tot = n_q_ * m;
    if (tot > 0) {
        k_rpim_interpolate<<<(tot + 255) / 256, 256, 0, st_>>>(n_q_, m, k_, group_, idx_.get(),
                                                                w_.get(), df.get(), off, dd.get());
        hip_check(hipGetLastError(), "rpim interpolate");
    }
    dd.download(dest, dd.size(), st_);
    NAGA_HIP(hipStreamSynchronize(st_));
}

void Rpim::weights(float* w) const {
    require(w != nullptr, "null argument");
    w_.download(w, w_.size(), st_);
    NAGA_HIP(hipStreamSynchronize(st_));
}

}  // namespace naga_amd
