// Standalone operators off the engine: the nonlocal advection operator
// (nonlocal_calculus/advection.cuh:88-524) and CSR matrices built from
// index stencils with their matrix-vector product
// (linalg/matrix.cuh:300-530).
#include "calculus.hpp"

#include <vector>

#include "setup_kernels.hpp"

namespace naga_amd {

namespace {

constexpr int kThreads = 256;

// rate_i = sum_s sum_d W[i][s][d] * (-v_d(j) (f_j - off)), j = idx[i][s]
// (divergence_field_map, advection.cuh:102-141); v constant (vstride 0)
// or per node.
__global__ void k_adv_rate(int64_t n, int k, int D, const uint32_t* idx, const float* W,
                           const float* v, int vstride, const float* f, float off,
                           float* rate) {
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
    for (int s = 0; s < k; ++s) {
        const int64_t j = idx[i * k + s];
        const float fj = -(f[j] - off);
        const float* vj = v + j * vstride;
        float t = 0.0f;
        for (int d = 0; d < D; ++d) t += W[(i * k + s) * D + d] * (vj[d] * fj);
        acc += t;
    }
    rate[i] = acc;
}

// f = f0 + c * r on [0, end)
__global__ void k_axpy(int64_t end, const float* f0, const float* r, float c, float* f) {
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i < end) f[i] = f0[i] + c * r[i];
}

// f = f0 + dt (r1/6 + r2/3 + r3/3 + r4/6) on [0, end)
__global__ void k_rk4_final(int64_t end, const float* f0, const float* r1, const float* r2,
                            const float* r3, const float* r4, float dt, float* f) {
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= end) return;
    const float s = r1[i] * (1.0f / 6.0f) + r2[i] * (1.0f / 3.0f) + r3[i] * (1.0f / 3.0f) +
                    r4[i] * (1.0f / 6.0f);
    f[i] = f0[i] + dt * s;
}

// y = alpha A x + beta y, one thread per row, entries in row order
__global__ void k_csr_spmv(int64_t rows, const int64_t* rp, const uint32_t* cols,
                           const float* vals, const float* x, float alpha, float beta,
                           float* y) {
    const int64_t r = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (r >= rows) return;
    float acc = 0.0f;
    for (int64_t e = rp[r]; e < rp[r + 1]; ++e) acc += vals[e] * x[cols[e]];
    y[r] = beta == 0.0f ? alpha * acc : alpha * acc + beta * y[r];
}

int64_t blocks(int64_t n) { return (n + kThreads - 1) / kThreads; }

}  // namespace

Advection::Advection(const float* pts, int64_t n, int D, int k, int rbf) : n_(n), D_(D), k_(k) {
    require(D == 2 || D == 3, "dims must be 2 or 3");
    require(n > k && k >= 2 && k + D + 1 <= 64, "bad stencil width");
    require(rbf == 0 || rbf == 1, "rbf must be 0 or 1");
    NAGA_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    try {
        DeviceBuffer<float> dp(n * D);
        dp.upload(pts, n * D, st_);
        const DeviceCells cells = bin_points_device(D, dp.get(), n, st_);
        DeviceBuffer<float> d2(n * k);
        idx_.allocate(n * k);
        launch_knn(cells.bins, dp.get(), dp.get(), n, k, d2.get(), idx_.get(), st_);
        // multiquadric shape parameter per 16384-query batch (divergence.cuh:118-160)
        const int64_t nbatch = (n + 16383) / 16384;
        std::vector<double> hbar(nbatch, 1.0);
        if (rbf == 0) {
            std::vector<float> hd2(n * k);
            d2.download(hd2.data(), n * k, st_);
            NAGA_HIP(hipStreamSynchronize(st_));
            for (int64_t b = 0; b < nbatch; ++b) {
                const int64_t e = std::min<int64_t>((b + 1) * 16384, n);
                double s = 0.0;
                for (int64_t i = b * 16384; i < e; ++i) s += std::sqrt(double(hd2[i * k + 1]));
                hbar[b] = s / double(e - b * 16384);
            }
        }
        auto dh = to_device(hbar, st_);
        W_.allocate(n * k * D);
        DeviceBuffer<int> info(n);
        launch_divergence_weights(D, dp.get(), dp.get(), n, idx_.get(), d2.get(), k, rbf, dh.get(),
                                  W_.get(), nullptr, info.get(), st_);
        std::vector<int> hi(n);
        info.download(hi.data(), n, st_);
        NAGA_HIP(hipStreamSynchronize(st_));
        for (int v : hi) require(v == 0, "singular RBF system while building the operator");
    } catch (...) {
        (void)hipStreamDestroy(st_);
        throw;
    }
}

Advection::Advection(const uint32_t* idx, const float* W, int64_t n, int D, int k)
    : n_(n), D_(D), k_(k) {
    require(idx && W, "null argument");
    require(D == 2 || D == 3, "dims must be 2 or 3");
    require(n > 0 && k >= 1, "bad operator shape");
    for (int64_t t = 0; t < n * k; ++t) require(int64_t(idx[t]) < n, "stencil index out of range");
    NAGA_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    try {
        idx_.allocate(n * k);
        idx_.upload(idx, n * k, st_);
        W_.allocate(n * k * D);
        W_.upload(W, n * k * D, st_);
        NAGA_HIP(hipStreamSynchronize(st_));
    } catch (...) {
        (void)hipStreamDestroy(st_);
        throw;
    }
}

Advection::~Advection() {
    if (st_) {
        (void)hipStreamSynchronize(st_);
        (void)hipStreamDestroy(st_);
    }
}

// advection_task::impl (advection.cuh:207-276): classic RK4 with stage
// weights 1/2, 1/2, 1 and summation weights 1/6, 1/3, 1/3, 1/6; nodes at and
// past end keep their value.
void Advection::step(const float* velocity, bool per_node, float* f, double dt, double off,
                     int64_t end, int nsteps) {
    require(velocity && f, "null argument");
    require(nsteps >= 0, "nsteps must be >= 0");
    const int64_t n = n_;
    const int64_t e = end > 0 ? std::min(end, n) : n;
    DeviceBuffer<float> v(per_node ? n * D_ : D_), f0(n), fs(n), r1(n), r2(n), r3(n), r4(n);
    v.upload(velocity, per_node ? n * D_ : D_, st_);
    f0.upload(f, n, st_);
    const int vs = per_node ? D_ : 0;
    const float fdt = float(dt), fo = float(off);
    NAGA_HIP(hipMemcpyAsync(fs.get(), f0.get(), n * sizeof(float), hipMemcpyDeviceToDevice, st_));
    for (int s = 0; s < nsteps; ++s) {
        float* rates[4] = {r1.get(), r2.get(), r3.get(), r4.get()};
        const float cw[3] = {0.5f * fdt, 0.5f * fdt, fdt};
        k_adv_rate<<<blocks(n), kThreads, 0, st_>>>(n, k_, D_, idx_.get(), W_.get(), v.get(), vs,
                                                    f0.get(), fo, rates[0]);
        for (int i = 0; i < 3; ++i) {
            k_axpy<<<blocks(e), kThreads, 0, st_>>>(e, f0.get(), rates[i], cw[i], fs.get());
            k_adv_rate<<<blocks(n), kThreads, 0, st_>>>(n, k_, D_, idx_.get(), W_.get(), v.get(),
                                                        vs, fs.get(), fo, rates[i + 1]);
        }
        k_rk4_final<<<blocks(e), kThreads, 0, st_>>>(e, f0.get(), r1.get(), r2.get(), r3.get(),
                                                      r4.get(), fdt, fs.get());
        NAGA_HIP(hipMemcpyAsync(f0.get(), fs.get(), n * sizeof(float), hipMemcpyDeviceToDevice,
                                st_));
    }
    hip_check(hipGetLastError(), "advection step");
    f0.download(f, n, st_);
    NAGA_HIP(hipStreamSynchronize(st_));
}

void Advection::weights(uint32_t* idx, float* W) {
    if (idx) idx_.download(idx, n_ * k_, st_);
    if (W) W_.download(W, n_ * k_ * D_, st_);
    NAGA_HIP(hipStreamSynchronize(st_));
}

Csr::Csr(int64_t rows, int64_t cols, const int64_t* rp, const uint32_t* ci, const float* v)
    : rows_(rows), cols_(cols) {
    require(rows > 0 && cols > 0 && rp && ci && v, "bad CSR arguments");
    require(rp[0] == 0, "row offsets must start at 0");
    for (int64_t r = 0; r < rows; ++r) require(rp[r + 1] >= rp[r], "row offsets must not decrease");
    const int64_t nnz = rp[rows];
    for (int64_t e = 0; e < nnz; ++e) require(int64_t(ci[e]) < cols, "column index out of range");
    NAGA_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    rp_.allocate(rows + 1);
    rp_.upload(rp, rows + 1, st_);
    ci_.allocate(std::max<int64_t>(nnz, 1));
    vals_.allocate(std::max<int64_t>(nnz, 1));
    if (nnz) {
        ci_.upload(ci, nnz, st_);
        vals_.upload(v, nnz, st_);
    }
    NAGA_HIP(hipStreamSynchronize(st_));
}

Csr::~Csr() {
    if (st_) {
        (void)hipStreamSynchronize(st_);
        (void)hipStreamDestroy(st_);
    }
}

void Csr::mult(float alpha, float beta, const float* x, float* y) {
    require(x && y, "null argument");
    DeviceBuffer<float> dx(cols_), dy(rows_);
    dx.upload(x, cols_, st_);
    if (beta != 0.0f) dy.upload(y, rows_, st_);
    k_csr_spmv<<<blocks(rows_), kThreads, 0, st_>>>(rows_, rp_.get(), ci_.get(), vals_.get(),
                                                    dx.get(), alpha, beta, dy.get());
    hip_check(hipGetLastError(), "csr spmv");
    dy.download(y, rows_, st_);
    NAGA_HIP(hipStreamSynchronize(st_));
}

}  // namespace naga_amd
