// Standalone nonlocal advection operator and CSR matrices (calculus.hip).
#pragma once

#include "device_buffer.hpp"

namespace naga_amd {

class Advection {
  public:
    // advection_operator::create(domain) (advection.cuh:143-151): kNN
    // stencils and divergence weights over the points
    Advection(const float* pts, int64_t n, int D, int k, int rbf);
    // advection_operator::create(divergence_op) (advection.cuh:153-159):
    // from existing stencils (n x k) and weights (n x k x D)
    Advection(const uint32_t* idx, const float* W, int64_t n, int D, int k);
    ~Advection();
    Advection(const Advection&) = delete;
    Advection& operator=(const Advection&) = delete;
    // nsteps RK4 steps of df/dt = div(-v (f - off)) on [0, end) (all nodes
    // when end <= 0); velocity constant (D values) or per node (n x D)
    void step(const float* velocity, bool per_node, float* f, double dt, double off,
              int64_t end, int nsteps);
    void weights(uint32_t* idx, float* W);
    int64_t size() const { return n_; }
    int stencil() const { return k_; }
    int dims() const { return D_; }

  private:
    int64_t n_;
    int D_, k_;
    hipStream_t st_ = nullptr;
    DeviceBuffer<uint32_t> idx_;
    DeviceBuffer<float> W_;
};

class Csr {
  public:
    Csr(int64_t rows, int64_t cols, const int64_t* row_ptr, const uint32_t* cols_idx,
        const float* vals);
    ~Csr();
    Csr(const Csr&) = delete;
    Csr& operator=(const Csr&) = delete;
    // y = alpha A x + beta y
    void mult(float alpha, float beta, const float* x, float* y);

  private:
    int64_t rows_, cols_;
    hipStream_t st_ = nullptr;
    DeviceBuffer<int64_t> rp_;
    DeviceBuffer<uint32_t> ci_;
    DeviceBuffer<float> vals_;
};

}  // namespace naga_amd
