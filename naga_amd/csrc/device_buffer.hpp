// Owning device allocation (hipMalloc/hipFree) with typed access.
#pragma once

#include <utility>
#include <vector>

#include "naga_common.hpp"

namespace naga_amd {

template <class T>
class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { allocate(n); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept { swap(o); }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        DeviceBuffer t(std::move(o));
        swap(t);
        return *this;
    }
    ~DeviceBuffer() { release(); }

    void allocate(size_t n) {
        release();
        if (n == 0) return;
        void* p = nullptr;
        NAGA_HIP(hipMalloc(&p, n * sizeof(T)));
        ptr_ = static_cast<T*>(p);
        n_ = n;
    }
    void release() {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        n_ = 0;
    }
    void zero(hipStream_t s) {
        if (n_) NAGA_HIP(hipMemsetAsync(ptr_, 0, n_ * sizeof(T), s));
    }
    void upload(const T* src, size_t n, hipStream_t s, size_t offset = 0) {
        require(offset + n <= n_, "upload out of range");
        if (n) NAGA_HIP(hipMemcpyAsync(ptr_ + offset, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* dst, size_t n, hipStream_t s, size_t offset = 0) const {
        require(offset + n <= n_, "download out of range");
        if (n) NAGA_HIP(hipMemcpyAsync(dst, ptr_ + offset, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    T* get() const { return ptr_; }
    size_t size() const { return n_; }
    void swap(DeviceBuffer& o) noexcept {
        std::swap(ptr_, o.ptr_);
        std::swap(n_, o.n_);
    }

  private:
    T* ptr_ = nullptr;
    size_t n_ = 0;
};

template <class T>
DeviceBuffer<T> to_device(const std::vector<T>& v, hipStream_t s) {
    DeviceBuffer<T> b(v.size());
    b.upload(v.data(), v.size(), s);
    return b;
}

}  // namespace naga_amd
