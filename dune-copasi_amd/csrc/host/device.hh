// Owning device buffers for the model's HBM-resident data.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "error.hh"

namespace dcopasi {

template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  ~DeviceArray() { release(); }

  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) DCG_HIP(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)));
    n_ = n;
  }
  void upload(const std::vector<T>& h, hipStream_t s) {
    resize(h.size());
    if (!h.empty()) DCG_HIP(hipMemcpyAsync(p_, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace dcopasi
