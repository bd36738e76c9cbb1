// Host runtime glue of the C++ API over the C ABI (include/hpsp.h):
// status -> exception mapping and an owning device buffer.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>

#include "hpsp.h"

namespace hydra {
namespace detail {

// The reference reports invalid masses with std::invalid_argument and too
// little energy with std::runtime_error (PhaseSpace.inl:49-76); device
// failures surface as std::runtime_error (thrust::system_error derives from it).
inline void hpsp_check(int rc, const char* where) {
  if (rc == HPSP_OK) return;
  const std::string msg = std::string("[hydra::") + where + "]: " + hpsp_error_string(rc);
  if (rc == HPSP_ERR_MASS || rc == HPSP_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// Calls are synchronous to the caller, as the reference's are
// (parallel_for.h:164-166 synchronises after each launch).
inline void hpsp_sync(const char* where) { hpsp_check(hpsp_stream_synchronize(nullptr), where); }

template <typename T>
class device_buffer {
 public:
  device_buffer() = default;
  explicit device_buffer(size_t n) { allocate(n); }
  device_buffer(const device_buffer& o) {
    allocate(o.n_);
    if (n_) hpsp_check(hpsp_memcpy_device(p_, o.p_, n_ * sizeof(T), nullptr), "device_buffer");
    hpsp_sync("device_buffer");
  }
  device_buffer(device_buffer&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  device_buffer& operator=(device_buffer o) noexcept {
    swap(o);
    return *this;
  }
  ~device_buffer() {
    if (p_) hpsp_free(p_);
  }
  void swap(device_buffer& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
  }
  void allocate(size_t n) {
    if (p_) hpsp_free(p_);
    p_ = nullptr;
    n_ = n;
    if (n) {
      void* v = nullptr;
      hpsp_check(hpsp_malloc(&v, n * sizeof(T)), "device_buffer::allocate");
      p_ = static_cast<T*>(v);
    }
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace detail
}  // namespace hydra
