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

// Column stride of the device containers: the event count rounded up to a
// multiple of 64, so that every column starts 512 bytes after the previous
// one and the kernels' groups of 64 events fill whole cache lines.  With a
// stride of n itself every column but the first starts mid-line unless n is a
// multiple of 16, and each group's stores then leave partial lines, which
// HBM completes by read-modify-write: a 3-body Generate of 2^27 events took
// 3.0-4.7 ms at a stride of n + 1 against 2.35 ms at n.
inline size_t column_stride(size_t n) { return n ? (n + 63) / 64 * 64 : 64; }

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
