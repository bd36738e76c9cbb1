// hydra::device::vector<Vector4R>: mother four-vectors on the device for
// Generate(mothers, daughters) and AverageOn over mothers (reference: the
// Thrust device vectors of Vector4R the examples pass, e.g.
// examples/phase_space/phsp_chain.inl:217-245).  Stored as four FP64 columns
// (E, px, py, pz) so the kernels read them coalesced.
#pragma once

#include <cstddef>
#include <vector>

#include <hydra/Vector4R.h>
#include <hydra/detail/Runtime.h>

namespace hydra {
namespace device {

template <typename T>
class vector;

template <>
class vector<hydra::Vector4R> {
 public:
  vector() = default;
  explicit vector(size_t n) : n_(n), buf_(4 * (n ? n : 1)) {}
  explicit vector(const std::vector<hydra::Vector4R>& host) : vector(host.size()) {
    assign(host);
  }
  void assign(const std::vector<hydra::Vector4R>& host) {
    if (host.size() != n_) {
      n_ = host.size();
      buf_.allocate(4 * (n_ ? n_ : 1));
    }
    std::vector<double> soa(4 * n_);
    for (size_t k = 0; k < n_; ++k)
      for (int c = 0; c < 4; ++c) soa[c * n_ + k] = host[k].get(c);
    if (n_)
      detail::hpsp_check(hpsp_memcpy_to_device(buf_.data(), soa.data(), soa.size() * sizeof(double)),
                         "device::vector");
  }
  std::vector<hydra::Vector4R> to_host() const {
    std::vector<double> soa(4 * n_);
    if (n_)
      detail::hpsp_check(hpsp_memcpy_to_host(soa.data(), buf_.data(), soa.size() * sizeof(double)),
                         "device::vector");
    std::vector<hydra::Vector4R> h(n_);
    for (size_t k = 0; k < n_; ++k)
      h[k] = hydra::Vector4R(soa[k], soa[n_ + k], soa[2 * n_ + k], soa[3 * n_ + k]);
    return h;
  }
  hydra::Vector4R operator[](size_t k) const {
    double v[4];
    detail::hpsp_check(hpsp_memcpy2d_to_host(v, sizeof(double), buf_.data() + k,
                                             ld() * sizeof(double), sizeof(double), 4),
                       "device::vector");
    return hydra::Vector4R(v[0], v[1], v[2], v[3]);
  }
  size_t size() const { return n_; }
  size_t ld() const { return n_ ? n_ : 1; }
  const double* data() const { return buf_.data(); }
  double* data() { return buf_.data(); }

 private:
  size_t n_ = 0;
  detail::device_buffer<double> buf_;
};

}  // namespace device
}  // namespace hydra
