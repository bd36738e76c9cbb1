// hydra::phase_space_range(mother, masses, seed, length) (reference:
// hydra/detail/PhaseSpaceRange.inl:44-72, detail/functors/GenerateDecay.h).
//
// A lazy range of `length` elements (0: unbounded, as in the reference);
// element k is (weight, p_1, ..., p_N), the GENBOD event k of GenerateDecay
// with thrust's default_random_engine (minstd_rand) discarded by k + 3N.  The
// reference's size_t seed passes through an int into the engine's uint32
// state, so only its low 32 bits count.  Nothing is stored: consumers
// generate the elements where they use them.
//   range | hydra::make_axes(f...)    axes for DenseHistogram::Fill, with the
//                                     elements generated inside the fill
//                                     (examples/misc/range_semantics.inl:154-167)
//   range.GetEventWeightFunctor()     the element's weight, as Fill's weights
//   range[i]                          one element, computed on the device
//   hydra::copy(range, events[, w])   elements 0 .. events.size() - 1 into a
//                                     Decays container (momenta) and
//                                     EventWeights (the elements' weights)
#pragma once

#include <array>
#include <cstddef>
#include <limits>
#include <stdexcept>
#include <utility>

#include <hydra/Decays.h>
#include <hydra/Tuple.h>
#include <hydra/Vector4R.h>
#include <hydra/detail/Runtime.h>
#include <hydra/device/System.h>

#include "hpsp.h"

namespace hydra {

// The weight element of a phase_space_range element.
struct RangeEventWeight {};

template <size_t N>
class PhaseSpaceRange {
  template <size_t... I>
  static auto tuple_type(std::index_sequence<I...>)
      -> hydra::tuple<double, decltype((void)I, Vector4R())...>;

 public:
  using value_type = decltype(tuple_type(std::make_index_sequence<N>{}));

  PhaseSpaceRange(Vector4R const& mother, std::array<double, N> const& masses, size_t seed,
                  size_t length)
      : seed_(seed), size_(length == 0 ? std::numeric_limits<size_t>::max() : length) {
    const double m4[4] = {mother.get(0), mother.get(1), mother.get(2), mother.get(3)};
    detail::hpsp_check(hpsp_range_init(&d_, m4, (int)N, masses.data()), "phase_space_range");
  }

  size_t size() const { return size_; }
  const hpsp_decay& descriptor() const { return d_; }
  uint64_t seed() const { return (uint64_t)seed_; }
  RangeEventWeight GetEventWeightFunctor() const { return {}; }

  // Elements [first, first + n) into device columns: momenta in 4N columns
  // of ld doubles, weights in wt (may be null).  Asynchronous.
  void generate(size_t first, size_t n, double* cols, size_t ld, double* wt) const {
    if (first > size_ || n > size_ - first)
      throw std::out_of_range("[hydra::phase_space_range]: elements past the end");
    detail::hpsp_check(hpsp_range_generate(&d_, (uint64_t)seed_, first, n, cols, ld, wt, nullptr),
                       "phase_space_range");
  }

  value_type operator[](size_t i) const {
    detail::device_buffer<double> buf(4 * N + 1);
    generate(i, 1, buf.data(), 1, buf.data() + 4 * N);
    std::array<double, 4 * N + 1> h;
    detail::hpsp_check(hpsp_memcpy_to_host(h.data(), buf.data(), h.size() * sizeof(double)),
                       "phase_space_range");
    return make_element(h, std::make_index_sequence<N>{});
  }

 private:
  template <size_t... I>
  static value_type make_element(const std::array<double, 4 * N + 1>& h,
                                 std::index_sequence<I...>) {
    return value_type(h[4 * N],
                      Vector4R(h[4 * I], h[4 * I + 1], h[4 * I + 2], h[4 * I + 3])...);
  }

  hpsp_decay d_{};
  size_t seed_;
  size_t size_;
};

template <size_t N>
PhaseSpaceRange<N> phase_space_range(Vector4R const& mother, std::array<double, N> masses,
                                     size_t seed, size_t length = 0) {
  return PhaseSpaceRange<N>(mother, masses, seed, length);
}

// The first events.size() elements of the range into a Decays container of the
// same N (its PhaseSpaceWeight is recomputed from them when asked), and the
// elements' own weights into w when given.
template <size_t N, typename... Particles>
void copy(PhaseSpaceRange<N> const& range,
          Decays<hydra::tuple<Particles...>, hydra::device::sys_t>& events,
          EventWeights* w = nullptr) {
  static_assert(sizeof...(Particles) == N, "the container needs N daughters");
  const size_t n = events.size();
  if (w && w->size() < n) throw std::invalid_argument("[hydra::copy]: too few weight slots");
  range.generate(0, n, events.columns_for_overwrite(), events.ld(), w ? w->data() : nullptr);
  events.set_weights_cache_valid(false);
  detail::hpsp_sync("hydra::copy");
}

}  // namespace hydra
