// Functor values of stored events: `events | hydra::make_functors(f...)`,
// the materialisable form of the reference's lazy `events | functor`
// (Range2.inl:213-238), over a Decays container or melded containers
// (Decays::Meld, phsp_chain.inl:244).  The functors are the device-described
// ones of PhaseSpaceFunctors.h; indices count the particles of the source.
//   view[i]                       the i-th row of values, computed on the device
//   hydra::copy(view, result)     all rows into a device multivector with one
//                                 column per functor (phsp_chain.inl:260-262)
#pragma once

#include <array>
#include <cstddef>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include <hydra/Decays.h>
#include <hydra/PhaseSpaceFunctors.h>
#include <hydra/Tuple.h>
#include <hydra/detail/Runtime.h>
#include <hydra/device/System.h>
#include <hydra/multivector.h>

#include "hpsp.h"

namespace hydra {

template <typename... F>
struct FunctorsSpec {
  std::array<hpsp_functor, sizeof...(F)> functors;
};

template <typename... F>
FunctorsSpec<F...> make_functors(F const&... f) {
  static_assert(sizeof...(F) >= 1 && sizeof...(F) <= HPSP_MAX_EVAL_FUNCTORS,
                "one to HPSP_MAX_EVAL_FUNCTORS functors");
  static_assert(std::conjunction<is_phase_space_functor<F>...>::value,
                "functors must be device-described phase-space functors");
  return FunctorsSpec<F...>{{detail::to_hpsp(f)...}};
}

// K functor values per stored event, not materialised.
template <size_t K>
class FunctorColumns {
  template <size_t... I>
  static auto row_type(std::index_sequence<I...>) -> hydra::tuple<decltype((void)I, double())...>;

 public:
  using value_type = decltype(row_type(std::make_index_sequence<K>{}));

  FunctorColumns(hpsp_particles const& p, size_t n, std::array<hpsp_functor, K> const& f)
      : p_(p), n_(n), f_(f) {}
  size_t size() const { return n_; }

  // rows [first, first + n) into K device columns of ld doubles (asynchronous)
  void evaluate(size_t first, size_t n, double* out, size_t ld) const {
    if (first > n_ || n > n_ - first)
      throw std::out_of_range("[hydra::make_functors]: rows past the end");
    hpsp_particles q = p_;
    for (int i = 0; i < q.n; ++i) q.cols[i] += first;
    detail::hpsp_check(hpsp_particles_evaluate(&q, f_.data(), (int)K, n, out, ld, nullptr),
                       "make_functors");
  }

  value_type operator[](size_t i) const {
    detail::device_buffer<double> buf(K);
    evaluate(i, 1, buf.data(), 1);
    std::array<double, K> h;
    detail::hpsp_check(hpsp_memcpy_to_host(h.data(), buf.data(), K * sizeof(double)),
                       "make_functors");
    return make(h, std::make_index_sequence<K>{});
  }

 private:
  template <size_t... I>
  static value_type make(const std::array<double, K>& h, std::index_sequence<I...>) {
    return value_type(h[I]...);
  }

  hpsp_particles p_;
  size_t n_;
  std::array<hpsp_functor, K> f_;
};

template <typename... Particles, typename... F>
FunctorColumns<sizeof...(F)> operator|(
    Decays<hydra::tuple<Particles...>, hydra::device::sys_t> const& events,
    FunctorsSpec<F...> const& s) {
  return FunctorColumns<sizeof...(F)>(events.particles(), events.size(), s.functors);
}

template <size_t P, typename... F>
FunctorColumns<sizeof...(F)> operator|(MeldedDecays<P> const& events, FunctorsSpec<F...> const& s) {
  return FunctorColumns<sizeof...(F)>(events.particles(), events.size(), s.functors);
}

// hydra::copy(events | make_functors(...), result): one column per functor;
// result holds at least as many rows as the source.
template <size_t K, typename... T>
void copy(FunctorColumns<K> const& src, multivector<hydra::tuple<T...>, hydra::device::sys_t>& dst) {
  static_assert(sizeof...(T) == K, "one result column per functor");
  if (dst.size() < src.size()) throw std::invalid_argument("[hydra::copy]: result too short");
  src.evaluate(0, src.size(), dst.data(), dst.ld());
  detail::hpsp_sync("hydra::copy");
}

}  // namespace hydra
