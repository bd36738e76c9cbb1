// Entry to the user-functor path.  Under a HIP compiler the kernels that run
// a user's callable are instantiated in the caller's translation unit
// (UserKernels.h); under any other compiler a call with such a functor fails
// to compile with a message that says so, because the callable must be
// compiled for the device (the reference compiles every functor into its
// Thrust kernels, hydra/Lambda.h:133-200).
#pragma once

#include <cstddef>
#include <cstdint>
#include <type_traits>

#include <hydra/Lambda.h>
#include <hydra/Tuple.h>

#include "hpsp.h"

#if defined(__HIP__)
#include <hydra/detail/UserKernels.h>
#else
namespace hydra {
namespace detail {

template <typename... F>
struct needs_hip_compiler : std::false_type {};

#define HYDRA_USER_FUNCTOR_NEEDS_HIP                                                     \
  "this functor is not one of the device-described phase-space functors "              \
  "(hydra/PhaseSpaceFunctors.h): it is compiled into the kernels, so this translation " \
  "unit must be built with hipcc"

template <size_t N, int ENGINE, bool MOTHERS, typename F>
hpsp_stats user_average_on(hpsp_decay const&, uint64_t, F const&, uint64_t, uint64_t,
                           const double*, uint64_t, int) {
  static_assert(needs_hip_compiler<F>::value, HYDRA_USER_FUNCTOR_NEEDS_HIP);
  return hpsp_stats{};
}

template <size_t N, int ENGINE, bool MOTHERS, typename... F>
void user_evaluate(hpsp_decay const&, uint64_t, const double*, uint64_t, uint64_t, uint64_t,
                   double*, uint64_t, F const&...) {
  static_assert(needs_hip_compiler<F...>::value, HYDRA_USER_FUNCTOR_NEEDS_HIP);
}

template <typename... P, typename... F>
void user_event_values(hydra::tuple<P...>*, hpsp_decay const&, const double*, uint64_t, uint64_t,
                       bool, double*, uint64_t, F const&...) {
  static_assert(needs_hip_compiler<F...>::value, HYDRA_USER_FUNCTOR_NEEDS_HIP);
}

template <typename Types, typename... F>
constexpr int user_width() {
  static_assert(needs_hip_compiler<F...>::value, HYDRA_USER_FUNCTOR_NEEDS_HIP);
  return 1;
}

}  // namespace detail
}  // namespace hydra
#endif
