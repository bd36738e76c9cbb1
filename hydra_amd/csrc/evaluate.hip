// PhaseSpace::Evaluate on gfx950: per event the GENBOD weight and the values
// of up to HPSP_MAX_EVAL_FUNCTORS functors, written as SoA columns; the
// momenta stay in registers.  Built with -ffp-contract=off.
//
//   Evaluate(mother, result, functors...)   EvalMother.h:283-306
//     the event's engine is the PhaseSpace seed discarded by evt + 3N, as in
//     Generate: the wave shares its squares words (EventWords).
//   Evaluate(mothers, result, functors...)  EvalMothers.h:145-273
//     each event seeds its own engine with pair_hash(evt, seed) and does not
//     discard, so every lane evaluates its own words (LaneWords); fECM and the
//     weight normalisation come from that event's mother.
#include <array>
#include <utility>

#include "genbod_f64.h"

namespace hpsp {

constexpr int kNumN = HPSP_MAX_N - 1;
constexpr int kNumE = 6;

template <int N, int E>
static void eval_one(bool mothers, const EvalArgs& in, hipStream_t s) {
  EvalArgs a = in;
  // runs over one mother (as Generate); over mothers each lane seeds its own
  // engine, one group per wave
  const uint64_t ngroups = (a.n + kWave - 1) / kWave;
  a.run = mothers ? 1u : run_groups(ngroups);
  const unsigned g = grid_for_runs((ngroups + a.run - 1) / a.run);
  if (mothers)
    hipLaunchKernelGGL((k_evaluate_f64<N, E, true>), dim3(g), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_evaluate_f64<N, E, false>), dim3(g), dim3(kBlock), 0, s, a);
}

template <int I>
static constexpr auto eval_at = &eval_one<2 + I / kNumE, I % kNumE>;

template <int... I>
static constexpr auto eval_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(bool, const EvalArgs&, hipStream_t), sizeof...(I)>{eval_at<I>...};
}

static const auto kEval = eval_table(std::make_integer_sequence<int, kNumN * kNumE>{});

int launch_evaluate_f64(int N, int engine, bool mothers, const EvalArgs& a, hipStream_t s) {
  kEval[(N - 2) * kNumE + engine](mothers, a, s);
  return (int)hipGetLastError();
}

}  // namespace hpsp
