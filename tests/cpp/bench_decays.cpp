// Times the headline step through the C++ drop-in, as the reference's
// examples/phase_space/phsp_unweighting.inl runs it: PhaseSpace<3>::Generate
// into a device Decays, then Decays::Unweight(), repeated on the same
// container.  Prints one JSON object (bench.py reports it beside the C ABI's
// figure for the same step).
//   bench_decays [events] [steps] [warmup]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <hydra/Decays.h>
#include <hydra/PhaseSpace.h>
#include <hydra/Vector4R.h>
#include <hydra/device/System.h>

declarg(A, hydra::Vector4R)
declarg(B, hydra::Vector4R)
declarg(C, hydra::Vector4R)

using namespace hydra::arguments;

int main(int argc, char** argv) {
  const size_t n = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : (size_t(1) << 27);
  const int steps = argc > 2 ? std::atoi(argv[2]) : 10;
  const int warmup = argc > 3 ? std::atoi(argv[3]) : 2;
  const double P_mass = 5.279;
  double masses[3]{3.0969, 0.493677, 0.13957};
  hydra::PhaseSpace<3> phsp{P_mass, masses};
  phsp.SetSeed(159753);
  const hydra::Vector4R B0(P_mass, 0.0, 0.0, 0.0);
  hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> Events(P_mass, masses, n);

  using clock = std::chrono::steady_clock;
  auto step = [&]() {
    phsp.Generate(B0, Events);
    return Events.Unweight().size();  // returns after the accepted count is on the host
  };
  const auto f0 = clock::now();
  size_t nacc = step();
  const double first_ms = std::chrono::duration<double, std::milli>(clock::now() - f0).count();
  for (int i = 1; i < warmup; ++i) nacc = step();
  hydra::detail::hpsp_sync("bench_decays");
  std::vector<double> ms;
  for (int i = 0; i < steps; ++i) {
    const auto t0 = clock::now();
    nacc = step();
    hydra::detail::hpsp_sync("bench_decays");
    ms.push_back(std::chrono::duration<double, std::milli>(clock::now() - t0).count());
  }
  double sum = 0.0, best = ms.empty() ? 0.0 : ms[0];
  for (double v : ms) {
    sum += v;
    best = v < best ? v : best;
  }
  std::printf("{\"events\": %zu, \"steps\": %d, \"warmup\": %d, \"n_acc\": %zu, "
              "\"ms_per_step\": %.6f, \"ms_min\": %.6f, \"first_step_ms\": %.6f}\n",
              n, steps, warmup, nacc, ms.empty() ? 0.0 : sum / ms.size(), best, first_ms);
  return 0;
}
