// Drives the C++ drop-in API (hydra_amd/include/hydra) the way the
// reference's phase-space examples do (examples/phase_space/phsp_basic.inl,
// phsp_unweighting.inl, phsp_averaging_functor.inl) and prints what it sees as
// one JSON object; tests/test_cpp_api.py checks it against the golden vectors
// and the oracle.
//   test_api host   -- host-only parts (engines, constructor, seeds, errors)
//   test_api gpu    -- plus Generate / weights / Unweight / AverageOn on the GPU
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include <hydra/Decays.h>
#include <hydra/DenseHistogram.h>
#include <hydra/Function.h>
#include <hydra/PhaseSpace.h>
#include <hydra/Vector4R.h>
#include <hydra/device/System.h>
#include <hydra/multivector.h>

declarg(A, hydra::Vector4R)
declarg(B, hydra::Vector4R)
declarg(C, hydra::Vector4R)
// the columns of examples/phase_space/phsp_evaluating_functor.inl
declarg(Weight, double)
declarg(M13Sq, double)
declarg(M23Sq, double)
declarg(CosTheta, double)
// the decay chain of examples/phase_space/phsp_chain.inl
declarg(Jpsi, hydra::Vector4R)
declarg(Kaon, hydra::Vector4R)
declarg(Pion, hydra::Vector4R)
declarg(MuonP, hydra::Vector4R)
declarg(MuonM, hydra::Vector4R)
declarg(DeltaPhi, double)

using namespace hydra::arguments;

static void emit_array(const char* key, const std::vector<double>& v, bool comma = true) {
  std::printf("\"%s\": [", key);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
  std::printf("]%s\n", comma ? "," : "");
}

// device-to-device copy of n doubles, synchronous
static void detail_copy(double* dst, const double* src, size_t n) {
  hydra::detail::hpsp_check(hpsp_memcpy_device(dst, src, n * sizeof(double), nullptr), "copy");
  hydra::detail::hpsp_sync("copy");
}

static void emit_words(const char* key, const std::vector<unsigned long long>& v) {
  std::printf("\"%s\": [", key);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s\"%llu\"", i ? ", " : "", v[i]);
  std::printf("],\n");
}

template <typename E>
static std::vector<unsigned long long> words(unsigned long long seed, unsigned long long disc, int n) {
  E eng(seed);
  eng.discard(disc);
  std::vector<unsigned long long> w;
  for (int i = 0; i < n; ++i) w.push_back((unsigned long long)eng());
  return w;
}

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
  const double P_mass = 5.279, A_mass = 3.0969, B_mass = 0.493677, C_mass = 0.13957;
  double masses[3]{A_mass, B_mass, C_mass};

  std::printf("{\n");
  // ---- engines
  const unsigned long long seed = 159753ULL, big = (1ULL << 40) + 12345ULL;
  emit_words("rng_squares3_d0", words<hydra::random::squares3>(seed, 0, 64));
  emit_words("rng_squares3_dbig", words<hydra::random::squares3>(seed, big, 64));
  emit_words("rng_squares4_d0", words<hydra::random::squares4>(seed, 0, 64));
  emit_words("rng_philox_d0", words<hydra::random::philox>(seed, 0, 64));
  emit_words("rng_philox_long_dbig", words<hydra::random::philox_long>(seed, big, 64));
  emit_words("rng_threefry_d0", words<hydra::random::threefry>(seed, 0, 64));
  emit_words("rng_threefry_long_d0", words<hydra::random::threefry_long>(seed, 0, 64));
  emit_words("rng_default_d0", words<hydra::default_random_engine>(seed, 0, 64));

  // ---- constructor products, default seed and its int truncation
  hydra::PhaseSpace<3> phsp{P_mass, masses};
  std::printf("\"ecm\": %.17g,\n\"max_weight\": %.17g,\n", phsp.GetECM(), phsp.GetMaxWeight());
  std::printf("\"seed_full\": \"%llu\",\n\"seed_get\": %d,\n",
              (unsigned long long)phsp.GetFullSeed(), phsp.GetSeed());
  hydra::PhaseSpace<3> copy(phsp);
  std::printf("\"copy_seed_full\": \"%llu\",\n", (unsigned long long)copy.GetFullSeed());

  // ---- constructor errors (PhaseSpace.inl:49-76)
  int err_mass = 0, err_energy = 0, err_mother = 0;
  try {
    double bad[3]{A_mass, -1.0, C_mass};
    hydra::PhaseSpace<3> p(P_mass, bad);
  } catch (std::invalid_argument const&) {
    err_mass = 1;
  }
  try {
    hydra::PhaseSpace<3> p(3.5, masses);
  } catch (std::runtime_error const&) {
    err_energy = 1;
  }
  try {
    hydra::PhaseSpace<3> p(0.0, masses);
  } catch (std::invalid_argument const&) {
    err_mother = 1;
  }
  std::printf("\"err_mass\": %d,\n\"err_energy\": %d,\n\"err_mother\": %d,\n", err_mass,
              err_energy, err_mother);

  if (gpu) {
    hydra::Vector4R Parent(P_mass, 0.0, 0.0, 0.0);
    phsp.SetSeed(159753);
    const size_t n = 4096;
    hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> Events(P_mass, masses, n);
    phsp.Generate(Parent, Events);
    // the momenta are formed when first read (here, by the rows below)
    std::printf("\"lazy_after_generate\": %d,\n", Events.lazy() ? 1 : 0);
    std::vector<double> rows;
    for (size_t i = 0; i < 16; ++i) {
      auto ev = Events[i];
      const hydra::Vector4R a = std::get<0>(ev), b = std::get<1>(ev), c = std::get<2>(ev);
      for (const hydra::Vector4R* v : {&a, &b, &c})
        for (int q = 0; q < 4; ++q) rows.push_back(v->get(q));
    }
    emit_array("gen_rows16", rows);
    std::printf("\"lazy_after_read\": %d,\n", Events.lazy() ? 1 : 0);
    // a copy taken while lazy forms the same momenta when read, and a
    // container of other masses is filled eagerly with the same momenta
    hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> L1(P_mass, masses, n);
    phsp.Generate(Parent, L1);
    auto L2 = L1;
    double masses_b[3]{A_mass, B_mass, C_mass * 1.0000001};
    hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> Eager(P_mass, masses_b, n);
    phsp.Generate(Parent, Eager);
    hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> h_l2(L2), h_eager(Eager), h_ev(Events);
    int lazy_same = (L1.lazy() && !Eager.lazy()) ? 1 : 0;
    for (size_t c = 0; c < 12 && lazy_same; ++c)
      for (size_t i = 0; i < n; ++i)
        if (std::memcmp(h_l2.data() + c * h_l2.ld() + i, h_eager.data() + c * h_eager.ld() + i,
                        sizeof(double)) != 0 ||
            std::memcmp(h_l2.data() + c * h_l2.ld() + i, h_ev.data() + c * h_ev.ld() + i,
                        sizeof(double)) != 0) {
          lazy_same = 0;
          break;
        }
    std::printf("\"lazy_same\": %d,\n", lazy_same);
    auto weights = Events | Events.GetEventWeightFunctor();
    emit_array("gen_weights", weights.to_host());
    auto pi_range = Events.GetDaugtherRange(hydra::placeholders::_2);
    const hydra::Vector4R pi7 = pi_range[7];
    emit_array("pion_event7", {pi7.get(0), pi7.get(1), pi7.get(2), pi7.get(3)});

    // Unweight() on a copy, Unweight(functor) on another
    auto E2 = Events;
    auto accepted = E2.Unweight();
    std::printf("\"unweight_nacc\": %zu,\n", accepted.size());
    std::vector<double> acc_rows;
    for (size_t i = 0; i < 4; ++i) {
      auto ev = accepted[i];
      const hydra::Vector4R a = std::get<0>(ev);
      for (int q = 0; q < 4; ++q) acc_rows.push_back(a.get(q));
    }
    emit_array("unweight_first_A", acc_rows);
    // a container whose columns were handed out for writing forgets that it
    // holds Generate's events and moves them; the result is the same
    auto E5 = Events;
    const int gen_before = E5.generated() ? 1 : 0;
    (void)E5.data();
    const int gen_after = E5.generated() ? 1 : 0;
    auto acc5 = E5.Unweight();
    hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> h2(E2), h5(E5);
    int paths_same = (acc5.size() == accepted.size()) ? 1 : 0;
    for (size_t c = 0; c < 12 && paths_same; ++c)
      for (size_t i = 0; i < h2.size(); ++i)
        if (std::memcmp(h2.data() + c * h2.ld() + i, h5.data() + c * h5.ld() + i,
                        sizeof(double)) != 0) {
          paths_same = 0;
          break;
        }
    std::printf("\"unweight_paths_same\": %d,\n\"gen_flags\": [%d, %d],\n", paths_same, gen_before,
                gen_after);
    // momenta written through a column pointer: Unweight() flags the events
    // with the PhaseSpaceWeight of the new momenta (Decays.inl:428-431), as
    // hpsp_unweight without a weight column does, not with the weights the
    // container kept from Generate
    {
      auto E6 = Events;
      double* px_a = E6.column(0, 1);
      detail_copy(px_a, E6.column(1, 1), n);
      hpsp_decay d6;
      hydra::detail::hpsp_check(hpsp_decay_init(&d6, 3, P_mass, masses, HPSP_SQUARES3), "init");
      hydra::detail::device_buffer<double> out6(12 * n);
      const size_t wsb = hpsp_unweight_workspace_size(n);
      hydra::detail::device_buffer<unsigned char> ws6(wsb);
      hydra::detail::device_buffer<uint64_t> nacc6(1);
      hydra::detail::hpsp_check(
          hpsp_unweight(&d6, nullptr, -1.0, 0x180ec6d33cfd0abaULL, E6.data(), E6.ld(), n, nullptr,
                        out6.data(), n, nullptr, nullptr, ws6.data(), wsb, nacc6.data(), nullptr),
          "hpsp_unweight");
      std::vector<double> want(12 * n);
      uint64_t want_n = 0;
      hydra::detail::hpsp_check(hpsp_memcpy_to_host(want.data(), out6.data(), 12 * n * 8), "copy");
      hydra::detail::hpsp_check(hpsp_memcpy_to_host(&want_n, nacc6.data(), 8), "copy");
      const int cache_before = E6.weights_cache() ? 1 : 0;
      auto acc6 = E6.Unweight();
      hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> h6(E6);
      int same6 = (acc6.size() == want_n && want_n != accepted.size()) ? 1 : 0;
      for (size_t c = 0; c < 12 && same6; ++c)
        if (std::memcmp(h6.data() + c * h6.ld(), want.data() + c * n, n * sizeof(double)) != 0)
          same6 = 0;
      std::printf("\"unweight_after_write\": [%d, %d],\n", cache_before, same6);
    }
    // a size that is not a multiple of 64: the container's columns are
    // ld() = 1024 apart, and Generate and Unweight() give the bits of the C
    // ABI calls on columns exactly n apart
    {
      const size_t n9 = 1000;
      hpsp_decay d9;
      hydra::detail::hpsp_check(hpsp_decay_init(&d9, 3, P_mass, masses, HPSP_SQUARES3), "init");
      hydra::detail::device_buffer<double> direct(12 * n9), direct_uw(12 * n9);
      hydra::detail::device_buffer<uint64_t> nacc9(1);
      const size_t wsb = hpsp_generate_unweight_workspace_size(n9);
      hydra::detail::device_buffer<unsigned char> ws9(wsb);
      hydra::detail::hpsp_check(hpsp_generate(&d9, 159753, 0, n9, HPSP_FP64, direct.data(), n9,
                                              0, nullptr, nullptr, nullptr),
                                "hpsp_generate");
      hydra::detail::hpsp_check(
          hpsp_generate_unweight(&d9, 159753, 0, n9, 0x180ec6d33cfd0abaULL, direct_uw.data(), n9,
                                 nullptr, nullptr, ws9.data(), wsb, nacc9.data(), nullptr),
          "hpsp_generate_unweight");
      std::vector<double> want(12 * n9), want_uw(12 * n9);
      uint64_t want_n = 0;
      hydra::detail::hpsp_check(hpsp_memcpy_to_host(want.data(), direct.data(), 12 * n9 * 8), "copy");
      hydra::detail::hpsp_check(hpsp_memcpy_to_host(want_uw.data(), direct_uw.data(), 12 * n9 * 8),
                                "copy");
      hydra::detail::hpsp_check(hpsp_memcpy_to_host(&want_n, nacc9.data(), 8), "copy");
      hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> E9(P_mass, masses, n9);
      const int stride_ok = (E9.ld() % 64 == 0 && E9.ld() >= n9 && E9.capacity() == E9.ld()) ? 1 : 0;
      phsp.Generate(Parent, E9);
      hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> h9(E9);
      int same_gen = 1;
      for (size_t c = 0; c < 12; ++c)
        if (std::memcmp(h9.data() + c * h9.ld(), want.data() + c * n9, n9 * sizeof(double)) != 0)
          same_gen = 0;
      hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> U9(P_mass, masses, n9);
      phsp.Generate(Parent, U9);
      auto acc9 = U9.Unweight();
      hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> hu9(U9);
      int same_uw = (acc9.size() == want_n && U9.ld() % 64 == 0) ? 1 : 0;
      for (size_t c = 0; c < 12 && same_uw; ++c)
        if (std::memcmp(hu9.data() + c * hu9.ld(), want_uw.data() + c * n9, n9 * sizeof(double)) !=
            0)
          same_uw = 0;
      std::printf("\"padded_stride\": [%d, %d, %d],\n", stride_ok, same_gen, same_uw);
    }
    // the opt-in deferred-path check: a write through a column pointer taken
    // before Generate is caught at Unweight(); an untouched container passes
    {
      hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> G(P_mass, masses, n);
      double* stale = G.column(0, 1);
      G.SetDeferredChecks(true);
      phsp.Generate(Parent, G);
      detail_copy(stale + 5, stale + 7, 1);
      int caught = 0;
      try {
        G.Unweight();
      } catch (std::logic_error const&) {
        caught = 1;
      }
      hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> H(P_mass, masses, n);
      H.SetDeferredChecks(true);
      phsp.Generate(Parent, H);
      (void)H[3];  // forms the deferred momenta: the checksum follows
      int passed = 1;
      size_t nacc_h = 0;
      try {
        nacc_h = H.Unweight().size();
      } catch (std::logic_error const&) {
        passed = 0;
      }
      std::printf("\"deferred_checks\": [%d, %d, %d],\n", caught, passed,
                  nacc_h == accepted.size() ? 1 : 0);
    }
    // Generate + Unweight() three times on one container, as a loop does (each
    // Unweight writes into the buffers the previous one left as spares, and
    // the weight column is formed again after each): every cycle gives the
    // rows of a fresh container's first cycle, bit for bit
    {
      hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> L(P_mass, masses, n);
      hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> h_ref(E2);
      int cycles_same = 1, weights_same = 1;
      auto w_ref = (E2 | E2.GetEventWeightFunctor()).to_host();
      for (int cyc = 0; cyc < 3; ++cyc) {
        phsp.Generate(Parent, L);
        const size_t na = L.Unweight().size();
        hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> h(L);
        if (na != accepted.size()) cycles_same = 0;
        for (size_t c = 0; c < 12 && cycles_same; ++c)
          if (std::memcmp(h.data() + c * h.ld(), h_ref.data() + c * h_ref.ld(),
                          n * sizeof(double)) != 0)
            cycles_same = 0;
        auto w = (L | L.GetEventWeightFunctor()).to_host();
        if (w.size() != w_ref.size() || std::memcmp(w.data(), w_ref.data(),
                                                    w.size() * sizeof(double)) != 0)
          weights_same = 0;
      }
      std::printf("\"unweight_cycles\": [%d, %d],\n", cycles_same, weights_same);
    }
    // Events was read after its deferred Generate: the next Generate into it
    // writes momenta and weights in one pass, with the same bits, and its
    // Unweight() still regenerates (the result of E2's above)
    const int read_after = Events.read_after_generate() ? 1 : 0;
    phsp.Generate(Parent, Events);
    const int lazy_again = Events.lazy() ? 1 : 0;
    const int gen_again = Events.generated() ? 1 : 0;
    hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> h_again(Events);
    int eager_same = 1;
    for (size_t c = 0; c < 12 && eager_same; ++c)
      for (size_t i = 0; i < n; ++i)
        if (std::memcmp(h_again.data() + c * h_again.ld() + i, h_ev.data() + c * h_ev.ld() + i,
                        sizeof(double)) != 0) {
          eager_same = 0;
          break;
        }
    auto E7 = Events;
    auto acc7 = E7.Unweight();
    hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> h7(E7);
    int unweight7_same = (acc7.size() == accepted.size()) ? 1 : 0;
    for (size_t c = 0; c < 12 && unweight7_same; ++c)
      for (size_t i = 0; i < h2.size(); ++i)
        if (std::memcmp(h2.data() + c * h2.ld() + i, h7.data() + c * h7.ld() + i,
                        sizeof(double)) != 0) {
          unweight7_same = 0;
          break;
        }
    std::printf("\"eager_after_read\": [%d, %d, %d, %d, %d],\n", read_after, lazy_again, gen_again,
                eager_same, unweight7_same);
    auto E3 = Events;
    auto acc_f = E3.Unweight(hydra::InvariantMass2{}, -1.0, 0x39abdc4529b1661cULL);
    std::printf("\"unweight_f_nacc\": %zu,\n", acc_f.size());

    // AverageOn with the three-resonance Dalitz functor, and its integrator
    hydra::BreitWignerSum bw{{0.892, 0.05, {0.4967141530, -0.1382643012}, 1, 2},
                             {1.43, 0.11, {0.6476885381, 1.5230298564}, 1, 2},
                             {0.775, 0.149, {-0.2341533747, -0.2341368981}, 0, 2}};
    auto avg = phsp.AverageOn(hydra::device::sys, Parent, bw, 100000);
    emit_array("avg_bw3", {avg.first, avg.second});
    auto avg1 = phsp.AverageOn(hydra::device::sys, Parent, hydra::ConstantFunctor{}, 10000);
    emit_array("avg_one", {avg1.first, avg1.second});
    hydra::PhaseSpaceIntegrator<3, hydra::device::sys_t> integrator(P_mass, masses, 100000);
    integrator.GetGenerator().SetSeed(159753);
    auto integ = integrator.Integrate(hydra::InvariantMass2{});
    emit_array("integrate_m2", {integ.first, integ.second});

    // Generate(mothers, daughters)
    std::vector<hydra::Vector4R> moms;
    for (int k = 0; k < 256; ++k) {
      const double px = 0.01 * k, py = -0.02 * k, pz = 0.005 * k;
      moms.emplace_back(std::sqrt(P_mass * P_mass + px * px + py * py + pz * pz), px, py, pz);
    }
    hydra::device::vector<hydra::Vector4R> mothers(moms);
    hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> Daughters(P_mass, masses, 256);
    phsp.Generate(mothers, Daughters);
    std::vector<double> mrows;
    for (size_t i : {0u, 100u, 255u}) {
      auto ev = Daughters[i];
      const hydra::Vector4R c = std::get<2>(ev);
      for (int q = 0; q < 4; ++q) mrows.push_back(c.get(q));
    }
    emit_array("mothers_rows_C", mrows);

    // Evaluate(mother, dataset, M13^2, M23^2, cos(theta_K*)) as the example does
    hydra::multivector<hydra::tuple<Weight, M13Sq, M23Sq, CosTheta>, hydra::device::sys_t>
        dataset(4096);
    phsp.Evaluate(Parent, dataset, hydra::InvariantMass2{0, 2}, hydra::InvariantMass2{1, 2},
                  hydra::CosHelicity{1, 2});
    hydra::multivector<hydra::tuple<Weight, M13Sq, M23Sq, CosTheta>, hydra::host::sys_t> data(
        dataset);
    std::vector<double> erows;
    for (size_t i : {0u, 1u, 2u, 4095u}) {
      auto row = data[i];
      erows.push_back(std::get<Weight>(row));
      erows.push_back(std::get<M13Sq>(row));
      erows.push_back(std::get<M23Sq>(row));
      erows.push_back(std::get<CosTheta>(row));
    }
    emit_array("evaluate_rows", erows);
    // over the 256 mothers: functors see (mother, A, B, C)
    hydra::multivector<hydra::tuple<Weight, M13Sq>, hydra::device::sys_t> dm(256);
    phsp.Evaluate(mothers, dm, hydra::InvariantMass2{1, 3});
    std::vector<double> emrows;
    for (size_t i : {0u, 100u, 255u}) {
      auto row = dm[i];
      emrows.push_back(std::get<0>(row).Value());
      emrows.push_back(std::get<1>(row).Value());
    }
    emit_array("evaluate_mothers_rows", emrows);

    // hydra::unweight(device::sys, events, events.GetEventWeightFunctor(bw))
    // (phsp_unweighting_functor.inl:209) on a copy
    auto E4 = Events;
    auto uw_free = hydra::unweight(hydra::device::sys, E4, E4.GetEventWeightFunctor(bw));
    std::printf("\"free_unweight_nacc\": %zu,\n", uw_free.size());
    std::vector<double> fu_rows;
    for (size_t i : {0ul, 1ul, (size_t)uw_free.size(), (size_t)4095}) {
      auto ev = E4[i];
      const hydra::Vector4R c = std::get<2>(ev);
      for (int q = 0; q < 4; ++q) fu_rows.push_back(c.get(q));
    }
    emit_array("free_unweight_rows_C", fu_rows);

    // trans-backend copies: device -> host -> device
    hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> HostEvents(Events);
    hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> Back(HostEvents);
    int same = (int)(HostEvents.size() == Events.size() && Back.size() == Events.size());
    for (size_t i : {0ul, 17ul, 4095ul}) {
      auto h = HostEvents[i];
      auto d = Events[i];
      auto b = Back[i];
      for (int q = 0; q < 4; ++q) {
        const hydra::Vector4R hc = std::get<2>(h), dc = std::get<2>(d), bc = std::get<2>(b);
        same &= (int)(hc.get(q) == dc.get(q) && bc.get(q) == dc.get(q));
      }
    }
    std::printf("\"host_copy_same\": %d,\n", same);
    auto wback = Back | Back.GetEventWeightFunctor();
    emit_array("host_copy_weights3", {wback[0], wback[17], wback[4095]});

    // Dalitz plot of the 4096 events (phsp_basic.inl:196-214): the fused fill
    // with the container's weight column, and with PhaseSpaceWeight computed
    // inside the fill, and from Evaluate's value columns
    const std::array<double, 2> lo{(A_mass + B_mass) * (A_mass + B_mass),
                                   (B_mass + C_mass) * (B_mass + C_mass)};
    const std::array<double, 2> hi{(P_mass - C_mass) * (P_mass - C_mass),
                                   (P_mass - A_mass) * (P_mass - A_mass)};
    auto dalitz_variables =
        Events | hydra::make_axes(hydra::InvariantMass2{0, 1}, hydra::InvariantMass2{1, 2});
    auto dalitz_weights = Events | Events.GetEventWeightFunctor();
    auto Hist_Dalitz = hydra::make_dense_histogram<double, 2>(
        hydra::device::sys, {20, 20}, lo, hi, dalitz_variables, dalitz_weights);
    emit_array("hist_dalitz_w", Hist_Dalitz.GetBinsContents());
    hydra::DenseHistogram<double, 2, hydra::device::sys_t> H2({20, 20}, lo, hi);
    H2.Fill(dalitz_variables, Events.GetEventWeightFunctor());
    emit_array("hist_dalitz_phsp", H2.GetBinsContents());
    H2.Fill(dalitz_variables);
    emit_array("hist_dalitz_n", H2.GetBinsContents());
    std::printf("\"hist_bin_3_7\": %.17g,\n", H2.GetBinContent(std::array<size_t, 2>{3, 7}));

    // phase_space_range as examples/misc/range_semantics.inl:154-167 uses it:
    // K+ -> 3 pi at rest, seed 321, the m(pi pi) histogram of the lazy range
    std::array<double, 3> pions{0.13957061, 0.13957061, 0.13957061};
    auto range = hydra::phase_space_range(hydra::Vector4R(0.493677, 0.0, 0.0, 0.0), pions, 321,
                                          100000);
    hydra::DenseHistogram<double, 1, hydra::device::sys_t> Hist_Mass(
        10, pions[0] + pions[1], 0.493677 - pions[0]);
    Hist_Mass.Fill(range | hydra::make_axes(hydra::InvariantMass{0, 1}));
    emit_array("range_hist_n", Hist_Mass.GetBinsContents());
    Hist_Mass.Fill(range | hydra::make_axes(hydra::InvariantMass{0, 1}),
                   range.GetEventWeightFunctor());
    emit_array("range_hist_w", Hist_Mass.GetBinsContents());
    // single elements and a copy of the first 512 into a Decays container
    std::vector<double> el;
    for (size_t i = 0; i < 4; ++i) {
      auto e = range[i];
      el.push_back(hydra::get<0>(e));
      const hydra::Vector4R v[3] = {hydra::get<1>(e), hydra::get<2>(e), hydra::get<3>(e)};
      for (int j = 0; j < 3; ++j)
        for (int c = 0; c < 4; ++c) el.push_back(v[j].get(c));
    }
    emit_array("range_elements4", el);
    hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> Kpi(0.493677, pions, 512);
    hydra::EventWeights kw(512);
    hydra::copy(range, Kpi, &kw);
    std::vector<double> krows;
    for (size_t i = 0; i < 512; ++i) {
      auto row = Kpi[i];
      const hydra::Vector4R v[3] = {std::get<0>(row), std::get<1>(row), std::get<2>(row)};
      for (int j = 0; j < 3; ++j)
        for (int c = 0; c < 4; ++c) krows.push_back(v[j].get(c));
    }
    emit_array("range_copy_rows", krows);
    emit_array("range_copy_w", kw.to_host());
    int err_unbounded = 0;
    try {
      auto inf = hydra::phase_space_range(hydra::Vector4R(0.493677, 0.0, 0.0, 0.0), pions, 321);
      Hist_Mass.Fill(inf | hydra::make_axes(hydra::InvariantMass{0, 1}));
    } catch (std::invalid_argument const&) {
      err_unbounded = 1;
    }
    std::printf("\"range_err_unbounded\": %d,\n", err_unbounded);

    // the decay chain of examples/phase_space/phsp_chain.inl:216-262
    {
      const double B0_mass = 5.27955, Jpsi_mass = 3.0969, K_mass = 0.493677,
                   pi_mass = 0.13957061, mup_mass = 0.1056583745, mum_mass = 0.1056583745;
      hydra::Vector4R B0(B0_mass, 0.0, 0.0, 0.0);
      hydra::PhaseSpace<3> ThreeBodyPHSP(B0_mass, {Jpsi_mass, K_mass, pi_mass});
      hydra::PhaseSpace<2> TwoBodyPHSP(Jpsi_mass, {mup_mass, mum_mass});
      const size_t nentries = 512;
      auto ThreeBodyDecay = hydra::Decays<hydra::tuple<Jpsi, Kaon, Pion>, hydra::device::sys_t>(
          B0_mass, {Jpsi_mass, K_mass, pi_mass}, nentries);
      auto TwoBodyDecay = hydra::Decays<hydra::tuple<MuonP, MuonM>, hydra::device::sys_t>(
          Jpsi_mass, {mup_mass, mup_mass}, nentries);
      ThreeBodyPHSP.Generate(B0, ThreeBodyDecay);
      TwoBodyPHSP.Generate(ThreeBodyDecay.GetDaugtherRange(hydra::placeholders::_0), TwoBodyDecay);
      auto melded = ThreeBodyDecay.Meld(TwoBodyDecay);
      auto Variables = hydra::make_functors(
          hydra::InvariantMass2{0, 2}, hydra::InvariantMass2{1, 2}, hydra::CosHelicity{1, 2, 0, 3},
          hydra::DecayPlaneAngle{1, 2, 3});
      auto chain = melded | Variables;
      auto weights = ThreeBodyDecay | ThreeBodyDecay.GetEventWeightFunctor();
      hydra::multivector<hydra::tuple<M13Sq, M23Sq, CosTheta, DeltaPhi>, hydra::device::sys_t>
          variables(chain.size());
      hydra::copy(chain, variables);
      std::vector<double> vars;
      for (size_t i = 0; i < nentries; ++i) {
        auto row = variables[i];
        vars.push_back(double(std::get<0>(row)));
        vars.push_back(double(std::get<1>(row)));
        vars.push_back(double(std::get<2>(row)));
        vars.push_back(double(std::get<3>(row)));
      }
      emit_array("chain_vars", vars);
      auto c7 = chain[7];
      emit_array("chain_row7", {std::get<0>(c7), std::get<1>(c7), std::get<2>(c7), std::get<3>(c7)});
      std::vector<double> mrow;
      for (size_t i : {size_t(0), size_t(300)}) {
        auto m = melded[i];
        const hydra::Vector4R v[5] = {std::get<0>(m), std::get<1>(m), std::get<2>(m),
                                      std::get<3>(m), std::get<4>(m)};
        for (int j = 0; j < 5; ++j)
          for (int c = 0; c < 4; ++c) mrow.push_back(v[j].get(c));
      }
      emit_array("chain_melded_0_300", mrow);
      emit_array("chain_weights", weights.to_host());
      // histograms of the angles, weighted (the example's TH1D fills)
      hydra::DenseHistogram<double, 2, hydra::device::sys_t> Angles({20, 16}, {-1.0, 0.0},
                                                                   {1.0, 6.283185307179586});
      Angles.Fill(melded | hydra::make_axes(hydra::CosHelicity{1, 2, 0, 3},
                                            hydra::DecayPlaneAngle{1, 2, 3}),
                  weights);
      emit_array("chain_hist_w", Angles.GetBinsContents());
    }
  }
  std::printf("\"ok\": 1\n}\n");
  return 0;
}
