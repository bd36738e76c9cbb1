// hydra::PhaseSpace<N, GRND> -- the phase-space generator
// (reference: hydra/PhaseSpace.h:86-345, hydra/detail/PhaseSpace.inl:42-455).
//
// Same constructor checks, seed handling and member functions as the
// reference; Generate, AverageOn and the container's Unweight run the gfx950
// kernels of libhpsp.so through the C ABI (include/hpsp.h).  Every call is
// synchronous to the caller, as in the reference.  AverageOn and Evaluate take
// the described functors of PhaseSpaceFunctors.h (run by the library's
// kernels) or, in a translation unit built with hipcc, any callable of the
// event's four-vectors (compiled into the library's kernel templates there,
// hydra/detail/UserKernels.h).
#pragma once

#include <array>
#include <cmath>
#include <functional>
#include <initializer_list>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include <hydra/Decays.h>
#include <hydra/EventFunctors.h>
#include <hydra/PhaseSpaceFunctors.h>
#include <hydra/PhaseSpaceRange.h>
#include <hydra/Types.h>
#include <hydra/Vector4R.h>
#include <hydra/detail/Runtime.h>
#include <hydra/detail/UserFunctors.h>
#include <hydra/detail/random/engines.h>
#include <hydra/device/System.h>
#include <hydra/device/vector.h>
#include <hydra/multivector.h>

#include "hpsp.h"

namespace hydra {

namespace detail {

// boost-style hash_combine over std::hash<double> (hydra/detail/Hash.h:45-60).
template <class It>
inline std::size_t hash_range(It first, It last) {
  std::size_t seed = 0;
  for (; first != last; ++first) {
    std::hash<typename std::decay<decltype(*first)>::type> hasher;
    seed ^= hasher(*first) + 0x9e3779b9 + (seed << 6) + (seed >> 2);
  }
  return seed;
}

}  // namespace detail

template <size_t N, typename GRND = hydra::default_random_engine>
class PhaseSpace {
  static_assert(N >= 2 && N <= HPSP_MAX_N, "PhaseSpace<N>: N must be in 2..HPSP_MAX_N");

 public:
  PhaseSpace(double motherMass, const double (&daughtersMasses)[N])
      : fSeed(1), fMotherMass(motherMass), fMaxWeight(0.), fECM(0.) {
    for (size_t i = 0; i < N; ++i) fMasses[i] = daughtersMasses[i];
    init();
  }
  PhaseSpace(double motherMass, std::array<GReal_t, N> const& daughtersMasses)
      : fSeed(1), fMotherMass(motherMass), fMaxWeight(0.), fECM(0.) {
    for (size_t i = 0; i < N; ++i) fMasses[i] = daughtersMasses[i];
    init();
  }
  PhaseSpace(double motherMass, std::initializer_list<GReal_t> const& daughtersMasses)
      : fSeed(1), fMotherMass(motherMass), fMaxWeight(0.), fECM(0.) {
    if (daughtersMasses.size() < N)
      throw std::invalid_argument("[hydra::PhaseSpace]: too few daughter masses");
    for (size_t i = 0; i < N; ++i) fMasses[i] = *(daughtersMasses.begin() + i);
    init();
  }

  // The copy goes through GetSeed(), which returns an int (PhaseSpace.inl:194-216).
  PhaseSpace(PhaseSpace const& other)
      : fSeed((size_t)other.GetSeed()), fMotherMass(other.GetMotherMass()),
        fMaxWeight(other.GetMaxWeight()), fECM(other.GetECM()) {
    for (size_t i = 0; i < N; ++i) fMasses[i] = other.GetMasses()[i];
  }
  template <typename GRND2>
  PhaseSpace(PhaseSpace<N, GRND2> const& other)
      : fSeed((size_t)other.GetSeed()), fMotherMass(other.GetMotherMass()),
        fMaxWeight(other.GetMaxWeight()), fECM(other.GetECM()) {
    for (size_t i = 0; i < N; ++i) fMasses[i] = other.GetMasses()[i];
  }
  PhaseSpace& operator=(PhaseSpace const& other) {
    if (this == &other) return *this;
    fSeed = (size_t)other.GetSeed();
    fMotherMass = other.GetMotherMass();
    fMaxWeight = other.GetMaxWeight();
    fECM = other.GetECM();
    for (size_t i = 0; i < N; ++i) fMasses[i] = other.GetMasses()[i];
    return *this;
  }

  // ---- AverageOn (PhaseSpace.inl:255-287) -------------------------------
  template <typename FUNCTOR>
  std::pair<GReal_t, GReal_t> AverageOn(hydra::device::sys_t const&, Vector4R const& mother,
                                        FUNCTOR const& functor, size_t n) {
    hpsp_stats st = AverageOnStats(mother, functor, 0, n, HPSP_FP64);
    return std::make_pair(st.mean, ::sqrt(st.m2) / st.w);
  }

  // Over a container of mothers; the functor sees the mother as particle 0.
  template <typename FUNCTOR>
  std::pair<GReal_t, GReal_t> AverageOn(hydra::device::vector<Vector4R> const& mothers,
                                        FUNCTOR const& functor) {
    const hpsp_decay d = descriptor(Vector4R(fMotherMass, 0, 0, 0));
    const size_t n = mothers.size();
    hpsp_stats st;
    if constexpr (!is_phase_space_functor<FUNCTOR>::value) {
      st = detail::user_average_on<N, GRND::engine_id, true>(d, fSeed, functor, 0, n,
                                                             mothers.data(), mothers.ld(),
                                                             HPSP_FP64);
    } else {
      const hpsp_functor f = detail::to_hpsp(functor);
      const size_t wsb = hpsp_average_on_workspace_size(n);
      detail::device_buffer<unsigned char> ws(wsb);
      detail::device_buffer<hpsp_stats> out(1);
      detail::hpsp_check(hpsp_average_on_mothers(&d, &f, fSeed, mothers.data(), mothers.ld(), n,
                                                 HPSP_FP64, ws.data(), wsb, out.data(), nullptr),
                         "PhaseSpace::AverageOn");
      detail::hpsp_check(hpsp_memcpy_to_host(&st, out.data(), sizeof(st)),
                         "PhaseSpace::AverageOn");
    }
    return std::make_pair(st.mean, ::sqrt(st.m2) / st.w);
  }

  // Partial statistics of events first..first+n-1 at a chosen precision; the
  // partials of consecutive ranges merge with hpsp_stats_merge into the result
  // of one call over their union.
  template <typename FUNCTOR>
  hpsp_stats AverageOnStats(Vector4R const& mother, FUNCTOR const& functor, size_t first,
                            size_t n, int precision = HPSP_FP64) {
    const hpsp_decay d = descriptor(mother);
    if constexpr (!is_phase_space_functor<FUNCTOR>::value) {
      return detail::user_average_on<N, GRND::engine_id, false>(d, fSeed, functor, first, n,
                                                                nullptr, 0, precision);
    } else {
      const hpsp_functor f = detail::to_hpsp(functor);
      const size_t wsb = hpsp_average_on_workspace_size(n);
      detail::device_buffer<unsigned char> ws(wsb);
      detail::device_buffer<hpsp_stats> out(1);
      detail::hpsp_check(hpsp_average_on(&d, &f, fSeed, first, n, precision, ws.data(), wsb,
                                         out.data(), nullptr),
                         "PhaseSpace::AverageOn");
      hpsp_stats st;
      detail::hpsp_check(hpsp_memcpy_to_host(&st, out.data(), sizeof(st)),
                         "PhaseSpace::AverageOn");
      return st;
    }
  }

  // ---- Generate (PhaseSpace.inl:351-444) --------------------------------
  template <typename... P>
  Range<typename Decays<hydra::tuple<P...>, hydra::device::sys_t>::iterator> Generate(
      Vector4R const& mother, Decays<hydra::tuple<P...>, hydra::device::sys_t>& events) {
    static_assert(sizeof...(P) == N, "container arity differs from PhaseSpace<N>");
    const hpsp_decay d = descriptor(mother);
    // keep the PhaseSpaceWeight column when the container's weight functor is
    // this generator's (same masses and fMaxWeight, bit for bit)
    const bool same = same_decay(events.GetMotherMass(), events.GetMasses()) &&
                      events.GetMaxWeight() == fMaxWeight;
    double* cols = events.columns_for_overwrite();
    if (same && events.size() > 0 && !events.read_after_generate()) {
      // nothing formed yet: the container keeps (decay, seed); the momenta
      // are formed when first read, or Unweight() places every event once
      // (hpsp_generate_unweight), its flags from the weight chain
      events.set_weights_cache_valid(false);
      events.set_generated_lazy(d, (uint64_t)fSeed);
    } else if (same && events.size() > 0) {
      // a container that is read after Generate: momenta and weights in one
      // pass (deferring would form the momenta twice); Unweight() still
      // regenerates them at their places instead of moving them
      detail::hpsp_check(hpsp_generate(&d, fSeed, 0, events.size(), HPSP_FP64, cols, events.ld(),
                                       0, nullptr, events.weights_cache_for_write(), nullptr),
                         "PhaseSpace::Generate");
      detail::hpsp_sync("PhaseSpace::Generate");
      events.set_weights_cache_valid(true);
      events.set_generated(d, (uint64_t)fSeed);
    } else {
      detail::hpsp_check(hpsp_generate(&d, fSeed, 0, events.size(), HPSP_FP64, cols, events.ld(),
                                       0, nullptr, nullptr, nullptr),
                         "PhaseSpace::Generate");
      detail::hpsp_sync("PhaseSpace::Generate");
      events.set_weights_cache_valid(false);
    }
    return make_range(events.begin(), events.end());
  }

  // Generate(mothers, daughters): event k boosted by mother k (DecayMothers).
  template <typename... P>
  Range<typename Decays<hydra::tuple<P...>, hydra::device::sys_t>::iterator> Generate(
      hydra::device::vector<Vector4R> const& mothers,
      Decays<hydra::tuple<P...>, hydra::device::sys_t>& daughters) {
    static_assert(sizeof...(P) == N, "container arity differs from PhaseSpace<N>");
    const hpsp_decay d = descriptor(Vector4R(fMotherMass, 0, 0, 0));
    generate_over(d, mothers.size(), mothers.data(), mothers.ld(), daughters);
    return make_range(daughters.begin(), daughters.end());
  }

  // Generate(mothers, daughters) with the mothers one particle of another
  // container (Generate(ThreeBodyDecay.GetDaugtherRange(_0), TwoBodyDecay),
  // phsp_chain.inl:234-238): its columns are read in place.
  template <typename MotherRange, typename... P,
            typename = decltype(std::declval<MotherRange const&>().data(0))>
  Range<typename Decays<hydra::tuple<P...>, hydra::device::sys_t>::iterator> Generate(
      MotherRange const& mothers, Decays<hydra::tuple<P...>, hydra::device::sys_t>& daughters) {
    static_assert(sizeof...(P) == N, "container arity differs from PhaseSpace<N>");
    const hpsp_decay d = descriptor(Vector4R(fMotherMass, 0, 0, 0));
    generate_over(d, mothers.size(), mothers.data(0), mothers.ld(), daughters);
    return make_range(daughters.begin(), daughters.end());
  }

  // ---- Evaluate (PhaseSpace.inl:289-348) --------------------------------
  // result rows: (weight, functors(daughters)...), one row per event.
  template <typename... T, typename... FUNCTORS>
  Range<typename multivector<hydra::tuple<T...>, hydra::device::sys_t>::iterator> Evaluate(
      Vector4R const& mother, multivector<hydra::tuple<T...>, hydra::device::sys_t>& result,
      FUNCTORS const&... functors) {
    static_assert(sizeof...(T) == sizeof...(FUNCTORS) + 1,
                  "result needs one column for the weight and one per functor");
    const hpsp_decay d = descriptor(mother);
    if constexpr (check_functors<FUNCTORS...>()) {
      const hpsp_functor fs[] = {detail::to_hpsp(functors)...};
      detail::hpsp_check(hpsp_evaluate(&d, fs, (int)sizeof...(FUNCTORS), fSeed, 0, result.size(),
                                       result.data(), result.ld(), nullptr),
                         "PhaseSpace::Evaluate");
      detail::hpsp_sync("PhaseSpace::Evaluate");
    } else {
      detail::user_evaluate<N, GRND::engine_id, false>(d, fSeed, nullptr, 0, 0, result.size(),
                                                       result.data(), result.ld(), functors...);
    }
    return make_range(result.begin(), result.end());
  }

  // Over a container of mothers (EvalMothers): each event's engine is seeded
  // with the pairing hash of its index and the seed; functors see the mother
  // as particle 0.
  template <typename... T, typename... FUNCTORS>
  Range<typename multivector<hydra::tuple<T...>, hydra::device::sys_t>::iterator> Evaluate(
      hydra::device::vector<Vector4R> const& mothers,
      multivector<hydra::tuple<T...>, hydra::device::sys_t>& result,
      FUNCTORS const&... functors) {
    static_assert(sizeof...(T) == sizeof...(FUNCTORS) + 1,
                  "result needs one column for the weight and one per functor");
    if (result.size() < mothers.size())
      throw std::invalid_argument("[hydra::PhaseSpace::Evaluate]: result shorter than mothers");
    const hpsp_decay d = descriptor(Vector4R(fMotherMass, 0, 0, 0));
    if constexpr (check_functors<FUNCTORS...>()) {
      const hpsp_functor fs[] = {detail::to_hpsp(functors)...};
      detail::hpsp_check(
          hpsp_evaluate_mothers(&d, fs, (int)sizeof...(FUNCTORS), fSeed, mothers.data(),
                                mothers.ld(), 0, mothers.size(), result.data(), result.ld(),
                                nullptr),
          "PhaseSpace::Evaluate");
      detail::hpsp_sync("PhaseSpace::Evaluate");
    } else {
      detail::user_evaluate<N, GRND::engine_id, true>(d, fSeed, mothers.data(), mothers.ld(), 0,
                                                      mothers.size(), result.data(), result.ld(),
                                                      functors...);
    }
    return make_range(result.begin(), result.begin() + (std::ptrdiff_t)mothers.size());
  }

  // ---- seeds and parameters (PhaseSpace.h:281-326) -----------------------
  // GetSeed/SetSeed go through int (GInt_t), truncating 64-bit seeds.
  GInt_t GetSeed() const { return (GInt_t)fSeed; }
  void SetSeed(GInt_t seed) { fSeed = (size_t)seed; }
  size_t GetFullSeed() const { return fSeed; }

  const GReal_t* GetMasses() const { return fMasses; }
  double GetMaxWeight() const { return fMaxWeight; }
  void SetMaxWeight(double w) { fMaxWeight = w; }
  double GetMotherMass() const { return fMotherMass; }
  void SetMotherMass(double m) { fMotherMass = m; }
  double GetECM() const { return fECM; }
  void SetECM(double e) { fECM = e; }

  static double PDK(double a, double b, double c) {
    double x = (a - b - c) * (a + b + c) * (a - b + c) * (a + b - c);
    x = ::sqrt(x) / (2 * a);
    return x;
  }

  // C ABI description of this generator for a mother four-vector.
  hpsp_decay descriptor(Vector4R const& mother) const {
    hpsp_decay d{};
    d.n = (int32_t)N;
    d.engine = GRND::engine_id;
    d.mother_mass = fMotherMass;
    for (size_t i = 0; i < N; ++i) d.masses[i] = fMasses[i];
    d.ecm = fECM;
    d.max_weight = fMaxWeight;
    const double m4[4] = {mother.get(0), mother.get(1), mother.get(2), mother.get(3)};
    detail::hpsp_check(hpsp_decay_set_mother(&d, m4), "PhaseSpace");
    return d;
  }

 private:
  void init() {
    if (fMotherMass <= 0.0)
      throw std::invalid_argument(
          "[hydra::PhaseSpace]: Mass of mother particle is negative or zero. (fMotherMass <=0)");
    for (size_t i = 0; i < N; ++i)
      if (fMasses[i] <= 0.0)
        throw std::invalid_argument(
            "[hydra::PhaseSpace]: Mass of daughter particle is negative or zero. (fMass[] <=0)");
    fSeed *= (1 + detail::hash_range(&fMasses[0], &fMasses[0] + N));
    hpsp_decay d;
    const int rc = hpsp_decay_init(&d, (int)N, fMotherMass, fMasses, GRND::engine_id);
    if (rc == HPSP_ERR_ENERGY)
      throw std::runtime_error(
          "[hydra::PhaseSpace]: Not enough energy to generate this decay. ( fECM <=0.");
    detail::hpsp_check(rc, "PhaseSpace");
    fECM = d.ecm;
    fMaxWeight = d.max_weight;
  }

  // True when every functor is a described one (the library's kernels),
  // false when none is (user functors, compiled here); one call takes one kind.
  template <typename... F>
  static constexpr bool check_functors() {
    static_assert(sizeof...(F) >= 1 && sizeof...(F) <= HPSP_MAX_EVAL_FUNCTORS,
                  "Evaluate takes 1..HPSP_MAX_EVAL_FUNCTORS functors");
    constexpr bool all = std::conjunction<is_phase_space_functor<F>...>::value;
    constexpr bool any = std::disjunction<is_phase_space_functor<F>...>::value;
    static_assert(all || !any,
                  "Evaluate: described functors and user functors cannot be mixed in one call");
    return all;
  }

  // DecayMothers over rows 0..n-1 of daughters (launch_decayers.inl:153-182):
  // rows from n on keep their events, as the reference's transform leaves
  // them, and the container is grown to n rows if shorter.  The
  // PhaseSpaceWeight column is kept when the container's decay is this one
  // and every row's weight is then known.
  template <typename Container>
  void generate_over(hpsp_decay const& d, size_t n, const double* mothers, size_t mld,
                     Container& daughters) {
    const bool same = same_decay(daughters.GetMotherMass(), daughters.GetMasses()) &&
                      daughters.GetMaxWeight() == fMaxWeight;
    const bool partial = daughters.size() > n;
    const bool rest_known = partial && daughters.weights_cache() != nullptr;
    // rows n.. stay: form their deferred momenta before the columns are taken
    if (partial) daughters.materialize_rows();
    double* cols = daughters.columns_for_overwrite();
    if (daughters.size() < n) {
      daughters.resize(n);
      cols = daughters.columns_for_overwrite();
    }
    const bool keep_w = same && n > 0 && (!partial || rest_known);
    detail::hpsp_check(hpsp_generate_mothers(&d, fSeed, 0, n, mothers, mld, HPSP_FP64, cols,
                                             daughters.ld(), 0, nullptr,
                                             keep_w ? daughters.weights_cache_for_write() : nullptr,
                                             nullptr),
                       "PhaseSpace::Generate");
    detail::hpsp_sync("PhaseSpace::Generate");
    daughters.set_weights_cache_valid(keep_w);
  }

  bool same_decay(double M, std::array<double, N> const& m) const {
    if (M != fMotherMass) return false;
    for (size_t i = 0; i < N; ++i)
      if (m[i] != fMasses[i]) return false;
    return true;
  }

  size_t fSeed;
  GReal_t fMotherMass;
  GReal_t fMaxWeight;
  GReal_t fECM;
  GReal_t fMasses[N];
};

// PhaseSpaceIntegrator<N, Backend, GRND>::Integrate (PhaseSpaceIntegrator.h:57-160,
// PhaseSpaceIntegrator.inl:37-40): AverageOn with the mother at rest.
template <size_t N, typename Backend = hydra::device::sys_t,
          typename GRND = hydra::default_random_engine>
class PhaseSpaceIntegrator {
 public:
  PhaseSpaceIntegrator(const GReal_t motherMass, const GReal_t (&daughtersMasses)[N], size_t n)
      : fGenerator(motherMass, daughtersMasses), fMother(motherMass, 0, 0, 0), fNSamples(n) {}
  PhaseSpaceIntegrator(const GReal_t motherMass, std::array<GReal_t, N> const& daughtersMasses,
                       size_t n)
      : fGenerator(motherMass, daughtersMasses), fMother(motherMass, 0, 0, 0), fNSamples(n) {}
  PhaseSpaceIntegrator(const GReal_t motherMass,
                       std::initializer_list<GReal_t> const& daughtersMasses, size_t n)
      : fGenerator(motherMass, daughtersMasses), fMother(motherMass, 0, 0, 0), fNSamples(n) {}

  const PhaseSpace<N, GRND>& GetGenerator() const { return fGenerator; }
  PhaseSpace<N, GRND>& GetGenerator() { return fGenerator; }
  void SetGenerator(const PhaseSpace<N, GRND>& g) { fGenerator = g; }
  const Vector4R& GetMother() const { return fMother; }
  void SetMother(const Vector4R& m) { fMother = m; }
  size_t GetNSamples() const { return fNSamples; }
  void SetNSamples(size_t n) { fNSamples = n; }

  template <typename FUNCTOR>
  std::pair<GReal_t, GReal_t> Integrate(FUNCTOR const& functor) {
    return fGenerator.AverageOn(hydra::device::sys, fMother, functor, fNSamples);
  }

 private:
  PhaseSpace<N, GRND> fGenerator;
  Vector4R fMother;
  size_t fNSamples;
};

}  // namespace hydra
