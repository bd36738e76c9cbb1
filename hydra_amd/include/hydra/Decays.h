// hydra::Decays -- the N-particle event container of the phase-space path
// (reference: hydra/Decays.h:67-651, hydra/detail/Decays.inl:114-511).
//
// Device storage is structure-of-arrays by component: 4N FP64 columns of
// capacity() events each (column 4*i + c holds component c of particle i;
// capacity() is a multiple of 64, detail::column_stride), so every kernel
// access is coalesced and every column starts on a cache line.  Rows are assembled into tuples of
// Vector4R on access.  The container also keeps the PhaseSpaceWeight of each
// event when it is known exactly (written by PhaseSpace::Generate, permuted by
// Unweight), so Unweight need not re-read the momenta to flag events.
//
// Storage is allocated without zero-filling (the reference value-initialises,
// Decays.h:92; the difference is not observable after Generate).
#pragma once

#include <array>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <tuple>
#include <utility>
#include <vector>

#include <hydra/Function.h>
#include <hydra/PhaseSpaceFunctors.h>
#include <hydra/Range.h>
#include <hydra/Tuple.h>
#include <hydra/Types.h>
#include <hydra/Vector4R.h>
#include <hydra/detail/Runtime.h>
#include <hydra/detail/UserFunctors.h>
#include <hydra/device/System.h>
#include <hydra/host/System.h>
#include <hydra/multivector.h>

#include "hpsp.h"

namespace hydra {

namespace placeholders {
template <unsigned I>
struct placeholder {};
constexpr placeholder<0> _0{};
constexpr placeholder<1> _1{};
constexpr placeholder<2> _2{};
constexpr placeholder<3> _3{};
constexpr placeholder<4> _4{};
constexpr placeholder<5> _5{};
constexpr placeholder<6> _6{};
constexpr placeholder<7> _7{};
constexpr placeholder<8> _8{};
constexpr placeholder<9> _9{};
constexpr placeholder<10> _10{};
constexpr placeholder<11> _11{};
}  // namespace placeholders

namespace detail {

// PhaseSpace<N> constructor products (PhaseSpace.inl:67-89), also computed
// by Decays and PhaseSpaceWeight (Decays.h:100-118, Decays.inl:139-157).
template <size_t N>
inline hpsp_decay make_decay(double mother_mass, const std::array<double, N>& masses,
                             const char* where, int engine = HPSP_SQUARES3) {
  hpsp_decay d;
  hpsp_check(hpsp_decay_init(&d, (int)N, mother_mass, masses.data(), engine), where);
  return d;
}

}  // namespace detail

// Weights of a container's events, computed on the device
// (events | events.GetEventWeightFunctor(), Range2.inl:213-238).
class EventWeights {
 public:
  explicit EventWeights(size_t n) : w_(n) {}
  size_t size() const { return w_.size(); }
  double operator[](size_t i) const {
    double v = 0;
    detail::hpsp_check(hpsp_memcpy_to_host(&v, w_.data() + i, sizeof(double)), "EventWeights");
    return v;
  }
  std::vector<double> to_host() const {
    std::vector<double> h(w_.size());
    if (!h.empty())
      detail::hpsp_check(hpsp_memcpy_to_host(h.data(), w_.data(), h.size() * sizeof(double)),
                         "EventWeights");
    return h;
  }
  double* data() { return w_.data(); }
  const double* data() const { return w_.data(); }

 private:
  detail::device_buffer<double> w_;
};

// PhaseSpaceWeight<P...> (Decays.inl:114-254): w = fMaxWeight * prod pdk over
// masses of running sums, evaluated on the device.
template <typename... Particles>
class PhaseSpaceWeight {
 public:
  static constexpr size_t N = sizeof...(Particles);
  PhaseSpaceWeight(double motherMass, std::array<double, N> const& masses)
      : d_(detail::make_decay<N>(motherMass, masses, "PhaseSpaceWeight")) {}
  double GetMaxWeight() const { return d_.max_weight; }
  const double* GetMasses() const { return d_.masses; }
  const hpsp_decay& descriptor() const { return d_; }

 private:
  hpsp_decay d_;
};

// PhaseSpaceReweight<F, P...> (Decays.inl:258-411): PhaseSpaceWeight * f.
template <typename Functor, typename... Particles>
class PhaseSpaceReweight {
 public:
  static constexpr size_t N = sizeof...(Particles);
  PhaseSpaceReweight(Functor const& f, double motherMass, std::array<double, N> const& masses)
      : d_(detail::make_decay<N>(motherMass, masses, "PhaseSpaceReweight")),
        f_(detail::to_hpsp(f)) {}
  double GetMaxWeight() const { return d_.max_weight; }
  const hpsp_decay& descriptor() const { return d_; }
  const hpsp_functor& functor() const { return f_; }

 private:
  hpsp_decay d_;
  hpsp_functor f_;
};

// PhaseSpaceReweight<F, P...> with a user's callable F (compiled by hipcc into
// the caller's kernels, hydra/detail/UserKernels.h): `events | reweight`
// gives PhaseSpaceWeight * F of each event, F receiving the event's particles
// by their declarg types.
template <typename Functor, typename... Particles>
class UserReweight {
 public:
  static constexpr size_t N = sizeof...(Particles);
  UserReweight(Functor const& f, double motherMass, std::array<double, N> const& masses)
      : d_(detail::make_decay<N>(motherMass, masses, "PhaseSpaceReweight")), f_(f) {}
  double GetMaxWeight() const { return d_.max_weight; }
  const hpsp_decay& descriptor() const { return d_; }
  const Functor& functor() const { return f_; }

 private:
  hpsp_decay d_;
  Functor f_;
};

// Decays::Meld(others...) (Decays.h:286-293): the particles of one container
// followed by those of the others, row by row, as the decay chain of
// examples/phase_space/phsp_chain.inl:244 joins B0 -> J/psi K pi with
// J/psi -> mu+ mu-.  A view over the containers' device columns (they must
// outlive it); rows are read as tuples of P four-vectors.
template <size_t P>
class MeldedDecays {
  template <size_t... I>
  static auto row_type(std::index_sequence<I...>) -> hydra::tuple<decltype((void)I, Vector4R())...>;

 public:
  using value_type = decltype(row_type(std::make_index_sequence<P>{}));
  static constexpr size_t nparticles = P;

  MeldedDecays(hpsp_particles const& p, size_t n) : p_(p), n_(n) {}
  size_t size() const { return n_; }
  const hpsp_particles& particles() const { return p_; }

  value_type operator[](size_t k) const {
    std::array<double, 4 * P> v;
    for (size_t i = 0; i < P; ++i)
      detail::hpsp_check(hpsp_memcpy2d_to_host(v.data() + 4 * i, sizeof(double), p_.cols[i] + k,
                                               p_.ld[i] * sizeof(double), sizeof(double), 4),
                         "Decays::Meld");
    return make(v, std::make_index_sequence<P>{});
  }

 private:
  template <size_t... I>
  static value_type make(const std::array<double, 4 * P>& v, std::index_sequence<I...>) {
    return value_type(Vector4R(v[4 * I], v[4 * I + 1], v[4 * I + 2], v[4 * I + 3])...);
  }

  hpsp_particles p_;
  size_t n_;
};

template <typename Tuple, typename Backend>
class Decays;

template <typename... Particles>
class Decays<hydra::tuple<Particles...>, hydra::device::sys_t> {
 public:
  static constexpr size_t N = sizeof...(Particles);
  enum { nparticles = N };
  typedef hydra::tuple<Particles...> value_type;
  typedef hydra::tuple<Particles...> tuple_type;
  typedef hydra::device::sys_t system_type;

  // Random-access iterator over rows; dereferencing copies one row to the host.
  class iterator {
   public:
    typedef std::ptrdiff_t difference_type;
    iterator() = default;
    iterator(const Decays* c, size_t i) : c_(c), i_(i) {}
    value_type operator*() const { return (*c_)[i_]; }
    iterator& operator++() { ++i_; return *this; }
    iterator operator+(difference_type d) const { return iterator(c_, i_ + d); }
    const Decays* container() const { return c_; }
    iterator operator-(difference_type d) const { return iterator(c_, i_ - d); }
    difference_type operator-(const iterator& o) const { return (difference_type)i_ - (difference_type)o.i_; }
    bool operator==(const iterator& o) const { return i_ == o.i_ && c_ == o.c_; }
    bool operator!=(const iterator& o) const { return !(*this == o); }
    size_t index() const { return i_; }

   private:
    const Decays* c_ = nullptr;
    size_t i_ = 0;
  };
  typedef iterator const_iterator;

  // One particle's column view (GetDaugtherRange, Decays.h:279-284).
  class ParticleRange {
   public:
    ParticleRange(const Decays* c, unsigned i) : c_(c), i_(i) {}
    size_t size() const { return c_->size(); }
    Vector4R operator[](size_t k) const {
      double v[4];
      detail::hpsp_check(hpsp_memcpy2d_to_host(v, sizeof(double), c_->column(i_, 0) + k,
                                               c_->ld() * sizeof(double), sizeof(double), 4),
                         "Decays::GetDaugtherRange");
      return Vector4R(v[0], v[1], v[2], v[3]);
    }
    // device pointer of component c (0:E 1:px 2:py 3:pz); components ld() apart
    const double* data(int c) const { return c_->column(i_, c); }
    size_t ld() const { return c_->ld(); }

   private:
    const Decays* c_;
    unsigned i_;
  };

  Decays() = delete;

  Decays(double motherMass, std::array<double, N> const& daughtersMasses, size_t nentries = 0)
      : fMotherMass(motherMass), fMasses(daughtersMasses) {
    init(nentries);
  }

  Decays(double motherMass, const double (&daughtersMasses)[N], size_t nentries = 0)
      : fMotherMass(motherMass) {
    for (size_t i = 0; i < N; ++i) fMasses[i] = daughtersMasses[i];
    init(nentries);
  }

  Decays(const Decays& other) = default;
  Decays(Decays&& other) noexcept = default;
  Decays& operator=(const Decays& other) = default;
  Decays& operator=(Decays&& other) noexcept = default;

  // trans-backend copy from host storage (Decays.h:180-193, 606-618)
  explicit Decays(Decays<hydra::tuple<Particles...>, hydra::host::sys_t> const& other);
  Decays& operator=(Decays<hydra::tuple<Particles...>, hydra::host::sys_t> const& other);

  // ---- STL-like interface (Decays.h:298-394)
  size_t size() const { return fSize; }
  size_t capacity() const { return fCap; }
  bool empty() const { return fSize == 0; }
  void clear() {
    fSize = 0;
    fWValid = false;
    fGenValid = false;
    fLazy = false;
  }
  void reserve(size_t n) {
    if (n > fCap) regrow(n);
  }
  void resize(size_t n) {
    materialize();
    fGenValid = false;
    if (n > fCap) regrow(n);
    fSize = n;
    fWValid = false;
  }
  void shrink_to_fit() {
    if (fCap > detail::column_stride(fSize)) regrow(fSize);
  }

  value_type operator[](size_t k) const {
    std::array<double, 4 * N> v;
    materialize();
    if (fSize)
      detail::hpsp_check(hpsp_memcpy2d_to_host(v.data(), sizeof(double), fCols.data() + k,
                                               fCap * sizeof(double), sizeof(double), 4 * N),
                         "Decays::operator[]");
    return make_row(v, std::make_index_sequence<N>{});
  }
  value_type GetEntry(size_t k) const { return (*this)[k]; }

  iterator begin() const { return iterator(this, 0); }
  iterator end() const { return iterator(this, fSize); }
  iterator cbegin() const { return begin(); }
  iterator cend() const { return end(); }

  template <unsigned I>
  ParticleRange GetDaugtherRange(placeholders::placeholder<I>) const {
    static_assert(I < N, "particle index out of range");
    return ParticleRange(this, I);
  }

  // ---- the particles of this container, and Meld (Decays.h:286-293)
  void append_particles(hpsp_particles& p) const {
    for (size_t i = 0; i < N; ++i) {
      p.cols[p.n] = column(i, 0);
      p.ld[p.n] = fCap;
      ++p.n;
    }
  }
  hpsp_particles particles() const {
    hpsp_particles p{};
    append_particles(p);
    return p;
  }
  template <typename... Others>
  MeldedDecays<(N + ... + Others::N)> Meld(Others const&... others) const {
    static_assert((N + ... + Others::N) <= HPSP_MAX_PARTICLES, "too many particles to meld");
    hpsp_particles p{};
    append_particles(p);
    (others.append_particles(p), ...);
    if (((others.size() < fSize) || ...))
      throw std::invalid_argument("[hydra::Decays::Meld]: a melded container is shorter");
    return MeldedDecays<(N + ... + Others::N)>(p, fSize);
  }

  // ---- weights (Decays.h:235-250)
  PhaseSpaceWeight<Particles...> GetEventWeightFunctor() const {
    return PhaseSpaceWeight<Particles...>(fMotherMass, fMasses);
  }
  // a described functor (the library's kernels) or a user's callable
  // (compiled by hipcc in the caller's translation unit)
  template <typename Functor>
  auto GetEventWeightFunctor(Functor const& f) const {
    if constexpr (is_phase_space_functor<Functor>::value)
      return PhaseSpaceReweight<Functor, Particles...>(f, fMotherMass, fMasses);
    else
      return UserReweight<Functor, Particles...>(f, fMotherMass, fMasses);
  }

  // ---- unweighting (Decays.inl:415-511)
  Range<iterator> Unweight(size_t seed = 0x180ec6d33cfd0aba) {
    if (fGenValid) {
      if (fCheck) guard_verify();
      return fWValid ? unweight_regenerate(seed) : generate_unweight(seed);
    }
    const hpsp_decay d = detail::make_decay<N>(fMotherMass, fMasses, "Decays::Unweight");
    return unweight(d, nullptr, -1.0, seed);
  }

  template <typename Functor>
  Range<iterator> Unweight(Functor const& functor, double weight = -1.0,
                           size_t seed = 0x39abdc4529b1661c) {
    const hpsp_decay d = detail::make_decay<N>(fMotherMass, fMasses, "Decays::Unweight");
    if constexpr (is_phase_space_functor<Functor>::value) {
      const hpsp_functor f = detail::to_hpsp(functor);
      return unweight(d, &f, weight, seed);
    } else {
      return unweight_user(d, functor, weight, seed);
    }
  }

  std::array<double, N> GetMasses() const { return fMasses; }
  double GetMaxWeight() const { return fMaxWeight; }
  double GetMotherMass() const { return fMotherMass; }

  // ---- raw device access (SoA)
  // Non-const access may write the momenta: the container then no longer
  // knows which Generate's events it holds, nor their PhaseSpaceWeight.
  double* column(size_t particle, int component) {
    materialize();
    fGenValid = false;
    fWValid = false;
    return fCols.data() + (4 * particle + component) * fCap;
  }
  const double* column(size_t particle, int component) const {
    materialize();
    return fCols.data() + (4 * particle + component) * fCap;
  }
  double* data() {
    materialize();
    fGenValid = false;
    fWValid = false;
    return fCols.data();
  }
  const double* data() const {
    materialize();
    return fCols.data();
  }
  size_t ld() const { return fCap; }

  // For a writer that replaces every column: no materialisation, and the
  // container forgets which events it held and their weights (the writer
  // states them again with set_generated / set_weights_cache_valid).
  double* columns_for_overwrite() {
    fLazy = false;
    fGenValid = false;
    fWValid = false;
    return fCols.data();
  }
  // Materialise deferred momenta without forgetting anything: for a writer
  // that replaces only some rows and must keep the others.
  void materialize_rows() const { materialize(); }

  // PhaseSpace::Generate(mother, events) with this container's decay:
  // nothing has been written; the momenta of events 0 .. size() - 1 of
  // (d, seed) are formed when first read (any access to the columns, a row, a
  // particle range, a copy to the host) or written straight to their places
  // by Unweight().  A function of (d, seed, index) alone, so the deferred
  // write gives the bits an eager one would.
  void set_generated_lazy(hpsp_decay const& d, uint64_t seed) {
    set_generated(d, seed);
    fLazy = true;
  }
  bool lazy() const { return fLazy; }
  // True once a read has had to form deferred momenta: this container is
  // read after Generate, so PhaseSpace::Generate writes its momenta with the
  // weights in one pass from then on instead of deferring them.
  bool read_after_generate() const { return fReadAfterGen; }

  // The events are 0 .. size() - 1 of one FP64 PhaseSpace::Generate(mother,
  // ...) with decay d and seed, untouched since (every non-const access to
  // the columns forgets it).  Unweight() then generates them again at their
  // destinations instead of moving them: the same bits, 8 bytes read per
  // event instead of 32N.
  void set_generated(hpsp_decay const& d, uint64_t seed) {
    fGen = d;
    fGenSeed = seed;
    fGenValid = true;
    if (fCheck) guard_record();
  }

  // Opt-in check of the deferred path (also on when the environment variable
  // HYDRA_DEFERRED_CHECKS is set to 1 when the container is made): a
  // checksum of the columns taken when Generate marks them as its events
  // (and after deferred momenta are formed) is compared before Unweight()
  // regenerates them; a write through a column pointer taken before the
  // Generate, which the regeneration would silently discard, throws
  // std::logic_error instead.  Costs two reads of the columns per
  // Generate/Unweight pair.
  void SetDeferredChecks(bool on) {
    fCheck = on;
    if (on && fGenValid) guard_record();
  }
  bool DeferredChecks() const { return fCheck; }
  bool generated() const { return fGenValid; }

  // PhaseSpaceWeight column kept alongside the momenta (null when unknown).
  const double* weights_cache() const { return fWValid ? fW.data() : nullptr; }
  double* weights_cache_for_write() { return fW.data(); }
  void set_weights_cache_valid(bool v) { fWValid = v; }

 private:
  template <size_t... I>
  static value_type make_row(const std::array<double, 4 * N>& v, std::index_sequence<I...>) {
    // column (4*i + c) -> row entry i
    return value_type(Vector4R(v[4 * I], v[4 * I + 1], v[4 * I + 2], v[4 * I + 3])...);
  }

  static bool checks_from_env() {
    const char* e = std::getenv("HYDRA_DEFERRED_CHECKS");
    return e && e[0] == '1';
  }

  uint64_t columns_checksum() const {
    detail::device_buffer<uint64_t> sum(1);
    uint64_t h = 0;
    detail::hpsp_check(hpsp_checksum(fCols.data(), (uint64_t)(4 * N * fCap * sizeof(double)),
                                     sum.data(), nullptr),
                       "Decays (deferred checks)");
    detail::hpsp_check(hpsp_memcpy_to_host(&h, sum.data(), sizeof(h)), "Decays (deferred checks)");
    return h;
  }
  void guard_record() const { fGuard = columns_checksum(); }
  void guard_verify() const {
    if (columns_checksum() != fGuard)
      throw std::logic_error(
          "[hydra::Decays::Unweight]: the columns of Generate's events were written through a "
          "pointer taken before that Generate; Unweight() regenerates those events and would "
          "discard the writes (take column()/data() after Generate to write momenta)");
  }

  void init(size_t n) {
    const hpsp_decay d = detail::make_decay<N>(fMotherMass, fMasses, "Decays");
    fMaxWeight = d.max_weight;
    regrow(n);
    fSize = n;
  }

  // Write the deferred momenta of a lazy Generate.  Const accessors call it,
  // so concurrent const use of one container from several host threads needs
  // the caller's synchronisation (as any use of the reference's containers).
  void materialize() const {
    if (!fLazy) return;
    if (fSize > 0) {
      detail::hpsp_check(hpsp_generate(&fGen, fGenSeed, 0, fSize, HPSP_FP64, fCols.data(), fCap,
                                       0, nullptr, nullptr, nullptr),
                         "Decays (deferred Generate)");
      detail::hpsp_sync("Decays (deferred Generate)");
      fReadAfterGen = true;
    }
    fLazy = false;
    if (fCheck && fGenValid) guard_record();
  }

  // Columns for at least cap events, column_stride(cap) apart.
  void regrow(size_t cap) {
    materialize();
    const size_t ld = detail::column_stride(cap);
    detail::device_buffer<double> cols(4 * N * ld);
    detail::device_buffer<double> w(ld);
    const size_t keep = fSize < cap ? fSize : cap;
    if (keep && fCols.data())
      for (size_t c = 0; c < 4 * N; ++c)
        detail::hpsp_check(hpsp_memcpy_device(cols.data() + c * ld, fCols.data() + c * fCap,
                                              keep * sizeof(double), nullptr),
                           "Decays::regrow");
    if (keep && fWValid)
      detail::hpsp_check(hpsp_memcpy_device(w.data(), fW.data(), keep * sizeof(double), nullptr),
                         "Decays::regrow");
    detail::hpsp_sync("Decays::regrow");
    fCols.swap(cols);
    fW.swap(w);
    fCap = ld;
  }

  // hydra::unweight(device::sys, events, functor, ...) (Random.inl:134-230):
  // u * max <= w with w the functor's values, kept events first; the
  // rejected ones in the order of the reference's sequential partition.
  // Permutes the container out of place.
 public:
  Range<iterator> UnweightBy(EventWeights const& w, double max_pdf, size_t rng_seed,
                             size_t rng_jump) {
    materialize();
    const size_t n = fSize;
    if (w.size() < n) throw std::invalid_argument("[hydra::unweight]: fewer weights than events");
    auto& out = spare_cols(n);
    const size_t wsb = hpsp_unweight_sample_workspace_size(n);
    unsigned char* ws = spare_ws(wsb);
    const double* wc = weights_cache();
    // the container's own PhaseSpaceWeight column moves with the events
    detail::hpsp_check(hpsp_unweight_sample(4 * (int)N, fCols.data(), fCap, n, w.data(), max_pdf,
                                            (uint64_t)rng_seed, (uint64_t)rng_jump,
                                            HPSP_ORDER_SEQUENTIAL, out.data(),
                                            detail::column_stride(n), nullptr, nullptr, ws, wsb,
                                            spare_nacc(), nullptr),
                       "hydra::unweight");
    detail::device_buffer<double>* wout = nullptr;
    if (wc) {
      // reorder the cached weights with the same permutation (as a one-column sample)
      wout = &spare_w(n);
      detail::hpsp_check(hpsp_unweight_sample(1, wc, n, n, w.data(), max_pdf, (uint64_t)rng_seed,
                                              (uint64_t)rng_jump, HPSP_ORDER_SEQUENTIAL,
                                              wout->data(), n ? n : 1, nullptr, nullptr, ws, wsb,
                                              spare_nacc(), nullptr),
                         "hydra::unweight");
    }
    fWValid = wc != nullptr;
    return adopt(out, wout);
  }

 private:
  // The events written once, at their places (the PhaseSpaceWeight column
  // is not moved with them, as the reference's Decays holds none: it is
  // formed again from the momenta if it is needed again).
  Range<iterator> unweight_regenerate(size_t seed) {
    const size_t n = fSize;
    auto& out = spare_cols(n);
    const size_t wsb = hpsp_unweight_workspace_size(n);
    detail::hpsp_check(hpsp_unweight_regenerate(&fGen, fGenSeed, 0, n, fW.data(), (uint64_t)seed,
                                                out.data(), detail::column_stride(n), nullptr,
                                                nullptr, spare_ws(wsb), wsb, spare_nacc(),
                                                nullptr),
                       "Decays::Unweight");
    fWValid = false;
    return adopt(out, nullptr);
  }

  // Events that Generate has not formed (or whose weight column is not
  // kept): flags from each event's weight chain, PhaseSpaceWeight only where
  // the chain leaves the test in doubt, then every event generated at its
  // place; the same container as unweight_regenerate's.
  Range<iterator> generate_unweight(size_t seed) {
    const size_t n = fSize;
    auto& out = spare_cols(n);
    const size_t wsb = hpsp_generate_unweight_workspace_size(n);
    detail::hpsp_check(hpsp_generate_unweight(&fGen, fGenSeed, 0, n, (uint64_t)seed, out.data(),
                                              detail::column_stride(n), nullptr, nullptr,
                                              spare_ws(wsb), wsb, spare_nacc(), nullptr),
                       "Decays::Unweight");
    fWValid = false;
    return adopt(out, nullptr);
  }

  // Takes the unweighted columns (and, when given, the weight column moved
  // with them); the container's previous buffers become the spares.
  Range<iterator> adopt(detail::device_buffer<double>& out, detail::device_buffer<double>* wout) {
    uint64_t acc = 0;
    detail::hpsp_check(hpsp_memcpy_to_host(&acc, fSpare.nacc.data(), sizeof(acc)),
                       "Decays::Unweight");
    fCols.swap(out);
    if (wout) fW.swap(*wout);
    fCap = detail::column_stride(fSize);
    fGenValid = false;
    fLazy = false;
    return Range<iterator>(begin(), begin() + (std::ptrdiff_t)acc);
  }

  // Buffers that the Unweight calls fill and then swap with the container's
  // own (output columns, weight column) or use as workspace: kept from call to
  // call, so that repeated Generate/Unweight cycles of one size allocate
  // nothing after the first.  A copy of the container starts without them.
  struct Spare {
    detail::device_buffer<double> cols, w;
    detail::device_buffer<unsigned char> ws;
    detail::device_buffer<uint64_t> nacc;
    Spare() = default;
    Spare(const Spare&) {}
    Spare(Spare&&) noexcept = default;
    Spare& operator=(const Spare&) { return *this; }
    Spare& operator=(Spare&&) noexcept = default;
  };
  // sized for n events: reallocated when too small or more than twice too large
  static void fit(detail::device_buffer<double>& b, size_t need) {
    if (b.size() < need || b.size() / 2 > need) b.allocate(need);
  }
  detail::device_buffer<double>& spare_cols(size_t n) {
    fit(fSpare.cols, 4 * N * detail::column_stride(n));
    return fSpare.cols;
  }
  detail::device_buffer<double>& spare_w(size_t n) {
    fit(fSpare.w, n ? n : 1);
    return fSpare.w;
  }
  unsigned char* spare_ws(size_t bytes) {
    if (fSpare.ws.size() < bytes) fSpare.ws.allocate(bytes);
    return fSpare.ws.data();
  }
  uint64_t* spare_nacc() {
    if (fSpare.nacc.size() < 1) fSpare.nacc.allocate(1);
    return fSpare.nacc.data();
  }

  // Unweight(functor) with a user's callable: its PhaseSpaceReweight values
  // from a kernel compiled here, the partition by the library.  The kept
  // PhaseSpaceWeight column is not carried through (it is recomputed when
  // next needed).
  template <typename Functor>
  Range<iterator> unweight_user(const hpsp_decay& d, Functor const& functor, double max_value,
                                size_t seed) {
    materialize();
    const size_t n = fSize;
    static_assert(detail::user_width<hydra::tuple<Particles...>, Functor>() == 1,
                  "Unweight(functor): the functor returns one value per event");
    detail::device_buffer<double> vals(n ? n : 1);
    detail::user_event_values((hydra::tuple<Particles...>*)nullptr, d, fCols.data(), fCap, n,
                              true, vals.data(), n ? n : 1, functor);
    auto& out = spare_cols(n);
    const size_t wsb = hpsp_unweight_workspace_size(n);
    detail::hpsp_check(hpsp_unweight_values((int)N, vals.data(), max_value, (uint64_t)seed,
                                            fCols.data(), fCap, n, out.data(),
                                            detail::column_stride(n),
                                            nullptr, nullptr, spare_ws(wsb), wsb, spare_nacc(),
                                            nullptr),
                       "Decays::Unweight");
    fWValid = false;
    return adopt(out, nullptr);
  }

  Range<iterator> unweight(const hpsp_decay& d, const hpsp_functor* f, double max_value,
                           size_t seed) {
    materialize();
    const size_t n = fSize;
    auto& out = spare_cols(n);
    const size_t wsb = hpsp_unweight_workspace_size(n);
    const double* win = weights_cache();
    detail::device_buffer<double>* wout = win ? &spare_w(n) : nullptr;
    detail::hpsp_check(hpsp_unweight(&d, f, max_value, (uint64_t)seed, fCols.data(), fCap, n,
                                     win, out.data(), detail::column_stride(n), nullptr,
                                     wout ? wout->data() : nullptr, spare_ws(wsb), wsb,
                                     spare_nacc(), nullptr),
                       "Decays::Unweight");
    return adopt(out, wout);
  }

  double fMotherMass = 0.0;
  double fMaxWeight = 0.0;
  std::array<double, N> fMasses{};
  size_t fSize = 0;
  size_t fCap = 0;
  mutable detail::device_buffer<double> fCols;
  detail::device_buffer<double> fW;
  bool fWValid = false;
  hpsp_decay fGen{};
  uint64_t fGenSeed = 0;
  bool fGenValid = false;
  mutable bool fLazy = false;
  mutable bool fReadAfterGen = false;
  bool fCheck = checks_from_env();
  mutable uint64_t fGuard = 0;
  Spare fSpare;
};

// Decays in host memory (the reference's Decays<tuple<P...>, host::sys_t>):
// the same structure-of-arrays columns, copied to and from the device in one
// strided transfer.  Rows read as tuples of the particle types.
template <typename... Particles>
class Decays<hydra::tuple<Particles...>, hydra::host::sys_t> {
 public:
  static constexpr size_t N = sizeof...(Particles);
  enum { nparticles = N };
  typedef hydra::tuple<Particles...> value_type;
  typedef hydra::tuple<Particles...> tuple_type;
  typedef hydra::host::sys_t system_type;

  Decays(double motherMass, std::array<double, N> const& daughtersMasses, size_t nentries = 0)
      : fMotherMass(motherMass), fMasses(daughtersMasses), fSize(nentries),
        fCols(4 * N * nentries) {
    fMaxWeight = detail::make_decay<N>(fMotherMass, fMasses, "Decays").max_weight;
  }
  // trans-backend copy from the device (Decays.h:180-193)
  explicit Decays(Decays<hydra::tuple<Particles...>, hydra::device::sys_t> const& other)
      : fMotherMass(other.GetMotherMass()), fMaxWeight(other.GetMaxWeight()),
        fMasses(other.GetMasses()), fSize(other.size()), fCols(4 * N * other.size()) {
    if (fSize)
      detail::hpsp_check(hpsp_memcpy2d_to_host(fCols.data(), fSize * sizeof(double), other.data(),
                                               other.ld() * sizeof(double),
                                               fSize * sizeof(double), 4 * N),
                         "Decays(host <- device)");
  }

  size_t size() const { return fSize; }
  size_t ld() const { return fSize; }
  double GetMotherMass() const { return fMotherMass; }
  double GetMaxWeight() const { return fMaxWeight; }
  std::array<double, N> GetMasses() const { return fMasses; }
  const double* data() const { return fCols.data(); }
  double* data() { return fCols.data(); }
  const double* column(size_t particle, int component) const {
    return fCols.data() + (4 * particle + component) * fSize;
  }
  value_type operator[](size_t k) const {
    return make_row(k, std::make_index_sequence<N>{});
  }

 private:
  template <size_t... I>
  value_type make_row(size_t k, std::index_sequence<I...>) const {
    return value_type(Vector4R(column(I, 0)[k], column(I, 1)[k], column(I, 2)[k],
                               column(I, 3)[k])...);
  }

  double fMotherMass;
  double fMaxWeight;
  std::array<double, N> fMasses;
  size_t fSize;
  std::vector<double> fCols;
};

template <typename... Particles>
Decays<hydra::tuple<Particles...>, hydra::device::sys_t>::Decays(
    Decays<hydra::tuple<Particles...>, hydra::host::sys_t> const& other)
    : fMotherMass(other.GetMotherMass()), fMasses(other.GetMasses()) {
  init(other.size());
  fMaxWeight = other.GetMaxWeight();
  if (fSize)
    detail::hpsp_check(hpsp_memcpy2d_to_device(fCols.data(), fCap * sizeof(double), other.data(),
                                               other.ld() * sizeof(double),
                                               fSize * sizeof(double), 4 * N),
                       "Decays(device <- host)");
}

// assignment keeps this container's masses, as the reference's operator=
// copies rows only (Decays.h:606-618)
template <typename... Particles>
Decays<hydra::tuple<Particles...>, hydra::device::sys_t>&
Decays<hydra::tuple<Particles...>, hydra::device::sys_t>::operator=(
    Decays<hydra::tuple<Particles...>, hydra::host::sys_t> const& other) {
  columns_for_overwrite();  // every column is replaced: nothing to keep
  resize(other.size());
  fWValid = false;
  if (fSize)
    detail::hpsp_check(hpsp_memcpy2d_to_device(fCols.data(), fCap * sizeof(double), other.data(),
                                               other.ld() * sizeof(double),
                                               fSize * sizeof(double), 4 * N),
                       "Decays(device <- host)");
  return *this;
}

// events | weight functor  (Range2.inl:213-238): weights computed on the device.
template <typename... Particles>
EventWeights operator|(const Decays<hydra::tuple<Particles...>, hydra::device::sys_t>& ev,
                       const PhaseSpaceWeight<Particles...>& w) {
  EventWeights out(ev.size());
  if (ev.size()) {
    detail::hpsp_check(hpsp_event_weight(&w.descriptor(), nullptr, ev.data(), ev.ld(),
                                         ev.size(), out.data(), nullptr),
                       "PhaseSpaceWeight");
    detail::hpsp_sync("PhaseSpaceWeight");
  }
  return out;
}

template <typename Functor, typename... Particles>
EventWeights operator|(const Decays<hydra::tuple<Particles...>, hydra::device::sys_t>& ev,
                       const PhaseSpaceReweight<Functor, Particles...>& w) {
  EventWeights out(ev.size());
  if (ev.size()) {
    detail::hpsp_check(hpsp_event_weight(&w.descriptor(), &w.functor(), ev.data(), ev.ld(),
                                         ev.size(), out.data(), nullptr),
                       "PhaseSpaceReweight");
    detail::hpsp_sync("PhaseSpaceReweight");
  }
  return out;
}

// events | events.GetEventWeightFunctor(f) with a user's callable f.
template <typename Functor, typename... Particles>
EventWeights operator|(const Decays<hydra::tuple<Particles...>, hydra::device::sys_t>& ev,
                       const UserReweight<Functor, Particles...>& w) {
  static_assert(detail::user_width<hydra::tuple<Particles...>, Functor>() == 1,
                "GetEventWeightFunctor(functor): the functor returns one value per event");
  EventWeights out(ev.size());
  detail::user_event_values((hydra::tuple<Particles...>*)nullptr, w.descriptor(), ev.data(),
                            ev.ld(), ev.size(), true, out.data(), ev.size() ? ev.size() : 1,
                            w.functor());
  return out;
}

namespace detail {

template <size_t K>
struct value_columns {
  template <size_t... I>
  static auto make(std::index_sequence<I...>) -> hydra::tuple<decltype((void)I, double())...>;
  using type = multivector<decltype(make(std::make_index_sequence<K>{})), hydra::device::sys_t>;
};

// rows [first, first + n) of a container through a wrapped lambda
template <typename L, typename... Particles>
auto lambda_rows(const Decays<hydra::tuple<Particles...>, hydra::device::sys_t>& ev, size_t first,
                 size_t n, hydra::Lambda<L, 0> const& f) {
  constexpr int K = user_width<hydra::tuple<Particles...>, hydra::Lambda<L, 0>>();
  const hpsp_decay d = make_decay<sizeof...(Particles)>(ev.GetMotherMass(), ev.GetMasses(),
                                                        "events | functor");
  if constexpr (K == 1) {
    EventWeights out(n);
    user_event_values((hydra::tuple<Particles...>*)nullptr, d, ev.data() + first, ev.ld(), n,
                      false, out.data(), n ? n : 1, f);
    return out;
  } else {
    typename value_columns<(size_t)K>::type out(n);
    user_event_values((hydra::tuple<Particles...>*)nullptr, d, ev.data() + first, ev.ld(), n,
                      false, out.data(), out.ld(), f);
    return out;
  }
}

}  // namespace detail

// events | lambda (Range2.inl:213-238) materialised on the device: a lambda
// of one value gives a column (EventWeights), of a tuple of K values a
// multivector of K columns (the Dalitz variables of
// examples/phase_space/phsp_basic.inl:159-163 and 194).
template <typename L, typename... Particles>
auto operator|(const Decays<hydra::tuple<Particles...>, hydra::device::sys_t>& ev,
               hydra::Lambda<L, 0> const& f) {
  return detail::lambda_rows(ev, 0, ev.size(), f);
}

// A range of a container's rows (what Unweight returns) through a lambda
// (examples/phase_space/phsp_unweighting_functor.inl:209).
template <typename It, typename L, typename = decltype(std::declval<It const&>().container())>
auto operator|(Range<It> const& r, hydra::Lambda<L, 0> const& f) {
  return detail::lambda_rows(*r.begin().container(), r.begin().index(), r.size(), f);
}

// hydra::unweight(device::sys, events, functor[, max_pdf, rng_seed, rng_jump])
// (hydra/Random.h:133-330, detail/Random.inl:134-230), as
// examples/phase_space/phsp_unweighting_functor.inl:209 calls it with
// events.GetEventWeightFunctor(f): the functor's values weight the events.
template <typename... Particles, typename W>
auto unweight(hydra::device::sys_t const&,
              Decays<hydra::tuple<Particles...>, hydra::device::sys_t>& events, W const& functor,
              double max_pdf = -1.0, size_t rng_seed = 0x8ec74d321e6b5a27, size_t rng_jump = 0)
    -> decltype(events.UnweightBy(events | functor, max_pdf, rng_seed, rng_jump)) {
  EventWeights w = events | functor;
  return events.UnweightBy(w, max_pdf, rng_seed, rng_jump);
}

}  // namespace hydra
