// hydra::DenseHistogram<double, N, device::sys_t> and make_dense_histogram
// (reference: hydra/DenseHistogram.h:54-636, detail/DenseHistogram.inl), as
// the phase-space examples use them on Generate's output
// (examples/phase_space/phsp_basic.inl:196-214).
//
// Contents live on the device: nbins + 2 doubles, bins in row-major order,
// then the underflow and overflow slots; a Fill overwrites them, as the
// reference's does.  Data come from
//   * value columns: a multivector of N columns (for example Evaluate's
//     functor columns) with optional weights;
//   * melded events: `events.Meld(others) | hydra::make_axes(f1, ..., fN)`,
//     each axis read from the particles it uses, weights optional;
//   * phase_space_range elements: `range | hydra::make_axes(f1, ..., fN)`,
//     generated inside the fill, unweighted or weighted by
//     range.GetEventWeightFunctor() (each element's own weight);
//   * the events themselves: `events | hydra::make_axes(f1, ..., fN)` names N
//     device-described functors of the daughters (the reference's
//     `events | dalitz_calculator`); the fill computes them on the fly, with
//     weights from `events | events.GetEventWeightFunctor()` (a weight
//     column) or from `events.GetEventWeightFunctor()` itself (PhaseSpaceWeight
//     recomputed inside the fill).
// Bins follow GetGlobalBin exactly; sums inside a bin are accumulated in
// another order than the reference's sorted reduce_by_key (equal to rounding;
// unit weights exactly).
#pragma once

#include <array>
#include <cstddef>
#include <limits>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include <hydra/Decays.h>
#include <hydra/PhaseSpaceFunctors.h>
#include <hydra/PhaseSpaceRange.h>
#include <hydra/detail/Runtime.h>
#include <hydra/device/System.h>
#include <hydra/multivector.h>

#include "hpsp.h"

namespace hydra {

// events | make_axes(f...): the N functor values of every event, not
// materialised; consumed by DenseHistogram::Fill.
template <size_t N, typename EventsT>
class EventAxes {
 public:
  EventAxes(const EventsT& events, std::array<hpsp_functor, N> const& axes)
      : events_(&events), axes_(axes) {}
  const EventsT& events() const { return *events_; }
  const std::array<hpsp_functor, N>& axes() const { return axes_; }

 private:
  const EventsT* events_;
  std::array<hpsp_functor, N> axes_;
};

template <typename... F>
struct AxesSpec {
  std::array<hpsp_functor, sizeof...(F)> axes;
};

template <typename... F>
AxesSpec<F...> make_axes(F const&... f) {
  static_assert(sizeof...(F) >= 1 && sizeof...(F) <= HPSP_HIST_MAX_DIM,
                "one to HPSP_HIST_MAX_DIM axes");
  static_assert(std::conjunction<is_phase_space_functor<F>...>::value,
                "axes must be device-described phase-space functors");
  return AxesSpec<F...>{{detail::to_hpsp(f)...}};
}

template <typename... Particles, typename... F>
EventAxes<sizeof...(F), Decays<hydra::tuple<Particles...>, hydra::device::sys_t>> operator|(
    const Decays<hydra::tuple<Particles...>, hydra::device::sys_t>& ev, AxesSpec<F...> const& s) {
  return EventAxes<sizeof...(F), Decays<hydra::tuple<Particles...>, hydra::device::sys_t>>(
      ev, s.axes);
}

// melded | make_axes(f...): axes of stored events given particle by particle
// (Decays::Meld), read by DenseHistogram::Fill where the functors use them.
template <size_t D>
class ParticleAxes {
 public:
  ParticleAxes(hpsp_particles const& p, size_t n, std::array<hpsp_functor, D> const& axes)
      : p_(p), n_(n), axes_(axes) {}
  const hpsp_particles& particles() const { return p_; }
  size_t size() const { return n_; }
  const std::array<hpsp_functor, D>& axes() const { return axes_; }

 private:
  hpsp_particles p_;
  size_t n_;
  std::array<hpsp_functor, D> axes_;
};

template <size_t P, typename... F>
ParticleAxes<sizeof...(F)> operator|(MeldedDecays<P> const& m, AxesSpec<F...> const& s) {
  return ParticleAxes<sizeof...(F)>(m.particles(), m.size(), s.axes);
}

// range | make_axes(f...): the axes of phase_space_range elements, generated
// inside DenseHistogram::Fill.
template <size_t D, size_t N>
class RangeAxes {
 public:
  RangeAxes(const PhaseSpaceRange<N>& r, std::array<hpsp_functor, D> const& axes)
      : range_(&r), axes_(axes) {}
  const PhaseSpaceRange<N>& range() const { return *range_; }
  const std::array<hpsp_functor, D>& axes() const { return axes_; }

 private:
  const PhaseSpaceRange<N>* range_;
  std::array<hpsp_functor, D> axes_;
};

template <size_t N, typename... F>
RangeAxes<sizeof...(F), N> operator|(const PhaseSpaceRange<N>& r, AxesSpec<F...> const& s) {
  return RangeAxes<sizeof...(F), N>(r, s.axes);
}

template <typename T, size_t N, typename Backend = hydra::device::sys_t>
class DenseHistogram;

template <size_t N>
class DenseHistogram<double, N, hydra::device::sys_t> {
  static_assert(N >= 1 && N <= HPSP_HIST_MAX_DIM, "DenseHistogram: 1..HPSP_HIST_MAX_DIM axes");

 public:
  DenseHistogram(std::array<size_t, N> const& grid, std::array<double, N> const& lowerlimits,
                 std::array<double, N> const& upperlimits) {
    init(grid.data(), lowerlimits.data(), upperlimits.data());
  }
  template <size_t M = N, typename = typename std::enable_if<M == 1>::type>
  DenseHistogram(size_t grid, double lowerlimit, double upperlimit) {
    init(&grid, &lowerlimit, &upperlimit);
  }

  // Fill(data) / Fill(data, weights) from N value columns
  template <typename... T>
  DenseHistogram& Fill(multivector<hydra::tuple<T...>, hydra::device::sys_t> const& data) {
    return fill_columns(data, nullptr);
  }
  template <typename... T>
  DenseHistogram& Fill(multivector<hydra::tuple<T...>, hydra::device::sys_t> const& data,
                       EventWeights const& weights) {
    if (weights.size() < data.size())
      throw std::invalid_argument("[hydra::DenseHistogram::Fill]: fewer weights than data");
    return fill_columns(data, weights.data());
  }

  // Fill(data) / Fill(data, weights) from one value column (a one-value
  // lambda over the events, `events | lambda`)
  template <size_t M = N, typename = typename std::enable_if<M == 1>::type>
  DenseHistogram& Fill(EventWeights const& data) {
    return fill_column(data, nullptr);
  }
  template <size_t M = N, typename = typename std::enable_if<M == 1>::type>
  DenseHistogram& Fill(EventWeights const& data, EventWeights const& weights) {
    if (weights.size() < data.size())
      throw std::invalid_argument("[hydra::DenseHistogram::Fill]: fewer weights than data");
    return fill_column(data, weights.data());
  }

  // Fused fills from the events
  template <typename EventsT>
  DenseHistogram& Fill(EventAxes<N, EventsT> const& axes) {
    return fill_events(axes, HPSP_HIST_WEIGHT_ONE, nullptr);
  }
  template <typename EventsT>
  DenseHistogram& Fill(EventAxes<N, EventsT> const& axes, EventWeights const& weights) {
    if (weights.size() < axes.events().size())
      throw std::invalid_argument("[hydra::DenseHistogram::Fill]: fewer weights than events");
    return fill_events(axes, HPSP_HIST_WEIGHT_COLUMN, weights.data());
  }
  template <typename EventsT, typename... Particles>
  DenseHistogram& Fill(EventAxes<N, EventsT> const& axes,
                       PhaseSpaceWeight<Particles...> const& weight) {
    const hpsp_decay& d = weight.descriptor();
    return fill_events(axes, HPSP_HIST_WEIGHT_PHSP, nullptr, &d);
  }

  // Fills over phase_space_range elements (range_semantics.inl:154-167):
  // unweighted, or weighted by each element's weight
  template <size_t NR>
  DenseHistogram& Fill(RangeAxes<N, NR> const& axes) {
    return fill_range(axes, HPSP_HIST_WEIGHT_ONE);
  }
  template <size_t NR>
  DenseHistogram& Fill(RangeAxes<N, NR> const& axes, RangeEventWeight) {
    return fill_range(axes, HPSP_HIST_WEIGHT_GENERATION);
  }

  // Fills over melded events (phsp_chain.inl:244-281), weights optional
  DenseHistogram& Fill(ParticleAxes<N> const& axes) { return fill_particles(axes, nullptr); }
  DenseHistogram& Fill(ParticleAxes<N> const& axes, EventWeights const& weights) {
    if (weights.size() < axes.size())
      throw std::invalid_argument("[hydra::DenseHistogram::Fill]: fewer weights than events");
    return fill_particles(axes, weights.data());
  }

  size_t GetNBins() const { return (size_t)nbins_; }
  std::array<size_t, N> GetGrid() const { return grid_; }
  double GetBinContent(size_t bin) const {
    if (bin > nbins_ + 1) return std::numeric_limits<double>::max();
    double v;
    detail::hpsp_check(hpsp_memcpy_to_host(&v, contents_.data() + bin, sizeof(double)),
                       "DenseHistogram");
    return v;
  }
  double GetBinContent(std::array<size_t, N> const& bins) const {
    size_t bin = 0;
    for (size_t i = 0; i < N; ++i) {
      size_t prod = 1;
      for (size_t j = N - 1; j > i; --j) prod *= grid_[j];
      bin += prod * bins[i];
    }
    return bin < nbins_ ? GetBinContent(bin) : std::numeric_limits<double>::max();
  }
  // all nbins + 2 contents, copied to the host
  std::vector<double> GetBinsContents() const {
    std::vector<double> h(nbins_ + 2);
    detail::hpsp_check(
        hpsp_memcpy_to_host(h.data(), contents_.data(), h.size() * sizeof(double)),
        "DenseHistogram");
    return h;
  }
  const double* data() const { return contents_.data(); }

 private:
  void init(const size_t* grid, const double* lo, const double* hi) {
    h_ = hpsp_histogram{};
    h_.ndim = (int32_t)N;
    for (size_t i = 0; i < N; ++i) {
      grid_[i] = grid[i];
      h_.grid[i] = grid[i];
      h_.lower[i] = lo[i];
      h_.upper[i] = hi[i];
    }
    nbins_ = hpsp_histogram_bins(&h_);
    if (nbins_ == 0) throw std::invalid_argument("[hydra::DenseHistogram]: bad grid or limits");
    contents_.allocate(nbins_ + 2);
    detail::hpsp_check(hpsp_histogram_fill(&h_, nullptr, 0, 0, nullptr, contents_.data(), nullptr),
                       "DenseHistogram");
    detail::hpsp_sync("DenseHistogram");
  }

  template <typename MV>
  DenseHistogram& fill_columns(MV const& data, const double* w) {
    static_assert(MV::ncols == N, "data needs one column per axis");
    detail::hpsp_check(hpsp_histogram_fill(&h_, data.data(), data.ld(), data.size(), w,
                                           contents_.data(), nullptr),
                       "DenseHistogram::Fill");
    detail::hpsp_sync("DenseHistogram::Fill");
    return *this;
  }

  DenseHistogram& fill_column(EventWeights const& data, const double* w) {
    detail::hpsp_check(hpsp_histogram_fill(&h_, data.data(), data.size() ? data.size() : 1,
                                           data.size(), w, contents_.data(), nullptr),
                       "DenseHistogram::Fill");
    detail::hpsp_sync("DenseHistogram::Fill");
    return *this;
  }

  template <typename EventsT>
  DenseHistogram& fill_events(EventAxes<N, EventsT> const& ax, int mode, const double* w,
                              const hpsp_decay* dw = nullptr) {
    const EventsT& ev = ax.events();
    const auto masses = ev.GetMasses();
    constexpr size_t ND = std::tuple_size<decltype(masses)>::value;
    const hpsp_decay d =
        dw ? *dw : detail::make_decay<ND>(ev.GetMotherMass(), masses, "DenseHistogram");
    detail::hpsp_check(hpsp_histogram_fill_events(&h_, &d, ax.axes().data(), ev.data(), ev.ld(),
                                                  ev.size(), mode, w, contents_.data(), nullptr),
                       "DenseHistogram::Fill");
    detail::hpsp_sync("DenseHistogram::Fill");
    return *this;
  }

  DenseHistogram& fill_particles(ParticleAxes<N> const& ax, const double* w) {
    detail::hpsp_check(hpsp_histogram_fill_particles(&h_, &ax.particles(), ax.axes().data(),
                                                     ax.size(), w, contents_.data(), nullptr),
                       "DenseHistogram::Fill");
    detail::hpsp_sync("DenseHistogram::Fill");
    return *this;
  }

  template <size_t NR>
  DenseHistogram& fill_range(RangeAxes<N, NR> const& ax, int mode) {
    const PhaseSpaceRange<NR>& r = ax.range();
    if (r.size() == std::numeric_limits<size_t>::max())
      throw std::invalid_argument("[hydra::DenseHistogram::Fill]: unbounded range");
    detail::hpsp_check(hpsp_histogram_fill_range(&h_, &r.descriptor(), r.seed(), ax.axes().data(),
                                                 0, r.size(), mode, contents_.data(), nullptr),
                       "DenseHistogram::Fill");
    detail::hpsp_sync("DenseHistogram::Fill");
    return *this;
  }

  hpsp_histogram h_{};
  std::array<size_t, N> grid_{};
  uint64_t nbins_ = 0;
  detail::device_buffer<double> contents_;
};

// make_dense_histogram<double, N>(device::sys, grid, lower, upper, data[, weights])
// (DenseHistogram.inl:555-607)
template <typename T, size_t N, typename Data>
DenseHistogram<T, N, hydra::device::sys_t> make_dense_histogram(
    hydra::device::sys_t const&, std::array<size_t, N> const& grid,
    std::array<double, N> const& lower, std::array<double, N> const& upper, Data const& data) {
  DenseHistogram<T, N, hydra::device::sys_t> h(grid, lower, upper);
  h.Fill(data);
  return h;
}

template <typename T, size_t N, typename Data, typename Weights>
DenseHistogram<T, N, hydra::device::sys_t> make_dense_histogram(
    hydra::device::sys_t const&, std::array<size_t, N> const& grid,
    std::array<double, N> const& lower, std::array<double, N> const& upper, Data const& data,
    Weights const& weights) {
  DenseHistogram<T, N, hydra::device::sys_t> h(grid, lower, upper);
  h.Fill(data, weights);
  return h;
}

}  // namespace hydra
