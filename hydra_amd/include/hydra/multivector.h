// hydra::multivector<hydra::tuple<T...>, System> of scalar columns
// (reference: hydra/multivector.h, the container PhaseSpace::Evaluate fills,
// examples/phase_space/phsp_evaluating_functor.inl:204-210).  Each T is a
// double or a declarg of double.  The device version keeps one FP64 column of
// ld() values per tuple element in device memory (column j at data() +
// j * ld(), ld() = detail::column_stride(size())); the host version keeps
// its columns size() apart in host memory and is constructed from a device
// one.  Rows are read as tuples by value.
#pragma once

#include <cstddef>
#include <tuple>
#include <utility>
#include <vector>

#include <hydra/Range.h>
#include <hydra/Tuple.h>
#include <hydra/detail/Runtime.h>
#include <hydra/device/System.h>
#include <hydra/host/System.h>

namespace hydra {

template <typename Tuple, typename System>
class multivector;

namespace detail {

template <typename Row, typename Source, size_t... I>
Row make_row(const Source& src, size_t i, std::index_sequence<I...>) {
  return Row(typename std::tuple_element<I, Row>::type(src.value(I, i))...);
}

// Random-access iterator over rows of a multivector; dereferencing reads one
// row.
template <typename Container>
class mv_iterator {
 public:
  using value_type = typename Container::value_type;
  using difference_type = std::ptrdiff_t;
  mv_iterator() = default;
  mv_iterator(const Container* c, size_t i) : c_(c), i_(i) {}
  value_type operator*() const { return (*c_)[i_]; }
  mv_iterator& operator++() { ++i_; return *this; }
  mv_iterator operator+(difference_type d) const { return mv_iterator(c_, i_ + d); }
  mv_iterator operator-(difference_type d) const { return mv_iterator(c_, i_ - d); }
  difference_type operator-(const mv_iterator& o) const {
    return (difference_type)i_ - (difference_type)o.i_;
  }
  bool operator==(const mv_iterator& o) const { return i_ == o.i_ && c_ == o.c_; }
  bool operator!=(const mv_iterator& o) const { return !(*this == o); }

 private:
  const Container* c_ = nullptr;
  size_t i_ = 0;
};

}  // namespace detail

template <typename... T>
class multivector<hydra::tuple<T...>, hydra::device::sys_t> {
 public:
  using value_type = hydra::tuple<T...>;
  using iterator = detail::mv_iterator<multivector>;
  static constexpr size_t ncols = sizeof...(T);

  multivector() = default;
  explicit multivector(size_t n) : n_(n), buf_(ncols * detail::column_stride(n)) {}

  size_t size() const { return n_; }
  size_t ld() const { return detail::column_stride(n_); }
  double* data() { return buf_.data(); }
  const double* data() const { return buf_.data(); }
  double* column(size_t j) { return buf_.data() + j * ld(); }
  const double* column(size_t j) const { return buf_.data() + j * ld(); }

  double value(size_t j, size_t i) const {
    double v;
    detail::hpsp_check(hpsp_memcpy_to_host(&v, column(j) + i, sizeof(double)), "multivector");
    return v;
  }
  value_type operator[](size_t i) const {
    return detail::make_row<value_type>(*this, i, std::index_sequence_for<T...>{});
  }
  iterator begin() const { return iterator(this, 0); }
  iterator end() const { return iterator(this, n_); }

 private:
  size_t n_ = 0;
  detail::device_buffer<double> buf_;
};

template <typename... T>
class multivector<hydra::tuple<T...>, hydra::host::sys_t> {
 public:
  using value_type = hydra::tuple<T...>;
  using iterator = detail::mv_iterator<multivector>;
  static constexpr size_t ncols = sizeof...(T);

  explicit multivector(size_t n = 0) : n_(n), data_(ncols * (n ? n : 1)) {}
  // copy from the device (one transfer of all columns)
  explicit multivector(multivector<hydra::tuple<T...>, hydra::device::sys_t> const& d)
      : n_(d.size()), data_(ncols * (d.size() ? d.size() : 1)) {
    if (n_)
      detail::hpsp_check(hpsp_memcpy2d_to_host(data_.data(), n_ * sizeof(double), d.data(),
                                               d.ld() * sizeof(double), n_ * sizeof(double),
                                               ncols),
                         "multivector");
  }

  size_t size() const { return n_; }
  size_t ld() const { return n_ ? n_ : 1; }
  const double* column(size_t j) const { return data_.data() + j * ld(); }
  double value(size_t j, size_t i) const { return column(j)[i]; }
  value_type operator[](size_t i) const {
    return detail::make_row<value_type>(*this, i, std::index_sequence_for<T...>{});
  }
  iterator begin() const { return iterator(this, 0); }
  iterator end() const { return iterator(this, n_); }

 private:
  size_t n_;
  std::vector<double> data_;
};

}  // namespace hydra
