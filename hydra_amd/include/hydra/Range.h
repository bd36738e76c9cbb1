// A [begin, end) view (reference: hydra/Range.h).
#pragma once

#include <cstddef>
#include <utility>

namespace hydra {

template <typename Iterator>
class Range {
 public:
  Range(Iterator b, Iterator e) : b_(b), e_(e) {}
  Iterator begin() const { return b_; }
  Iterator end() const { return e_; }
  size_t size() const { return (size_t)(e_ - b_); }
  auto operator[](size_t i) const -> decltype(*std::declval<Iterator>()) { return *(b_ + i); }

 private:
  Iterator b_, e_;
};

template <typename Iterator>
Range<Iterator> make_range(Iterator b, Iterator e) {
  return Range<Iterator>(b, e);
}

}  // namespace hydra
