// Per-event functors for AverageOn, Evaluate and Unweight(functor).
//
// The reference compiles any callable of the N daughters into its Thrust
// kernels (hydra/Lambda.h:133-200).  Here the kernels are prebuilt for gfx950,
// so a functor is a description the kernels evaluate on the device (the
// hpsp_functor of include/hpsp.h):
//   ConstantFunctor          f = 1
//   InvariantMass2{a, b}     f = (p[a] + p[b])^2 (default the pair (0, 1))
//   InvariantMass{a, b}      f = (p[a] + p[b]).mass() (Vector4R::mass)
//   CosHelicity{d, e}        helicity cosine of p[d] in the pair (d, e) within
//                            the sum of the daughters: the cos(theta_K*) of
//                            examples/phase_space/phsp_evaluating_functor.inl:160-176;
//                            CosHelicity{d, e, first, count}: within
//                            p[first] + ... + p[first + count - 1] (a sub-decay
//                            of melded events)
//   DecayPlaneAngle{d, e, h} angle in [0, 2 pi) between the plane of (d, e) and
//                            the plane of h about d + e, the delta phi of
//                            examples/phase_space/phsp_chain.inl:191-207; over
//                            stored events only (events | make_functors,
//                            melded events, their histograms)
//   BreitWignerSum           f = |sum_k c_k / (M_k^2 - s_k - i M_k G_k)|^2,
//                            s_k = (p[a_k] + p[b_k])^2
// Indices count the four-vectors the functor sees: the daughters, or for the
// calls over a container of mothers the mother (0) then the daughters.
// Any other functor type is rejected at compile time.
#pragma once

#include <complex>
#include <stdexcept>
#include <initializer_list>
#include <type_traits>
#include <vector>

#include "hpsp.h"

namespace hydra {

struct ConstantFunctor {};

struct InvariantMass2 {
  int a = 0;
  int b = 1;
};

struct InvariantMass {
  int a = 0;
  int b = 1;
};

struct CosHelicity {
  int d;
  int e;
  int first = 0;
  int count = 0;  // 0: all the daughters
};

struct DecayPlaneAngle {
  int d;
  int e;
  int h;
};

// One relativistic Breit-Wigner term c / (M^2 - s - i M G) of the pair (a, b).
struct BreitWignerTerm {
  double mass;
  double width;
  std::complex<double> coupling;
  int a;
  int b;
};

class BreitWignerSum {
 public:
  BreitWignerSum(std::initializer_list<BreitWignerTerm> terms) : terms_(terms) {}
  explicit BreitWignerSum(std::vector<BreitWignerTerm> terms) : terms_(std::move(terms)) {}
  const std::vector<BreitWignerTerm>& terms() const { return terms_; }

 private:
  std::vector<BreitWignerTerm> terms_;
};

template <typename F>
struct is_phase_space_functor
    : std::integral_constant<bool, std::is_same<F, ConstantFunctor>::value ||
                                       std::is_same<F, InvariantMass2>::value ||
                                       std::is_same<F, InvariantMass>::value ||
                                       std::is_same<F, CosHelicity>::value ||
                                       std::is_same<F, DecayPlaneAngle>::value ||
                                       std::is_same<F, BreitWignerSum>::value> {};

namespace detail {

inline hpsp_functor to_hpsp(const ConstantFunctor&) {
  hpsp_functor f{};
  f.kind = HPSP_F_ONE;
  return f;
}

inline hpsp_functor to_hpsp(const InvariantMass2& m) {
  hpsp_functor f{};
  f.kind = HPSP_F_MASS2;
  f.a[0] = m.a;
  f.b[0] = m.b;
  return f;
}

inline hpsp_functor to_hpsp(const InvariantMass& m) {
  hpsp_functor f{};
  f.kind = HPSP_F_MASS;
  f.a[0] = m.a;
  f.b[0] = m.b;
  return f;
}

inline hpsp_functor to_hpsp(const CosHelicity& h) {
  hpsp_functor f{};
  f.kind = HPSP_F_COS_HELICITY;
  f.a[0] = h.d;
  f.b[0] = h.e;
  f.a[1] = h.first;
  f.b[1] = h.count;
  return f;
}

inline hpsp_functor to_hpsp(const DecayPlaneAngle& g) {
  hpsp_functor f{};
  f.kind = HPSP_F_DECAY_PLANE_ANGLE;
  f.a[0] = g.d;
  f.b[0] = g.e;
  f.a[1] = g.h;
  return f;
}

inline hpsp_functor to_hpsp(const BreitWignerSum& s) {
  hpsp_functor f{};
  f.kind = HPSP_F_BW_SUM;
  if (s.terms().size() > HPSP_MAX_TERMS)
    throw std::invalid_argument("[hydra::BreitWignerSum]: too many terms");
  f.nterms = (int32_t)s.terms().size();
  for (size_t k = 0; k < s.terms().size(); ++k) {
    const BreitWignerTerm& t = s.terms()[k];
    f.m[k] = t.mass;
    f.g[k] = t.width;
    f.cr[k] = t.coupling.real();
    f.ci[k] = t.coupling.imag();
    f.a[k] = t.a;
    f.b[k] = t.b;
  }
  return f;
}

}  // namespace detail
}  // namespace hydra
