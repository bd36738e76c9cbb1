// User functors compiled into the phase-space kernels by hipcc, written as
// the reference's examples write them, beside the same computations through
// the library's described functors.  Prints one JSON object; checked by
// tests/test_user_functors.py against the described-functor results and the
// oracle.
//   * AverageOn with the cosTheta lambda of
//     examples/phase_space/phsp_averaging_functor.inl:126-143 (verbatim body)
//   * Evaluate with the three lambdas of
//     examples/phase_space/phsp_evaluating_functor.inl:150-182
//   * Unweight(functor, -1, seed) with a three-Breit-Wigner lambda
//     (examples/phys/dalitz_plot.inl:1188-1198 calls Unweight(model, -1, seed))
//   * events | dalitz_calculator and GetEventWeightFunctor(breit_wigner)
//     into a Dalitz histogram (examples/phase_space/phsp_reweighting.inl:164-198)
//   * AverageOn and Evaluate over a container of mothers
// Built twice: with -ffp-contract=off (the reference's operations one by one,
// so functor values equal the described functors' bit for bit) and with
// hipcc's default contraction (generation stays bit-exact; functor bodies may
// fuse multiply-adds).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include <hydra/Decays.h>
#include <hydra/DenseHistogram.h>
#include <hydra/Function.h>
#include <hydra/Lambda.h>
#include <hydra/PhaseSpace.h>
#include <hydra/Vector4R.h>
#include <hydra/device/System.h>
#include <hydra/functions/CosHelicityAngle.h>
#include <hydra/multivector.h>

declarg(A, hydra::Vector4R)
declarg(B, hydra::Vector4R)
declarg(C, hydra::Vector4R)
declarg(Jpsi, hydra::Vector4R)
declarg(Kaon, hydra::Vector4R)
declarg(Pion, hydra::Vector4R)
declarg(Weight, double)
declarg(M13Sq, double)
declarg(M23Sq, double)
declarg(CosTheta, double)

using namespace hydra::arguments;

namespace {

void emit(const char* key, const std::vector<double>& v, bool last = false) {
  std::printf("\"%s\": [", key);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
  std::printf("]%s\n", last ? "" : ",");
}

std::vector<double> host_column(const double* dev, size_t n) {
  std::vector<double> h(n);
  if (n) hydra::detail::hpsp_check(hpsp_memcpy_to_host(h.data(), dev, n * sizeof(double)), "copy");
  return h;
}

// The columns of a device multivector, size() values each (its columns are
// ld() apart on the device)
template <typename MV>
std::vector<double> host_all(const MV& mv) {
  const size_t n = mv.size();
  std::vector<double> h(MV::ncols * n);
  if (n)
    hydra::detail::hpsp_check(hpsp_memcpy2d_to_host(h.data(), n * sizeof(double), mv.data(),
                                                    mv.ld() * sizeof(double), n * sizeof(double),
                                                    MV::ncols),
                              "copy");
  return h;
}

// |sum_k c_k / (M_k^2 - s_k - i M_k G_k)|^2 with the operations of the
// described BreitWignerSum (exact divisions), written out in double
struct BWTerm {
  double M, G, cr, ci;
};

__hydra_dual__ inline void bw_add(const BWTerm& t, double s, double& Ar, double& Ai) {
  const double dr = t.M * t.M - s;
  const double di = -(t.M * t.G);
  const double den = dr * dr + di * di;
  const double br = dr / den;
  const double bi = -di / den;
  const double tr = t.cr * br - t.ci * bi;
  const double ti = t.cr * bi + t.ci * br;
  Ar += tr;
  Ai += ti;
}

}  // namespace

int main() {
  std::printf("{\n");

  // ---- AverageOn with the example's cosTheta lambda (B0 -> J/psi K pi)
  const double P_mass = 5.27955, A_mass = 3.0969, B_mass = 0.493677, C_mass = 0.13957061;
  {
    auto cosTheta = hydra::wrap_lambda(  [] __hydra_dual__ ( A a, B b, C c) {
      auto coshelang = hydra::CosHelicityAngle();

      return coshelang(a+b+c, b+c, c);
    });

    hydra::Vector4R Parent(P_mass, 0.0, 0.0, 0.0);
    double masses[3]{A_mass, B_mass, C_mass};
    hydra::PhaseSpace<3> phsp{P_mass, masses};
    const size_t nentries = 1000000;
    auto user = phsp.AverageOn(hydra::device::sys, Parent, cosTheta, nentries);
    auto desc = phsp.AverageOn(hydra::device::sys, Parent, hydra::CosHelicity{2, 1}, nentries);
    emit("avg_cos_user", {user.first, user.second});
    emit("avg_cos_desc", {desc.first, desc.second});
    std::printf("\"avg_cos_seed\": \"%llu\",\n", (unsigned long long)phsp.GetFullSeed());
  }

  // ---- Evaluate with the phsp_evaluating_functor.inl lambdas
  const double B0_mass = 5.27955, Jpsi_mass = 3.0969, K_mass = 0.493677, pi_mass = 0.13957061;
  {
    auto M13_Sq = hydra::wrap_lambda(
        [] __hydra_dual__ (Jpsi jpsi, Kaon kaon, Pion pion  ){
           return  ( jpsi + pion ).mass2();
        });
    auto M23_Sq = hydra::wrap_lambda(
        [] __hydra_dual__ (  Jpsi jpsi, Kaon kaon, Pion pion ){
           return  ( kaon + pion ).mass2();
        });
    auto cosTheta = hydra::wrap_lambda(
        [] __hydra_dual__ ( Jpsi jpsi, Kaon kaon, Pion pion )
    {
      hydra::Vector4R p =  jpsi+kaon+pion;
      hydra::Vector4R q = kaon+pion;

      double pd = p * kaon;
      double pq = p * q;
      double qd = q * kaon;
      double mp2 = p.mass2();
      double mq2 = q.mass2();
      double md2 = kaon().mass2();

      return (pd * mq2 - pq * qd)
          / sqrt((pq * pq - mq2 * mp2) * (qd * qd - mq2 * md2));
    });

    hydra::Vector4R B0(B0_mass, 0.0, 0.0, 0.0);
    hydra::PhaseSpace<3> phsp{B0_mass, {Jpsi_mass, K_mass, pi_mass}};
    const size_t n = 20000;
    hydra::multivector<hydra::tuple<Weight, M13Sq, M23Sq, CosTheta>, hydra::device::sys_t> u(n),
        d(n);
    phsp.Evaluate(B0, u, M13_Sq, M23_Sq, cosTheta);
    phsp.Evaluate(B0, d, hydra::InvariantMass2{0, 2}, hydra::InvariantMass2{1, 2},
                  hydra::CosHelicity{1, 2});
    emit("eval_user", host_all(u));
    emit("eval_desc", host_all(d));
    std::printf("\"eval_seed\": \"%llu\",\n\"eval_n\": %zu,\n",
                (unsigned long long)phsp.GetFullSeed(), n);

    // over mothers: the functor sees the mother first
    std::vector<hydra::Vector4R> mh;
    for (size_t k = 0; k < 3000; ++k) {
      const double px = 0.3 * std::sin(0.01 * k), py = 0.2 * std::cos(0.013 * k), pz = 0.1;
      mh.push_back(hydra::Vector4R(std::sqrt(B0_mass * B0_mass + px * px + py * py + pz * pz), px,
                                   py, pz));
    }
    hydra::device::vector<hydra::Vector4R> mothers(mh);
    auto m12 = hydra::wrap_lambda([] __hydra_dual__ (hydra::Vector4R m, hydra::Vector4R a,
                                                     hydra::Vector4R b, hydra::Vector4R c) {
      return (a + c).mass2();
    });
    hydra::multivector<hydra::tuple<Weight, M13Sq>, hydra::device::sys_t> um(mh.size()),
        dm(mh.size());
    phsp.Evaluate(mothers, um, m12);
    phsp.Evaluate(mothers, dm, hydra::InvariantMass2{1, 3});
    emit("evalm_user", host_all(um));
    emit("evalm_desc", host_all(dm));
    auto avm_u = phsp.AverageOn(mothers, m12);
    auto avm_d = phsp.AverageOn(mothers, hydra::InvariantMass2{1, 3});
    emit("avgm_user", {avm_u.first, avm_u.second});
    emit("avgm_desc", {avm_d.first, avm_d.second});
  }

  // ---- Unweight(functor) with a three-Breit-Wigner lambda, and the Dalitz
  // histograms of phsp_reweighting.inl, on the C1 decay
  {
    const double M = 5.279;
    double masses[3]{3.0969, 0.493677, 0.13957};
    hydra::PhaseSpace<3> phsp{M, masses};
    phsp.SetSeed(159753);
    hydra::Vector4R Parent(M, 0.0, 0.0, 0.0);
    const size_t n = 50000;
    const BWTerm t0{0.892, 0.05, 0.4967141530, -0.1382643012};
    const BWTerm t1{1.43, 0.11, 0.6476885381, 1.5230298564};
    const BWTerm t2{0.775, 0.149, -0.2341533747, -0.2341368981};
    auto model = hydra::wrap_lambda([t0, t1, t2] __hydra_dual__ (A a, B b, C c) {
      double Ar = 0.0, Ai = 0.0;
      bw_add(t0, (b + c).mass2(), Ar, Ai);
      bw_add(t1, (b + c).mass2(), Ar, Ai);
      bw_add(t2, (a + c).mass2(), Ar, Ai);
      return Ar * Ar + Ai * Ai;
    });
    hydra::BreitWignerSum bw{{t0.M, t0.G, {t0.cr, t0.ci}, 1, 2},
                             {t1.M, t1.G, {t1.cr, t1.ci}, 1, 2},
                             {t2.M, t2.G, {t2.cr, t2.ci}, 0, 2}};

    hydra::Decays<hydra::tuple<A, B, C>, hydra::device::sys_t> Events(M, masses, n);
    phsp.Generate(Parent, Events);
    hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> h_gen(Events);
    emit("uw_events", std::vector<double>(h_gen.data(), h_gen.data() + 12 * n));

    auto E1 = Events;
    auto E2 = Events;
    auto acc_u = E1.Unweight(model, -1.0, 0x39abdc4529b1661cULL);
    auto acc_d = E2.Unweight(bw, -1.0, 0x39abdc4529b1661cULL);
    hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> h1(E1), h2(E2);
    std::printf("\"uw_nacc\": [%zu, %zu],\n", acc_u.size(), acc_d.size());
    int same = acc_u.size() == acc_d.size() ? 1 : 0;
    if (same && std::memcmp(h1.data(), h2.data(), 12 * n * sizeof(double)) != 0) same = 0;
    std::printf("\"uw_same\": %d,\n", same);
    // a fixed maximum as well
    auto E3 = Events;
    auto E4 = Events;
    auto acc3 = E3.Unweight(model, 40.0, 99);
    auto acc4 = E4.Unweight(bw, 40.0, 99);
    hydra::Decays<hydra::tuple<A, B, C>, hydra::host::sys_t> h3(E3), h4(E4);
    int same2 = acc3.size() == acc4.size() &&
                std::memcmp(h3.data(), h4.data(), 12 * n * sizeof(double)) == 0;
    std::printf("\"uw_fixed_max\": [%zu, %d],\n", acc3.size(), same2);

    // events | lambda (two Dalitz variables), weights from a (B, C) lambda
    auto dalitz_calculator = hydra::wrap_lambda(	[] __hydra_dual__ (A a, B b, C c) {

      return hydra::make_tuple( (a + b).mass2(), (b + c).mass2());
    });
    const BWTerm k892{0.895, 0.055, 1.0, 0.0};
    auto breit_wigner = hydra::wrap_lambda(	[ k892 ] __hydra_dual__ ( B b, C c ){
      double Ar = 0.0, Ai = 0.0;
      bw_add(k892, (b + c).mass2(), Ar, Ai);
      return Ar * Ar + Ai * Ai;
    } );
    hydra::BreitWignerSum bw1{{k892.M, k892.G, {1.0, 0.0}, 1, 2}};

    auto dalitz_variables = Events | dalitz_calculator;
    auto dalitz_flat_weights = Events | Events.GetEventWeightFunctor();
    auto dalitz_resonant_weights = Events | Events.GetEventWeightFunctor(breit_wigner);
    auto resonant_desc = Events | Events.GetEventWeightFunctor(bw1);
    emit("dalitz_vars", host_all(dalitz_variables));
    emit("res_w_user", host_column(dalitz_resonant_weights.data(), n));
    emit("res_w_desc", host_column(resonant_desc.data(), n));
    const double lo0 = std::pow(masses[0] + masses[1], 2), lo1 = std::pow(masses[1] + masses[2], 2);
    const double hi0 = std::pow(M - masses[2], 2), hi1 = std::pow(M - masses[0], 2);
    auto hist_user = hydra::make_dense_histogram<double, 2>(
        hydra::device::sys, {100, 100}, {lo0, lo1}, {hi0, hi1}, dalitz_variables,
        dalitz_resonant_weights);
    auto hist_flat = hydra::make_dense_histogram<double, 2>(
        hydra::device::sys, {100, 100}, {lo0, lo1}, {hi0, hi1}, dalitz_variables,
        dalitz_flat_weights);
    auto hist_desc = hydra::make_dense_histogram<double, 2>(
        hydra::device::sys, {100, 100}, {lo0, lo1}, {hi0, hi1},
        Events | hydra::make_axes(hydra::InvariantMass2{0, 1}, hydra::InvariantMass2{1, 2}),
        Events | Events.GetEventWeightFunctor());
    emit("hist_user_res", hist_user.GetBinsContents());
    emit("hist_user_flat", hist_flat.GetBinsContents());
    emit("hist_desc_flat", hist_desc.GetBinsContents());

    // the unweighted sample through the lambda (phsp_unweighting_functor.inl:209)
    auto E5 = Events;
    auto sample = E5.Unweight(model, -1.0, 0x39abdc4529b1661cULL) | dalitz_calculator;
    emit("sample_vars", host_all(sample), true);
  }
  std::printf("}\n");
  return 0;
}
