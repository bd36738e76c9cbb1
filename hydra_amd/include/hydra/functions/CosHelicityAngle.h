// hydra::CosHelicityAngle (reference: hydra/functions/CosHelicityAngle.h:54-101):
// cosine of the helicity angle of a particle with four-momentum D, daughter
// of Q, itself a daughter of P:
//   (P.D Q^2 - P.Q Q.D) / sqrt((P.Q^2 - Q^2 P^2) (Q.D^2 - Q^2 D^2)).
// Usable in device code (user functors compiled by a HIP compiler), with the
// reference's operation order.
#pragma once

#include <cmath>

#include <hydra/Types.h>
#include <hydra/Vector4R.h>

namespace hydra {

class CosHelicityAngle {
 public:
  CosHelicityAngle() = default;

  __hydra_dual__ double Evaluate(Vector4R const& P, Vector4R const& Q, Vector4R const& D) const {
    HYDRA_EXACT_FP
    const GReal_t pd = P * D;
    const GReal_t pq = P * Q;
    const GReal_t qd = Q * D;
    const GReal_t mp2 = P.mass2();
    const GReal_t mq2 = Q.mass2();
    const GReal_t md2 = D.mass2();
    return (pd * mq2 - pq * qd) / ::sqrt((pq * pq - mq2 * mp2) * (qd * qd - mq2 * md2));
  }

  __hydra_dual__ double operator()(Vector4R const& P, Vector4R const& Q, Vector4R const& D) const {
    return Evaluate(P, Q, D);
  }
};

}  // namespace hydra
