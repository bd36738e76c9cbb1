// Four-vector of the hydra API (reference: hydra/Vector4R.h:59-130,
// hydra/detail/Vector4R.inl).  Events live on the device as SoA columns;
// this class is what rows are assembled into on the host, and, compiled by a
// HIP compiler, what user functors receive in device code.  The arithmetic
// keeps the reference's operation order (mass, mass2, applyBoostTo).
#pragma once

#include <cmath>
#include <ostream>

#include <hydra/Types.h>

namespace hydra {

class Vector3R {
 public:
  __hydra_dual__ Vector3R() : v{0, 0, 0} {}
  __hydra_dual__ Vector3R(GReal_t x, GReal_t y, GReal_t z) : v{x, y, z} {}
  __hydra_dual__ GReal_t get(GInt_t i) const { return v[i]; }
  __hydra_dual__ void set(GInt_t i, GReal_t d) { v[i] = d; }
  __hydra_dual__ void set(GReal_t x, GReal_t y, GReal_t z) {
    v[0] = x;
    v[1] = y;
    v[2] = z;
  }
  __hydra_dual__ GReal_t d3mag() const { HYDRA_EXACT_FP return ::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

 private:
  GReal_t v[3];
};

class alignas(16) Vector4R {
 public:
  Vector4R() = default;
  __hydra_dual__ Vector4R(GReal_t e, GReal_t px, GReal_t py, GReal_t pz) : v{e, px, py, pz} {}
  __hydra_dual__ Vector4R(GReal_t e, const Vector3R& p) : v{e, p.get(0), p.get(1), p.get(2)} {}

  __hydra_dual__ void set(GInt_t i, GReal_t d) { v[i] = d; }
  __hydra_dual__ void set(GReal_t e, GReal_t px, GReal_t py, GReal_t pz) {
    v[0] = e;
    v[1] = px;
    v[2] = py;
    v[3] = pz;
  }
  __hydra_dual__ GReal_t get(GInt_t i) const { return v[i]; }
  __hydra_dual__ void swap(Vector4R& o) { HYDRA_EXACT_FP
    Vector4R t(*this);
    *this = o;
    o = t;
  }

  __hydra_dual__ Vector4R& operator+=(const Vector4R& o) { HYDRA_EXACT_FP
    for (int i = 0; i < 4; ++i) v[i] += o.v[i];
    return *this;
  }
  __hydra_dual__ Vector4R& operator-=(const Vector4R& o) { HYDRA_EXACT_FP
    for (int i = 0; i < 4; ++i) v[i] -= o.v[i];
    return *this;
  }
  __hydra_dual__ Vector4R& operator*=(GReal_t c) { HYDRA_EXACT_FP
    for (int i = 0; i < 4; ++i) v[i] *= c;
    return *this;
  }
  __hydra_dual__ Vector4R& operator/=(GReal_t c) { HYDRA_EXACT_FP
    const GReal_t ci = 1.0 / c;
    for (int i = 0; i < 4; ++i) v[i] *= ci;
    return *this;
  }

  __hydra_dual__ GReal_t mass2() const { HYDRA_EXACT_FP return v[0] * v[0] - v[1] * v[1] - v[2] * v[2] - v[3] * v[3]; }
  __hydra_dual__ GReal_t mass() const { HYDRA_EXACT_FP
    const GReal_t m2 = v[0] * v[0] - v[1] * v[1] - v[2] * v[2] - v[3] * v[3];
    return (m2 > 0.0) ? ::sqrt(m2) : 0.0;
  }
  __hydra_dual__ GReal_t p2() const { HYDRA_EXACT_FP return v[1] * v[1] + v[2] * v[2] + v[3] * v[3]; }
  __hydra_dual__ GReal_t p() const { HYDRA_EXACT_FP return ::sqrt(p2()); }
  __hydra_dual__ GReal_t d3mag() const { HYDRA_EXACT_FP return ::sqrt(p2()); }
  __hydra_dual__ GReal_t dot(const Vector4R& o) const { HYDRA_EXACT_FP
    return v[0] * o.v[0] - v[1] * o.v[1] - v[2] * o.v[2] - v[3] * o.v[3];
  }
  __hydra_dual__ GReal_t cont(const Vector4R& o) const { HYDRA_EXACT_FP return dot(o); }
  __hydra_dual__ Vector3R vector3() const { return Vector3R(v[1], v[2], v[3]); }

  // Vector4R::applyBoostTo (Vector4R.inl:250-369)
  __hydra_dual__ void applyBoostTo(GReal_t bx, GReal_t by, GReal_t bz, bool inverse = false) { HYDRA_EXACT_FP
    const GReal_t bxx = bx * bx, byy = by * by, bzz = bz * bz;
    const GReal_t b2 = bxx + byy + bzz;
    if (b2 > 0.0 && b2 < 1.0) {
      const GReal_t gamma = 1.0 / ::sqrt(1.0 - b2);
      const GReal_t gb2 = (gamma - 1.0) / b2;
      const GReal_t gb2xy = gb2 * bx * by;
      const GReal_t gb2xz = gb2 * bx * bz;
      const GReal_t gb2yz = gb2 * by * bz;
      const GReal_t gbx = gamma * bx;
      const GReal_t gby = gamma * by;
      const GReal_t gbz = gamma * bz;
      const GReal_t e2 = v[0], px2 = v[1], py2 = v[2], pz2 = v[3];
      if (inverse) {
        v[0] = gamma * e2 - gbx * px2 - gby * py2 - gbz * pz2;
        v[1] = -gbx * e2 + gb2 * bxx * px2 + px2 + gb2xy * py2 + gb2xz * pz2;
        v[2] = -gby * e2 + gb2 * byy * py2 + py2 + gb2xy * px2 + gb2yz * pz2;
        v[3] = -gbz * e2 + gb2 * bzz * pz2 + pz2 + gb2yz * py2 + gb2xz * px2;
      } else {
        v[0] = gamma * e2 + gbx * px2 + gby * py2 + gbz * pz2;
        v[1] = gbx * e2 + gb2 * bxx * px2 + px2 + gb2xy * py2 + gb2xz * pz2;
        v[2] = gby * e2 + gb2 * byy * py2 + py2 + gb2xy * px2 + gb2yz * pz2;
        v[3] = gbz * e2 + gb2 * bzz * pz2 + pz2 + gb2yz * py2 + gb2xz * px2;
      }
    }
  }
  __hydra_dual__ void applyBoostTo(const Vector3R& b, bool inverse = false) { HYDRA_EXACT_FP
    applyBoostTo(b.get(0), b.get(1), b.get(2), inverse);
  }
  __hydra_dual__ void applyBoostTo(const Vector4R& p4, bool inverse = false) { HYDRA_EXACT_FP
    const GReal_t e = p4.get(0);
    applyBoostTo(p4.get(1) / e, p4.get(2) / e, p4.get(3) / e, inverse);
  }

  friend __hydra_dual__ Vector4R operator+(const Vector4R& a, const Vector4R& b) { HYDRA_EXACT_FP
    return Vector4R(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]);
  }
  friend __hydra_dual__ Vector4R operator-(const Vector4R& a, const Vector4R& b) { HYDRA_EXACT_FP
    return Vector4R(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2], a.v[3] - b.v[3]);
  }
  friend __hydra_dual__ Vector4R operator*(GReal_t d, const Vector4R& a) { HYDRA_EXACT_FP
    return Vector4R(d * a.v[0], d * a.v[1], d * a.v[2], d * a.v[3]);
  }
  friend __hydra_dual__ Vector4R operator*(const Vector4R& a, GReal_t d) { HYDRA_EXACT_FP return d * a; }
  friend __hydra_dual__ Vector4R operator/(const Vector4R& a, GReal_t d) { HYDRA_EXACT_FP
    Vector4R r(a);
    r /= d;
    return r;
  }
  friend __hydra_dual__ GReal_t operator*(const Vector4R& a, const Vector4R& b) { HYDRA_EXACT_FP return a.dot(b); }
  friend std::ostream& operator<<(std::ostream& s, const Vector4R& a) {
    return s << "(" << a.v[0] << "," << a.v[1] << "," << a.v[2] << "," << a.v[3] << ")";
  }

 private:
  GReal_t v[4];
};

}  // namespace hydra
