// Host-side helpers shared by the C ABI (capi.hip) and the server
// (serve.hip): validation and folding of per-decay constants.  Host
// arithmetic follows the reference's order (built with -ffp-contract=off).
#pragma once

#include <math.h>
#include <string.h>

#include "hpsp.h"
#include "hpsp_launch.h"

namespace hpsp {

inline uint64_t splitmix64_host(uint64_t* x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Engine keys as the engines' constructors derive them from a seed:
// squares3/4: splitmix64(seed) (squares3.h:81-84); Random123: successive
// splitmix64 draws, one per key word (EngineR123.h:97-112).
inline void engine_keys(int engine, uint64_t seed, uint64_t key[4]) {
  uint64_t s = seed;
  key[0] = key[1] = key[2] = key[3] = 0;
  int nkey = 1;
  if (engine == HPSP_PHILOX_LONG || engine == HPSP_THREEFRY) nkey = 2;
  if (engine == HPSP_THREEFRY_LONG) nkey = 4;
  for (int i = 0; i < nkey; ++i) key[i] = splitmix64_host(&s);
}

inline bool valid_engine(int e) { return e >= HPSP_SQUARES3 && e <= HPSP_THREEFRY_LONG; }

inline int check_decay(const hpsp_decay* d) {
  if (!d) return HPSP_ERR_INVALID_ARGUMENT;
  if (d->n < 2 || d->n > HPSP_MAX_N) return HPSP_ERR_INVALID_ARGUMENT;
  if (!valid_engine(d->engine)) return HPSP_ERR_INVALID_ARGUMENT;
  return HPSP_OK;
}

// The entry points that read only the masses and fMaxWeight of a decay
// (PhaseSpaceWeight, unweighting, fills from stored events) also take a
// phase_space_range description.
inline int check_decay_masses(const hpsp_decay* d) {
  if (!d) return HPSP_ERR_INVALID_ARGUMENT;
  if (d->n < 2 || d->n > HPSP_MAX_N) return HPSP_ERR_INVALID_ARGUMENT;
  if (!valid_engine(d->engine) && d->engine != HPSP_MINSTD_RAND) return HPSP_ERR_INVALID_ARGUMENT;
  return HPSP_OK;
}

inline DecayK make_decayk(const hpsp_decay* d, uint64_t seed) {
  DecayK k;
  memset(&k, 0, sizeof(k));
  for (int i = 0; i < HPSP_MAX_N; ++i) k.m[i] = (i < d->n) ? d->masses[i] : 0.0;
  k.ecm = d->ecm;
  k.maxw = d->max_weight;
  engine_keys(d->engine, seed, k.key);
  // Vector4R::applyBoostTo(bx,by,bz) constants (Vector4R.inl:318-337).
  const double bx = d->beta[0], by = d->beta[1], bz = d->beta[2];
  const double bxx = bx * bx, byy = by * by, bzz = bz * bz;
  const double b2 = bxx + byy + bzz;
  k.do_boost = (b2 > 0.0 && b2 < 1.0) ? 1 : 0;
  if (k.do_boost) {
    k.bx = bx;
    k.by = by;
    k.bz = bz;
    k.bxx = bxx;
    k.byy = byy;
    k.bzz = bzz;
    k.gamma = 1.0 / sqrt(1.0 - b2);
    k.gb2 = (k.gamma - 1.0) / b2;
    k.gb2xy = k.gb2 * bx * by;
    k.gb2xz = k.gb2 * bx * bz;
    k.gb2yz = k.gb2 * by * bz;
    k.gbx = k.gamma * bx;
    k.gby = k.gamma * by;
    k.gbz = k.gamma * bz;
  }
  return k;
}

// nvec: four-vectors the functor sees; first_daughter: index of the first
// decay product among them (0, or 1 when the mother comes first).
// The decay-plane angle only where the kernels compile it (allow_plane: the
// entry points over stored particles).
inline int make_functork(const hpsp_functor* f, int nvec, FunctorK* out,
                         int first_daughter = 0, bool allow_plane = false) {
  memset(out, 0, sizeof(*out));
  if (!f) {
    out->kind = HPSP_F_ONE;
    return HPSP_OK;
  }
  const int last_kind = allow_plane ? HPSP_F_DECAY_PLANE_ANGLE : HPSP_F_MASS;
  if (f->kind < HPSP_F_ONE || f->kind > last_kind) return HPSP_ERR_INVALID_ARGUMENT;
  out->kind = f->kind;
  if (f->kind == HPSP_F_MASS2 || f->kind == HPSP_F_COS_HELICITY || f->kind == HPSP_F_MASS ||
      f->kind == HPSP_F_DECAY_PLANE_ANGLE) {
    if (f->a[0] < 0 || f->a[0] >= nvec || f->b[0] < 0 || f->b[0] >= nvec)
      return HPSP_ERR_INVALID_ARGUMENT;
    out->a[0] = f->a[0];
    out->b[0] = f->b[0];
    if (f->kind == HPSP_F_COS_HELICITY) {
      if (f->b[1] > 0) {  // an explicit parent p[a1] .. p[a1 + b1 - 1]
        if (f->a[1] < 0 || f->a[1] >= nvec || f->b[1] > nvec - f->a[1])
          return HPSP_ERR_INVALID_ARGUMENT;
        out->a[1] = f->a[1];
        out->b[1] = f->b[1];
      } else {  // the sum of the daughters
        out->a[1] = first_daughter;
        out->b[1] = nvec - first_daughter;
      }
    }
    if (f->kind == HPSP_F_DECAY_PLANE_ANGLE) {
      if (f->a[1] < 0 || f->a[1] >= nvec) return HPSP_ERR_INVALID_ARGUMENT;
      out->a[1] = f->a[1];
    }
  }
  if (f->kind == HPSP_F_BW_SUM) {
    if (f->nterms < 0 || f->nterms > HPSP_MAX_TERMS) return HPSP_ERR_INVALID_ARGUMENT;
    out->nterms = f->nterms;
    for (int k = 0; k < f->nterms; ++k) {
      if (f->a[k] < 0 || f->a[k] >= nvec || f->b[k] < 0 || f->b[k] >= nvec)
        return HPSP_ERR_INVALID_ARGUMENT;
      out->m[k] = f->m[k];
      out->g[k] = f->g[k];
      out->cr[k] = f->cr[k];
      out->ci[k] = f->ci[k];
      out->a[k] = f->a[k];
      out->b[k] = f->b[k];
    }
  }
  if (f->kind == HPSP_F_MASS2_01 && nvec < 2) return HPSP_ERR_INVALID_ARGUMENT;
  return HPSP_OK;
}

inline int hip_status(int e) { return e == (int)hipSuccess ? HPSP_OK : HPSP_ERR_HIP; }

}  // namespace hpsp
