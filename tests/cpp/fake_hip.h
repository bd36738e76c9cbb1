// Patterns written by the model kernels of fake_hip.cpp, shared with the
// test that checks what the serving loop returns.
#pragma once

#include <stdint.h>

#include "hpsp_launch.h"

inline uint64_t fake_mix(uint64_t x) {
  x ^= x >> 31;
  x *= 0x7fb5d329728ea185ULL;
  x ^= x >> 27;
  x *= 0x81dadef4bc2dd44dULL;
  return x ^ (x >> 33);
}

// Word written for column `col` of event `evt` (col == 4*HPSP_MAX_N: the
// weight), for a request keyed `key` at precision `prec`.
inline uint64_t fake_event_word(uint64_t key, uint64_t evt, int col, int prec) {
  return fake_mix(key ^ fake_mix(evt * 64 + (uint64_t)col) ^ ((uint64_t)prec << 62));
}

// Partial statistics of the `count` events starting at `evt0`.
inline hpsp::Stats fake_partial(uint64_t key, uint64_t evt0, uint64_t count, int prec) {
  hpsp::Stats s;
  s.w = (double)count;
  s.mean = (double)(fake_event_word(key, evt0, 0, prec) >> 40);
  s.m2 = (double)(fake_event_word(key, evt0, 1, prec) >> 40);
  return s;
}

// In-order fold of partials, as a reduction kernel would produce it.
inline hpsp::Stats fake_fold(const hpsp::Stats* p, uint64_t m) {
  hpsp::Stats s{0.0, 0.0, 0.0};
  for (uint64_t j = 0; j < m; ++j) {
    s.w += p[j].w;
    s.mean = s.mean * 0.5 + p[j].mean;
    s.m2 += p[j].m2;
  }
  return s;
}

// Operations the model device has executed so far.
uint64_t fake_hip_ops_run();
