// Host-logic test of the serving loop (hydra_amd/csrc/serve.hip), linked
// against the HIP model of fake_hip.cpp instead of the HIP runtime, built
// with the address/undefined-behaviour sanitizers or the thread sanitizer.
// It drives the server through the C ABI the way a client does -- paced
// arrivals, a saturating burst, admission by footprint, shutdown with work
// pending -- and checks every returned result against what the model kernels
// write for that request.  Prints one line per scenario; exits non-zero on
// the first mismatch.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <map>
#include <random>
#include <thread>
#include <vector>

#include "fake_hip.h"
#include "hpsp.h"
#include "hpsp_host.h"

namespace {

struct Case {
  hpsp_request r;
  int N;
  uint64_t key;
};

#define CHECK(cond, ...)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);                       \
      fprintf(stderr, "\n");                              \
      exit(1);                                            \
    }                                                     \
  } while (0)

Case make_case(std::mt19937_64& rng, uint64_t nmax, uint64_t tag) {
  std::uniform_real_distribution<double> U(0.0, 1.0);
  Case c;
  memset(&c.r, 0, sizeof(c.r));
  c.N = 2 + (int)(rng() % 11);
  hpsp_decay& d = c.r.decay;
  d.n = c.N;
  d.engine = (int)(rng() % 6);
  double msum = 0;
  for (int i = 0; i < c.N; ++i) {
    d.masses[i] = 0.1 + 0.9 * U(rng);
    msum += d.masses[i];
  }
  d.mother_mass = msum * (1.05 + 2.0 * U(rng));
  d.ecm = d.mother_mass - msum;
  d.max_weight = 1.0;
  c.r.kind = U(rng) < 0.3 ? HPSP_REQ_AVERAGE : HPSP_REQ_GENERATE;
  c.r.prec = U(rng) < 0.5 ? HPSP_MIXED : HPSP_FP64;
  c.r.out_f32 = c.r.prec == HPSP_MIXED;
  c.r.want_weights = U(rng) < 0.5;
  c.r.seed = rng();
  c.r.first = rng() % (1ULL << 33);
  c.r.n = (uint64_t)exp(U(rng) * log((double)nmax));
  if (tag % 97 == 3) c.r.n = 0;
  c.r.user_tag = tag;
  c.r.functor.kind = HPSP_F_ONE;
  uint64_t key[4];
  hpsp::engine_keys(d.engine, c.r.seed, key);
  c.key = key[0];
  return c;
}

void verify(const Case& c, const hpsp_completion& got) {
  const hpsp_request& r = c.r;
  CHECK(got.status == HPSP_OK, "tag %llu status %d", (unsigned long long)r.user_tag, got.status);
  CHECK(got.n == r.n && got.kind == r.kind, "tag %llu header", (unsigned long long)r.user_tag);
  CHECK(r.kind != HPSP_REQ_GENERATE || (got.ld >= r.n && got.ld % 64 == 0 && got.ld > 0),
        "tag %llu column stride %llu", (unsigned long long)r.user_tag,
        (unsigned long long)got.ld);
  CHECK(got.submit_s <= got.admit_s && got.admit_s <= got.done_s, "tag %llu times",
        (unsigned long long)r.user_tag);
  if (r.kind == HPSP_REQ_AVERAGE) {
    const uint64_t m = (r.n + hpsp::kAvgChunk - 1) / hpsp::kAvgChunk;
    std::vector<hpsp::Stats> p(m);
    for (uint64_t j = 0; j < m; ++j)
      p[j] = fake_partial(c.key, r.first + j * hpsp::kAvgChunk,
                          std::min<uint64_t>(hpsp::kAvgChunk, r.n - j * hpsp::kAvgChunk), r.prec);
    const hpsp::Stats want = fake_fold(p.data(), m);
    CHECK(memcmp(&want, &got.stats, sizeof(want)) == 0, "tag %llu stats",
          (unsigned long long)r.user_tag);
    return;
  }
  for (uint64_t i = 0; i < r.n; ++i) {
    const uint64_t evt = r.first + i;
    for (int col = 0; col < 4 * c.N; ++col) {
      const uint64_t v = fake_event_word(c.key, evt, col, r.prec);
      if (r.out_f32) {
        const uint32_t g = static_cast<const uint32_t*>(got.cols)[col * got.ld + i];
        CHECK(g == (uint32_t)v, "tag %llu event %llu col %d", (unsigned long long)r.user_tag,
              (unsigned long long)i, col);
      } else {
        const uint64_t g = static_cast<const uint64_t*>(got.cols)[col * got.ld + i];
        CHECK(g == v, "tag %llu event %llu col %d", (unsigned long long)r.user_tag,
              (unsigned long long)i, col);
      }
    }
    if (r.want_weights) {
      uint64_t g;
      memcpy(&g, &got.weights[i], 8);
      CHECK(g == fake_event_word(c.key, evt, 4 * HPSP_MAX_N, r.prec), "tag %llu weight %llu",
            (unsigned long long)r.user_tag, (unsigned long long)i);
    } else {
      CHECK(got.weights == nullptr, "tag %llu weights pointer", (unsigned long long)r.user_tag);
    }
  }
}

struct Run {
  uint64_t completed = 0, batches = 0, launches = 0;
};

// Submits `cases` (paced by `gap_us` between submissions; 0 = all at once),
// polls, checks and releases every result; returns the server's counters.
Run drive(hpsp_server* s, const std::vector<Case>& cases, int gap_us) {
  std::map<uint64_t, size_t> by_ticket;
  size_t next = 0, done = 0;
  std::vector<hpsp_completion> buf(64);
  const auto t_end = std::chrono::steady_clock::now() + std::chrono::seconds(240);
  while (done < cases.size()) {
    CHECK(std::chrono::steady_clock::now() < t_end, "no progress: %zu of %zu done", done,
          cases.size());
    if (next < cases.size()) {
      const size_t burst = gap_us ? 1 : cases.size();
      for (size_t k = 0; k < burst && next < cases.size(); ++k, ++next) {
        uint64_t t = 0;
        const int rc = hpsp_server_submit(s, &cases[next].r, &t);
        CHECK(rc == HPSP_OK, "submit %zu rc %d", next, rc);
        by_ticket[t] = next;
      }
    }
    int n = 0;
    CHECK(hpsp_server_poll(s, buf.data(), (int)buf.size(), &n,
                           next < cases.size() ? 0 : 5) == HPSP_OK, "poll");
    for (int k = 0; k < n; ++k) {
      const hpsp_completion& c = buf[k];
      auto it = by_ticket.find(c.ticket);
      CHECK(it != by_ticket.end(), "unknown ticket %llu", (unsigned long long)c.ticket);
      const Case& cs = cases[it->second];
      CHECK(c.user_tag == cs.r.user_tag, "tag mismatch");
      verify(cs, c);
      if (c.kind == HPSP_REQ_GENERATE) {
        CHECK(hpsp_server_release(s, c.ticket) == HPSP_OK, "release");
        CHECK(hpsp_server_release(s, c.ticket) == HPSP_ERR_INVALID_ARGUMENT, "double release");
      }
      by_ticket.erase(it);
      ++done;
    }
    if (gap_us && next < cases.size())
      std::this_thread::sleep_for(std::chrono::microseconds(gap_us));
  }
  hpsp_server_info info;
  CHECK(hpsp_server_info_get(s, &info) == HPSP_OK, "info");
  CHECK(info.arena_in_use == 0, "arena not empty after all releases: %llu",
        (unsigned long long)info.arena_in_use);
  CHECK(info.arena_peak <= info.arena_bytes, "arena peak");
  Run r;
  r.completed = info.completed;
  r.batches = info.batches;
  r.launches = info.launches;
  return r;
}

std::vector<Case> mix(uint64_t seed, size_t count, uint64_t nmax, uint64_t tag0) {
  std::mt19937_64 rng(seed);
  std::vector<Case> v;
  for (size_t i = 0; i < count; ++i) v.push_back(make_case(rng, nmax, tag0 + i));
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  const double scale = argc > 1 ? atof(argv[1]) : 1.0;
  const size_t nreq = (size_t)(600 * scale);
  const uint64_t chunk = 16384;

  // 1. warm-up then paced arrivals; 2. a second server, one burst
  for (int pass = 0; pass < 2; ++pass) {
    hpsp_server* s = nullptr;
    CHECK(hpsp_server_create(&s, 96u << 20, 4, chunk) == HPSP_OK, "create");
    std::vector<Case> warm = mix(100 + pass, 100, 2, 0);
    for (auto& c : warm) c.r.n = 4096;
    drive(s, warm, 0);
    const std::vector<Case> cases = mix(7 + pass, nreq, 200000, 1000);
    const Run r = drive(s, cases, pass == 0 ? 100 : 0);
    CHECK(hpsp_server_destroy(s) == HPSP_OK, "destroy");
    printf("%s: %zu requests ok, %llu batches, %llu launches\n",
           pass == 0 ? "paced" : "burst", cases.size(), (unsigned long long)r.batches,
           (unsigned long long)r.launches);
    CHECK(r.completed == warm.size() + cases.size(), "completed count");
    CHECK(r.launches > r.batches, "chunks of different requests should share batches");
  }

  // 3. admission by footprint: a request larger than the arena is refused;
  //    a stream of requests larger in total than the arena completes.
  {
    hpsp_server* s = nullptr;
    const size_t arena = 8u << 20;
    CHECK(hpsp_server_create(&s, arena, 2, 4096) == HPSP_OK, "create");
    std::vector<Case> cases = mix(31, 40, 2, 5000);
    for (auto& c : cases) {
      c.r.kind = HPSP_REQ_GENERATE;
      c.r.n = 20000;
      c.r.decay.n = c.N = 3;
      c.r.prec = HPSP_FP64;
      c.r.out_f32 = 0;
      uint64_t key[4];
      hpsp::engine_keys(c.r.decay.engine, c.r.seed, key);
      c.key = key[0];
    }
    Case big = cases[0];
    big.r.n = arena / 96 + 1000;
    uint64_t t = 0;
    CHECK(hpsp_server_submit(s, &big.r, &t) == HPSP_ERR_WORKSPACE, "oversized request accepted");
    drive(s, cases, 0);
    hpsp_server_info info;
    CHECK(hpsp_server_info_get(s, &info) == HPSP_OK, "info");
    CHECK(info.rejected == 1 && info.arena_peak <= arena, "admission counters");
    CHECK(hpsp_server_destroy(s) == HPSP_OK, "destroy");
    printf("admission: %zu requests through a %zu-byte arena ok\n", cases.size(), arena);
  }

  // 3b. sizes that do not fit size_t are refused, never admitted: an event
  //     count whose footprint wraps around (a negative count cast to
  //     uint64), and chunk sizes that round up past 2^64
  {
    hpsp_server* s = nullptr;
    CHECK(hpsp_server_create(&s, 8u << 20, 2, 4096) == HPSP_OK, "create");
    std::vector<Case> cases = mix(53, 4, 2, 7000);
    const uint64_t wraps[] = {(uint64_t)-1, (uint64_t)-96, 1ULL << 61, (1ULL << 64) / 96 + 1,
                              ((uint64_t)-1) / 8 + 1};
    for (int kind = 0; kind < 2; ++kind) {
      for (uint64_t n : wraps) {
        Case c = cases[0];
        c.r.kind = kind ? HPSP_REQ_AVERAGE : HPSP_REQ_GENERATE;
        c.r.want_weights = 1;
        c.r.n = n;
        uint64_t t = 0;
        const int rc = hpsp_server_submit(s, &c.r, &t);
        CHECK(rc == HPSP_ERR_WORKSPACE, "kind %d n %llu admitted (rc %d)", kind,
              (unsigned long long)n, rc);
      }
    }
    drive(s, cases, 0);  // the server still works afterwards
    CHECK(hpsp_server_destroy(s) == HPSP_OK, "destroy");
    hpsp_server* bad = nullptr;
    CHECK(hpsp_server_create(&bad, 8u << 20, 2, (uint64_t)-1) == HPSP_ERR_INVALID_ARGUMENT &&
              bad == nullptr,
          "chunk size rounding to 0 accepted");
    CHECK(hpsp_server_create(&bad, 8u << 20, 2, (uint64_t)-4095) == HPSP_ERR_INVALID_ARGUMENT,
          "chunk size rounding past 2^64 accepted");
    CHECK(hpsp_server_create(&bad, (size_t)-1, 2, 4096) == HPSP_ERR_INVALID_ARGUMENT,
          "arena size rounding past 2^64 accepted");
    printf("oversized and wrapping sizes refused ok\n");
  }

  // 4. shutdown with requests pending, admitted and held: destroy returns,
  //    frees everything (the leak checker sees the rest)
  {
    hpsp_server* s = nullptr;
    CHECK(hpsp_server_create(&s, 8u << 20, 2, 4096) == HPSP_OK, "create");
    std::vector<Case> cases = mix(47, 60, 100000, 9000);
    for (auto& c : cases) {
      uint64_t t = 0;
      const int rc = hpsp_server_submit(s, &c.r, &t);
      CHECK(rc == HPSP_OK || rc == HPSP_ERR_WORKSPACE, "submit rc %d", rc);
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    CHECK(hpsp_server_destroy(s) == HPSP_OK, "destroy");
    printf("shutdown with work pending ok\n");
  }
  printf("model device ran %llu operations\n", (unsigned long long)fake_hip_ops_run());
  return 0;
}
