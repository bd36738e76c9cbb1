// Continuous-batching server for phase-space requests (declared in
// include/hpsp.h, "serving").  No counterpart in the reference, whose calls
// are synchronous and one at a time.
//
// * Admission by HBM footprint: every request's buffers come from one arena
//   (first-fit, coalescing free list).  Requests are admitted in arrival
//   order; a later request may pass one that does not fit yet, until the
//   head has waited kHeadMaxWait, after which admission holds for it.
// * Requests are split into chunks of chunk_events (a multiple of the
//   AverageOn reduction chunk), so each chunk's output -- event columns, or
//   the 4096-event partials of AverageOn -- is exactly what one direct call
//   produces for that range; a request's result is bit-identical alone and in
//   any mix.
// * Each batch takes one chunk from every active request (round robin), so
//   requests that arrive while others run join the next batch.  A batch's
//   launches are spread over n_streams default-priority streams.  Batches
//   run one after another on the device (each waits for the previous one's
//   completion event); the worker prepares the next batch while the current
//   one runs.
// * Launches are direct, not captured into HIP graphs: a batch's launch list
//   changes from one batch to the next, so a graph would be instantiated per
//   batch, which costs more host time than the launches it replaces for
//   requests of 10^3..10^7 events (a launch costs a few microseconds, a
//   request's kernel tens of microseconds to milliseconds).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hpsp.h"
#include "hpsp_host.h"
#include "hpsp_launch.h"

namespace hpsp {
namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

constexpr size_t kArenaAlign = 256;
constexpr int kMaxInflight = 2;
constexpr int kMaxLaunchesPerBatch = 256;
constexpr uint64_t kDefaultChunk = 1ULL << 22;
constexpr double kHeadMaxWait = 0.020;  // seconds
constexpr int kStatSlots = 8192;

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Bytes x rounded up to a multiple of a; false if that does not fit size_t.
bool round_up_checked(size_t x, size_t a, size_t* out) {
  size_t t;
  if (__builtin_add_overflow(x, a - 1, &t)) return false;
  *out = t / a * a;
  return true;
}

class Arena {
 public:
  int init(size_t bytes) {
    if (!round_up_checked(bytes, kArenaAlign, &size_)) return HPSP_ERR_INVALID_ARGUMENT;
    void* p = nullptr;
    if (hipMalloc(&p, size_) != hipSuccess) return HPSP_ERR_HIP;
    base_ = static_cast<char*>(p);
    free_[0] = size_;
    return HPSP_OK;
  }
  void destroy() {
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = round_up(bytes ? bytes : 1, kArenaAlign);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t roff = it->first + bytes;
        free_.erase(it);
        if (rest) free_[roff] = rest;
        in_use_ += bytes;
        peak_ = std::max(peak_, in_use_);
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = round_up(bytes ? bytes : 1, kArenaAlign);
    in_use_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  char* base() const { return base_; }
  size_t size() const { return size_; }
  size_t in_use() const { return in_use_; }
  size_t peak() const { return peak_; }

 private:
  char* base_ = nullptr;
  size_t size_ = 0, in_use_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

struct Request {
  uint64_t ticket = 0;
  hpsp_request r{};
  DecayK dk{};
  FunctorK fk{};
  size_t bytes = 0, off = 0;
  uint64_t next = 0;        // next event offset to launch
  bool launched_all = false;
  int stream = 0;
  int slot = -1;            // pinned stats slot (AVERAGE)
  double t_submit = 0, t_admit = 0;
  void* cols = nullptr;
  uint64_t ld = 0;          // column stride of cols (GENERATE)
  double* weights = nullptr;
  Stats* partials = nullptr;
  hpsp_stats* out_dev = nullptr;
  int status = HPSP_OK;
};

struct Item {
  Request* q;
  uint64_t off, len;
  bool last;
};

struct Batch {
  std::vector<Item> items;
  hipEvent_t done = nullptr;
};

// Column stride of a GENERATE request's output: n rounded up to a multiple
// of 64, so that every column starts on a cache line and each group of 64
// events fills whole lines (a stride that leaves lines partial makes HBM
// complete them by read-modify-write: Generate about 1.8x slower).  False
// when that does not fit.
bool col_stride(uint64_t n, uint64_t* ld) {
  size_t t;
  if (!round_up_checked((size_t)(n ? n : 1), 64, &t)) return false;
  *ld = t;
  return true;
}

// Arena bytes of a request: its columns (+ weights) for GENERATE, its
// partials and result for AVERAGE.  False when the size does not fit size_t
// (a huge or wrapped-around event count), which the caller rejects.
bool footprint(const hpsp_request& r, size_t* out) {
  if (r.kind == HPSP_REQ_GENERATE) {
    const size_t esz = r.out_f32 ? 4 : 8;
    size_t cols, b;
    uint64_t ld;
    if (!col_stride(r.n, &ld)) return false;
    if (__builtin_mul_overflow((size_t)ld, 4 * (size_t)r.decay.n * esz, &cols)) return false;
    if (!round_up_checked(cols, kArenaAlign, &b)) return false;
    if (r.want_weights) {
      size_t w, wr;
      if (__builtin_mul_overflow((size_t)r.n, (size_t)8, &w)) return false;
      if (!round_up_checked(w, kArenaAlign, &wr)) return false;
      if (__builtin_add_overflow(b, wr, &b)) return false;
    }
    *out = b;
    return true;
  }
  const size_t m = r.n / kAvgChunk + (r.n % kAvgChunk ? 1 : 0);
  size_t p, b;
  if (__builtin_mul_overflow(m ? m : (size_t)1, sizeof(Stats), &p)) return false;
  if (!round_up_checked(p, kArenaAlign, &b)) return false;
  if (__builtin_add_overflow(b, kArenaAlign, &b)) return false;
  *out = b;
  return true;
}

}  // namespace
}  // namespace hpsp

using namespace hpsp;

struct hpsp_server {
  Arena arena;
  std::vector<hipStream_t> streams;
  hipStream_t main = nullptr;
  hipEvent_t last_done = nullptr;  // completion of the last launched batch
  std::vector<hipEvent_t> joins;
  hpsp_stats* pinned = nullptr;
  std::vector<int> free_slots;
  uint64_t chunk = kDefaultChunk;
  int device = 0;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::unique_ptr<Request>> pending;
  std::vector<std::unique_ptr<Request>> active;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> held;  // completed GENERATE outputs
  std::deque<hpsp_completion> done;
  std::deque<Batch> inflight;
  size_t rr = 0;
  uint64_t next_ticket = 1;
  bool stop = false;
  hpsp_server_info info{};
  std::thread worker;

  // Streams, events and pinned result slots; on failure the caller releases
  // whatever was created.
  int create_resources(int n_streams) {
    streams.assign(n_streams, nullptr);
    for (auto& st : streams)
      if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return HPSP_ERR_HIP;
    if (hipStreamCreateWithFlags(&main, hipStreamNonBlocking) != hipSuccess) return HPSP_ERR_HIP;
    joins.assign(n_streams, nullptr);
    for (auto& e : joins)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return HPSP_ERR_HIP;
    if (hipEventCreateWithFlags(&last_done, hipEventDisableTiming) != hipSuccess)
      return HPSP_ERR_HIP;
    if (hipEventRecord(last_done, main) != hipSuccess) return HPSP_ERR_HIP;
    void* p = nullptr;
    if (hipHostMalloc(&p, kStatSlots * sizeof(hpsp_stats), hipHostMallocDefault) != hipSuccess)
      return HPSP_ERR_HIP;
    pinned = static_cast<hpsp_stats*>(p);
    return HPSP_OK;
  }

  void release_resources() {
    for (auto& e : joins)
      if (e) (void)hipEventDestroy(e);
    joins.clear();
    if (last_done) (void)hipEventDestroy(last_done);
    last_done = nullptr;
    for (auto& st : streams)
      if (st) (void)hipStreamDestroy(st);
    streams.clear();
    if (main) (void)hipStreamDestroy(main);
    main = nullptr;
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    arena.destroy();
  }

  // ---- worker side ------------------------------------------------------
  void admit_locked() {
    const double t = now_s();
    for (auto it = pending.begin(); it != pending.end();) {
      Request* q = it->get();
      bool ok = true;
      int slot = -1;
      if (q->r.kind == HPSP_REQ_AVERAGE) {
        if (free_slots.empty()) ok = false;
        else slot = free_slots.back();
      }
      size_t off = 0;
      if (ok && !arena.alloc(q->bytes, &off)) ok = false;
      if (!ok) {
        // the oldest request has waited long enough: nobody passes it
        if (it == pending.begin() && t - q->t_submit > kHeadMaxWait) break;
        ++it;
        continue;
      }
      if (slot >= 0) free_slots.pop_back();
      q->slot = slot;
      q->off = off;
      q->t_admit = t;
      char* base = arena.base() + off;
      if (q->r.kind == HPSP_REQ_GENERATE) {
        const size_t esz = q->r.out_f32 ? 4 : 8;
        q->cols = base;
        col_stride(q->r.n, &q->ld);  // fits: footprint checked it
        if (q->r.want_weights)
          q->weights = reinterpret_cast<double*>(
              base + round_up(q->ld * 4 * (size_t)q->r.decay.n * esz, kArenaAlign));
      } else {
        q->out_dev = reinterpret_cast<hpsp_stats*>(base);
        q->partials = reinterpret_cast<Stats*>(base + kArenaAlign);
      }
      q->stream = (int)(q->ticket % streams.size());
      ++info.admitted;
      active.push_back(std::move(*it));
      it = pending.erase(it);
    }
  }

  Batch build_batch_locked() {
    Batch b;
    const size_t na = active.size();
    for (size_t k = 0; k < na && (int)b.items.size() < kMaxLaunchesPerBatch; ++k) {
      Request* q = active[(rr + k) % na].get();
      if (q->launched_all) continue;
      const uint64_t len = std::min<uint64_t>(chunk, q->r.n - q->next);
      Item it{q, q->next, len, false};
      q->next += len;
      if (q->next >= q->r.n) {
        q->launched_all = true;
        it.last = true;
      }
      b.items.push_back(it);
    }
    if (na) rr = (rr + 1) % na;
    return b;
  }

  // Launches one chunk (and, after a request's last chunk, the AverageOn
  // finalisation and the copy of its result); returns the hipError_t value
  // and adds the number of kernel launches to *nl.
  int launch_item(const Item& it, hipStream_t s, uint64_t* nl) {
    Request* q = it.q;
    const int N = q->r.decay.n, E = q->r.decay.engine;
    int e = 0;
    if (it.len > 0) {
      if (q->r.kind == HPSP_REQ_GENERATE) {
        GenArgs a;
        memset(&a, 0, sizeof(a));
        a.d = q->dk;
        a.first = q->r.first + it.off;
        a.n = it.len;
        const size_t esz = q->r.out_f32 ? 4 : 8;
        a.cols = static_cast<char*>(q->cols) + it.off * esz;
        a.ld = q->ld;
        a.out_f32 = q->r.out_f32;
        a.evw = q->weights ? q->weights + it.off : nullptr;
        // chunks of several streams share the chip: one group per wave (runs
        // lengthen the waves, and with them the tails of concurrent chunks;
        // 9% less throughput at saturation)
        a.run_cap = 1;
        e = (q->r.prec == HPSP_FP64) ? launch_generate_f64(N, E, a, s)
                                     : launch_generate_mixed(N, E, a, s);
      } else {
        AvgArgs a;
        memset(&a, 0, sizeof(a));
        a.d = q->dk;
        a.f = q->fk;
        a.first = q->r.first + it.off;
        a.n = it.len;
        a.partials = q->partials + it.off / kAvgChunk;
        e = (q->r.prec == HPSP_FP64) ? launch_average_f64(N, E, a, s)
                                     : launch_average_mixed(N, E, a, s);
      }
      ++*nl;
    }
    if (!e && it.last && q->r.kind == HPSP_REQ_AVERAGE) {
      const uint64_t m = (q->r.n + kAvgChunk - 1) / kAvgChunk;
      e = launch_stats_finalize(q->partials, m, reinterpret_cast<Stats*>(q->out_dev), s);
      if (!e)
        e = (int)hipMemcpyAsync(pinned + q->slot, q->out_dev, sizeof(hpsp_stats),
                                hipMemcpyDeviceToHost, s);
      ++*nl;
    }
    return e;
  }

  // Launches a batch on the server's streams; returns the number of kernel
  // launches.
  uint64_t launch_batch(Batch& b) {
    uint64_t nl = 0;
    (void)hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
    const int K = (int)streams.size();
    // batches run in order: this one starts when the previous one is done
    for (int k = 0; k < K; ++k) (void)hipStreamWaitEvent(streams[k], last_done, 0);
    for (const Item& it : b.items) {
      const int e = launch_item(it, streams[it.q->stream], &nl);
      if (e) it.q->status = HPSP_ERR_HIP;
    }
    for (int k = 0; k < K; ++k) {
      (void)hipEventRecord(joins[k], streams[k]);
      (void)hipStreamWaitEvent(main, joins[k], 0);
    }
    (void)hipEventRecord(b.done, main);
    (void)hipEventRecord(last_done, main);
    return nl;
  }

  void retire_locked(Batch& b) {
    const double t = now_s();
    for (const Item& it : b.items) {
      if (!it.last) continue;
      Request* q = it.q;
      hpsp_completion c;
      memset(&c, 0, sizeof(c));
      c.ticket = q->ticket;
      c.user_tag = q->r.user_tag;
      c.status = q->status;
      c.kind = q->r.kind;
      c.n = q->r.n;
      c.submit_s = q->t_submit;
      c.admit_s = q->t_admit;
      c.done_s = t;
      auto pos = std::find_if(active.begin(), active.end(),
                              [q](const std::unique_ptr<Request>& p) { return p.get() == q; });
      std::unique_ptr<Request> owned = std::move(*pos);
      active.erase(pos);
      if (q->r.kind == HPSP_REQ_AVERAGE) {
        c.stats = pinned[q->slot];
        free_slots.push_back(q->slot);
        arena.release(q->off, q->bytes);
      } else {
        c.cols = q->cols;
        c.weights = q->weights;
        c.ld = q->ld;
        held[q->ticket] = std::move(owned);
      }
      done.push_back(c);
      ++info.completed;
    }
    (void)hipEventDestroy(b.done);
    if (!active.empty()) rr %= active.size();
    else rr = 0;
  }

  void run() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      while (!inflight.empty() && hipEventQuery(inflight.front().done) == hipSuccess) {
        retire_locked(inflight.front());
        inflight.pop_front();
        cv_done.notify_all();
      }
      admit_locked();
      if ((int)inflight.size() < kMaxInflight) {
        Batch b = build_batch_locked();
        if (!b.items.empty()) {
          lk.unlock();
          const uint64_t nl = launch_batch(b);
          lk.lock();
          ++info.batches;
          info.launches += nl;
          inflight.push_back(std::move(b));
          continue;
        }
      }
      if (stop && !pending.empty()) {
        // shutting down: requests not yet admitted are answered, not run
        const double t = now_s();
        for (auto& q : pending) {
          hpsp_completion c;
          memset(&c, 0, sizeof(c));
          c.ticket = q->ticket;
          c.user_tag = q->r.user_tag;
          c.status = HPSP_ERR_SHUTDOWN;
          c.kind = q->r.kind;
          c.n = q->r.n;
          c.submit_s = q->t_submit;
          c.done_s = t;
          done.push_back(c);
        }
        pending.clear();
        cv_done.notify_all();
      }
      if (stop && active.empty() && inflight.empty()) break;
      // in-flight batches are polled; new submissions wake the worker
      cv_work.wait_for(lk, std::chrono::microseconds(inflight.empty() ? 2000 : 20));
    }
  }
};

extern "C" {

int hpsp_server_create(hpsp_server** out, size_t hbm_bytes, int n_streams,
                       uint64_t chunk_events) {
  if (!out || n_streams < 1 || n_streams > 16 || hbm_bytes == 0)
    return HPSP_ERR_INVALID_ARGUMENT;
  auto s = std::make_unique<hpsp_server>();
  if (hipGetDevice(&s->device) != hipSuccess) return HPSP_ERR_NO_DEVICE;
  if (chunk_events) {
    size_t c = 0;
    if (!round_up_checked(chunk_events, kAvgChunk, &c) || c == 0)
      return HPSP_ERR_INVALID_ARGUMENT;
    s->chunk = c;
  }
  int rc = s->arena.init(hbm_bytes);
  if (rc) return rc;
  rc = s->create_resources(n_streams);
  if (rc) {
    s->release_resources();
    return rc;
  }
  for (int i = kStatSlots - 1; i >= 0; --i) s->free_slots.push_back(i);
  s->info.arena_bytes = s->arena.size();
  hpsp_server* raw = s.release();
  raw->worker = std::thread([raw] { raw->run(); });
  *out = raw;
  return HPSP_OK;
}

int hpsp_server_submit(hpsp_server* s, const hpsp_request* r, uint64_t* ticket) {
  if (!s || !r || !ticket) return HPSP_ERR_INVALID_ARGUMENT;
  int rc = check_decay(&r->decay);
  if (rc) return rc;
  if (r->kind != HPSP_REQ_GENERATE && r->kind != HPSP_REQ_AVERAGE) return HPSP_ERR_INVALID_ARGUMENT;
  if (r->prec != HPSP_FP64 && r->prec != HPSP_MIXED) return HPSP_ERR_INVALID_ARGUMENT;
  auto q = std::make_unique<Request>();
  q->r = *r;
  if (r->kind == HPSP_REQ_AVERAGE) {
    rc = make_functork(&r->functor, r->decay.n, &q->fk);
    if (rc) return rc;
  }
  q->dk = make_decayk(&r->decay, r->seed);
  if (!footprint(*r, &q->bytes) || q->bytes > s->arena.size()) {
    std::lock_guard<std::mutex> g(s->mu);
    ++s->info.rejected;
    return HPSP_ERR_WORKSPACE;
  }
  q->t_submit = now_s();
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) return HPSP_ERR_SHUTDOWN;
    q->ticket = s->next_ticket++;
    *ticket = q->ticket;
    s->pending.push_back(std::move(q));
    ++s->info.submitted;
  }
  s->cv_work.notify_one();
  return HPSP_OK;
}

int hpsp_server_poll(hpsp_server* s, hpsp_completion* out, int max, int* n_out, int wait_ms) {
  if (!s || !n_out || (max > 0 && !out)) return HPSP_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->done.empty() && wait_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [s] { return !s->done.empty(); });
  int k = 0;
  while (k < max && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *n_out = k;
  return HPSP_OK;
}

int hpsp_server_release(hpsp_server* s, uint64_t ticket) {
  if (!s) return HPSP_ERR_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->held.find(ticket);
    if (it == s->held.end()) return HPSP_ERR_INVALID_ARGUMENT;
    s->arena.release(it->second->off, it->second->bytes);
    s->held.erase(it);
  }
  s->cv_work.notify_one();
  return HPSP_OK;
}

int hpsp_server_info_get(hpsp_server* s, hpsp_server_info* info) {
  if (!s || !info) return HPSP_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> g(s->mu);
  *info = s->info;
  info->arena_in_use = s->arena.in_use();
  info->arena_peak = s->arena.peak();
  return HPSP_OK;
}

int hpsp_server_destroy(hpsp_server* s) {
  if (!s) return HPSP_ERR_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipDeviceSynchronize();
  s->release_resources();
  delete s;
  return HPSP_OK;
}

}  // extern "C"
