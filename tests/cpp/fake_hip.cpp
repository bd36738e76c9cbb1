// Host model of the HIP runtime calls made by the serving loop
// (hydra_amd/csrc/serve.hip) and of the kernel launchers it calls, so that
// its scheduling logic runs on the CPU under the address, undefined-behaviour
// and thread sanitizers.  Not a GPU emulator: "device" memory is host memory,
// a kernel is a host function that writes a pattern fixed by its arguments.
//
// What the model enforces (a violation prints the reason and aborts):
// * streams execute in order on one model device thread, asynchronously to
//   the host and in a randomised interleaving across streams;
// * events complete when the device reaches their record; a wait holds its
//   stream until the event's record as of the wait call has completed;
//   using a destroyed event or stream is an error;
// * hipFree, hipHostFree and hipStreamDestroy wait for the work they could
//   still be touching, as the HIP runtime does;
// * stream capture: operations on a capturing stream become nodes of the
//   capture's graph instead of running; a stream joins a capture by waiting
//   on an event recorded in it and must rejoin the origin stream (the origin
//   waits on an event recorded after its last operation) before
//   hipStreamEndCapture, else hipErrorStreamCaptureUnjoined; capturing from
//   another thread, a capturing stream waiting on an event from outside the
//   capture, waiting on or querying an event last recorded in a capture, and
//   destroying a graph exec with a launch pending are errors; hipGraphLaunch
//   runs the nodes in capture order as one stream operation.  (The serving
//   loop launches directly; the model of capture replays the graph-capturing
//   loop of the project's history, see the Makefile's graph_replay.)
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <map>
#include <random>
#include <thread>
#include <vector>

#include "fake_hip.h"
#include "hpsp_launch.h"

namespace {

[[noreturn]] void model_fail(const char* what) {
  fprintf(stderr, "fake HIP model: %s\n", what);
  fflush(stderr);
  abort();
}

struct Op {
  enum Kind { FN, WAIT, RECORD } kind = FN;
  std::function<void()> fn;
  ihipEvent_t* ev = nullptr;
  uint64_t gen = 0;
};

}  // namespace

struct Capture;

struct ihipEvent_t {
  uint64_t recorded = 0, completed = 0;
  bool destroyed = false;
  Capture* cap = nullptr;  // recorded inside this capture (a capture event)
  uint64_t cap_pos = 0;    // position of that record in the capture
};

struct ihipStream_t {
  std::deque<Op> q;
  bool destroyed = false;
  Capture* cap = nullptr;  // capturing into this capture
  uint64_t last_pos = 0;   // position of this stream's last captured operation
};

struct ihipGraph {
  std::vector<std::function<void()>> nodes;
  bool destroyed = false;
};

struct hipGraphExec {
  std::vector<std::function<void()>> nodes;
  bool destroyed = false;
  uint64_t pending = 0;
};

struct Capture {
  ihipStream_t* origin = nullptr;
  ihipGraph* g = nullptr;
  std::vector<ihipStream_t*> joined;  // besides the origin
  std::map<ihipStream_t*, uint64_t> joined_back;  // stream -> position of its rejoin
  uint64_t pos = 0;
  std::thread::id tid;
  bool invalid = false;
  bool ended = false;
};

namespace {

struct Device {
  std::mutex mu;
  std::condition_variable cv_host, cv_dev;
  std::vector<ihipStream_t*> streams;
  std::thread th;
  bool quit = false;
  std::mt19937_64 rng{12345};
  std::atomic<uint64_t> ops_run{0};

  Device() { th = std::thread([this] { loop(); }); }
  ~Device() {
    {
      std::lock_guard<std::mutex> g(mu);
      quit = true;
    }
    cv_dev.notify_all();
    th.join();
  }

  // One runnable head operation is taken at a time, from a stream chosen at
  // random among those whose head can run.
  void loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      std::vector<ihipStream_t*> ready;
      for (ihipStream_t* s : streams) {
        if (s->q.empty()) continue;
        const Op& h = s->q.front();
        if (h.kind == Op::WAIT && h.ev->completed < h.gen) continue;
        ready.push_back(s);
      }
      if (ready.empty()) {
        if (quit) return;
        cv_dev.wait_for(lk, std::chrono::microseconds(200));
        continue;
      }
      ihipStream_t* s = ready[rng() % ready.size()];
      Op op = std::move(s->q.front());
      s->q.pop_front();
      switch (op.kind) {
        case Op::WAIT:
          break;
        case Op::RECORD:
          if (op.ev->completed < op.gen) op.ev->completed = op.gen;
          break;
        case Op::FN:
          op.fn();
          break;
      }
      ++ops_run;
      cv_host.notify_all();
      if (rng() % 4 == 0) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(rng() % 50));
        lk.lock();
      }
    }
  }

  bool idle_locked() const {
    for (ihipStream_t* s : streams)
      if (!s->q.empty()) return false;
    return true;
  }
};

Device& dev() {
  static Device d;
  return d;
}

// Appends an operation to a stream (the caller holds the device lock).
void enqueue(ihipStream_t* s, Op op) {
  Device& d = dev();
  if (!s) model_fail("operation on the null stream");
  if (s->destroyed) model_fail("operation on a destroyed stream");
  if (s->cap) {
    Capture* c = s->cap;
    if (std::this_thread::get_id() != c->tid) model_fail("capture from another thread");
    ++c->pos;
    s->last_pos = c->pos;
    switch (op.kind) {
      case Op::FN:
        c->g->nodes.push_back(std::move(op.fn));
        break;
      case Op::RECORD:
        op.ev->cap = c;
        op.ev->cap_pos = c->pos;
        break;
      case Op::WAIT:
        if (op.ev->cap != c) model_fail("capturing stream waits on an event from outside the capture");
        // the origin waiting on an event recorded by a joined stream after
        // that stream's last operation rejoins it
        if (s == c->origin)
          for (ihipStream_t* j : c->joined)
            if (j->last_pos <= op.ev->cap_pos) c->joined_back[j] = c->pos;
        break;
    }
    return;
  }
  if (op.kind == Op::WAIT && op.ev->cap) model_fail("wait outside a capture on a capture event");
  if (op.kind == Op::RECORD) op.ev->cap = nullptr;
  s->q.push_back(std::move(op));
  d.cv_dev.notify_all();
}

}  // namespace

uint64_t fake_hip_ops_run() { return dev().ops_run.load(); }

extern "C" {

hipError_t hipMalloc(void** ptr, size_t size) {
  if (!ptr) return hipErrorInvalidValue;
  *ptr = aligned_alloc(256, (size + 255) / 256 * 256);
  return *ptr ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipFree(void* ptr) {
  Device& d = dev();
  std::unique_lock<std::mutex> lk(d.mu);
  d.cv_host.wait(lk, [&] { return d.idle_locked(); });  // hipFree synchronises
  free(ptr);
  return hipSuccess;
}

hipError_t hipHostMalloc(void** ptr, size_t size, unsigned int) {
  if (!ptr) return hipErrorInvalidValue;
  *ptr = calloc(1, size ? size : 1);
  return *ptr ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipHostFree(void* ptr) {
  Device& d = dev();
  std::unique_lock<std::mutex> lk(d.mu);
  d.cv_host.wait(lk, [&] { return d.idle_locked(); });
  free(ptr);
  return hipSuccess;
}

hipError_t hipSetDevice(int id) { return id == 0 ? hipSuccess : hipErrorInvalidDevice; }

hipError_t hipGetDevice(int* id) {
  if (!id) return hipErrorInvalidValue;
  *id = 0;
  return hipSuccess;
}

hipError_t hipDeviceSynchronize(void) {
  Device& d = dev();
  std::unique_lock<std::mutex> lk(d.mu);
  d.cv_host.wait(lk, [&] { return d.idle_locked(); });
  return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  if (!s) return hipErrorInvalidValue;
  Device& d = dev();
  std::lock_guard<std::mutex> g(d.mu);
  *s = new ihipStream_t;
  d.streams.push_back(*s);
  return hipSuccess;
}

hipError_t hipStreamDestroy(hipStream_t s) {
  Device& d = dev();
  std::unique_lock<std::mutex> lk(d.mu);
  if (!s || s->destroyed) model_fail("hipStreamDestroy of an invalid stream");
  d.cv_host.wait(lk, [&] { return s->q.empty(); });
  s->destroyed = true;
  d.streams.erase(std::find(d.streams.begin(), d.streams.end(), s));
  delete s;
  return hipSuccess;
}

hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) {
  if (!e) return hipErrorInvalidValue;
  // kept (reachable) until the process ends, to catch use after destroy
  static std::mutex reg_mu;
  static std::vector<std::unique_ptr<ihipEvent_t>> registry;
  std::lock_guard<std::mutex> g(reg_mu);
  registry.emplace_back(new ihipEvent_t);
  *e = registry.back().get();
  return hipSuccess;
}

hipError_t hipEventDestroy(hipEvent_t e) {
  std::lock_guard<std::mutex> g(dev().mu);
  if (!e || e->destroyed) model_fail("hipEventDestroy of an invalid event");
  e->destroyed = true;
  return hipSuccess;
}

hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  std::lock_guard<std::mutex> g(dev().mu);
  if (!e || e->destroyed) model_fail("hipEventRecord of a destroyed event");
  Op op;
  op.kind = Op::RECORD;
  op.ev = e;
  op.gen = ++e->recorded;
  enqueue(s, std::move(op));
  return hipSuccess;
}

hipError_t hipEventQuery(hipEvent_t e) {
  std::lock_guard<std::mutex> g(dev().mu);
  if (!e || e->destroyed) model_fail("hipEventQuery of a destroyed event");
  if (e->cap) model_fail("hipEventQuery of a capture event");
  return e->completed >= e->recorded ? hipSuccess : hipErrorNotReady;
}

hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int) {
  std::lock_guard<std::mutex> g(dev().mu);
  if (!e || e->destroyed) model_fail("hipStreamWaitEvent on a destroyed event");
  if (!s) model_fail("hipStreamWaitEvent on the null stream");
  if (!s->cap && e->cap) {
    // joins the capture of the event's record
    Capture* c = e->cap;
    if (c->ended) model_fail("wait on an event last recorded in an ended capture");
    s->cap = c;
    c->joined.push_back(s);
    s->last_pos = ++c->pos;
    return hipSuccess;
  }
  Op op;
  op.kind = Op::WAIT;
  op.ev = e;
  op.gen = e->recorded;
  enqueue(s, std::move(op));
  return hipSuccess;
}

hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind,
                          hipStream_t s) {
  std::lock_guard<std::mutex> g(dev().mu);
  Op op;
  op.fn = [dst, src, bytes] { memcpy(dst, src, bytes); };
  enqueue(s, std::move(op));
  return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  std::lock_guard<std::mutex> g(dev().mu);
  if (!s || s->destroyed) model_fail("capture on an invalid stream");
  if (s->cap) model_fail("capture on a capturing stream");
  Capture* c = new Capture;
  c->origin = s;
  c->g = new ihipGraph;
  c->tid = std::this_thread::get_id();
  s->cap = c;
  return hipSuccess;
}

hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* out) {
  std::lock_guard<std::mutex> g(dev().mu);
  if (!s || !s->cap || s->cap->origin != s) model_fail("end capture on a stream not capturing");
  Capture* c = s->cap;
  if (std::this_thread::get_id() != c->tid) model_fail("end capture from another thread");
  bool unjoined = false;
  for (ihipStream_t* j : c->joined) {
    auto it = c->joined_back.find(j);
    if (it == c->joined_back.end() || j->last_pos > it->second) unjoined = true;
    j->cap = nullptr;
  }
  s->cap = nullptr;
  c->ended = true;
  if (unjoined) {
    delete c->g;
    delete c;
    *out = nullptr;
    return hipErrorStreamCaptureUnjoined;
  }
  *out = c->g;
  // capture events keep pointing at the capture: they stay capture events
  // until recorded again outside one (a model of what may be waited on)
  static std::vector<std::unique_ptr<Capture>> ended;  // reachable, as events may refer to it
  ended.emplace_back(c);
  return hipSuccess;
}

hipError_t hipGraphInstantiate(hipGraphExec_t* out, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  std::lock_guard<std::mutex> lk(dev().mu);
  if (!g || g->destroyed) model_fail("instantiate of an invalid graph");
  hipGraphExec* x = new hipGraphExec;
  x->nodes = g->nodes;
  *out = x;
  return hipSuccess;
}

hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s) {
  std::lock_guard<std::mutex> lk(dev().mu);
  if (!x || x->destroyed) model_fail("launch of an invalid graph exec");
  ++x->pending;
  Op op;
  op.fn = [x] {
    for (auto& f : x->nodes) f();
    --x->pending;
  };
  enqueue(s, std::move(op));
  return hipSuccess;
}

hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
  std::lock_guard<std::mutex> lk(dev().mu);
  if (!x || x->destroyed) model_fail("destroy of an invalid graph exec");
  if (x->pending) model_fail("graph exec destroyed with a launch pending");
  x->destroyed = true;
  delete x;
  return hipSuccess;
}

hipError_t hipGraphDestroy(hipGraph_t gr) {
  std::lock_guard<std::mutex> lk(dev().mu);
  if (!gr || gr->destroyed) model_fail("destroy of an invalid graph");
  delete gr;
  return hipSuccess;
}

}  // extern "C"

// ---- model kernels ---------------------------------------------------------
namespace hpsp {

namespace {

int launch_generate(int N, int engine, const GenArgs& a, hipStream_t s, int prec) {
  if (N < 2 || N > HPSP_MAX_N || engine < 0 || engine > 5) return (int)hipErrorInvalidValue;
  const GenArgs c = a;
  std::lock_guard<std::mutex> g(dev().mu);
  Op op;
  op.fn = [c, N, prec] {
    for (uint64_t i = 0; i < c.n; ++i) {
      const uint64_t evt = c.first + i;
      for (int col = 0; col < 4 * N; ++col) {
        const uint64_t v = fake_event_word(c.d.key[0], evt, col, prec);
        if (c.out_f32)
          static_cast<uint32_t*>(c.cols)[col * c.ld + i] = (uint32_t)v;
        else
          static_cast<uint64_t*>(c.cols)[col * c.ld + i] = v;
      }
      if (c.evw) {
        const uint64_t w = fake_event_word(c.d.key[0], evt, 4 * HPSP_MAX_N, prec);
        memcpy(&c.evw[i], &w, 8);
      }
    }
  };
  enqueue(s, std::move(op));
  return 0;
}

int launch_average(int N, int engine, const AvgArgs& a, hipStream_t s, int prec) {
  if (N < 2 || N > HPSP_MAX_N || engine < 0 || engine > 5) return (int)hipErrorInvalidValue;
  const AvgArgs c = a;
  std::lock_guard<std::mutex> g(dev().mu);
  Op op;
  op.fn = [c, prec] {
    const uint64_t m = (c.n + kAvgChunk - 1) / kAvgChunk;
    for (uint64_t j = 0; j < m; ++j)
      c.partials[j] = fake_partial(c.d.key[0], c.first + j * kAvgChunk,
                                   std::min<uint64_t>(kAvgChunk, c.n - j * kAvgChunk), prec);
  };
  enqueue(s, std::move(op));
  return 0;
}

}  // namespace

int launch_generate_f64(int N, int engine, const GenArgs& a, hipStream_t s) {
  return launch_generate(N, engine, a, s, HPSP_FP64);
}
int launch_generate_mixed(int N, int engine, const GenArgs& a, hipStream_t s) {
  return launch_generate(N, engine, a, s, HPSP_MIXED);
}
int launch_average_f64(int N, int engine, const AvgArgs& a, hipStream_t s) {
  return launch_average(N, engine, a, s, HPSP_FP64);
}
int launch_average_mixed(int N, int engine, const AvgArgs& a, hipStream_t s) {
  return launch_average(N, engine, a, s, HPSP_MIXED);
}

int launch_stats_finalize(const Stats* partials, uint64_t m, Stats* out, hipStream_t s) {
  std::lock_guard<std::mutex> g(dev().mu);
  Op op;
  op.fn = [partials, m, out] { *out = fake_fold(partials, m); };
  enqueue(s, std::move(op));
  return 0;
}

}  // namespace hpsp
