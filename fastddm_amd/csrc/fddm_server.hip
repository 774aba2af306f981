// Request server: many image structure function analyses in flight on one
// GPU at once (continuous batching).
//
//   * One HBM arena is allocated when the server starts; every request's
//     workspace, output (and device copy of a host input) are carved from it
//     by a first-fit allocator, so serving never calls hipMalloc / hipFree
//     (hipFree waits for the whole device).
//   * A request is admitted as soon as a slot is free and its footprint
//     (workspace + output [+ input]) fits in the arena.  Requests are taken
//     in arrival order; a later request may pass one that does not fit yet,
//     unless the oldest pending request has waited longer than the aging
//     limit, which then holds admissions until it fits.
//   * Each slot is a worker thread with its own non-blocking stream: it
//     copies a host input (if any), runs the plan's kernels and waits for its
//     own stream only, so admitted requests overlap on the device.
//   * Plans (twiddles, lags, window tables) are cached per request shape for
//     the server's lifetime.
//   * Results stay in the arena until the client fetches and releases them.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include <hip/hip_runtime.h>

#include "fddm_hip.h"
#include "fddm_plan.hpp"

namespace {

int set_err(char* err, size_t len, int code, const char* fmt, ...) {
  if (err && len) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, len, fmt, ap);
    va_end(ap);
  }
  return code;
}

size_t pixel_bytes(int dtype) {
  switch (dtype) {
    case FDDM_U8: return 1;
    case FDDM_I16:
    case FDDM_U16: return 2;
    case FDDM_I32:
    case FDDM_U32:
    case FDDM_F32: return 4;
    case FDDM_I64:
    case FDDM_U64:
    case FDDM_F64: return 8;
    default: return 0;
  }
}

constexpr uint64_t kAlign = 256;
constexpr uint64_t kPieceBytes = uint64_t(256) << 20;  // spectra per launch of a large request
// Host inputs of at least kStageMin bytes cross the host link through two
// pinned buffers of kStageBytes per slot (allocated at the slot's first such
// request): host threads fill one while the DMA engine drains the other.
// Smaller inputs are copied by the runtime from the caller's pages directly.
constexpr uint64_t kStageBytes = uint64_t(32) << 20;
constexpr uint64_t kStageMin = uint64_t(8) << 20;
uint64_t align_up(uint64_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device range; blocks are kept sorted by
// offset and neighbours are merged on release.
class Arena {
 public:
  void reset(uint64_t size) {
    free_.clear();
    free_[0] = size;
    size_ = size;
    used_ = peak_ = 0;
  }
  bool alloc(uint64_t bytes, uint64_t* off) {
    bytes = align_up(std::max<uint64_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const uint64_t rest = it->second - bytes, at = it->first + bytes;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = align_up(std::max<uint64_t>(bytes, 1));
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  uint64_t size() const { return size_; }
  uint64_t used() const { return used_; }
  uint64_t peak() const { return peak_; }

 private:
  std::map<uint64_t, uint64_t> free_;
  uint64_t size_ = 0, used_ = 0, peak_ = 0;
};

struct PlanKey {
  int mode, dtype, precision, out_dtype;
  uint64_t L, H, W, nx, ny, nt;
  std::vector<uint32_t> lags;
  std::vector<unsigned char> window;  // raw bytes (empty: no window)
  int window_f32;
  bool operator<(const PlanKey& o) const {
    return std::tie(mode, dtype, precision, out_dtype, L, H, W, nx, ny, nt, window_f32, lags, window) <
           std::tie(o.mode, o.dtype, o.precision, o.out_dtype, o.L, o.H, o.W, o.nx, o.ny, o.nt, o.window_f32,
                    o.lags, o.window);
  }
};

enum State { kPending = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Request {
  uint64_t id = 0;
  PlanKey key;
  const void* img = nullptr;
  bool img_on_device = false;
  fddm_plan* plan = nullptr;
  uint64_t ws_bytes = 0, out_bytes = 0, in_bytes = 0;
  uint64_t ws_off = 0, out_off = 0, in_off = 0;
  int state = kPending;
  double t_submit = 0, t_admit = 0, t_done = 0;  // seconds since server start
  std::string error;
};

using Clock = std::chrono::steady_clock;

}  // namespace

struct fddm_server {
  int device = 0;
  int nslots = 0;
  double aging_s = 2.0;
  unsigned char* base = nullptr;
  Arena arena;
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;
  std::deque<uint64_t> finished;  // ids in completion order, not yet reported
  std::map<uint64_t, std::unique_ptr<Request>> requests;
  std::map<PlanKey, fddm_plan*> plans;
  std::mutex plan_mu;
  std::vector<std::thread> workers;
  std::vector<hipStream_t> streams;
  struct Stage {
    void* pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
  };
  std::vector<Stage> stages;  // per slot; touched only by the slot's worker until destroy
  uint64_t next_id = 1;
  bool stop = false;
  int running = 0;
  int waiters = 0;  // threads inside fddm_server_wait / fddm_server_next_done
  uint64_t completed = 0;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  // Picks the next request to admit and reserves its memory (mu held).
  Request* take_admissible() {
    if (pending.empty()) return nullptr;
    const bool head_starving = now() - pending.front()->t_submit > aging_s;
    for (auto it = pending.begin(); it != pending.end(); ++it) {
      Request* r = *it;
      uint64_t ws = 0, out = 0, in = 0;
      if (!arena.alloc(r->ws_bytes, &ws)) {
        if (head_starving) return nullptr;
        continue;
      }
      if (!arena.alloc(r->out_bytes, &out)) {
        arena.release(ws, r->ws_bytes);
        if (head_starving) return nullptr;
        continue;
      }
      if (r->in_bytes && !arena.alloc(r->in_bytes, &in)) {
        arena.release(ws, r->ws_bytes);
        arena.release(out, r->out_bytes);
        if (head_starving) return nullptr;
        continue;
      }
      r->ws_off = ws;
      r->out_off = out;
      r->in_off = in;
      pending.erase(it);
      return r;
    }
    return nullptr;
  }

  int plan_for(Request* r, char* err, size_t errlen) {
    std::lock_guard<std::mutex> g(plan_mu);
    auto it = plans.find(r->key);
    if (it != plans.end()) {
      r->plan = it->second;
      return FDDM_OK;
    }
    const PlanKey& k = r->key;
    fddm_plan* p = nullptr;
    const int rc = fddm_plan_create(&p, k.mode, k.dtype, k.L, k.H, k.W, k.lags.data(), k.lags.size(), k.nx, k.ny,
                                    k.nt, k.window.empty() ? nullptr : k.window.data(), k.window_f32, k.precision,
                                    k.out_dtype, err, errlen);
    if (rc) return rc;
    plans.emplace(k, p);
    r->plan = p;
    return FDDM_OK;
  }

  // Host input -> device on stream st (asynchronous: the caller synchronises).
  hipError_t copy_in(int slot, void* d_dst, const void* h_src, uint64_t n, hipStream_t st) {
    Stage& g = stages[slot];
    if (n >= kStageMin && !g.pin[0]) {
      for (int k = 0; k < 2; ++k) {
        if (hipHostMalloc(&g.pin[k], kStageBytes, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&g.ev[k], hipEventDisableTiming) != hipSuccess) {
          (void)hipGetLastError();
          free_stage(g);  // pinned memory short: the runtime copies this input itself
          break;
        }
      }
    }
    if (n < kStageMin || !g.pin[0]) return hipMemcpyAsync(d_dst, h_src, n, hipMemcpyHostToDevice, st);
    uint64_t off = 0;
    for (int k = 0; off < n; k ^= 1) {
      const uint64_t m = std::min(kStageBytes, n - off);
      hipError_t e = hipEventSynchronize(g.ev[k]);  // the buffer's previous DMA is done
      if (e != hipSuccess) return e;
      fddm::parallel_memcpy(g.pin[k], static_cast<const unsigned char*>(h_src) + off, m);
      if ((e = hipMemcpyAsync(static_cast<unsigned char*>(d_dst) + off, g.pin[k], m, hipMemcpyHostToDevice, st)) !=
              hipSuccess ||
          (e = hipEventRecord(g.ev[k], st)) != hipSuccess)
        return e;
      off += m;
    }
    return hipSuccess;
  }
  static void free_stage(Stage& g) {
    for (int k = 0; k < 2; ++k) {
      if (g.pin[k]) (void)hipHostFree(g.pin[k]);
      if (g.ev[k]) (void)hipEventDestroy(g.ev[k]);
      g.pin[k] = nullptr;
      g.ev[k] = nullptr;
    }
  }

  void worker(int slot) {
    (void)hipSetDevice(device);  // validated by fddm_server_create
    hipStream_t s = streams[slot];
    char err[512];
    for (;;) {
      Request* r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || (r = take_admissible()) != nullptr; });
        if (!r) return;  // stopping and nothing admissible
        r->state = kRunning;
        r->t_admit = now();
        ++running;
      }
      err[0] = 0;
      int rc = FDDM_OK;
      const void* img = r->img;
      if (!r->img_on_device) {
        if (copy_in(slot, base + r->in_off, r->img, r->in_bytes, s) != hipSuccess)
          rc = set_err(err, sizeof err, FDDM_ERR_HIP, "input copy failed");
        img = base + r->in_off;
      }
      // the plan was created at submission; execute only.  Large requests run
      // as launches of about kPieceBytes of spectra each, so the kernels of
      // requests on other slots are dispatched between them.
      if (!rc) {
        const uint64_t sb = fddm::plan_spectra_bytes(r->plan);
        if (sb > 2 * kPieceBytes) {
          const uint64_t L = r->key.L, nq = r->key.ny * (r->key.nx / 2 + 1);
          const uint64_t frames = std::max<uint64_t>(1, L * kPieceBytes / sb);
          const int cols = static_cast<int>(std::max<uint64_t>(256, nq * kPieceBytes / sb));
          rc = fddm::plan_execute_pieces(r->plan, img, base + r->out_off, base + r->ws_off, s, frames, cols, err,
                                         sizeof err);
        } else {
          rc = fddm_plan_execute(r->plan, img, base + r->out_off, base + r->ws_off, s, err, sizeof err);
        }
      }
      if (hipStreamSynchronize(s) != hipSuccess && !rc)
        rc = set_err(err, sizeof err, FDDM_ERR_HIP, "stream synchronisation failed");
      {
        std::lock_guard<std::mutex> g(mu);
        arena.release(r->ws_off, r->ws_bytes);
        if (r->in_bytes) arena.release(r->in_off, r->in_bytes);
        r->in_bytes = 0;
        r->t_done = now();
        if (rc) {
          arena.release(r->out_off, r->out_bytes);
          r->out_bytes = 0;
          r->state = kFailed;
          r->error = err;
        } else {
          r->state = kDone;
        }
        --running;
        ++completed;
        finished.push_back(r->id);
      }
      cv_done.notify_all();
      cv_work.notify_all();  // memory came back: others may now fit
    }
  }
};

extern "C" {

int fddm_server_create(fddm_server** out, int max_slots, uint64_t hbm_bytes, char* err, size_t errlen) {
  if (!out || max_slots < 1 || max_slots > 64)
    return set_err(err, errlen, FDDM_ERR_ARG, "max_slots must be in [1, 64]");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_err(err, errlen, FDDM_ERR_DEVICE, "no current device");
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return set_err(err, errlen, FDDM_ERR_HIP, "hipMemGetInfo failed");
  if (hbm_bytes == 0) hbm_bytes = static_cast<uint64_t>(0.85 * static_cast<double>(fr));
  hbm_bytes &= ~(kAlign - 1);
  if (hbm_bytes > fr) return set_err(err, errlen, FDDM_ERR_MEMORY, "arena of %llu bytes exceeds free memory (%llu)",
                                      (unsigned long long)hbm_bytes, (unsigned long long)fr);
  auto* s = new (std::nothrow) fddm_server();
  if (!s) return set_err(err, errlen, FDDM_ERR_MEMORY, "host allocation failed");
  s->device = dev;
  s->nslots = max_slots;
  s->t0 = Clock::now();
  if (hipMalloc(reinterpret_cast<void**>(&s->base), hbm_bytes) != hipSuccess) {
    delete s;
    return set_err(err, errlen, FDDM_ERR_MEMORY, "arena allocation of %llu bytes failed",
                   (unsigned long long)hbm_bytes);
  }
  s->arena.reset(hbm_bytes);
  s->streams.resize(max_slots);
  s->stages.resize(max_slots);
  for (int i = 0; i < max_slots; ++i) {
    if (hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking) != hipSuccess) {
      for (int j = 0; j < i; ++j) (void)hipStreamDestroy(s->streams[j]);
      (void)hipFree(s->base);
      delete s;
      return set_err(err, errlen, FDDM_ERR_HIP, "stream creation failed");
    }
  }
  for (int i = 0; i < max_slots; ++i) s->workers.emplace_back(&fddm_server::worker, s, i);
  *out = s;
  return FDDM_OK;
}

int fddm_server_destroy(fddm_server* s) {
  if (!s) return FDDM_OK;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // pending requests that were never admitted are dropped
    for (Request* r : s->pending) {
      r->state = kFailed;
      r->error = "server destroyed before admission";
      r->t_done = s->now();
      s->finished.push_back(r->id);
    }
    s->pending.clear();
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();  // waiters on dropped requests return (failed)
  for (auto& t : s->workers) t.join();
  {
    // no thread may still be blocked on (or about to re-lock) the server's
    // mutex and condition variables when they are freed
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.notify_all();
    s->cv_done.wait(lk, [&] { return s->waiters == 0; });
  }
  (void)hipSetDevice(s->device);
  for (auto& kv : s->plans) fddm_plan_destroy(kv.second);
  for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
  for (auto& g : s->stages) fddm_server::free_stage(g);
  (void)hipFree(s->base);
  delete s;
  return FDDM_OK;
}

int fddm_server_set_aging(fddm_server* s, double seconds) {
  if (!s || !(seconds >= 0)) return FDDM_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  s->aging_s = seconds;
  return FDDM_OK;
}

int fddm_server_submit(fddm_server* s, int mode, const void* img, int img_on_device, int dtype, uint64_t length,
                       uint64_t height, uint64_t width, const uint32_t* lags, uint64_t num_lags, uint64_t nx,
                       uint64_t ny, uint64_t nt, const void* window, int window_f32, int precision, int out_dtype,
                       uint64_t* id, char* err, size_t errlen) {
  if (!s || !img || !id) return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument to submit");
  const double t_entry = s->now();  // latency includes plan creation
  const size_t pb = pixel_bytes(dtype);
  if (!pb) return set_err(err, errlen, FDDM_ERR_ARG, "unknown input dtype %d", dtype);
  auto r = std::make_unique<Request>();
  PlanKey& k = r->key;
  k.mode = mode;
  k.dtype = dtype;
  k.precision = precision;
  k.out_dtype = out_dtype;
  k.L = length;
  k.H = height;
  k.W = width;
  k.nx = nx;
  k.ny = ny;
  k.nt = nt;
  k.window_f32 = window_f32;
  if (num_lags) k.lags.assign(lags, lags + num_lags);
  if (window) {
    const size_t wb = height * width * (window_f32 ? 4 : 8);
    k.window.assign(static_cast<const unsigned char*>(window), static_cast<const unsigned char*>(window) + wb);
  }
  r->img = img;
  r->img_on_device = img_on_device != 0;
  // validates every argument and builds (or finds) the plan
  if (hipSetDevice(s->device) != hipSuccess)
    return set_err(err, errlen, FDDM_ERR_DEVICE, "cannot select device %d", s->device);
  int rc = s->plan_for(r.get(), err, errlen);
  if (rc) return rc;
  fddm_plan_workspace_size(r->plan, &r->ws_bytes);
  fddm_plan_output_size(r->plan, &r->out_bytes);
  r->in_bytes = r->img_on_device ? 0 : length * height * width * pb;
  const uint64_t need = align_up(r->ws_bytes) + align_up(r->out_bytes) + align_up(r->in_bytes);
  if (need > s->arena.size())
    return set_err(err, errlen, FDDM_ERR_MEMORY, "request needs %llu bytes, arena holds %llu",
                   (unsigned long long)need, (unsigned long long)s->arena.size());
  {
    std::lock_guard<std::mutex> g(s->mu);
    r->id = s->next_id++;
    r->t_submit = t_entry;
    *id = r->id;
    s->pending.push_back(r.get());
    s->requests.emplace(r->id, std::move(r));
  }
  s->cv_work.notify_all();
  return FDDM_OK;
}

int fddm_server_status(fddm_server* s, uint64_t id, int* state, double* times, char* err, size_t errlen) {
  if (!s || !state) return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument to status");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return set_err(err, errlen, FDDM_ERR_ARG, "unknown request %llu", (unsigned long long)id);
  const Request* r = it->second.get();
  *state = r->state;
  if (times) {
    times[0] = r->t_submit;
    times[1] = r->t_admit;
    times[2] = r->t_done;
  }
  if (r->state == kFailed) set_err(err, errlen, FDDM_ERR_HIP, "%s", r->error.c_str());
  return FDDM_OK;
}

int fddm_server_wait(fddm_server* s, uint64_t id, double timeout_s, char* err, size_t errlen) {
  if (!s) return set_err(err, errlen, FDDM_ERR_ARG, "NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return set_err(err, errlen, FDDM_ERR_ARG, "unknown request %llu", (unsigned long long)id);
  Request* r = it->second.get();
  const auto done = [&] { return r->state == kDone || r->state == kFailed; };
  ++s->waiters;
  bool ok = true;
  if (timeout_s < 0) {
    s->cv_done.wait(lk, done);
  } else {
    ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
  }
  const int state = r->state;
  const std::string error = r->error;
  if (--s->waiters == 0 && s->stop) s->cv_done.notify_all();
  if (!ok)
    return set_err(err, errlen, FDDM_ERR_ARG, "request %llu not finished after %.3f s", (unsigned long long)id,
                   timeout_s);
  if (state == kFailed) return set_err(err, errlen, FDDM_ERR_HIP, "%s", error.c_str());
  return FDDM_OK;
}

int fddm_server_next_done(fddm_server* s, double timeout_s, uint64_t* id, char* err, size_t errlen) {
  if (!s || !id) return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument to next_done");
  std::unique_lock<std::mutex> lk(s->mu);
  const auto ready = [&] { return !s->finished.empty() || s->stop; };
  ++s->waiters;
  bool ok = true;
  if (timeout_s < 0) {
    s->cv_done.wait(lk, ready);
  } else {
    ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
  }
  *id = 0;
  if (ok && !s->finished.empty()) {
    *id = s->finished.front();
    s->finished.pop_front();
  }
  const bool stopped = s->stop;
  if (--s->waiters == 0 && s->stop) s->cv_done.notify_all();
  if (!*id && stopped) return set_err(err, errlen, FDDM_ERR_ARG, "server is shutting down");
  return FDDM_OK;
}

int fddm_server_result(fddm_server* s, uint64_t id, void** d_out, uint64_t* bytes, char* err, size_t errlen) {
  if (!s || !d_out || !bytes) return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument to result");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return set_err(err, errlen, FDDM_ERR_ARG, "unknown request %llu", (unsigned long long)id);
  const Request* r = it->second.get();
  if (r->state != kDone) return set_err(err, errlen, FDDM_ERR_ARG, "request %llu has no result", (unsigned long long)id);
  *d_out = s->base + r->out_off;
  *bytes = r->out_bytes;
  return FDDM_OK;
}

int fddm_server_fetch(fddm_server* s, uint64_t id, void* host_out, char* err, size_t errlen) {
  void* d = nullptr;
  uint64_t bytes = 0;
  int rc = fddm_server_result(s, id, &d, &bytes, err, errlen);
  if (rc) return rc;
  if (hipSetDevice(s->device) != hipSuccess) return set_err(err, errlen, FDDM_ERR_HIP, "result copy failed");
  return fddm::copy_to_host(s->device, host_out, d, bytes, err, errlen);
}

int fddm_server_release(fddm_server* s, uint64_t id) {
  if (!s) return FDDM_ERR_ARG;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return FDDM_ERR_ARG;
    Request* r = it->second.get();
    if (r->state == kPending || r->state == kRunning) return FDDM_ERR_ARG;
    if (r->out_bytes) s->arena.release(r->out_off, r->out_bytes);
    // a released id is never reported by fddm_server_next_done
    s->finished.erase(std::remove(s->finished.begin(), s->finished.end(), id), s->finished.end());
    s->requests.erase(it);
  }
  s->cv_work.notify_all();
  return FDDM_OK;
}

int fddm_server_stats(fddm_server* s, uint64_t* stats) {
  if (!s || !stats) return FDDM_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  stats[0] = s->pending.size();
  stats[1] = static_cast<uint64_t>(s->running);
  stats[2] = s->completed;
  stats[3] = s->arena.size();
  stats[4] = s->arena.used();
  stats[5] = s->arena.peak();
  stats[6] = s->plans.size();
  return FDDM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Synthetic device-side stacks for load generation: uniform integers in
// [lo, hi) from a counter-based hash (splitmix64 of seed + index), so a
// large stack never crosses the host link.
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill_uniform(T* p, uint64_t n, uint64_t seed, uint32_t lo, uint32_t span) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t h = splitmix64(seed ^ (i * 0xD1B54A32D192ED03ull));
    p[i] = static_cast<T>(lo + static_cast<uint32_t>((h >> 32) % span));
  }
}

}  // namespace

extern "C" int fddm_fill_uniform(void* d, uint64_t count, int dtype, uint64_t seed, uint32_t lo, uint32_t hi,
                                 void* stream, char* err, size_t errlen) {
  if (!d || hi <= lo) return set_err(err, errlen, FDDM_ERR_ARG, "fill: NULL buffer or empty range");
  if (count == 0) return FDDM_OK;
  auto s = static_cast<hipStream_t>(stream);
  const uint64_t blocks = std::min<uint64_t>((count + 255) / 256, 1u << 16);
  if (dtype == FDDM_U8 && hi <= 256) {
    hipLaunchKernelGGL(k_fill_uniform<uint8_t>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                       static_cast<uint8_t*>(d), count, seed, lo, hi - lo);
  } else if (dtype == FDDM_U16 && hi <= 65536) {
    hipLaunchKernelGGL(k_fill_uniform<uint16_t>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                       static_cast<uint16_t*>(d), count, seed, lo, hi - lo);
  } else {
    return set_err(err, errlen, FDDM_ERR_ARG, "fill: only uint8 / uint16 ranges are supported");
  }
  if (hipGetLastError() != hipSuccess) return set_err(err, errlen, FDDM_ERR_HIP, "fill kernel launch failed");
  return FDDM_OK;
}
