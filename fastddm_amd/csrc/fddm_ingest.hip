// Streamed ingest: frames arrive piece by piece from the caller (a reader
// decoding a TIFF stack, a camera) and are transformed while the next piece
// is read.
//
// Pushed frames are gathered into chunks of whole frame blocks in two pinned
// host buffers.  A full chunk is copied to one of two device input buffers on
// a copy stream, and the row and column passes of its frames run on a
// compute stream into the q-major spectra at the chunk's frame offset -- the
// same launches as the chunked host path, so the result equals the one-shot
// result bit for bit.  Host copies, DMA and kernels of consecutive chunks
// overlap: the host fills one pinned buffer while the other one's frames move
// and the previous chunk is transformed.  Frame 0 (the subtracted frame) and
// the edge frames of the compensated path are kept on the device as they
// pass.  finish runs the time stage (and its corrections) over the whole
// plane and copies the result to the host.
#include <algorithm>
#include <cstdarg>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "fddm_hip.h"
#include "fddm_kernels.hpp"
#include "fddm_plan.hpp"

using namespace fddm;

struct fddm_ingest {
  fddm_plan* plan = nullptr;
  int device = 0;
  PlanInfo info{};
  uint64_t frame_b = 0, chunk_frames = 0;
  uint64_t pushed = 0;   // frames received
  uint64_t staged = 0;   // frames in the pinned buffer being filled
  uint64_t issued = 0;   // frames whose chunk has been handed to the device
  int cur = 0;           // pinned / device buffer being filled
  void* pin[2] = {nullptr, nullptr};
  bool pooled = false;  // pin[] belong to the device's pool (returned, not freed, at the end)
  void* din[2] = {nullptr, nullptr};
  void* frame0 = nullptr;
  void* edge_in = nullptr;
  void* ws = nullptr;
  void* out = nullptr;
  hipStream_t copy = nullptr, comp = nullptr;
  hipEvent_t copied[2] = {nullptr, nullptr}, done[2] = {nullptr, nullptr};
  bool failed = false;
};

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

#define FDDM_HIP(expr)                                                                              \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) {                                                                         \
      h->failed = true;                                                                             \
      return set_err(err, errlen, e_ == hipErrorOutOfMemory ? FDDM_ERR_MEMORY : FDDM_ERR_HIP, "%s: %s", \
                     #expr, hipGetErrorString(e_));                                                 \
    }                                                                                               \
  } while (0)

constexpr uint64_t kChunkBytes = uint64_t(64) << 20;     // input bytes per chunk (target)

// Pinned chunk buffers kept per device between ingests, until the process
// ends, so an ingest pins no host memory again: one ingest at a time per
// device reuses the pool; a concurrent one pins its own.
struct PinPool {
  std::mutex mu;
  bool busy = false;
  void* buf[2] = {nullptr, nullptr};
  uint64_t cap = 0;
  // the workspace and result buffers of the last ingest, kept while they
  // total at most kKeepDeviceBytes
  void* ws = nullptr;
  void* out = nullptr;
  uint64_t ws_cap = 0, out_cap = 0;
};
PinPool g_pool[64];
constexpr uint64_t kKeepDeviceBytes = uint64_t(16) << 30;

// Frees the pool's device buffers (pool lock held).
void drop_device(PinPool& pool) {
  for (void** b : {&pool.ws, &pool.out})
    if (*b) (void)hipFree(*b);
  pool.ws = pool.out = nullptr;
  pool.ws_cap = pool.out_cap = 0;
}

// Device buffer of at least `need` bytes in *buf (capacity *cap), grown by
// a new allocation when smaller.
hipError_t grow_device(void** buf, uint64_t* cap, uint64_t need) {
  if (*buf && *cap >= need) return hipSuccess;
  if (*buf) (void)hipFree(*buf);
  *buf = nullptr;
  *cap = 0;
  const hipError_t e = hipMalloc(buf, need ? need : 1);
  if (e == hipSuccess) *cap = need;
  return e;
}
constexpr uint64_t kMinStageBytes = uint64_t(4) << 20;   // smaller chunk buffers: result copied directly

void free_all(fddm_ingest* h) {
  if (h->comp) (void)hipStreamSynchronize(h->comp);
  if (h->copy) (void)hipStreamSynchronize(h->copy);
  if (h->pooled) {
    PinPool& pool = g_pool[h->device];
    std::lock_guard<std::mutex> lk(pool.mu);
    pool.busy = false;
    h->pin[0] = h->pin[1] = nullptr;
    h->ws = h->out = nullptr;
    if (pool.ws_cap + pool.out_cap > kKeepDeviceBytes) drop_device(pool);
  }
  for (int k = 0; k < 2; ++k) {
    if (h->pin[k]) (void)hipHostFree(h->pin[k]);
    if (h->din[k]) (void)hipFree(h->din[k]);
    if (h->copied[k]) (void)hipEventDestroy(h->copied[k]);
    if (h->done[k]) (void)hipEventDestroy(h->done[k]);
  }
  for (void* p : {h->frame0, h->edge_in, h->ws, h->out})
    if (p) (void)hipFree(p);
  if (h->copy) (void)hipStreamDestroy(h->copy);
  if (h->comp) (void)hipStreamDestroy(h->comp);
  if (h->plan) fddm_plan_destroy(h->plan);
}

// Hands the pinned buffer `cur` (staged frames starting at frame `issued`) to
// the device: DMA on the copy stream, rows and columns on the compute stream.
int issue_chunk(fddm_ingest* h, char* err, size_t errlen) {
  const int k = h->cur;
  const uint64_t fr = h->staged, t0 = h->issued;
  if (fr == 0) return FDDM_OK;
  // the device buffer is free once the chunk that used it last is transformed
  FDDM_HIP(hipStreamWaitEvent(h->copy, h->done[k], 0));
  FDDM_HIP(hipMemcpyAsync(h->din[k], h->pin[k], fr * h->frame_b, hipMemcpyHostToDevice, h->copy));
  FDDM_HIP(hipEventRecord(h->copied[k], h->copy));
  FDDM_HIP(hipStreamWaitEvent(h->comp, h->copied[k], 0));
  const auto* din = static_cast<const unsigned char*>(h->din[k]);
  if (t0 == 0) FDDM_HIP(hipMemcpyAsync(h->frame0, din, h->frame_b, hipMemcpyDeviceToDevice, h->comp));
  // edge frames of the compensated path: the first E and the last E
  if (h->edge_in) {
    // slots of the child's input: frames [0, E) then [L - E, L) (every frame
    // of a stack of at most 2E)
    const uint64_t E = kEdgeFrames, L = h->info.L;
    for (uint64_t t = t0; t < t0 + fr; ++t) {
      uint64_t slot = UINT64_MAX;
      if (L <= 2 * E || t < E) slot = t;
      else if (t >= L - E) slot = E + (t - (L - E));
      if (slot != UINT64_MAX)
        FDDM_HIP(hipMemcpyAsync(static_cast<unsigned char*>(h->edge_in) + slot * h->frame_b, din + (t - t0) * h->frame_b,
                                h->frame_b, hipMemcpyDeviceToDevice, h->comp));
    }
  }
  const int rc = plan_spatial_frames(h->plan, din, h->frame0, t0, fr, h->ws, h->comp, err, errlen);
  if (rc) {
    h->failed = true;
    return rc;
  }
  FDDM_HIP(hipEventRecord(h->done[k], h->comp));
  h->issued += fr;
  h->staged = 0;
  h->cur ^= 1;
  return FDDM_OK;
}

}  // namespace

void fddm::ingest_release_pool(int dev) {
  if (dev < 0 || dev >= 64) return;
  PinPool& pool = g_pool[dev];
  std::lock_guard<std::mutex> lk(pool.mu);
  if (pool.busy) return;  // an ingest still uses it; freed by a later release
  drop_device(pool);       // the pinned buffers stay (see PinPool)
}

extern "C" {

int fddm_ingest_begin(fddm_ingest** out, int mode, int dtype, uint64_t length, uint64_t height, uint64_t width,
                      const uint32_t* lags, uint64_t num_lags, uint64_t nx, uint64_t ny, uint64_t nt,
                      const void* window, int window_f32, int precision, int out_dtype, char* err, size_t errlen) {
  if (!out) return set_err(err, errlen, FDDM_ERR_ARG, "NULL handle pointer");
  *out = nullptr;
  auto* h = new (std::nothrow) fddm_ingest();
  if (!h) return set_err(err, errlen, FDDM_ERR_MEMORY, "host allocation failed");
  struct Guard {
    fddm_ingest* h;
    bool keep = false;
    ~Guard() {
      if (!keep) {
        free_all(h);
        delete h;
      }
    }
  } guard{h};
  int rc = plan_create_ex(&h->plan, mode, dtype, length, height, width, lags, num_lags, nx, ny, nt, window, window_f32,
                          precision, out_dtype, 0, err, errlen);
  if (rc) return rc;
  FDDM_HIP(hipGetDevice(&h->device));
  if (h->device < 0 || h->device >= 64) return set_err(err, errlen, FDDM_ERR_DEVICE, "device id %d not supported", h->device);
  h->info = plan_info(h->plan);
  h->frame_b = height * width * h->info.in_bytes;
  // chunks of whole frame blocks, about kChunkBytes of input each
  const uint64_t tb = static_cast<uint64_t>(h->info.Tb);
  h->chunk_frames = std::max<uint64_t>(tb, kChunkBytes / std::max<uint64_t>(1, h->frame_b) / tb * tb);
  h->chunk_frames = std::min(h->chunk_frames, (length + tb - 1) / tb * tb);
  FDDM_HIP(hipStreamCreateWithFlags(&h->copy, hipStreamNonBlocking));
  FDDM_HIP(hipStreamCreateWithFlags(&h->comp, hipStreamNonBlocking));
  const uint64_t pin_b = h->chunk_frames * h->frame_b;
  {
    PinPool& pool = g_pool[h->device];
    std::lock_guard<std::mutex> lk(pool.mu);
    if (!pool.busy) {
      if (pool.cap < pin_b) {
        for (void*& b : pool.buf)
          if (b) (void)hipHostFree(b);
        pool.buf[0] = pool.buf[1] = nullptr;
        pool.cap = 0;
        for (void*& b : pool.buf) FDDM_HIP(hipHostMalloc(&b, pin_b, hipHostMallocDefault));
        pool.cap = pin_b;
      }
      pool.busy = true;
      h->pooled = true;
      h->pin[0] = pool.buf[0];
      h->pin[1] = pool.buf[1];
    }
  }
  for (int k = 0; k < 2; ++k) {
    if (!h->pooled) FDDM_HIP(hipHostMalloc(&h->pin[k], pin_b, hipHostMallocDefault));
    FDDM_HIP(hipMalloc(&h->din[k], pin_b));
    FDDM_HIP(hipEventCreateWithFlags(&h->copied[k], hipEventDisableTiming));
    FDDM_HIP(hipEventCreateWithFlags(&h->done[k], hipEventDisableTiming));
    FDDM_HIP(hipEventRecord(h->done[k], h->comp));
  }
  FDDM_HIP(hipMalloc(&h->frame0, h->frame_b));
  if (h->info.edge_frames) FDDM_HIP(hipMalloc(&h->edge_in, h->info.edge_frames * h->frame_b));
  if (h->pooled) {
    PinPool& pool = g_pool[h->device];
    std::lock_guard<std::mutex> lk(pool.mu);
    FDDM_HIP(grow_device(&pool.ws, &pool.ws_cap, h->info.ws_bytes));
    FDDM_HIP(grow_device(&pool.out, &pool.out_cap, h->info.out_bytes));
    h->ws = pool.ws;
    h->out = pool.out;
  } else {
    FDDM_HIP(hipMalloc(&h->ws, h->info.ws_bytes));
    FDDM_HIP(hipMalloc(&h->out, h->info.out_bytes));
  }
  guard.keep = true;
  *out = h;
  return FDDM_OK;
}

int fddm_ingest_push(fddm_ingest* h, const void* frames, uint64_t count, char* err, size_t errlen) {
  if (!h || (!frames && count)) return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument to push");
  if (h->failed) return set_err(err, errlen, FDDM_ERR_ARG, "ingest failed earlier");
  if (h->pushed + count > h->info.L)
    return set_err(err, errlen, FDDM_ERR_ARG, "%llu frames pushed, the stack has %llu",
                   static_cast<unsigned long long>(h->pushed + count), static_cast<unsigned long long>(h->info.L));
  FDDM_HIP(hipSetDevice(h->device));
  const auto* src = static_cast<const unsigned char*>(frames);
  while (count) {
    if (h->staged == 0) FDDM_HIP(hipEventSynchronize(h->copied[h->cur]));  // its last DMA is done
    const uint64_t take = std::min(count, h->chunk_frames - h->staged);
    parallel_memcpy(static_cast<unsigned char*>(h->pin[h->cur]) + h->staged * h->frame_b, src, take * h->frame_b);
    h->staged += take;
    h->pushed += take;
    src += take * h->frame_b;
    count -= take;
    if (h->staged == h->chunk_frames || h->pushed == h->info.L) {
      const int rc = issue_chunk(h, err, errlen);
      if (rc) return rc;
    }
  }
  return FDDM_OK;
}

int fddm_ingest_finish(fddm_ingest* h, void* out, char* err, size_t errlen) {
  if (!h || !out) return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument to finish");
  if (h->failed) return set_err(err, errlen, FDDM_ERR_ARG, "ingest failed earlier");
  if (h->pushed != h->info.L)
    return set_err(err, errlen, FDDM_ERR_ARG, "%llu of %llu frames pushed", static_cast<unsigned long long>(h->pushed),
                   static_cast<unsigned long long>(h->info.L));
  FDDM_HIP(hipSetDevice(h->device));
  int rc = plan_finish_from_spectra(h->plan, h->frame0, h->edge_in, h->ws, h->out, h->comp, err, errlen);
  if (rc) {
    h->failed = true;
    return rc;
  }
  // the result leaves through the two pinned chunk buffers (their frames are
  // on the device by now: the compute stream waited for every copy): host
  // threads drain one while the DMA engine fills the other
  const uint64_t n = h->info.out_bytes, stage = h->chunk_frames * h->frame_b;
  const auto* src = static_cast<const unsigned char*>(h->out);
  auto* dst = static_cast<unsigned char*>(out);
  if (stage < kMinStageBytes || n <= stage) {
    FDDM_HIP(hipMemcpyAsync(out, h->out, n, hipMemcpyDeviceToHost, h->comp));
    FDDM_HIP(hipStreamSynchronize(h->comp));
    return FDDM_OK;
  }
  const uint64_t nchunk = (n + stage - 1) / stage;
  for (uint64_t i = 0; i < nchunk + 2; ++i) {
    if (i >= 2) {  // chunk i - 2 has arrived: copy it out, freeing its buffer
      const uint64_t j = i - 2, off = j * stage, m = std::min(stage, n - off);
      FDDM_HIP(hipEventSynchronize(h->copied[j & 1]));
      parallel_memcpy(dst + off, h->pin[j & 1], m);
    }
    if (i < nchunk) {
      const uint64_t off = i * stage, m = std::min(stage, n - off);
      FDDM_HIP(copy_to_pinned(h->pin[i & 1], src + off, m, h->comp));
      FDDM_HIP(hipEventRecord(h->copied[i & 1], h->comp));
    }
  }
  return FDDM_OK;
}

int fddm_ingest_destroy(fddm_ingest* h) {
  if (!h) return FDDM_OK;
  (void)hipSetDevice(h->device);
  free_all(h);
  delete h;
  return FDDM_OK;
}

}  // extern "C"
