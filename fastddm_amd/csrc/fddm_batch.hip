// Batch executor: many stacks of uneven length and image size analysed by
// one graph launch.
//
// At creation every stack gets a plan (frames per block chosen by the
// autotuner below), one device workspace block holds all their workspaces,
// and the stage chains of all stacks are captured into a single hipGraph on
// parallel branches (fork from one event, join into one).  Each launch
// replays the graph: the device sees every stack's kernels at once, with no
// host round trips between the 4 x n kernel launches.
//
// Autotuner: for each distinct (mode, dtype, H, W, nx, ny, precision,
// window) the candidate frame-block sizes that fit LDS, and for each the
// mixed-radix or the generic form of the row and column kernels where the
// mixed-radix ones apply, are timed on a sample of the shape's longest stack
// (the spatial stages' HIP-event times, best of two rounds); the choice is
// cached for the life of the process.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <new>
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

uint64_t align256(uint64_t b) { return (b + 255) & ~uint64_t(255); }

struct TuneKey {
  int mode, dtype, precision, out_dtype, has_window;
  uint64_t H, W, nx, ny;
  bool operator<(const TuneKey& o) const {
    return std::tie(mode, dtype, precision, out_dtype, has_window, H, W, nx, ny) <
           std::tie(o.mode, o.dtype, o.precision, o.out_dtype, o.has_window, o.H, o.W, o.nx, o.ny);
  }
};

// Tuned choice for one shape: frames per block and, where the mixed-radix
// row / column kernels apply, whether the generic workgroup kernels run
// instead.
struct TuneChoice {
  int tb = 0, rows_generic = 0, cols_generic = 0;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneChoice> g_tuned;

constexpr uint64_t kTuneFrames = 512;

// Spatial-stage times of one candidate over the first `frames` frames of
// `it`: frame-0 spectra, row pass and column pass, the best of two rounds of
// three executions.  *forms receives whether the mixed-radix row / column
// kernels ran (with rows_generic / cols_generic 0).
int time_candidate(const fddm_batch_item& it, uint64_t frames, int tb, int rows_generic, int cols_generic,
                   hipStream_t s, double* f0_ms, double* row_ms, double* col_ms, int* rows_mr, int* cols_mr,
                   char* err, size_t errlen) {
  std::vector<uint32_t> lags;
  for (uint64_t d = 1; d < frames && lags.size() < 64; ++d) lags.push_back(static_cast<uint32_t>(d));
  uint64_t nt = 2 * frames;
  fddm_plan* p = nullptr;
  int rc = fddm::plan_create_ex(&p, it.mode, it.dtype, frames, it.height, it.width, lags.data(), lags.size(), it.nx,
                                it.ny, nt, it.window, it.window_f32, it.precision, it.out_dtype, tb, err, errlen);
  if (rc) return rc;
  fddm::plan_set_generic_spatial(p, rows_generic, cols_generic);
  fddm::plan_spatial_forms(p, rows_mr, cols_mr);
  uint64_t ws = 0, out = 0;
  fddm_plan_workspace_size(p, &ws);
  fddm_plan_output_size(p, &out);
  void* buf = nullptr;
  if (hipMalloc(&buf, align256(ws) + align256(out)) != hipSuccess) {
    fddm_plan_destroy(p);
    return set_err(err, errlen, FDDM_ERR_MEMORY, "autotune buffers");
  }
  void* d_ws = buf;
  void* d_out = static_cast<unsigned char*>(buf) + align256(ws);
  if (!rc) rc = fddm_plan_execute(p, it.d_img, d_out, d_ws, s, err, errlen);  // warm-up
  *f0_ms = *row_ms = *col_ms = 1e30;
  for (int round = 0; round < 2 && !rc; ++round) {
    if (!rc) rc = fddm_plan_timing_begin(p, 3, err, errlen);
    for (int i = 0; i < 3 && !rc; ++i) rc = fddm_plan_execute(p, it.d_img, d_out, d_ws, s, err, errlen);
    double stage[5] = {0, 0, 0, 0, 0};
    int steps = 0;
    if (!rc) rc = fddm_plan_timing_end(p, stage, &steps, err, errlen);
    if (!rc) {
      const double n = std::max(steps, 1);
      *f0_ms = std::min(*f0_ms, stage[0] / n);
      *row_ms = std::min(*row_ms, stage[1] / n);
      *col_ms = std::min(*col_ms, stage[2] / n);
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(buf);
  fddm_plan_destroy(p);
  return rc;
}

// The choice for the shape of it: for each frame-block size (16, 8, 4) the row and
// the column kernel form are chosen apart (the generic form must win by more
// than 3 %), then the block size with the shortest spatial stage (a smaller
// block must win by more than 3 %, so noise does not pick more, shorter runs).
int tuned_choice(const fddm_batch_item& it, hipStream_t s, TuneChoice* out, char* err, size_t errlen) {
  const TuneKey k{it.mode, it.dtype, it.precision, it.out_dtype, it.window != nullptr, it.height, it.width, it.nx,
                  it.ny};
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto f = g_tuned.find(k);
    if (f != g_tuned.end()) {
      *out = f->second;
      return FDDM_OK;
    }
  }
  const uint64_t frames = std::min<uint64_t>(it.length, kTuneFrames);
  TuneChoice best;
  double best_ms = 0;
  for (int tb : {16, 8, 4}) {
    double f0 = 0, rows = 0, cols = 0, f0g = 0, rowsg = 0, colsg = 0;
    int rows_mr = 0, cols_mr = 0, r2 = 0, c2 = 0;
    char e2[256];
    if (time_candidate(it, frames, tb, 0, 0, s, &f0, &rows, &cols, &rows_mr, &cols_mr, e2, sizeof e2) != FDDM_OK)
      continue;  // does not fit: skip
    TuneChoice c;
    c.tb = tb;
    if ((rows_mr || cols_mr) &&
        time_candidate(it, frames, tb, 1, 1, s, &f0g, &rowsg, &colsg, &r2, &c2, e2, sizeof e2) == FDDM_OK) {
      if (rows_mr && rowsg < 0.97 * rows) {
        c.rows_generic = 1;
        rows = rowsg;
      }
      if (cols_mr && colsg < 0.97 * cols) {
        c.cols_generic = 1;
        cols = colsg;
      }
    }
    const double ms = f0 + rows + cols;
    if (!best.tb || ms < 0.97 * best_ms) {
      best = c;
      best_ms = ms;
    }
  }
  if (!best.tb) return set_err(err, errlen, FDDM_ERR_SIZE, "no frame-block size fits %llux%llu",
                               (unsigned long long)it.height, (unsigned long long)it.width);
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tuned[k] = best;
  *out = best;
  return FDDM_OK;
}

}  // namespace

struct fddm_batch {
  std::vector<fddm_batch_item> items;
  std::vector<fddm_plan*> plans;
  std::vector<void*> ws;
  void* ws_base = nullptr;
  std::vector<hipStream_t> branches;
  std::vector<hipEvent_t> events;  // [0] fork, [1..] joins
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    for (hipStream_t s : branches) (void)hipStreamDestroy(s);
    for (fddm_plan* p : plans) fddm_plan_destroy(p);
    if (ws_base) (void)hipFree(ws_base);
    exec = nullptr;
    graph = nullptr;
    events.clear();
    branches.clear();
    plans.clear();
    ws_base = nullptr;
  }
};

extern "C" {

int fddm_batch_create(fddm_batch** out, const fddm_batch_item* items, int n, int autotune, char* err,
                      size_t errlen) {
  if (!out || !items || n < 1) return set_err(err, errlen, FDDM_ERR_ARG, "batch needs at least one item");
  *out = nullptr;
  auto* b = new (std::nothrow) fddm_batch();
  if (!b) return set_err(err, errlen, FDDM_ERR_MEMORY, "host allocation failed");
  b->items.assign(items, items + n);
  auto fail = [&](int rc) {
    b->release();
    delete b;
    return rc;
  };
  hipStream_t tune_stream = nullptr;
  if (hipStreamCreateWithFlags(&tune_stream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_err(err, errlen, FDDM_ERR_HIP, "stream creation failed"));
  uint64_t total = 0;
  // each image shape is tuned on its longest stack (a short first stack
  // would favour small blocks for all of them)
  std::map<TuneKey, int> longest;
  for (int i = 0; i < n; ++i) {
    const fddm_batch_item& it = items[i];
    if (!it.d_img || !it.d_out) {
      (void)hipStreamDestroy(tune_stream);
      return fail(set_err(err, errlen, FDDM_ERR_ARG, "item %d: NULL device pointer", i));
    }
    const TuneKey k{it.mode, it.dtype, it.precision, it.out_dtype, it.window != nullptr, it.height, it.width, it.nx,
                    it.ny};
    auto f = longest.find(k);
    if (f == longest.end() || items[f->second].length < it.length) longest[k] = i;
  }
  for (int i = 0; i < n; ++i) {
    const fddm_batch_item& it = items[i];
    TuneChoice choice;
    if (autotune) {
      const TuneKey k{it.mode, it.dtype, it.precision, it.out_dtype, it.window != nullptr, it.height, it.width, it.nx,
                      it.ny};
      const int rc = tuned_choice(items[longest[k]], tune_stream, &choice, err, errlen);
      if (rc) {
        (void)hipStreamDestroy(tune_stream);
        return fail(rc);
      }
    }
    fddm_plan* p = nullptr;
    const int rc = fddm::plan_create_ex(&p, it.mode, it.dtype, it.length, it.height, it.width, it.lags,
                                        it.num_lags, it.nx, it.ny, it.nt, it.window, it.window_f32, it.precision,
                                        it.out_dtype, choice.tb, err, errlen);
    if (rc) {
      (void)hipStreamDestroy(tune_stream);
      return fail(rc);
    }
    fddm::plan_set_generic_spatial(p, choice.rows_generic, choice.cols_generic);
    b->plans.push_back(p);
    uint64_t ws = 0;
    fddm_plan_workspace_size(p, &ws);
    total += align256(ws);
  }
  (void)hipStreamDestroy(tune_stream);
  if (hipMalloc(&b->ws_base, std::max<uint64_t>(total, 256)) != hipSuccess)
    return fail(set_err(err, errlen, FDDM_ERR_MEMORY, "batch workspace of %llu bytes", (unsigned long long)total));
  uint64_t off = 0;
  for (fddm_plan* p : b->plans) {
    uint64_t ws = 0;
    fddm_plan_workspace_size(p, &ws);
    b->ws.push_back(static_cast<unsigned char*>(b->ws_base) + off);
    off += align256(ws);
  }

  // capture: fork on the origin stream, one branch stream per item (up to
  // 16, reused round-robin), join back
  const int nb = std::min(n, 16);
  hipStream_t origin = nullptr;
  if (hipStreamCreateWithFlags(&origin, hipStreamNonBlocking) != hipSuccess)
    return fail(set_err(err, errlen, FDDM_ERR_HIP, "stream creation failed"));
  b->branches.push_back(origin);
  for (int i = 0; i < nb; ++i) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
      return fail(set_err(err, errlen, FDDM_ERR_HIP, "stream creation failed"));
    b->branches.push_back(s);
  }
  for (int i = 0; i <= nb; ++i) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return fail(set_err(err, errlen, FDDM_ERR_HIP, "event creation failed"));
    b->events.push_back(e);
  }
  if (hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(set_err(err, errlen, FDDM_ERR_HIP, "graph capture could not start"));
  int rc = FDDM_OK;
  hipError_t ge = hipEventRecord(b->events[0], origin);
  for (int j = 0; j < nb && ge == hipSuccess; ++j) ge = hipStreamWaitEvent(b->branches[1 + j], b->events[0], 0);
  for (int i = 0; i < n && !rc && ge == hipSuccess; ++i) {
    hipStream_t s = b->branches[1 + (i % nb)];
    rc = fddm_plan_execute(b->plans[i], items[i].d_img, items[i].d_out, b->ws[i], s, err, errlen);
  }
  for (int j = 0; j < nb && ge == hipSuccess; ++j) {
    ge = hipEventRecord(b->events[1 + j], b->branches[1 + j]);
    if (ge == hipSuccess) ge = hipStreamWaitEvent(origin, b->events[1 + j], 0);
  }
  const hipError_t ce = hipStreamEndCapture(origin, &b->graph);
  if (rc) return fail(rc);
  if (ge != hipSuccess)
    return fail(set_err(err, errlen, FDDM_ERR_HIP, "graph capture fork/join failed: %s", hipGetErrorString(ge)));
  if (ce != hipSuccess) return fail(set_err(err, errlen, FDDM_ERR_HIP, "graph capture failed: %s",
                                            hipGetErrorString(ce)));
  if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
    return fail(set_err(err, errlen, FDDM_ERR_HIP, "graph instantiation failed"));
  *out = b;
  return FDDM_OK;
}

int fddm_batch_launch(fddm_batch* b, void* stream, char* err, size_t errlen) {
  if (!b || !b->exec) return set_err(err, errlen, FDDM_ERR_ARG, "NULL batch");
  if (hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)) != hipSuccess)
    return set_err(err, errlen, FDDM_ERR_HIP, "graph launch failed");
  return FDDM_OK;
}

int fddm_batch_launch_eager(fddm_batch* b, void* stream, char* err, size_t errlen) {
  if (!b) return set_err(err, errlen, FDDM_ERR_ARG, "NULL batch");
  for (size_t i = 0; i < b->plans.size(); ++i) {
    const int rc = fddm_plan_execute(b->plans[i], b->items[i].d_img, b->items[i].d_out, b->ws[i],
                                     static_cast<hipStream_t>(stream), err, errlen);
    if (rc) return rc;
  }
  return FDDM_OK;
}

int fddm_batch_frames_per_block(const fddm_batch* b, int* tb) {
  if (!b || !tb) return FDDM_ERR_ARG;
  for (size_t i = 0; i < b->plans.size(); ++i) tb[i] = fddm::plan_frames_per_block(b->plans[i]);
  return FDDM_OK;
}

int fddm_batch_spatial_kernels(const fddm_batch* b, int* rows_mr, int* cols_mr) {
  if (!b || !rows_mr || !cols_mr) return FDDM_ERR_ARG;
  for (size_t i = 0; i < b->plans.size(); ++i) fddm::plan_spatial_forms(b->plans[i], &rows_mr[i], &cols_mr[i]);
  return FDDM_OK;
}

int fddm_batch_destroy(fddm_batch* b) {
  if (!b) return FDDM_OK;
  b->release();
  delete b;
  return FDDM_OK;
}

}  // extern "C"
