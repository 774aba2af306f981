// C ABI and host orchestration of the image structure function engine.
//
// One plan holds the sizes, the mixed-radix factorisations, device twiddle
// tables, lags and window; execute runs every stage back to back on one
// stream with no host synchronisation:
//   [frame-0 spectra] -> row pass -> column pass (q-major) -> time / diff kernel.
// The host-buffer entry points (fddm_ddm_fft / fddm_ddm_diff) wrap a plan with
// cached device buffers, one H2D copy of the raw pixels and one D2H copy of
// the result.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "fddm_copy_pool.hpp"
#include "fddm_hip.h"
#include "fddm_kernels.hpp"
#include "fddm_plan.hpp"

using namespace fddm;

struct fddm_plan {
  int mode = MODE_FFT;
  int precision = PREC_FP64;
  int in_dtype = IN_U16;
  int out_dtype = OUT_F64;
  uint64_t L = 0, H = 0, W = 0, nx = 0, ny = 0, nt = 0, nxh = 0, nq = 0, nl = 0;
  int Tb = 2, Gx = 1, Gy = 1;
  int logn_t = 0;  // time transform length 2^logn_t (fft mode)
  bool long_t = false;  // four-step time transform with device scratch (fddm_time_long.hip)
  int long_slots = 0;
  uint64_t long_slot = 0, bytes_long = 0;
  int logn_x = 0;  // > 0: power-of-two row kernel
  int logn_y = 0;  // > 0: power-of-two column kernel
  // 1: the generic workgroup row / column kernel even where the mixed-radix
  // one applies (set by the batch autotuner)
  int rows_generic = 0, cols_generic = 0;
  uint64_t Lp = 0, nblocks = 0;
  bool subtract0 = false;
  FftDesc dx{}, dy{}, dt{};
  void* d_lags = nullptr;
  void* d_window = nullptr;
  void* d_twx = nullptr;
  void* d_twy = nullptr;
  void* d_twt = nullptr;
  uint64_t bytes_R = 0, bytes_S = 0, bytes_S0 = 0, bytes_flags = 0, ws_bytes = 0;
  // Compensated FP32: child FP64 plan over the first and last kEdgeFrames
  // frames (gathered into its own input unless the stack is that short); the
  // lags from index edge_i0 on have at most kEdgeFrames pairs.
  fddm_plan* edge = nullptr;
  uint64_t edge_frames = 0, bytes_edge_in = 0, bytes_edge_ws = 0;
  int edge_i0 = 0;
  // timing
  std::vector<hipEvent_t> events;
  int timing_cap = 0;
  int timing_steps = 0;
};

namespace {

constexpr int kStages = 5;
constexpr int kEventsPerStep = kStages + 1;
constexpr size_t kLdsLimit = 160 * 1024 - 1024;

int set_err(char* err, size_t len, int code, const char* fmt, ...) {
  if (err && len) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, len, fmt, ap);
    va_end(ap);
  }
  return code;
}

int hip_err(char* err, size_t len, hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return set_err(err, len, FDDM_ERR_MEMORY, "Not enough space on GPU (%s: %s)", what, hipGetErrorString(e));
  return set_err(err, len, FDDM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define FDDM_HIP(expr)                                          \
  do {                                                          \
    hipError_t e_ = (expr);                                     \
    if (e_ != hipSuccess) return hip_err(err, errlen, e_, #expr); \
  } while (0)

uint64_t align256(uint64_t b) { return (b + 255) & ~uint64_t(255); }

size_t in_bytes(int dtype) {
  switch (dtype) {
    case IN_U8: return 1;
    case IN_I16:
    case IN_U16: return 2;
    case IN_I32:
    case IN_U32:
    case IN_F32: return 4;
    default: return 8;
  }
}

// Radix plan: 2^a as 8s (two 4s or one 4 for the remainder), then 3s, 5s, 7s.
bool factorize(uint64_t n, FftDesc& d) {
  std::memset(&d, 0, sizeof(d));
  if (n == 0 || n > (1u << 20)) return false;
  d.n = static_cast<int>(n);
  uint64_t r = n;
  int a = 0;
  while (r % 2 == 0) {
    r /= 2;
    ++a;
  }
  d.c8 = a / 3;
  const int rem = a % 3;
  if (rem == 1 && d.c8 > 0) {
    --d.c8;
    d.c4 = 2;
  } else if (rem == 1) {
    d.c2 = 1;
  } else if (rem == 2) {
    d.c4 = 1;
  }
  while (r % 3 == 0) {
    r /= 3;
    ++d.c3;
  }
  while (r % 5 == 0) {
    r /= 5;
    ++d.c5;
  }
  while (r % 7 == 0) {
    r /= 7;
    ++d.c7;
  }
  return r == 1;
}

// Per-pass twiddle tables in execution order (radix 8s, 4s, 2s, 3s, 5s, 7s):
// pass with radix R after Ns points stores exp(-2 pi i r k / (Ns R)) at
// k (R-1) + r - 1 for k < Ns, r = 1 .. R-1; n - 1 entries in all.  Evaluated
// in long double and rounded once.
template <typename T>
std::vector<T> twiddles(const FftDesc& d) {
  std::vector<T> w;
  w.reserve(2 * static_cast<size_t>(d.n));
  const long double two_pi = 6.283185307179586476925286766559005768L;
  const int counts[6] = {d.c8, d.c4, d.c2, d.c3, d.c5, d.c7};
  const int radix[6] = {8, 4, 2, 3, 5, 7};
  long long ns = 1;
  for (int g = 0; g < 6; ++g)
    for (int i = 0; i < counts[g]; ++i) {
      const int R = radix[g];
      for (long long k = 0; k < ns; ++k)
        for (int r = 1; r < R; ++r) {
          const long double ang = -two_pi * static_cast<long double>(r * k) / static_cast<long double>(ns * R);
          w.push_back(static_cast<T>(cosl(ang)));
          w.push_back(static_cast<T>(sinl(ang)));
        }
      ns *= R;
    }
  if (w.empty()) w.assign(2, T(0));
  return w;
}

// Twiddles of the power-of-two time plans: passes 1 .. np-1 of p2_plan, then
// of p2_half_plan (length N/2), each [k][r] = exp(-2 pi i r k / (NS R)), then
// (cos, sin)(2 pi k / N) for k < N/2.
template <typename T>
std::vector<T> p2_twiddles(int logn) {
  const bool f64 = sizeof(T) == 8;
  const P2Plan P = p2_plan(logn, f64), Q = p2_half_plan(logn, f64);
  std::vector<T> w;
  const long double two_pi = 6.283185307179586476925286766559005768L;
  auto pass = [&](long long R, long long NS) {
    for (long long k = 0; k < NS; ++k)
      for (long long r = 0; r < R; ++r) {
        const long double ang = -two_pi * static_cast<long double>(r * k) / static_cast<long double>(NS * R);
        w.push_back(static_cast<T>(cosl(ang)));
        w.push_back(static_cast<T>(sinl(ang)));
      }
  };
  for (int i = 1; i < P.np; ++i) pass(p2_radix(P, i), p2_ns(P, i));
  for (int i = 1; i < Q.np; ++i) pass(p2_radix(Q, i), p2_ns(Q, i));
  const long long n = 1ll << logn;
  for (long long k = 0; k < n / 2; ++k) {
    const long double ang = two_pi * static_cast<long double>(k) / static_cast<long double>(n);
    w.push_back(static_cast<T>(cosl(ang)));
    w.push_back(static_cast<T>(sinl(ang)));
  }
  return w;
}

// Twiddles of a three-pass power-of-two spatial plan: passes 1 and 2,
// [k][r] = exp(-2 pi i r k / (NS R)).
template <typename T>
std::vector<T> sp_twiddles(int r0, int r1, int r2) {
  std::vector<T> w;
  const long double two_pi = 6.283185307179586476925286766559005768L;
  auto pass = [&](long long R, long long NS) {
    for (long long k = 0; k < NS; ++k)
      for (long long r = 0; r < R; ++r) {
        const long double ang = -two_pi * static_cast<long double>(r * k) / static_cast<long double>(NS * R);
        w.push_back(static_cast<T>(cosl(ang)));
        w.push_back(static_cast<T>(sinl(ang)));
      }
  };
  pass(r1, r0);
  pass(r2, static_cast<long long>(r0) * r1);
  return w;
}

int ilog2_exact(uint64_t n) {
  if (n == 0 || (n & (n - 1))) return 0;
  int l = 0;
  while ((1ull << l) < n) ++l;
  return l;
}

int upload(void** dst, const void* src, size_t bytes, char* err, size_t errlen) {
  FDDM_HIP(hipMalloc(dst, bytes ? bytes : 1));
  if (bytes) FDDM_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
  return FDDM_OK;
}

int upload_twiddles(void** dst, const FftDesc& d, bool f64, char* err, size_t errlen) {
  if (f64) {
    auto w = twiddles<double>(d);
    return upload(dst, w.data(), w.size() * sizeof(double), err, errlen);
  }
  auto w = twiddles<float>(d);
  return upload(dst, w.data(), w.size() * sizeof(float), err, errlen);
}

void plan_free_device(fddm_plan* p) {
  if (p->edge) {
    plan_free_device(p->edge);
    delete p->edge;
    p->edge = nullptr;
  }
  for (void* ptr : {p->d_lags, p->d_window, p->d_twx, p->d_twy, p->d_twt})
    if (ptr) (void)hipFree(ptr);
  p->d_lags = p->d_window = p->d_twx = p->d_twy = p->d_twt = nullptr;
  for (hipEvent_t e : p->events) (void)hipEventDestroy(e);
  p->events.clear();
}

}  // namespace

int fddm::plan_frames_per_block(const fddm_plan* p) { return p ? p->Tb : 0; }
uint64_t fddm::plan_spectra_bytes(const fddm_plan* p) { return p ? p->bytes_S : 0; }

// frames_per_block: 0 picks the default block size; 2, 4, 8 or 16 forces it
// (the autotuner of fddm_batch.hip tries several).
int fddm::plan_create_ex(fddm_plan** out, int mode, int dtype, uint64_t L, uint64_t H, uint64_t W,
                         const uint32_t* lags, uint64_t nl, uint64_t nx, uint64_t ny, uint64_t nt,
                         const void* window, int window_f32, int precision, int out_dtype, int frames_per_block,
                         char* err, size_t errlen) {
  if (!out) return set_err(err, errlen, FDDM_ERR_ARG, "plan pointer is NULL");
  *out = nullptr;
  if (mode != MODE_FFT && mode != MODE_DIFF) return set_err(err, errlen, FDDM_ERR_ARG, "unknown mode %d", mode);
  if (dtype < IN_U8 || dtype > IN_F64) return set_err(err, errlen, FDDM_ERR_ARG, "unknown input dtype %d", dtype);
  if (precision < PREC_FP64 || precision > PREC_FP32C)
    return set_err(err, errlen, FDDM_ERR_ARG, "unknown precision %d", precision);
  if (out_dtype != OUT_F64 && out_dtype != OUT_F32)
    return set_err(err, errlen, FDDM_ERR_ARG, "unknown output dtype %d", out_dtype);
  if (L == 0 || H == 0 || W == 0) return set_err(err, errlen, FDDM_ERR_ARG, "empty image stack");
  if (nx < W || ny < H) return set_err(err, errlen, FDDM_ERR_ARG, "nx, ny must be >= image width, height");
  if (nl > 0 && !lags) return set_err(err, errlen, FDDM_ERR_ARG, "lags pointer is NULL");
  for (uint64_t i = 0; i < nl; ++i) {
    if (lags[i] < 1 || lags[i] > L - 1)
      return set_err(err, errlen, FDDM_ERR_ARG, "lag %u outside [1, %llu]", lags[i],
                     static_cast<unsigned long long>(L - 1));
    if (i > 0 && lags[i] <= lags[i - 1])
      return set_err(err, errlen, FDDM_ERR_ARG, "lags must be sorted and unique");
  }
  if (L > (1ull << 30) || nl > (1ull << 30)) return set_err(err, errlen, FDDM_ERR_SIZE, "too many frames");

  auto* p = new (std::nothrow) fddm_plan();
  if (!p) return set_err(err, errlen, FDDM_ERR_MEMORY, "host allocation failed");
  p->mode = mode;
  p->precision = precision;
  p->in_dtype = dtype;
  p->out_dtype = out_dtype;
  p->L = L;
  p->H = H;
  p->W = W;
  p->nx = nx;
  p->ny = ny;
  p->nt = mode == MODE_FFT ? nt : 0;
  p->nxh = nx / 2 + 1;
  p->nq = p->nxh * ny;
  p->nl = nl;
  p->subtract0 = precision != PREC_FP32;
  const bool f64 = precision == PREC_FP64;
  const size_t csz = f64 ? 16 : 8;

  auto fail = [&](int code) {
    plan_free_device(p);
    delete p;
    return code;
  };
  if (!factorize(nx, p->dx) || !factorize(ny, p->dy))
    return fail(set_err(err, errlen, FDDM_ERR_SIZE, "nx=%llu, ny=%llu must be 2^a 3^b 5^c 7^d",
                        (unsigned long long)nx, (unsigned long long)ny));
  if (p->nq > (1ull << 31)) return fail(set_err(err, errlen, FDDM_ERR_SIZE, "image too large"));
  // Frames per block: runs of Tb * sizeof(complex) = 128 B in the q-major
  // store, halved while a column slab (ny x Tb complex) exceeds 80 KiB (two
  // slabs per CU).
  if (frames_per_block) {
    if (frames_per_block != 2 && frames_per_block != 4 && frames_per_block != 8 && frames_per_block != 16)
      return fail(set_err(err, errlen, FDDM_ERR_ARG, "frames per block must be 2, 4, 8 or 16"));
    p->Tb = frames_per_block;
  } else {
    p->Tb = f64 ? 8 : 16;
    while (p->Tb > 2 && lds_padded_size(static_cast<int>(ny) * p->Tb) * csz > 80 * 1024) p->Tb /= 2;
  }
  if (lds_padded_size(static_cast<int>(ny) * p->Tb) * csz > kLdsLimit)
    return fail(set_err(err, errlen, FDDM_ERR_SIZE, "ny=%llu too large for the column pass", (unsigned long long)ny));
  // Transforms handled at once by the row / column kernels (<= 16 or 32
  // elements per thread).
  const uint64_t span = 4096;
  p->Gx = 1;
  while (p->Gx * 2 <= p->Tb / 2 && (p->Gx * 2) * nx <= span) p->Gx *= 2;
  p->Gy = 1;
  while (p->Gy * 2 <= p->Tb && (p->Gy * 2) * ny <= span) p->Gy *= 2;
  if (fft_bucket(static_cast<long long>(p->Gx) * nx) == 0 || fft_bucket(p->Gx * nx) > 32 ||
      fft_bucket(static_cast<long long>(p->Gy) * ny) > 32 || fft_bucket(p->Gy * ny) == 0)
    return fail(set_err(err, errlen, FDDM_ERR_SIZE, "nx or ny above 8192"));
  if (lds_padded_size(static_cast<int>(p->Gx * (nx + 1))) * csz > kLdsLimit)
    return fail(set_err(err, errlen, FDDM_ERR_SIZE, "nx=%llu too large for the row pass", (unsigned long long)nx));
  p->nblocks = (L + p->Tb - 1) / p->Tb;
  p->Lp = p->nblocks * p->Tb;

  if (mode == MODE_FFT) {
    // The caller's nt (>= 2L - 1) fixes no result: D is the same for every
    // zero-padded length >= 2L - 1.  The time kernel runs at the power of two
    // 2^logn_t >= max(2L - 1, 256).
    if (nt < 2 * L - 1) return fail(set_err(err, errlen, FDDM_ERR_ARG, "nt must be >= 2*length-1"));
    int logn = kP2MinLog;
    while ((1ull << logn) < 2 * L - 1) ++logn;
    const int maxlog = f64 ? kP2MaxLog - 1 : kP2MaxLog;
    if (logn > maxlog || p2_lds_bytes(logn, precision) > kLdsLimit) {
      // longer series: four-step transform through device scratch
      const int lmax = long_max_log(precision);
      if (logn > lmax)
        return fail(set_err(err, errlen, FDDM_ERR_SIZE, "nt=%llu too long for the time kernels (at most %d frames in %s)",
                            (unsigned long long)nt, 1 << (lmax - 1), f64 ? "fp64" : "fp32"));
      p->long_t = true;
    }
    p->logn_t = logn;
  }

  // Power-of-two spatial kernels where the size and frame block allow them.
  {
    const int lx = ilog2_exact(nx), ly = ilog2_exact(ny);
    if (p2_spatial_ok(lx) && (p->Tb == 4 || p->Tb == 8 || p->Tb == 16) &&
        lds_padded_size((p->Tb / 2) * static_cast<int>(nx)) * csz <= 96 * 1024)
      p->logn_x = lx;
    if (p2_spatial_ok(ly) && p->Tb >= 4) p->logn_y = ly;
  }
  int rc;
  auto upload_sp = [&](void** dst, int logn, bool cols) {
    int r0, r1, r2;
    p2_spatial_plan(logn, cols, &r0, &r1, &r2);
    if (f64) {
      auto w = sp_twiddles<double>(r0, r1, r2);
      return upload(dst, w.data(), w.size() * sizeof(double), err, errlen);
    }
    auto w = sp_twiddles<float>(r0, r1, r2);
    return upload(dst, w.data(), w.size() * sizeof(float), err, errlen);
  };
  if ((rc = p->logn_x ? upload_sp(&p->d_twx, p->logn_x, false) : upload_twiddles(&p->d_twx, p->dx, f64, err, errlen)))
    return fail(rc);
  if ((rc = p->logn_y ? upload_sp(&p->d_twy, p->logn_y, true) : upload_twiddles(&p->d_twy, p->dy, f64, err, errlen)))
    return fail(rc);
  if (mode == MODE_FFT && p->long_t) {
    const std::vector<double> wd = long_twiddles(p->logn_t);
    if (f64) {
      rc = upload(&p->d_twt, wd.data(), wd.size() * sizeof(double), err, errlen);
    } else {
      std::vector<float> wf(wd.begin(), wd.end());
      rc = upload(&p->d_twt, wf.data(), wf.size() * sizeof(float), err, errlen);
    }
    if (rc) return fail(rc);
  } else if (mode == MODE_FFT) {
    if (f64) {
      auto w = p2_twiddles<double>(p->logn_t);
      rc = upload(&p->d_twt, w.data(), w.size() * sizeof(double), err, errlen);
    } else {
      auto w = p2_twiddles<float>(p->logn_t);
      rc = upload(&p->d_twt, w.data(), w.size() * sizeof(float), err, errlen);
    }
    if (rc) return fail(rc);
  }
  if ((rc = upload(&p->d_lags, lags, nl * sizeof(uint32_t), err, errlen))) return fail(rc);
  if (window) {
    const size_t npx = H * W;
    if (f64) {
      std::vector<double> w(npx);
      for (size_t i = 0; i < npx; ++i)
        w[i] = window_f32 ? static_cast<const float*>(window)[i] : static_cast<const double*>(window)[i];
      rc = upload(&p->d_window, w.data(), npx * sizeof(double), err, errlen);
    } else {
      std::vector<float> w(npx);
      for (size_t i = 0; i < npx; ++i)
        w[i] = window_f32 ? static_cast<const float*>(window)[i]
                          : static_cast<float>(static_cast<const double*>(window)[i]);
      rc = upload(&p->d_window, w.data(), npx * sizeof(float), err, errlen);
    }
    if (rc) return fail(rc);
  }
  p->bytes_R = align256(p->nblocks * p->nxh * H * p->Tb * csz);
  p->bytes_S = align256(p->nq * p->Lp * csz);
  p->bytes_S0 = p->subtract0 ? align256(p->nq * p->Tb * csz) : 0;
  // fft mode, compensated: one byte per q marks lags for the direct recomputation
  p->bytes_flags = (mode == MODE_FFT && precision == PREC_FP32C) ? align256(p->nq) : 0;
  if (p->long_t) {
    // one scratch slot per resident workgroup of the four-step kernel
    p->long_slots = static_cast<int>(std::min<uint64_t>(p->nq, f64 ? 512 : 1024));
    p->long_slot = long_slot_bytes(p->logn_t, precision, L);
    p->bytes_long = align256(p->long_slot * static_cast<uint64_t>(p->long_slots));
  }
  p->ws_bytes = p->bytes_R + p->bytes_S + p->bytes_S0 + p->bytes_flags + p->bytes_long;
  if (precision == PREC_FP32C && nl > 0 && L - lags[nl - 1] <= static_cast<uint64_t>(kEdgeFrames)) {
    // FP64 spectra of the edge frames for the lags with few pairs
    const uint64_t E = kEdgeFrames;
    p->edge_frames = std::min<uint64_t>(L, 2 * E);
    if ((rc = plan_create_ex(&p->edge, MODE_DIFF, dtype, p->edge_frames, H, W, nullptr, 0, nx, ny, 0, window,
                             window_f32, PREC_FP64, OUT_F64, 0, err, errlen)))
      return fail(rc);
    p->bytes_edge_in = L > 2 * E ? align256(2 * E * H * W * in_bytes(dtype)) : 0;
    p->bytes_edge_ws = p->edge->ws_bytes;
    p->ws_bytes += p->bytes_edge_in + p->bytes_edge_ws;
    p->edge_i0 = static_cast<int>(nl);
    while (p->edge_i0 > 0 && L - lags[p->edge_i0 - 1] <= E) --p->edge_i0;
  }
  *out = p;
  return FDDM_OK;
}

namespace {

uint64_t output_bytes(const fddm_plan* p) {
  return (p->nl + 2) * p->nq * (p->out_dtype == OUT_F32 ? 4 : 8);
}

int record(fddm_plan* p, int slot, hipStream_t s, char* err, size_t errlen) {
  if (p->timing_cap == 0 || p->timing_steps >= p->timing_cap) return FDDM_OK;
  FDDM_HIP(hipEventRecord(p->events[p->timing_steps * kEventsPerStep + slot], s));
  return FDDM_OK;
}

int run_spatial(fddm_plan* p, const void* d_img, void* ws, hipStream_t s, char* err, size_t errlen) {
  auto* base = static_cast<unsigned char*>(ws);
  void* R = base;
  void* S = base + p->bytes_R;
  void* S0 = base + p->bytes_R + p->bytes_S;
  const double norm = 1.0 / std::sqrt(static_cast<double>(p->nx) * static_cast<double>(p->ny));
  RowArgs ra{};
  ra.img = d_img;
  ra.img0 = d_img;
  ra.in_dtype = p->in_dtype;
  ra.L = p->L;
  ra.H = static_cast<int>(p->H);
  ra.W = static_cast<int>(p->W);
  ra.nx = static_cast<int>(p->nx);
  ra.Tb = p->Tb;
  ra.G = p->Gx;
  ra.window = p->d_window;
  ra.logn = p->logn_x;
  ra.norm = norm;
  ra.dx = p->dx;
  ra.twx = p->d_twx;
  ra.R = R;
  ra.generic = p->rows_generic;
  ColArgs ca{};
  ca.R = R;
  ca.H = static_cast<int>(p->H);
  ca.ny = static_cast<int>(p->ny);
  ca.nxh = static_cast<int>(p->nxh);
  ca.Tb = p->Tb;
  ca.G = p->Gy;
  ca.logn = p->logn_y;
  ca.dy = p->dy;
  ca.twy = p->d_twy;
  ca.generic = p->cols_generic;
  int rc;
  if ((rc = record(p, 0, s, err, errlen))) return rc;
  if (p->subtract0) {
    // F(0) of the raw first frame, for the power spectrum.
    ra.frames = 1;
    ra.subtract0 = 0;
    FDDM_HIP(ra.logn ? launch_rows_p2(ra, p->precision, s) : launch_rows(ra, p->precision, s));
    ca.nblocks = 1;
    ca.Lp = static_cast<uint64_t>(p->Tb);
    ca.S = S0;
    FDDM_HIP(ca.logn ? launch_cols_p2(ca, p->precision, s) : launch_cols(ca, p->precision, s));
  }
  if ((rc = record(p, 1, s, err, errlen))) return rc;
  ra.frames = p->L;
  ra.subtract0 = p->subtract0 ? 1 : 0;
  FDDM_HIP(ra.logn ? launch_rows_p2(ra, p->precision, s) : launch_rows(ra, p->precision, s));
  if ((rc = record(p, 2, s, err, errlen))) return rc;
  ca.nblocks = p->nblocks;
  ca.Lp = p->Lp;
  ca.S = S;
  FDDM_HIP(ca.logn ? launch_cols_p2(ca, p->precision, s) : launch_cols(ca, p->precision, s));
  if ((rc = record(p, 3, s, err, errlen))) return rc;
  return FDDM_OK;
}

// The kernels read pixels in 8-byte vectors and write spectra in 16-byte
// vectors; the output is written per element.
int check_alignment(const fddm_plan* p, const void* d_img, const void* d_out, const void* ws, char* err,
                    size_t errlen) {
  const uintptr_t out_elem = p->out_dtype == OUT_F32 ? 4 : 8;
  if (reinterpret_cast<uintptr_t>(d_img) % 8 || (ws && reinterpret_cast<uintptr_t>(ws) % 16) ||
      (d_out && reinterpret_cast<uintptr_t>(d_out) % out_elem))
    return set_err(err, errlen, FDDM_ERR_ARG,
                   "device buffers must be aligned: input to 8 bytes, workspace to 16, output to its element");
  return FDDM_OK;
}

int plan_execute(fddm_plan* p, const void* d_img, void* d_out, void* ws, hipStream_t s, char* err,
                 size_t errlen) {
  if (!p || !d_img || !d_out || !ws) return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument to execute");
  if (const int rc = check_alignment(p, d_img, d_out, ws, err, errlen)) return rc;
  int rc = run_spatial(p, d_img, ws, s, err, errlen);
  if (rc) return rc;
  auto* base = static_cast<unsigned char*>(ws);
  TimeArgs ta{};
  ta.S = base + p->bytes_R;
  ta.Lp = p->Lp;
  ta.L = static_cast<int>(p->L);
  ta.ny = static_cast<int>(p->ny);
  ta.nxh = static_cast<int>(p->nxh);
  ta.S0 = p->subtract0 ? base + p->bytes_R + p->bytes_S : nullptr;
  ta.s0_stride = static_cast<uint64_t>(p->Tb);
  ta.lags = static_cast<const uint32_t*>(p->d_lags);
  ta.nl = static_cast<int>(p->nl);
  ta.dt = p->dt;
  ta.twt = p->d_twt;
  ta.out = d_out;
  ta.flags = p->bytes_flags ? base + p->bytes_R + p->bytes_S + p->bytes_S0 : nullptr;
  time_full(ta);
  if (p->mode == MODE_FFT && p->long_t) {
    FDDM_HIP(launch_time_long(ta, p->logn_t, p->precision, p->out_dtype,
                              base + p->bytes_R + p->bytes_S + p->bytes_S0 + p->bytes_flags, p->long_slot,
                              p->long_slots, s));
  } else if (p->mode == MODE_FFT) {
    FDDM_HIP(launch_time_p2(ta, p->logn_t, p->precision, p->out_dtype, s));
  } else {
    FDDM_HIP(launch_diff(ta, p->precision, p->out_dtype, s));
  }
  if ((rc = record(p, 4, s, err, errlen))) return rc;
  // compensated FP32 corrections: flagged Wiener-Khinchin lags, then the lags
  // with at most kEdgeFrames pairs from FP64 spectra
  if (ta.flags) FDDM_HIP(launch_wk_fix(ta, p->out_dtype, s));
  if (p->edge) {
    // FP64 spectra of the edge frames, then the few-pair lags from them
    unsigned char* ein = base + p->ws_bytes - p->bytes_edge_ws - p->bytes_edge_in;
    unsigned char* ews = base + p->ws_bytes - p->bytes_edge_ws;
    const void* src = d_img;
    if (p->bytes_edge_in) {
      FDDM_HIP(launch_gather_edges(d_img, ein, p->H * p->W * in_bytes(p->in_dtype), p->L, kEdgeFrames, s));
      src = ein;
    }
    const int cap = p->edge->timing_cap;
    p->edge->timing_cap = 0;
    rc = run_spatial(p->edge, src, ews, s, err, errlen);
    p->edge->timing_cap = cap;
    if (rc) return rc;
    FDDM_HIP(launch_edge_fix(ta, ews + p->edge->bytes_R, p->edge->Lp, kEdgeFrames, p->edge_i0, p->out_dtype, s));
  }
  if ((rc = record(p, 5, s, err, errlen))) return rc;
  if (p->timing_cap > 0 && p->timing_steps < p->timing_cap) ++p->timing_steps;
  return FDDM_OK;
}

// Device buffers reused across the host-buffer entry points, per device,
// and two pinned staging buffers for the host link.
constexpr size_t kStageBytes = size_t(64) << 20;
struct HostPathCache {
  void* img = nullptr;
  uint64_t img_cap = 0;
  void* out = nullptr;
  uint64_t out_cap = 0;
  void* ws = nullptr;
  uint64_t ws_cap = 0;
  hipStream_t stream = nullptr;
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
};

}  // namespace

// memcpy split over the copy workers (pageable <-> pinned staging); the
// caller's pages are touched by several threads at once.
void fddm::parallel_memcpy(void* dst, const void* src, size_t n) {
  const size_t per = size_t(4) << 20;
  const unsigned parts = static_cast<unsigned>(std::min<size_t>((n + per - 1) / per, 16));
  if (parts <= 1) {
    std::memcpy(dst, src, n);
    return;
  }
  CopyPool& pool = CopyPool::get();
  pool.run(static_cast<char*>(dst), static_cast<const char*>(src), n, std::min(parts, pool.workers() + 1));
}

// Done by a kernel where n and the addresses are multiples of 16, else by the
// DMA engine.  The DMA engine's device-to-host rate fell from about 45 to
// 28 GB/s for source buffers allocated where large buffers had just been freed
// (a result buffer allocated after another analysis's buffers were released);
// the kernel copies ran at 45 GB/s either way.
hipError_t fddm::copy_to_pinned(void* pin, const void* d_src, size_t n, hipStream_t s) {
  void* dp = nullptr;
  if (n % 16 == 0 && reinterpret_cast<uintptr_t>(d_src) % 16 == 0) {
    if (hipHostGetDevicePointer(&dp, pin, 0) == hipSuccess && dp && reinterpret_cast<uintptr_t>(dp) % 16 == 0)
      return launch_copy_out(dp, d_src, n, s);
    (void)hipGetLastError();  // not a device-addressable buffer: the DMA engine copies it
  }
  return hipMemcpyAsync(pin, d_src, n, hipMemcpyDeviceToHost, s);
}

namespace {

// Host -> device through the staging pair: host threads fill one buffer while
// the DMA engine drains the other.
int h2d_staged(HostPathCache& c, void* d_dst, const void* h_src, size_t n, char* err, size_t errlen) {
  size_t off = 0;
  for (int k = 0; off < n; k ^= 1) {
    const size_t m = std::min(kStageBytes, n - off);
    FDDM_HIP(hipEventSynchronize(c.ev[k]));
    parallel_memcpy(c.pin[k], static_cast<const char*>(h_src) + off, m);
    FDDM_HIP(hipMemcpyAsync(static_cast<char*>(d_dst) + off, c.pin[k], m, hipMemcpyHostToDevice, c.stream));
    FDDM_HIP(hipEventRecord(c.ev[k], c.stream));
    off += m;
  }
  return FDDM_OK;
}

// Device -> host: two chunks in flight; each finished chunk is copied out by
// host threads while the next one transfers.
int d2h_staged(HostPathCache& c, void* h_dst, const void* d_src, size_t n, char* err, size_t errlen) {
  const size_t nchunk = (n + kStageBytes - 1) / kStageBytes;
  auto issue = [&](size_t i) -> int {
    const size_t off = i * kStageBytes, m = std::min(kStageBytes, n - off);
    FDDM_HIP(copy_to_pinned(c.pin[i & 1], static_cast<const char*>(d_src) + off, m, c.stream));
    FDDM_HIP(hipEventRecord(c.ev[i & 1], c.stream));
    return FDDM_OK;
  };
  int rc;
  for (size_t i = 0; i < std::min<size_t>(2, nchunk); ++i)
    if ((rc = issue(i))) return rc;
  for (size_t i = 0; i < nchunk; ++i) {
    const size_t off = i * kStageBytes, m = std::min(kStageBytes, n - off);
    FDDM_HIP(hipEventSynchronize(c.ev[i & 1]));
    parallel_memcpy(static_cast<char*>(h_dst) + off, c.pin[i & 1], m);
    if (i + 2 < nchunk && (rc = issue(i + 2))) return rc;
  }
  return FDDM_OK;
}
std::mutex g_cache_mu[64];  // one per device: calls on different devices do not wait for each other
HostPathCache g_cache[64];
// Device bytes of input, output and workspace buffers kept for the next
// host-buffer call on a device (kCacheKeepBytes until set).  Larger sets are
// freed when the call returns.
constexpr uint64_t kCacheKeepBytes = uint64_t(16) << 30;
bool g_cache_limit_set[64] = {};
uint64_t g_cache_limit[64] = {};

uint64_t cache_limit(int dev) { return g_cache_limit_set[dev] ? g_cache_limit[dev] : kCacheKeepBytes; }

// Frees the cached input, output and workspace buffers (staging buffers and
// the stream stay); the caller holds the device's cache lock.
void drop_buffers(HostPathCache& c) {
  for (void** ptr : {&c.img, &c.out, &c.ws})
    if (*ptr) (void)hipFree(*ptr);
  c.img = c.out = c.ws = nullptr;
  c.img_cap = c.out_cap = c.ws_cap = 0;
}

int grow(void** ptr, uint64_t* cap, uint64_t need, char* err, size_t errlen) {
  if (*cap >= need && *ptr) return FDDM_OK;
  if (*ptr) (void)hipFree(*ptr);
  *ptr = nullptr;
  *cap = 0;
  FDDM_HIP(hipMalloc(ptr, need ? need : 1));
  *cap = need;
  return FDDM_OK;
}

// ---- host path within a device-memory budget ----
//
// When the one-shot footprint (input + workspace + output) exceeds the
// budget, the call runs in pieces, as the reference's planner does with its
// loop counts (memchk_gpu.cu:141-702, ddm_cuda.cu:40-261, :535-923):
//   * the spatial stage over chunks of whole frame blocks: each chunk's frames
//     are copied in and transformed (frame 0, the subtracted frame, stays on
//     the device), and its columns land in the q-major spectra at their
//     frame offset -- in device memory (mode 1) or, when the spectra
//     themselves do not fit, in host memory through a device piece (mode 2);
//   * the time stage over chunks of whole ky rows (q ranges), each stored
//     unshifted in a device piece and copied to its y-shifted rows of the
//     host result (two contiguous runs at most).
// Every kernel computes each (frame block, row) and each q exactly as in one
// launch, so the result is bit-identical to the one-shot path.
uint64_t g_budget[64] = {};
uint64_t g_last_stats[64][3] = {};

struct ChunkPlan {
  int mode = 0;                 // 1: spectra on the device, 2: spectra on the host
  uint64_t frames_per = 0;      // frames per spatial chunk (multiple of Tb)
  uint64_t rows_per = 0;        // ky rows per time chunk
  uint64_t fixed = 0, work = 0;  // device bytes
};

bool plan_chunks(const fddm_plan* p, uint64_t budget, ChunkPlan* cp) {
  const uint64_t csz = p->precision == PREC_FP64 ? 16 : 8;
  const uint64_t osz = p->out_dtype == OUT_F32 ? 4 : 8;
  const uint64_t frame_b = p->H * p->W * in_bytes(p->in_dtype);
  const uint64_t block_b = static_cast<uint64_t>(p->Tb) * frame_b + align256(p->nxh * p->H * p->Tb * csz);
  const uint64_t row_out = p->nxh * (p->nl + 2) * osz;
  const uint64_t edge_in = p->edge ? align256(p->edge_frames * frame_b) : 0;
  const uint64_t fixed = align256(frame_b) + p->bytes_S0 + p->bytes_flags + p->bytes_long + edge_in +
                         p->bytes_edge_ws + 8 * 256;
  for (int mode = 1; mode <= 2; ++mode) {
    const uint64_t fx = fixed + (mode == 1 ? p->bytes_S : 0);
    if (fx >= budget) continue;
    const uint64_t avail = budget - fx;
    const uint64_t per_block = block_b + (mode == 2 ? p->nq * p->Tb * csz : 0);
    const uint64_t per_row = row_out + (mode == 2 ? p->nxh * p->Lp * csz : 0);
    const uint64_t nb = std::min<uint64_t>(avail / per_block, p->nblocks);
    const uint64_t rows = std::min<uint64_t>(avail / per_row, p->ny);
    if (nb == 0 || rows == 0) continue;
    cp->mode = mode;
    cp->frames_per = nb * p->Tb;
    cp->rows_per = rows;
    cp->fixed = fx;
    cp->work = std::max(nb * per_block, rows * per_row) + 4 * 256;
    return true;
  }
  return false;
}

RowArgs row_args(const fddm_plan* p) {
  RowArgs ra{};
  ra.in_dtype = p->in_dtype;
  ra.H = static_cast<int>(p->H);
  ra.W = static_cast<int>(p->W);
  ra.nx = static_cast<int>(p->nx);
  ra.Tb = p->Tb;
  ra.G = p->Gx;
  ra.window = p->d_window;
  ra.logn = p->logn_x;
  ra.norm = 1.0 / std::sqrt(static_cast<double>(p->nx) * static_cast<double>(p->ny));
  ra.dx = p->dx;
  ra.twx = p->d_twx;
  ra.generic = p->rows_generic;
  return ra;
}

ColArgs col_args(const fddm_plan* p) {
  ColArgs ca{};
  ca.H = static_cast<int>(p->H);
  ca.ny = static_cast<int>(p->ny);
  ca.nxh = static_cast<int>(p->nxh);
  ca.Tb = p->Tb;
  ca.G = p->Gy;
  ca.logn = p->logn_y;
  ca.dy = p->dy;
  ca.twy = p->d_twy;
  ca.generic = p->cols_generic;
  return ca;
}

// Rows and columns of `frames` frames at img (frame I0 at img0 subtracted when
// sub), blocks of Tb frames into S_dst[q][t] with row pitch `pitch`.
int spatial_frames(const fddm_plan* p, const void* img, const void* img0, uint64_t frames, int sub, void* R,
                   void* S_dst, uint64_t pitch, hipStream_t s, char* err, size_t errlen) {
  RowArgs ra = row_args(p);
  ra.img = img;
  ra.img0 = img0;
  ra.L = frames;
  ra.frames = frames;
  ra.subtract0 = sub;
  ra.R = R;
  FDDM_HIP(ra.logn ? launch_rows_p2(ra, p->precision, s) : launch_rows(ra, p->precision, s));
  ColArgs ca = col_args(p);
  ca.R = R;
  ca.nblocks = (frames + p->Tb - 1) / p->Tb;
  ca.Lp = pitch;
  ca.S = S_dst;
  FDDM_HIP(ca.logn ? launch_cols_p2(ca, p->precision, s) : launch_cols(ca, p->precision, s));
  return FDDM_OK;
}

// Time stage and corrections of one launch range (ta.c0, ta.cols, ...).
int time_launch(const fddm_plan* p, const TimeArgs& ta, void* lng, const void* edge_S64, hipStream_t s, char* err,
                size_t errlen) {
  if (p->mode == MODE_FFT && p->long_t) {
    FDDM_HIP(launch_time_long(ta, p->logn_t, p->precision, p->out_dtype, lng, p->long_slot, p->long_slots, s));
  } else if (p->mode == MODE_FFT) {
    FDDM_HIP(launch_time_p2(ta, p->logn_t, p->precision, p->out_dtype, s));
  } else {
    FDDM_HIP(launch_diff(ta, p->precision, p->out_dtype, s));
  }
  if (ta.flags) FDDM_HIP(launch_wk_fix(ta, p->out_dtype, s));
  if (p->edge) FDDM_HIP(launch_edge_fix(ta, edge_S64, p->edge->Lp, kEdgeFrames, p->edge_i0, p->out_dtype, s));
  return FDDM_OK;
}

}  // namespace

// Device -> pageable host memory.  From 8 MiB up the bytes cross through the
// device's staging pair (the host-buffer calls' pinned buffers and stream,
// whose lock is held for the copy: a host-buffer call on the same device
// waits); smaller copies are left to the runtime.  The device must be current.
int fddm::copy_to_host(int dev, void* h_dst, const void* d_src, size_t n, char* err, size_t errlen) {
  if (dev < 0 || dev >= 64) return set_err(err, errlen, FDDM_ERR_DEVICE, "device id %d not supported", dev);
  if (n < (size_t(8) << 20)) {
    FDDM_HIP(hipMemcpy(h_dst, d_src, n, hipMemcpyDeviceToHost));
    return FDDM_OK;
  }
  std::lock_guard<std::mutex> lk(g_cache_mu[dev]);
  HostPathCache& c = g_cache[dev];
  if (!c.stream) FDDM_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  for (int k = 0; k < 2; ++k) {
    if (!c.pin[k]) FDDM_HIP(hipHostMalloc(&c.pin[k], kStageBytes, hipHostMallocDefault));
    if (!c.ev[k]) FDDM_HIP(hipEventCreateWithFlags(&c.ev[k], hipEventDisableTiming));
  }
  return d2h_staged(c, h_dst, d_src, n, err, errlen);  // returns once the last piece is in h_dst
}

// plan_execute in smaller launches: the spatial stage over chunks of
// frames_per frames (whole frame blocks) into the same workspace, the time
// stage and its corrections over chunks of cols_per output columns.  Each
// launch is short, so kernels of other streams (other requests of the
// server) are dispatched between them; the result is bit-identical.
int fddm::plan_spatial_frames(fddm_plan* p, const void* img, const void* img0, uint64_t t0, uint64_t frames, void* ws,
                              hipStream_t s, char* err, size_t errlen) {
  const uint64_t csz = p->precision == PREC_FP64 ? 16 : 8;
  const uint64_t rblock = p->nxh * p->H * p->Tb * csz;  // R bytes per frame block
  auto* base = static_cast<unsigned char*>(ws);
  return spatial_frames(p, img, img0, frames, p->subtract0 ? 1 : 0, base + t0 / p->Tb * rblock,
                        base + p->bytes_R + t0 * csz, p->Lp, s, err, errlen);
}

void fddm::plan_spatial_forms(const fddm_plan* p, int* rows_mr, int* cols_mr) {
  const RowArgs ra = row_args(p);
  const ColArgs ca = col_args(p);
  *rows_mr = ra.logn == 0 && !ra.generic && rows_mr_ok(ra, p->precision);
  *cols_mr = ca.logn == 0 && !ca.generic && cols_mr_ok(ca, p->precision);
}

void fddm::plan_set_generic_spatial(fddm_plan* p, int rows_generic, int cols_generic) {
  p->rows_generic = rows_generic ? 1 : 0;
  p->cols_generic = cols_generic ? 1 : 0;
}

PlanInfo fddm::plan_info(const fddm_plan* p) {
  PlanInfo i{};
  i.L = p->L;
  i.Tb = p->Tb;
  i.in_bytes = in_bytes(p->in_dtype);
  i.ws_bytes = p->ws_bytes;
  i.out_bytes = output_bytes(p);
  i.edge_frames = p->edge ? p->edge_frames : 0;
  return i;
}

// The stages after the spatial pass over every frame: F(0) from frame0, the
// FP64 edge spectra from edge_in (the child's frames, or gathered from img
// when edge_in is null), then the time stage and its corrections over the
// whole plane in launches of cols_per columns.
int finish_from_spectra(fddm_plan* p, const void* frame0, const void* img, const void* edge_in, void* ws,
                        void* d_out, int cols_per, hipStream_t s, char* err, size_t errlen) {
  auto* base = static_cast<unsigned char*>(ws);
  unsigned char* R = base;
  unsigned char* S = base + p->bytes_R;
  unsigned char* S0 = S + p->bytes_S;
  unsigned char* flags = S0 + p->bytes_S0;
  unsigned char* lng = flags + p->bytes_flags;
  const int nq = static_cast<int>(p->nq);
  const int sub = p->subtract0 ? 1 : 0;
  int rc;
  if (sub && (rc = spatial_frames(p, frame0, frame0, 1, 0, R, S0, p->Tb, s, err, errlen))) return rc;
  const void* S64 = nullptr;
  if (p->edge) {
    unsigned char* ein = base + p->ws_bytes - p->bytes_edge_ws - p->bytes_edge_in;
    unsigned char* ews = base + p->ws_bytes - p->bytes_edge_ws;
    const void* src = edge_in ? edge_in : img;
    if (!edge_in && p->bytes_edge_in) {
      FDDM_HIP(launch_gather_edges(img, ein, p->H * p->W * in_bytes(p->in_dtype), p->L, kEdgeFrames, s));
      src = ein;
    }
    const int cap = p->edge->timing_cap;
    p->edge->timing_cap = 0;
    rc = run_spatial(p->edge, src, ews, s, err, errlen);
    p->edge->timing_cap = cap;
    if (rc) return rc;
    S64 = ews + p->edge->bytes_R;
  }
  TimeArgs ta{};
  ta.S = S;
  ta.Lp = p->Lp;
  ta.L = static_cast<int>(p->L);
  ta.ny = static_cast<int>(p->ny);
  ta.nxh = static_cast<int>(p->nxh);
  ta.S0 = sub ? S0 : nullptr;
  ta.s0_stride = static_cast<uint64_t>(p->Tb);
  ta.lags = static_cast<const uint32_t*>(p->d_lags);
  ta.nl = static_cast<int>(p->nl);
  ta.dt = p->dt;
  ta.twt = p->d_twt;
  ta.out = d_out;
  ta.flags = p->bytes_flags ? flags : nullptr;
  time_full(ta);
  if (cols_per <= 0) cols_per = nq;
  for (int c0 = 0; c0 < nq; c0 += cols_per) {
    ta.c0 = c0;
    ta.cols = std::min(cols_per, nq - c0);
    if ((rc = time_launch(p, ta, lng, S64, s, err, errlen))) return rc;
  }
  return FDDM_OK;
}

int fddm::plan_finish_from_spectra(fddm_plan* p, const void* frame0, const void* edge_in, void* ws, void* d_out,
                                   hipStream_t s, char* err, size_t errlen) {
  return finish_from_spectra(p, frame0, nullptr, edge_in, ws, d_out, 0, s, err, errlen);
}

// plan_execute in smaller launches: the spatial stage over chunks of
// frames_per frames (whole frame blocks) into the same workspace, the time
// stage and its corrections over chunks of cols_per output columns.  Each
// launch is short, so kernels of other streams (other requests of the
// server) are dispatched between them; the result is bit-identical.
int fddm::plan_execute_pieces(fddm_plan* p, const void* d_img, void* d_out, void* ws, hipStream_t s,
                              uint64_t frames_per, int cols_per, char* err, size_t errlen) {
  if (!p || !d_img || !d_out || !ws) return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument to execute");
  if (const int rc = check_alignment(p, d_img, d_out, ws, err, errlen)) return rc;
  const int nq = static_cast<int>(p->nq);
  if ((frames_per == 0 || frames_per >= p->L) && (cols_per <= 0 || cols_per >= nq))
    return plan_execute(p, d_img, d_out, ws, s, err, errlen);
  frames_per = frames_per == 0 ? p->L : std::max<uint64_t>(p->Tb, frames_per / p->Tb * p->Tb);
  const uint64_t frame_b = p->H * p->W * in_bytes(p->in_dtype);
  const auto* img = static_cast<const unsigned char*>(d_img);
  int rc;
  for (uint64_t t0 = 0; t0 < p->L; t0 += frames_per) {
    const uint64_t fr = std::min<uint64_t>(frames_per, p->L - t0);
    if ((rc = plan_spatial_frames(p, img + t0 * frame_b, img, t0, fr, ws, s, err, errlen))) return rc;
  }
  return finish_from_spectra(p, img, img, nullptr, ws, d_out, cols_per, s, err, errlen);
}

namespace {

int host_path_chunked(fddm_plan* p, HostPathCache& c, const ChunkPlan& cp, const void* img, void* out,
                      uint64_t* stats, char* err, size_t errlen) {
  const hipStream_t s = c.stream;
  const uint64_t csz = p->precision == PREC_FP64 ? 16 : 8;
  const uint64_t osz = p->out_dtype == OUT_F32 ? 4 : 8;
  const uint64_t frame_b = p->H * p->W * in_bytes(p->in_dtype);
  unsigned char* dev = nullptr;
  FDDM_HIP(hipMalloc(reinterpret_cast<void**>(&dev), cp.fixed + cp.work));
  struct Free {
    unsigned char* d;
    ~Free() { (void)hipFree(d); }
  } guard{dev};
  // fixed pieces, then the spectra (mode 1), then the reused work region
  unsigned char* at = dev;
  auto take = [&](uint64_t b) {
    unsigned char* r = at;
    at += align256(b);
    return r;
  };
  unsigned char* frame0 = take(frame_b);
  unsigned char* S0 = take(p->bytes_S0);
  unsigned char* flags = take(p->bytes_flags);
  unsigned char* lng = take(p->bytes_long);
  unsigned char* ein = take(p->edge ? p->edge_frames * frame_b : 0);
  unsigned char* ews = take(p->bytes_edge_ws);
  unsigned char* S = cp.mode == 1 ? take(p->bytes_S) : nullptr;
  unsigned char* work = at;
  std::vector<unsigned char> hostS;
  if (cp.mode == 2) hostS.resize(p->nq * p->Lp * csz);
  const int sub = p->subtract0 ? 1 : 0;
  int rc;

  // frame 0 (the subtracted frame) and its raw spectrum F(0) for PS
  FDDM_HIP(hipMemcpyAsync(frame0, img, frame_b, hipMemcpyHostToDevice, s));
  if (sub && (rc = spatial_frames(p, frame0, frame0, 1, 0, work, S0, p->Tb, s, err, errlen))) return rc;
  // spatial stage in chunks of whole frame blocks
  uint64_t nchunks = 0;
  for (uint64_t t0 = 0; t0 < p->L; t0 += cp.frames_per, ++nchunks) {
    const uint64_t fr = std::min<uint64_t>(cp.frames_per, p->L - t0);
    unsigned char* cin = work;
    unsigned char* R = work + align256(cp.frames_per * frame_b);
    if ((rc = h2d_staged(c, cin, static_cast<const unsigned char*>(img) + t0 * frame_b, fr * frame_b, err, errlen)))
      return rc;
    if (cp.mode == 1) {
      rc = spatial_frames(p, cin, frame0, fr, sub, R, S + t0 * csz, p->Lp, s, err, errlen);
    } else {
      unsigned char* piece = R + align256(cp.frames_per / p->Tb * p->nxh * p->H * p->Tb * csz);
      rc = spatial_frames(p, cin, frame0, fr, sub, R, piece, cp.frames_per, s, err, errlen);
      if (!rc) {
        const uint64_t width = (fr + p->Tb - 1) / p->Tb * p->Tb * csz;
        FDDM_HIP(hipMemcpy2DAsync(hostS.data() + t0 * csz, p->Lp * csz, piece, cp.frames_per * csz, width, p->nq,
                                  hipMemcpyDeviceToHost, s));
      }
    }
    if (rc) return rc;
  }
  // FP64 spectra of the edge frames (compensated path)
  if (p->edge) {
    // the first and last kEdgeFrames frames (every frame of a short stack)
    const uint64_t E = kEdgeFrames;
    if (p->L > 2 * E) {
      FDDM_HIP(hipMemcpyAsync(ein, img, E * frame_b, hipMemcpyHostToDevice, s));
      FDDM_HIP(hipMemcpyAsync(ein + E * frame_b, static_cast<const unsigned char*>(img) + (p->L - E) * frame_b,
                              E * frame_b, hipMemcpyHostToDevice, s));
    } else {
      FDDM_HIP(hipMemcpyAsync(ein, img, p->L * frame_b, hipMemcpyHostToDevice, s));
    }
    const int cap = p->edge->timing_cap;
    p->edge->timing_cap = 0;
    rc = run_spatial(p->edge, ein, ews, s, err, errlen);
    p->edge->timing_cap = cap;
    if (rc) return rc;
  }
  FDDM_HIP(hipStreamSynchronize(s));  // the work region is reused by the time chunks

  // time stage in chunks of whole ky rows, local (unshifted) pieces
  TimeArgs ta{};
  ta.Lp = p->Lp;
  ta.L = static_cast<int>(p->L);
  ta.ny = static_cast<int>(p->ny);
  ta.nxh = static_cast<int>(p->nxh);
  ta.S0 = sub ? S0 : nullptr;
  ta.s0_stride = static_cast<uint64_t>(p->Tb);
  ta.lags = static_cast<const uint32_t*>(p->d_lags);
  ta.nl = static_cast<int>(p->nl);
  ta.dt = p->dt;
  ta.twt = p->d_twt;
  ta.flags = p->bytes_flags ? flags : nullptr;
  ta.local = 1;
  const void* S64 = p->edge ? ews + p->edge->bytes_R : nullptr;
  const uint64_t ny = p->ny, nxh = p->nxh, kyw = ny - ny / 2;  // rows ky >= kyw wrap to the top
  uint64_t qchunks = 0;
  for (uint64_t ky0 = 0; ky0 < ny; ky0 += cp.rows_per, ++qchunks) {
    const uint64_t ky1 = std::min(ny, ky0 + cp.rows_per), cols = (ky1 - ky0) * nxh;
    unsigned char* opiece = work;
    if (cp.mode == 2) {
      unsigned char* Sc = work + align256(cp.rows_per * nxh * (p->nl + 2) * osz);
      FDDM_HIP(hipMemcpyAsync(Sc, hostS.data() + ky0 * nxh * p->Lp * csz, cols * p->Lp * csz,
                              hipMemcpyHostToDevice, s));
      ta.S = Sc;
      ta.s_q0 = static_cast<int>(ky0 * nxh);
    } else {
      ta.S = S;
      ta.s_q0 = 0;
    }
    ta.c0 = static_cast<int>(ky0 * nxh);
    ta.cols = static_cast<int>(cols);
    ta.ldo = cols;
    ta.out = opiece;
    if ((rc = time_launch(p, ta, lng, S64, s, err, errlen))) return rc;
    // rows ky -> output rows (ky + ny/2) mod ny: at most two contiguous runs
    for (int part = 0; part < 2; ++part) {
      const uint64_t a0 = part == 0 ? ky0 : std::max(ky0, kyw), a1 = part == 0 ? std::min(ky1, kyw) : ky1;
      if (a0 >= a1) continue;
      const uint64_t yo = part == 0 ? a0 + ny / 2 : a0 - kyw;
      FDDM_HIP(hipMemcpy2DAsync(static_cast<unsigned char*>(out) + yo * nxh * osz, p->nq * osz,
                                opiece + (a0 - ky0) * nxh * osz, cols * osz, (a1 - a0) * nxh * osz, p->nl + 2,
                                hipMemcpyDeviceToHost, s));
    }
    FDDM_HIP(hipStreamSynchronize(s));  // the piece is rewritten by the next chunk
  }
  stats[0] = static_cast<uint64_t>(cp.mode);
  stats[1] = nchunks;
  stats[2] = qchunks;
  return FDDM_OK;
}

int host_path(int mode, const void* img, int dtype, uint64_t L, uint64_t H, uint64_t W,
              const uint32_t* lags, uint64_t nl, uint64_t nx, uint64_t ny, uint64_t nt,
              const void* window, int window_f32, int precision, int out_dtype, void* out,
              char* err, size_t errlen) {
  if (!img || !out) return set_err(err, errlen, FDDM_ERR_ARG, "NULL image or output pointer");
  int dev = 0;
  FDDM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(err, errlen, FDDM_ERR_DEVICE, "device id %d not supported", dev);
  fddm_plan* p = nullptr;
  int rc = plan_create_ex(&p, mode, dtype, L, H, W, lags, nl, nx, ny, nt, window, window_f32, precision,
                          out_dtype, 0, err, errlen);
  if (rc) return rc;
  struct PlanGuard {
    fddm_plan* p;
    ~PlanGuard() {
      plan_free_device(p);
      delete p;
    }
  } guard{p};
  const uint64_t img_bytes = L * H * W * in_bytes(dtype);
  const uint64_t out_b = output_bytes(p);
  std::lock_guard<std::mutex> lk(g_cache_mu[dev]);
  HostPathCache& c = g_cache[dev];
  if (!c.stream) FDDM_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  for (int k = 0; k < 2; ++k) {
    if (!c.pin[k]) FDDM_HIP(hipHostMalloc(&c.pin[k], kStageBytes, hipHostMallocDefault));
    if (!c.ev[k]) FDDM_HIP(hipEventCreateWithFlags(&c.ev[k], hipEventDisableTiming));
  }
  size_t free_b = 0, total_b = 0;
  FDDM_HIP(hipMemGetInfo(&free_b, &total_b));
  // the device memory this call may use: the caller's budget, else 95 % of
  // what is free plus what the cache already holds
  const uint64_t cached = c.img_cap + c.out_cap + c.ws_cap;
  const uint64_t budget = g_budget[dev] ? g_budget[dev]
                                        : static_cast<uint64_t>(0.95 * static_cast<double>(free_b + cached));
  if (img_bytes + out_b + p->ws_bytes > budget) {
    // in pieces: drop the cached buffers first so the pieces fit beside the rest
    FDDM_HIP(hipStreamSynchronize(c.stream));
    drop_buffers(c);
    FDDM_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t avail = std::min<uint64_t>(budget, static_cast<uint64_t>(0.95 * static_cast<double>(free_b)));
    ChunkPlan cp;
    if (!plan_chunks(p, avail, &cp))
      return set_err(err, errlen, FDDM_ERR_MEMORY,
                     "Not enough space on GPU: %llu bytes available, even one frame block and one row of q "
                     "need more",
                     static_cast<unsigned long long>(avail));
    return host_path_chunked(p, c, cp, img, out, g_last_stats[dev], err, errlen);
  }
  const uint64_t need = (c.img_cap >= img_bytes ? 0 : img_bytes) + (c.out_cap >= out_b ? 0 : out_b) +
                        (c.ws_cap >= p->ws_bytes ? 0 : p->ws_bytes);
  if (need > static_cast<uint64_t>(0.95 * static_cast<double>(free_b)))
    return set_err(err, errlen, FDDM_ERR_MEMORY,
                   "Not enough space on GPU: %llu bytes needed, %llu free",
                   static_cast<unsigned long long>(need), static_cast<unsigned long long>(free_b));
  g_last_stats[dev][0] = 0;
  g_last_stats[dev][1] = 1;
  g_last_stats[dev][2] = 1;
  if ((rc = grow(&c.img, &c.img_cap, img_bytes, err, errlen))) return rc;
  if ((rc = grow(&c.out, &c.out_cap, out_b, err, errlen))) return rc;
  if ((rc = grow(&c.ws, &c.ws_cap, p->ws_bytes, err, errlen))) return rc;
  if ((rc = h2d_staged(c, c.img, img, img_bytes, err, errlen))) return rc;
  if ((rc = plan_execute(p, c.img, c.out, c.ws, c.stream, err, errlen))) return rc;
  if ((rc = d2h_staged(c, out, c.out, out_b, err, errlen))) return rc;
  FDDM_HIP(hipStreamSynchronize(c.stream));
  if (c.img_cap + c.out_cap + c.ws_cap > cache_limit(dev)) drop_buffers(c);
  return FDDM_OK;
}

// Workspace of fddm_plan_execute_static: one array per device, held from
// the library's load to the end of the process.  Its device address comes
// from hipGetSymbolAddress, once per device (the array's name is only a host
// stand-in); nothing in it has to start at zero, since every execute writes
// each workspace element it later reads.
constexpr uint64_t kStaticScratchBytes = uint64_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

int static_scratch(void** ws, char* err, size_t errlen) {
  int dev = 0;
  FDDM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(err, errlen, FDDM_ERR_DEVICE, "device id %d not supported", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    FDDM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_err(err, errlen, FDDM_ERR_HIP, "static scratch has no device address");
    g_scratch_ptr[dev] = p;
  }
  *ws = g_scratch_ptr[dev];
  return FDDM_OK;
}

}  // namespace

extern "C" {

int fddm_get_num_devices(int* count, char* err, size_t errlen) {
  if (!count) return set_err(err, errlen, FDDM_ERR_ARG, "NULL count");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e == hipErrorNoDevice) {
    *count = 0;
    return FDDM_OK;
  }
  if (e != hipSuccess) return hip_err(err, errlen, e, "hipGetDeviceCount");
  return FDDM_OK;
}

int fddm_set_device(int gpu_id, char* err, size_t errlen) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess && e != hipErrorNoDevice) return hip_err(err, errlen, e, "hipGetDeviceCount");
  if (gpu_id < 0 || gpu_id >= n)
    return set_err(err, errlen, FDDM_ERR_DEVICE, "Device id %d out of range [0, %d)", gpu_id, n);
  FDDM_HIP(hipSetDevice(gpu_id));
  return FDDM_OK;
}

int fddm_get_device(int* gpu_id, char* err, size_t errlen) {
  if (!gpu_id) return set_err(err, errlen, FDDM_ERR_ARG, "NULL gpu_id");
  FDDM_HIP(hipGetDevice(gpu_id));
  return FDDM_OK;
}

int fddm_get_free_device_memory(uint64_t* free_bytes, char* err, size_t errlen) {
  if (!free_bytes) return set_err(err, errlen, FDDM_ERR_ARG, "NULL free_bytes");
  size_t f = 0, t = 0;
  FDDM_HIP(hipMemGetInfo(&f, &t));
  *free_bytes = f;
  return FDDM_OK;
}

int fddm_get_free_host_memory(uint64_t* free_bytes, char* err, size_t errlen) {
  if (!free_bytes) return set_err(err, errlen, FDDM_ERR_ARG, "NULL free_bytes");
  FILE* f = fopen("/proc/meminfo", "r");
  if (!f) return set_err(err, errlen, FDDM_ERR_ARG, "Failed to retrieve memory information.");
  char line[256];
  unsigned long long kib = 0;
  bool found = false;
  while (fgets(line, sizeof line, f))
    if (sscanf(line, "MemAvailable: %llu kB", &kib) == 1) {
      found = true;
      break;
    }
  fclose(f);
  if (!found) return set_err(err, errlen, FDDM_ERR_ARG, "Failed to retrieve memory information.");
  *free_bytes = static_cast<uint64_t>(kib) * 1024;
  return FDDM_OK;
}

int fddm_get_host_memory(int mode, uint64_t nx, uint64_t ny, uint64_t length, uint64_t num_lags, int out_dtype,
                         uint64_t* bytes, char* err, size_t errlen) {
  (void)length;  // the frames stay in the caller's array; the spectra never come to the host
  if (!bytes) return set_err(err, errlen, FDDM_ERR_ARG, "NULL bytes");
  if (mode != MODE_FFT && mode != MODE_DIFF) return set_err(err, errlen, FDDM_ERR_ARG, "unknown mode %d", mode);
  if (out_dtype != OUT_F64 && out_dtype != OUT_F32)
    return set_err(err, errlen, FDDM_ERR_ARG, "unknown output dtype %d", out_dtype);
  const uint64_t scalar = out_dtype == OUT_F32 ? 4 : 8;
  *bytes = (nx / 2 + 1) * ny * (num_lags + 2) * scalar + 2 * static_cast<uint64_t>(kStageBytes);
  return FDDM_OK;
}

int fddm_check_host_memory(int mode, uint64_t nx, uint64_t ny, uint64_t length, uint64_t num_lags, int out_dtype,
                           int* ok, char* err, size_t errlen) {
  if (!ok) return set_err(err, errlen, FDDM_ERR_ARG, "NULL ok");
  uint64_t need = 0, avail = 0;
  int rc = fddm_get_host_memory(mode, nx, ny, length, num_lags, out_dtype, &need, err, errlen);
  if (!rc) rc = fddm_get_free_host_memory(&avail, err, errlen);
  if (rc) return rc;
  *ok = static_cast<uint64_t>(0.95 * static_cast<double>(avail)) > need ? 1 : 0;  // the reference's margin
  return FDDM_OK;
}

int fddm_ddm_fft(const void* img, int dtype, uint64_t length, uint64_t height, uint64_t width,
                 const uint32_t* lags, uint64_t num_lags, uint64_t nx, uint64_t ny, uint64_t nt,
                 const void* window, int window_f32, int precision, int out_dtype, void* out, char* err,
                 size_t errlen) {
  return host_path(MODE_FFT, img, dtype, length, height, width, lags, num_lags, nx, ny, nt, window,
                   window_f32, precision, out_dtype, out, err, errlen);
}

int fddm_ddm_diff(const void* img, int dtype, uint64_t length, uint64_t height, uint64_t width,
                  const uint32_t* lags, uint64_t num_lags, uint64_t nx, uint64_t ny, const void* window,
                  int window_f32, int precision, int out_dtype, void* out, char* err, size_t errlen) {
  return host_path(MODE_DIFF, img, dtype, length, height, width, lags, num_lags, nx, ny, 0, window,
                   window_f32, precision, out_dtype, out, err, errlen);
}

int fddm_release_cache(char* err, size_t errlen) {
  int dev = 0;
  FDDM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return FDDM_OK;
  std::lock_guard<std::mutex> lk(g_cache_mu[dev]);
  HostPathCache& c = g_cache[dev];
  if (c.stream) FDDM_HIP(hipStreamSynchronize(c.stream));
  // device memory only: the two pinned staging buffers (and their stream and
  // events) stay until the process ends, so later calls pin nothing again
  drop_buffers(c);
  azavg_release_scratch(dev);
  ingest_release_pool(dev);
  return FDDM_OK;
}

int fddm_get_cached_bytes(uint64_t* bytes, char* err, size_t errlen) {
  if (!bytes) return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument");
  int dev = 0;
  FDDM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(err, errlen, FDDM_ERR_DEVICE, "device id %d not supported", dev);
  std::lock_guard<std::mutex> lk(g_cache_mu[dev]);
  const HostPathCache& c = g_cache[dev];
  *bytes = c.img_cap + c.out_cap + c.ws_cap;
  return FDDM_OK;
}

int fddm_set_cache_limit(int64_t bytes, char* err, size_t errlen) {
  int dev = 0;
  FDDM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(err, errlen, FDDM_ERR_DEVICE, "device id %d not supported", dev);
  std::lock_guard<std::mutex> lk(g_cache_mu[dev]);
  g_cache_limit_set[dev] = bytes >= 0;
  g_cache_limit[dev] = bytes >= 0 ? static_cast<uint64_t>(bytes) : 0;
  HostPathCache& c = g_cache[dev];
  if (c.img_cap + c.out_cap + c.ws_cap > cache_limit(dev)) {
    if (c.stream) FDDM_HIP(hipStreamSynchronize(c.stream));
    drop_buffers(c);
  }
  return FDDM_OK;
}

int fddm_set_memory_budget(uint64_t bytes, char* err, size_t errlen) {
  int dev = 0;
  FDDM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(err, errlen, FDDM_ERR_DEVICE, "device id %d not supported", dev);
  std::lock_guard<std::mutex> lk(g_cache_mu[dev]);
  g_budget[dev] = bytes;
  return FDDM_OK;
}

int fddm_get_last_call_chunks(uint64_t* stats, char* err, size_t errlen) {
  if (!stats) return set_err(err, errlen, FDDM_ERR_ARG, "NULL stats");
  int dev = 0;
  FDDM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(err, errlen, FDDM_ERR_DEVICE, "device id %d not supported", dev);
  std::lock_guard<std::mutex> lk(g_cache_mu[dev]);
  for (int k = 0; k < 3; ++k) stats[k] = g_last_stats[dev][k];
  return FDDM_OK;
}

int fddm_plan_create(fddm_plan** plan, int mode, int dtype, uint64_t length, uint64_t height, uint64_t width,
                     const uint32_t* lags, uint64_t num_lags, uint64_t nx, uint64_t ny, uint64_t nt,
                     const void* window, int window_f32, int precision, int out_dtype, char* err,
                     size_t errlen) {
  return plan_create_ex(plan, mode, dtype, length, height, width, lags, num_lags, nx, ny, nt, window, window_f32,
                        precision, out_dtype, 0, err, errlen);
}

int fddm_plan_destroy(fddm_plan* plan) {
  if (!plan) return FDDM_OK;
  plan_free_device(plan);
  delete plan;
  return FDDM_OK;
}

int fddm_plan_workspace_size(const fddm_plan* plan, uint64_t* bytes) {
  if (!plan || !bytes) return FDDM_ERR_ARG;
  *bytes = plan->ws_bytes;
  return FDDM_OK;
}

int fddm_plan_output_size(const fddm_plan* plan, uint64_t* bytes) {
  if (!plan || !bytes) return FDDM_ERR_ARG;
  *bytes = output_bytes(plan);
  return FDDM_OK;
}

int fddm_plan_execute(fddm_plan* plan, const void* d_img, void* d_out, void* d_workspace, void* stream,
                      char* err, size_t errlen) {
  return plan_execute(plan, d_img, d_out, d_workspace, static_cast<hipStream_t>(stream), err, errlen);
}

int fddm_plan_execute_static(fddm_plan* plan, const void* d_img, void* d_out, void* stream, char* err,
                             size_t errlen) {
  if (!plan || !d_img || !d_out) return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument to execute");
  if (plan->ws_bytes > kStaticScratchBytes)
    return set_err(err, errlen, FDDM_ERR_SIZE, "workspace of %llu bytes exceeds the static scratch (%llu bytes)",
                   static_cast<unsigned long long>(plan->ws_bytes),
                   static_cast<unsigned long long>(kStaticScratchBytes));
  void* ws = nullptr;
  const int rc = static_scratch(&ws, err, errlen);
  if (rc) return rc;
  return plan_execute(plan, d_img, d_out, ws, static_cast<hipStream_t>(stream), err, errlen);
}

int fddm_plan_timing_begin(fddm_plan* plan, int max_steps, char* err, size_t errlen) {
  if (!plan || max_steps <= 0 || max_steps > 100000)
    return set_err(err, errlen, FDDM_ERR_ARG, "bad timing request");
  for (hipEvent_t e : plan->events) (void)hipEventDestroy(e);
  plan->events.assign(static_cast<size_t>(max_steps) * kEventsPerStep, nullptr);
  for (auto& e : plan->events) FDDM_HIP(hipEventCreate(&e));
  plan->timing_cap = max_steps;
  plan->timing_steps = 0;
  return FDDM_OK;
}

int fddm_plan_timing_end(fddm_plan* plan, double* stage_ms, int* steps, char* err, size_t errlen) {
  if (!plan || !stage_ms || !steps) return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument");
  for (int k = 0; k < kStages; ++k) stage_ms[k] = 0.0;
  *steps = plan->timing_steps;
  for (int i = 0; i < plan->timing_steps; ++i) {
    hipEvent_t* ev = &plan->events[static_cast<size_t>(i) * kEventsPerStep];
    FDDM_HIP(hipEventSynchronize(ev[kEventsPerStep - 1]));
    for (int k = 0; k < kStages; ++k) {
      float ms = 0.f;
      FDDM_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      stage_ms[k] += ms;
    }
  }
  for (hipEvent_t e : plan->events) (void)hipEventDestroy(e);
  plan->events.clear();
  plan->timing_cap = 0;
  plan->timing_steps = 0;
  return FDDM_OK;
}

int fddm_plan_spectra(fddm_plan* plan, const void* d_img, void* d_workspace, double* host_spectra, char* err,
                      size_t errlen) {
  if (!plan || !d_img || !d_workspace || !host_spectra)
    return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument");
  if (const int rc = check_alignment(plan, d_img, nullptr, d_workspace, err, errlen)) return rc;
  const int cap = plan->timing_cap;
  plan->timing_cap = 0;
  int rc = run_spatial(plan, d_img, d_workspace, nullptr, err, errlen);
  plan->timing_cap = cap;
  if (rc) return rc;
  FDDM_HIP(hipDeviceSynchronize());
  const bool f64 = plan->precision == PREC_FP64;
  const size_t csz = f64 ? 16 : 8;
  std::vector<unsigned char> tmp(plan->nq * plan->Lp * csz);
  FDDM_HIP(hipMemcpy(tmp.data(), static_cast<unsigned char*>(d_workspace) + plan->bytes_R, tmp.size(),
                     hipMemcpyDeviceToHost));
  for (uint64_t q = 0; q < plan->nq; ++q)
    for (uint64_t t = 0; t < plan->L; ++t) {
      const uint64_t i = q * plan->Lp + t, o = 2 * (q * plan->L + t);
      if (f64) {
        host_spectra[o] = reinterpret_cast<const double*>(tmp.data())[2 * i];
        host_spectra[o + 1] = reinterpret_cast<const double*>(tmp.data())[2 * i + 1];
      } else {
        host_spectra[o] = reinterpret_cast<const float*>(tmp.data())[2 * i];
        host_spectra[o + 1] = reinterpret_cast<const float*>(tmp.data())[2 * i + 1];
      }
    }
  return FDDM_OK;
}

int fddm_malloc(void** ptr, uint64_t bytes, char* err, size_t errlen) {
  if (!ptr) return set_err(err, errlen, FDDM_ERR_ARG, "NULL ptr");
  FDDM_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return FDDM_OK;
}

int fddm_free(void* ptr) { return ptr && hipFree(ptr) != hipSuccess ? FDDM_ERR_HIP : FDDM_OK; }

int fddm_memcpy_htod(void* dst, const void* src, uint64_t bytes, char* err, size_t errlen) {
  FDDM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return FDDM_OK;
}

int fddm_memcpy_dtoh(void* dst, const void* src, uint64_t bytes, char* err, size_t errlen) {
  FDDM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return FDDM_OK;
}

int fddm_memcpy_dtod(void* dst, const void* src, uint64_t bytes, char* err, size_t errlen) {
  FDDM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return FDDM_OK;
}

int fddm_stream_create(void** stream, char* err, size_t errlen) {
  if (!stream) return set_err(err, errlen, FDDM_ERR_ARG, "NULL stream");
  hipStream_t s = nullptr;
  FDDM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return FDDM_OK;
}

int fddm_stream_destroy(void* stream) {
  return stream && hipStreamDestroy(static_cast<hipStream_t>(stream)) != hipSuccess ? FDDM_ERR_HIP : FDDM_OK;
}

int fddm_stream_synchronize(void* stream, char* err, size_t errlen) {
  FDDM_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return FDDM_OK;
}

int fddm_test_fft(double* data, uint64_t n, uint64_t batch, int precision, char* err, size_t errlen) {
  if (!data || n == 0 || batch == 0) return set_err(err, errlen, FDDM_ERR_ARG, "bad FFT test arguments");
  FftDesc d{};
  if (!factorize(n, d)) return set_err(err, errlen, FDDM_ERR_SIZE, "n=%llu is not 7-smooth", (unsigned long long)n);
  const bool f64 = precision == PREC_FP64;
  const size_t csz = f64 ? 16 : 8;
  const int e = fft_bucket(static_cast<long long>(n));
  if (e == 0 || (f64 && e > 32) || lds_padded_size(static_cast<int>(n)) * csz > kLdsLimit)
    return set_err(err, errlen, FDDM_ERR_SIZE, "n too large");
  if (batch > (1u << 30)) return set_err(err, errlen, FDDM_ERR_SIZE, "batch too large");
  void* tw = nullptr;
  int rc = upload_twiddles(&tw, d, f64, err, errlen);
  if (rc) return rc;
  void* buf = nullptr;
  hipError_t he = hipMalloc(&buf, n * batch * csz);
  if (he != hipSuccess) {
    (void)hipFree(tw);
    return hip_err(err, errlen, he, "hipMalloc");
  }
  std::vector<float> f;
  if (f64) {
    he = hipMemcpy(buf, data, n * batch * csz, hipMemcpyHostToDevice);
  } else {
    f.resize(2 * n * batch);
    for (size_t i = 0; i < f.size(); ++i) f[i] = static_cast<float>(data[i]);
    he = hipMemcpy(buf, f.data(), n * batch * csz, hipMemcpyHostToDevice);
  }
  if (he == hipSuccess) he = launch_fft_test(buf, static_cast<int>(n), static_cast<int>(batch), d, tw, precision, nullptr);
  if (he == hipSuccess) he = hipDeviceSynchronize();
  if (he == hipSuccess) {
    if (f64) {
      he = hipMemcpy(data, buf, n * batch * csz, hipMemcpyDeviceToHost);
    } else {
      he = hipMemcpy(f.data(), buf, n * batch * csz, hipMemcpyDeviceToHost);
      for (size_t i = 0; i < f.size(); ++i) data[i] = f[i];
    }
  }
  (void)hipFree(buf);
  (void)hipFree(tw);
  if (he != hipSuccess) return hip_err(err, errlen, he, "fft test");
  return FDDM_OK;
}

}  // extern "C"
