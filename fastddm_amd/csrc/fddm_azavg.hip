// Azimuthal average of an image structure function on the device.
//
// Semantics of azimuthal_average_array (src/python/azimuthalaverage.py:277-480
// of the reference): for every bin b and row t (lag, then PS, then VAR)
//   avg[b, t] = sum_i w_i x_i(t) / sum_i w_i
//   err[b, t] = sqrt( sum_i w_i (x_i(t) - avg[b, t])^2 / (sum_i w_i - sum_i w2_i / sum_i w_i) )
// over the pixels i of bin b, where w_i = counts_i * weights_i (zero for
// masked pixels) and w2_i = counts_i * weights_i^2.  The host computes the
// bin index and weights of every pixel; the device reads the (rows, ny, nxh)
// result where it already lies, once per pass (twice with errors: the
// deviations are taken from the finished average, as the reference does).
//
// Work split (deterministic, no atomics).  Pixels are taken in flat q order
// in tiles of 64 (one per wave instruction); a workgroup owns T consecutive
// rows t and a contiguous range of tiles.  Each lane loads its pixel's bin
// and weight once and the T values x_q(t0 .. t0+T-1) together, so T loads per
// lane are in flight.  Along one ky row the modulus |k| grows with kx, so the
// key (row within the tile, bin) never decreases across a tile's lanes: runs
// of equal key are reduced with a segmented wave scan built from DPP row
// shifts and row broadcasts, and the last lane of each run adds the run's sum
// to the wave's own LDS accumulator.  Runs of one row have distinct bins, and
// runs of different rows are added in separate steps, so no two lanes of one
// LDS write share an address.  Waves' accumulators are summed in fixed order
// into partial[t][chunk][bin]; a second kernel sums the chunks in order.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <vector>

#include <hip/hip_runtime.h>

#include "fddm_hip.h"
#include "fddm_plan.hpp"
#include "fddm_wave.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxBins = 1024;
constexpr int kTargetBlocks = 4096;  // workgroups aimed for per pass (256 CUs x 16)
constexpr size_t kLdsBudget = 64 * 1024;

int set_err(char* err, size_t len, int code, const char* fmt, ...) {
  if (err && len) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, len, fmt, ap);
    va_end(ap);
  }
  return code;
}

using fddm::wave::dpp_zero;

// Keys moved by DPP: lanes with no source read -1, which no real key equals.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ bool same_key(int key) {
  return __builtin_amdgcn_update_dpp(-1, key, CTRL, ROW_MASK, 0xf, false) == key;
}

// Run flags of one tile: bit s set when the lane's run reaches back to the
// source lane of scan step s.
__device__ __forceinline__ unsigned run_flags(int key) {
  unsigned f = 0;
  f |= same_key<0x111, 0xf>(key) ? 1u : 0u;
  f |= same_key<0x112, 0xf>(key) ? 2u : 0u;
  f |= same_key<0x114, 0xf>(key) ? 4u : 0u;
  f |= same_key<0x118, 0xf>(key) ? 8u : 0u;
  f |= same_key<0x142, 0xa>(key) ? 16u : 0u;
  f |= same_key<0x143, 0xc>(key) ? 32u : 0u;
  return f;
}

// Segmented inclusive scan over the wave: within 16-lane rows by shifts of
// 1, 2, 4, 8, then across rows by broadcasts of lanes 15/47 and 31.
__device__ __forceinline__ double seg_scan(double v, unsigned f) {
  double u;
  u = dpp_zero<0x111, 0xf>(v);
  v += (f & 1u) ? u : 0.0;
  u = dpp_zero<0x112, 0xf>(v);
  v += (f & 2u) ? u : 0.0;
  u = dpp_zero<0x114, 0xf>(v);
  v += (f & 4u) ? u : 0.0;
  u = dpp_zero<0x118, 0xf>(v);
  v += (f & 8u) ? u : 0.0;
  u = dpp_zero<0x142, 0xa>(v);
  v += (f & 16u) ? u : 0.0;
  u = dpp_zero<0x143, 0xc>(v);
  v += (f & 32u) ? u : 0.0;
  return v;
}

// Pass 1 (SQ = false): sum w x.  Pass 2 (SQ = true): sum w (x - avg[b, t])^2.
// LDS: acc[kWaves][T][nbins], then (SQ) mean[T][nbins].
template <typename In, bool SQ, int T>
__global__ void __launch_bounds__(kThreads)
    k_azavg_partial(const In* __restrict__ data, uint32_t nq, int nxh, int nbins, int chunk_tiles, int nchunks,
                    const int* __restrict__ bin, const double* __restrict__ w, const double* __restrict__ avg,
                    int rows, double* __restrict__ partial) {
  extern __shared__ double acc[];
  const int ttile = blockIdx.x / nchunks, chunk = blockIdx.x - ttile * nchunks;
  const int t0 = ttile * T, nt = min(T, rows - t0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* const my = acc + wave * T * nbins;
  double* const mean = acc + kWaves * T * nbins;
  for (int i = threadIdx.x; i < kWaves * T * nbins; i += kThreads) acc[i] = 0.0;
  if constexpr (SQ) {
    for (int i = threadIdx.x; i < T * nbins; i += kThreads) {
      const int tt = i / nbins, b = i - tt * nbins;
      mean[i] = tt < nt ? avg[static_cast<size_t>(b) * rows + t0 + tt] : 0.0;
    }
  }
  __syncthreads();

  const In* __restrict__ rowt[T];
#pragma unroll
  for (int tt = 0; tt < T; ++tt) rowt[tt] = data + static_cast<size_t>(min(t0 + tt, rows - 1)) * nq;
  const uint32_t tile_end = min(static_cast<uint32_t>(chunk + 1) * chunk_tiles * 64u, nq);
  for (uint32_t q0 = (static_cast<uint32_t>(chunk) * chunk_tiles + wave) * 64u; q0 < tile_end; q0 += kWaves * 64u) {
    const uint32_t qlast = min(q0 + 63u, tile_end - 1u);
    const bool in = q0 + lane <= qlast;
    const uint32_t q = in ? q0 + lane : qlast;  // lanes past the end repeat the last pixel with weight 0
    In x[T];
#pragma unroll
    for (int tt = 0; tt < T; ++tt) x[tt] = rowt[tt][q];
    const int b = bin[q];
    const double wq = in ? w[q] : 0.0;
    const int y0 = static_cast<int>(q0 / nxh), yrow = static_cast<int>(q / nxh) - y0;
    const int nrows = static_cast<int>(qlast / nxh) - y0 + 1;
    const int key = yrow * kMaxBins + b;
    const unsigned f = run_flags(key);
    const int key_next = __shfl_down(key, 1, 64);  // every lane takes part
    const bool run_end = (lane == 63) | (key_next != key);
#pragma unroll
    for (int tt = 0; tt < T; ++tt) {
      double v = 0.0;
      if (wq != 0.0) {  // masked pixels add nothing, not even a NaN
        const double xv = static_cast<double>(x[tt]);
        if constexpr (SQ) {
          const double d = xv - mean[tt * nbins + b];
          v = wq * d * d;
        } else {
          v = wq * xv;
        }
      }
      v = seg_scan(v, f);
      if (tt < nt) {
        for (int r = 0; r < nrows; ++r)
          if (run_end && yrow == r) my[tt * nbins + b] += v;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nt * nbins; i += kThreads) {
    const int tt = i / nbins, b = i - tt * nbins;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) s += acc[(k * T + tt) * nbins + b];
    partial[(static_cast<size_t>(t0 + tt) * nchunks + chunk) * nbins + b] = s;
  }
}

// Sums the chunks of every (t, b) in order; pass 1 divides by sum_w, pass 2
// by the bias factor and takes the square root.  Output [b][t].
template <bool SQ>
__global__ void __launch_bounds__(kThreads)
    k_azavg_finish(const double* __restrict__ partial, int rows, int nchunks, int nbins,
                   const double* __restrict__ denom, double* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= static_cast<size_t>(rows) * nbins) return;
  const int t = static_cast<int>(i / nbins), b = static_cast<int>(i - static_cast<size_t>(t) * nbins);
  const double* p = partial + static_cast<size_t>(t) * nchunks * nbins + b;
  double s = 0.0;
  for (int c = 0; c < nchunks; ++c) s += p[static_cast<size_t>(c) * nbins];
  const double d = denom[b];
  double r;
  if constexpr (SQ) {
    r = sqrt(s / d);  // 0/0 -> NaN, x/0 -> inf, as numpy
  } else {
    r = s / d;
  }
  out[static_cast<size_t>(b) * rows + t] = r;
}

// Rows per workgroup: the largest of 8, 4, 2, 1 whose accumulators (and, in
// the second pass, the staged averages) fit the LDS budget.
int rows_per_block(int nbins) {
  for (int t : {8, 4, 2})
    if (static_cast<size_t>(kWaves + 1) * t * nbins * sizeof(double) <= kLdsBudget) return t;
  return 1;
}

template <typename In, bool SQ>
void launch_partial(int T, unsigned grid, size_t lds, hipStream_t s, const In* data, uint32_t nq, int nxh, int nbins,
                    int chunk_tiles, int nchunks, const int* bin, const double* w, const double* avg, int rows,
                    double* partial) {
  switch (T) {
    case 8:
      hipLaunchKernelGGL((k_azavg_partial<In, SQ, 8>), dim3(grid), dim3(kThreads), lds, s, data, nq, nxh, nbins,
                         chunk_tiles, nchunks, bin, w, avg, rows, partial);
      break;
    case 4:
      hipLaunchKernelGGL((k_azavg_partial<In, SQ, 4>), dim3(grid), dim3(kThreads), lds, s, data, nq, nxh, nbins,
                         chunk_tiles, nchunks, bin, w, avg, rows, partial);
      break;
    case 2:
      hipLaunchKernelGGL((k_azavg_partial<In, SQ, 2>), dim3(grid), dim3(kThreads), lds, s, data, nq, nxh, nbins,
                         chunk_tiles, nchunks, bin, w, avg, rows, partial);
      break;
    default:
      hipLaunchKernelGGL((k_azavg_partial<In, SQ, 1>), dim3(grid), dim3(kThreads), lds, s, data, nq, nxh, nbins,
                         chunk_tiles, nchunks, bin, w, avg, rows, partial);
      break;
  }
}

// Device scratch kept between calls (grown on demand), one per device; the
// mutex serialises calls, which share it.
struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::mutex g_scratch_mu;
Scratch g_scratch[64];

}  // namespace

namespace fddm {
void azavg_release_scratch(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0 || dev >= 64 || !g_scratch[dev].ptr) return;
  (void)hipDeviceSynchronize();
  (void)hipFree(g_scratch[dev].ptr);
  g_scratch[dev] = Scratch{};
}
}  // namespace fddm

extern "C" int fddm_azimuthal_average(const void* d_data, int data_f32, uint64_t rows, uint64_t ny, uint64_t nxh,
                                      const int32_t* bin_index, const double* w, const double* w2, int nbins,
                                      int eval_err, double* avg_out, double* err_out, void* stream, char* err,
                                      size_t errlen) {
  if (!d_data || !bin_index || !w || !avg_out || (eval_err && (!w2 || !err_out)))
    return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument to azimuthal average");
  if (nbins < 1 || nbins > kMaxBins) return set_err(err, errlen, FDDM_ERR_ARG, "nbins must be in [1, %d]", kMaxBins);
  if (rows == 0 || ny == 0 || nxh == 0 || rows > (1u << 30) || ny > (1u << 20) || nxh > (1u << 20) ||
      ny * nxh > 0xffffffc0ull)
    return set_err(err, errlen, FDDM_ERR_ARG, "bad data shape");
  const size_t nq = ny * nxh;
  // sums of weights per bin and the denominators, in FP64 on the host
  std::vector<double> sw(nbins, 0.0), sw2(nbins, 0.0);
  for (size_t q = 0; q < nq; ++q) {
    const int b = bin_index[q];
    if (b < 0 || b >= nbins) return set_err(err, errlen, FDDM_ERR_ARG, "bin index %d of pixel %zu outside [0, %d)",
                                            b, q, nbins);
    if (q % nxh && b < bin_index[q - 1])
      return set_err(err, errlen, FDDM_ERR_ARG, "bin indices must not decrease along kx (pixel %zu)", q);
    sw[b] += w[q];
    if (eval_err) sw2[b] += w2[q];
  }
  std::vector<double> bias(nbins);
  for (int b = 0; b < nbins; ++b) bias[b] = sw[b] > 0 ? sw[b] - sw2[b] / sw[b] : 0.0;

  // grid: T rows per workgroup, chunks of chunk_tiles tiles of 64 pixels
  const int T = rows_per_block(nbins);
  const uint64_t ntiles = (nq + 63) / 64, ttiles = (rows + T - 1) / T;
  const uint64_t want = (ntiles * ttiles + kTargetBlocks - 1) / kTargetBlocks;
  const int chunk_tiles = static_cast<int>(std::min<uint64_t>(std::max<uint64_t>(want, kWaves), 1024));
  const int nchunks = static_cast<int>((ntiles + chunk_tiles - 1) / chunk_tiles);
  if (ttiles * nchunks > 0x7fffffffull) return set_err(err, errlen, FDDM_ERR_ARG, "bad data shape");

  const size_t b_bin = nq * sizeof(int32_t), b_w = nq * sizeof(double);
  const size_t b_part = rows * nchunks * nbins * sizeof(double), b_res = rows * nbins * sizeof(double);
  const size_t b_den = nbins * sizeof(double);
  const size_t total = b_bin + b_w + b_part + 2 * b_res + 2 * b_den + 7 * 256;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return set_err(err, errlen, FDDM_ERR_HIP, "no current device");
  auto s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  Scratch& sc = g_scratch[dev];
  if (sc.bytes < total) {
    if (sc.ptr) {
      (void)hipDeviceSynchronize();
      (void)hipFree(sc.ptr);
      sc = Scratch{};
    }
    if (hipMalloc(&sc.ptr, total) != hipSuccess) {
      sc = Scratch{};
      return set_err(err, errlen, FDDM_ERR_MEMORY, "azimuthal average scratch of %zu bytes", total);
    }
    sc.bytes = total;
  }
  auto carve = [base = static_cast<unsigned char*>(sc.ptr), off = size_t(0)](size_t n) mutable {
    void* p = base + off;
    off += (n + 255) & ~size_t(255);
    return p;
  };
  auto* d_bin = static_cast<int*>(carve(b_bin));
  auto* d_w = static_cast<double*>(carve(b_w));
  auto* d_part = static_cast<double*>(carve(b_part));
  auto* d_avg = static_cast<double*>(carve(b_res));
  auto* d_err = static_cast<double*>(carve(b_res));
  auto* d_sw = static_cast<double*>(carve(b_den));
  auto* d_bias = static_cast<double*>(carve(b_den));
  int rc = FDDM_OK;
  const unsigned grid1 = static_cast<unsigned>(ttiles * nchunks);
  const unsigned grid2 = static_cast<unsigned>((rows * nbins + kThreads - 1) / kThreads);
  auto pass = [&](auto sq) {
    constexpr bool SQ = decltype(sq)::value;
    const size_t lds = static_cast<size_t>(kWaves + (SQ ? 1 : 0)) * T * nbins * sizeof(double);
    if (data_f32)
      launch_partial<float, SQ>(T, grid1, lds, s, static_cast<const float*>(d_data), static_cast<uint32_t>(nq),
                                static_cast<int>(nxh), nbins, chunk_tiles, nchunks, d_bin, d_w, d_avg,
                                static_cast<int>(rows), d_part);
    else
      launch_partial<double, SQ>(T, grid1, lds, s, static_cast<const double*>(d_data), static_cast<uint32_t>(nq),
                                 static_cast<int>(nxh), nbins, chunk_tiles, nchunks, d_bin, d_w, d_avg,
                                 static_cast<int>(rows), d_part);
    hipLaunchKernelGGL(k_azavg_finish<SQ>, dim3(grid2), dim3(kThreads), 0, s, d_part, static_cast<int>(rows),
                       nchunks, nbins, SQ ? d_bias : d_sw, SQ ? d_err : d_avg);
  };
  if (hipMemcpyAsync(d_bin, bin_index, b_bin, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d_w, w, b_w, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d_sw, sw.data(), b_den, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d_bias, bias.data(), b_den, hipMemcpyHostToDevice, s) != hipSuccess) {
    rc = set_err(err, errlen, FDDM_ERR_HIP, "upload of the bin maps failed");
  }
  if (!rc) {
    pass(std::false_type{});
    if (eval_err) pass(std::true_type{});
    if (hipGetLastError() != hipSuccess) rc = set_err(err, errlen, FDDM_ERR_HIP, "azimuthal average launch failed");
  }
  if (!rc && (hipMemcpyAsync(avg_out, d_avg, b_res, hipMemcpyDeviceToHost, s) != hipSuccess ||
              (eval_err && hipMemcpyAsync(err_out, d_err, b_res, hipMemcpyDeviceToHost, s) != hipSuccess) ||
              hipStreamSynchronize(s) != hipSuccess))
    rc = set_err(err, errlen, FDDM_ERR_HIP, "azimuthal average failed");
  (void)hipStreamSynchronize(s);
  return rc;
}
