// Azimuthal average of an image structure function on the device.
//
// Semantics of azimuthal_average_array (src/python/azimuthalaverage.py:277-480
// of the reference): for every bin b and row t (lag, then PS, then VAR)
//   avg[b, t] = sum_i w_i x_i(t) / sum_i w_i
//   err[b, t] = sqrt( sum_i w_i (x_i(t) - avg[b, t])^2 / (sum_i w_i - sum_i w2_i / sum_i w_i) )
// over the pixels i of bin b, where w_i = counts_i * weights_i (zero for
// masked pixels) and w2_i = counts_i * weights_i^2.  The host computes the
// bin index and weights of every pixel; the device reads the (rows, ny, nxh)
// result once per pass (twice with errors) where it already lies.
//
// Data layout and reduction order (deterministic, no atomics): along one ky
// row the modulus |k| grows with kx, so a row's bin indices never decrease.
// A workgroup takes one row t of the data and a chunk of kChunkRows ky rows;
// each wave walks its ky rows 64 kx at a time, reduces equal-bin runs with a
// segmented wave scan and adds each run's sum to its own LDS accumulators;
// the waves' accumulators are added in fixed order into
// partial[t][chunk][bin], and a second kernel sums the chunks in order.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <vector>

#include <hip/hip_runtime.h>

#include "fddm_hip.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kChunkRows = 64;
constexpr int kMaxBins = 1024;

int set_err(char* err, size_t len, int code, const char* fmt, ...) {
  if (err && len) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, len, fmt, ap);
    va_end(ap);
  }
  return code;
}

// Pass 1 (SQ = false): sum w x.  Pass 2 (SQ = true): sum w (x - avg[b, t])^2.
template <typename In, bool SQ>
__global__ void __launch_bounds__(kThreads)
    k_azavg_partial(const In* __restrict__ data, int ny, int nxh, int nbins, int nchunks,
                    const int* __restrict__ bin, const double* __restrict__ w, const double* __restrict__ avg,
                    int rows, double* __restrict__ partial) {
  extern __shared__ double acc[];  // [kWaves][nbins]
  const int t = blockIdx.x / nchunks, chunk = blockIdx.x - t * nchunks;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* my = acc + wave * nbins;
  for (int b = threadIdx.x; b < kWaves * nbins; b += kThreads) acc[b] = 0.0;
  __syncthreads();
  const size_t nq = static_cast<size_t>(ny) * nxh;
  const In* __restrict__ row = data + static_cast<size_t>(t) * nq;
  const int y0 = chunk * kChunkRows, y1 = min(ny, y0 + kChunkRows);
  for (int y = y0 + wave; y < y1; y += kWaves) {
    const size_t base = static_cast<size_t>(y) * nxh;
    for (int x0 = 0; x0 < nxh; x0 += 64) {
      const int x = x0 + lane;
      const bool in = x < nxh;
      const size_t q = base + (in ? x : 0);
      int b = in ? bin[q] : -1;
      double v = 0.0;
      if (in) {
        const double wq = w[q];
        if (wq != 0.0) {
          const double xv = static_cast<double>(row[q]);
          if constexpr (SQ) {
            const double d = xv - avg[static_cast<size_t>(b) * rows + t];
            v = wq * d * d;
          } else {
            v = wq * xv;
          }
        }
      }
      // lanes past the row end take the last bin so the run structure holds
      const int blast = __shfl(b, min(63, nxh - 1 - x0), 64);
      if (!in) b = blast;
      // segmented inclusive scan over runs of equal b (b is non-decreasing)
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const double u = __shfl_up(v, o, 64);
        const int bu = __shfl_up(b, o, 64);
        if (lane >= o && bu == b) v += u;
      }
      const int bn = __shfl_down(b, 1, 64);
      if (lane == 63 || bn != b) my[b] += v;  // one lane per run: no conflicts
    }
  }
  __syncthreads();
  double* out = partial + (static_cast<size_t>(t) * nchunks + chunk) * nbins;
  for (int b = threadIdx.x; b < nbins; b += kThreads) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) s += acc[k * nbins + b];
    out[b] = s;
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

}  // namespace

extern "C" int fddm_azimuthal_average(const void* d_data, int data_f32, uint64_t rows, uint64_t ny, uint64_t nxh,
                                      const int32_t* bin_index, const double* w, const double* w2, int nbins,
                                      int eval_err, double* avg_out, double* err_out, void* stream, char* err,
                                      size_t errlen) {
  if (!d_data || !bin_index || !w || !avg_out || (eval_err && (!w2 || !err_out)))
    return set_err(err, errlen, FDDM_ERR_ARG, "NULL argument to azimuthal average");
  if (nbins < 1 || nbins > kMaxBins) return set_err(err, errlen, FDDM_ERR_ARG, "nbins must be in [1, %d]", kMaxBins);
  if (rows == 0 || ny == 0 || nxh == 0 || rows > (1u << 30) || ny > (1u << 20) || nxh > (1u << 20))
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

  const int nchunks = static_cast<int>((ny + kChunkRows - 1) / kChunkRows);
  const size_t b_bin = nq * sizeof(int32_t), b_w = nq * sizeof(double);
  const size_t b_part = rows * nchunks * nbins * sizeof(double), b_res = rows * nbins * sizeof(double);
  const size_t b_den = nbins * sizeof(double);
  auto s = static_cast<hipStream_t>(stream);
  unsigned char* buf = nullptr;
  const size_t total = b_bin + b_w + b_part + 2 * b_res + 2 * b_den + 5 * 256;
  if (hipMalloc(&buf, total) != hipSuccess)
    return set_err(err, errlen, FDDM_ERR_MEMORY, "azimuthal average scratch of %zu bytes", total);
  auto carve = [&, off = size_t(0)](size_t n) mutable {
    void* p = buf + off;
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
  const unsigned grid1 = static_cast<unsigned>(rows * nchunks);
  const unsigned grid2 = static_cast<unsigned>((rows * nbins + kThreads - 1) / kThreads);
  const size_t lds = kWaves * nbins * sizeof(double);
  auto pass = [&](auto sq) {
    constexpr bool SQ = decltype(sq)::value;
    if (data_f32)
      hipLaunchKernelGGL((k_azavg_partial<float, SQ>), dim3(grid1), dim3(kThreads), lds, s,
                         static_cast<const float*>(d_data), static_cast<int>(ny), static_cast<int>(nxh), nbins,
                         nchunks, d_bin, d_w, d_avg, static_cast<int>(rows), d_part);
    else
      hipLaunchKernelGGL((k_azavg_partial<double, SQ>), dim3(grid1), dim3(kThreads), lds, s,
                         static_cast<const double*>(d_data), static_cast<int>(ny), static_cast<int>(nxh), nbins,
                         nchunks, d_bin, d_w, d_avg, static_cast<int>(rows), d_part);
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
  hipStreamSynchronize(s);
  hipFree(buf);
  return rc;
}
