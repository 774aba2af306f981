// FP64 edge correction of the compensated FP32 path.
//
// A lag dt with n = L - dt frame pairs sums n squared differences.  For n = 1
// the single difference |F(L-1) - F(0)|^2 can be arbitrarily small compared
// with |F|^2 (the two frames happen to agree at that q), and the float
// rounding of the spatial transform -- about 1e-6 of the spectrum's rms --
// then dominates it: at 1024^2 x 4000 about one q in 10^6 misses rtol 1e-3.
// For n >= 2 all n differences would have to be small at once, which does
// not happen at these sizes.  So the first and last kEdgeFrames frames are
// transformed once more in FP64 (a child plan over those frames, frame-0
// subtracted like the main pass), and the lags with n <= kEdgeFrames are
// recomputed from them by direct sums in FP64, after the time / diff kernel
// has written its own values.
//
// The time kernel's Wiener-Khinchin values carry an absolute error of a few
// 1e-7 of sum_t |G|^2 in (L - dt) D.  Where (L - dt) D is small against that
// sum -- slow low-q modes and chance near-recurrences -- the kernel marks the
// q, and k_wk_fix recomputes the affected lags of the marked q by direct
// differences of the stored float series in double: each workgroup lists its
// marked columns (flags read in parallel), each marked one's lags are shared
// by the workgroup's waves, lanes run over t and a wave sum closes each lag.
#include <algorithm>

#include "fddm_kernels.hpp"
#include "fddm_wave.hpp"

namespace fddm {
namespace {

constexpr int kThreads = 256;

// Copies frames [0, e) and [L - e, L) of an (L, frame_bytes) stack into dst,
// 8 bytes per lane when the frame size allows it.
__global__ void __launch_bounds__(kThreads)
    k_gather_edges(const unsigned char* __restrict__ src, unsigned char* __restrict__ dst, uint64_t head_bytes,
                   uint64_t tail_src_off, uint64_t total_bytes, int wide) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kThreads;
  const uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (wide) {
    const uint64_t n = total_bytes / 8, head = head_bytes / 8, toff = tail_src_off / 8;
    const auto* s = reinterpret_cast<const uint64_t*>(src);
    auto* d = reinterpret_cast<uint64_t*>(dst);
    for (uint64_t i = i0; i < n; i += stride) d[i] = s[i < head ? i : i - head + toff];
  } else {
    for (uint64_t i = i0; i < total_bytes; i += stride)
      dst[i] = src[i < head_bytes ? i : i - head_bytes + tail_src_off];
  }
}

// One thread per output column o (y-shifted q): for every lag index i >= i0
// (those with n = L - lags[i] <= E), D = sum_{t < n} |F(t + dt) - F(t)|^2 / n
// from the FP64 edge spectra S64[q][c] (row pitch lp), where frame f sits in
// column f if f < E or the stack is short (L <= 2E), else f - (L - 2E).
// For E = kEdgeFrames and L > 2E the 2E values are loaded once (16 loads in
// flight) and D depends on n alone: frame t + dt = L - n + t sits in column
// 2E - n + t, so the E sums run over registers with compile-time indices
// and are then stored to the lag rows they belong to (same order of
// additions as the general loop, so the same values).
template <typename OutT>
__global__ void __launch_bounds__(kThreads)
    k_edge_fix(TimeArgs a, const double2* __restrict__ S64, uint64_t lp, int E, int i0) {
  const int c = blockIdx.x * kThreads + threadIdx.x;  // column of the launch
  if (c >= a.cols) return;
  const int q = col_q(a, c), L = a.L, nl = a.nl;
  const uint32_t* __restrict__ lags = a.lags;
  OutT* __restrict__ out = static_cast<OutT*>(a.out);
  const size_t o = col_out(a, c), ldo = a.ldo;
  const double2* __restrict__ f = S64 + static_cast<size_t>(q) * lp;
  if (E == kEdgeFrames && L > 2 * kEdgeFrames) {
    constexpr int EE = kEdgeFrames;
    double2 v[2 * EE];
#pragma unroll
    for (int k = 0; k < 2 * EE; ++k) v[k] = f[k];
    double dn[EE];  // dn[n - 1]: D of the lag with n pairs
#pragma unroll
    for (int n = 1; n <= EE; ++n) {
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < n; ++t) {
        const double2 x = v[2 * EE - n + t], y = v[t];
        const double dr = x.x - y.x, di = x.y - y.y;
        acc += dr * dr + di * di;
      }
      dn[n - 1] = acc / n;
    }
    for (int i = i0; i < nl; ++i) {
      const int n = L - static_cast<int>(lags[i]);  // 1 <= n <= EE
      double D = dn[0];
#pragma unroll
      for (int k = 1; k < EE; ++k)
        if (n == k + 1) D = dn[k];
      out[static_cast<size_t>(i) * ldo + o] = static_cast<OutT>(D);
    }
    return;
  }
  const int shift = L <= 2 * E ? 0 : L - 2 * E;
  for (int i = i0; i < nl; ++i) {
    const int d = static_cast<int>(lags[i]), n = L - d;
    double acc = 0.0;
    for (int t = 0; t < n; ++t) {
      const int e = t + d;
      const double2 x = f[e < E ? e : e - shift];
      const double2 y = f[t];  // t < n <= E
      const double dr = x.x - y.x, di = x.y - y.y;
      acc += dr * dr + di * di;
    }
    out[static_cast<size_t>(i) * ldo + o] = static_cast<OutT>(acc / n);
  }
}

template <typename OutT>
__global__ void __launch_bounds__(kThreads) k_wk_fix(TimeArgs a) {
  // workgroup g owns the columns g + m G (G workgroups); their flags are read
  // at once, one per thread, the marked columns listed in LDS, and the
  // waves of the workgroup share the lags of each listed column
  __shared__ int list[kThreads];
  __shared__ int count;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L = a.L, nl = a.nl, G = static_cast<int>(gridDim.x), g = static_cast<int>(blockIdx.x);
  const uint32_t* __restrict__ lags = a.lags;
  OutT* __restrict__ out = static_cast<OutT*>(a.out);
  const size_t ldo = a.ldo;
  const int M = (a.cols - g + G - 1) / G;  // columns of this workgroup (g < cols)
  for (int m0 = 0; m0 < M; m0 += kThreads) {
    if (threadIdx.x == 0) count = 0;
    __syncthreads();
    const int m = m0 + static_cast<int>(threadIdx.x);
    if (m < M) {
      const int c = g + m * G;
      if (a.flags[col_q(a, c)] != 0) list[atomicAdd(&count, 1)] = c;
    }
    __syncthreads();
    const int nmark = count;
    for (int k = 0; k < nmark; ++k) {
      const int c = list[k];
      const int q = col_q(a, c);
      const size_t oq = col_out(a, c);
      const double var = static_cast<double>(out[static_cast<size_t>(nl + 1) * ldo + oq]);
      const double thr = kWkFlag * static_cast<double>(L) * var;
      const float2* __restrict__ f = static_cast<const float2*>(a.S) + static_cast<size_t>(q - a.s_q0) * a.Lp;
      for (int b = wave * 64; b < nl; b += kThreads) {
        const int i = b + lane;
        bool need = false;
        if (i < nl) {
          const int n = L - static_cast<int>(lags[i]);
          need = n > 32 && static_cast<double>(n) * static_cast<double>(out[static_cast<size_t>(i) * ldo + oq]) < thr;
        }
        uint64_t todo = __ballot(need);
        while (todo) {
          const int ii = b + __ffsll(static_cast<long long>(todo)) - 1;
          todo &= todo - 1;
          const int d = static_cast<int>(lags[ii]), n = L - d;
          double acc = 0.0;
          for (int t = lane; t < n; t += 64) {
            const float2 x = f[t + d], y = f[t];
            const double dr = static_cast<double>(x.x) - static_cast<double>(y.x);
            const double di = static_cast<double>(x.y) - static_cast<double>(y.y);
            acc += dr * dr + di * di;
          }
          acc = wave::incl_scan(acc);
          if (lane == 63) out[static_cast<size_t>(ii) * ldo + oq] = static_cast<OutT>(acc / n);
        }
      }
    }
    __syncthreads();  // list and count are rewritten by the next round
  }
}

}  // namespace

hipError_t launch_wk_fix(const TimeArgs& a, int out_dtype, hipStream_t s) {
  if (a.nl <= 0 || a.cols <= 0 || !a.flags) return hipSuccess;
  const dim3 grid(static_cast<unsigned>(std::min(a.cols, 2048)));
  if (out_dtype == OUT_F32) {
    hipLaunchKernelGGL(k_wk_fix<float>, grid, dim3(kThreads), 0, s, a);
  } else {
    hipLaunchKernelGGL(k_wk_fix<double>, grid, dim3(kThreads), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_gather_edges(const void* img, void* dst, uint64_t frame_bytes, uint64_t L, uint64_t E,
                               hipStream_t s) {
  const uint64_t head = E * frame_bytes, total = 2 * E * frame_bytes, toff = (L - E) * frame_bytes;
  const int wide = (frame_bytes % 8 == 0 && reinterpret_cast<uintptr_t>(img) % 8 == 0 &&
                    reinterpret_cast<uintptr_t>(dst) % 8 == 0)
                       ? 1
                       : 0;
  const uint64_t items = wide ? total / 8 : total;
  const unsigned blocks = static_cast<unsigned>(std::min<uint64_t>((items + kThreads - 1) / kThreads, 4096));
  hipLaunchKernelGGL(k_gather_edges, dim3(blocks ? blocks : 1), dim3(kThreads), 0, s,
                     static_cast<const unsigned char*>(img), static_cast<unsigned char*>(dst), head, toff, total,
                     wide);
  return hipGetLastError();
}

hipError_t launch_edge_fix(const TimeArgs& a, const void* S64, uint64_t lp, int E, int i0, int out_dtype,
                           hipStream_t s) {
  if (i0 >= a.nl || a.cols <= 0) return hipSuccess;
  const dim3 grid((a.cols + kThreads - 1) / kThreads);
  if (out_dtype == OUT_F32) {
    hipLaunchKernelGGL(k_edge_fix<float>, grid, dim3(kThreads), 0, s, a, static_cast<const double2*>(S64), lp, E, i0);
  } else {
    hipLaunchKernelGGL(k_edge_fix<double>, grid, dim3(kThreads), 0, s, a, static_cast<const double2*>(S64), lp, E, i0);
  }
  return hipGetLastError();
}

}  // namespace fddm
