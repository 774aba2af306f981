// Stage (c) + (e) for series too long for an LDS-resident time transform.
//
// The time transform of length N = N1 x N2 (N2 = 4096, N1 = 2 .. 32) runs as
// a four-step FFT: with n = n1 + N1 n2 and k = k2 + N2 k1,
//   X[k2 + N2 k1] = sum_n1 W_N1^(n1 k1) W_N^(n1 k2) sum_n2 x[n1 + N1 n2] W_N2^(n2 k2),
// so step A transforms each decimated subsequence (n1 fixed) in LDS and
// multiplies by W_N^(n1 k2), and step B combines the N1 results of each k2
// with one radix-N1 butterfly in registers.  Per q, one workgroup of 256
// threads (persistent: workgroup w takes q = w, w + grid, ...) runs
//   0. the temporal mean (compensated paths) and the prefix sums of |G|^2,
//      G = F - mean, into its scratch slot;
//   1. G -> X (steps A and B), P = |X|^2 written in the order the second
//      transform's step A reads it;
//   2. P -> N C(dt) (steps A and B), keeping Re C(dt) for dt < L;
//   3. D for the requested lags (direct sums when L - dt <= 32), PS, VAR,
//      stored fftshifted along y.
// Each slot holds two N-point complex buffers and L + 1 prefix sums in
// device memory; they stay in the L2 / MALL while the workgroup works on
// them.  The compensated FP32 path marks q with small D against sum |G|^2 in
// a.flags exactly as the LDS-resident kernel does.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "fddm_kernels.hpp"
#include "fddm_p2.hpp"
#include "fddm_wave.hpp"

namespace fddm {
namespace {
using namespace p2;

constexpr int kLongThreads = 256;
constexpr int kLongLogN2 = 12;
constexpr int kN2 = 1 << kLongLogN2;
constexpr int kLongEdge = 32;

// Twiddle table layout (complex): pass 1 [k][r] (16 x 16), pass 2 [k][r]
// (256 x 16) of the radix-16 N2 transform, then W_N^e for e < N.
constexpr int kTwPass1 = 0, kTwPass2 = 256, kTwFour = 256 + 4096;

template <typename A>
__device__ __forceinline__ A group_sum(A x, A* red) {
  const int wave = threadIdx.x >> 6;
  x = wave::incl_scan(x);
  __syncthreads();  // red may still be read from an earlier call
  if ((threadIdx.x & 63) == 63) red[wave] = x;
  __syncthreads();
  A s = 0;
#pragma unroll
  for (int w = 0; w < kLongThreads / 64; ++w) s += red[w];
  return s;
}

// Exclusive scan of one value per thread over the workgroup; *total = sum.
template <typename A>
__device__ __forceinline__ A group_exclusive(A x, A* red, A* total) {
  const int wave = threadIdx.x >> 6;
  const A inc = wave::incl_scan(x);
  __syncthreads();
  if ((threadIdx.x & 63) == 63) red[wave] = inc;
  __syncthreads();
  A before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kLongThreads / 64; ++w) {
    if (w < wave) before += red[w];
    all += red[w];
  }
  *total = all;
  return before + inc - x;
}

// Step A of one four-step transform for subsequence s: the 16 first-pass
// inputs of thread `tid` are in v[0][*] already (elements tid + 256 r of the
// subsequence); three radix-16 passes (the last one kept in registers), then
// the outputs k2 = tid + 256 r are multiplied by W_N^(s k2) and stored to
// dst[s N2 + k2].
template <typename T>
__device__ __forceinline__ void step_a(cplx_t<T> (&v)[1][16], cplx_t<T>* lds, const cplx_t<T>* __restrict__ tw,
                                       int s, int logn, cplx_t<T>* __restrict__ dst, int tid) {
  using C = cplx_t<T>;
  using P0 = Pass<T, kN2, 16, 1, kLongThreads>;
  using P1 = Pass<T, kN2, 16, 16, kLongThreads>;
  using P2 = Pass<T, kN2, 16, 256, kLongThreads>;
  P0::compute_regs(v);
  __syncthreads();  // the previous user of lds is done
  P0::store(lds, v, tid);
  __syncthreads();
  {
    C w[1][16];
    P1::compute(lds, w, tw + kTwPass1, tid);
    __syncthreads();
    P1::store(lds, w, tid);
    __syncthreads();
  }
  P2::compute(lds, v, tw + kTwPass2, tid);
  const int nmask = (1 << logn) - 1;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int k2 = tid + 256 * r;
    const C t = tw[kTwFour + ((s * k2) & nmask)];
    dst[s * kN2 + k2] = cmul(v[0][r], t);
  }
}

template <typename T, typename OutT, bool SUB, int N1>
__global__ void __launch_bounds__(kLongThreads) k_time_long(TimeArgs a, int nq, int logn, void* scratch,
                                                            uint64_t slot_bytes) {
  using C = cplx_t<T>;
  using Acc = std::conditional_t<SUB, double, float>;
  constexpr bool F64 = sizeof(T) == 8;
  constexpr int N = N1 * kN2;
  extern __shared__ __align__(16) unsigned char smem[];
  C* const lds = reinterpret_cast<C*>(smem);
  Acc* const red = reinterpret_cast<Acc*>(smem + lds_padded_size(kN2) * sizeof(C));
  const int tid = threadIdx.x;
  const int L = a.L;
  const C* __restrict__ tw = static_cast<const C*>(a.twt);
  OutT* __restrict__ out = static_cast<OutT*>(a.out);
  unsigned char* const slot = static_cast<unsigned char*>(scratch) + blockIdx.x * slot_bytes;
  C* const bufA = reinterpret_cast<C*>(slot);
  C* const bufB = bufA + N;
  Acc* const psum = reinterpret_cast<Acc*>(bufB + N);  // psum[t] = sum_{s < t} |G(s)|^2
  Acc* const cr = reinterpret_cast<Acc*>(bufB);         // Re C(dt), dt < L (after step 2)
  const int seg = (L + kLongThreads - 1) / kLongThreads;
  const size_t nqs = a.ldo;

  for (int c = blockIdx.x; c < nq; c += gridDim.x) {  // columns of the launch
    const int q = col_q(a, c);
    const C* __restrict__ src = static_cast<const C*>(a.S) + static_cast<size_t>(q - a.s_q0) * a.Lp;
    // 0. mean and prefix sums over contiguous segments
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < seg; ++i) {
      const int t = tid * seg + i;
      if (t < L) {
        sr += static_cast<double>(src[t].x);
        si += static_cast<double>(src[t].y);
      }
    }
    sr = group_sum(sr, reinterpret_cast<double*>(red));
    si = group_sum(si, reinterpret_cast<double*>(red));
    const T m_r = SUB ? static_cast<T>(sr / L) : T(0), m_i = SUB ? static_cast<T>(si / L) : T(0);
    const double fsr = sr - static_cast<double>(L) * static_cast<double>(m_r);
    const double fsi = si - static_cast<double>(L) * static_cast<double>(m_i);
    Acc run = 0;
    for (int i = 0; i < seg; ++i) {
      const int t = tid * seg + i;
      if (t < L) {
        const Acc gr = static_cast<Acc>(src[t].x - m_r), gi = static_cast<Acc>(src[t].y - m_i);
        run += gr * gr + gi * gi;
      }
    }
    Acc ptot;
    Acc acc = group_exclusive(run, red, &ptot);
    for (int i = 0; i < seg; ++i) {
      const int t = tid * seg + i;
      if (t < L) {
        psum[t] = acc;
        const Acc gr = static_cast<Acc>(src[t].x - m_r), gi = static_cast<Acc>(src[t].y - m_i);
        acc += gr * gr + gi * gi;
      }
    }
    if (tid == 0) psum[L] = ptot;

    // 1. X = FFT_N(G): step A per subsequence n1, step B per k2
    for (int n1 = 0; n1 < N1; ++n1) {
      C v[1][16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int t = n1 + N1 * (tid + 256 * r);
        v[0][r] = t < L ? C{src[t].x - m_r, src[t].y - m_i} : C{0, 0};
      }
      step_a<T>(v, lds, tw, n1, logn, bufA, tid);
    }
    __syncthreads();  // bufA complete (global writes of the workgroup)
    for (int k2 = tid; k2 < kN2; k2 += kLongThreads) {
      C x[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) x[n1] = bufA[n1 * kN2 + k2];
      dft<T, N1>(x);
      // P[k], k = k2 + N2 k1, goes to the second transform's subsequence
      // m1 = k mod N1 at position m2 = k div N1
      const int m1 = k2 % N1, m2 = k2 / N1;
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1)
        bufB[m1 * kN2 + m2 + (kN2 / N1) * k1] = C{x[k1].x * x[k1].x + x[k1].y * x[k1].y, T(0)};
    }
    __syncthreads();

    // 2. FFT_N(P) = N conj C: step A per subsequence m1, step B per j2
    for (int m1 = 0; m1 < N1; ++m1) {
      C v[1][16];
#pragma unroll
      for (int r = 0; r < 16; ++r) v[0][r] = bufB[m1 * kN2 + tid + 256 * r];
      step_a<T>(v, lds, tw, m1, logn, bufA, tid);
    }
    __syncthreads();
    const Acc inv_n = static_cast<Acc>(1) / static_cast<Acc>(N);
    for (int j2 = tid; j2 < kN2; j2 += kLongThreads) {
      C x[N1];
#pragma unroll
      for (int m1 = 0; m1 < N1; ++m1) x[m1] = bufA[m1 * kN2 + j2];
      dft<T, N1>(x);
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        const int dt = j2 + kN2 * j1;
        if (dt < L) cr[dt] = static_cast<Acc>(x[j1].x) * inv_n;
      }
    }
    __syncthreads();

    // 3. lags, PS, VAR
    const size_t oq = col_out(a, c);
    bool flag = false;
    for (int i = tid; i < a.nl; i += kLongThreads) {
      const int d = static_cast<int>(a.lags[i]), n = L - d;
      Acc D;
      if (n <= kLongEdge) {
        Acc s = 0;
        for (int t = 0; t < n; ++t) {
          const C x = src[t + d], y = src[t];
          const Acc dr = static_cast<Acc>(x.x) - static_cast<Acc>(y.x);
          const Acc di = static_cast<Acc>(x.y) - static_cast<Acc>(y.y);
          s += dr * dr + di * di;
        }
        D = s / n;
      } else {
        D = (psum[n] + (psum[L] - psum[d]) - 2 * cr[d]) / n;
        if constexpr (SUB && !F64) flag |= static_cast<Acc>(n) * D * static_cast<Acc>(1.0 / kWkFlag) < psum[L];
      }
      out[static_cast<size_t>(i) * nqs + oq] = static_cast<OutT>(D);
    }
    if constexpr (SUB && !F64) {
      const bool any = __syncthreads_or(flag);
      if (a.flags && tid == 0) a.flags[q] = any ? 1 : 0;
    }
    if (tid == 0) {
      const Acc ptot_l = psum[L];
      const Acc mr2 = static_cast<Acc>(fsr / L), mi2 = static_cast<Acc>(fsi / L);
      Acc ps, var;
      if constexpr (SUB) {
        var = ptot_l / L - (mr2 * mr2 + mi2 * mi2);
        Acc fr = static_cast<Acc>(m_r) + mr2, fi = static_cast<Acc>(m_i) + mi2;
        if (a.S0) {
          const C z = static_cast<const C*>(a.S0)[static_cast<size_t>(q) * a.s0_stride];
          fr += static_cast<Acc>(z.x);
          fi += static_cast<Acc>(z.y);
        }
        ps = var + fr * fr + fi * fi;
      } else {
        ps = ptot_l / L;
        var = ps - (mr2 * mr2 + mi2 * mi2);
      }
      out[static_cast<size_t>(a.nl) * nqs + oq] = static_cast<OutT>(ps);
      out[static_cast<size_t>(a.nl + 1) * nqs + oq] = static_cast<OutT>(var);
    }
    __syncthreads();  // the slot is rewritten by the next q
  }
}

template <typename T, typename OutT, bool SUB>
hipError_t launch_long_t(const TimeArgs& a, int logn, void* scratch, uint64_t slot_bytes, int slots, hipStream_t s) {
  const int nq = a.cols;
  const size_t lds = lds_padded_size(kN2) * sizeof(cplx_t<T>) + 8 * sizeof(double);
  const dim3 grid(static_cast<unsigned>(std::min(slots, nq)));
  auto go = [&](auto kernel) {
    if (lds > 64 * 1024) {
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kernel, grid, dim3(kLongThreads), lds, s, a, nq, logn, scratch, slot_bytes);
    return hipGetLastError();
  };
  switch (logn - kLongLogN2) {
    case 1: return go(k_time_long<T, OutT, SUB, 2>);
    case 2: return go(k_time_long<T, OutT, SUB, 4>);
    case 3: return go(k_time_long<T, OutT, SUB, 8>);
    case 4: return go(k_time_long<T, OutT, SUB, 16>);
    case 5:
      if constexpr (sizeof(T) == 4) return go(k_time_long<T, OutT, SUB, 32>);
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

int long_max_log(int precision) { return precision == PREC_FP64 ? kLongLogN2 + 4 : kLongLogN2 + 5; }

uint64_t long_slot_bytes(int logn, int precision, uint64_t L) {
  const uint64_t c = precision == PREC_FP64 ? 16 : 8;
  const uint64_t b = 2 * (uint64_t(1) << logn) * c + (L + 1) * 8;
  return (b + 255) & ~uint64_t(255);
}

std::vector<double> long_twiddles(int logn) {
  // pass tables [k][r] = exp(-2 pi i r k / (NS R)) for (R, NS) = (16, 16) and
  // (16, 256), then exp(-2 pi i e / N) for e < N; extended precision
  std::vector<double> w;
  const long double two_pi = 6.283185307179586476925286766559005768L;
  auto pass = [&](long long R, long long NS) {
    for (long long k = 0; k < NS; ++k)
      for (long long r = 0; r < R; ++r) {
        const long double ang = -two_pi * static_cast<long double>(r * k) / static_cast<long double>(NS * R);
        w.push_back(static_cast<double>(cosl(ang)));
        w.push_back(static_cast<double>(sinl(ang)));
      }
  };
  pass(16, 16);
  pass(16, 256);
  const long long n = 1ll << logn;
  for (long long e = 0; e < n; ++e) {
    const long double ang = -two_pi * static_cast<long double>(e) / static_cast<long double>(n);
    w.push_back(static_cast<double>(cosl(ang)));
    w.push_back(static_cast<double>(sinl(ang)));
  }
  return w;
}

hipError_t launch_time_long(const TimeArgs& a, int logn, int precision, int out_dtype, void* scratch,
                            uint64_t slot_bytes, int slots, hipStream_t s) {
  const bool f32 = out_dtype == OUT_F32;
  switch (precision) {
    case PREC_FP64:
      return f32 ? launch_long_t<double, float, true>(a, logn, scratch, slot_bytes, slots, s)
                 : launch_long_t<double, double, true>(a, logn, scratch, slot_bytes, slots, s);
    case PREC_FP32:
      return f32 ? launch_long_t<float, float, false>(a, logn, scratch, slot_bytes, slots, s)
                 : launch_long_t<float, double, false>(a, logn, scratch, slot_bytes, slots, s);
    case PREC_FP32C:
      return f32 ? launch_long_t<float, float, true>(a, logn, scratch, slot_bytes, slots, s)
                 : launch_long_t<float, double, true>(a, logn, scratch, slot_bytes, slots, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fddm
