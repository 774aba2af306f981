// Stage (c) + (e) with power-of-two time transforms.
//
// D(q, dt) does not depend on the zero-padded transform length N as long as
// N >= 2L - 1, so the time stage always runs at N = 2^ceil(log2(2L - 1))
// (N >= 512) with plans fixed at compile time (p2_plan in fddm_common.hpp):
// three passes per transform with radix 4..32 butterflies, all strides and
// LDS offsets known to the compiler, and padded LDS addresses (one slot per
// 32 elements) that stay linear in the butterfly index.
//
// Per q, one group of T threads:
//   1. loads its first-pass inputs F(t), t = u + b T + r N/R0, straight from
//      HBM (t >= N/2 >= L is the zero padding and is never read);
//   2. compensated paths: subtracts the temporal mean, records the first and
//      last kEdge samples, forms prefix sums of |F(t)|^2 in double (chunked
//      wave scans);
//   3. FFT1 (R0, R1, R2); the last pass stays in registers, |G|^2 is formed
//      there and FFT2's first pass (radix R2, same distribution) runs on it;
//   4. FFT2 (R2, R1, R0); the last pass leaves Re C(t) (times N) in the same
//      registers as the loaded F(t);
//   5. writes C and the prefix sums to LDS, then evaluates D for every
//      requested lag (direct sums when L - dt <= kEdge), PS and VAR, and
//      stores them fftshifted along y.
#include <type_traits>
#include <utility>

#include "fddm_kernels.hpp"
#include "fddm_p2.hpp"

namespace fddm {
namespace {
using namespace p2;

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kEdge = 32;

template <typename A>
__device__ __forceinline__ A wave_sum(A v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename A>
__device__ __forceinline__ A wave_incl_scan(A v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const A w = __shfl_up(v, o, 64);
    if (lane >= o) v += w;
  }
  return v;
}

// Sums of x and y over the WPT waves of this thread's group (one barrier);
// red holds 2 kWaves doubles and is written once per launch.
template <int WPT>
__device__ __forceinline__ void group_sum2(double& x, double& y, double* red) {
  const int wave = threadIdx.x >> 6;
  x = wave_sum(x);
  y = wave_sum(y);
  if ((threadIdx.x & 63) == 0) {
    red[2 * wave] = x;
    red[2 * wave + 1] = y;
  }
  __syncthreads();
  const int first = (wave / WPT) * WPT;
  double sx = 0, sy = 0;
#pragma unroll
  for (int w = 0; w < WPT; ++w) {
    sx += red[2 * (first + w)];
    sy += red[2 * (first + w) + 1];
  }
  x = sx;
  y = sy;
}

// Twiddle table offsets (complex entries): FFT1 passes 1, 2 then FFT2 passes 1, 2;
// the pass with radix R after NS points holds exp(-2 pi i r k / (NS R)) at k R + r.
template <int LOGN>
struct P2Tw {
  static constexpr P2Plan P = p2_plan(LOGN);
  static constexpr int N = 1 << LOGN;
  static constexpr int a1 = 0;
  static constexpr int a2 = P.r0 * P.r1;
  static constexpr int b1 = a2 + N;
  static constexpr int b2 = b1 + P.r2 * P.r1;
  static constexpr int total = b2 + N;
};

template <int LOGN, typename T, bool SUB>
struct P2Lds {
  static constexpr P2Plan P = p2_plan(LOGN);
  static constexpr int N = 1 << LOGN, TPT = P.threads, G = kThreads / TPT, WPT = TPT / 64;
  static constexpr int MB0 = N / P.r0 / TPT;
  static constexpr int NCH = (P.r0 / 2) * MB0;
  using Acc = std::conditional_t<SUB, double, float>;
  static constexpr size_t csz = sizeof(cplx_t<T>);
  static constexpr size_t buf = static_cast<size_t>(G) * lds_padded_size(N) * csz;
  static constexpr size_t edge = static_cast<size_t>(G) * 2 * kEdge * csz;
  static constexpr size_t red = 2 * kWaves * sizeof(double);
  static constexpr size_t chunk = static_cast<size_t>(G) * WPT * sizeof(Acc);
  static constexpr size_t bytes = buf + edge + red + chunk;
};

template <typename T, typename OutT, int LOGN, bool SUB>
__global__ void __launch_bounds__(kThreads) k_time_p2(TimeArgs a, int wg_per_xcd, int nq) {
  using C = cplx_t<T>;
  using Lay = P2Lds<LOGN, T, SUB>;
  using Acc = typename Lay::Acc;
  using Tw = P2Tw<LOGN>;
  constexpr P2Plan P = p2_plan(LOGN);
  constexpr int N = 1 << LOGN, TPT = P.threads, G = Lay::G, WPT = Lay::WPT;
  constexpr int R0 = P.r0, R1 = P.r1, R2 = P.r2;
  using A0 = Pass<T, N, R0, 1, TPT>;
  using A1 = Pass<T, N, R1, R0, TPT>;
  using A2 = Pass<T, N, R2, R0 * R1, TPT>;
  using B0 = Pass<T, N, R2, 1, TPT>;
  using B1 = Pass<T, N, R1, R2, TPT>;
  using B2 = Pass<T, N, R0, R2 * R1, TPT>;
  static_assert(A2::MB == B0::MB && B2::MB == A0::MB, "register hand-offs");
  constexpr int MB0 = A0::MB, M0 = A0::M, RH = R0 / 2;

  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int g = tid / TPT, u = tid - g * TPT, lane = tid & 63, wg = u >> 6;
  C* const buf = reinterpret_cast<C*>(smem) + g * lds_padded_size(N);
  C* const edge = reinterpret_cast<C*>(smem + Lay::buf) + g * 2 * kEdge;
  double* const red = reinterpret_cast<double*>(smem + Lay::buf + Lay::edge);
  Acc* const chunk = reinterpret_cast<Acc*>(smem + Lay::buf + Lay::edge + Lay::red) + g * WPT;

  // Workgroups are dealt round-robin to the 8 XCDs: each XCD gets a
  // contiguous range of q so neighbouring output columns share its L2.
  const int q = ((blockIdx.x & 7) * wg_per_xcd + (blockIdx.x >> 3)) * G + g;
  const bool active = q < nq;
  const int L = a.L;
  const C* __restrict__ src = static_cast<const C*>(a.S) + static_cast<size_t>(active ? q : 0) * a.Lp;
  const C* __restrict__ tw = static_cast<const C*>(a.twt);

  // 1. first-pass inputs straight from HBM
  C v0[MB0][R0];
#pragma unroll
  for (int b = 0; b < MB0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int t = u + b * TPT + r * M0;
      v0[b][r] = (r < RH && active && t < L) ? src[t] : C{0, 0};
    }

  // 2. temporal mean (the compensated paths transform F - mean)
  double sr = 0.0, si = 0.0;
#pragma unroll
  for (int b = 0; b < MB0; ++b)
#pragma unroll
    for (int r = 0; r < RH; ++r) {
      sr += static_cast<double>(v0[b][r].x);
      si += static_cast<double>(v0[b][r].y);
    }
  group_sum2<WPT>(sr, si, red);
  const double mr = sr / L, mi = si / L;
  const T m_r = SUB ? static_cast<T>(mr) : T(0), m_i = SUB ? static_cast<T>(mi) : T(0);
  if constexpr (SUB) {
#pragma unroll
    for (int b = 0; b < MB0; ++b)
#pragma unroll
      for (int r = 0; r < RH; ++r) {
        const int t = u + b * TPT + r * M0;
        if (t < L) {
          v0[b][r].x -= m_r;
          v0[b][r].y -= m_i;
        }
      }
  }
  // sum of the (subtracted) series, for VAR and PS
  const double fsr = sr - static_cast<double>(L) * static_cast<double>(m_r);
  const double fsi = si - static_cast<double>(L) * static_cast<double>(m_i);

  // 3. FFT1; its last pass stays in registers
  A0::compute_regs(v0);
  A0::store(buf, v0, u);
  __syncthreads();
  C v1[A1::MB][R1];
  A1::compute(buf, v1, tw + Tw::a1, u);
  __syncthreads();
  A1::store(buf, v1, u);
  __syncthreads();
  C v2[A2::MB][R2];
  A2::compute(buf, v2, tw + Tw::a2, u);
#pragma unroll
  for (int b = 0; b < A2::MB; ++b)
#pragma unroll
    for (int r = 0; r < R2; ++r) v2[b][r] = C{v2[b][r].x * v2[b][r].x + v2[b][r].y * v2[b][r].y, 0};
  // 4. FFT2; first pass on the registers, last pass back into registers
  B0::compute_regs(v2);
  __syncthreads();
  B0::store(buf, v2, u);
  __syncthreads();
  B1::compute(buf, v1, tw + Tw::b1, u);
  __syncthreads();
  B1::store(buf, v1, u);
  __syncthreads();
  B2::compute(buf, v0, tw + Tw::b2, u);
  __syncthreads();

  // 5. C(t) N into LDS; prefix sums psum(t + 1) of |F(t) - mean|^2 from a
  // second (cache-resident) read of the series, in chunks of T consecutive t
  // (wave scans combined across the group's waves), into the free tail of the
  // buffer; first and last kEdge samples into `edge`.
  Acc* const tail = reinterpret_cast<Acc*>(buf + padi(L));
#pragma unroll
  for (int b = 0; b < MB0; ++b)
#pragma unroll
    for (int r = 0; r < RH; ++r) {
      const int t = u + b * TPT + r * M0;
      if (t < L) buf[padi(t)].x = v0[b][r].x;
    }
  // each thread: a contiguous segment of seg <= SEGMAX samples, local running
  // sums, then one exclusive scan over the group's threads; seg is odd, so
  // the 8-byte stores of neighbouring threads fall in different LDS banks
  constexpr int SEGMAX = ((N / 2 + TPT - 1) / TPT) | 1;
  const int seg = ((L + TPT - 1) / TPT) | 1;
  const int ts = u * seg;
  Acc pre[SEGMAX];
  Acc run = 0;
#pragma unroll
  for (int i = 0; i < SEGMAX; ++i) {
    const int t = ts + i;
    Acc val = 0;
    if (i < seg && active && t < L) {
      C f = src[t];
      f.x -= m_r;
      f.y -= m_i;
      if constexpr (SUB) {
        if (t < kEdge) edge[t] = f;
        if (t >= L - kEdge) edge[t - L + 2 * kEdge] = f;
      }
      val = static_cast<Acc>(f.x) * static_cast<Acc>(f.x) + static_cast<Acc>(f.y) * static_cast<Acc>(f.y);
    }
    run += val;
    pre[i] = run;
  }
  const Acc inc = wave_incl_scan(run);
  Acc exc = __shfl_up(inc, 1, 64);
  if (lane == 0) exc = 0;
  if (lane == 63) chunk[wg] = inc;
  __syncthreads();
  Acc ptot = 0;
#pragma unroll
  for (int w = 0; w < WPT; ++w) {
    const Acc x = chunk[w];
    if (w < wg) exc += x;
    ptot += x;
  }
#pragma unroll
  for (int i = 0; i < SEGMAX; ++i) {
    const int t = ts + i;
    if (i < seg && t < L - 1) tail[t] = pre[i] + exc;
  }
  __syncthreads();
  if (!active) return;

  OutT* __restrict__ out = static_cast<OutT*>(a.out);
  const size_t nqs = static_cast<size_t>(a.ny) * a.nxh;
  const int ky = q / a.nxh, kx = q - ky * a.nxh;
  const size_t oq = static_cast<size_t>((ky + a.ny / 2) % a.ny) * a.nxh + kx;
  const Acc inv_n = static_cast<Acc>(1) / static_cast<Acc>(N);
  for (int i = u; i < a.nl; i += TPT) {
    const int d = static_cast<int>(a.lags[i]);
    Acc D;
    if (SUB && L - d <= kEdge) {
      Acc acc = 0;
      for (int t = 0; t < L - d; ++t) {
        const int s = t + d;
        const C x = s < kEdge ? edge[s] : edge[s - L + 2 * kEdge];
        const C y = edge[t];
        const Acc dr = static_cast<Acc>(x.x) - static_cast<Acc>(y.x);
        const Acc di = static_cast<Acc>(x.y) - static_cast<Acc>(y.y);
        acc += dr * dr + di * di;
      }
      D = acc / static_cast<Acc>(L - d);
    } else {
      const Acc head = tail[L - d - 1];   // psum(L - d)
      const Acc tl = ptot - tail[d - 1];  // psum(L) - psum(d)
      const Acc c = static_cast<Acc>(buf[padi(d)].x) * inv_n;
      D = (head + tl - 2 * c) / static_cast<Acc>(L - d);
    }
    out[static_cast<size_t>(i) * nqs + oq] = static_cast<OutT>(D);
  }
  if (u == 0) {
    const Acc mr2 = static_cast<Acc>(fsr / L), mi2 = static_cast<Acc>(fsi / L);
    Acc ps, var;
    if constexpr (SUB) {
      var = ptot / L - (mr2 * mr2 + mi2 * mi2);
      Acc fr = static_cast<Acc>(m_r) + mr2, fi = static_cast<Acc>(m_i) + mi2;
      if (a.S0) {
        const C z = static_cast<const C*>(a.S0)[static_cast<size_t>(q) * a.s0_stride];
        fr += static_cast<Acc>(z.x);
        fi += static_cast<Acc>(z.y);
      }
      ps = var + fr * fr + fi * fi;
    } else {
      ps = ptot / L;
      var = ps - (mr2 * mr2 + mi2 * mi2);
    }
    out[static_cast<size_t>(a.nl) * nqs + oq] = static_cast<OutT>(ps);
    out[static_cast<size_t>(a.nl + 1) * nqs + oq] = static_cast<OutT>(var);
  }
}

template <typename T, typename OutT, int LOGN, bool SUB>
hipError_t launch_one(const TimeArgs& a, hipStream_t s) {
  using Lay = P2Lds<LOGN, T, SUB>;
  const int nq = a.ny * a.nxh;
  const int wgs = (nq + Lay::G - 1) / Lay::G;
  const int wpx = (wgs + 7) / 8;
  const size_t lds = Lay::bytes;
  auto kernel = k_time_p2<T, OutT, LOGN, SUB>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kernel, dim3(8u * wpx), dim3(kThreads), lds, s, a, wpx, nq);
  return hipGetLastError();
}

template <typename T, typename OutT, bool SUB>
hipError_t launch_logn(const TimeArgs& a, int logn, hipStream_t s) {
  switch (logn) {
    case 9: return launch_one<T, OutT, 9, SUB>(a, s);
    case 10: return launch_one<T, OutT, 10, SUB>(a, s);
    case 11: return launch_one<T, OutT, 11, SUB>(a, s);
    case 12: return launch_one<T, OutT, 12, SUB>(a, s);
    case 13: return launch_one<T, OutT, 13, SUB>(a, s);
    case 14:
      if constexpr (sizeof(T) == 4) return launch_one<T, OutT, 14, SUB>(a, s);
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

size_t p2_lds_bytes(int logn, int precision) {
  auto pick = [&](auto tag, auto sub) -> size_t {
    using T = decltype(tag);
    constexpr bool S = decltype(sub)::value;
    switch (logn) {
      case 9: return P2Lds<9, T, S>::bytes;
      case 10: return P2Lds<10, T, S>::bytes;
      case 11: return P2Lds<11, T, S>::bytes;
      case 12: return P2Lds<12, T, S>::bytes;
      case 13: return P2Lds<13, T, S>::bytes;
      case 14: return P2Lds<14, T, S>::bytes;
      default: return 0;
    }
  };
  if (precision == PREC_FP64) return pick(double{}, std::true_type{});
  if (precision == PREC_FP32) return pick(float{}, std::false_type{});
  return pick(float{}, std::true_type{});
}

int p2_twiddle_count(int logn) {
  switch (logn) {
    case 9: return P2Tw<9>::total;
    case 10: return P2Tw<10>::total;
    case 11: return P2Tw<11>::total;
    case 12: return P2Tw<12>::total;
    case 13: return P2Tw<13>::total;
    case 14: return P2Tw<14>::total;
    default: return 0;
  }
}

hipError_t launch_time_p2(const TimeArgs& a, int logn, int precision, int out_dtype, hipStream_t s) {
  const bool f32 = out_dtype == OUT_F32;
  switch (precision) {
    case PREC_FP64:
      return f32 ? launch_logn<double, float, true>(a, logn, s) : launch_logn<double, double, true>(a, logn, s);
    case PREC_FP32:
      return f32 ? launch_logn<float, float, false>(a, logn, s) : launch_logn<float, double, false>(a, logn, s);
    case PREC_FP32C:
      return f32 ? launch_logn<float, float, true>(a, logn, s) : launch_logn<float, double, true>(a, logn, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fddm
