// Stage (c) + (e) with power-of-two time transforms.
//
// D(q, dt) does not depend on the zero-padded transform length N as long as
// N >= 2L - 1, so the time stage always runs at N = 2^ceil(log2(2L - 1))
// (N >= 512) with plans fixed at compile time (p2_plan in fddm_common.hpp):
// three or four passes per transform with radix 4..16 butterflies, all
// strides and LDS offsets known to the compiler, and padded LDS addresses
// (one slot per 32 elements) that stay linear in the butterfly index.  One
// transform is worked on by 64 .. 1024 threads; plans of fewer than 256
// threads put several q in one workgroup.
//
// Per q, one group of T threads:
//   1. loads its first-pass inputs F(t), t = u + b T + r N/R0, straight from
//      HBM (t >= N/2 >= L is the zero padding and is never read);
//   2. compensated paths: subtracts the temporal mean (one group reduction);
//   3. FFT1 (radices R0 .. RL); the last pass stays in registers, |G|^2 is
//      formed there and written to LDS as z(n) = P(2n) + i P(2n+1);
//   4. forms the prefix sums of |F(t)|^2 (double on the compensated paths)
//      and keeps the first and last kEdge samples: from the first pass's
//      registers before step 3 (wave scans per segment of the register
//      layout, held in registers until a buffer half is free), or, on the
//      paired N = 4096 plans and at N = 16384, whose registers are full, from
//      a second (cache-resident) read of the series with contiguous segments
//      per thread and one group scan;
//   5. the second forward transform of the real P runs as one complex
//      transform of length N/2 of z (p2_half_plan) in the other half (in
//      place, or alternating between the halves before step 4 on the paired
//      N = 4096 plan); for each lag Re FFT_N(P)(d) = N C(d) is untangled
//      from Z(d) and Z(N/2 - d);
//   6. evaluates D for every requested lag (direct sums when L - dt <= kEdge),
//      PS and VAR, and stores them fftshifted along y.
//
// The output is lag-major, so each lag store of one q writes one value into
// its own line.  Workgroups of several q interleave them across lanes; the
// plans from N = 8192 (FP64 from 4096), whose LDS holds a single q, instead
// run steps 1-6 for two neighbouring output columns in turn, keep the first
// column's D in registers and store each lag's two values together.
#include <type_traits>
#include <utility>

#include "fddm_kernels.hpp"
#include "fddm_p2.hpp"
#include "fddm_wave.hpp"

namespace fddm {
namespace {
using namespace p2;

constexpr int kMinThreads = 256;  // workgroup size for plans of fewer threads (G q per group)
constexpr int kEdge = 32;
// Compensated FP32: the float transforms leave an absolute error of at most
// about 3e-7 x sum_t |G|^2 in (L - dt) D (measured worst case over 2 x 10^9
// cells at N = 8192).  A q with a lag whose (L - dt) D falls below
// kWkFlag x sum |G|^2 (2e-3: a relative error bound above ~1.5e-4) -- slow
// low-q modes at short lags, chance near-recurrences at large lags -- is
// marked in a.flags; launch_wk_fix then recomputes those lags by direct
// differences.

// 1 / n for a positive integer n: hardware reciprocal refined by Newton's
// iteration r <- r (2 - n r), twice for double (error below 2 ulp).
template <typename A>
__device__ __forceinline__ A recip(int n) {
  if constexpr (sizeof(A) == 8) {
    const double x = static_cast<double>(n);
    double r = __builtin_amdgcn_rcp(x);
    r = r * (2.0 - x * r);
    r = r * (2.0 - x * r);
    return r;
  } else {
    return 1.0f / static_cast<float>(n);
  }
}

// Sums of x and y over the WPT waves of this thread's group (one barrier);
// red holds two doubles per wave of the workgroup and is written once.
template <int WPT>
__device__ __forceinline__ void group_sum2(double& x, double& y, double* red) {
  const int wave = threadIdx.x >> 6;
  x = wave::incl_scan(x);
  y = wave::incl_scan(y);
  if ((threadIdx.x & 63) == 63) {  // the wave totals
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

// Two neighbouring output values, stored as one access.  A vector type keeps
// the store whole: the store of a two-member struct is split into two scalar
// stores, which the compiler may then merge with the unpaired fallback path.
typedef double OutPairD __attribute__((ext_vector_type(2)));
typedef float OutPairF __attribute__((ext_vector_type(2)));
template <typename O>
using OutPair = std::conditional_t<sizeof(O) == 8, OutPairD, OutPairF>;

// Workgroup shape and LDS layout of the time kernel at N = 2^LOGN.  Plans of
// fewer than kMinThreads threads put G = kMinThreads / TPT q in a workgroup;
// plans of 256 or 512 threads put two q in one when two such workgroups
// still fit a CU's LDS.  The q of one workgroup are neighbours, and the lag
// loop interleaves them across lanes, so each lag store covers G adjacent
// output values instead of one.
template <int LOGN, typename T, bool SUB>
struct P2Lds {
  static constexpr P2Plan P = p2_plan(LOGN, sizeof(T) == 8);
  static constexpr int N = 1 << LOGN, TPT = P.threads, WPT = TPT / 64;
  using Acc = std::conditional_t<SUB, double, float>;
  static constexpr size_t csz = sizeof(cplx_t<T>);
  static constexpr int S = N / 2 / TPT;  // first-pass segments of TPT samples below N/2
  static constexpr size_t group_bytes0 =
      lds_padded_size(N) * csz + 2 * kEdge * csz + 2 * WPT * sizeof(double) + (S * WPT + 1) * sizeof(Acc);
  // (the uncompensated FP32 N = 4096 kernel spills 12 VGPRs at 8 waves per
  // SIMD and is still faster paired: 3.28 -> 3.09 ms at 512^2 x 2000)
  static constexpr bool kPair = TPT >= 256 && TPT <= 512 && 4 * group_bytes0 <= 160 * 1024;
  // Prefix sums formed from the first pass's registers (no second read of the
  // series) where the registers allow it: not on the paired plans, whose 8
  // waves per SIMD leave no room to hold S sums across the first transform,
  // nor at N = 16384.  Those re-read the series from cache.
  static constexpr bool REG = !kPair && LOGN <= 13;
  // per q group: wave totals (REG: by segment and wave, then their offsets)
  // and, last, their sum
  static constexpr int CH = (REG ? S : 1) * WPT + 1;
  static constexpr int NT = TPT < kMinThreads ? kMinThreads : (kPair ? 2 * TPT : TPT);  // workgroup size
  static constexpr int G = NT / TPT;
  // Plans from N = 8192 (FP64 from N = 4096) hold one q per workgroup (the
  // LDS has no room for a second) and leave registers free: the workgroup
  // transforms the q of two neighbouring output columns one after the other
  // and keeps the first one's D in registers (KMAX lag iterations per
  // thread), so each lag store writes both columns at once.  (Plain FP32 at
  // N = 16384 would spill.)
  static constexpr int SEQ =
      (!kPair && TPT >= kMinThreads && (LOGN >= 13 || sizeof(T) == 8) && (SUB || LOGN == 13)) ? 2 : 1;
  static constexpr int KMAX = (N / 2 + NT - 1) / NT;
  // Register bound: paired workgroups, two per CU, hold 8 waves per SIMD when
  // at most 64 VGPRs are used; two-column FP32 plans get the 4 waves per SIMD
  // their LDS allows (FP64 at N = 8192 has one workgroup of 4 waves per CU).
  static constexpr int kMinWaves = kPair ? 8 : (SEQ > 1 && sizeof(T) == 4 ? 4 : 1);
  static constexpr size_t buf = static_cast<size_t>(G) * lds_padded_size(N) * csz;
  static constexpr size_t edge = static_cast<size_t>(G) * 2 * kEdge * csz;
  static constexpr size_t red = 2 * (NT / 64) * sizeof(double);
  static constexpr size_t chunk = static_cast<size_t>(G) * CH * sizeof(Acc);
  static constexpr size_t bytes = buf + edge + red + chunk;
};

// LDS-to-LDS passes I .. LAST-1 of a transform (HALF = false: the first
// transform, plan p2_plan, length 2^LOGN; HALF = true: the half-length second
// transform, plan p2_half_plan, length 2^(LOGN-1)).
template <typename T, int LOGN, bool HALF, int I, int LAST>
__device__ __forceinline__ void p2_passes(cplx_t<T>* buf, const cplx_t<T>* __restrict__ tw, int u) {
  constexpr bool F64 = sizeof(T) == 8;
  constexpr P2Plan P = HALF ? p2_half_plan(LOGN, F64) : p2_plan(LOGN, F64);
  if constexpr (I < LAST) {
    constexpr int N = HALF ? 1 << (LOGN - 1) : 1 << LOGN;
    constexpr int R = p2_radix(P, I);
    constexpr int NS = p2_ns(P, I);
    constexpr int OFF = HALF ? p2_tw_h(LOGN, F64, I) : p2_tw_a(LOGN, F64, I);
    using Ps = Pass<T, N, R, NS, P.threads>;
    cplx_t<T> v[Ps::MB][R];
    Ps::compute(buf, v, tw + OFF, u);
    __syncthreads();
    Ps::store(buf, v, u);
    __syncthreads();
    p2_passes<T, LOGN, HALF, I + 1, LAST>(buf, tw, u);
  }
}

// Passes I .. LAST-1 of the half-length transform, alternating between the
// buffer halves a and b: one barrier per pass instead of two.
template <typename T, int LOGN, int I, int LAST>
__device__ __forceinline__ void p2_half_pp(cplx_t<T>* a, cplx_t<T>* b, const cplx_t<T>* __restrict__ tw, int u) {
  constexpr bool F64 = sizeof(T) == 8;
  constexpr P2Plan P = p2_half_plan(LOGN, F64);
  if constexpr (I < LAST) {
    constexpr int N = 1 << (LOGN - 1);
    constexpr int R = p2_radix(P, I);
    constexpr int NS = p2_ns(P, I);
    constexpr int OFF = p2_tw_h(LOGN, F64, I);
    using Ps = Pass<T, N, R, NS, P.threads>;
    cplx_t<T> v[Ps::MB][R];
    Ps::compute(a, v, tw + OFF, u);
    Ps::store(b, v, u);
    __syncthreads();
    p2_half_pp<T, LOGN, I + 1, LAST>(b, a, tw, u);
  }
}

template <typename T, typename OutT, int LOGN, bool SUB>
__global__ void __launch_bounds__((P2Lds<LOGN, T, SUB>::NT), (P2Lds<LOGN, T, SUB>::kMinWaves))
    k_time_p2(TimeArgs a, int wg_per_xcd, int nq) {
  using C = cplx_t<T>;
  using Lay = P2Lds<LOGN, T, SUB>;
  using Acc = typename Lay::Acc;
  constexpr P2Plan P = Lay::P;
  constexpr int N = 1 << LOGN, TPT = P.threads, G = Lay::G, WPT = Lay::WPT, NP = P.np;
  constexpr bool F64 = sizeof(T) == 8;
  constexpr int R0 = P.r0, RL = p2_radix(P, NP - 1), NH = N / 2;
  using A0 = Pass<T, N, R0, 1, TPT>;                          // first pass (registers)
  using AL = Pass<T, N, RL, N / RL, TPT>;                     // last pass (to registers)
  static_assert(N % (R0 * TPT) == 0, "whole first-pass butterflies per thread");
  static_assert((N / RL) % TPT == 0, "whole last-pass butterflies per thread (|G|^2 is formed from registers)");
  constexpr int TWA = p2_tw_a(LOGN, F64, NP - 1), TWP = p2_tw_post(LOGN, F64);
  constexpr int MB0 = A0::MB, M0 = A0::M, RH = R0 / 2;

  extern __shared__ __align__(16) unsigned char smem[];

  // Workgroups are dealt round-robin to the 8 XCDs: each XCD gets a
  // contiguous range of q (of output columns when SEQ > 1) so neighbouring
  // output columns share its L2.
  constexpr int SEQ = Lay::SEQ;
  const int wgi = (blockIdx.x & 7) * wg_per_xcd + (blockIdx.x >> 3);
  if (SEQ > 1 && SEQ * wgi >= nq) return;  // whole workgroup: before any barrier
  const int cbase = wgi * G;  // first column of this workgroup (nq: columns of the launch)
  const int L = a.L;
  const C* __restrict__ tw = static_cast<const C*>(a.twt);
  OutT* __restrict__ out = static_cast<OutT*>(a.out);
  const size_t nqs = a.ldo;
  const Acc inv_n = static_cast<Acc>(1) / static_cast<Acc>(N);
  Acc dsave[SEQ > 1 ? Lay::KMAX : 1];  // SEQ > 1: D of the first column per lag iteration

#pragma nounroll
  for (int sq = 0; sq < SEQ; ++sq) {
    // thread-derived addresses are rebuilt per column rather than kept live
    // across the loop (the thread index is opaque to the optimiser there)
    int tid = threadIdx.x;
    if constexpr (SEQ > 1) asm volatile("" : "+v"(tid));
    const int g = tid / TPT, u = tid - g * TPT, lane = tid & 63, wg = u >> 6;
    C* const buf0 = reinterpret_cast<C*>(smem);
    C* const edge0 = reinterpret_cast<C*>(smem + Lay::buf);
    double* const red = reinterpret_cast<double*>(smem + Lay::buf + Lay::edge);
    Acc* const chunk0 = reinterpret_cast<Acc*>(smem + Lay::buf + Lay::edge + Lay::red);
    C* const buf = buf0 + g * lds_padded_size(N);
    C* const edge = edge0 + g * 2 * kEdge;
    Acc* const chunk = chunk0 + g * Lay::CH;
    const int c = SEQ > 1 ? SEQ * wgi + sq : cbase + g;  // column of the launch
    const bool active = c < nq;
    const int q = active ? col_q(a, c) : a.s_q0;
    const C* __restrict__ src = static_cast<const C*>(a.S) + static_cast<size_t>(q - a.s_q0) * a.Lp;
    if constexpr (SUB && !F64) {
      // cleared here, set after the lag loop (several barriers later) when needed
      if (a.flags && active && u == 0) a.flags[q] = 0;
    }

    // 1. first-pass inputs straight from HBM (clamped indices, no branches;
    // samples t >= L are replaced by the zero padding afterwards)
    C v0[MB0][R0];
#pragma unroll
    for (int b = 0; b < MB0; ++b)
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const int t = u + b * TPT + r * M0;
        if (r < RH) {
          const C f = src[min(t, L - 1)];
          v0[b][r] = (active && t < L) ? f : C{0, 0};
        } else {
          v0[b][r] = C{0, 0};
        }
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
    const double rl = recip<double>(L);
    const double mr = sr * rl, mi = si * rl;
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

    // 4. (REG) prefix sums psum(t + 1) of |F(t) - mean|^2, from the first
    // pass's registers: they hold t = u + b TPT + r M0 = s TPT + u (segment
    // s = b + r MB0), so each segment is one wave scan per wave, then the
    // wave totals are offset in t order by one wave of the group, and the
    // sums stay in registers (ps[s]) until a buffer half is free for them.
    // The first and last kEdge samples go to `edge` for the direct sums.
    constexpr int S = Lay::S;
    static_assert(S == RH * MB0, "first-pass registers cover t < N/2");
    Acc ps[Lay::REG ? S : 1];
    if constexpr (Lay::REG) {
#pragma unroll
      for (int b = 0; b < MB0; ++b)
#pragma unroll
        for (int r = 0; r < RH; ++r) {
          const int sg = b + r * MB0, t = sg * TPT + u;
          const C f = v0[b][r];  // zero for t >= L and for an inactive column
          if constexpr (SUB) {
            if (t < kEdge && t < L) edge[t] = f;
            if (t >= L - kEdge && t < L) edge[t - L + 2 * kEdge] = f;
          }
          ps[sg] = wave::incl_scan(static_cast<Acc>(f.x) * static_cast<Acc>(f.x) +
                                   static_cast<Acc>(f.y) * static_cast<Acc>(f.y));
          if (lane == 63) chunk[sg * WPT + wg] = ps[sg];
        }
      __syncthreads();
      if (wg == 0) {
        // exclusive offsets of the (segment, wave) totals in t order, and their sum
        constexpr int E = S * WPT;
        Acc carry = 0;
#pragma unroll
        for (int e0 = 0; e0 < E; e0 += 64) {
          const int e = e0 + lane;
          const Acc x = e < E ? chunk[e] : Acc(0);
          const Acc inc = wave::incl_scan(x);
          const Acc exc = carry + wave::from_prev_lane(inc);
          carry += wave::lane63(inc);
          if (e < E) chunk[e] = exc;
        }
        if (lane == 0) chunk[E] = carry;
      }
      __syncthreads();
#pragma unroll
      for (int sg = 0; sg < S; ++sg) ps[sg] += chunk[sg * WPT + wg];
    }

    // 3. first transform G = FFT(F): first pass on the registers, last pass
    // back into registers; P = |G|^2 goes to LDS as z(n) = P(2n) + i P(2n+1)
    A0::compute_regs(v0);
    A0::store(buf, v0, u);
    __syncthreads();
    p2_passes<T, LOGN, false, 1, NP - 1>(buf, tw, u);
    {
      C v2[AL::MB][RL];
      AL::compute(buf, v2, tw + TWA, u);
      __syncthreads();
      T* const zf = reinterpret_cast<T*>(buf);
#pragma unroll
      for (int b = 0; b < AL::MB; ++b)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const int w = u + b * TPT + r * (N / RL);  // output position of the last pass
          zf[2 * padi(w >> 1) + (w & 1)] = v2[b][r].x * v2[b][r].x + v2[b][r].y * v2[b][r].y;
        }
    }
    __syncthreads();
    // 4-5. (not REG) prefix sums psum(t + 1) of |F(t) - mean|^2 from a second
    // (cache-resident) read of the series into the buffer half the second
    // transform leaves free, first and last kEdge samples into `edge`; and the
    // second transform at half length, Z = FFT_{N/2}(z).  PP (the paired
    // N = 4096 plan): the transform alternates between the halves (one barrier
    // per pass) and the sums follow it into the half it leaves; otherwise the
    // sums come first, so the re-read follows the first read closely, and the
    // transform runs in place in the first half.  Both orders were measured.
    constexpr bool PP = Lay::kPair;
    constexpr int NPH = p2_half_plan(LOGN, F64).np;
    constexpr bool ZA = !PP || NPH % 2 == 0;  // Z ends in the first half
    auto prefix_sums = [&](Acc* const tail) {
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
      const Acc inc = wave::incl_scan(run);
      Acc exc = wave::from_prev_lane(inc);  // 0 in lane 0
      if (lane == 63) chunk[wg] = inc;
      __syncthreads();
#pragma unroll
      for (int w = 0; w < WPT; ++w)
        if (w < wg) exc += chunk[w];
#pragma unroll
      for (int i = 0; i < SEGMAX; ++i) {
        const int t = ts + i;
        if (i < seg && t < L - 1) tail[t] = pre[i] + exc;
      }
      __syncthreads();
    };
    Acc* const tail = reinterpret_cast<Acc*>(ZA ? buf + padi(NH) : buf);
    if constexpr (Lay::REG) {
      // tail[t] = psum(t + 1), t < L - 1, into the second half, read after the
      // in-place transform's barriers
#pragma unroll
      for (int sg = 0; sg < S; ++sg) {
        const int t = sg * TPT + u;
        if (t < L - 1) tail[t] = ps[sg];
      }
      p2_passes<T, LOGN, true, 0, NPH>(buf, tw, u);
    } else if constexpr (PP) {
      p2_half_pp<T, LOGN, 0, NPH>(buf, buf + padi(NH), tw, u);
      prefix_sums(tail);
    } else {
      prefix_sums(tail);
      p2_passes<T, LOGN, true, 0, NPH>(buf, tw, u);
    }

    // 6. lags: thread tid serves q group h = tid % G and lag indices
    // nl - 1 - tid / G, stepping by NT / G, so neighbouring lanes store the same
    // lag of neighbouring q.  Lags in reverse order: the largest lags (direct
    // sums, longest loops) all fall in the first wave(s).
    {
      const int h = tid % G, v = tid / G, ch = SEQ > 1 ? c : cbase + h;
      if (SEQ > 1 || ch < nq) {
        const C* const hgrp = buf0 + h * lds_padded_size(N);
        const C* const hbuf = ZA ? hgrp : hgrp + padi(NH);  // Z
        const Acc* const htail = reinterpret_cast<const Acc*>(ZA ? hgrp + padi(NH) : hgrp);
        const C* const hedge = edge0 + h * 2 * kEdge;
        bool flag = false;  // some lag of this q needs the direct recomputation
        Acc hptot = 0;  // psum(L)
        if constexpr (Lay::REG) {
          hptot = chunk0[h * Lay::CH + S * WPT];
        } else {
#pragma unroll
          for (int w = 0; w < WPT; ++w) hptot += chunk0[h * Lay::CH + w];
        }
        // This thread's lag iterations k (index nl - 1 - v - k STEP, at most KL
        // of them since nl < N/2).  On the plans with registers to spare (the
        // paired and two-column ones, and FP64, whose LDS allows at most 4
        // waves per SIMD) their lag values and (cos, sin) entries are loaded
        // first, so those loads overlap instead of opening every iteration;
        // the other plans would lose occupancy and load per iteration.
        constexpr int STEP = Lay::NT / G;
        constexpr int KL = (NH + STEP - 1) / STEP;
        constexpr bool PRE = Lay::kPair || SEQ > 1 || F64;
        // (cos, sin)(2 pi d / N) of lag d: unsigned 32-bit indices, so each
        // entry address is one 64-bit shift-add from this base
        const C* __restrict__ const twp = tw + TWP;
        int dk[PRE ? KL : 1];
        C csk[PRE ? KL : 1];
        if constexpr (PRE) {
#pragma unroll
          for (int k = 0; k < KL; ++k) {
            const int i = a.nl - 1 - v - k * STEP;
            dk[k] = i >= 0 ? static_cast<int>(a.lags[static_cast<uint32_t>(i)]) : 1;
          }
#pragma unroll
          for (int k = 0; k < KL; ++k) csk[k] = twp[static_cast<uint32_t>(dk[k])];
        }
        auto lag_value = [&](int d, C cs) -> Acc {
          Acc D;
          if (SUB && L - d <= kEdge) {
            Acc acc = 0;
            for (int t = 0; t < L - d; ++t) {
              const int e = t + d;
              const C x = e < kEdge ? hedge[e] : hedge[e - L + 2 * kEdge];
              const C y = hedge[t];
              const Acc dr = static_cast<Acc>(x.x) - static_cast<Acc>(y.x);
              const Acc di = static_cast<Acc>(x.y) - static_cast<Acc>(y.y);
              acc += dr * dr + di * di;
            }
            D = acc * recip<Acc>(L - d);
          } else {
            const Acc head = htail[L - d - 1];   // psum(L - d)
            const Acc tl = hptot - htail[d - 1];  // psum(L) - psum(d)
            // N C(d) = Re FFT_N(P)(d) from A = Z(d), B = Z(N/2 - d) and
            // (c, s) = (cos, sin)(2 pi d / N): 2 Re X = A.x + B.x + c (A.y + B.y) - s (A.x - B.x)
            const C za = hbuf[padi(d)], zb = hbuf[padi((NH - d) & (NH - 1))];
            const Acc ax = za.x, ay = za.y, bx = zb.x, by = zb.y;
            const Acc re = (ax + bx) + static_cast<Acc>(cs.x) * (ay + by) - static_cast<Acc>(cs.y) * (ax - bx);
            const Acc c = re * (Acc(0.5) * inv_n);
            const Acc num = head + tl - 2 * c;  // (L - dt) D
            D = num * recip<Acc>(L - d);
            if constexpr (SUB && !F64) flag |= num * static_cast<Acc>(1.0 / kWkFlag) < hptot;
          }
          return D;
        };
        if constexpr (SEQ > 1) {
          // the second column stores both values (one access where the pair is
          // aligned: the first column is even, so that needs an even row
          // length and an aligned base)
          const size_t o0 = col_out(a, SEQ * wgi);
          const bool vec = ((nqs | o0) & 1) == 0 && reinterpret_cast<uintptr_t>(out) % sizeof(OutPair<OutT>) == 0;
          static_assert(Lay::KMAX == KL, "one lag iteration count");
#pragma unroll
          for (int k = 0; k < KL; ++k) {
            const int i = a.nl - 1 - v - k * STEP;
            if (i < 0) break;
            const Acc D = lag_value(dk[k], csk[k]);
            if (sq == 0) {
              dsave[k] = D;
            } else {
              OutT* const dst = out + static_cast<size_t>(i) * nqs + o0;
              if (active && vec) {
                *reinterpret_cast<OutPair<OutT>*>(dst) = OutPair<OutT>{static_cast<OutT>(dsave[k]), static_cast<OutT>(D)};
              } else {
                dst[0] = static_cast<OutT>(dsave[k]);
                if (active) dst[1] = static_cast<OutT>(D);
              }
            }
          }
        } else {
          const size_t oq = col_out(a, ch);
          if constexpr (PRE) {
#pragma unroll
            for (int k = 0; k < KL; ++k) {
              const int i = a.nl - 1 - v - k * STEP;
              if (i < 0) break;
              out[static_cast<size_t>(i) * nqs + oq] = static_cast<OutT>(lag_value(dk[k], csk[k]));
            }
          } else {
            for (int i = a.nl - 1 - v; i >= 0; i -= STEP) {
              const int d = static_cast<int>(a.lags[i]);
              out[static_cast<size_t>(i) * nqs + oq] = static_cast<OutT>(lag_value(d, twp[static_cast<uint32_t>(d)]));
            }
          }
        }
        if constexpr (SUB && !F64) {
          if (flag && a.flags && (SEQ == 1 || active)) a.flags[SEQ > 1 ? q : col_q(a, ch)] = 1;
        }
      }
    }
    if (active && u == 0) {
      Acc ptot = 0;
      if constexpr (Lay::REG) {
        ptot = chunk[S * WPT];
      } else {
#pragma unroll
        for (int w = 0; w < WPT; ++w) ptot += chunk[w];
      }
      const size_t oq = col_out(a, c);
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
    if (SEQ > 1 && sq + 1 < SEQ) __syncthreads();  // the next column reuses the LDS
  }
}

template <typename T, typename OutT, int LOGN, bool SUB>
hipError_t launch_one(const TimeArgs& a, hipStream_t s) {
  using Lay = P2Lds<LOGN, T, SUB>;
  const int nq = a.cols;
  // two-column plans keep at most KMAX lags per thread in registers (nl < L <= N/2 fits)
  if (Lay::SEQ > 1 && a.nl > Lay::KMAX * Lay::NT) return hipErrorInvalidValue;
  const int wgs = (nq + Lay::G * Lay::SEQ - 1) / (Lay::G * Lay::SEQ);
  const int wpx = (wgs + 7) / 8;
  const size_t lds = Lay::bytes;
  auto kernel = k_time_p2<T, OutT, LOGN, SUB>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kernel, dim3(8u * wpx), dim3(Lay::NT), lds, s, a, wpx, nq);
  return hipGetLastError();
}

template <typename T, typename OutT, bool SUB>
hipError_t launch_logn(const TimeArgs& a, int logn, hipStream_t s) {
  switch (logn) {
    case 8: return launch_one<T, OutT, 8, SUB>(a, s);
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
      case 8: return P2Lds<8, T, S>::bytes;
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
