// Compile-time power-of-two FFT machinery shared by the time kernel and the
// power-of-two spatial kernels: constant-folded twiddle factors, radix
// 2..32 butterflies (16 and 32 as Cooley-Tukey composites), and a Stockham
// pass whose size, radix, stride and data layout are template parameters, so
// every LDS address is one per-butterfly base plus immediate offsets.
#pragma once

#include <type_traits>
#include <utility>

#include "fddm_fft.hpp"

namespace fddm {
namespace p2 {

// ---------------------------------------------------------------------------
// Compile-time cos / sin of 2 pi a / b: exact octant reduction of the
// rational angle, then Taylor series on [0, pi/4].
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double taylor_sin(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int i = 1; i < 14; ++i) {
    term *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int i = 1; i < 14; ++i) {
    term *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
struct CosSin {
  double c, s;
};
constexpr CosSin cossin2pi(long long a, long long b) {
  a %= b;
  if (a < 0) a += b;
  const long long o = (8 * a) / b;
  const long long rem = 8 * a - o * b;
  const double th = kPi / 4 * static_cast<double>(rem) / static_cast<double>(b);
  const double ph = kPi / 4 * static_cast<double>(b - rem) / static_cast<double>(b);
  const double ct = taylor_cos(th), st = taylor_sin(th), cp = taylor_cos(ph), sp = taylor_sin(ph);
  switch (o) {
    case 0: return {ct, st};
    case 1: return {sp, cp};
    case 2: return {-st, ct};
    case 3: return {-cp, sp};
    case 4: return {-ct, -st};
    case 5: return {-sp, -cp};
    case 6: return {st, -ct};
    default: return {cp, -sp};
  }
}

// a * exp(-2 pi i E / N), constant folded.
template <typename T, int E, int N>
__device__ __forceinline__ cplx_t<T> mul_w(cplx_t<T> a) {
  using C = cplx_t<T>;
  constexpr int e = ((E % N) + N) % N;
  constexpr T h = static_cast<T>(0.70710678118654752440084436210485);
  if constexpr (e == 0) {
    return a;
  } else if constexpr (4 * e == N) {
    return C{a.y, -a.x};
  } else if constexpr (2 * e == N) {
    return C{-a.x, -a.y};
  } else if constexpr (4 * e == 3 * N) {
    return C{-a.y, a.x};
  } else if constexpr (8 * e == N) {
    return C{(a.x + a.y) * h, (a.y - a.x) * h};
  } else if constexpr (8 * e == 3 * N) {
    return C{(a.y - a.x) * h, -(a.x + a.y) * h};
  } else if constexpr (8 * e == 5 * N) {
    return C{-(a.x + a.y) * h, (a.x - a.y) * h};
  } else if constexpr (8 * e == 7 * N) {
    return C{(a.x - a.y) * h, (a.x + a.y) * h};
  } else {
    constexpr CosSin w = cossin2pi(e, N);
    constexpr T c = static_cast<T>(w.c), s = static_cast<T>(w.s);
    return C{a.x * c + a.y * s, a.y * c - a.x * s};
  }
}

template <int I, int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    sfor<I + 1, N>(f);
  }
}

template <typename T, int R>
__device__ __forceinline__ void dft(cplx_t<T> (&v)[R]);

// Length R1 R2 DFT in place, natural order in and out (Cooley-Tukey):
// X[k1 + R1 k2] = sum_n2 W_R2^(n2 k2) W_N^(n2 k1) sum_n1 x[R2 n1 + n2] W_R1^(n1 k1).
template <typename T, int R1, int R2>
__device__ __forceinline__ void dft_composite(cplx_t<T> (&v)[R1 * R2]) {
  using C = cplx_t<T>;
  constexpr int N = R1 * R2;
  C a[R2][R1];
  sfor<0, R2>([&](auto n2c) {
    constexpr int n2 = decltype(n2c)::value;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) a[n2][n1] = v[R2 * n1 + n2];
    dft<T, R1>(a[n2]);
    sfor<1, R1>([&](auto k1c) {
      constexpr int k1 = decltype(k1c)::value;
      a[n2][k1] = mul_w<T, n2 * k1, N>(a[n2][k1]);
    });
  });
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    C b[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) b[n2] = a[n2][k1];
    dft<T, R2>(b);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
  }
}

template <typename T, int R>
__device__ __forceinline__ void dft(cplx_t<T> (&v)[R]) {
  if constexpr (R <= 8) {
    butterfly<T, R>(v);
  } else if constexpr (R == 16) {
    dft_composite<T, 4, 4>(v);
  } else {
    static_assert(R == 32, "radix");
    dft_composite<T, 8, 4>(v);
  }
}

__device__ __forceinline__ int padi(int e) { return e + (e >> 5); }
constexpr int cpad(int e) { return e + (e >> 5); }

// One Stockham pass with compile-time parameters: nb transforms of N points,
// radix R after NS points, worked on by NT threads (thread index `tid` in
// [0, NT)).  Element idx of transform ti sits at logical LDS offset
// ti * BS + idx * E (physical: padi of that).  TIF: consecutive items walk the
// transform index first (interleaved transforms, BS = 1, E = NB); otherwise
// the butterfly index first (BS a multiple of 32).  Butterfly j reads
// x[j + r M] and writes y[(j - k) R + k + r NS], k = j mod NS; with all sizes
// powers of two both are one padded base per butterfly plus constant offsets.
template <typename T, int N, int R, int NS, int NT, int NB = 1, bool TIF = false, int E = 1, int BS = 0>
struct Pass {
  using C = cplx_t<T>;
  static constexpr int M = N / R;
  static constexpr int ITEMS = M * NB;
  static constexpr int MB = (ITEMS + NT - 1) / NT;
  static_assert((M * E) % 32 == 0, "linear padded reads");
  static_assert(TIF ? (BS == 1 && E == NB && (32 % E == 0 || E % 32 == 0)) : (NB == 1 || BS % 32 == 0),
                "layout keeps padded writes linear");

  static __device__ __forceinline__ void item(int w, int& j, int& ti) {
    if constexpr (TIF) {
      ti = w % NB;
      j = w / NB;
    } else {
      j = w % M;
      ti = w / M;
    }
  }
  static __device__ __forceinline__ void compute(const C* buf, C (&v)[MB][R], const C* __restrict__ tw, int tid) {
    // opaque thread index: the address arithmetic of each pass stays in that
    // pass instead of being hoisted to the kernel entry and kept live
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      const int w = tid + b * NT;
      if (ITEMS % NT != 0 && w >= ITEMS) break;
      int j, ti;
      item(w, j, ti);
      const C* p = buf + padi(ti * BS + j * E);
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = p[cpad(r * M * E)];
      if constexpr (NS > 1) {
        const int k = j & (NS - 1);
        const C* t = tw + k * R;
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], t[r]);
      }
      dft<T, R>(v[b]);
    }
  }
  static __device__ __forceinline__ void compute_regs(C (&v)[MB][R]) {
#pragma unroll
    for (int b = 0; b < MB; ++b) dft<T, R>(v[b]);
  }
  static __device__ __forceinline__ void store(C* buf, const C (&v)[MB][R], int tid) {
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      const int w = tid + b * NT;
      if (ITEMS % NT != 0 && w >= ITEMS) break;
      int j, ti;
      item(w, j, ti);
      const int k = j & (NS - 1);
      C* p = buf + padi(ti * BS + ((j - k) * R + k) * E);
#pragma unroll
      for (int r = 0; r < R; ++r) p[cpad(r * NS * E)] = v[b][r];
    }
  }
};

// Three-pass plans for power-of-two spatial transforms (radices in order).
// Columns of 1024 start with radix 8 so their first pass, read straight from
// HBM, has a butterfly for every thread (measured faster there; rows keep
// radix 16 first).
struct SpPlan {
  int r0, r1, r2;
};
__host__ __device__ constexpr SpPlan sp_plan(int logn, bool cols = false) {
  return logn == 8    ? SpPlan{8, 8, 4}
         : logn == 9  ? SpPlan{8, 8, 8}
         : logn == 10 ? (cols ? SpPlan{8, 16, 8} : SpPlan{16, 8, 8})
         : logn == 11 ? SpPlan{16, 16, 8}
                      : SpPlan{16, 16, 16};
}

// Whole forward transform in LDS, three passes, block-wide synchronisation.
// tw: concatenated [k][r] tables of passes 1 and 2.
template <typename T, int N, int R0, int R1, int R2, int NT, int NB, bool TIF, int E, int BS>
__device__ __forceinline__ void fft3_lds(cplx_t<T>* buf, const cplx_t<T>* __restrict__ tw, int tid) {
  using C = cplx_t<T>;
  using A0 = Pass<T, N, R0, 1, NT, NB, TIF, E, BS>;
  using A1 = Pass<T, N, R1, R0, NT, NB, TIF, E, BS>;
  using A2 = Pass<T, N, R2, R0 * R1, NT, NB, TIF, E, BS>;
  {
    C v[A0::MB][R0];
    A0::compute(buf, v, tw, tid);
    __syncthreads();
    A0::store(buf, v, tid);
    __syncthreads();
  }
  {
    C v[A1::MB][R1];
    A1::compute(buf, v, tw, tid);
    __syncthreads();
    A1::store(buf, v, tid);
    __syncthreads();
  }
  {
    C v[A2::MB][R2];
    A2::compute(buf, v, tw + R0 * R1, tid);
    __syncthreads();
    A2::store(buf, v, tid);
    __syncthreads();
  }
}

// LDS writes of this wave complete and visible to its own later reads, and
// no memory access moved across this point by the compiler.  (Each pass of
// fft3_wave reads all its inputs into registers before any lane writes, so
// in-place passes need this only between passes.)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0), other counters left alone
  __builtin_amdgcn_wave_barrier();
}

// Whole forward transform of one length-N row in LDS by one wave (lane in
// [0, 64)), three passes, no workgroup barriers: the row's LDS region is
// this wave's own.
template <typename T, int N, int R0, int R1, int R2>
__device__ __forceinline__ void fft3_wave(cplx_t<T>* row, const cplx_t<T>* __restrict__ tw, int lane) {
  using C = cplx_t<T>;
  using A0 = Pass<T, N, R0, 1, 64>;
  using A1 = Pass<T, N, R1, R0, 64>;
  using A2 = Pass<T, N, R2, R0 * R1, 64>;
  {
    C v[A0::MB][R0];
    A0::compute(row, v, tw, lane);
    A0::store(row, v, lane);
    wave_lds_sync();
  }
  {
    C v[A1::MB][R1];
    A1::compute(row, v, tw, lane);
    A1::store(row, v, lane);
    wave_lds_sync();
  }
  {
    C v[A2::MB][R2];
    A2::compute(row, v, tw + R0 * R1, lane);
    A2::store(row, v, lane);
  }
}

}  // namespace p2
}  // namespace fddm
