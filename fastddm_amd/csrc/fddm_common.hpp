// Shared definitions of the fastddm_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fddm {

// Input pixel types, in the order of the reference's nine pybind11 overloads
// (src/python/bindings/ddm_bindings.h:34-43).
enum InDtype : int {
  IN_U8 = 0,
  IN_I16 = 1,
  IN_U16 = 2,
  IN_I32 = 3,
  IN_U32 = 4,
  IN_I64 = 5,
  IN_U64 = 6,
  IN_F32 = 7,
  IN_F64 = 8,
};

enum Precision : int {
  PREC_FP64 = 0,   // spectra, FFTs and sums in double
  PREC_FP32 = 1,   // spectra and FFTs in float, plain float sums
  PREC_FP32C = 2,  // float spectra/FFTs; frame-0 and temporal-mean subtraction,
                   // double prefix sums and combination
};

enum OutDtype : int { OUT_F64 = 0, OUT_F32 = 1 };
enum Mode : int { MODE_DIFF = 0, MODE_FFT = 1 };

// Mixed-radix factorisation of one FFT length, passed to kernels by value:
// n = 8^c8 4^c4 2^c2 3^c3 5^c5 7^c7, passes applied in that order.
struct FftDesc {
  int n;
  int c8, c4, c2, c3, c5, c7;
};

// Elements of an LDS buffer holding n logical elements with one pad slot
// every 32 (see lds_pad in fddm_fft.hpp).
__host__ __device__ constexpr int lds_padded_size(int n) { return n + (n >> 5) + 1; }

// Time-axis plans for N = 2^logn: threads per transform, number of passes
// (3 or 4) and the radices of the first transform in order; the first and
// last passes have a whole number of butterflies per thread.  For
// FP32 at N >= 4096 the transform is spread over 512 or 1024 threads (radix
// 8 or 16 passes) so that a CU holds 32 (N = 4096) or 16 waves despite the
// LDS buffer; FP64 keeps three passes on 256 threads, which measured faster
// than the four-pass plans at its register and LDS footprint.
struct P2Plan {
  int threads, np, r0, r1, r2, r3;
};
__host__ __device__ constexpr P2Plan p2_plan(int logn, bool f64) {
  return logn == 8    ? P2Plan{64, 4, 4, 4, 4, 4}
         : logn == 9  ? P2Plan{64, 3, 8, 8, 8, 1}
         : logn == 10 ? P2Plan{64, 3, 16, 16, 4, 1}
         : logn == 11 ? P2Plan{128, 3, 16, 16, 8, 1}
         : logn == 12 ? (f64 ? P2Plan{256, 3, 16, 16, 16, 1} : P2Plan{512, 4, 8, 8, 8, 8})
         : logn == 13 ? (f64 ? P2Plan{256, 3, 16, 16, 32, 1} : P2Plan{512, 4, 16, 8, 8, 8})
                      : P2Plan{1024, 4, 16, 8, 8, 16};
}
__host__ __device__ constexpr int p2_radix(const P2Plan& P, int i) {
  return i == 0 ? P.r0 : i == 1 ? P.r1 : i == 2 ? P.r2 : P.r3;
}
// Plan of the second transform: |G|^2 is real, so its length-N transform
// runs as one length-N/2 complex transform of z(n) = P(2n) + i P(2n+1) on the
// same threads (passes with fewer butterflies than threads leave whole waves
// idle), followed by a per-lag untangling step.
__host__ __device__ constexpr P2Plan p2_half_plan(int logn, bool f64) {
  return logn == 8    ? P2Plan{64, 4, 4, 4, 4, 2}
         : logn == 9  ? P2Plan{64, 3, 8, 8, 4, 1}
         : logn == 10 ? P2Plan{64, 3, 8, 8, 8, 1}
         : logn == 11 ? P2Plan{128, 3, 16, 16, 4, 1}
         : logn == 12 ? (f64 ? P2Plan{256, 3, 16, 16, 8, 1} : P2Plan{512, 4, 8, 8, 8, 4})
         : logn == 13 ? (f64 ? P2Plan{256, 3, 16, 16, 16, 1} : P2Plan{512, 4, 8, 8, 8, 8})
                      : P2Plan{1024, 4, 16, 8, 8, 8};
}
// Points already combined before pass i (product of the earlier radices).
__host__ __device__ constexpr int p2_ns(const P2Plan& P, int i) { return i == 0 ? 1 : p2_ns(P, i - 1) * p2_radix(P, i - 1); }
// Twiddle table layout (complex entries): FFT1 passes 1 .. np-1, then the
// half-length passes 1 .. np-1 of the second transform, then N/2 entries
// (cos, sin)(2 pi k / N) for the untangling step.  The pass with radix R
// after NS points holds exp(-2 pi i r k / (NS R)) at offset k R + r of its
// table (NS R entries).
__host__ __device__ constexpr int p2_tw_pass(const P2Plan& P, int i, int base) {
  return i <= 1 ? base : p2_tw_pass(P, i - 1, base) + p2_ns(P, i - 1) * p2_radix(P, i - 1);
}
__host__ __device__ constexpr int p2_tw_a(int logn, bool f64, int i) { return p2_tw_pass(p2_plan(logn, f64), i, 0); }
__host__ __device__ constexpr int p2_tw_h(int logn, bool f64, int i) {
  return p2_tw_pass(p2_half_plan(logn, f64), i, p2_tw_a(logn, f64, p2_plan(logn, f64).np));
}
__host__ __device__ constexpr int p2_tw_post(int logn, bool f64) {
  return p2_tw_h(logn, f64, p2_half_plan(logn, f64).np);
}
__host__ __device__ constexpr int p2_tw_total(int logn, bool f64) { return p2_tw_post(logn, f64) + (1 << (logn - 1)); }
constexpr int kP2MinLog = 8, kP2MaxLog = 14;

template <typename T> struct Cplx;
template <> struct Cplx<float> { using type = float2; };
template <> struct Cplx<double> { using type = double2; };
template <typename T> using cplx_t = typename Cplx<T>::type;

}  // namespace fddm
