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

// Power-of-two plans of the time kernel: N = 2^logn points, T threads per
// transform, three passes of radices R0, R1, R2 (forward order; the second
// transform runs them in reverse).  Every pass has a whole number of
// butterflies per thread: N / (R * T) >= 1.
struct P2Plan {
  int threads, r0, r1, r2;
};
__host__ __device__ constexpr P2Plan p2_plan(int logn) {
  return logn == 9    ? P2Plan{64, 8, 8, 8}
         : logn == 10 ? P2Plan{64, 16, 16, 4}
         : logn == 11 ? P2Plan{128, 16, 16, 8}
         : logn == 12 ? P2Plan{256, 16, 16, 16}
         : logn == 13 ? P2Plan{256, 16, 16, 32}
                      : P2Plan{256, 16, 32, 32};
}
constexpr int kP2MinLog = 9, kP2MaxLog = 14;

template <typename T> struct Cplx;
template <> struct Cplx<float> { using type = float2; };
template <> struct Cplx<double> { using type = double2; };
template <typename T> using cplx_t = typename Cplx<T>::type;

}  // namespace fddm
