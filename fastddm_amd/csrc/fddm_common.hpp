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

template <typename T> struct Cplx;
template <> struct Cplx<float> { using type = float2; };
template <> struct Cplx<double> { using type = double2; };
template <typename T> using cplx_t = typename Cplx<T>::type;

}  // namespace fddm
