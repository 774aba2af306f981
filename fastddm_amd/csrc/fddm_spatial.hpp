// Helpers shared by the spatial (row-pass) kernels of every size class.
#pragma once

#include <type_traits>

#include "fddm_common.hpp"

namespace fddm {
namespace sp {

// Pixel (p, x) of a pair row: converted, frame 0 subtracted in integer
// arithmetic where the type allows, times the window value.
template <typename Tin, typename T>
__device__ __forceinline__ T pixel_value(Tin v, Tin v0, bool sub, T w) {
  if constexpr (std::is_integral_v<Tin> && sizeof(Tin) <= 2) {
    return static_cast<T>(static_cast<int>(v) - (sub ? static_cast<int>(v0) : 0)) * w;
  } else if constexpr (std::is_integral_v<Tin> && sizeof(Tin) == 4) {
    return static_cast<T>(static_cast<long long>(v) - (sub ? static_cast<long long>(v0) : 0)) * w;
  } else if constexpr (std::is_integral_v<Tin>) {
    return static_cast<T>(static_cast<double>(v) - (sub ? static_cast<double>(v0) : 0.0)) * w;
  } else {
    // floating pixels: the difference in the wider of the two types (a
    // float32 difference would round where the FP64 transform is exact)
    using Wd = std::conditional_t<(sizeof(Tin) > sizeof(T)), Tin, T>;
    return static_cast<T>(sub ? static_cast<Wd>(v) - static_cast<Wd>(v0) : static_cast<Wd>(v)) * w;
  }
}

// Split of pair row p at frequency k, Z = FFT(frame 2p + i frame 2p+1):
// A(k) = (Z(k) + conj Z(-k)) / 2 and B(k) = (Z(k) - conj Z(-k)) / 2i (times
// the norm), stored together at dst (frames 2p and 2p+1 are neighbours in R):
// one 16-byte store in FP32, and one read of the two Z values per pair.
template <typename T>
__device__ __forceinline__ void split_store(cplx_t<T>* dst, cplx_t<T> zk, cplx_t<T> zn, T half_norm) {
  const T ax = (zk.x + zn.x) * half_norm, ay = (zk.y - zn.y) * half_norm;
  const T bx = (zk.y + zn.y) * half_norm, by = (zn.x - zk.x) * half_norm;
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4*>(dst) = make_float4(ax, ay, bx, by);
  } else {
    dst[0] = cplx_t<T>{ax, ay};
    dst[1] = cplx_t<T>{bx, by};
  }
}

}  // namespace sp
}  // namespace fddm
