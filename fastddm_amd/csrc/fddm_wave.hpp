// Cross-lane helpers for 64-lane waves built on DPP moves (no LDS traffic).
//
// DPP controls in the GFX9 encoding: row_shr:n = 0x110 + n moves lane l - n
// of the same 16-lane row into lane l; row_bcast:15 (0x142) gives every lane
// of rows 1..3 lane 15 of the row before; row_bcast:31 (0x143) gives rows 2
// and 3 lane 31; wave_shr:1 (0x138) moves lane l - 1 into lane l across the
// whole wave.  dpp_zero fills lanes with no source, and rows left out by the
// row mask, with 0.
#pragma once

#include <hip/hip_runtime.h>

namespace fddm {
namespace wave {

constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_zero(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xf, true));
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_zero(double x) {
  const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(x));
  const unsigned lo = static_cast<unsigned>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(u & 0xffffffffull), CTRL, ROW_MASK, 0xf, true));
  const unsigned hi =
      static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(u >> 32), CTRL, ROW_MASK, 0xf, true));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Inclusive scan over the wave: shifts of 1, 2, 4, 8 inside each 16-lane row,
// then the row totals carried across rows by the two broadcasts.  Lane 63
// ends with the sum of all 64 lanes.
template <typename A>
__device__ __forceinline__ A incl_scan(A v) {
  v += dpp_zero<kRowShr1, 0xf>(v);
  v += dpp_zero<kRowShr2, 0xf>(v);
  v += dpp_zero<kRowShr4, 0xf>(v);
  v += dpp_zero<kRowShr8, 0xf>(v);
  v += dpp_zero<kRowBcast15, 0xa>(v);
  v += dpp_zero<kRowBcast31, 0xc>(v);
  return v;
}

// Lane 63's value, in every lane.
__device__ __forceinline__ float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }
__device__ __forceinline__ double lane63(double v) {
  const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u & 0xffffffffull), 63));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), 63));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// The previous lane's value (0 in lane 0).
template <typename A>
__device__ __forceinline__ A from_prev_lane(A v) {
  return dpp_zero<kWaveShr1, 0xf>(v);
}

}  // namespace wave
}  // namespace fddm
