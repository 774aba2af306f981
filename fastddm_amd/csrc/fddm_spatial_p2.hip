// Stage (a)+(b) for power-of-two image sizes (256 .. 4096): the same data
// flow as k_rows / k_cols (fddm_kernels.hip) with compile-time three-pass
// plans (p2::sp_plan), constant LDS offsets and constant twiddle strides.
//
//   k_rows_p2  Tb frames of one image row: frame pairs (2p, 2p+1) packed as
//              one complex row, convert (+ frame-0 subtraction) x window,
//              length-NX FFTs of all Tb/2 pairs at once, split into half spectra,
//              store R[tb][kx][y][ti] with the frame index fastest.
//   k_cols_p2  one kx column of Tb frames: length-NY FFTs of the interleaved
//              slab [y][ti]; the last pass writes the q-major spectra
//              S[ky * nxh + kx][t] straight from registers.  Slabs above
//              64 KiB leave room for two workgroups per CU, which then run
//              1024 threads each to keep 32 waves resident.
#include <algorithm>

#include "fddm_kernels.hpp"
#include "fddm_p2.hpp"
#include "fddm_spatial.hpp"

namespace fddm {
namespace {
using namespace p2;
using sp::pixel_value;
using sp::split_store;

// 512 threads per workgroup: the LDS slabs (up to 64 KiB) bound the number
// of resident workgroups, so each one brings 8 waves.
constexpr int kThreads = 512;

// Item of this workgroup.  When a frame block's run of TB values is shorter
// than a 128-byte line (8 FP32 frames: 64 bytes), neighbouring items write
// the two halves of the same lines; the XCD-aware order puts them on one XCD
// so the halves meet in its L2 instead of reaching memory as partial lines
// from two XCDs.  Workgroups are dealt to the 8 XCDs round-robin (block b
// runs on XCD b mod 8) and the grid is a multiple of 8, so item
// (b mod 8) (grid / 8) + b / 8 gives each XCD a contiguous run of items.
// Full-line runs keep the launch order (measured neutral to 1.5 % slower
// with the XCD order).
template <bool XCD>
__device__ __forceinline__ uint32_t block_item() {
  if constexpr (XCD) return (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  return blockIdx.x;
}
template <typename T, int TB>
constexpr bool kPartialLines = TB * 2 * sizeof(T) < 128;

// Launch grid for `items` workgroups: rounded up to a multiple of 8 (the
// blocks past `items` return at once).
inline unsigned xcd_grid(unsigned items) { return (items + 7u) & ~7u; }

template <typename Tin, typename T, int LOGX, int NP>
__global__ void __launch_bounds__(kThreads) k_rows_p2(RowArgs a, uint32_t nitems) {
  using C = cplx_t<T>;
  constexpr int NX = 1 << LOGX, NXH = NX / 2 + 1, TB = 2 * NP;
  constexpr SpPlan P = sp_plan(LOGX);
  extern __shared__ __align__(16) unsigned char smem[];
  C* const buf = reinterpret_cast<C*>(smem);
  const int tid = threadIdx.x;
  const int H = a.H, W = a.W;
  const uint32_t item = block_item<kPartialLines<T, TB>>();  // tb * H + y
  if (item >= nitems) return;
  const int y = static_cast<int>(item % H);
  const uint64_t tb = item / H;
  const uint64_t t0 = tb * TB;
  const Tin* __restrict__ img = static_cast<const Tin*>(a.img);
  const T* __restrict__ win = static_cast<const T*>(a.window);
  const size_t plane = static_cast<size_t>(H) * W;
  const bool sub = a.subtract0 != 0;
  const C* __restrict__ tw = static_cast<const C*>(a.twx);
  const Tin* __restrict__ row0 = img + static_cast<size_t>(y) * W;  // first frame of the launch, row y
  const Tin* __restrict__ ref0 = static_cast<const Tin*>(a.img0) + static_cast<size_t>(y) * W;  // subtracted frame

  // pairs (2p, 2p+1) -> LDS, element (p, x) at logical p NX + x
  if (W == NX && (static_cast<size_t>(W) * sizeof(Tin)) % 8 == 0) {
    // full rows: one 8-byte load brings VX consecutive pixels
    constexpr int VX = 8 / sizeof(Tin), G = NX / VX;
    struct alignas(8) Vec {
      Tin v[VX];
    };
    struct alignas(VX * sizeof(T) >= 16 ? 16 : VX * sizeof(T)) WVec {
      T v[VX];
    };
    for (int e = tid; e < NP * G; e += kThreads) {
      const int p = e / G, g = e - p * G, x0 = g * VX;
      const uint64_t ta = t0 + 2 * p;
      Vec va{}, vb{}, v0{};
      WVec wv;
      if (ta < a.frames) va = *reinterpret_cast<const Vec*>(row0 + ta * plane + x0);
      if (ta + 1 < a.frames) vb = *reinterpret_cast<const Vec*>(row0 + (ta + 1) * plane + x0);
      if (sub) v0 = *reinterpret_cast<const Vec*>(ref0 + x0);
      // the VX window values in one or two vector loads issued with the
      // pixels (one load per value: row pass 9.3 ms at 1024^2 x 4000, this
      // way 7.9 ms)
      if (win) wv = *reinterpret_cast<const WVec*>(win + static_cast<size_t>(y) * W + x0);
      C* d = buf + padi(p * NX + x0);  // VX divides 32: the VX slots are contiguous
#pragma unroll
      for (int v = 0; v < VX; ++v) {
        const T w = win ? wv.v[v] : T(1);
        const T fa = ta < a.frames ? pixel_value<Tin, T>(va.v[v], v0.v[v], sub, w) : T(0);
        const T fb = ta + 1 < a.frames ? pixel_value<Tin, T>(vb.v[v], v0.v[v], sub, w) : T(0);
        d[v] = C{fa, fb};
      }
    }
  } else {
    for (int e = tid; e < NP * NX; e += kThreads) {
      const int p = e >> LOGX, x = e & (NX - 1);
      const uint64_t ta = t0 + 2 * p;
      T fa = 0, fb = 0;
      if (x < W) {
        const T w = win ? win[static_cast<size_t>(y) * W + x] : T(1);
        const Tin x0v = sub ? ref0[x] : Tin(0);
        if (ta < a.frames) fa = pixel_value<Tin, T>(row0[ta * plane + x], x0v, sub, w);
        if (ta + 1 < a.frames) fb = pixel_value<Tin, T>(row0[(ta + 1) * plane + x], x0v, sub, w);
      }
      buf[padi(e)] = C{fa, fb};
    }
  }
  __syncthreads();
  fft3_lds<T, NX, P.r0, P.r1, P.r2, kThreads, NP, false, 1, NX>(buf, tw, tid);

  // A(k) = (Z(k) + conj Z(-k)) / 2, B(k) = (Z(k) - conj Z(-k)) / 2i; store
  // R[tb][k][y][ti] with ti fastest (runs of TB complex per k)
  const T half_norm = static_cast<T>(0.5 * a.norm);
  C* __restrict__ R = static_cast<C*>(a.R) + (tb * NXH * static_cast<size_t>(H) + y) * TB;
  const uint32_t kstride = static_cast<uint32_t>(H) * TB;
  for (int w = tid; w < NXH * NP; w += kThreads) {
    const int p = w & (NP - 1), k = w / NP;
    split_store<T>(R + static_cast<uint32_t>(k) * kstride + 2 * p, buf[padi(p * NX + k)],
                   buf[padi(p * NX + ((NX - k) & (NX - 1)))], half_norm);
  }
}

// Full-width rows (W == NX <= 512) of pixels of up to 4 bytes, FP32 and FP64
// (held to 64 and 128 VGPRs): a persistent form of k_rows_p2 with one wave
// per pair row.  Workgroup b takes items b,
// b + grid, ... (item = tb * H + y); wave p loads, converts and transforms
// pair row p in its own LDS region with no workgroup barrier (the 32 waves
// of the four resident workgroups per CU overlap one another's loads).  Two
// barriers per item remain: before the split into half spectra (which reads
// every pair row) and before the next item rewrites the rows.  The twiddles
// of passes 1 and 2 are copied into LDS once per workgroup.  Loads are not
// carried across items: holding the next item's pixels through the
// transform needed more than the 64 VGPRs of 8 waves per SIMD, and the
// spill reloads waited (vmcnt) for every load and store in flight
// (0.73 ms at 512^2 x 2000 with that prefetch, 0.68-0.69 ms without).
template <typename Tin, typename T, int LOGX, int NP>
__global__ void __launch_bounds__(64 * NP, (sizeof(T) == 8 ? 4 : 8)) k_rows_p2_persist(RowArgs a, uint32_t nitems) {
  using C = cplx_t<T>;
  constexpr int NT = 64 * NP;
  constexpr int NX = 1 << LOGX, NXH = NX / 2 + 1, TB = 2 * NP;
  constexpr SpPlan P = sp_plan(LOGX);
  constexpr int TWN = P.r0 * P.r1 * (1 + P.r2);  // [k][r] tables of passes 1 and 2
  constexpr int VX = 8 / sizeof(Tin), GV = NX / VX;
  constexpr int PER = GV / 64;  // 8-byte vectors per lane and frame row
  static_assert(GV % 64 == 0, "whole vectors per lane");
  struct alignas(8) Vec {
    Tin v[VX];
  };
  struct alignas(VX * sizeof(T)) WVec {
    T v[VX];
  };
  extern __shared__ __align__(16) unsigned char smem[];
  C* const buf = reinterpret_cast<C*>(smem);
  C* const twl = buf + lds_padded_size(NP * NX);
  const int tid = threadIdx.x, lane = tid & 63;
  const int p = __builtin_amdgcn_readfirstlane(tid >> 6);  // this wave's pair row (uniform)
  C* const row = buf + padi(p * NX);  // p NX is a multiple of 32: padded offsets add
  const int H = a.H;
  const Tin* __restrict__ img = static_cast<const Tin*>(a.img);
  const T* __restrict__ win = static_cast<const T*>(a.window);
  const size_t plane = static_cast<size_t>(H) * NX;
  const bool sub = a.subtract0 != 0;
  const T half_norm = static_cast<T>(0.5 * a.norm);
  const uint32_t kstride = static_cast<uint32_t>(H) * TB;
  {
    const C* __restrict__ tw = static_cast<const C*>(a.twx);
    for (int i = tid; i < TWN; i += NT) twl[i] = tw[i];
  }

  Vec va[PER], vb[PER], v0[PER];
  WVec wv[PER];
  auto load = [&](uint32_t item) {
    const int y = static_cast<int>(item % H);
    const uint64_t ta = static_cast<uint64_t>(item / H) * TB + 2 * p;
    const Tin* __restrict__ row0 = img + static_cast<size_t>(y) * NX;
    const Tin* __restrict__ ref0 = static_cast<const Tin*>(a.img0) + static_cast<size_t>(y) * NX;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int x0 = (lane + 64 * k) * VX;
      va[k] = ta < a.frames ? *reinterpret_cast<const Vec*>(row0 + ta * plane + x0) : Vec{};
      vb[k] = ta + 1 < a.frames ? *reinterpret_cast<const Vec*>(row0 + (ta + 1) * plane + x0) : Vec{};
      v0[k] = sub ? *reinterpret_cast<const Vec*>(ref0 + x0) : Vec{};
      if (win) wv[k] = *reinterpret_cast<const WVec*>(win + static_cast<size_t>(y) * NX + x0);
    }
  };

  uint32_t item = blockIdx.x;
  __syncthreads();  // twiddles in LDS
  for (; item < nitems; item += gridDim.x) {
    load(item);
    const int y = static_cast<int>(item % H);
    const uint64_t tb = item / H, ta = tb * TB + 2 * p;
    const bool has_a = ta < a.frames, has_b = ta + 1 < a.frames;  // uniform over the wave
    // registers -> this wave's row: frames (2p, 2p+1) as one complex row
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int x0 = (lane + 64 * k) * VX;
      C* d = row + padi(x0);  // VX divides 32: the VX slots are contiguous
#pragma unroll
      for (int v = 0; v < VX; ++v) {
        const T w = win ? wv[k].v[v] : T(1);
        const T fa = has_a ? pixel_value<Tin, T>(va[k].v[v], v0[k].v[v], sub, w) : T(0);
        const T fb = has_b ? pixel_value<Tin, T>(vb[k].v[v], v0[k].v[v], sub, w) : T(0);
        d[v] = C{fa, fb};
      }
    }
    wave_lds_sync();
    fft3_wave<T, NX, P.r0, P.r1, P.r2>(row, twl, lane);
    __syncthreads();  // every pair row transformed

    // split into the two half spectra and store R[tb][k][y][ti] (ti fastest)
    C* __restrict__ R = static_cast<C*>(a.R) + (tb * NXH * static_cast<size_t>(H) + y) * TB;
    for (int w = tid; w < NXH * NP; w += NT) {
      const int pp = w & (NP - 1), k = w / NP;
      split_store<T>(R + static_cast<uint32_t>(k) * kstride + 2 * pp, buf[padi(pp * NX + k)],
                     buf[padi(pp * NX + ((NX - k) & (NX - 1)))], half_norm);
    }
    __syncthreads();  // the LDS rows are rewritten by the next item
  }
}

template <typename T, int LOGY, int TB, int NT>
__global__ void __launch_bounds__(NT) k_cols_p2(ColArgs a, uint32_t nitems) {
  using C = cplx_t<T>;
  constexpr int NY = 1 << LOGY;
  constexpr SpPlan P = sp_plan(LOGY, true);
  using A0 = Pass<T, NY, P.r0, 1, NT, TB, true, TB, 1>;
  using A1 = Pass<T, NY, P.r1, P.r0, NT, TB, true, TB, 1>;
  using A2 = Pass<T, NY, P.r2, P.r0 * P.r1, NT, TB, true, TB, 1>;
  extern __shared__ __align__(16) unsigned char smem[];
  C* const buf = reinterpret_cast<C*>(smem);
  const int tid = threadIdx.x;
  const int nxh = a.nxh, H = a.H;
  // partial-line runs: item = kx * nblocks + tb, so the frame blocks of one
  // kx, whose S runs share lines, are neighbours on one XCD; otherwise
  // item = tb * nxh + kx
  constexpr bool XCD = kPartialLines<T, TB>;
  const uint32_t item = block_item<XCD>();
  if (item >= nitems) return;
  const uint32_t nb = static_cast<uint32_t>(a.nblocks);
  const int kx = static_cast<int>(XCD ? item / nb : item % static_cast<uint32_t>(nxh));
  const uint64_t tb = XCD ? item - static_cast<uint32_t>(kx) * nb : item / static_cast<uint32_t>(nxh);
  const C* __restrict__ src = static_cast<const C*>(a.R) + (tb * nxh + kx) * static_cast<size_t>(H) * TB;
  const int nin = H * TB;
  const C* __restrict__ tw = static_cast<const C*>(a.twy);
  if constexpr (LOGY <= 10) {
    // first pass straight from the slab: butterfly j of transform ti takes
    // y = j + r NY/R0, element y TB + ti (rows y >= H are the zero padding);
    // neighbouring lanes read neighbouring elements.  (Measured faster for
    // columns up to 1024, whose first pass is radix 8: FP64 1024^2 x 4000
    // 18.2 -> 15.2 ms; longer columns keep the staged copy below.)
    constexpr int M0 = NY / P.r0;
    C v[A0::MB][P.r0];
#pragma unroll
    for (int b = 0; b < A0::MB; ++b) {
      const int w = tid + b * NT;
      if (A0::ITEMS % NT != 0 && w >= A0::ITEMS) break;
      const int ti = w % TB, j = w / TB;
#pragma unroll
      for (int r = 0; r < P.r0; ++r) {
        const int e = (j + r * M0) * TB + ti;
        v[b][r] = e < nin ? src[e] : C{0, 0};
      }
    }
    A0::compute_regs(v);
    A0::store(buf, v, tid);
    __syncthreads();
  } else {
    for (int e = tid; e < NY * TB; e += NT) buf[padi(e)] = e < nin ? src[e] : C{0, 0};
    __syncthreads();
    C v[A0::MB][P.r0];
    A0::compute(buf, v, tw, tid);
    __syncthreads();
    A0::store(buf, v, tid);
    __syncthreads();
  }
  {
    C v[A1::MB][P.r1];
    A1::compute(buf, v, tw, tid);
    __syncthreads();
    A1::store(buf, v, tid);
    __syncthreads();
  }
  // last pass: outputs ky = j + r NY/R2 of frame ti go straight to S
  C v[A2::MB][P.r2];
  A2::compute(buf, v, tw + P.r0 * P.r1, tid);
  C* __restrict__ S = static_cast<C*>(a.S);
  constexpr int M2 = NY / P.r2;
#pragma unroll
  for (int b = 0; b < A2::MB; ++b) {
    const int w = tid + b * NT;
    if (A2::ITEMS % NT != 0 && w >= A2::ITEMS) break;
    const int ti = w % TB, j = w / TB;
#pragma unroll
    for (int r = 0; r < P.r2; ++r) {
      const int ky = j + r * M2;
      S[(static_cast<size_t>(ky) * nxh + kx) * a.Lp + tb * TB + ti] = v[b][r];
    }
  }
}

template <int LOGX, typename Tin, typename T, int NP>
hipError_t rows_np(const RowArgs& a, unsigned grid, hipStream_t s) {
  const size_t lds_rows = sizeof(cplx_t<T>) * lds_padded_size(NP * (1 << LOGX));
  const size_t lds = lds_rows;
  // persistent form (pixels of up to 4 bytes in whole 8-byte vectors per
  // lane, full rows up to 512 wide, FP32 and FP64: 512^2 x 2000 fp64 rows
  // 1.75 -> 1.49 ms, uint8 fp32c 0.78 -> 0.61 ms, float32 fp64 1.07 -> 0.77
  // ms at 1000 frames); as many workgroups as stay resident
  if constexpr (sizeof(Tin) <= 4 && LOGX <= 9 && ((1 << LOGX) * sizeof(Tin) / 8) % 64 == 0) {
    if (a.W == (1 << LOGX)) {
      constexpr SpPlan P = sp_plan(LOGX);
      const size_t lds = lds_rows + sizeof(cplx_t<T>) * P.r0 * P.r1 * (1 + P.r2);  // + the LDS twiddles
      auto kernel = k_rows_p2_persist<Tin, T, LOGX, NP>;
      if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
      }
      int dev = 0, cus = 0, per_cu = 0;
      hipError_t e = hipGetDevice(&dev);
      if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel), 64 * NP, lds);
      if (e != hipSuccess) return e;
      const unsigned resident = static_cast<unsigned>(std::max(1, cus) * std::max(1, per_cu));
      hipLaunchKernelGGL(kernel, dim3(std::min(grid, resident)), dim3(64 * NP), lds, s, a, static_cast<uint32_t>(grid));
      return hipGetLastError();
    }
  }
  auto kernel = k_rows_p2<Tin, T, LOGX, NP>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kernel, dim3(xcd_grid(grid)), dim3(kThreads), lds, s, a, static_cast<uint32_t>(grid));
  return hipGetLastError();
}

template <int LOGX, typename Tin, typename T>
hipError_t rows_one(const RowArgs& a, unsigned grid, hipStream_t s) {
  if (a.Tb == 16) return rows_np<LOGX, Tin, T, 8>(a, grid, s);
  if (a.Tb == 8) return rows_np<LOGX, Tin, T, 4>(a, grid, s);
  if (a.Tb == 4) return rows_np<LOGX, Tin, T, 2>(a, grid, s);
  return hipErrorInvalidValue;
}

template <typename Tin, typename T>
hipError_t rows_logx(const RowArgs& a, unsigned grid, hipStream_t s) {
  switch (a.logn) {
    case 8: return rows_one<8, Tin, T>(a, grid, s);
    case 9: return rows_one<9, Tin, T>(a, grid, s);
    case 10: return rows_one<10, Tin, T>(a, grid, s);
    case 11: return rows_one<11, Tin, T>(a, grid, s);
    case 12: return rows_one<12, Tin, T>(a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t rows_in(const RowArgs& a, unsigned grid, hipStream_t s) {
  switch (a.in_dtype) {
    case IN_U8: return rows_logx<uint8_t, T>(a, grid, s);
    case IN_I16: return rows_logx<int16_t, T>(a, grid, s);
    case IN_U16: return rows_logx<uint16_t, T>(a, grid, s);
    case IN_I32: return rows_logx<int32_t, T>(a, grid, s);
    case IN_U32: return rows_logx<uint32_t, T>(a, grid, s);
    case IN_I64: return rows_logx<int64_t, T>(a, grid, s);
    case IN_U64: return rows_logx<uint64_t, T>(a, grid, s);
    case IN_F32: return rows_logx<float, T>(a, grid, s);
    case IN_F64: return rows_logx<double, T>(a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T, int LOGY, int TB>
hipError_t cols_one(const ColArgs& a, unsigned grid, hipStream_t s) {
  constexpr size_t lds = sizeof(cplx_t<T>) * lds_padded_size((1 << LOGY) * TB);
  constexpr int NT = lds > 64 * 1024 ? 1024 : kThreads;
  auto kernel = k_cols_p2<T, LOGY, TB, NT>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kernel, dim3(xcd_grid(grid)), dim3(NT), lds, s, a, static_cast<uint32_t>(grid));
  return hipGetLastError();
}

template <typename T, int TB>
hipError_t cols_logy(const ColArgs& a, unsigned grid, hipStream_t s) {
  switch (a.logn) {
    case 8: return cols_one<T, 8, TB>(a, grid, s);
    case 9: return cols_one<T, 9, TB>(a, grid, s);
    case 10: return cols_one<T, 10, TB>(a, grid, s);
    case 11: return cols_one<T, 11, TB>(a, grid, s);
    case 12: return cols_one<T, 12, TB>(a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t cols_tb(const ColArgs& a, unsigned grid, hipStream_t s) {
  switch (a.Tb) {
    case 4: return cols_logy<T, 4>(a, grid, s);
    case 8: return cols_logy<T, 8>(a, grid, s);
    case 16: return cols_logy<T, 16>(a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

bool p2_spatial_ok(int logn) { return logn >= 8 && logn <= 12; }

// Twiddles of the spatial power-of-two plan: passes 1 and 2 of sp_plan,
// [k][r] = exp(-2 pi i r k / (NS R)).
int p2_spatial_plan(int logn, bool cols, int* r0, int* r1, int* r2) {
  if (!p2_spatial_ok(logn)) return 0;
  const SpPlan P = sp_plan(logn, cols);
  *r0 = P.r0;
  *r1 = P.r1;
  *r2 = P.r2;
  return 1;
}

hipError_t launch_rows_p2(const RowArgs& a, int precision, hipStream_t s) {
  const uint64_t nblocks = (a.frames + a.Tb - 1) / a.Tb;
  const unsigned grid = static_cast<unsigned>(nblocks * a.H);
  if (grid == 0) return hipSuccess;
  return precision == PREC_FP64 ? rows_in<double>(a, grid, s) : rows_in<float>(a, grid, s);
}

hipError_t launch_cols_p2(const ColArgs& a, int precision, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>(a.nblocks * a.nxh);
  if (grid == 0) return hipSuccess;
  return precision == PREC_FP64 ? cols_tb<double>(a, grid, s) : cols_tb<float>(a, grid, s);
}

}  // namespace fddm
