// Stage (a) row and column passes for transform lengths that are not powers
// of two (7-smooth, up to 2048: 96, 150, 200, 384, 480, 640, 1280, 1920, ...):
// the data flow of k_rows / k_cols (fddm_kernels.hip) reorganised the way the
// power-of-two persistent row kernel is:
//
//   * one wave per transform (rows: frames 2p and 2p+1 of one image row as
//     one complex row; columns: one frame of the block), held in the wave's
//     own LDS row and transformed with no workgroup barrier (runtime radix
//     sequence 8s, 4s, 2s, 3s, 5s, 7s; each pass reads all of a lane's
//     butterflies into registers before it writes, and the wave's LDS
//     operations complete in order, so passes run in place);
//   * the pass twiddles are copied into LDS once per workgroup, which then
//     loops over items ((frame block, image row) or (kx, frame block));
//   * pixels and window values come in 8-byte / 16-byte vector loads, all of
//     a lane's loads of a round issued before the first conversion;
//   * 16-byte stores: the row split writes both frames of a pair at once
//     from one read of Z(k) and Z(-k), the columns two frames of a q run.
//
// Items are dealt so that each XCD works through a contiguous range: the
// runs of neighbouring items share 128-byte lines when a frame block holds
// fewer than 16 FP32 (8 FP64) frames, and those halves then meet in one L2.
#include <algorithm>

#include "fddm_fft.hpp"
#include "fddm_kernels.hpp"
#include "fddm_p2.hpp"
#include "fddm_spatial.hpp"

namespace fddm {
namespace {

// Longest transforms of the two kernel forms: up to kMrShort points the
// FP32 kernels fit 64 VGPRs (8 waves per SIMD, the occupancy their LDS
// allows); longer ones, up to kMrLong, hold fewer workgroups per CU anyway
// (their LDS rows) and get the registers for it.
constexpr int kMrShort = 768, kMrLong = 2048;
template <int KMAX, typename T>
constexpr int mr_min_waves() {
  return KMAX <= kMrShort ? (sizeof(T) == 4 ? 8 : 4) : (sizeof(T) == 4 ? 3 : 2);
}

// LDS slot of element e of a transform row.  FP32 rows are not padded:
// those kernels spend most of their time issuing VALU instructions, and the
// pad slot every 32 elements of lds_pad (fddm_fft.hpp) costs three more
// instructions per access than the bank conflicts of the early passes'
// strided writes cost (row pass
// 7.72 -> 7.21 ms, columns 8.76 -> 8.63 ms over the 48-stack batch's
// shapes).  FP64 rows (16-byte elements) keep the pad (480x640 x 1000: rows
// 1.15 ms padded, 1.18 ms not; columns 1.40 and 1.48 ms).
template <typename T>
__device__ __forceinline__ int mr_pad(int e) {
  return sizeof(T) == 8 ? e + (e >> 5) : e;
}

// One radix-R Stockham pass (formulation of fddm_fft.hpp) of a length-n
// transform held by one wave: element e at row[mr_pad<T>(e)], lane handles
// butterflies j = lane + 64 b.  ns: product of the radices already applied;
// tw: this pass's table, tw[k (R-1) + r - 1] = exp(-2 pi i r k / (ns R)).
template <typename T, int R, int KMAX>
__device__ __forceinline__ void wave_pass(cplx_t<T>* row, int n, int ns, const cplx_t<T>* tw, int lane) {
  using C = cplx_t<T>;
  constexpr int MB = (KMAX / R + 63) / 64;
  // opaque lane index: the lane-derived address arithmetic stays in this
  // pass instead of being hoisted to the kernel entry and kept live
  asm volatile("" : "+v"(lane));
  const int m = n / R;
  // k = j mod ns, advanced by 64 mod ns per item; ns is uniform, and a power
  // of two for every pass before the first odd radix (a mask, not a division)
  const bool p2ns = (ns & (ns - 1)) == 0;
  int k = p2ns ? (lane & (ns - 1)) : lane % ns;
  const int dk = p2ns ? (64 & (ns - 1)) : 64 % ns;
  C v[MB][R];
  int ob[MB];
#pragma unroll
  for (int b = 0; b < MB; ++b) ob[b] = -1;
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    const int j = lane + 64 * b;
    if (64 * b < m && j < m) {  // (the first test is uniform)
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = row[mr_pad<T>(j + r * m)];
      if (ns > 1) {
        const C* t = tw + k * (R - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], t[r - 1]);
      }
      butterfly<T, R>(v[b]);
      ob[b] = j * R - k * (R - 1);
    }
    k += dk;
    if (k >= ns) k -= ns;
  }
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    if (ob[b] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) row[mr_pad<T>(ob[b] + r * ns)] = v[b][r];
    }
  }
  p2::wave_lds_sync();
}

// Forward FFT (unnormalised) of one length-d.n row by one wave; passes in
// the order of the twiddle table (8s, 4s, 2s, 3s, 5s, 7s).
template <typename T, int KMAX>
__device__ __forceinline__ void fft_wave_mr(cplx_t<T>* row, const FftDesc& d, const cplx_t<T>* tw, int lane) {
  const int n = d.n;
  int ns = 1;
#define FDDM_WAVE_PASSES(RAD, COUNT)                  \
  for (int i = 0; i < d.COUNT; ++i) {                 \
    wave_pass<T, RAD, KMAX>(row, n, ns, tw, lane);    \
    tw += ns * (RAD - 1);                             \
    ns *= RAD;                                        \
  }
  FDDM_WAVE_PASSES(8, c8)
  FDDM_WAVE_PASSES(4, c4)
  FDDM_WAVE_PASSES(2, c2)
  FDDM_WAVE_PASSES(3, c3)
  FDDM_WAVE_PASSES(5, c5)
  FDDM_WAVE_PASSES(7, c7)
#undef FDDM_WAVE_PASSES
}

// Row pass, Tb = 2 NP frames per block, NP waves.  rs: LDS slots per pair
// row (lds_padded_size(nx)); vec: pixel rows start 8-byte aligned and the
// loaded width is a multiple of the vector (host-checked).
template <typename Tin, typename T, int NP, int KMAX>
__global__ void __launch_bounds__(64 * NP, (mr_min_waves<KMAX, T>())) k_rows_mr(RowArgs a, uint32_t nitems, int rs, int vec) {
  using C = cplx_t<T>;
  constexpr int NT = 64 * NP, TB = 2 * NP;
  constexpr int VX = sizeof(Tin) <= 4 ? 8 / static_cast<int>(sizeof(Tin)) : 1;
  // vectors per lane and frame row, at most (even: taken two at a time)
  constexpr int KV = ((KMAX / VX + 63) / 64 + 1) & ~1;
  struct alignas(8) Vec {
    Tin v[VX];
  };
  struct alignas(VX * sizeof(T) >= 16 ? 16 : VX * sizeof(T)) WVec {
    T v[VX];
  };
  extern __shared__ __align__(16) unsigned char smem[];
  C* const buf = reinterpret_cast<C*>(smem);
  C* const twl = buf + NP * rs;
  const int tid = threadIdx.x;
  const int p = __builtin_amdgcn_readfirstlane(tid >> 6);  // this wave's pair row (uniform)
  C* const row = buf + p * rs;
  const int H = a.H, W = a.W, nx = a.nx, nxh = nx / 2 + 1;
  const int wl = W < nx ? W : nx;  // pixels loaded per row; x in [wl, nx) is zero padding
  const Tin* __restrict__ img = static_cast<const Tin*>(a.img);
  const Tin* __restrict__ img0 = static_cast<const Tin*>(a.img0);
  const T* __restrict__ win = static_cast<const T*>(a.window);
  const size_t plane = static_cast<size_t>(H) * W;
  const bool sub = a.subtract0 != 0;
  const T half_norm = static_cast<T>(0.5 * a.norm);
  const uint32_t kstride = static_cast<uint32_t>(H) * TB;
  {
    const C* __restrict__ tw = static_cast<const C*>(a.twx);
    const int twn = nx > 1 ? nx - 1 : 1;
    for (int i = tid; i < twn; i += NT) twl[i] = tw[i];
  }
  // items of this workgroup's XCD: [x per, (x + 1) per), stepped by the
  // XCD's workgroup count
  const uint64_t per = (static_cast<uint64_t>(nitems) + 7u) >> 3, x = blockIdx.x & 7u;
  const uint64_t end = (x + 1u) * per < nitems ? (x + 1u) * per : nitems;
  __syncthreads();  // twiddles in LDS
  for (uint64_t item = x * per + (blockIdx.x >> 3); item < end; item += gridDim.x >> 3) {
    // thread-derived addresses are rebuilt per item rather than hoisted out
    // of the loop and kept live through the transform
    int lane = tid & 63, tidi = tid;
    asm volatile("" : "+v"(lane), "+v"(tidi));
    const int y = static_cast<int>(static_cast<uint32_t>(item) % static_cast<uint32_t>(H));
    const uint64_t tb = static_cast<uint32_t>(item) / static_cast<uint32_t>(H), ta = tb * TB + 2 * p;
    const bool has_a = ta < a.frames, has_b = ta + 1 < a.frames;  // uniform over the wave
    const size_t yo = static_cast<size_t>(y) * W;
    const Tin* __restrict__ ra = img + ta * plane + yo;
    const Tin* __restrict__ r0 = img0 + yo;
    auto scalar_rows = [&]() {
#pragma unroll 4
      for (int x0 = lane; x0 < wl; x0 += 64) {
        const T w = win ? win[yo + x0] : T(1);
        const Tin p0 = sub ? r0[x0] : Tin(0);
        const T fa = has_a ? sp::pixel_value<Tin, T>(ra[x0], p0, sub, w) : T(0);
        const T fb = has_b ? sp::pixel_value<Tin, T>(ra[plane + x0], p0, sub, w) : T(0);
        row[mr_pad<T>(x0)] = C{fa, fb};
      }
    };
    if constexpr (sizeof(Tin) <= 4) {
      if (vec) {
        // two vectors per lane at a time: their loads (clamped addresses) are
        // all issued before the first conversion
#pragma unroll
        for (int k0 = 0; k0 < KV; k0 += 2) {
          if (64 * k0 * VX < wl) {  // uniform
            Vec va[2], vb[2], v0[2];
            WVec wv[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const int x0 = min((lane + 64 * (k0 + h)) * VX, wl - VX);
              va[h] = has_a ? *reinterpret_cast<const Vec*>(ra + x0) : Vec{};
              vb[h] = has_b ? *reinterpret_cast<const Vec*>(ra + plane + x0) : Vec{};
              v0[h] = sub ? *reinterpret_cast<const Vec*>(r0 + x0) : Vec{};
              if (win) wv[h] = *reinterpret_cast<const WVec*>(win + yo + x0);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const int x0 = (lane + 64 * (k0 + h)) * VX;
              if (x0 < wl) {
#pragma unroll
                for (int v = 0; v < VX; ++v) {
                  const T w = win ? wv[h].v[v] : T(1);
                  const T fa = has_a ? sp::pixel_value<Tin, T>(va[h].v[v], v0[h].v[v], sub, w) : T(0);
                  const T fb = has_b ? sp::pixel_value<Tin, T>(vb[h].v[v], v0[h].v[v], sub, w) : T(0);
                  row[mr_pad<T>(x0 + v)] = C{fa, fb};
                }
              }
            }
          }
        }
      } else {
        scalar_rows();
      }
    } else {
      scalar_rows();
    }
    for (int xp = wl + lane; xp < nx; xp += 64) row[mr_pad<T>(xp)] = C{0, 0};
    p2::wave_lds_sync();
    fft_wave_mr<T, KMAX>(row, a.dx, twl, lane);
    __syncthreads();  // every pair row transformed

    // split into the two half spectra, store R[tb][k][y][ti] (ti fastest)
    C* __restrict__ R = static_cast<C*>(a.R) + (tb * nxh * static_cast<size_t>(H) + y) * TB;
    // thread tidi takes pair row pp = tidi mod NP and k = tidi / NP + 64 i
    // (NT = 64 NP threads): the row is fixed and the address advances
    {
      const int pp = tidi & (NP - 1);
      const C* const rp = buf + pp * rs;
      C* dst = R + static_cast<uint32_t>(tidi / NP) * kstride + 2 * pp;
      for (int k = tidi / NP; k < nxh; k += 64, dst += 64 * kstride)
        sp::split_store<T>(dst, rp[mr_pad<T>(k)], rp[mr_pad<T>(k == 0 ? 0 : nx - k)], half_norm);
    }
    __syncthreads();  // the LDS rows are rewritten by the next item
  }
}

// Column pass + q-major store, Tb = TB frames per block, TB waves: one
// (kx, frame block) item is the slab R[tb][kx][y][ti]; it is loaded with the
// frame index turned outermost (transform ti is LDS row ti), wave ti
// transforms row ti with no workgroup barrier, and the store writes
// S[ky nxh + kx][tb TB + ti] in runs of TB frames (two frames per 16-byte
// store in FP32).  Items run kx-major (item = kx nblocks + tb), so the frame
// blocks whose runs share lines are neighbours in one XCD's range.
template <typename T, int TB, int KMAX>
__global__ void __launch_bounds__(64 * TB, (mr_min_waves<KMAX, T>())) k_cols_mr(ColArgs a, uint32_t nitems, int rs) {
  using C = cplx_t<T>;
  constexpr int NT = 64 * TB;
  constexpr int KC = 4;                                // slab elements per thread and load round
  constexpr int KL = (KMAX * TB + NT * KC - 1) / (NT * KC);  // load rounds, at most
  extern __shared__ __align__(16) unsigned char smem[];
  C* const buf = reinterpret_cast<C*>(smem);
  C* const twl = buf + TB * rs;
  const int tid = threadIdx.x;
  const int tw_i = __builtin_amdgcn_readfirstlane(tid >> 6);  // this wave's transform (uniform)
  C* const col = buf + tw_i * rs;
  const int H = a.H, ny = a.ny, nxh = a.nxh;
  const int hl = H < ny ? H : ny;  // rows loaded; y in [hl, ny) is zero padding
  const int nin = hl * TB;
  const uint32_t nb = static_cast<uint32_t>(a.nblocks);
  {
    const C* __restrict__ tw = static_cast<const C*>(a.twy);
    const int twn = ny > 1 ? ny - 1 : 1;
    for (int i = tid; i < twn; i += NT) twl[i] = tw[i];
  }
  const uint64_t per = (static_cast<uint64_t>(nitems) + 7u) >> 3, x = blockIdx.x & 7u;
  const uint64_t end = (x + 1u) * per < nitems ? (x + 1u) * per : nitems;
  for (uint64_t item = x * per + (blockIdx.x >> 3); item < end; item += gridDim.x >> 3) {
    int tidi = tid;
    asm volatile("" : "+v"(tidi));
    const int lane = tidi & 63;
    const uint32_t kx = static_cast<uint32_t>(item) / nb, tb = static_cast<uint32_t>(item) - kx * nb;
    const C* __restrict__ src = static_cast<const C*>(a.R) + (static_cast<size_t>(tb) * nxh + kx) * H * TB;
    __syncthreads();  // the previous item's rows are read; twiddles in LDS
    // slab -> LDS rows: element e = y TB + ti goes to row ti, slot y; KC loads
    // per thread (clamped addresses) issued before their LDS writes
#pragma unroll
    for (int l = 0; l < KL; ++l) {
      if (l * NT * KC < nin) {  // uniform
        C v[KC];
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int e = tidi + (l * KC + c) * NT;
          v[c] = src[e < nin ? e : nin - 1];
        }
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int e = tidi + (l * KC + c) * NT;
          if (e < nin) buf[(e & (TB - 1)) * rs + mr_pad<T>(e / TB)] = v[c];
        }
      }
    }
    for (int y = hl + lane; y < ny; y += 64) col[mr_pad<T>(y)] = C{0, 0};
    __syncthreads();
    fft_wave_mr<T, KMAX>(col, a.dy, twl, lane);
    __syncthreads();  // every transform done
    C* __restrict__ S = static_cast<C*>(a.S) + static_cast<size_t>(kx) * a.Lp + static_cast<size_t>(tb) * TB;
    // thread tidi stores frames 2 tp, 2 tp + 1 (tp = tidi mod TB/2) of the q
    // rows ky = tidi / (TB/2) + 128 i (NT = 128 TB/2 threads)
    constexpr int TH = TB / 2, KS = NT / TH;
    {
      const int tp = tidi & (TH - 1);
      const C* const r0 = buf + (2 * tp) * rs;
      const C* const r1 = r0 + rs;
      const size_t qrow = static_cast<size_t>(nxh) * a.Lp;
      C* dst = S + static_cast<size_t>(tidi / TH) * qrow + 2 * tp;
      for (int ky = tidi / TH; ky < ny; ky += KS, dst += KS * qrow) {
        const C z0 = r0[mr_pad<T>(ky)], z1 = r1[mr_pad<T>(ky)];
        if constexpr (sizeof(T) == 4) {
          *reinterpret_cast<float4*>(dst) = make_float4(z0.x, z0.y, z1.x, z1.y);
        } else {
          dst[0] = z0;
          dst[1] = z1;
        }
      }
    }
  }
}

size_t rows_mr_lds(const RowArgs& a, int precision, int* rs) {
  *rs = lds_padded_size(a.nx);
  const size_t csz = precision == PREC_FP64 ? 16 : 8;
  return csz * (static_cast<size_t>(a.Tb / 2) * *rs + std::max(a.nx - 1, 1));
}

template <typename Tin, typename T, int NP>
hipError_t rows_mr_np(const RowArgs& a, uint32_t items, size_t lds, int rs, int vec, hipStream_t s) {
  auto kernel = a.nx <= kMrShort ? k_rows_mr<Tin, T, NP, kMrShort> : k_rows_mr<Tin, T, NP, kMrLong>;
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
  // as many workgroups as stay resident, a multiple of 8 (one share per XCD),
  // and no more than the items need
  const uint32_t resident = static_cast<uint32_t>(std::max(1, cus) * std::max(1, per_cu)) & ~7u;
  const uint32_t grid = std::max(8u, std::min(resident, (items + 7u) & ~7u));
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(64 * NP), lds, s, a, items, rs, vec);
  return hipGetLastError();
}

template <typename Tin, typename T>
hipError_t rows_mr_in(const RowArgs& a, uint32_t items, size_t lds, int rs, hipStream_t s) {
  int vec = 0;
  if constexpr (sizeof(Tin) <= 4) {
    constexpr int VX = 8 / sizeof(Tin);
    const int wl = std::min(a.W, a.nx);
    vec = (static_cast<size_t>(a.W) * sizeof(Tin)) % 8 == 0 && wl % VX == 0 &&
          reinterpret_cast<uintptr_t>(a.img) % 8 == 0 && reinterpret_cast<uintptr_t>(a.img0) % 8 == 0 &&
          reinterpret_cast<uintptr_t>(a.window) % 16 == 0;
  }
  switch (a.Tb) {
    case 4: return rows_mr_np<Tin, T, 2>(a, items, lds, rs, vec, s);
    case 8: return rows_mr_np<Tin, T, 4>(a, items, lds, rs, vec, s);
    case 16: return rows_mr_np<Tin, T, 8>(a, items, lds, rs, vec, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t rows_mr_t(const RowArgs& a, uint32_t items, size_t lds, int rs, hipStream_t s) {
  switch (a.in_dtype) {
    case IN_U8: return rows_mr_in<uint8_t, T>(a, items, lds, rs, s);
    case IN_I16: return rows_mr_in<int16_t, T>(a, items, lds, rs, s);
    case IN_U16: return rows_mr_in<uint16_t, T>(a, items, lds, rs, s);
    case IN_I32: return rows_mr_in<int32_t, T>(a, items, lds, rs, s);
    case IN_U32: return rows_mr_in<uint32_t, T>(a, items, lds, rs, s);
    case IN_I64: return rows_mr_in<int64_t, T>(a, items, lds, rs, s);
    case IN_U64: return rows_mr_in<uint64_t, T>(a, items, lds, rs, s);
    case IN_F32: return rows_mr_in<float, T>(a, items, lds, rs, s);
    case IN_F64: return rows_mr_in<double, T>(a, items, lds, rs, s);
    default: return hipErrorInvalidValue;
  }
}

size_t cols_mr_lds(const ColArgs& a, int precision, int* rs) {
  *rs = lds_padded_size(a.ny);
  const size_t csz = precision == PREC_FP64 ? 16 : 8;
  return csz * (static_cast<size_t>(a.Tb) * *rs + std::max(a.ny - 1, 1));
}

template <typename T, int TB>
hipError_t cols_mr_tb(const ColArgs& a, uint32_t items, size_t lds, int rs, hipStream_t s) {
  // (the long form is FP32 only, see cols_mr_ok)
  constexpr int KL = sizeof(T) == 4 ? kMrLong : kMrShort;
  auto kernel = a.ny <= kMrShort ? k_cols_mr<T, TB, kMrShort> : k_cols_mr<T, TB, KL>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel), 64 * TB, lds);
  if (e != hipSuccess) return e;
  const uint32_t resident = static_cast<uint32_t>(std::max(1, cus) * std::max(1, per_cu)) & ~7u;
  const uint32_t grid = std::max(8u, std::min(resident, (items + 7u) & ~7u));
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(64 * TB), lds, s, a, items, rs);
  return hipGetLastError();
}

template <typename T>
hipError_t cols_mr_t(const ColArgs& a, uint32_t items, size_t lds, int rs, hipStream_t s) {
  switch (a.Tb) {
    case 4: return cols_mr_tb<T, 4>(a, items, lds, rs, s);
    case 8: return cols_mr_tb<T, 8>(a, items, lds, rs, s);
    case 16: return cols_mr_tb<T, 16>(a, items, lds, rs, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

bool rows_mr_ok(const RowArgs& a, int precision) {
  if (a.nx < 2 || a.nx > kMrLong || !(a.Tb == 4 || a.Tb == 8 || a.Tb == 16)) return false;
  int rs = 0;
  const size_t lds = rows_mr_lds(a, precision, &rs);
  // FP64 rows longer than kMrShort only while two workgroups fit a CU: with
  // one (two waves) the generic kernel was as fast (1080 x 1920 x 400: 8.35
  // vs 8.08 ms; 1000^2 x 800, three workgroups: 4.5 vs 6.3 ms)
  if (precision == PREC_FP64 && a.nx > kMrShort && lds > 80 * 1024) return false;
  return lds <= 160 * 1024;
}

hipError_t launch_rows_mr(const RowArgs& a, int precision, hipStream_t s) {
  const uint64_t nblocks = (a.frames + a.Tb - 1) / a.Tb;
  const uint64_t items = nblocks * static_cast<uint64_t>(a.H);
  if (items == 0) return hipSuccess;
  if (items > 0xfffffff0ull) return hipErrorInvalidValue;
  int rs = 0;
  const size_t lds = rows_mr_lds(a, precision, &rs);
  return precision == PREC_FP64 ? rows_mr_t<double>(a, static_cast<uint32_t>(items), lds, rs, s)
                                : rows_mr_t<float>(a, static_cast<uint32_t>(items), lds, rs, s);
}

bool cols_mr_ok(const ColArgs& a, int precision) {
  // Columns shorter than 256 stay on the generic kernel, whose workgroup
  // shares its lanes among all the block's transforms (one wave per 96- or
  // 150-point column leaves most lanes idle: 48-stack batch, 96^2 columns
  // 0.20 ms generic vs 0.35 ms here, 150 x 200 0.52 vs 0.67 ms)
  if (a.ny < 256 || a.ny > kMrLong || !(a.Tb == 4 || a.Tb == 8 || a.Tb == 16)) return false;
  // FP64 columns longer than kMrShort measured slower than the generic
  // kernel (1000^2 x 800: 8.76 vs 5.85 ms; 1080 x 1920 x 400: 10.2 vs 6.8 ms)
  if (precision == PREC_FP64 && a.ny > kMrShort) return false;
  int rs = 0;
  return cols_mr_lds(a, precision, &rs) <= 160 * 1024;
}

hipError_t launch_cols_mr(const ColArgs& a, int precision, hipStream_t s) {
  const uint64_t items = a.nblocks * static_cast<uint64_t>(a.nxh);
  if (items == 0) return hipSuccess;
  if (items > 0xfffffff0ull) return hipErrorInvalidValue;
  int rs = 0;
  const size_t lds = cols_mr_lds(a, precision, &rs);
  return precision == PREC_FP64 ? cols_mr_t<double>(a, static_cast<uint32_t>(items), lds, rs, s)
                                : cols_mr_t<float>(a, static_cast<uint32_t>(items), lds, rs, s);
}

}  // namespace fddm
