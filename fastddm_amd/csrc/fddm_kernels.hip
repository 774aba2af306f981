// Kernels of the image structure function path (gfx950, wave64).
//
//   k_rows  stage (a), row pass: pixel convert (+ frame-0 subtraction) x window,
//           two frames packed as one complex row, length-nx FFT, split into the
//           two half spectra, blocked store R[tb][kx][y][ti].
//   k_cols  stage (a) column pass + stage (b) transpose: length-ny FFT of Tb
//           frames of one kx column at once, store q-major S[q][t] in runs of
//           Tb consecutive frames.
//   k_diff  stage (d) + (e): direct lag differences, one wave per q, PS and VAR.
//
// The fft-mode time stage lives in fddm_time_p2.hip (power-of-two lengths).
//
// Semantics follow the reference cores: spectra and normalisation
// src/cpp/ddm.cc:288-323, correlation and two-ended sums ddm.cc:344-426,
// differences ddm.cc:117-171, output shift src/cpp/helper_ddm.cc:41-103.
#include <type_traits>

#include "fddm_fft.hpp"
#include "fddm_kernels.hpp"
#include "fddm_wave.hpp"

namespace fddm {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

int fft_bucket(long long work) {
  const long long e = (work + kThreads - 1) / kThreads;
  if (e <= 4) return 4;
  if (e <= 8) return 8;
  if (e <= 16) return 16;
  if (e <= 32) return 32;
  if (e <= 64) return 64;
  return 0;
}

// ---------------------------------------------------------------------------
// Pixel conversion.  Integer frame differences are formed exactly in integer
// arithmetic before the conversion to T.
// ---------------------------------------------------------------------------
template <typename Tin, typename T>
__device__ __forceinline__ T load_pixel(const Tin* __restrict__ img, const Tin* __restrict__ ref, size_t idx, size_t idx0,
                                        bool sub) {
  if constexpr (std::is_integral_v<Tin> && sizeof(Tin) <= 2) {
    int v = static_cast<int>(img[idx]);
    if (sub) v -= static_cast<int>(ref[idx0]);
    return static_cast<T>(v);
  } else if constexpr (std::is_integral_v<Tin> && sizeof(Tin) == 4) {
    long long v = static_cast<long long>(img[idx]);
    if (sub) v -= static_cast<long long>(ref[idx0]);
    return static_cast<T>(v);
  } else if constexpr (std::is_integral_v<Tin>) {
    double v = static_cast<double>(img[idx]);
    if (sub) v -= static_cast<double>(ref[idx0]);
    return static_cast<T>(v);
  } else {
    // floating pixels: the difference in the wider of the two types
    using Wd = std::conditional_t<(sizeof(Tin) > sizeof(T)), Tin, T>;
    Wd v = static_cast<Wd>(img[idx]);
    if (sub) v -= static_cast<Wd>(ref[idx0]);
    return static_cast<T>(v);
  }
}

// ---------------------------------------------------------------------------
// Stage (a): row pass.
// ---------------------------------------------------------------------------
template <typename Tin, typename T, int E>
__global__ void __launch_bounds__(kThreads) k_rows(RowArgs a) {
  using C = cplx_t<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  C* buf = reinterpret_cast<C*>(smem);
  const int tid = threadIdx.x;
  const int H = a.H, W = a.W, nx = a.nx, Tb = a.Tb, G = a.G;
  const int nxh = nx / 2 + 1;
  const int y = blockIdx.x % H;
  const uint64_t tb = blockIdx.x / H;
  const uint64_t t0 = tb * Tb;
  const Tin* __restrict__ img = static_cast<const Tin*>(a.img);
  const Tin* __restrict__ ref = static_cast<const Tin*>(a.img0);  // the subtracted frame
  const T* __restrict__ win = static_cast<const T*>(a.window);
  const T half_norm = static_cast<T>(0.5 * a.norm);
  const int bstride = nx + 1;
  const size_t plane = static_cast<size_t>(H) * W;
  C* __restrict__ R = static_cast<C*>(a.R);
  const bool sub = a.subtract0 != 0;
  const C* __restrict__ tw = static_cast<const C*>(a.twx);

  for (int g0 = 0; g0 < Tb / 2; g0 += G) {
    for (int w = tid; w < G * nx; w += kThreads) {
      const int p = w / nx, x = w - p * nx;
      T va = 0, vb = 0;
      if (x < W) {
        const uint64_t ta = t0 + 2 * (g0 + p), tb2 = ta + 1;
        const size_t i0 = static_cast<size_t>(y) * W + x;
        if (ta < a.frames) va = load_pixel<Tin, T>(img, ref, ta * plane + i0, i0, sub);
        if (tb2 < a.frames) vb = load_pixel<Tin, T>(img, ref, tb2 * plane + i0, i0, sub);
        if (win) {
          const T wv = win[i0];
          va *= wv;
          vb *= wv;
        }
      }
      buf[lds_pad(p * bstride + x)] = C{va, vb};
    }
    __syncthreads();
    fft_lds<T, E, false>(buf, 0, a.dx, G, 0, bstride, 1, tw, tid, kThreads);
    // Z = A + i B with A, B the spectra of the two real rows:
    // A(k) = (Z(k) + conj Z(-k)) / 2, B(k) = (Z(k) - conj Z(-k)) / 2i.
    const int nti = 2 * G;
    for (int w = tid; w < nxh * nti; w += kThreads) {
      const int tl = w % nti, k = w / nti;
      const int p = tl >> 1;
      const C zk = buf[lds_pad(p * bstride + k)];
      const C zn = buf[lds_pad(p * bstride + (k == 0 ? 0 : nx - k))];
      C o;
      if ((tl & 1) == 0) {
        o = C{(zk.x + zn.x) * half_norm, (zk.y - zn.y) * half_norm};
      } else {
        o = C{(zk.y + zn.y) * half_norm, (zn.x - zk.x) * half_norm};
      }
      const int ti = 2 * g0 + tl;
      R[((tb * nxh + k) * static_cast<size_t>(H) + y) * Tb + ti] = o;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Stage (a) columns + (b) q-major store.
// ---------------------------------------------------------------------------
template <typename T, int E>
__global__ void __launch_bounds__(kThreads) k_cols(ColArgs a) {
  using C = cplx_t<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  C* buf = reinterpret_cast<C*>(smem);
  const int tid = threadIdx.x;
  const int nxh = a.nxh, H = a.H, ny = a.ny, Tb = a.Tb;
  const int kx = blockIdx.x % nxh;
  const uint64_t tb = blockIdx.x / nxh;
  const C* __restrict__ src = static_cast<const C*>(a.R) + (tb * nxh + kx) * static_cast<size_t>(H) * Tb;
  const int nin = H * Tb, nall = ny * Tb;
  for (int w = tid; w < nall; w += kThreads) buf[lds_pad(w)] = w < nin ? src[w] : C{0, 0};
  __syncthreads();
  const C* __restrict__ tw = static_cast<const C*>(a.twy);
  const int logG = __builtin_ctz(a.G);
  for (int g0 = 0; g0 < Tb; g0 += a.G) fft_lds<T, E, true>(buf, g0, a.dy, a.G, logG, 1, Tb, tw, tid, kThreads);
  C* __restrict__ S = static_cast<C*>(a.S);
  for (int w = tid; w < nall; w += kThreads) {
    const int ti = w % Tb, ky = w / Tb;
    S[(static_cast<size_t>(ky) * nxh + kx) * a.Lp + tb * Tb + ti] = buf[lds_pad(w)];
  }
}

// ---------------------------------------------------------------------------
// Stage (d)+(e): direct differences.  A wave works on QPW q at once, one per
// segment of SEG = 64 / QPW lanes (QPW = 4: one q per 16-lane DPP row, 2: one
// per half wave); the
// segment's series sits in its own LDS slice when the workgroup's slices fit
// 64 KiB (else it is read from L2).  Lane l of a segment takes t = l, l + SEG,
// ...; lags go in batches of kLagBatch, so each F(t) load serves kLagBatch
// differences, and each lag's SEG partial sums are added by a DPP scan inside
// the segment, whose last lane writes D: one 4-step scan serves four q, one
// 5-step scan two.  Differences are formed after widening to Acc, so they are exact on
// the compensated paths.
// ---------------------------------------------------------------------------
constexpr int kLagBatch = 4;

// Inclusive scan inside segments of SEG lanes (16: DPP rows; 32: wave
// halves, rows 1 and 3 adding the last lane of rows 0 and 2; 64: the wave).
template <int SEG, typename A>
__device__ __forceinline__ A seg_total_scan(A v) {
  if constexpr (SEG == 64) {
    return wave::incl_scan(v);
  } else {
    static_assert(SEG == 16 || SEG == 32, "segment");
    v += wave::dpp_zero<wave::kRowShr1, 0xf>(v);
    v += wave::dpp_zero<wave::kRowShr2, 0xf>(v);
    v += wave::dpp_zero<wave::kRowShr4, 0xf>(v);
    v += wave::dpp_zero<wave::kRowShr8, 0xf>(v);
    if constexpr (SEG == 32) v += wave::dpp_zero<wave::kRowBcast15, 0xa>(v);
    return v;
  }
}

template <typename T, typename OutT, int PREC, bool USE_LDS, int QPW>
__global__ void __launch_bounds__(kThreads) k_diff(TimeArgs a, int wg_per_xcd, int nq) {
  using C = cplx_t<T>;
  using Acc = std::conditional_t<PREC == PREC_FP32, float, double>;
  constexpr int SEG = 64 / QPW;
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int sl = lane & (SEG - 1), seg = lane / SEG;
  // columns of the launch (nq of them): cw is this wave's first
  const int cw = ((blockIdx.x & 7) * wg_per_xcd + (blockIdx.x >> 3)) * kWaves * QPW + wid * QPW;
  if (cw >= nq) return;  // whole waves: nothing below synchronises the workgroup
  const int c = cw + seg;
  const bool valid = c < nq;
  const int q = col_q(a, valid ? c : cw);
  const int L = a.L, Lv = valid ? L : 0;  // segments past the last column run empty loops
  const C* __restrict__ src = static_cast<const C*>(a.S) + static_cast<size_t>(q - a.s_q0) * a.Lp;
  const C* ser = src;
  if constexpr (USE_LDS) {
    C* slice = reinterpret_cast<C*>(smem) + static_cast<size_t>(wid * QPW + seg) * L;
    for (int t0 = sl; t0 < Lv; t0 += 4 * SEG) {  // four loads in flight per lane
      C v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) v[k] = t0 + SEG * k < Lv ? src[t0 + SEG * k] : C{0, 0};
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (t0 + SEG * k < Lv) slice[t0 + SEG * k] = v[k];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    ser = slice;
  }
  Acc sr = 0, si = 0, s2 = 0;
  for (int t = sl; t < Lv; t += SEG) {
    const C f = ser[t];
    const Acc fx = static_cast<Acc>(f.x), fy = static_cast<Acc>(f.y);
    sr += fx;
    si += fy;
    s2 += fx * fx + fy * fy;
  }
  sr = seg_total_scan<SEG>(sr);  // the segment's last lane holds the sums
  si = seg_total_scan<SEG>(si);
  s2 = seg_total_scan<SEG>(s2);
  const bool writer = valid && sl == SEG - 1;

  OutT* __restrict__ out = static_cast<OutT*>(a.out);
  const size_t nqs = a.ldo;
  const size_t oq = col_out(a, valid ? c : cw);
  for (int i0 = 0; i0 < a.nl; i0 += kLagBatch) {
    int d[kLagBatch];
#pragma unroll
    for (int k = 0; k < kLagBatch; ++k) d[k] = i0 + k < a.nl ? static_cast<int>(a.lags[i0 + k]) : L;
    int dmin = d[0];
#pragma unroll
    for (int k = 1; k < kLagBatch; ++k) dmin = d[k] < dmin ? d[k] : dmin;
    Acc acc[kLagBatch] = {};
    for (int t = sl; t + dmin < Lv; t += SEG) {
      const C v = ser[t];
      const Acc vx = static_cast<Acc>(v.x), vy = static_cast<Acc>(v.y);
#pragma unroll
      for (int k = 0; k < kLagBatch; ++k) {
        if (t + d[k] < L) {
          const C u = ser[t + d[k]];
          const Acc dr = static_cast<Acc>(u.x) - vx;
          const Acc di = static_cast<Acc>(u.y) - vy;
          acc[k] += dr * dr + di * di;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kLagBatch; ++k) {
      const Acc tot = seg_total_scan<SEG>(acc[k]);
      if (writer && i0 + k < a.nl)
        out[static_cast<size_t>(i0 + k) * nqs + oq] = static_cast<OutT>(tot / static_cast<Acc>(L - d[k]));
    }
  }
  if (writer) {
    const Acc mr = sr / L, mi = si / L;
    Acc ps, var;
    if (a.S0) {
      // spectra hold F(t) - F(0): VAR is shift invariant, PS adds |F(0) + mean|^2
      var = s2 / L - (mr * mr + mi * mi);
      const C z = static_cast<const C*>(a.S0)[static_cast<size_t>(q) * a.s0_stride];
      const Acc fr = mr + static_cast<Acc>(z.x), fi = mi + static_cast<Acc>(z.y);
      ps = var + fr * fr + fi * fi;
    } else {
      ps = s2 / L;
      var = ps - mr * mr - mi * mi;
    }
    out[static_cast<size_t>(a.nl) * nqs + oq] = static_cast<OutT>(ps);
    out[static_cast<size_t>(a.nl + 1) * nqs + oq] = static_cast<OutT>(var);
  }
}

// ---------------------------------------------------------------------------
// FFT engine test kernel: one forward transform per workgroup, in place.
// ---------------------------------------------------------------------------
template <typename T, int E>
__global__ void __launch_bounds__(kThreads) k_fft_test(cplx_t<T>* data, FftDesc d, const cplx_t<T>* tw) {
  using C = cplx_t<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  C* buf = reinterpret_cast<C*>(smem);
  C* p = data + static_cast<size_t>(blockIdx.x) * d.n;
  for (int t = threadIdx.x; t < d.n; t += kThreads) buf[lds_pad(t)] = p[t];
  __syncthreads();
  fft_lds<T, E, false>(buf, 0, d, 1, 0, d.n, 1, tw, threadIdx.x, kThreads);
  for (int t = threadIdx.x; t < d.n; t += kThreads) p[t] = buf[lds_pad(t)];
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
namespace {

template <typename K, typename... Args>
hipError_t launch(K kernel, unsigned grid, size_t lds, hipStream_t s, Args... args) {
  if (grid == 0) return hipSuccess;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), lds, s, args...);
  return hipGetLastError();
}

template <typename Tin, typename T>
hipError_t rows_e(const RowArgs& a, unsigned grid, size_t lds, int e, hipStream_t s) {
  switch (e) {
    case 16: return launch(k_rows<Tin, T, 16>, grid, lds, s, a);
    case 32: return launch(k_rows<Tin, T, 32>, grid, lds, s, a);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t rows_t(const RowArgs& a, unsigned grid, size_t lds, int e, hipStream_t s) {
  switch (a.in_dtype) {
    case IN_U8: return rows_e<uint8_t, T>(a, grid, lds, e, s);
    case IN_I16: return rows_e<int16_t, T>(a, grid, lds, e, s);
    case IN_U16: return rows_e<uint16_t, T>(a, grid, lds, e, s);
    case IN_I32: return rows_e<int32_t, T>(a, grid, lds, e, s);
    case IN_U32: return rows_e<uint32_t, T>(a, grid, lds, e, s);
    case IN_I64: return rows_e<int64_t, T>(a, grid, lds, e, s);
    case IN_U64: return rows_e<uint64_t, T>(a, grid, lds, e, s);
    case IN_F32: return rows_e<float, T>(a, grid, lds, e, s);
    case IN_F64: return rows_e<double, T>(a, grid, lds, e, s);
    default: return hipErrorInvalidValue;
  }
}

int rc_bucket(long long work) {
  const int e = fft_bucket(work);
  return e <= 16 ? 16 : e;
}

}  // namespace

hipError_t launch_rows(const RowArgs& a, int precision, hipStream_t s) {
  if (!a.generic && rows_mr_ok(a, precision)) return launch_rows_mr(a, precision, s);
  const uint64_t nblocks = (a.frames + a.Tb - 1) / a.Tb;
  const unsigned grid = static_cast<unsigned>(nblocks * a.H);
  const int e = rc_bucket(static_cast<long long>(a.G) * a.nx);
  if (precision == PREC_FP64) {
    return rows_t<double>(a, grid, sizeof(double2) * lds_padded_size(a.G * (a.nx + 1)), e, s);
  }
  return rows_t<float>(a, grid, sizeof(float2) * lds_padded_size(a.G * (a.nx + 1)), e, s);
}

hipError_t launch_cols(const ColArgs& a, int precision, hipStream_t s) {
  if (!a.generic && cols_mr_ok(a, precision)) return launch_cols_mr(a, precision, s);
  const unsigned grid = static_cast<unsigned>(a.nblocks * a.nxh);
  const int e = rc_bucket(static_cast<long long>(a.G) * a.ny);
  if (precision == PREC_FP64) {
    const size_t lds = sizeof(double2) * lds_padded_size(a.ny * a.Tb);
    if (e == 16) return launch(k_cols<double, 16>, grid, lds, s, a);
    if (e == 32) return launch(k_cols<double, 32>, grid, lds, s, a);
  } else {
    const size_t lds = sizeof(float2) * lds_padded_size(a.ny * a.Tb);
    if (e == 16) return launch(k_cols<float, 16>, grid, lds, s, a);
    if (e == 32) return launch(k_cols<float, 32>, grid, lds, s, a);
  }
  return hipErrorInvalidValue;
}

namespace {
// The most q per wave (4, 2, 1) whose staged series leave room for 8
// workgroups (32 waves) per CU; else one q per wave, LDS-staged while the
// workgroup's four series fit 64 KiB, else read from L2.
template <typename T, typename OutT, int PREC>
hipError_t diff_l(const TimeArgs& a, hipStream_t s) {
  const int nq = a.cols;
  const size_t series = sizeof(cplx_t<T>) * a.L;
  constexpr size_t kFull = 160 * 1024 / 8;  // LDS per workgroup at 8 workgroups per CU
  auto go = [&](auto kernel, int qpw, size_t lds) {
    const int wgs = (nq + kWaves * qpw - 1) / (kWaves * qpw);
    const int wpx = (wgs + 7) / 8;
    return launch(kernel, static_cast<unsigned>(wpx) * 8u, lds, s, a, wpx, nq);
  };
  if (series * kWaves * 4 <= kFull) return go(k_diff<T, OutT, PREC, true, 4>, 4, series * kWaves * 4);
  if (series * kWaves * 2 <= kFull) return go(k_diff<T, OutT, PREC, true, 2>, 2, series * kWaves * 2);
  if (series * kWaves <= 64 * 1024) return go(k_diff<T, OutT, PREC, true, 1>, 1, series * kWaves);
  return go(k_diff<T, OutT, PREC, false, 1>, 1, 0);
}
}  // namespace

hipError_t launch_diff(const TimeArgs& a, int precision, int out_dtype, hipStream_t s) {
  const bool f32 = out_dtype == OUT_F32;
  switch (precision) {
    case PREC_FP64: return f32 ? diff_l<double, float, PREC_FP64>(a, s) : diff_l<double, double, PREC_FP64>(a, s);
    case PREC_FP32: return f32 ? diff_l<float, float, PREC_FP32>(a, s) : diff_l<float, double, PREC_FP32>(a, s);
    case PREC_FP32C: return f32 ? diff_l<float, float, PREC_FP32C>(a, s) : diff_l<float, double, PREC_FP32C>(a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fft_test(void* data, int n, int batch, const FftDesc& d, const void* tw, int precision,
                           hipStream_t s) {
  const int e = fft_bucket(n);
  if (precision == PREC_FP64) {
    using C = double2;
    const size_t lds = sizeof(C) * lds_padded_size(n);
    C* p = static_cast<C*>(data);
    const C* t = static_cast<const C*>(tw);
    switch (e) {
      case 4: return launch(k_fft_test<double, 4>, batch, lds, s, p, d, t);
      case 8: return launch(k_fft_test<double, 8>, batch, lds, s, p, d, t);
      case 16: return launch(k_fft_test<double, 16>, batch, lds, s, p, d, t);
      case 32: return launch(k_fft_test<double, 32>, batch, lds, s, p, d, t);
      default: return hipErrorInvalidValue;
    }
  }
  using C = float2;
  const size_t lds = sizeof(C) * lds_padded_size(n);
  C* p = static_cast<C*>(data);
  const C* t = static_cast<const C*>(tw);
  switch (e) {
    case 4: return launch(k_fft_test<float, 4>, batch, lds, s, p, d, t);
    case 8: return launch(k_fft_test<float, 8>, batch, lds, s, p, d, t);
    case 16: return launch(k_fft_test<float, 16>, batch, lds, s, p, d, t);
    case 32: return launch(k_fft_test<float, 32>, batch, lds, s, p, d, t);
    case 64: return launch(k_fft_test<float, 64>, batch, lds, s, p, d, t);
    default: return hipErrorInvalidValue;
  }
}

namespace {

// Device buffer -> pinned host buffer, 16 bytes per lane and four in flight
// per lane: the GPU writes the host pages itself over the host link.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy_out(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                  uint64_t n16) {
  const uint64_t stride = uint64_t(gridDim.x) * 256;
  uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const u32x4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    dst[i] = a;
    dst[i + stride] = b;
    dst[i + 2 * stride] = c;
    dst[i + 3 * stride] = d;
  }
  for (; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace

hipError_t launch_copy_out(void* dst, const void* src, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (n % 16 || reinterpret_cast<uintptr_t>(src) % 16 || reinterpret_cast<uintptr_t>(dst) % 16)
    return hipErrorInvalidValue;
  const uint64_t n16 = n / 16;
  const unsigned blocks = static_cast<unsigned>(std::min<uint64_t>((n16 + 255) / 256, 1024));
  hipLaunchKernelGGL(k_copy_out, dim3(blocks), dim3(256), 0, s, static_cast<const u32x4*>(src),
                     static_cast<u32x4*>(dst), n16);
  return hipGetLastError();
}

}  // namespace fddm
