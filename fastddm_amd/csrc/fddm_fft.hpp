// Device FFT engine: mixed-radix Stockham autosort FFT in LDS, executed
// cooperatively by one workgroup on a batch of transforms.
//
// Formulation: the radix-R Stockham pass of Govindaraju et al., "High
// Performance Discrete Fourier Transforms on Graphics Processors" (SC 2008):
// with Ns the product of the radices already applied, butterfly j reads
// x[j + r N/R] (r < R), multiplies by w^(r k), k = j mod Ns, w = exp(-2 pi i /
// (Ns R)), applies a length-R DFT and writes y[(j - k) R + k + r Ns].  The
// output is in natural order after the last pass.  Every pass is done in
// place: each thread first reads all of its butterflies into registers, the
// workgroup synchronises, then the results are written back.
//
// Butterflies: radix 2, 4, 8 in closed form; radix 3, 5, 7 from the
// symmetric-pair form of the DFT,
//   X_s = v_0 + sum_r (v_r + v_{R-r}) cos(2 pi r s / R) - i (v_r - v_{R-r}) sin(2 pi r s / R).
// Twiddles come from per-pass tables exp(-2 pi i r k / (Ns R)) built on the
// host in extended precision and stored contiguously per butterfly.
#pragma once

#include "fddm_common.hpp"

namespace fddm {

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
  return C{a.x + b.x, a.y + b.y};
}
template <typename C>
__device__ __forceinline__ C csub(C a, C b) {
  return C{a.x - b.x, a.y - b.y};
}
// -i * a
template <typename C>
__device__ __forceinline__ C cmul_mi(C a) {
  return C{a.y, -a.x};
}

// cos(2 pi m / R) and sin(2 pi m / R) for the odd radices, m = 0 .. (R - 1) / 2.
template <int R>
__host__ __device__ constexpr double odd_cos(int m) {
  if constexpr (R == 3) {
    return m == 0 ? 1.0 : -0.5;
  } else if constexpr (R == 5) {
    return m == 0 ? 1.0 : (m == 1 ? 0.30901699437494742410229341718282 : -0.80901699437494742410229341718282);
  } else {
    return m == 0 ? 1.0
                  : (m == 1 ? 0.62348980185873353052500488400424
                            : (m == 2 ? -0.22252093395631440428890256449680 : -0.90096886790241912623610231950745));
  }
}
template <int R>
__host__ __device__ constexpr double odd_sin(int m) {
  if constexpr (R == 3) {
    return m == 0 ? 0.0 : 0.86602540378443864676372317075294;
  } else if constexpr (R == 5) {
    return m == 0 ? 0.0 : (m == 1 ? 0.95105651629515357211643933337938 : 0.58778525229247312916870595463907);
  } else {
    return m == 0 ? 0.0
                  : (m == 1 ? 0.78183148246802980870844452667406
                            : (m == 2 ? 0.97492791218182360701813168299393 : 0.43388373911755812047576833284836));
  }
}

template <typename T, int R>
__device__ __forceinline__ void butterfly_odd(cplx_t<T> (&v)[R]) {
  using C = cplx_t<T>;
  constexpr int H = (R - 1) / 2;
  C a[H + 1], b[H + 1];  // a_r = v_r + v_{R-r}, b_r = v_r - v_{R-r}
#pragma unroll
  for (int r = 1; r <= H; ++r) {
    a[r] = cadd(v[r], v[R - r]);
    b[r] = csub(v[r], v[R - r]);
  }
  C x0 = v[0];
#pragma unroll
  for (int r = 1; r <= H; ++r) x0 = cadd(x0, a[r]);
  C out[R];
  out[0] = x0;
#pragma unroll
  for (int s = 1; s <= H; ++s) {
    T ar = v[0].x, ai = v[0].y, br = 0, bi = 0;
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      const int m = (r * s) % R;
      const int mm = m <= H ? m : R - m;
      const T cc = static_cast<T>(odd_cos<R>(mm));
      const T ss = m <= H ? static_cast<T>(odd_sin<R>(mm)) : -static_cast<T>(odd_sin<R>(mm));
      ar += a[r].x * cc;
      ai += a[r].y * cc;
      br += b[r].x * ss;
      bi += b[r].y * ss;
    }
    // X_s = A - i B, X_{R-s} = A + i B
    out[s] = C{ar + bi, ai - br};
    out[R - s] = C{ar - bi, ai + br};
  }
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = out[r];
}

template <typename T>
__device__ __forceinline__ void dft4(cplx_t<T>& v0, cplx_t<T>& v1, cplx_t<T>& v2, cplx_t<T>& v3) {
  using C = cplx_t<T>;
  const C a = cadd(v0, v2), b = csub(v0, v2), c = cadd(v1, v3), d = csub(v1, v3);
  v0 = cadd(a, c);
  v2 = csub(a, c);
  v1 = C{b.x + d.y, b.y - d.x};
  v3 = C{b.x - d.y, b.y + d.x};
}

template <typename T, int R>
__device__ __forceinline__ void butterfly(cplx_t<T> (&v)[R]) {
  using C = cplx_t<T>;
  if constexpr (R == 2) {
    const C a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  } else if constexpr (R == 4) {
    dft4<T>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    constexpr T h = static_cast<T>(0.70710678118654752440084436210485);
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<T>(e0, e1, e2, e3);
    dft4<T>(o0, o1, o2, o3);
    const C w1 = C{(o1.x + o1.y) * h, (o1.y - o1.x) * h};
    const C w2 = cmul_mi(o2);
    const C w3 = C{(o3.y - o3.x) * h, -(o3.x + o3.y) * h};
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, w1);
    v[5] = csub(e1, w1);
    v[2] = cadd(e2, w2);
    v[6] = csub(e2, w2);
    v[3] = cadd(e3, w3);
    v[7] = csub(e3, w3);
  } else {
    butterfly_odd<T, R>(v);
  }
}

// LDS address of logical element e: one pad slot every 32 elements spreads
// the strided Stockham writes of the early passes over the banks.
__device__ __forceinline__ int lds_pad(int e) { return e + (e >> 5); }

// One in-place radix-R pass over nb transforms of length n held in LDS.
// Element idx of transform ti lives at buf[lds_pad(off + ti * bstride + idx * estride)].
// TIF: consecutive work items walk the transform index first (nb a power of
// two, lognb = log2 nb; used when the transforms are interleaved); otherwise
// the butterfly index first.  MB work items per thread must cover n / R * nb.
// tw: this pass's table, tw[k (R-1) + r - 1] = exp(-2 pi i r k / (ns R)).
// Item indices advance incrementally (m = n / R is a multiple of ns), so no
// per-item integer division is needed.
template <typename T, int R, int MB, bool TIF>
__device__ __forceinline__ void stockham_pass(cplx_t<T>* buf, int off, int n, int ns, int nb,
                                              int lognb, int bstride, int estride,
                                              const cplx_t<T>* __restrict__ tw, int tid,
                                              int nthr) {
  using C = cplx_t<T>;
  // Opaque copy of the thread index: keeps the per-item index arithmetic
  // inside the pass instead of letting it be hoisted (and kept live) across
  // all passes of the transform.
  asm volatile("" : "+v"(tid));
  const int m = n / R;
  const int items = m * nb;
  int j, ti, dj, dti;
  if constexpr (TIF) {
    ti = tid & (nb - 1);
    j = tid >> lognb;
    dj = nthr >> lognb;
    dti = 0;
  } else {
    j = tid % m;
    ti = tid / m;
    dj = nthr % m;
    dti = nthr / m;
  }
  int k = j % ns;
  const int dk = dj % ns;
  C v[MB][R];
  int obase[MB];
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    obase[b] = -1;
    if (tid + b * nthr < items) {
      const int base = off + ti * bstride;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = buf[lds_pad(base + (j + r * m) * estride)];
      if (ns > 1) {
        const C* __restrict__ t = tw + k * (R - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], t[r - 1]);
      }
      butterfly<T, R>(v[b]);
      obase[b] = base + (j * R - k * (R - 1)) * estride;
    }
    j += dj;
    k += dk;
    if (k >= ns) k -= ns;
    if constexpr (!TIF) {
      ti += dti;
      if (j >= m) {
        j -= m;
        ++ti;
      }
    }
  }
  __syncthreads();
  const int ostep = ns * estride;
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    if (obase[b] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[lds_pad(obase[b] + r * ostep)] = v[b][r];
    }
  }
  __syncthreads();
}

// Forward FFT (unnormalised) of nb transforms in LDS.  E bounds the elements
// per thread: E * nthr >= n * nb.  `tw` is the plan's concatenated per-pass
// twiddle table (pass order: 8s, 4s, 2s, 3s, 5s, 7s).  The caller
// synchronises before the call; the data are synchronised on return.  One
// loop per radix (rather than a switch over a pass list) keeps the register
// allocation of the six pass variants disjoint.
template <typename T, int E, bool TIF>
__device__ void fft_lds(cplx_t<T>* buf, int off, const FftDesc& d, int nb, int lognb, int bstride,
                        int estride, const cplx_t<T>* __restrict__ tw, int tid, int nthr) {
  const int n = d.n;
  int ns = 1;
#define FDDM_PASSES(RAD, COUNT)                                                              \
  for (int i = 0; i < d.COUNT; ++i) {                                                        \
    stockham_pass<T, RAD, (E + RAD - 1) / RAD, TIF>(buf, off, n, ns, nb, lognb, bstride,     \
                                                    estride, tw, tid, nthr);                 \
    tw += ns * (RAD - 1);                                                                    \
    ns *= RAD;                                                                               \
  }
  FDDM_PASSES(8, c8)
  FDDM_PASSES(4, c4)
  FDDM_PASSES(2, c2)
  FDDM_PASSES(3, c3)
  FDDM_PASSES(5, c5)
  FDDM_PASSES(7, c7)
#undef FDDM_PASSES
}

}  // namespace fddm
