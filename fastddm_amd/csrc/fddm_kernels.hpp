// Host-side launchers of the fastddm_amd kernels.
#pragma once

#include "fddm_common.hpp"

namespace fddm {

// Stage (a), row pass.  Frames [0, L) of the (L, H, W) stack `img` (pixel
// type in_dtype) -> per-row half spectra along x, scaled by `norm`, stored in
// the blocked layout R[tb][kx][y][ti] (t = tb * Tb + ti, kx < nx/2+1, y < H).
// `window` (H*W of T) may be null.  subtract0: use I(t) - I(0) in place of I(t).
// `frames` limits how many leading frames are transformed (>= 1).
struct RowArgs {
  const void* img;
  int in_dtype;
  uint64_t L;       // frames present in img
  uint64_t frames;  // frames to transform (<= L)
  int H, W, nx, Tb, G;
  int logn;  // > 0: power-of-two kernel for nx = 2^logn (twx holds its tables)
  const void* window;
  int subtract0;
  double norm;
  FftDesc dx;
  const void* twx;
  void* R;
};

// Stage (a) column pass + q-major store: R -> S[q][t], q = ky * nxh + kx,
// row pitch Lp (multiple of Tb).  Transforms blocks tb in [0, nblocks).
struct ColArgs {
  const void* R;
  uint64_t nblocks;
  int H, ny, nxh, Tb, G;
  int logn;  // > 0: power-of-two kernel for ny = 2^logn (twy holds its tables)
  uint64_t Lp;
  FftDesc dy;
  const void* twy;
  void* S;
};

// Stage (c)/(d): per-q time series -> D(q, lags), PS, VAR in the y-shifted
// output (NL+2, ny, nxh).
struct TimeArgs {
  const void* S;
  uint64_t Lp;
  int L, ny, nxh;
  const void* S0;  // frame-0 spectra (stride s0_stride) or null
  uint64_t s0_stride;
  const uint32_t* lags;
  int nl;
  FftDesc dt;
  const void* twt;
  void* out;
};

hipError_t launch_rows(const RowArgs& a, int precision, hipStream_t s);
hipError_t launch_cols(const ColArgs& a, int precision, hipStream_t s);
hipError_t launch_diff(const TimeArgs& a, int precision, int out_dtype, hipStream_t s);
// Time kernel at N = 2^logn (p2_plan): LDS bytes per workgroup and launcher.
size_t p2_lds_bytes(int logn, int precision);
hipError_t launch_time_p2(const TimeArgs& a, int logn, int precision, int out_dtype, hipStream_t s);
// Power-of-two spatial kernels (256 <= n <= 4096): plan radices and launchers.
bool p2_spatial_ok(int logn);
int p2_spatial_plan(int logn, bool cols, int* r0, int* r1, int* r2);
hipError_t launch_rows_p2(const RowArgs& a, int precision, hipStream_t s);
hipError_t launch_cols_p2(const ColArgs& a, int precision, hipStream_t s);
hipError_t launch_fft_test(void* data, int n, int batch, const FftDesc& d, const void* tw,
                           int precision, hipStream_t s);

// Elements-per-thread template bucket for a transform of `work` points
// cooperatively done by 256 threads; 0 if too large.
int fft_bucket(long long work);

}  // namespace fddm
