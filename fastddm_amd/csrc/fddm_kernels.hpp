// Host-side launchers of the fastddm_amd kernels.
#pragma once

#include <vector>

#include "fddm_common.hpp"

namespace fddm {

// Stage (a), row pass.  Frames [0, L) of the (L, H, W) stack `img` (pixel
// type in_dtype) -> per-row half spectra along x, scaled by `norm`, stored in
// the blocked layout R[tb][kx][y][ti] (t = tb * Tb + ti, kx < nx/2+1, y < H).
// `window` (H*W of T) may be null.  subtract0: use I(t) - I0 in place of I(t),
// where I0 is the frame at img0 (the stack's frame 0; img itself when the
// launch starts at frame 0).  `frames` limits how many leading frames of img
// are transformed (>= 1).
struct RowArgs {
  const void* img;
  const void* img0;
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
  int generic;  // 1: the generic workgroup kernel even where k_rows_mr applies
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
  int generic;  // 1: the generic workgroup kernel even where k_cols_mr applies
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
  unsigned char* flags;  // fft mode, compensated FP32: per q, 1 = some lag needs launch_wk_fix
  // Columns of this launch, c in [0, cols).  local = 0: column c is output
  // column c0 + c of the y-shifted plane (row pitch ldo) and reads the q at
  // that position; local = 1: column c reads q = c0 + c and is stored
  // unshifted at out[i ldo + c].  S holds the rows of q >= s_q0 (row q - s_q0).
  int c0, cols, local;
  uint64_t ldo;
  int s_q0;
};

// Whole-plane launch: every q, y-shifted output of row pitch ny (nx/2+1).
inline void time_full(TimeArgs& a) {
  a.c0 = 0;
  a.cols = a.ny * a.nxh;
  a.local = 0;
  a.ldo = static_cast<uint64_t>(a.ny) * a.nxh;
  a.s_q0 = 0;
}

// q read by column c of a launch, and the column's output offset in a row.
__device__ __forceinline__ int col_q(const TimeArgs& a, int c) {
  if (a.local) return a.c0 + c;
  const int o = a.c0 + c, yo = o / a.nxh, kx = o - yo * a.nxh;
  return ((yo + a.ny - a.ny / 2) % a.ny) * a.nxh + kx;
}
__device__ __forceinline__ size_t col_out(const TimeArgs& a, int c) {
  return static_cast<size_t>(a.local ? c : a.c0 + c);
}

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
// Mixed-radix row kernel (one wave per frame pair, nx <= 768, Tb 4, 8 or
// 16); launch_rows uses it where rows_mr_ok holds.
bool rows_mr_ok(const RowArgs& a, int precision);
hipError_t launch_rows_mr(const RowArgs& a, int precision, hipStream_t s);
// Mixed-radix column kernel (one wave per frame of the block, ny <= 768);
// launch_cols uses it where cols_mr_ok holds.
bool cols_mr_ok(const ColArgs& a, int precision);
hipError_t launch_cols_mr(const ColArgs& a, int precision, hipStream_t s);
hipError_t launch_fft_test(void* data, int n, int batch, const FftDesc& d, const void* tw,
                           int precision, hipStream_t s);

// n bytes from device memory to pinned host memory addressed by the device
// (hipHostGetDevicePointer), by a kernel; n and both addresses multiples of 16.
hipError_t launch_copy_out(void* dst, const void* src, uint64_t n, hipStream_t s);

// FP64 edge correction of the compensated path (fddm_edge.hip): frames
// [0, E) and [L - E, L) of an (L, frame_bytes) stack copied to dst, and the
// lags lags[i0 .. nl) (n = L - dt <= E pairs) recomputed by direct sums from
// the FP64 spectra S64[q][c] (pitch lp) of those frames into the y-shifted
// output.
hipError_t launch_gather_edges(const void* img, void* dst, uint64_t frame_bytes, uint64_t L, uint64_t E,
                               hipStream_t s);
hipError_t launch_edge_fix(const TimeArgs& a, const void* S64, uint64_t lp, int E, int i0, int out_dtype,
                           hipStream_t s);
constexpr int kEdgeFrames = 8;
// Lags of the q marked in flags whose (L - dt) D < kWkFlag L VAR and
// L - dt > 32 recomputed by direct differences of the float series S[q][t]
// (pitch Lp) in double, written to the y-shifted output.
constexpr double kWkFlag = 2e-3;
hipError_t launch_wk_fix(const TimeArgs& a, int out_dtype, hipStream_t s);

// Time stage for series longer than the LDS-resident plans allow
// (fddm_time_long.hip): N = 2^logn in (2^12, 2^long_max_log] as a four-step
// transform, one workgroup per q at a time, each with a scratch slot of
// long_slot_bytes in `scratch` (slots slots).  long_twiddles: the table the
// kernel reads from a.twt (complex, converted to the compute type).
int long_max_log(int precision);
uint64_t long_slot_bytes(int logn, int precision, uint64_t L);
std::vector<double> long_twiddles(int logn);
hipError_t launch_time_long(const TimeArgs& a, int logn, int precision, int out_dtype, void* scratch,
                            uint64_t slot_bytes, int slots, hipStream_t s);

// Elements-per-thread template bucket for a transform of `work` points
// cooperatively done by 256 threads; 0 if too large.
int fft_bucket(long long work);

}  // namespace fddm
