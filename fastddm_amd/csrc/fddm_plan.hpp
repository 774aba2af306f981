// Library-internal plan entry points shared by the C ABI, the request server
// and the batch executor.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

#include "fddm_hip.h"

namespace fddm {

// fddm_plan_create with an explicit frames-per-block (0: default choice).
int plan_create_ex(fddm_plan** out, int mode, int dtype, uint64_t L, uint64_t H, uint64_t W, const uint32_t* lags,
                   uint64_t nl, uint64_t nx, uint64_t ny, uint64_t nt, const void* window, int window_f32,
                   int precision, int out_dtype, int frames_per_block, char* err, size_t errlen);
// Frames per block chosen for a plan.
int plan_frames_per_block(const fddm_plan* p);
// Spatial kernel forms: whether the mixed-radix row / column kernels run for
// a plan (else the power-of-two or generic ones), and a switch to the generic
// workgroup kernels where they would (the batch autotuner times both).
void plan_spatial_forms(const fddm_plan* p, int* rows_mr, int* cols_mr);
void plan_set_generic_spatial(fddm_plan* p, int rows_generic, int cols_generic);
// fddm_plan_execute as many shorter launches (spatial stage over chunks of
// frames_per frames, time stage over chunks of cols_per output columns; 0 =
// no split), bit-identical to it.  Used by the request server so that large
// requests leave gaps for the kernels of small ones.
int plan_execute_pieces(fddm_plan* p, const void* d_img, void* d_out, void* ws, hipStream_t s, uint64_t frames_per,
                        int cols_per, char* err, size_t errlen);
// Spectra bytes of a plan (q-major, all frames): the size of its work.
uint64_t plan_spectra_bytes(const fddm_plan* p);

// Pieces of an execute for callers that bring the frames in themselves (the
// streamed ingest): sizes of the plan, the row and column passes of frames
// [t0, t0 + frames) held at img (img0: frame 0, subtracted on the
// compensated paths) into the workspace's spectra, and the remaining stages
// once every frame has passed (frame0: frame 0; edge_in: the edge child's
// frames, the first and last kEdgeFrames frames).
struct PlanInfo {
  uint64_t L;
  int Tb;
  uint64_t in_bytes, ws_bytes, out_bytes, edge_frames;
};
PlanInfo plan_info(const fddm_plan* p);
int plan_spatial_frames(fddm_plan* p, const void* img, const void* img0, uint64_t t0, uint64_t frames, void* ws,
                        hipStream_t s, char* err, size_t errlen);
int plan_finish_from_spectra(fddm_plan* p, const void* frame0, const void* edge_in, void* ws, void* d_out,
                             hipStream_t s, char* err, size_t errlen);
// memcpy split over up to 16 host threads (4 MiB or more each): fills and
// drains pinned staging buffers at host-memory speed, and first-touches a
// fresh destination from several threads.
void parallel_memcpy(void* dst, const void* src, size_t n);
// Device -> pinned host (a hipHostMalloc buffer), asynchronous on s.
hipError_t copy_to_pinned(void* pin, const void* d_src, size_t n, hipStream_t s);
// Device -> pageable host, synchronous, through the device's pinned staging
// pair from 8 MiB up (device dev current).
int copy_to_host(int dev, void* h_dst, const void* d_src, size_t n, char* err, size_t errlen);
// Frees the device buffers the streamed ingest keeps for device dev (unless
// an ingest is using them); its pinned chunk buffers stay.
void ingest_release_pool(int dev);
// Frees the azimuthal-average scratch kept for device dev.
void azavg_release_scratch(int dev);

}  // namespace fddm
