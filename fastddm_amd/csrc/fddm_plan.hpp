// Library-internal plan entry points shared by the C ABI, the request server
// and the batch executor.
#pragma once

#include <cstddef>
#include <cstdint>

#include "fddm_hip.h"

namespace fddm {

// fddm_plan_create with an explicit frames-per-block (0: default choice).
int plan_create_ex(fddm_plan** out, int mode, int dtype, uint64_t L, uint64_t H, uint64_t W, const uint32_t* lags,
                   uint64_t nl, uint64_t nx, uint64_t ny, uint64_t nt, const void* window, int window_f32,
                   int precision, int out_dtype, int frames_per_block, char* err, size_t errlen);
// Frames per block chosen for a plan.
int plan_frames_per_block(const fddm_plan* p);
// Frees the azimuthal-average scratch kept for device dev.
void azavg_release_scratch(int dev);

}  // namespace fddm
