/*
 * fastddm_amd -- C ABI of the MI355X (gfx950) image structure function engine
 * (libfddm_hip.so).
 *
 * Every entry point returns 0 on success and a non-zero FDDM_ERR_* code on
 * failure, with a NUL-terminated message written to err[0..errlen) (err may be
 * NULL).  No entry point exits or aborts the process.  Device memory handed to
 * the library is never freed by it; host memory is only read (inputs) or
 * written (outputs).  Calls are synchronous unless a stream is passed.
 *
 * Reference interfaces replaced (paths in somexlab/fastddm):
 *   - ddm_fft_cuda<T> / ddm_diff_cuda<T>   src/cuda/ddm_cuda.h:37-58,
 *     bound as _core_cuda.ddm_fft_cuda / ddm_diff_cuda in
 *     src/python/bindings/ddm_cuda_bindings.h:36-55 (nine input dtypes);
 *   - device utilities get_num_devices / set_device / get_device /
 *     get_free_device_memory  src/python/bindings/ddm_cuda_bindings.h:26-31,
 *     src/cuda/gpu_utils.cu:20-116.
 */
#ifndef FDDM_HIP_H
#define FDDM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Input pixel types: the nine overloads of ddm_cuda_bindings.h:36-55. */
enum {
    FDDM_U8 = 0,
    FDDM_I16 = 1,
    FDDM_U16 = 2,
    FDDM_I32 = 3,
    FDDM_U32 = 4,
    FDDM_I64 = 5,
    FDDM_U64 = 6,
    FDDM_F32 = 7,
    FDDM_F64 = 8
};

/* Compute precision.  FP64: all stages in double (the reference GPU core
 * computes in double, src/cuda/memchk_gpu.cu:156,290,397).  FP32: float
 * spectra, FFTs and sums.  FP32C: float spectra and FFTs, frame-0 and
 * temporal-mean subtraction, double prefix sums and combination. */
enum { FDDM_PREC_FP64 = 0, FDDM_PREC_FP32 = 1, FDDM_PREC_FP32C = 2 };

/* Output element type: the reference's Scalar / DTYPE (float64 by default,
 * float32 with SINGLE_PRECISION; src/python/_config.py.in:16). */
enum { FDDM_OUT_F64 = 0, FDDM_OUT_F32 = 1 };

enum { FDDM_MODE_DIFF = 0, FDDM_MODE_FFT = 1 };

enum {
    FDDM_OK = 0,
    FDDM_ERR_ARG = 1,      /* invalid argument */
    FDDM_ERR_SIZE = 2,     /* size outside what the kernels support */
    FDDM_ERR_MEMORY = 3,   /* not enough device memory */
    FDDM_ERR_HIP = 4,      /* HIP runtime error */
    FDDM_ERR_DEVICE = 5    /* bad or missing device */
};

/* ---- device utilities (ddm_cuda_bindings.h:26-31; gpu_utils.cu:20-116) ---- */
int fddm_get_num_devices(int *count, char *err, size_t errlen);
/* Out-of-range ids fail like set_device (gpu_utils.cu:40-60). */
int fddm_set_device(int gpu_id, char *err, size_t errlen);
int fddm_get_device(int *gpu_id, char *err, size_t errlen);
/* Free device memory in bytes (get_free_device_memory, gpu_utils.cu:83-116). */
int fddm_get_free_device_memory(uint64_t *free_bytes, char *err, size_t errlen);
/* Available host memory in bytes, MemAvailable of /proc/meminfo
 * (get_free_host_memory, src/cuda/memchk_gpu.cc:35-87). */
int fddm_get_free_host_memory(uint64_t *free_bytes, char *err, size_t errlen);
/* Host memory a fddm_ddm_fft / fddm_ddm_diff call needs besides its input:
 * the (num_lags+2, ny, nx/2+1) result in the output dtype plus the two pinned
 * staging buffers (get_host_memory_fft / get_host_memory_diff,
 * memchk_gpu.cc:89-140, whose estimate also holds the host-staged spectra
 * of the CUDA core).  mode: FDDM_MODE_FFT or FDDM_MODE_DIFF. */
int fddm_get_host_memory(int mode, uint64_t nx, uint64_t ny, uint64_t length, uint64_t num_lags,
                         int out_dtype, uint64_t *bytes, char *err, size_t errlen);
/* *ok = 1 when 0.95 x the available host memory exceeds that need
 * (check_host_memory_fft_py / check_host_memory_diff_py, memchk_gpu.cc:142-199). */
int fddm_check_host_memory(int mode, uint64_t nx, uint64_t ny, uint64_t length, uint64_t num_lags,
                           int out_dtype, int *ok, char *err, size_t errlen);

/* ---- whole path, host buffers in and out ----
 * Replaces ddm_fft_cuda<T> (ddm_cuda.h:52-58, ddm_cuda.cc:181-285) and
 * ddm_diff_cuda<T> (ddm_cuda.h:37-43, ddm_cuda.cc:25-127).
 *   img      (length, height, width) C-contiguous pixels of type `dtype`
 *   lags     num_lags sorted, unique values in [1, length-1]
 *   nx, ny   padded FFT sizes (>= width, >= height; 7-smooth)
 *   nt       fft mode only: time FFT length (>= 2*length-1, 7-smooth)
 *   window   NULL or height*width values of type `window_f32 ? float : double`
 *   out      (num_lags+2, ny, nx/2+1) of `out_dtype`: D slices, power
 *            spectrum, variance, fftshifted along y (helper_ddm.cc:41-103).
 * Uses the current device (fddm_set_device). */
int fddm_ddm_fft(const void *img, int dtype, uint64_t length, uint64_t height, uint64_t width,
                 const uint32_t *lags, uint64_t num_lags, uint64_t nx, uint64_t ny, uint64_t nt,
                 const void *window, int window_f32, int precision, int out_dtype, void *out,
                 char *err, size_t errlen);
int fddm_ddm_diff(const void *img, int dtype, uint64_t length, uint64_t height, uint64_t width,
                  const uint32_t *lags, uint64_t num_lags, uint64_t nx, uint64_t ny,
                  const void *window, int window_f32, int precision, int out_dtype, void *out,
                  char *err, size_t errlen);
/* Drops the device buffers cached by fddm_ddm_fft / fddm_ddm_diff, the
 * scratch kept by fddm_azimuthal_average and the device buffers kept by the
 * streamed ingest (when no ingest is using them), on the current device.  The
 * pinned host staging buffers of those calls (2 x 64 MiB each, at least) stay
 * until the process ends: host memory freed and pinned again measured slower
 * for device-to-host copies. */
int fddm_release_cache(char *err, size_t errlen);
/* Device bytes of input, output and workspace buffers that fddm_ddm_fft /
 * fddm_ddm_diff keep on the current device for the next call (a negative
 * value restores the default, 16 GiB; 0 keeps none).  A call whose buffers
 * exceed it frees them before returning, as the reference frees its device
 * buffers at the end of every call (ddm_cuda.cu:250-260, :524-529,
 * :917-922); a lower
 * limit frees what is cached now.  No counterpart in the reference (which
 * keeps nothing between calls). */
int fddm_set_cache_limit(int64_t bytes, char *err, size_t errlen);
/* Device bytes of input, output and workspace buffers cached on the current
 * device now (the runtime may keep freed device memory in its own pools, so
 * free-memory queries need not show a release at once). */
int fddm_get_cached_bytes(uint64_t *bytes, char *err, size_t errlen);
/* Device memory fddm_ddm_fft / fddm_ddm_diff may use on the current device
 * (0: 95 % of free memory, the default).  A call whose input, workspace and
 * output exceed it runs in pieces -- the spatial stage over chunks of frames,
 * the time stage over chunks of q, with the q-major spectra in device memory
 * or, when they do not fit either, in host memory -- and gives the same
 * result bit for bit.  This replaces the reference's fixed-point planner of
 * loop counts over frames and q (src/cuda/memchk_gpu.cu:141-702, driven by
 * free memory x 0.9) and lets a caller cap the footprint. */
int fddm_set_memory_budget(uint64_t bytes, char *err, size_t errlen);
/* How the last fddm_ddm_fft / fddm_ddm_diff call on the current device ran:
 * stats[3] = {0 one launch per stage | 1 pieces, spectra in device memory |
 * 2 pieces, spectra in host memory, frame chunks, q chunks}. */
int fddm_get_last_call_chunks(uint64_t *stats, char *err, size_t errlen);

/* ---- streamed ingest: frames pushed piece by piece ----
 * For callers that read or decode the stack while it is analysed (a TIFF
 * reader, a camera): replaces the reference's read-all-then-analyse flow
 * (src/python/utils.py:23-257 loading the whole stack, then the frame-batch
 * H2D of compute_fft2, src/cuda/ddm_cuda.cu:40-261).  Pushed frames are
 * staged in pinned host memory in chunks of whole frame blocks; each chunk's
 * DMA (copy stream) and row and column passes (compute stream) overlap the
 * caller's next pushes.  begin takes the arguments of fddm_ddm_fft /
 * fddm_ddm_diff (mode FDDM_MODE_FFT or FDDM_MODE_DIFF; nt ignored in diff
 * mode); push takes `count` frames of height*width pixels in stack order;
 * finish, once every frame is pushed, runs the remaining stages and writes
 * the (num_lags+2, ny, nx/2+1) result to host memory -- the same values as
 * fddm_ddm_fft / fddm_ddm_diff.  The two pinned chunk buffers and the
 * workspace and result device buffers come from a per-device pool that the
 * next ingest on the device reuses (a concurrent ingest allocates its own);
 * fddm_release_cache frees the pool's device buffers, and device buffers
 * above 16 GiB in total are freed at destroy. */
typedef struct fddm_ingest fddm_ingest;
int fddm_ingest_begin(fddm_ingest **ingest, int mode, int dtype, uint64_t length, uint64_t height, uint64_t width,
                      const uint32_t *lags, uint64_t num_lags, uint64_t nx, uint64_t ny, uint64_t nt,
                      const void *window, int window_f32, int precision, int out_dtype, char *err, size_t errlen);
int fddm_ingest_push(fddm_ingest *ingest, const void *frames, uint64_t count, char *err, size_t errlen);
int fddm_ingest_finish(fddm_ingest *ingest, void *out, char *err, size_t errlen);
int fddm_ingest_destroy(fddm_ingest *ingest);

/* ---- plans: device-resident inputs and outputs ----
 * The phases of ddm_fft_cuda / ddm_diff_cuda (compute_fft2,
 * structure_function_fft/diff, make_shift; ddm_cuda.cu:40-1007) as one
 * reusable plan whose execute runs all stages on one stream. */
typedef struct fddm_plan fddm_plan;
int fddm_plan_create(fddm_plan **plan, int mode, int dtype, uint64_t length, uint64_t height,
                     uint64_t width, const uint32_t *lags, uint64_t num_lags, uint64_t nx,
                     uint64_t ny, uint64_t nt, const void *window, int window_f32, int precision,
                     int out_dtype, char *err, size_t errlen);
int fddm_plan_destroy(fddm_plan *plan);
/* Workspace bytes that fddm_plan_execute needs (spectra and intermediates). */
int fddm_plan_workspace_size(const fddm_plan *plan, uint64_t *bytes);
/* Output bytes: (num_lags+2) * ny * (nx/2+1) * sizeof(out element). */
int fddm_plan_output_size(const fddm_plan *plan, uint64_t *bytes);
/* d_img, d_out, d_workspace: device pointers (d_img 8-byte aligned,
 * d_workspace 16-byte aligned, d_out aligned to its element; hipMalloc gives
 * 256 bytes); stream: hipStream_t or NULL. */
int fddm_plan_execute(fddm_plan *plan, const void *d_img, void *d_out, void *d_workspace,
                      void *stream, char *err, size_t errlen);
/* fddm_plan_execute without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  It
 * allocates nothing: the workspace is a static device array of 256 MiB per
 * GPU (held from library load to process exit); plans whose
 * fddm_plan_workspace_size exceeds it are rejected with FDDM_ERR_SIZE.  All
 * calls on one GPU share that array and it keeps the last call's contents,
 * so calls that overlap on one GPU (other streams, other threads) race. */
int fddm_plan_execute_static(fddm_plan *plan, const void *d_img, void *d_out, void *stream, char *err,
                             size_t errlen);
/* Per-stage timing with HIP events on the execute stream: begin reserves
 * events for up to max_steps executes; end waits for them and writes the
 * summed milliseconds of the stages {frame-0 spectra, row pass, column pass,
 * time/diff kernel, FP64 edge correction (FP32C only)} to stage_ms[5] and the
 * number of timed executes. */
int fddm_plan_timing_begin(fddm_plan *plan, int max_steps, char *err, size_t errlen);
int fddm_plan_timing_end(fddm_plan *plan, double *stage_ms, int *steps, char *err, size_t errlen);
/* Test entry: spatial stages only; copies the q-major spectra (ny*(nx/2+1),
 * length) to host as interleaved double complex (unshifted, q = ky*nxh+kx). */
int fddm_plan_spectra(fddm_plan *plan, const void *d_img, void *d_workspace, double *host_spectra,
                      char *err, size_t errlen);

/* ---- device memory and streams (so callers need no other GPU runtime) ---- */
int fddm_malloc(void **ptr, uint64_t bytes, char *err, size_t errlen);
int fddm_free(void *ptr);
int fddm_memcpy_htod(void *dst, const void *src, uint64_t bytes, char *err, size_t errlen);
int fddm_memcpy_dtoh(void *dst, const void *src, uint64_t bytes, char *err, size_t errlen);
int fddm_memcpy_dtod(void *dst, const void *src, uint64_t bytes, char *err, size_t errlen);
int fddm_stream_create(void **stream, char *err, size_t errlen);
int fddm_stream_destroy(void *stream);
int fddm_stream_synchronize(void *stream, char *err, size_t errlen);

/* ---- request server (continuous batching on one GPU) ----
 * No counterpart in the reference, whose cores run one analysis per blocking
 * call; this is the serving layer around fddm_plan_*.  A server owns one
 * device arena of hbm_bytes (0: 85 % of free memory) and max_slots worker
 * streams on the device current at creation.  A submitted request is
 * admitted as soon as a slot is free and its footprint (plan workspace +
 * output [+ a device copy of a host input]) fits in the arena; admitted
 * requests run concurrently.  Results stay in the arena until released.
 * img_on_device: img is a device pointer that stays valid until the request
 * finishes; otherwise a host pointer read when the request is admitted
 * (inputs of 8 MiB and more through two 32 MiB pinned buffers per slot,
 * allocated at the slot's first such request and freed by destroy).
 * States: 0 pending, 1 running, 2 done, 3 failed.  times[3] = seconds since
 * server start at submit, admission and completion. */
typedef struct fddm_server fddm_server;
int fddm_server_create(fddm_server **server, int max_slots, uint64_t hbm_bytes, char *err,
                       size_t errlen);
/* Fails the requests not yet admitted (fddm_server_next_done reports them),
 * waits for the running ones and for every thread blocked in
 * fddm_server_wait / fddm_server_next_done to return, then frees everything. */
int fddm_server_destroy(fddm_server *server);
/* A pending request older than `seconds` blocks later ones from passing it. */
int fddm_server_set_aging(fddm_server *server, double seconds);
int fddm_server_submit(fddm_server *server, int mode, const void *img, int img_on_device, int dtype,
                       uint64_t length, uint64_t height, uint64_t width, const uint32_t *lags,
                       uint64_t num_lags, uint64_t nx, uint64_t ny, uint64_t nt, const void *window,
                       int window_f32, int precision, int out_dtype, uint64_t *id, char *err,
                       size_t errlen);
int fddm_server_status(fddm_server *server, uint64_t id, int *state, double *times, char *err,
                       size_t errlen);
/* timeout_s < 0 waits without limit. */
int fddm_server_wait(fddm_server *server, uint64_t id, double timeout_s, char *err, size_t errlen);
/* Id of the next request that finished (done or failed), in completion
 * order, each reported once (released ids are not); *id = 0 when timeout_s
 * (>= 0) passes first, or with an error once the server is shutting down. */
int fddm_server_next_done(fddm_server *server, double timeout_s, uint64_t *id, char *err,
                          size_t errlen);
/* Device pointer and size of a finished result (valid until release). */
int fddm_server_result(fddm_server *server, uint64_t id, void **d_out, uint64_t *bytes, char *err,
                       size_t errlen);
/* Copies a finished result to host memory; from 8 MiB up through the
 * device's pinned staging pair shared with the host-buffer calls (a
 * host-buffer call on the same device waits for it, and it for them). */
int fddm_server_fetch(fddm_server *server, uint64_t id, void *host_out, char *err, size_t errlen);
int fddm_server_release(fddm_server *server, uint64_t id);
/* stats[7] = {pending, running, completed, arena bytes, arena used, arena
 * peak, cached plans}. */
int fddm_server_stats(fddm_server *server, uint64_t *stats);
/* Load generation: fills count uint8/uint16 elements of a device buffer with
 * uniform integers in [lo, hi) from a counter-based hash of (seed, index). */
int fddm_fill_uniform(void *d, uint64_t count, int dtype, uint64_t seed, uint32_t lo, uint32_t hi,
                      void *stream, char *err, size_t errlen);

/* ---- batches: many stacks of uneven size in one graph launch ----
 * No counterpart in the reference (one stack per call).  Each item is one
 * analysis with device-resident input and output (output of
 * fddm_plan_output_size bytes for the same arguments).  Creation builds one
 * plan per item -- with autotune != 0 the frames-per-block of each image
 * shape is chosen by timing the candidates' spatial stages on the first
 * frames of that shape's longest item (cached per shape for the process)
 * -- allocates one workspace for all
 * items and captures every item's stage chain into one hipGraph on parallel
 * branches.  fddm_batch_launch replays the graph on `stream`;
 * fddm_batch_launch_eager runs the same plans one after another without the
 * graph.  Inputs are read and outputs written at each launch. */
typedef struct {
    int mode, dtype;
    uint64_t length, height, width;
    const uint32_t *lags;
    uint64_t num_lags;
    uint64_t nx, ny, nt;
    const void *window;
    int window_f32;
    int precision, out_dtype;
    const void *d_img;
    void *d_out;
} fddm_batch_item;
typedef struct fddm_batch fddm_batch;
int fddm_batch_create(fddm_batch **batch, const fddm_batch_item *items, int num_items, int autotune,
                      char *err, size_t errlen);
int fddm_batch_launch(fddm_batch *batch, void *stream, char *err, size_t errlen);
int fddm_batch_launch_eager(fddm_batch *batch, void *stream, char *err, size_t errlen);
/* Frames per block of each item's plan (tb[num_items]). */
int fddm_batch_frames_per_block(const fddm_batch *batch, int *tb);
/* Spatial kernels of each item's plan: rows_mr[i] (cols_mr[i]) is 1 where the
 * mixed-radix row (column) kernel runs, 0 where the power-of-two or the
 * generic one does (with autotune, the generic form is chosen where it timed
 * faster).  No counterpart in the reference. */
int fddm_batch_spatial_kernels(const fddm_batch *batch, int *rows_mr, int *cols_mr);
int fddm_batch_destroy(fddm_batch *batch);

/* ---- azimuthal average of a device-resident result ----
 * Replaces the reduction loops of azimuthal_average_array
 * (src/python/azimuthalaverage.py:420-480; caller azimuthal_average :481-566).
 * d_data: (rows, ny, nxh) float64 (data_f32 = 0) or float32 values in device
 * memory, e.g. a plan output.  Per pixel q = y*nxh + x (host arrays):
 * bin_index[q] in [0, nbins), non-decreasing along x; w[q] = counts*weights
 * (0 excludes the pixel); w2[q] = counts*weights^2 (read when eval_err).
 * Writes avg_out[nbins*rows] = sum w x / sum w and, when eval_err,
 * err_out[nbins*rows] = sqrt(sum w (x - avg)^2 / (sum w - sum w2 / sum w)),
 * both [bin][row] in float64 on the host; NaN for bins with no weight.
 * nbins <= 1024.  stream: hipStream_t or NULL. */
int fddm_azimuthal_average(const void *d_data, int data_f32, uint64_t rows, uint64_t ny,
                           uint64_t nxh, const int32_t *bin_index, const double *w,
                           const double *w2, int nbins, int eval_err, double *avg_out,
                           double *err_out, void *stream, char *err, size_t errlen);

/* ---- FFT engine test entry: batch forward FFTs of length n, in place on
 * host data (interleaved double complex); precision FP64 or FP32. */
int fddm_test_fft(double *data, uint64_t n, uint64_t batch, int precision, char *err,
                  size_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* FDDM_HIP_H */
