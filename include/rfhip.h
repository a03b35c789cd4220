/*
 * rfhip.h — C ABI of the MI355X (gfx950) implementation of gpu-rf-pipeline's hot path.
 *
 * Every function is plain C: pointers, sizes and status codes, no C++ or HIP types. Device
 * pointers are ordinary `void*`/`rf_cf32*` addresses returned by rf_malloc (or any hipMalloc).
 * A stream is an `rf_stream` holding a hipStream_t; NULL means the default (null) stream.
 *
 * Conventions carried over from the reference's operator surface (src/device_ops.hpp:1-106,
 * src/ops.hpp:1-52):
 *   - compute entry points only enqueue work on the given stream; they never synchronise,
 *     never copy to or from the host and, once a stream has been used, never allocate: the
 *     constant tables are compiled into the library, and the one-shot FIR forms that derive a
 *     table from the taps (fft_ola's 4096-point tap spectrum from 226 taps, the MFMA variants'
 *     split taps; fft_ola's 512-point form for 17..225 taps makes its spectrum in registers)
 *     keep it in a 72 KiB workspace block per stream, allocated by the first such call on that
 *     stream (or by rf_stream_create) and freed by rf_stream_destroy;
 *   - the caller owns every input and output buffer; plan handles own their workspaces;
 *   - taps are read from the caller's pointer on every launch (no program-wide tap bank, unlike
 *     the reference's __constant__ c_taps, src/fir.cu:208-209);
 *   - no function throws: failures return an rf_status, and rf_last_error() gives the message
 *     (thread-local), in the reference's exception-text style (src/check.hpp:42-66).
 */
#ifndef RFHIP_H
#define RFHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One complex float32 sample, I then Q: the .cf32 file layout (src/types.hpp:7-12). */
typedef struct rf_cf32 {
    float re;
    float im;
} rf_cf32;

/* A hipStream_t. NULL is the default stream. */
typedef void* rf_stream;

typedef enum rf_status {
    RF_OK = 0,
    RF_INVALID_ARG = 1, /* bad argument: unknown impl, block size, channel count, NULL pointer */
    RF_LIMIT = 2,       /* a documented size limit (shared memory, taps, FFT length, batch) */
    RF_HIP_ERROR = 3,   /* the HIP runtime reported an error */
    RF_ALLOC = 4        /* device or host allocation failed */
} rf_status;

/* Name of a status code ("RF_OK", ...). */
const char* rf_status_string(rf_status status);
/* Message of the last failing call on this thread ("" if none). Replaces the text of the
 * std::runtime_error the reference throws (src/check.hpp:42-66, src/fir.cu:160-166). */
const char* rf_last_error(void);
/* Version of this ABI; bumped whenever a signature changes. */
int rf_abi_version(void);

/* ---------------------------------------------------------------------------------------
 * Runtime utilities. No reference counterpart: the reference calls the CUDA runtime directly
 * (src/device_buffer.hpp:11-76, src/bench.cpp:10-40). They let a binding without a GPU
 * framework own device memory, streams and events through this library alone.
 * ------------------------------------------------------------------------------------- */
rf_status rf_device_count(int* count);
rf_status rf_device_name(char* buffer, int length);
/* Total global memory of the current device, in bytes. */
rf_status rf_device_memory(size_t* bytes);
rf_status rf_malloc(void** ptr, size_t bytes);
rf_status rf_free(void* ptr);
rf_status rf_host_alloc(void** ptr, size_t bytes); /* page-locked host memory */
rf_status rf_host_free(void* ptr);
/* Page-locked host memory mapped into the device's address space (zero-copy): *host is the
 * CPU address, *device the address kernels use. Free with rf_host_free(*host). */
rf_status rf_host_alloc_mapped(void** host, void** device, size_t bytes);
/* Asynchronous on `stream`; pageable host memory makes them effectively synchronous.
 * rf_memcpy_d2d runs the library's streaming copy kernel (16-byte accesses). */
rf_status rf_memcpy_h2d(void* dst, const void* src, size_t bytes, rf_stream stream);
rf_status rf_memcpy_d2h(void* dst, const void* src, size_t bytes, rf_stream stream);
rf_status rf_memcpy_d2d(void* dst, const void* src, size_t bytes, rf_stream stream);
/* A fill kernel (not hipMemsetAsync), so it can be recorded in a graph. */
rf_status rf_memset(void* dst, int value, size_t bytes, rf_stream stream);
rf_status rf_stream_create(rf_stream* stream);
rf_status rf_stream_destroy(rf_stream stream);
rf_status rf_stream_synchronize(rf_stream stream);
rf_status rf_device_synchronize(void);
/* Timing events (hipEvent_t behind a void*). */
rf_status rf_event_create(void** event);
rf_status rf_event_destroy(void* event);
rf_status rf_event_record(void* event, rf_stream stream);
rf_status rf_event_synchronize(void* event);
rf_status rf_event_elapsed_ms(void* start, void* stop, float* ms);
/* Work enqueued on `stream` after this call waits for `event`'s last record (hipStreamWaitEvent),
 * so that several streams can pipeline one job (the streaming runner). */
rf_status rf_stream_wait_event(rf_stream stream, void* event);
/* HIP graphs by stream capture: the work enqueued on `stream` (a created stream, not NULL)
 * between begin and end is recorded instead of run; end instantiates it (*graph is a
 * hipGraphExec_t). rf_graph_launch replays it on a stream, rf_graph_destroy frees it. Every
 * rf_fir / rf_fir_batched / rf_fir_plan_run call only enqueues kernels and can be captured;
 * rf_fir and rf_fir_batched with fft_ola or an MFMA variant keep their tap spectrum / split
 * taps in the capture stream's workspace block, so a graph holding such a call is replayed on
 * its capture stream (or on another stream while no eager call on the capture stream runs). */
rf_status rf_graph_begin_capture(rf_stream stream);
rf_status rf_graph_end_capture(rf_stream stream, void** graph);
rf_status rf_graph_launch(void* graph, rf_stream stream);
rf_status rf_graph_destroy(void* graph);

/* ---------------------------------------------------------------------------------------
 * FIR:  y[n] = sum_{k=0}^{T-1} taps[k] * x[n-k],  x[m] = 0 for m < 0,  n < num_samples.
 * Definition: src/fir.cu:1-5, tools/reference.py:15-17,103-111.
 * ------------------------------------------------------------------------------------- */
enum {
    RF_FIR_NAIVE = 0,       /* one lane per output, global memory (src/fir.cu:20-43) */
    RF_FIR_TILED = 1,       /* LDS tile with a T-1 halo (src/fir.cu:100-144) */
    RF_FIR_TILED_CONST = 2, /* tiled, taps through the scalar (constant) cache (src/fir.cu:213-254) */
    RF_FIR_REGISTER = 3,    /* R outputs per lane from a register window (src/fir.cu:337-409) */
    RF_FIR_FFT_OLA = 4,     /* overlap-save with the library's FFT: 512-point segments per wavefront
                               for 17..225 taps, 4096-point for 226..2049, the register kernel
                               otherwise; no reference counterpart */
    RF_FIR_MFMA_F16X3 = 5,  /* Toeplitz GEMM, FP16 matrix cores, hi/lo split; no counterpart */
    RF_FIR_MFMA_BF16X3 = 6, /* the same in BF16 (reduced precision, error ~2^-16 relative) */
    RF_FIR_IMPL_COUNT = 7
};
#define RF_DEFAULT_FIR_BLOCK 256 /* kDefaultFirBlock, src/ops.hpp:35 */
#define RF_MAX_FIR_BATCH 8       /* kMaxFirBatch, src/device_ops.hpp:46 */
#define RF_MAX_CONST_TAPS 8192   /* kMaxConstTaps, src/fir.cu:208 */

/* The variant table, in the reference's order (src/fir.cu:725-734); entry 0 is the default. */
int rf_fir_impl_count(void);
const char* rf_fir_impl_name(int impl);
const char* rf_fir_impl_summary(int impl);
/* parse_fir_impl (src/fused.cu:76-90): RF_INVALID_ARG for an unknown name. */
rf_status rf_fir_parse_impl(const char* name, int* impl);

/* fir_launch / fir_device (src/fir.cu:513-556) with the block size of the host wrappers
 * (src/fir.cu:62-85,170-200,449-485): `block` threads per workgroup, a multiple of 32 up to
 * 1024, or 0 for RF_DEFAULT_FIR_BLOCK. Tiled variants return RF_LIMIT when the halo does not
 * fit in one workgroup's LDS; tiled_const and register return RF_LIMIT above
 * RF_MAX_CONST_TAPS taps, as the reference does. num_samples == 0 enqueues nothing. */
rf_status rf_fir(int impl, const rf_cf32* x, const float* taps, rf_cf32* y, int num_samples,
                 int num_taps, int block, rf_stream stream);

/* rf_fir that allocates nothing, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in
 * the reference. Same arguments, errors and bits as rf_fir; fft_ola's tap spectrum and the MFMA
 * variants' split taps go to a fixed 72 KiB scratch array compiled into the library (one per
 * GPU, kept until the process ends) instead of a stream's workspace block; RF_LIMIT if a filter
 * needed more. Every rf_fir_noalloc call on a GPU shares that array, so calls that overlap on
 * one GPU (different streams, or a stream and a graph replay) race: serialise them. */
rf_status rf_fir_noalloc(int impl, const rf_cf32* x, const float* taps, rf_cf32* y,
                         int num_samples, int num_taps, int block, rf_stream stream);

/* fir_batched_device (src/fused.cu:92-109): `count` (1..RF_MAX_FIR_BATCH) channels sharing one
 * set of taps in a single launch. x, y and num_samples are HOST arrays of `count` entries
 * holding device pointers and lengths. Results equal rf_fir with the same impl per channel. */
rf_status rf_fir_batched(int impl, const rf_cf32* const* x, rf_cf32* const* y,
                         const int* num_samples, int count, const float* taps, int num_taps,
                         rf_stream stream);

/* FIR autotuner. No counterpart in the reference (which fixes block 256 and R by tap count,
 * src/fir.cu:449-485): times every variant allowed by impl_mask (bit i = RF_FIR_* value i) on
 * the caller's buffers (y is overwritten) with its tile shapes, lanes per workgroup
 * {64, 128, 256, 512} x outputs per lane {4, 8, 16} for `register`, {128, 256, 512} for the
 * other direct variants, the overlap-save transform length {512, 4096} for fft_ola (each where
 * the tap count allows it), and reports the fastest; device events on `stream`, which it
 * synchronises. The direct variants give identical bits for every shape. */
typedef struct rf_fir_config {
    int32_t impl;             /* RF_FIR_* */
    int32_t block;            /* lanes per workgroup of the direct variants */
    int32_t outputs_per_lane; /* register variant: 4, 8 or 16; 0 = the library's choice */
    float ms;                 /* measured time of one call */
    int32_t transform_points; /* fft_ola: 512 (2..449 taps) or 4096 (2..2049 taps) points per
                                 overlap-save segment; 0 = the library's choice for the tap
                                 count (rf_fir's bits); 0 for every other variant */
} rf_fir_config;
rf_status rf_fir_tune(const rf_cf32* x, const float* taps, rf_cf32* y, int num_samples,
                      int num_taps, unsigned impl_mask, rf_stream stream, rf_fir_config* best);
/* rf_fir with a tuned configuration. RF_INVALID_ARG for a transform length other than 0, 512
 * or 4096 or one given to a variant other than fft_ola; RF_LIMIT when the tap count is outside
 * the chosen length's range. */
rf_status rf_fir_run_config(const rf_fir_config* config, const rf_cf32* x, const float* taps,
                            rf_cf32* y, int num_samples, int num_taps, rf_stream stream);

/* Prepared filter. No counterpart in the reference, whose host wrappers upload the taps on
 * every call (src/fir.cu:62-85): for one filter applied to many blocks, create copies the
 * device taps into memory the plan owns and, for fft_ola from 226 taps, makes their 4096-point
 * spectrum once
 * (then synchronises `stream`), so each run launches only the filtering kernel. Arguments and
 * errors as rf_fir (impl, taps, num_taps, block checked at create); a run gives the same bits
 * as rf_fir with that impl and block. Runs on one plan may overlap on different streams. */
typedef struct rf_fir_plan rf_fir_plan;
rf_status rf_fir_plan_create(int impl, const float* taps, int num_taps, int block,
                             rf_stream stream, rf_fir_plan** plan);
rf_status rf_fir_plan_run(const rf_fir_plan* plan, const rf_cf32* x, rf_cf32* y,
                          int num_samples, rf_stream stream);
rf_status rf_fir_plan_destroy(rf_fir_plan* plan);

/* Persistent batched FIR: any number of captures of uneven length sharing one filter, all
 * filtered by ONE launch of a persistent kernel that walks {capture, segment} work items
 * (XCD-major). No reference counterpart beyond fir_batched_device (src/fused.cu:92-109), which
 * takes at most 8 channels as grid.y. impl is RF_FIR_FFT_OLA (overlap-save segments of 512
 * points for 17..225 taps and 4096 for 226..2049, the register kernel's blocks otherwise) or
 * RF_FIR_REGISTER; each capture's
 * output is bit-identical to rf_fir with that impl. create copies the DEVICE taps, makes the
 * tap spectrum and the item map (synchronises `stream`); bind takes HOST arrays of `count`
 * device pointers (synchronous); run only enqueues one launch (no allocation: capturable). */
typedef struct rf_fir_batch rf_fir_batch;
rf_status rf_fir_batch_create(int impl, const float* taps, int num_taps, int count,
                              const int* num_samples, rf_stream stream, rf_fir_batch** batch);
rf_status rf_fir_batch_bind(rf_fir_batch* batch, const rf_cf32* const* x, rf_cf32* const* y);
rf_status rf_fir_batch_run(const rf_fir_batch* batch, rf_stream stream);
rf_status rf_fir_batch_destroy(rf_fir_batch* batch);

/* ---------------------------------------------------------------------------------------
 * Magnitude: mag[n] = |x[n]| (src/magnitude.cu:17-39, tools/reference.py:98-100).
 * ------------------------------------------------------------------------------------- */
rf_status rf_magnitude(const rf_cf32* x, float* mag, int num_samples, rf_stream stream);

/* ---------------------------------------------------------------------------------------
 * FFT plans: `batch` transforms of `nfft` points stored back to back; replaces FftPlan over
 * cufftPlanMany (src/fft_plan.hpp:11-37). Unnormalised in both directions, like cuFFT.
 * Powers of two run the library's LDS Stockham FFT (one workgroup up to 4096 points, a
 * two-pass four-step transform up to 2^24); other lengths run an exact DFT.
 * ------------------------------------------------------------------------------------- */
typedef struct rf_fft_plan rf_fft_plan;
#define RF_FFT_FORWARD (-1) /* CUFFT_FORWARD: exp(-2 pi i n k / N) */
#define RF_FFT_INVERSE (1)  /* CUFFT_INVERSE: exp(+2 pi i n k / N) */
rf_status rf_fft_plan_create(int nfft, int batch, rf_fft_plan** plan);
rf_status rf_fft_plan_destroy(rf_fft_plan* plan);
/* GPU memory the plan holds (tables and scratch), like cufftGetSize (src/xcorr.cu:331). */
rf_status rf_fft_plan_bytes(const rf_fft_plan* plan, size_t* bytes);
/* cufftExecC2C: in == out is allowed (in place). */
rf_status rf_fft_exec(const rf_fft_plan* plan, const rf_cf32* in, rf_cf32* out, int direction,
                      rf_stream stream);

/* spectrum_device (src/spectrum.cu:11-17): the plan's batched forward FFT from `in` (already
 * zero-padded to batch * nfft samples) to `out`. */
rf_status rf_spectrum(const rf_fft_plan* plan, const rf_cf32* in, rf_cf32* out,
                      rf_stream stream);
/* Spectrum then magnitude in one pass (no reference counterpart; the fusion deferred in
 * docs/ROADMAP.md:57-59 of the reference): mag (DEVICE, nfft * batch floats) = |rf_spectrum|,
 * sqrtf(re^2 + im^2) as rf_magnitude computes it. Plans with 8 <= nfft <= 4096 run as one
 * kernel whose store writes the magnitude; for other lengths the plan's last pass (the row
 * pass of the four-step transform, the DFT, the 1/2/4-point transform) stores the magnitude.
 * No allocation on any path; calls that share a four-step plan on two streams race on the
 * plan's scratch, as rf_fft_exec does. */
rf_status rf_spectrum_magnitude(const rf_fft_plan* plan, const rf_cf32* in, float* mag,
                                rf_stream stream);

/* ---------------------------------------------------------------------------------------
 * Polyphase channelizer (src/channelize.cu:1-72, tools/reference.py:24-30,128-146):
 *   z[t][k] = sum_l h[l] x[tM - l] exp(+2 pi i k l / M),  t < ceil(N/M), k < M.
 * ------------------------------------------------------------------------------------- */
/* channelize_frames (src/channelize.cu:51-54): ceil(num_samples / channels). */
int rf_channelize_frames(int num_samples, int channels);
/* channelize_device (src/channelize.cu:56-72). `plan` is a plan of `channels` points times
 * rf_channelize_frames() frames, or NULL: the library then uses its fused branch-sum + FFT
 * kernel where one exists for `channels` and an internal plan otherwise. */
rf_status rf_channelize(const rf_fft_plan* plan, const rf_cf32* x, const float* taps,
                        rf_cf32* out, int num_samples, int channels, int num_taps,
                        rf_stream stream);
/* The same channelizer with the filterbank's DFT on the matrix cores (no reference
 * counterpart; channelize_device's definition, src/channelize.cu:3-11): FP32 branch sums, then
 * the 64-point DFT as a split-FP16 GEMM (hi*hi + hi*lo + lo*hi, FP32 accumulation; power-of-two
 * scaling per 64 frames). |error| <= 2^-15 * sum_l |h[l] x[tM - l]| per output.
 * channels must be 64 and num_taps <= 1024 (RF_LIMIT otherwise). */
rf_status rf_channelize_mfma(const rf_cf32* x, const float* taps, rf_cf32* out,
                             int num_samples, int channels, int num_taps, rf_stream stream);

/* ---------------------------------------------------------------------------------------
 * Cross-correlation (src/xcorr.cu:1-7, tools/reference.py:19-22,114-125):
 *   r[i] = sum_n b[n + k] conj(a[n]),  k = i - (num_a - 1),  i < num_a + num_b - 1.
 * ------------------------------------------------------------------------------------- */
enum { RF_XCORR_NAIVE = 0, RF_XCORR_TILED = 1, RF_XCORR_FFT = 2, RF_XCORR_IMPL_COUNT = 3 };
int rf_xcorr_impl_count(void);
const char* rf_xcorr_impl_name(int impl);
const char* rf_xcorr_impl_summary(int impl);
rf_status rf_xcorr_parse_impl(const char* name, int* impl);

/* xcorr_naive_device / xcorr_tiled_device (src/xcorr.cu:50-64,174-188). */
rf_status rf_xcorr_direct(int impl, const rf_cf32* a, const rf_cf32* b, rf_cf32* r, int num_a,
                          int num_b, rf_stream stream);

/* class XcorrFft (src/device_ops.hpp:70-88, src/xcorr.cu:315-375): setup-owning FFT
 * cross-correlation for fixed lengths. With `phat`, GCC-PHAT: every frequency of B conj(A) is
 * scaled to unit magnitude (0 where it is 0) before the inverse FFT (src/xcorr.cu:219-233,
 * tools/reference.py:32-36,149-162). FFT length L = next power of two >= num_a + num_b - 1;
 * RF_LIMIT above 2^24. */
typedef struct rf_xcorr_fft rf_xcorr_fft;
rf_status rf_xcorr_fft_create(int num_a, int num_b, int phat, rf_xcorr_fft** handle);
rf_status rf_xcorr_fft_destroy(rf_xcorr_fft* handle);
rf_status rf_xcorr_fft_bytes(const rf_xcorr_fft* handle, size_t* bytes);
rf_status rf_xcorr_fft_length(const rf_xcorr_fft* handle, int* fft_length);
rf_status rf_xcorr_fft_run(const rf_xcorr_fft* handle, const rf_cf32* a, const rf_cf32* b,
                           rf_cf32* r, rf_stream stream);

/* ---------------------------------------------------------------------------------------
 * Peak search: the host loop of rfgpu doa (src/main.cpp:390-401) as a device reduction.
 * index = first k with the largest hypotf(r[k].re, r[k].im) (strict '>', like np.argmax).
 * `result` is DEVICE memory. `frac` is the sub-sample offset in [-0.5, 0.5] of the parabola
 * through |r| at index-1, index, index+1, evaluated in float64 (0 at either end).
 * ------------------------------------------------------------------------------------- */
typedef struct rf_peak_result {
    int32_t index;
    float magnitude;
    double frac;
} rf_peak_result;
rf_status rf_peak(const rf_cf32* r, int n, rf_peak_result* result, rf_stream stream);

/* ---------------------------------------------------------------------------------------
 * Batched GCC-PHAT with the peak search fused in. No single reference function: it does the
 * work of gcc_phat_cuda (src/xcorr.cu:379-399) plus the host peak loop of rfgpu doa
 * (src/main.cpp:390-401) for `count` pairs of uneven length in one pass sequence of
 * persistent kernels (column FFTs, row FFTs + PHAT weighting + inverse row FFTs, inverse
 * column FFTs + peak), optionally replayed as one hipGraph. Pairs whose FFT length is 2^15
 * run in one workgroup each instead (whole GCC-PHAT and peak in registers and LDS).
 * Each pair needs num_a + num_b - 1 > 2048 (FFT length 4096 .. 2^24); RF_LIMIT otherwise.
 * ------------------------------------------------------------------------------------- */
typedef struct rf_gcc_batch rf_gcc_batch;
/* Setup: per-pair FFT shapes, work buffers (2 L values per pair), twiddle tables. */
rf_status rf_gcc_batch_create(int count, const int* num_a, const int* num_b,
                              rf_gcc_batch** handle);
rf_status rf_gcc_batch_destroy(rf_gcc_batch* handle);
rf_status rf_gcc_batch_bytes(const rf_gcc_batch* handle, size_t* bytes);
/* Binds the inputs: a, b are HOST arrays of `count` device pointers; r is NULL (peaks only)
 * or a HOST array of `count` device pointers (each NULL or num_a + num_b - 1 values) that also
 * receive the full GCC-PHAT output. Synchronises the device (setup, not per run). */
rf_status rf_gcc_batch_bind(rf_gcc_batch* handle, const rf_cf32* const* a,
                            const rf_cf32* const* b, rf_cf32* const* r);
/* Enqueues the batch on `stream`. results: DEVICE array of `count` rf_peak_result; index is
 * the output index of pair i's peak (lag = index - (num_a - 1)). */
rf_status rf_gcc_batch_run(rf_gcc_batch* handle, rf_peak_result* results, rf_stream stream);
/* The same as one hipGraph: captured on first use (and again when `results` changes), then
 * replayed with a single launch. */
rf_status rf_gcc_batch_launch_graph(rf_gcc_batch* handle, rf_peak_result* results,
                                    rf_stream stream);
/* The passes run over consecutive groups of pairs whose work buffers (2 L complex values per
 * pair) fit `bytes` (0: one group), so that a group's intermediate data can stay in the
 * 256 MB Infinity Cache between its passes. Default 0 (measured fastest for 256 uneven
 * pairs: more groups add launch tails). */
rf_status rf_gcc_batch_set_group_bytes(rf_gcc_batch* handle, size_t bytes);
rf_status rf_gcc_batch_group_bytes(const rf_gcc_batch* handle, size_t* bytes, int* groups);
/* Autotuner: times the bound batch for 2, 3, 4, 6 and 8 persistent workgroups per CU and group
 * budgets of 0, 48, 96 and 160 MiB (3 runs each after a warm-up, device events on `stream`,
 * which it synchronises), keeps the fastest pair for later runs and reports the workgroups per
 * CU (rf_gcc_batch_group_bytes reports the budget). The results buffer is overwritten. */
rf_status rf_gcc_batch_tune(rf_gcc_batch* handle, rf_peak_result* results, rf_stream stream,
                            int* workgroups_per_cu);

/* ---------------------------------------------------------------------------------------
 * Learned delay search. No counterpart in the reference: a batched GCC-PHAT (the work of
 * gcc_phat_cuda, src/xcorr.cu:379-399, plus the peak loop of rfgpu doa, src/main.cpp:390-401)
 * over `count` equal-length pairs in which a small network proposes the delay bin from cheap
 * features of the PHAT cross-spectrum; where it is confident, the GCC-PHAT output is evaluated
 * exactly over a 64-lag window around the proposal instead of by the full inverse transform.
 * Delays are searched in [-512, 512) samples; 2^15 <= L <= 2^20 (8193 <= n <= 524288).
 * ------------------------------------------------------------------------------------- */
#define RF_DELAY_FEATURES 256 /* network input: pooled coarse |GCC-PHAT| over lags -512..511 */
#define RF_DELAY_BINS 64      /* network output: bins of 16 lags, bin b = lags -512+16b.. */
#define RF_DELAY_SPAN 512
#define RF_DELAY_MLP_PARAMS 148032

/* Synthetic capture pairs (the training and evaluation distribution: delay U[-512, 512),
 * SNR U[-10, 20] dB, AR(1) noise with coefficient U[0, 0.9] per channel, 65-tap windowed-sinc
 * fractional delay), a pure function of (seed, pair index) by Philox4x32-10. a, b: DEVICE,
 * count * n values (pair k at k * n); delays: DEVICE, count floats (true delay of b against a,
 * in samples), or NULL. Pairs first_pair .. first_pair + count - 1 of the seed. */
rf_status rf_synth_pairs(unsigned long long seed, int first_pair, int count, int n, rf_cf32* a,
                         rf_cf32* b, float* delays, rf_stream stream);

/* Delay-proposal network 256 -> 256 (GELU) -> 256 (GELU) -> 64 (softmax). FP32 master
 * parameters (layout W1[256][256], b1[256], W2[256][256], b2[256], W3[64][256], b3[64]); GEMM
 * operands rounded to BF16 on the matrix cores, FP32 accumulate. Workspaces for `max_batch`
 * rows. Weights He-uniform from Philox4x32-10 of `seed`, biases 0. */
typedef struct rf_delay_mlp rf_delay_mlp;
rf_status rf_delay_mlp_create(int max_batch, unsigned long long seed, rf_delay_mlp** handle);
rf_status rf_delay_mlp_destroy(rf_delay_mlp* handle);
/* DEVICE pointer to the RF_DELAY_MLP_PARAMS master parameters (read or overwrite them). */
rf_status rf_delay_mlp_params(rf_delay_mlp* handle, float** params, size_t* count);
/* probs: DEVICE batch x 64 softmax probabilities of DEVICE features (batch x 256). */
rf_status rf_delay_mlp_forward(rf_delay_mlp* handle, const float* features, int batch,
                               float* probs, rf_stream stream);
/* Gradients of the mean cross-entropy against DEVICE int labels (0..63): grads (DEVICE,
 * RF_DELAY_MLP_PARAMS floats, or NULL) and loss (DEVICE, 1 float). Parameters unchanged. */
rf_status rf_delay_mlp_gradients(rf_delay_mlp* handle, const float* features, const int* labels,
                                 int batch, float* grads, float* loss, rf_stream stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) on the batch. */
rf_status rf_delay_mlp_train_step(rf_delay_mlp* handle, const float* features,
                                  const int* labels, int batch, float learning_rate,
                                  float* loss, rf_stream stream);
/* Training driver: ceil(pairs / batch) steps, each synthesising `batch` fresh pairs of n
 * samples on the GPU (rf_synth_pairs of `seed`, pairs step * batch ..), computing their
 * features and taking one Adam step towards the bin of the rounded true delay. losses: HOST,
 * one mean loss per step, or NULL. Synchronises `stream`. */
rf_status rf_delay_mlp_train(rf_delay_mlp* handle, unsigned long long seed, int pairs,
                             int batch, int n, float learning_rate, float* losses,
                             rf_stream stream);

typedef struct rf_delay_search rf_delay_search;
/* First 16 bytes as rf_peak_result: index = output index of the peak (lag = index - (n - 1)),
 * magnitude = |r| there, frac = sub-sample offset (f = m1 / (m0 + m1) towards the larger
 * neighbour m1). proposal = proposed bin, confidence = its probability (1 for the coarse
 * proposal), path = 0 window refine, 1 full inverse transform. */
typedef struct rf_delay_result {
    int32_t index;
    float magnitude;
    double frac;
    int32_t proposal;
    float confidence;
    int32_t path;
    int32_t reserved;
} rf_delay_result;
rf_status rf_delay_search_create(int count, int n, rf_delay_search** handle);
rf_status rf_delay_search_destroy(rf_delay_search* handle);
rf_status rf_delay_search_bytes(const rf_delay_search* handle, size_t* bytes);
/* Forward transforms, PHAT weighting and features of DEVICE inputs a, b (count * n each);
 * features (DEVICE count x 256) and coarse proposals (DEVICE count ints) may be NULL. */
rf_status rf_delay_search_features(rf_delay_search* handle, const rf_cf32* a, const rf_cf32* b,
                                   float* features, int* coarse, rf_stream stream);
/* The search. mlp NULL: the coarse proposal (largest coarse |GCC-PHAT|) for every pair.
 * Pairs whose confidence is below `threshold` (threshold > 1: every pair) take the full
 * inverse transform, and so do pairs whose exact window peak sits at the window's edge or
 * below the PHAT noise floor sqrt((ln L + 5) / L) (all L lags of pure noise stay below it
 * with probability 1 - e^-5). results: DEVICE, count records. */
rf_status rf_delay_search_run(rf_delay_search* handle, const rf_cf32* a, const rf_cf32* b,
                              rf_delay_mlp* mlp, float threshold, rf_delay_result* results,
                              rf_stream stream);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* RFHIP_H */
