/*
 * svt_jpegxs_amd.h -- entry points of the MI355X library beyond the
 * reference API.  The reference has no counterpart for these: they serve
 * callers that keep frames resident in HBM (batch encoders, the serving loop,
 * benchmarks) and want one launch sequence for many frames.
 *
 * Device pointers are HIP device addresses; `stream` is a hipStream_t (NULL
 * for the default stream).
 */
#ifndef SVT_JPEGXS_AMD_EXT_H
#define SVT_JPEGXS_AMD_EXT_H

#include <stdint.h>
#include "SvtJpegxsEnc.h"
#include "SvtJpegxsDec.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One frame of a device-resident batch.  `encoder` is an initialised
 * encoder whose parameters define the coding configuration of this frame
 * (frames of one batch may use different encoders: resolution, format,
 * depth, bpp and decomposition may differ).  Planes follow the reference's
 * image-buffer convention (stride in samples; one interleaved plane for the
 * packed format). */
typedef struct svt_jpeg_xs_amd_frame_spec {
    svt_jpeg_xs_encoder_api_t* encoder;
    const void* d_planes[MAX_COMPONENTS_NUM];
    uint32_t strides[MAX_COMPONENTS_NUM];
    uint8_t* d_out; /* 4-byte aligned, >= bytes_per_frame of the encoder rounded up to a multiple of 4 */
} svt_jpeg_xs_amd_frame_spec_t;

typedef struct svt_jpeg_xs_amd_plan svt_jpeg_xs_amd_plan_t;

/* Builds the job lists, frame descriptors and work buffers of a batch of
 * `count` frames and keeps them in HBM.  Returns NULL on error. */
PREFIX_API svt_jpeg_xs_amd_plan_t* svt_jpeg_xs_amd_plan_create(const svt_jpeg_xs_amd_frame_spec_t* frames, uint32_t count);

/* Enqueues the encode of every frame of the plan on `stream`: one launch per
 * stage for the whole batch.  When `use_graph` is set the launch sequence is
 * captured into a HIP graph on the first call and replayed afterwards. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_plan_run(svt_jpeg_xs_amd_plan_t* plan, void* stream, uint32_t use_graph);

/* Same as plan_run without a graph, recording a HIP event around every
 * kernel on `stream`; waits and returns the milliseconds of each stage in
 * stage_ms[0 .. SVT_JPEG_XS_AMD_STAGES - 1] (transform level 1, transform
 * level 2, rate control and packet plan, pack). */
#define SVT_JPEG_XS_AMD_STAGES 4
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_plan_run_timed(svt_jpeg_xs_amd_plan_t* plan, void* stream, float* stage_ms);

/* After the stream has completed: per-frame error flags (0 = ok), count
 * entries.  Returns SvtJxsErrorEncodeFrameError if any frame failed. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_plan_status(svt_jpeg_xs_amd_plan_t* plan, uint32_t* frame_errors);

PREFIX_API void svt_jpeg_xs_amd_plan_destroy(svt_jpeg_xs_amd_plan_t* plan);

/* Device bytes a plan of these frames allocates (inputs and outputs excluded). */
PREFIX_API uint64_t svt_jpeg_xs_amd_plan_footprint(const svt_jpeg_xs_amd_frame_spec_t* frames, uint32_t count);

/* Encodes one frame whose planes and codestream buffer are in HBM, on
 * `stream`, without allocating: no hipMalloc, hipFree or other allocating
 * routine is called.  No counterpart in the reference: this entry point is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  The frame's work
 * block lives in a 64 MiB device array the library holds on every GPU it runs
 * on until the process ends; a frame that needs more (2160p and larger) is
 * refused with SvtJxsErrorBadParameter before anything is enqueued.
 * `d_status` is the device address of a 4-byte-aligned u32 that receives the
 * frame's error flags (0: the codestream is complete).  The result equals
 * svt_jpeg_xs_amd_plan_create + _run + _status for the same frame.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other threads) race and must be ordered by the caller; calls on one
 * stream are ordered by it. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_encode_noalloc(const svt_jpeg_xs_amd_frame_spec_t* frame, void* stream,
                                                            uint32_t* d_status);

/* Device address of the precinct-offset table the encode plan wrote for
 * frame `frame` (byte offset of every precinct header in its codestream,
 * one u32 every *stride u32).  Valid after the plan has run; a decode plan of
 * the same codestream can take it instead of walking the precinct chain. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_plan_precinct_offsets(svt_jpeg_xs_amd_plan_t* plan, uint32_t frame,
                                                                   const uint32_t** d_offsets, uint32_t* stride);

/* One codestream of a device-resident decode batch.  `decoder` is a decoder
 * initialised on a codestream of the same configuration; planes receive the
 * decoded components (stride in samples) at the decoder's resolution (its
 * proxy_mode: full, half or quarter).  d_precinct_offsets may be NULL, then
 * the precinct chain is walked on the device. */
typedef struct svt_jpeg_xs_amd_decode_spec {
    svt_jpeg_xs_decoder_api_t* decoder;
    const uint8_t* d_codestream;
    uint32_t codestream_size;
    void* d_planes[MAX_COMPONENTS_NUM];
    uint32_t strides[MAX_COMPONENTS_NUM];
    const uint32_t* d_precinct_offsets;
    uint32_t precinct_offsets_stride;
} svt_jpeg_xs_amd_decode_spec_t;

typedef struct svt_jpeg_xs_amd_decode_plan svt_jpeg_xs_amd_decode_plan_t;

PREFIX_API svt_jpeg_xs_amd_decode_plan_t* svt_jpeg_xs_amd_decode_plan_create(const svt_jpeg_xs_amd_decode_spec_t* frames,
                                                                           uint32_t count);
/* Enqueues unpack, inverse transform and output scaling of every codestream. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_decode_plan_run(svt_jpeg_xs_amd_decode_plan_t* plan, void* stream);
/* After completion: per-codestream error flags; SvtJxsErrorDecoderInvalidBitstream if any. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_decode_plan_status(svt_jpeg_xs_amd_decode_plan_t* plan, uint32_t* errors);
PREFIX_API void svt_jpeg_xs_amd_decode_plan_destroy(svt_jpeg_xs_amd_decode_plan_t* plan);

/* ---------------- transform tile autotuner ----------------
 * The transform kernels give each thread a chunk of 8, 16 or 32 consecutive
 * columns of a row (8: 512-thread blocks; 16 and 32: 256-thread blocks).
 * svt_jpeg_xs_amd_autotune times one frame of the encoder's configuration
 * with every width the kernels support for each of its transform job shapes
 * (input kind, vertical level, horizontal levels, width), and then with 4, 2
 * and 1 waves per precinct in the rate table kernel (rate control modes 0 and
 * 1), on `stream`, and records the fastest in a per-process table that every
 * plan, batch, server launch and encoder built afterwards uses; the encoder's
 * own plain-path plans are rebuilt.  Output is identical for every choice.
 * Call it with no frame of this encoder in flight.  Fills up to max_out
 * entries of out (may be NULL) and returns the number of choices made, or -1
 * on error.  The reference has no counterpart (its tiles are fixed by its
 * SIMD kernels). */
typedef struct svt_jpeg_xs_amd_tile {
    uint32_t width, levels, vertical; /* stage 0: job shape (plane width, horizontal levels, vertical level or not);
                                         stage 1: picture width */
    uint32_t chunk;                   /* stage 0: columns per thread; stage 1: waves per precinct */
    float frame_us;                   /* that stage's time of one frame with the choice */
    uint32_t stage;                   /* 0: transform tile, 1: rate table waves per precinct */
} svt_jpeg_xs_amd_tile_t;

PREFIX_API int32_t svt_jpeg_xs_amd_autotune(svt_jpeg_xs_encoder_api_t* encoder, void* stream,
                                            svt_jpeg_xs_amd_tile_t* out, uint32_t max_out);

/* ---------------- serving loop ----------------
 * A request loop for many live streams whose frames arrive in HBM (from a
 * capture DMA, a decoder, or a host copy the caller made).  Frames are
 * admitted while the server's HBM budget has room for their work buffers and
 * descriptors, and grouped into batches by continuous batching: whenever a
 * launch slot is free, every queued frame that fits (up to max_batch, in
 * arrival order) goes into one launch; frames that arrive meanwhile wait for
 * the next.  Up to max_inflight batches run at once, each on its own HIP
 * stream, so their kernels overlap on the device.  The per-stream state is
 * the encoder's (band layout, weight tables, slice budgets), resident in HBM
 * from svt_jpeg_xs_encoder_init; a frame's codestream is the same whatever
 * batch carries it.  The reference has no counterpart: its encoder takes one
 * stream per handle (EncHandle.c). */

typedef struct svt_jpeg_xs_amd_server_config {
    uint64_t hbm_budget;   /* device bytes for admitted frames' work buffers and batch descriptors */
    uint32_t max_batch;    /* most frames in one launch, 1 .. 4096 */
    uint32_t max_inflight; /* batches in flight at once, one HIP stream each, 1 .. 4 */
} svt_jpeg_xs_amd_server_config_t;

typedef struct svt_jpeg_xs_amd_completion {
    uint64_t tag;          /* as submitted */
    int32_t status;        /* SvtJxsErrorNone, SvtJxsErrorEncodeFrameError, or SvtJxsErrorUndefined (launch failed) */
    uint32_t batch_frames; /* frames in the launch that carried this frame */
    uint64_t submit_ns;    /* times on the svt_jpeg_xs_amd_server_now_ns clock: submitted, */
    uint64_t launch_ns;    /* its batch launched, */
    uint64_t done_ns;      /* completion observed (the codestream is in d_out) */
} svt_jpeg_xs_amd_completion_t;

typedef struct svt_jpeg_xs_amd_server_stats {
    uint64_t submitted, completed, batches;
    uint64_t largest_batch;
    uint64_t hbm_budget, hbm_in_use, hbm_peak; /* bytes of the budget held by batches in flight */
    uint32_t queued, inflight;
} svt_jpeg_xs_amd_server_stats_t;

typedef struct svt_jpeg_xs_amd_server svt_jpeg_xs_amd_server_t;

/* Allocates the HBM budget on the current device and starts the loop.
 * Returns NULL on a bad configuration or when the budget cannot be allocated. */
PREFIX_API svt_jpeg_xs_amd_server_t* svt_jpeg_xs_amd_server_create(const svt_jpeg_xs_amd_server_config_t* cfg);

/* Queues one frame; never blocks.  The encoder, planes and codestream buffer
 * must stay valid until the frame's completion has been polled.
 * SvtJxsErrorBadParameter: invalid spec (as for a plan).
 * SvtJxsErrorInsufficientResources: the frame alone exceeds the budget. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_server_submit(svt_jpeg_xs_amd_server_t* server,
                                                           const svt_jpeg_xs_amd_frame_spec_t* frame, uint64_t tag);

/* Moves up to max completions, in the order they were observed, into out.
 * With blocking set, waits until at least one is available or nothing is
 * queued or in flight.  Returns how many were written. */
PREFIX_API uint32_t svt_jpeg_xs_amd_server_poll(svt_jpeg_xs_amd_server_t* server, svt_jpeg_xs_amd_completion_t* out,
                                                uint32_t max, uint32_t blocking);

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_server_stats(svt_jpeg_xs_amd_server_t* server,
                                                          svt_jpeg_xs_amd_server_stats_t* stats);

/* Monotonic clock of the completion records, in nanoseconds. */
PREFIX_API uint64_t svt_jpeg_xs_amd_server_now_ns(void);

/* Finishes every queued and in-flight frame, stops the loop and releases the
 * budget.  Completions not yet polled are dropped. */
PREFIX_API void svt_jpeg_xs_amd_server_destroy(svt_jpeg_xs_amd_server_t* server);

/* Version / build string of the library. */
PREFIX_API const char* svt_jpeg_xs_amd_version(void);

#ifdef __cplusplus
}
#endif
#endif
