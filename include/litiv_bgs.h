/*
 * litiv_bgs.h -- C ABI of the MI355X (gfx950) SuBSENSE / LOBSTER background subtractors.
 *
 * Drop-in boundary for litiv's IBackgroundSubtractor surface
 * (modules/video/include/litiv/video/BackgroundSubtractionUtils.hpp:24-97): plain pointers and
 * sizes, no exceptions across the ABI (status codes + per-context message), one context per
 * video stream. Each entry point names the reference interface it replaces.
 *
 * Threading: calls on one context must be serialized; different contexts are independent
 * (their random streams are keyed by lvbgs_params.stream_id) and may be used concurrently.
 * Images are 8-bit, 1 or 3 channels, interleaved (HWC), row stride in bytes.
 */
#ifndef LITIV_BGS_H
#define LITIV_BGS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LVBGS_OK 0
#define LVBGS_ERR_ARG (-1)      /* invalid argument (the reference's lvAssert_ failures) */
#define LVBGS_ERR_STATE (-2)    /* "algo & model must be initialized first" */
#define LVBGS_ERR_HIP (-3)      /* HIP runtime error */
#define LVBGS_ERR_ALLOC (-4)    /* device memory allocation failed */
#define LVBGS_ERR_DEVICE (-5)   /* device-side consistency check failed */
#define LVBGS_ERR_BUDGET (-6)   /* a server's device-memory budget cannot hold the model */
#define LVBGS_ERR_BUSY (-7)     /* a served stream's host-ingest slots are all busy: poll, then retry */
#define LVBGS_ERR_TOO_LARGE (-8) /* input beyond a fixed-capacity entry point's stated maximum */

enum { LVBGS_SUBSENSE = 1, LVBGS_LOBSTER = 2 };

typedef struct lvbgs_ctx lvbgs_ctx;

/* Constructor parameters (BackgroundSubtractorSuBSENSE.hpp:52-57, BackgroundSubtractorLOBSTER.hpp:50-55,
 * BackgroundSubtractorLBSP.hpp:46-49), plus the counter-based RNG key and the device. */
typedef struct lvbgs_params {
    size_t desc_dist_threshold;   /* SuBSENSE nDescDistThresholdOffset (3) | LOBSTER nDescDistThreshold (4) */
    size_t color_dist_threshold;  /* SuBSENSE nMinColorDistThreshold (30) | LOBSTER nColorDistThreshold (30) */
    size_t bg_samples;            /* nBGSamples: 50 | 35 (this build: 1..255) */
    size_t required_bg_samples;   /* nRequiredBGSamples: 2 */
    size_t samples_for_mv_avgs;   /* SuBSENSE nSamplesForMovingAvgs (100) */
    size_t lbsp_threshold_offset; /* nLBSPThresholdOffset (0) */
    float rel_lbsp_threshold;     /* fRelLBSPThreshold (0.333f) */
    int median_blur_kernel;       /* nDefaultMedianBlurKernelSize (9), odd, 3..13 */
    uint64_t rng_seed;            /* replaces the global libc rand() state */
    uint32_t stream_id;           /* per-stream RNG counter word */
    int device;                   /* HIP device ordinal */
} lvbgs_params;

/* Fills the reference defaults of the given algorithm. */
int lvbgs_default_params(int algo, lvbgs_params* out);

/* Constructor (BackgroundSubtractorSuBSENSE.cpp:62-78 / BackgroundSubtractorLOBSTER.cpp:20-30). */
int lvbgs_create(int algo, const lvbgs_params* params, lvbgs_ctx** out);
void lvbgs_destroy(lvbgs_ctx* ctx);

/* IIBackgroundSubtractor::initialize(img, roi) (BackgroundSubtractionUtils.hpp:29-31;
 * SuBSENSE .cpp:107-186, LOBSTER .cpp:442-457). Host pointers. roi may be NULL (reuse the last ROI
 * of the same size, else the full frame); otherwise a 0/255 mask of the same size. */
int lvbgs_initialize(lvbgs_ctx* ctx, const uint8_t* img, int width, int height, int channels, size_t stride,
                     const uint8_t* roi, size_t roi_stride);

/* initialize from a device-resident, tightly packed frame (same contract otherwise). */
int lvbgs_initialize_device(lvbgs_ctx* ctx, const void* d_img, int width, int height, int channels,
                            const uint8_t* roi, size_t roi_stride);

/* Device bytes a context of this geometry allocates at initialize (model + per-step buffers);
 * 0 for invalid arguments. params may be NULL (defaults). */
size_t lvbgs_model_bytes(int algo, const lvbgs_params* params, int width, int height, int channels);

/* IIBackgroundSubtractor::apply(image, fgmask, learningRate) (BackgroundSubtractionUtils.hpp:35;
 * SuBSENSE .cpp:188-612, LOBSTER .cpp:459-581). Host pointers, synchronous. fgmask receives the
 * width x height 0/255 mask. SuBSENSE: lr <= 0 uses the per-pixel T(x), lr = +inf never updates;
 * LOBSTER: lr must be > 0. */
int lvbgs_apply(lvbgs_ctx* ctx, const uint8_t* img, size_t stride, uint8_t* fgmask, size_t mask_stride, double lr);

/* Same step on device-resident, tightly packed buffers; enqueued on the context's stream and
 * returns without waiting. d_mask may be NULL. */
int lvbgs_apply_device(lvbgs_ctx* ctx, const void* d_img, void* d_mask, double lr);

/* Waits for the context's stream; reports device-side errors of the steps since the last call. */
int lvbgs_synchronize(lvbgs_ctx* ctx);

/* getBackgroundImage (SuBSENSE .cpp:614-630, LOBSTER .cpp:583-600). Host pointer. */
int lvbgs_get_background_image(lvbgs_ctx* ctx, uint8_t* out, size_t stride);

/* getBackgroundDescriptorsImage (BackgroundSubtractorLBSP.hpp:41; SuBSENSE .cpp:632-649). Host pointer,
 * 16-bit values, stride in bytes. */
int lvbgs_get_background_descriptors_image(lvbgs_ctx* ctx, uint16_t* out, size_t stride);

/* setAutomaticModelReset (BackgroundSubtractionUtils.cpp:24-26). */
int lvbgs_set_automatic_model_reset(lvbgs_ctx* ctx, int enabled);

/* refreshModel(fSamplesRefreshFrac, bForceFGUpdate) (SuBSENSE .cpp:80-105, LOBSTER .cpp:410-440). */
int lvbgs_refresh_model(lvbgs_ctx* ctx, float frac, int force_fg);

/* getDefaultLearningRate (SuBSENSE .hpp:69 -> 0, LOBSTER .hpp:57 -> 16). */
double lvbgs_default_learning_rate(const lvbgs_ctx* ctx);

/* getROICopy (BackgroundSubtractionUtils.cpp:50-52). */
int lvbgs_get_roi(lvbgs_ctx* ctx, uint8_t* out, size_t stride);

/* validateROI (BackgroundSubtractionUtils.cpp:28-36): zeroes the 2-pixel LBSP border in place. */
int lvbgs_validate_roi(int width, int height, uint8_t* roi, size_t stride);

/* setROI (BackgroundSubtractionUtils.cpp:38-48): validates roi in place; if initialized, re-initializes
 * from the current background image with the new ROI, otherwise stores it for the next initialize. */
int lvbgs_set_roi(lvbgs_ctx* ctx, uint8_t* roi, size_t stride);

/* Last error message of the context ("" when none). */
const char* lvbgs_last_error(const lvbgs_ctx* ctx);

/* LBSP::compute2 with an absolute threshold (LBSP.cpp:102-123, :227-233): device pointers,
 * tightly packed; border pixels (2 px) are written as 0. stream: hipStream_t or NULL. */
int lvbgs_lbsp_compute_abs(const void* d_img, int width, int height, int channels, int threshold, void* d_out,
                           void* stream);

/* ---- introspection (tests, benchmarks) ---- */
typedef struct lvbgs_stats {
    uint64_t examined;        /* samples examined by the final ordered match, summed over pixels (k_p) */
    uint64_t writes;          /* sample writes landing on pixels (w_p) */
    uint32_t fixpoint_iters;  /* fix-point iterations that had work in the last step */
    uint32_t max_fixpoint_iter;
    uint32_t width, height, channels, n_roi;
    uint64_t examined_speculative; /* samples examined by the speculative match kernel alone */
    uint32_t fixpoint_work[8];     /* [0]: pixels finished by the cooperative match phase;
                                      [k]: worklist size of fix-point iteration k (1..7) */
    uint64_t examined_cooperative; /* samples examined by the cooperative match phase */
} lvbgs_stats;
int lvbgs_get_stats(lvbgs_ctx* ctx, lvbgs_stats* out);

/* Copies a named internal buffer to host memory (samples_color, samples_desc, last_color, last_desc,
 * last_fg, roi, lut, T, R, V, MLD, DminLT, DminST, RawLT, RawST, FinLT, FinST, unstable, blinks,
 * last_raw, last_raw_blink, last_dil_inv, dsLT, dsST). Returns bytes written or < 0. */
long long lvbgs_get_state(lvbgs_ctx* ctx, const char* name, void* out, size_t cap);

/* Frame-level scalars in the order frame_idx, frames_since_reset, cooldown, auto_reset, lr_scaling,
 * use3x3, median_k, n_roi, lower_cap, upper_cap, last_nz_ratio, refresh_seq, apply_count. */
int lvbgs_get_scalars(lvbgs_ctx* ctx, double* out, int cap);

/* hipStream_t of the context (for ordering device-pointer calls with other work). */
void* lvbgs_stream(lvbgs_ctx* ctx);

/* Kernel timing: stage = "match" (phase A of the sample match) | "update" (model update and feedback
 * maps) | "step" | "" (off). While on, every step records HIP events around the stage on the context's
 * stream; lvbgs_take_timing returns the summed milliseconds and count since
 * the last take (it synchronizes). */
int lvbgs_set_timing(lvbgs_ctx* ctx, const char* stage);
int lvbgs_take_timing(lvbgs_ctx* ctx, double* total_ms, int* count);

/* Fix-point schedule: full-grid iterations before the single-workgroup tail, and their grid size. */
int lvbgs_set_fixpoint_schedule(lvbgs_ctx* ctx, int launches, int blocks);

/* Match schedule: samples scanned per pixel by the sequential phase (1..255; 0, the default: 8 for
 * SuBSENSE RGB frames of 1280x720 pixels or more, else 255) before unfinished pixels move to the
 * continuation phases, and the grid of those phases. Results do not depend on it. */
int lvbgs_set_match_schedule(lvbgs_ctx* ctx, int seq_limit, int cont_blocks);

/* LBSP self-check: lvbgs_lbsp_compute_abs through the form of the threshold test the sample match
 * uses (taps packed four per register, masked quad SADs) instead of the per-tap form the descriptor
 * kernels use; the two must give the same bits. */
int lvbgs_debug_lbsp_packed(const void* d_img, int width, int height, int channels, int threshold, void* d_out,
                            void* stream);

/* RNG self-check: the four draws of group 0 of counter (pixel p, frame, phase, event) for p in [0, n),
 * written to host_out[4*p .. 4*p+3] (phase 1 = apply, 2 = refresh). */
int lvbgs_debug_rng(lvbgs_ctx* ctx, uint32_t frame, uint32_t phase, uint32_t event, uint32_t n, uint32_t* host_out);

/* Post-processing self-check (SuBSENSE): runs the hole filling and median stage of
 * BackgroundSubtractorSuBSENSE.cpp:540-547 on a caller-given W x H raw mask (0/255, tightly packed host
 * buffer) and writes the resulting median-filtered mask to host_last_fg. Replaces the context's
 * last foreground mask; for tests only. */
int lvbgs_debug_postprocess(lvbgs_ctx* ctx, const uint8_t* host_raw, uint8_t* host_last_fg);

/* ---- batches (no counterpart in the reference: its changedet app runs one sequence per host
 * thread, apps/changedet/src/main.cpp:148-152) ----
 * A batch steps several initialized SuBSENSE contexts (any mix of sizes, 1 or 3 channels) together:
 * one launch per kernel covers every stream. While a batch is in use its contexts must not be
 * stepped individually. Device pointers, tightly packed; asynchronous on the batch's stream. */
typedef struct lvbgs_batch lvbgs_batch;
int lvbgs_batch_create(lvbgs_ctx* const* ctxs, int n, lvbgs_batch** out);
void lvbgs_batch_destroy(lvbgs_batch* b);
const char* lvbgs_batch_last_error(const lvbgs_batch* b);
void* lvbgs_batch_stream(lvbgs_batch* b);
/* Launch options of a batch. Results never depend on them. */
typedef struct lvbgs_batch_options {
    int use_graph;          /* capture the step once as a hipGraph and replay it */
    int seq_limit;          /* samples scanned by the sequential match phase (1..255) */
    int cont_blocks;        /* cooperative match phase: grid of a stream, at most */
    int cont_px_per_block;  /*   ... and ceil(pixels / this) when > 0 */
    int fixpoint_launches;  /* full-grid fix-point iterations before the one-workgroup tail */
    int fixpoint_blocks;    /* fix-point grid of a stream, at most */
    int fix_px_per_block;   /*   ... and ceil(pixels / this) when > 0 */
    int fetch_depth;        /* sample loads in flight per pixel in the sequential match (2 or 4) */
    int groups;             /* 1..4: members split into groups stepped concurrently on their own streams */
} lvbgs_batch_options;
int lvbgs_batch_default_options(lvbgs_batch_options* out);
int lvbgs_batch_get_options(const lvbgs_batch* b, lvbgs_batch_options* out);
int lvbgs_batch_set_options(lvbgs_batch* b, const lvbgs_batch_options* opts);
/* Per-stream overrides of seq_limit, cont_px_per_block and fix_px_per_block for member `index`
 * (0 = the batch's value), so the launch shape can follow each stream's resolution. Results never
 * depend on them. */
int lvbgs_batch_set_stream_options(lvbgs_batch* b, int index, int seq_limit, int cont_px_per_block,
                                   int fix_px_per_block);
int lvbgs_batch_get_stream_options(const lvbgs_batch* b, int index, int* seq_limit, int* cont_px_per_block,
                                   int* fix_px_per_block);
/* one apply() of every stream: d_imgs[i] is stream i's frame, d_masks (nullable) its output mask,
 * lrs (nullable = 0) its learning-rate override */
int lvbgs_batch_apply_device(lvbgs_batch* b, const void* const* d_imgs, void* const* d_masks, const double* lrs);
int lvbgs_batch_synchronize(lvbgs_batch* b);
/* GPU time of batch steps: stage "match" (phase A of the match) | "update" (the update kernel) | "step"
 * (whole step) | "" (off); kernel stages need groups = 1;
 * HIP events on the batch stream, summed by lvbgs_batch_take_timing (which synchronizes) */
int lvbgs_batch_set_timing(lvbgs_batch* b, const char* stage);
int lvbgs_batch_take_timing(lvbgs_batch* b, double* total_ms, int* count);

/* ---- serving (no counterpart in the reference) ----
 * A server keeps SuBSENSE models resident on one device and batches whatever frames are waiting:
 * each step takes the oldest waiting frame of every stream that has one, and frames submitted
 * while steps run join the next step (continuous batching). Streams are admitted against a
 * device-memory budget; closed streams' contexts are kept for reuse by streams of the same
 * geometry. A stream's results equal those of the same stream run alone. Frames and masks are
 * device pointers that must stay valid until their completion is polled. Not thread-safe. */
typedef struct lvbgs_server lvbgs_server;
typedef struct lvbgs_server_options {
    size_t hbm_budget;            /* bytes of stream contexts (lvbgs_model_bytes) held at most */
    int max_inflight;             /* steps queued on the device at once (1..16) */
    int max_batch;                /* streams per step at most */
    lvbgs_batch_options launch;   /* launch options of each step (use_graph is ignored) */
    int host_ring;                /* 0..64 device frame slots per stream for lvbgs_server_submit_host, held in
                                     the stream's arena and counted in the budget (0: device frames only) */
} lvbgs_server_options;
typedef struct lvbgs_done {
    uint64_t tag;                 /* the submit call's tag */
    int32_t stream;
    int32_t status;               /* LVBGS_OK, or LVBGS_ERR_DEVICE: a device-side check failed for this
                                     stream (fix-point iteration cap); the server then closes the stream,
                                     completes its waiting frames with the same status and rejects
                                     further submits; lvbgs_server_close releases its id. Other streams
                                     of the same steps are unaffected. */
} lvbgs_done;
typedef struct lvbgs_server_stats {
    uint64_t steps, frames_submitted, frames_launched, frames_done;
    uint64_t bytes_used, pending_frames;
    uint32_t open_streams, pooled_contexts, queued_steps, reserved;
} lvbgs_server_stats;
int lvbgs_server_default_options(lvbgs_server_options* out);
int lvbgs_server_create(int device, const lvbgs_server_options* opts, lvbgs_server** out);
void lvbgs_server_destroy(lvbgs_server* s);
const char* lvbgs_server_last_error(const lvbgs_server* s);
void* lvbgs_server_stream(lvbgs_server* s);
/* admits a SuBSENSE stream initialized from d_img0 (full-frame ROI); LVBGS_ERR_BUDGET when it does not
 * fit. The initialization is queued on the device: d_img0 must stay valid until the stream's first
 * frame completes; its frames join steps once the initialization is done. */
int lvbgs_server_open(lvbgs_server* s, const lvbgs_params* params, const void* d_img0, int width, int height,
                      int channels, int* stream_out);
/* no further frames; the stream's waiting frames still complete, then its id may be reused.
 * On a stream the server closed after a device error it acknowledges the failure (LVBGS_OK). */
int lvbgs_server_close(lvbgs_server* s, int stream);
int lvbgs_server_submit(lvbgs_server* s, int stream, const void* d_img, void* d_mask, double lr, uint64_t tag);
/* Ingest from host memory (no counterpart in the reference, whose datasets module prefetches frames
 * on a host thread, modules/datasets/include/litiv/datasets/utils.hpp:456 DataPrecacher): the frame
 * (row stride in bytes, 0 = packed) is copied at once into the stream's next free device slot on the
 * server's copy stream, so the copy overlaps the steps already on the device; the step that takes the
 * frame waits for its copy on the device. Needs host_ring > 0. host_img must stay valid and unchanged
 * until the copy has run (at the latest, until the frame completes); pinned memory (lvbgs_host_alloc)
 * makes the copy asynchronous. LVBGS_ERR_BUSY when all the stream's slots hold frames not completed. */
int lvbgs_server_submit_host(lvbgs_server* s, int stream, const void* host_img, size_t stride, void* d_mask,
                             double lr, uint64_t tag);
/* lvbgs_server_open from a frame in host memory (row stride in bytes, 0 = packed) */
int lvbgs_server_open_host(lvbgs_server* s, const lvbgs_params* params, const void* img0, size_t stride, int width,
                           int height, int channels, int* stream_out);
/* device bytes a SuBSENSE stream of this geometry holds in this server (what admission counts against
 * the budget: lvbgs_model_bytes plus its host-ingest slots); 0 for invalid arguments */
size_t lvbgs_server_stream_bytes(const lvbgs_server* s, const lvbgs_params* params, int width, int height,
                                 int channels);
/* pinned (page-locked) host memory for frames handed to lvbgs_server_submit_host; NULL on failure */
void* lvbgs_host_alloc(size_t bytes);
void lvbgs_host_free(void* p);
/* non-blocking: collects finished steps, launches steps while the device queue has room;
 * returns the number of steps launched or < 0 */
int lvbgs_server_pump(lvbgs_server* s);
/* up to cap completions since the last call, in completion order; returns the count */
int lvbgs_server_poll(lvbgs_server* s, lvbgs_done* out, int cap);
/* blocks until every submitted frame has completed */
int lvbgs_server_drain(lvbgs_server* s);
int lvbgs_server_get_stats(lvbgs_server* s, lvbgs_server_stats* out);
/* the context of an open stream (getters and state reads after lvbgs_server_drain) */
lvbgs_ctx* lvbgs_server_context(lvbgs_server* s, int stream);
/* Test hook (no counterpart in the reference): iteration cap of the single-workgroup fix-point tail
 * of later steps (default 65536). A stream whose raster-order fix point needs more iterations than
 * fixpoint_launches + cap stops there with LVBGS_ERR_DEVICE instead of finishing; every kernel still
 * exits normally. Lets tests drive the failure path from the host. */
int lvbgs_server_debug_fixpoint_cap(lvbgs_server* s, int tail_cap);

/* ---- evaluation ----
 * Binary classification counts of a foreground mask against ground truth, on the device:
 * lv::BinClassif (modules/datasets/include/litiv/datasets/metrics.hpp:32-66) and its accumulate()
 * (modules/datasets/src/metrics.cpp:21-50). Labels as in the CDnet scripts (metrics.hpp:23-28):
 * ground truth 85 (out of scope) or 170 (unknown), or a ROI value of 0, make a pixel a don't-care;
 * otherwise 255 is positive and every other value negative; a positive result on shadow (50)
 * ground truth also counts as a special error. */
typedef struct lvbgs_binclassif {
    uint64_t tp, tn, fp, fn, se, dc;  /* nTP, nTN, nFP, nFN, nSE, nDC */
} lvbgs_binclassif;
/* BinClassif::accumulate(oClassif, oGT, oROI): adds the counts of d_classif against d_gt into *counts.
 * Device pointers, width x height 8-bit masks sharing one row stride in bytes (0 = width); d_gt NULL
 * counts every pixel as a don't-care (metrics.cpp:26-29), d_roi NULL means no ROI. Ordered on
 * `stream` (hipStream_t or NULL); synchronizes it before returning. Allocates a small device
 * buffer per call (about 8 bytes per 1024 pixels). */
int lvbgs_binclassif_accumulate(const void* d_classif, const void* d_gt, const void* d_roi, int width, int height,
                                size_t stride, lvbgs_binclassif* counts, void* stream);
/* The same counts without any device allocation. No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, serving loops and harnesses that forbid
 * hipMalloc). The partial counts live in a static 512 KiB device array of the library, one copy per
 * device that uses it, held until the process ends; so a frame may hold at most
 * LVBGS_BINCLASSIF_NOALLOC_MAX_PIXELS pixels (LVBGS_ERR_TOO_LARGE beyond, before any launch). Every
 * call on one device shares that array: calls that overlap on one device (from several host threads)
 * race and must be serialized by the caller. Arguments and results as lvbgs_binclassif_accumulate. */
#define LVBGS_BINCLASSIF_NOALLOC_MAX_PIXELS 67100672ull
int lvbgs_binclassif_accumulate_noalloc(const void* d_classif, const void* d_gt, const void* d_roi, int width,
                                        int height, size_t stride, lvbgs_binclassif* counts, void* stream);

/* ---- binary matrix archives and checkpoint / resume ----
 * lv::write / lv::read with lv::MatArchive_BINARY (modules/utils/src/opencv.cpp:514-531, :608-625):
 * int32 OpenCV type (depth | (channels - 1) << 3), uint64 element size, uint64 element count,
 * int32 dimension count, int32 sizes, then the dense elements. Files written here are read by the
 * reference's lv::read and the other way round. */
int lvbgs_archive_write(const char* path, int cv_type, int dims, const int32_t* sizes, const void* data);
/* Reads the header (type, dims, up to max_dims sizes) and, with data != NULL, the elements (at most
 * cap bytes). Returns the element bytes, or LVBGS_ERR_ARG for a missing / truncated / inconsistent
 * file or elements larger than cap. */
long long lvbgs_archive_read(const char* path, int* cv_type, int* dims, int32_t* sizes, int max_dims, void* data,
                             size_t cap);
/* Writes the whole model of an initialized context into the existing directory dir, one archive per
 * state buffer (the getters' layouts) plus meta.bin and lut.bin. The reference keeps its model only
 * in memory (BackgroundSubtractorSuBSENSE.hpp:71-131); a checkpoint lets a stream resume after a
 * restart. */
int lvbgs_checkpoint_save(lvbgs_ctx* ctx, const char* dir);
/* Restores a checkpoint into a context created with the same algorithm, parameters, rng_seed and
 * stream_id (else LVBGS_ERR_ARG): the context then continues the saved sequence bit for bit. A load
 * that fails after reading the geometry (a missing or truncated buffer) leaves the context
 * uninitialized: initialize it or load a checkpoint again. */
int lvbgs_checkpoint_load(lvbgs_ctx* ctx, const char* dir);

#ifdef __cplusplus
}
#endif
#endif
