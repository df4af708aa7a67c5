/*
 * orc.h -- CPU oracle for the SuBSENSE / LOBSTER background subtractors.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * sequential CPU algorithm, used as the parity checker for the HIP library and
 * as the timed CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (litiv_amd/) never does.
 *
 * Reference files restated here (paths relative to the litiv repository):
 *   modules/video/src/BackgroundSubtractorSuBSENSE.cpp        (all)
 *   modules/video/src/BackgroundSubtractorLOBSTER.cpp:410-620
 *   modules/video/src/BackgroundSubtractorLBSP.cpp:21-65
 *   modules/video/src/BackgroundSubtractionUtils.cpp:20-155
 *   modules/features2d/include/litiv/features2d/LBSP.hpp:143-319
 *   modules/utils/include/litiv/utils/math.hpp:199-234,628-735
 *   modules/utils/include/litiv/utils/opencv.hpp:859-966
 *   modules/datasets/src/metrics.cpp:21-50 (orc_eval.c)
 *
 * Deliberate, documented differences from the reference binary:
 *   - every libc rand() call is replaced by a counter-based Philox4x32-10 draw
 *     keyed on (seed, stream, frame, phase/event/group, pixel); see orc_rng.c;
 *   - OpenCV whole-image ops are restated by their closed forms (orc_post.c);
 *   - built with -ffp-contract=off (the reference build allows contraction).
 *
 * Pins: tests/test_oracle_pins.py checks each primitive against an outside reference (Random123
 * Philox vectors, the reference's LBSP golden data, the reference's sampling table, scipy's
 * morphology/median/hole filling, closed forms); tests/test_oracle_apply_pins.py checks the whole
 * apply(), refreshModel and initialize against closed forms derived by hand from the reference's
 * lines (threshold flip points, feedback trajectories, controller reset frame and caps, scaling
 * table); tests/test_oracle_behaviour.py checks behaviour on scenes with known answers.
 * Parity unpinned: the fp32 rounding of OpenCV's addWeighted / accumulateWeighted (restated here as
 * fixed unfused formulas), orc_resize_area for non-integer factors.
 */
#ifndef LITIV_ORACLE_ORC_H
#define LITIV_ORACLE_ORC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_SUBSENSE = 1, ORC_LOBSTER = 2 };

/* Constructor parameters, named after the reference members they set.
 * SuBSENSE: BackgroundSubtractorSuBSENSE.hpp:52-57; LOBSTER: BackgroundSubtractorLOBSTER.hpp:50-55. */
typedef struct orc_params {
    size_t desc_dist_threshold;    /* SuBSENSE m_nDescDistThresholdOffset (3) | LOBSTER m_nDescDistThreshold (4) */
    size_t color_dist_threshold;   /* SuBSENSE m_nMinColorDistThreshold (30)  | LOBSTER m_nColorDistThreshold (30) */
    size_t bg_samples;             /* N: 50 | 35 */
    size_t required_bg_samples;    /* 2 | 2 */
    size_t samples_for_mv_avgs;    /* SuBSENSE only: 100 */
    size_t lbsp_threshold_offset;  /* 0 */
    float  rel_lbsp_threshold;     /* 0.333f */
    int    median_blur_kernel;     /* default median kernel size (9), BackgroundSubtractorLBSP.hpp:28 */
    uint64_t rng_seed;             /* Philox key */
    uint32_t stream_id;            /* Philox counter word 3 */
} orc_params;

typedef struct orc_bgs orc_bgs;

void     orc_default_params(int algo, orc_params* out);
orc_bgs* orc_create(int algo, const orc_params* p);
void     orc_destroy(orc_bgs* o);
/* returns 0 on success, <0 on contract violation (message via orc_last_error) */
int      orc_initialize(orc_bgs* o, const uint8_t* img, int w, int h, int c, const uint8_t* roi /*nullable, w*h 0/255*/);
int      orc_apply(orc_bgs* o, const uint8_t* img, uint8_t* mask_out, double lr);
int      orc_get_background_image(const orc_bgs* o, uint8_t* out);
int      orc_get_background_descriptors_image(const orc_bgs* o, uint16_t* out);
int      orc_refresh_model(orc_bgs* o, float frac, int force_fg);
void     orc_set_automatic_model_reset(orc_bgs* o, int enabled);
const char* orc_last_error(const orc_bgs* o);
/* state access for parity tests: copies the named buffer; returns bytes copied or <0 */
long long orc_state(const orc_bgs* o, const char* name, void* out, size_t cap);
/* scalar state: frame_idx, frames_since_reset, cooldown, auto_reset, lr_scaling, use3x3, median_k,
 * n_roi, lower_cap*1000, upper_cap*1000 ... returned as doubles in a fixed order (see orc_bgs.c) */
int      orc_scalars(const orc_bgs* o, double* out, int cap);
/* deterministic workload counters of the last apply: samples examined by the final match (sum k_p)
 * and sample writes landing on pixels (sum w_p) */
void     orc_counters(const orc_bgs* o, uint64_t* examined, uint64_t* writes);

/* ---- primitives (exposed for unit tests) ---- */
void     orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t orc_draw(uint64_t seed, uint32_t stream, uint32_t frame, uint32_t phase, uint32_t event,
                  uint32_t pixel, uint32_t draw);
uint16_t orc_lbsp_threshold(const uint8_t vals[16], uint8_t ref, uint8_t t);
void     orc_lbsp_lookup(const uint8_t* img, int w, int c, int x, int y, int ch, uint8_t vals[16]);
void     orc_lbsp_compute_abs(const uint8_t* img, int w, int h, int c, int threshold, uint16_t* out);
void     orc_build_lut(int channels, float rel, size_t offset, uint8_t lut[256]);
uint8_t  orc_sat_u8_f(float v);
void     orc_sample_pos_7x7(int rnd, int x, int y, int border, int w, int h, int* ox, int* oy);
void     orc_neighbor_pos_3x3(int rnd, int x, int y, int border, int w, int h, int* ox, int* oy);
void     orc_neighbor_pos_5x5(int rnd, int x, int y, int border, int w, int h, int* ox, int* oy);
uint8_t  orc_l1dist3_u8(const uint8_t a[3], const uint8_t b[3]);
/* whole-image closed forms (u8 0/255 masks, w*h) */
void     orc_morph(const uint8_t* in, uint8_t* out, int w, int h, int radius, int is_max);
void     orc_median_binary(const uint8_t* in, uint8_t* out, int w, int h, int k);
void     orc_flood_holes(const uint8_t* in, uint8_t* holes, int w, int h);
void     orc_resize_area(const uint8_t* in, int w, int h, int c, uint8_t* out, int dw, int dh);
/* binary classification counts of a mask against ground truth (orc_eval.c), accumulated into
 * out = {TP, TN, FP, FN, SE, DC}; gt and roi may be NULL */
void     orc_binclassif(const uint8_t* classif, const uint8_t* gt, const uint8_t* roi, int w, int h, size_t stride,
                        uint64_t out[6]);

#ifdef __cplusplus
}
#endif
#endif
