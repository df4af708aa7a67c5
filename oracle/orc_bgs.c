/*
 * orc_bgs.c -- sequential CPU restatement of SuBSENSE and LOBSTER (initialize / apply /
 * refreshModel / getBackgroundImage / getBackgroundDescriptorsImage).
 *
 * TEST INFRASTRUCTURE (oracle): the parity checker of the HIP library and the timed CPU
 * baseline. Not part of the product. Every block cites the reference lines it follows
 * (BackgroundSubtractorSuBSENSE.cpp = "SUB", BackgroundSubtractorLOBSTER.cpp = "LOB",
 * BackgroundSubtractionUtils.cpp = "UTL", BackgroundSubtractorLBSP.cpp = "LBS").
 *
 * The pixel loop runs in raster order exactly like the reference (Gauss-Seidel): samples
 * diffused into a later pixel are seen by that pixel's match in the same frame, and the
 * ghost test reads the neighbour's maps as they are at that point of the loop.
 */
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { PH_APPLY = 1, PH_REFRESH = 2 };

struct orc_bgs {
    int algo;
    orc_params p;
    char err[256];
    /* geometry and ROI (UTL:70-155) */
    int W, H, C;
    size_t npx;
    uint8_t* roi;
    int roi_w, roi_h;
    size_t* px_lut;
    size_t n_roi, orig_roi;
    int initialized;
    /* LBSP base (LBS:21-65) */
    uint8_t lut[256];
    uint8_t* last_color;
    uint16_t* last_desc;
    uint8_t* last_fg;
    /* samples, sample-major: plane s holds npx*C values */
    uint8_t* s_color;
    uint16_t* s_desc;
    /* SuBSENSE maps (SUB:129-173) */
    float *T, *R, *V, *MLD, *DminLT, *DminST, *RawLT, *RawST, *FinLT, *FinST;
    uint8_t *unstable, *blinks, *last_raw, *last_raw_blink, *last_dil_inv;
    float *dsLT, *dsST;
    int dsW, dsH;
    uint8_t* ds_frame;
    /* scratch */
    uint8_t *mask_raw, *preflood, *holes, *tmp, *tmp2;
    /* scalars */
    size_t frame_idx, frames_since_reset, cooldown;
    int auto_reset, lr_scaling, use3x3, median_k;
    float lower_cap, upper_cap, last_nz_ratio;
    uint32_t refresh_seq, apply_count;
    uint64_t cnt_examined, cnt_writes;
};

static int fail(orc_bgs* o, const char* msg) {
    snprintf(o->err, sizeof(o->err), "%s", msg);
    return -1;
}

const char* orc_last_error(const orc_bgs* o) { return o->err; }

void orc_default_params(int algo, orc_params* out) {
    memset(out, 0, sizeof(*out));
    if (algo == ORC_LOBSTER) {
        /* BackgroundSubtractorLOBSTER.hpp:22-31 */
        out->desc_dist_threshold = 4;
        out->color_dist_threshold = 30;
        out->bg_samples = 35;
        out->required_bg_samples = 2;
    } else {
        /* BackgroundSubtractorSuBSENSE.hpp:22-31 */
        out->desc_dist_threshold = 3;
        out->color_dist_threshold = 30;
        out->bg_samples = 50;
        out->required_bg_samples = 2;
        out->samples_for_mv_avgs = 100;
    }
    out->lbsp_threshold_offset = 0;    /* BackgroundSubtractorLBSP.hpp:26 */
    out->rel_lbsp_threshold = 0.333f;  /* BackgroundSubtractorLBSP.hpp:24 */
    out->median_blur_kernel = 9;       /* BackgroundSubtractorLBSP.hpp:28 */
    out->rng_seed = 0xB65Eu;
    out->stream_id = 0;
}

static void free_buffers(orc_bgs* o) {
    void* ptrs[] = {o->px_lut, o->last_color, o->last_desc, o->last_fg, o->s_color, o->s_desc, o->T, o->R, o->V,
                    o->MLD, o->DminLT, o->DminST, o->RawLT, o->RawST, o->FinLT, o->FinST, o->unstable, o->blinks,
                    o->last_raw, o->last_raw_blink, o->last_dil_inv, o->dsLT, o->dsST, o->ds_frame, o->mask_raw,
                    o->preflood, o->holes, o->tmp, o->tmp2};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    o->px_lut = NULL; o->last_color = NULL; o->last_desc = NULL; o->last_fg = NULL; o->s_color = NULL;
    o->s_desc = NULL; o->T = o->R = o->V = o->MLD = o->DminLT = o->DminST = o->RawLT = o->RawST = NULL;
    o->FinLT = o->FinST = NULL; o->unstable = o->blinks = o->last_raw = o->last_raw_blink = NULL;
    o->last_dil_inv = NULL; o->dsLT = o->dsST = NULL; o->ds_frame = NULL; o->mask_raw = o->preflood = NULL;
    o->holes = o->tmp = o->tmp2 = NULL;
}

orc_bgs* orc_create(int algo, const orc_params* p) {
    if (algo != ORC_SUBSENSE && algo != ORC_LOBSTER) return NULL;
    orc_bgs* o = (orc_bgs*)calloc(1, sizeof(orc_bgs));
    o->algo = algo;
    if (p) o->p = *p; else orc_default_params(algo, &o->p);
    /* SUB:76-77, LOB:28-29 */
    if (o->p.bg_samples == 0 || o->p.required_bg_samples > o->p.bg_samples || o->p.bg_samples > 256 ||
        (o->p.color_dist_threshold == 0 && o->p.desc_dist_threshold == 0) || !(o->p.rel_lbsp_threshold >= 0.f) ||
        (algo == ORC_SUBSENSE && o->p.samples_for_mv_avgs < 4)) {
        free(o);
        return NULL;
    }
    return o;
}

void orc_destroy(orc_bgs* o) {
    if (!o) return;
    free_buffers(o);
    free(o->roi);
    free(o);
}

void orc_set_automatic_model_reset(orc_bgs* o, int enabled) { o->auto_reset = enabled ? 1 : 0; }

static uint32_t draw(const orc_bgs* o, uint32_t frame, uint32_t phase, uint32_t event, uint32_t pixel, uint32_t d) {
    return orc_draw(o->p.rng_seed, o->p.stream_id, frame, phase, event, pixel, d);
}

/* the four draws of group g (draws 4g..4g+3) from one Philox call; same values as draw() */
static void draw4(const orc_bgs* o, uint32_t frame, uint32_t phase, uint32_t event, uint32_t pixel, uint32_t g,
                  uint32_t out[4]) {
    const uint32_t key[2] = {(uint32_t)o->p.rng_seed, (uint32_t)(o->p.rng_seed >> 32)};
    const uint32_t ctr[4] = {pixel, frame, (phase << 24) | ((event & 0xFFFFu) << 8) | g, o->p.stream_id};
    orc_philox4x32_10(ctr, key, out);
    for (int i = 0; i < 4; ++i) out[i] >>= 1;
}

/* --------------------------------------------------------------------------------------------
 * refreshModel: SUB:80-105 (descriptor copied from the last descriptor frame) and LOB:410-440
 * (descriptor recomputed from the last colour frame at the sampled position).
 * ------------------------------------------------------------------------------------------ */
static int refresh_model(orc_bgs* o, float frac, int force_fg) {
    if (!o->initialized) return fail(o, "algo must be initialized first");
    if (!(frac > 0.0f && frac <= 1.0f)) return fail(o, "model refresh must be given as a non-null fraction");
    const size_t N = o->p.bg_samples, C = (size_t)o->C;
    const uint32_t ev = o->refresh_seq++;
    const uint32_t frame = (uint32_t)o->apply_count;
    const size_t nref = frac < 1.0f ? (size_t)(frac * (float)N) : N;
    const size_t start = frac < 1.0f ? (size_t)draw(o, frame, PH_REFRESH, ev, 0xFFFFFFFFu, 0) % N : 0;
    for (size_t m = 0; m < o->n_roi; ++m) {
        const size_t px = o->px_lut[m];
        if (!(force_fg || !o->last_fg[px])) continue;
        const int x = (int)(px % (size_t)o->W), y = (int)(px / (size_t)o->W);
        uint32_t dg[4];
        for (size_t j = start; j < start + nref; ++j) {
            const uint32_t k = (uint32_t)(j - start);
            if ((k & 3u) == 0) draw4(o, frame, PH_REFRESH, ev, (uint32_t)px, k >> 2, dg);
            int sx, sy;
            orc_sample_pos_7x7((int)dg[k & 3u], x, y, 2, o->W, o->H, &sx, &sy);
            const size_t spx = (size_t)o->W * (size_t)sy + (size_t)sx;
            if (!(force_fg || !o->last_fg[spx])) continue;
            const size_t slot = j % N;
            for (size_t c = 0; c < C; ++c) {
                if (o->algo == ORC_LOBSTER) {
                    uint8_t vals[16];
                    const uint8_t ref = o->last_color[spx * C + c];
                    orc_lbsp_lookup(o->last_color, o->W, (int)C, sx, sy, (int)c, vals);
                    o->last_desc[spx * C + c] = orc_lbsp_threshold(vals, ref, o->lut[ref]);
                }
                o->s_color[slot * o->npx * C + px * C + c] = o->last_color[spx * C + c];
                o->s_desc[slot * o->npx * C + px * C + c] = o->last_desc[spx * C + c];
            }
        }
    }
    return 0;
}

int orc_refresh_model(orc_bgs* o, float frac, int force_fg) { return refresh_model(o, frac, force_fg); }

/* --------------------------------------------------------------------------------------------
 * initialize: UTL:70-155 (ROI, LUTs, last colour), LBS:21-65 (LBSP LUT + initial descriptors),
 * SUB:107-186 / LOB:442-457.
 * ------------------------------------------------------------------------------------------ */
int orc_initialize(orc_bgs* o, const uint8_t* img, int w, int h, int c, const uint8_t* roi) {
    if (!img || w <= 0 || h <= 0) return fail(o, "provided image for initialization must be non-empty");
    if (c != 1 && c != 3) return fail(o, "provided image for initialization must be of type 8UC1/8UC3");
    const size_t npx = (size_t)w * (size_t)h;
    uint8_t* newroi = (uint8_t*)malloc(npx);
    if (!roi) {
        if (o->roi && o->roi_w == w && o->roi_h == h) memcpy(newroi, o->roi, npx); /* UTL:87-88 reuse */
        else memset(newroi, 255, npx);                                              /* UTL:83-86 */
    } else {
        for (size_t i = 0; i < npx; ++i)
            if (roi[i] != 0 && roi[i] != 255) { free(newroi); return fail(o, "provided ROI mat values must be 0 or 255 only"); }
        /* UTL:92-95: roi | (dilate(roi, 2 iterations)/2); 255*0.5 rounds half-to-even to 128 */
        uint8_t* dil = (uint8_t*)malloc(npx);
        orc_morph(roi, dil, w, h, 2, 1);
        for (size_t i = 0; i < npx; ++i) newroi[i] = roi[i] | (dil[i] ? 128 : 0);
        free(dil);
    }
    size_t orig = 0;
    for (size_t i = 0; i < npx; ++i) orig += newroi[i] ? 1 : 0;
    if (orig == 0) { free(newroi); return fail(o, "provided ROI mat contains no useful pixels"); }
    /* UTL:28-36 validateROI with border LBSP::PATCH_SIZE/2 = 2 */
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            if (x < 2 || y < 2 || x >= w - 2 || y >= h - 2) newroi[(size_t)y * w + x] = 0;
    size_t fin = 0;
    for (size_t i = 0; i < npx; ++i) fin += newroi[i] ? 1 : 0;
    if (fin == 0) { free(newroi); return fail(o, "provided ROI mat contains no useful pixels away from borders"); }

    free_buffers(o);
    free(o->roi);
    o->roi = newroi; o->roi_w = w; o->roi_h = h;
    o->W = w; o->H = h; o->C = c; o->npx = npx;
    o->orig_roi = orig; o->n_roi = fin;
    o->initialized = 0;
    o->frame_idx = 0; o->frames_since_reset = 0; o->cooldown = 0;
    o->apply_count = 0; o->refresh_seq = 0;
    const size_t C = (size_t)c, N = o->p.bg_samples;
    o->px_lut = (size_t*)malloc(sizeof(size_t) * fin);
    o->last_fg = (uint8_t*)calloc(npx, 1);
    o->last_color = (uint8_t*)calloc(npx * C, 1);
    for (size_t i = 0, m = 0; i < npx; ++i)
        if (newroi[i]) {
            o->px_lut[m++] = i;
            for (size_t ch = 0; ch < C; ++ch) o->last_color[i * C + ch] = img[i * C + ch];
        }
    /* LBS:21-65 */
    o->last_desc = (uint16_t*)calloc(npx * C, sizeof(uint16_t));
    orc_build_lut(c, o->p.rel_lbsp_threshold, o->p.lbsp_threshold_offset, o->lut);
    for (size_t i = 0; i < npx; ++i) {
        const int x = (int)(i % (size_t)w), y = (int)(i / (size_t)w);
        /* strict bounds: column x=2 and row y=2 keep descriptor 0 (LBS:34,47) */
        if (newroi[i] && x > 2 && y > 2 && x < w - 2 && y < h - 2) {
            uint8_t vals[16];
            for (size_t ch = 0; ch < C; ++ch) {
                orc_lbsp_lookup(img, w, c, x, y, (int)ch, vals);
                const uint8_t ref = img[i * C + ch];
                o->last_desc[i * C + ch] = orc_lbsp_threshold(vals, ref, o->lut[ref]);
            }
        }
    }
    o->s_color = (uint8_t*)calloc(N * npx * C, 1);
    o->s_desc = (uint16_t*)calloc(N * npx * C, sizeof(uint16_t));
    o->mask_raw = (uint8_t*)calloc(npx, 1);
    o->tmp = (uint8_t*)calloc(npx, 1);
    if (o->algo == ORC_SUBSENSE) {
        o->last_nz_ratio = 0.0f;
        /* SUB:112-128 */
        if (o->orig_roi >= npx / 2 && (int)npx >= 320 * 240) {
            o->lr_scaling = 1;
            o->auto_reset = 1;
            o->use3x3 = !((int)npx > 320 * 240 * 2);
            const int raw = (int)floorf((float)npx / (float)(320 * 240) + 0.5f) + o->p.median_blur_kernel;
            const int rk = raw < 14 ? raw : 14;
            o->median_k = (rk % 2) ? rk : rk - 1;
            o->lower_cap = 2.0f;
            o->upper_cap = 256.0f;
        } else {
            o->lr_scaling = 0;
            o->auto_reset = 0;
            o->use3x3 = 1;
            o->median_k = o->p.median_blur_kernel;
            o->lower_cap = 2.0f * 2;
            o->upper_cap = 256.0f * 2;
        }
        float** maps[] = {&o->T, &o->R, &o->V, &o->MLD, &o->DminLT, &o->DminST, &o->RawLT, &o->RawST, &o->FinLT, &o->FinST};
        for (size_t k = 0; k < 10; ++k) *maps[k] = (float*)calloc(npx, sizeof(float));
        for (size_t i = 0; i < npx; ++i) { o->T[i] = o->lower_cap; o->R[i] = 1.0f; o->V[i] = 10.0f; }
        o->dsW = w / 8; o->dsH = h / 8;
        const size_t nds = (size_t)o->dsW * (size_t)o->dsH * C;
        o->dsLT = (float*)calloc(nds ? nds : 1, sizeof(float));
        o->dsST = (float*)calloc(nds ? nds : 1, sizeof(float));
        o->ds_frame = (uint8_t*)calloc(nds ? nds : 1, 1);
        o->unstable = (uint8_t*)calloc(npx, 1);
        o->blinks = (uint8_t*)calloc(npx, 1);
        o->last_raw = (uint8_t*)calloc(npx, 1);
        o->last_raw_blink = (uint8_t*)calloc(npx, 1);
        o->last_dil_inv = (uint8_t*)calloc(npx, 1);
        o->preflood = (uint8_t*)calloc(npx, 1);
        o->holes = (uint8_t*)calloc(npx, 1);
        o->tmp2 = (uint8_t*)calloc(npx, 1);
        o->initialized = 1;
        return refresh_model(o, 1.0f, 0); /* SUB:184 */
    }
    o->median_k = o->p.median_blur_kernel;
    o->initialized = 1;
    return refresh_model(o, 1.0f, 1); /* LOB:455 */
}

/* --------------------------------------------------------------------------------------------
 * SuBSENSE apply: SUB:188-612
 * ------------------------------------------------------------------------------------------ */
static inline float fminf_ref(float a, float b) { return (b < a) ? b : a; } /* std::min */
static inline float fmaxf_ref(float a, float b) { return (a < b) ? b : a; } /* std::max */
static inline uint64_t umod(uint64_t a, uint64_t b) { return a % b; }

static void write_sample(orc_bgs* o, size_t slot, size_t px, const uint8_t* col, const uint16_t* desc) {
    const size_t C = (size_t)o->C;
    for (size_t c = 0; c < C; ++c) {
        o->s_color[slot * o->npx * C + px * C + c] = col[c];
        o->s_desc[slot * o->npx * C + px * C + c] = desc[c];
    }
    o->cnt_writes++;
}

static int subsense_apply(orc_bgs* o, const uint8_t* img, uint8_t* mask_out, double lr) {
    const size_t C = (size_t)o->C, N = o->p.bg_samples, REQ = o->p.required_bg_samples, npx = o->npx;
    const int W = o->W, H = o->H;
    uint8_t* mask = o->mask_raw;
    memset(mask, 0, npx); /* SUB:196 */
    size_t nonzero = 0;
    ++o->frame_idx;
    ++o->apply_count;
    const uint32_t frame = o->apply_count;
    const size_t mv = o->p.samples_for_mv_avgs;
    const float aLT = 1.0f / (float)(o->frame_idx < mv ? o->frame_idx : mv);             /* SUB:198 */
    const float aST = 1.0f / (float)(o->frame_idx < mv / 4 ? o->frame_idx : mv / 4);     /* SUB:199 */
    const size_t minc = o->p.color_dist_threshold, doff = o->p.desc_dist_threshold;
    const size_t stab_off = minc / 5;                   /* STAB_COLOR_DIST_OFFSET, SUB:53 */
    const size_t desc_max = C * 16, col_max = C * 255;  /* s_n*MaxDataRange_*ch, SUB:57-60 */
    const int isinf_lr = isinf(lr) ? 1 : 0;
    const uint64_t lower_cap_sz = (uint64_t)o->lower_cap;
    o->cnt_examined = 0; o->cnt_writes = 0;

    for (size_t m = 0; m < o->n_roi; ++m) {
        const size_t px = o->px_lut[m];
        const int x = (int)(px % (size_t)W), y = (int)(px / (size_t)W);
        const uint8_t* cur = img + px * C;
        float* pR = &o->R[px]; float* pV = &o->V[px]; float* pT = &o->T[px];
        float* pMLD = &o->MLD[px]; float* pDLT = &o->DminLT[px]; float* pDST = &o->DminST[px];
        float* pRLT = &o->RawLT[px]; float* pRST = &o->RawST[px];
        const float fLT = o->FinLT[px], fST = o->FinST[px];
        /* SUB:355-359 (3ch) / :222-223 (1ch): thresholds from the previous unstable flag */
        const int unst_old = o->unstable[px] ? 1 : 0;
        const float colf = (*pR) * (float)minc - (float)((size_t)(!unst_old) * stab_off);
        size_t col_thr = (size_t)colf;
        if (C == 1) col_thr /= 2;
        const size_t shift = (size_t)floorf(*pR + 0.5f);
        const size_t desc_thr = ((size_t)1 << (shift & 63)) + doff + (size_t)unst_old * doff;
        const size_t tot_col_thr = col_thr * 3, tot_desc_thr = desc_thr * 3, sc_col_thr = tot_col_thr / 2;
        /* SUB:360-364 / :224-226: LBSP taps and intra descriptors */
        uint8_t vals[3][16];
        uint16_t intra[3];
        for (size_t c = 0; c < C; ++c) {
            orc_lbsp_lookup(img, W, (int)C, x, y, (int)c, vals[c]);
            intra[c] = orc_lbsp_threshold(vals[c], cur[c], o->lut[cur[c]]);
        }
        /* SUB:365 / :227 */
        const int unst_new = ((*pR) > 3.0f || ((*pRLT) - fLT) > 0.1f || ((*pRST) - fST) > 0.1f) ? 1 : 0;
        o->unstable[px] = (uint8_t)unst_new;
        /* SUB:366-395 / :228-253: ordered sample scan */
        size_t good = 0, s = 0;
        size_t min_desc = desc_max, min_sum = col_max;
        while (good < REQ && s < N) {
            const uint8_t* bc = o->s_color + s * npx * C + px * C;
            const uint16_t* bd = o->s_desc + s * npx * C + px * C;
            int ok = 1;
            if (C == 1) {
                const size_t cd = (size_t)(cur[0] > bc[0] ? cur[0] - bc[0] : bc[0] - cur[0]);
                if (cd > col_thr) ok = 0;
                else {
                    const size_t intra_d = (size_t)__builtin_popcount((unsigned)(intra[0] ^ bd[0]));
                    const uint16_t inter = orc_lbsp_threshold(vals[0], bc[0], o->lut[bc[0]]);
                    const size_t inter_d = (size_t)__builtin_popcount((unsigned)(inter ^ bd[0]));
                    const size_t dd = (intra_d + inter_d) / 2;
                    if (dd > desc_thr) ok = 0;
                    else {
                        size_t sum = (dd / 4) * (255 / 16) + cd;
                        if (sum > 255) sum = 255;
                        if (sum > col_thr) ok = 0;
                        else {
                            if (min_desc > dd) min_desc = dd;
                            if (min_sum > sum) min_sum = sum;
                            ++good;
                        }
                    }
                }
            } else {
                size_t tot_desc = 0, tot_sum = 0;
                for (size_t c = 0; c < 3; ++c) {
                    const size_t cd = (size_t)(cur[c] > bc[c] ? cur[c] - bc[c] : bc[c] - cur[c]);
                    if (cd > sc_col_thr) { ok = 0; break; }
                    const size_t intra_d = (size_t)__builtin_popcount((unsigned)(intra[c] ^ bd[c]));
                    const uint16_t inter = orc_lbsp_threshold(vals[c], bc[c], o->lut[bc[c]]);
                    const size_t inter_d = (size_t)__builtin_popcount((unsigned)(inter ^ bd[c]));
                    const size_t dd = (intra_d + inter_d) / 2;
                    size_t sum = (dd / 2) * (255 / 16) + cd;
                    if (sum > 255) sum = 255;
                    if (sum > sc_col_thr) { ok = 0; break; }
                    tot_desc += dd;
                    tot_sum += sum;
                }
                if (ok && (tot_desc > tot_desc_thr || tot_sum > tot_col_thr)) ok = 0;
                if (ok) {
                    if (min_desc > tot_desc) min_desc = tot_desc;
                    if (min_sum > tot_sum) min_sum = tot_sum;
                    ++good;
                }
            }
            ++s;
        }
        o->cnt_examined += s;
        /* SUB:396-397 / :254-255 */
        uint8_t* lc = o->last_color + px * C;
        uint16_t* ld = o->last_desc + px * C;
        float nld;
        if (C == 1) {
            const uint8_t l1 = (uint8_t)(lc[0] > cur[0] ? lc[0] - cur[0] : cur[0] - lc[0]);
            const uint8_t hd = (uint8_t)__builtin_popcount((unsigned)(ld[0] ^ intra[0]));
            nld = ((float)l1 / 255.0f + (float)hd / 16.0f) / 2.0f;
        } else {
            const uint8_t l1 = orc_l1dist3_u8(lc, cur);
            uint8_t hd = 0;
            for (size_t c = 0; c < 3; ++c) hd = (uint8_t)(hd + __builtin_popcount((unsigned)(ld[c] ^ intra[c])));
            nld = ((float)l1 / (float)col_max + (float)hd / (float)desc_max) / 2.0f;
        }
        *pMLD = (*pMLD) * (1.0f - aST) + nld * aST;
        uint32_t dg[4];
        draw4(o, frame, PH_APPLY, 0, (uint32_t)px, 0, dg);
        const uint32_t d0 = dg[0], d1 = dg[1];
        if (good < REQ) {
            /* SUB:398-413 foreground */
            const float nmd = fminf_ref(1.0f, ((float)min_sum / (float)col_max + (float)min_desc / (float)desc_max) / 2.0f +
                                                  (float)(REQ - good) / (float)REQ);
            *pDLT = (*pDLT) * (1.0f - aLT) + nmd * aLT;
            *pDST = (*pDST) * (1.0f - aST) + nmd * aST;
            *pRLT = (*pRLT) * (1.0f - aLT) + aLT;
            *pRST = (*pRST) * (1.0f - aST) + aST;
            mask[px] = 255;
            if (o->cooldown && umod(d0, 2) == 0) write_sample(o, umod(d1, N), px, cur, intra);
        } else {
            /* SUB:414-450 background */
            const float nmd = ((float)min_sum / (float)col_max + (float)min_desc / (float)desc_max) / 2.0f;
            *pDLT = (*pDLT) * (1.0f - aLT) + nmd * aLT;
            *pDST = (*pDST) * (1.0f - aST) + nmd * aST;
            *pRLT = (*pRLT) * (1.0f - aLT);
            *pRST = (*pRST) * (1.0f - aST);
            const uint64_t L = isinf_lr ? UINT64_MAX : (lr > 0 ? (uint64_t)ceil(lr) : (uint64_t)ceilf(*pT));
            if (umod(d0, L) == 0) write_sample(o, umod(d1, N), px, cur, intra);
            const int use3 = o->use3x3 && !unst_new;
            const uint32_t d2 = dg[2];
            int nx, ny;
            if (use3) orc_neighbor_pos_3x3((int)d2, x, y, 2, W, H, &nx, &ny);
            else orc_neighbor_pos_5x5((int)d2, x, y, 2, W, H, &nx, &ny);
            const uint64_t n_rand = dg[3];
            const size_t q = (size_t)W * (size_t)ny + (size_t)nx;
            const float qMLD = o->MLD[q], qRST = o->RawST[q];
            if (umod(n_rand, use3 ? L : L / 2 + 1) == 0 ||
                (qRST > 0.995f && qMLD < 0.010f && umod(n_rand, lower_cap_sz) == 0)) {
                const uint32_t d4 = draw(o, frame, PH_APPLY, 0, (uint32_t)px, 4);
                write_sample(o, umod(d4, N), q, cur, intra);
            }
        }
        /* SUB:451-474 / :302-325 feedback */
        const float mn = fminf_ref(*pDLT, *pDST), mx = fmaxf_ref(*pDLT, *pDST);
        if (o->last_fg[px] || (mn < 0.1f && mask[px])) {
            if ((*pT) < o->upper_cap) *pT += 0.5f / (mx * (*pV));
        } else if ((*pT) > o->lower_cap) {
            *pT -= 0.25f * (*pV) / mx;
        }
        if ((*pT) < o->lower_cap) *pT = o->lower_cap;
        else if ((*pT) > o->upper_cap) *pT = o->upper_cap;
        if (mx > 0.1f && o->blinks[px]) {
            *pV += 1.0f;
        } else if ((*pV) > 0.1f) {
            *pV -= o->last_fg[px] ? 0.1f / 4 : unst_new ? 0.1f / 2 : 0.1f;
            if ((*pV) < 0.1f) *pV = 0.1f;
        }
        const float base = 1.0f + mn * 2.0f;
        if ((double)(*pR) < (double)base * (double)base) { /* std::pow(float,2) promoted to double: exact square */
            *pR += 0.01f * ((*pV) - 0.1f);
        } else {
            *pR -= 0.01f / (*pV);
            if ((*pR) < 1.0f) *pR = 1.0f;
        }
        /* SUB:475-480 / :326-329 */
        unsigned pc = 0;
        for (size_t c = 0; c < C; ++c) pc += (unsigned)__builtin_popcount((unsigned)intra[c]);
        if (pc >= (C == 1 ? 2u : 4u)) ++nonzero;
        for (size_t c = 0; c < C; ++c) { ld[c] = intra[c]; lc[c] = cur[c]; }
    }

    /* SUB:536-554 post-processing, whole image */
    uint8_t* currblink = o->tmp2;
    for (size_t i = 0; i < npx; ++i) {
        currblink[i] = mask[i] ^ o->last_raw[i];
        o->blinks[i] = currblink[i] | o->last_raw_blink[i];
        o->last_raw_blink[i] = currblink[i];
        o->last_raw[i] = mask[i];
    }
    orc_morph(mask, o->tmp, W, H, 1, 1);           /* close = erode(dilate) with 3x3 */
    orc_morph(o->tmp, o->preflood, W, H, 1, 0);
    orc_flood_holes(o->preflood, o->holes, W, H);
    orc_morph(o->preflood, o->tmp, W, H, 3, 0);    /* erode, 3 iterations */
    for (size_t i = 0; i < npx; ++i) mask[i] = mask[i] | o->holes[i] | o->tmp[i];
    orc_median_binary(mask, o->last_fg, W, H, o->median_k);
    orc_morph(o->last_fg, o->tmp, W, H, 3, 1);     /* dilate, 3 iterations */
    for (size_t i = 0; i < npx; ++i) {
        o->blinks[i] &= o->last_dil_inv[i];        /* stale inverted dilation (SUB:549) */
        o->last_dil_inv[i] = (uint8_t)~o->tmp[i];
        o->blinks[i] &= o->last_dil_inv[i];
    }
    if (mask_out) memcpy(mask_out, o->last_fg, npx);
    /* addWeighted(..., CV_32F): defined as the double-precision scalar form
     * dst = (float)(src1*alpha + src2*beta + gamma), alpha=(double)(1-a), beta=(1.0/255)*a, gamma=0 */
    {
        const double alt = (double)(1.0f - aLT), blt = (1.0 / 255) * (double)aLT;
        const double ast = (double)(1.0f - aST), bst = (1.0 / 255) * (double)aST;
        for (size_t i = 0; i < npx; ++i) {
            const double f = (double)o->last_fg[i];
            o->FinLT[i] = (float)((double)o->FinLT[i] * alt + f * blt + 0.0);
            o->FinST[i] = (float)((double)o->FinST[i] * ast + f * bst + 0.0);
        }
    }
    /* SUB:555-566 LBSP LUT adaptation */
    const float nz_ratio = (float)nonzero / (float)o->n_roi;
    if (nz_ratio < 0.1f && o->last_nz_ratio < 0.1f) {
        for (size_t t = 0; t <= 255; ++t) {
            const float lb = (float)o->p.lbsp_threshold_offset + ceilf((float)t * o->p.rel_lbsp_threshold / 4.0f);
            if (o->lut[t] > orc_sat_u8_f(lb)) --o->lut[t];
        }
    } else if (nz_ratio > 0.5f && o->last_nz_ratio > 0.5f) {
        const uint8_t ub = orc_sat_u8_f((float)o->p.lbsp_threshold_offset + 255.0f * o->p.rel_lbsp_threshold);
        for (size_t t = 0; t <= 255; ++t)
            if (o->lut[t] < ub) ++o->lut[t];
    }
    o->last_nz_ratio = nz_ratio;
    /* SUB:567-611 frame-level analysis */
    if (o->lr_scaling) {
        orc_resize_area(img, W, H, (int)C, o->ds_frame, o->dsW, o->dsH);
        const float aL = aLT, bL = 1.0f - aL, aS = aST, bS = 1.0f - aS; /* accumulateWeighted: dst = src*a + dst*(1-a) */
        const size_t nds = (size_t)o->dsW * (size_t)o->dsH;
        size_t tot = 0;
        for (size_t i = 0; i < nds * C; ++i) {
            o->dsLT[i] = (float)o->ds_frame[i] * aL + o->dsLT[i] * bL;
            o->dsST[i] = (float)o->ds_frame[i] * aS + o->dsST[i] * bS;
        }
        for (size_t i = 0; i < nds; ++i) {
            if (C == 1) {
                tot += (size_t)fabsf(o->dsST[i] - o->dsLT[i]) / 2;
            } else {
                size_t best = 0;
                for (size_t c = 0; c < 3; ++c) {
                    const size_t v = (size_t)fabsf(o->dsST[i * 3 + c] - o->dsLT[i * 3 + c]);
                    if (v > best) best = v;
                }
                tot += best;
            }
        }
        const float ratio = (float)tot / (float)(o->dsW * o->dsH);
        const size_t thr = minc / 2; /* FRAMELEVEL_MIN_COLOR_DIFF_THRESHOLD (integer) */
        if (o->auto_reset) {
            if (o->frames_since_reset > 1000) {
                o->auto_reset = 0;
            } else if (ratio >= (float)thr && o->cooldown == 0) {
                o->frames_since_reset = 0;
                refresh_model(o, 0.1f, 0);
                o->cooldown = mv / 4;
                for (size_t i = 0; i < npx; ++i) o->T[i] = 1.0f;
            } else {
                ++o->frames_since_reset;
            }
        } else if (ratio >= (float)(thr * 2)) {
            o->frames_since_reset = 0;
            o->auto_reset = 1;
        }
        if (ratio >= (float)(thr / 2)) {
            /* (int)2 >> (int)(ratio/2): the reference's x86 build shifts by the count mod 32 */
            const int k = (int)(ratio / 2.0f);
            const int lo = 2 >> (k & 31), hi = 256 >> (k & 31);
            o->lower_cap = (float)(lo > 1 ? lo : 1);
            o->upper_cap = (float)(hi > 1 ? hi : 1);
        } else {
            o->lower_cap = 2.0f;
            o->upper_cap = 256.0f;
        }
        if (o->cooldown > 0) --o->cooldown;
    }
    return 0;
}

/* --------------------------------------------------------------------------------------------
 * LOBSTER apply: LOB:459-581
 * ------------------------------------------------------------------------------------------ */
static int lobster_apply(orc_bgs* o, const uint8_t* img, uint8_t* mask_out, double lr) {
    if (!(lr > 0)) return fail(o, "learning rate must be a positive value; faster learning is achieved with smaller values");
    const size_t C = (size_t)o->C, N = o->p.bg_samples, REQ = o->p.required_bg_samples, npx = o->npx;
    const int W = o->W, H = o->H;
    uint8_t* mask = o->mask_raw;
    memset(mask, 0, npx);
    ++o->apply_count;
    const uint32_t frame = o->apply_count;
    const uint64_t L = isinf(lr) ? UINT64_MAX : (uint64_t)ceil(lr);
    const size_t dthr = o->p.desc_dist_threshold, cthr = o->p.color_dist_threshold;
    const size_t tot_d = dthr * 3, tot_c = cthr * 3, sc_d = tot_d / 2, sc_c = tot_c / 2;
    o->cnt_examined = 0; o->cnt_writes = 0;
    for (size_t m = 0; m < o->n_roi; ++m) {
        const size_t px = o->px_lut[m];
        const int x = (int)(px % (size_t)W), y = (int)(px / (size_t)W);
        const uint8_t* cur = img + px * C;
        uint8_t vals[3][16];
        for (size_t c = 0; c < C; ++c) orc_lbsp_lookup(img, W, (int)C, x, y, (int)c, vals[c]);
        size_t good = 0, s = 0;
        while (good < REQ && s < N) {
            const uint8_t* bc = o->s_color + s * npx * C + px * C;
            const uint16_t* bd = o->s_desc + s * npx * C + px * C;
            if (C == 1) {
                const size_t cd = (size_t)(cur[0] > bc[0] ? cur[0] - bc[0] : bc[0] - cur[0]);
                if (cd <= cthr / 2) {
                    const uint16_t inter = orc_lbsp_threshold(vals[0], bc[0], o->lut[bc[0]]);
                    if ((size_t)__builtin_popcount((unsigned)(inter ^ bd[0])) <= dthr) ++good;
                }
            } else {
                size_t tc = 0, td = 0;
                int ok = 1;
                for (size_t c = 0; c < 3; ++c) {
                    const size_t cd = (size_t)(cur[c] > bc[c] ? cur[c] - bc[c] : bc[c] - cur[c]);
                    if (cd > sc_c) { ok = 0; break; }
                    const uint16_t inter = orc_lbsp_threshold(vals[c], bc[c], o->lut[bc[c]]);
                    const size_t dd = (size_t)__builtin_popcount((unsigned)(inter ^ bd[c]));
                    if (dd > sc_d) { ok = 0; break; }
                    tc += cd;
                    td += dd;
                }
                if (ok && td <= tot_d && tc <= tot_c) ++good;
            }
            ++s;
        }
        o->cnt_examined += s;
        if (good < REQ) {
            mask[px] = 255;
            continue;
        }
        uint16_t intra[3];
        for (size_t c = 0; c < C; ++c) intra[c] = orc_lbsp_threshold(vals[c], cur[c], o->lut[cur[c]]);
        const uint32_t d0 = draw(o, frame, PH_APPLY, 0, (uint32_t)px, 0);
        if (umod(d0, L) == 0) write_sample(o, umod(draw(o, frame, PH_APPLY, 0, (uint32_t)px, 1), N), px, cur, intra);
        const uint32_t d2 = draw(o, frame, PH_APPLY, 0, (uint32_t)px, 2);
        if (umod(d2, L) == 0) {
            int nx, ny;
            orc_neighbor_pos_3x3((int)draw(o, frame, PH_APPLY, 0, (uint32_t)px, 3), x, y, 2, W, H, &nx, &ny);
            const size_t q = (size_t)W * (size_t)ny + (size_t)nx;
            write_sample(o, umod(draw(o, frame, PH_APPLY, 0, (uint32_t)px, 4), N), q, cur, intra);
        }
    }
    orc_median_binary(mask, o->last_fg, W, H, o->median_k); /* LOB:578 */
    if (mask_out) memcpy(mask_out, o->last_fg, npx);
    memcpy(o->last_color, img, npx * C);                    /* LOB:580 */
    return 0;
}

int orc_apply(orc_bgs* o, const uint8_t* img, uint8_t* mask_out, double lr) {
    if (!o->initialized) return fail(o, "algo & model must be initialized first");
    if (!img) return fail(o, "input image type/size mismatch with initialization type/size");
    return o->algo == ORC_SUBSENSE ? subsense_apply(o, img, mask_out, lr) : lobster_apply(o, img, mask_out, lr);
}

/* SUB:614-630 / LOB:583-600: float accumulation in sample order, then saturate_cast<uchar> */
int orc_get_background_image(const orc_bgs* o, uint8_t* out) {
    if (!o->initialized) return -1;
    const size_t C = (size_t)o->C, N = o->p.bg_samples;
    for (size_t i = 0; i < o->npx * C; ++i) {
        float acc = 0.0f;
        for (size_t s = 0; s < N; ++s) acc += (float)o->s_color[s * o->npx * C + i] / (float)N;
        out[i] = orc_sat_u8_f(acc);
    }
    return 0;
}

/* SUB:632-649 / LOB:602-620: same with 16-bit descriptors and saturate_cast<ushort> */
int orc_get_background_descriptors_image(const orc_bgs* o, uint16_t* out) {
    if (!o->initialized) return -1;
    const size_t C = (size_t)o->C, N = o->p.bg_samples;
    for (size_t i = 0; i < o->npx * C; ++i) {
        float acc = 0.0f;
        for (size_t s = 0; s < N; ++s) acc += (float)o->s_desc[s * o->npx * C + i] / (float)N;
        const long r = lrintf(acc);
        out[i] = (uint16_t)(r < 0 ? 0 : r > 65535 ? 65535 : r);
    }
    return 0;
}

static long long copy_out(void* out, size_t cap, const void* src, size_t bytes) {
    if (!src) return -2;
    if (bytes > cap) return -3;
    memcpy(out, src, bytes);
    return (long long)bytes;
}

long long orc_state(const orc_bgs* o, const char* name, void* out, size_t cap) {
    const size_t npx = o->npx, C = (size_t)o->C, N = o->p.bg_samples;
    const size_t fb = npx * sizeof(float);
    if (!strcmp(name, "samples_color")) return copy_out(out, cap, o->s_color, N * npx * C);
    if (!strcmp(name, "samples_desc")) return copy_out(out, cap, o->s_desc, N * npx * C * 2);
    if (!strcmp(name, "last_color")) return copy_out(out, cap, o->last_color, npx * C);
    if (!strcmp(name, "last_desc")) return copy_out(out, cap, o->last_desc, npx * C * 2);
    if (!strcmp(name, "last_fg")) return copy_out(out, cap, o->last_fg, npx);
    if (!strcmp(name, "roi")) return copy_out(out, cap, o->roi, npx);
    if (!strcmp(name, "lut")) return copy_out(out, cap, o->lut, 256);
    if (!strcmp(name, "T")) return copy_out(out, cap, o->T, fb);
    if (!strcmp(name, "R")) return copy_out(out, cap, o->R, fb);
    if (!strcmp(name, "V")) return copy_out(out, cap, o->V, fb);
    if (!strcmp(name, "MLD")) return copy_out(out, cap, o->MLD, fb);
    if (!strcmp(name, "DminLT")) return copy_out(out, cap, o->DminLT, fb);
    if (!strcmp(name, "DminST")) return copy_out(out, cap, o->DminST, fb);
    if (!strcmp(name, "RawLT")) return copy_out(out, cap, o->RawLT, fb);
    if (!strcmp(name, "RawST")) return copy_out(out, cap, o->RawST, fb);
    if (!strcmp(name, "FinLT")) return copy_out(out, cap, o->FinLT, fb);
    if (!strcmp(name, "FinST")) return copy_out(out, cap, o->FinST, fb);
    if (!strcmp(name, "unstable")) return copy_out(out, cap, o->unstable, npx);
    if (!strcmp(name, "blinks")) return copy_out(out, cap, o->blinks, npx);
    if (!strcmp(name, "last_raw")) return copy_out(out, cap, o->last_raw, npx);
    if (!strcmp(name, "last_raw_blink")) return copy_out(out, cap, o->last_raw_blink, npx);
    if (!strcmp(name, "last_dil_inv")) return copy_out(out, cap, o->last_dil_inv, npx);
    if (!strcmp(name, "dsLT")) return copy_out(out, cap, o->dsLT, (size_t)o->dsW * o->dsH * C * sizeof(float));
    if (!strcmp(name, "dsST")) return copy_out(out, cap, o->dsST, (size_t)o->dsW * o->dsH * C * sizeof(float));
    return -1;
}

int orc_scalars(const orc_bgs* o, double* out, int cap) {
    const double v[] = {(double)o->frame_idx, (double)o->frames_since_reset, (double)o->cooldown, (double)o->auto_reset,
                        (double)o->lr_scaling, (double)o->use3x3, (double)o->median_k, (double)o->n_roi,
                        (double)o->lower_cap, (double)o->upper_cap, (double)o->last_nz_ratio, (double)o->refresh_seq,
                        (double)o->apply_count};
    const int n = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < cap; ++i) out[i] = v[i];
    return n;
}

void orc_counters(const orc_bgs* o, uint64_t* examined, uint64_t* writes) {
    if (examined) *examined = o->cnt_examined;
    if (writes) *writes = o->cnt_writes;
}
