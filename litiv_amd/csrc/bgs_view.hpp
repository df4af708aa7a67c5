// bgs_view.hpp -- per-stream device state of a background subtractor, passed by value to every
// kernel. All arrays are full-frame, row-major, pixel index p = y*W + x.
#pragma once
#include <stdint.h>

#include "bgs_common.hpp"

namespace lvbgs {

enum Algo : int { kSuBSENSE = 1, kLOBSTER = 2 };

struct View {
    int W, H, C, N, REQ, algo;
    uint32_t npx;            // W*H
    uint32_t n_roi;          // number of ROI pixels
    uint32_t color_thr;      // SuBSENSE m_nMinColorDistThreshold | LOBSTER m_nColorDistThreshold
    uint32_t desc_thr;       // SuBSENSE m_nDescDistThresholdOffset | LOBSTER m_nDescDistThreshold
    uint32_t mv_avgs;        // SuBSENSE m_nSamplesForMovingAvgs
    uint32_t lbsp_offset;    // m_nLBSPThresholdOffset
    float rel_lbsp;          // m_fRelLBSPThreshold
    uint32_t seed_lo, seed_hi, stream_id;
    int use3x3, median_k, lr_scaling, dsW, dsH;
    // per-step inputs
    const uint8_t* frame;    // current input frame, device W*H*C
    uint8_t* out_mask;       // optional device W*H output mask
    double lr;               // learning-rate override of apply()
    uint32_t seq_limit;      // samples of the sequential match phase
    // roi and model
    const uint8_t* roi;      // W*H, 0 = outside
    uint8_t* last_color;     // W*H*C
    uint16_t* last_desc;     // W*H*C
    void* sA;                // N sample planes of one record per pixel: C==3 -> 3 x u32
                             // {d0 | d1<<16, d2 | c0<<16 | c1<<24, c2}; C==1 -> u32 {d0 | c0<<16}
    // SuBSENSE feedback maps (f32 W*H)
    float *T, *R, *V, *MLD, *DminLT, *DminST, *RawLT, *RawST, *FinLT, *FinST;
    // u8 W*H maps
    uint8_t *unstable, *blinks, *last_raw, *last_raw_blink, *last_fg, *last_dil_inv;
    // per-frame scratch
    uint8_t* raw;            // W*H raw segmentation (0/255)
    // hole filling on 64x64 zones (row-major zone grid): bit x of word 64*z+y is pixel (x, y) of
    // zone z; zbits = zero pixels of the closed mask, hbits = holes
    uint64_t *zbits, *hbits;
    uint8_t* zedge;          // 256 per zone: border component of each zone-border pixel, 255 = none
    uint32_t* zpar;          // 128 per zone: union-find parents of the zones' border components
    uint32_t* zcount;        // per zone: number of border components
    uint16_t* desc_cur;      // W*H*C intra descriptors of the current frame
    uint32_t* rec;           // W*H update records
    uint32_t* mres;          // W*H match results
    uint8_t* dec;            // W*H raw decisions (1 = foreground)
    uint32_t* wl0;           // worklists (n_roi entries each)
    uint32_t* wl1;
    uint32_t* cont;          // pixels left by the speculative match to the continuation: stripe k's
                             // entries at k * cont_region (kContStripes stripes)
    uint32_t* cont_ctx;      // match contexts of the first cont_ctx_region entries of each stripe
                             // (kCtxWords each, stripe k at k * cont_ctx_region)
    uint32_t cont_region, cont_ctx_region;
    uint32_t* cont2;         // pixels left by the continuation to the cooperative phase
    unsigned long long* stamp;  // W*H worklist de-duplication stamps
    float *dsLT, *dsST;      // (W/8)*(H/8)*C
    Scalars* sc;
};

}  // namespace lvbgs
