// bgs_kernels.hpp -- host-side launchers of the background-subtraction kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "bgs_view.hpp"

namespace lvbgs {

struct StepOptions {
    int fixpoint_launches = 6;   // fix-point iterations launched as full-grid kernels
    int fixpoint_blocks = 256;   // grid of each fix-point kernel (grid-stride over the worklist)
    int tail_cap = 1 << 16;      // iteration cap of the single-workgroup tail (error flag if reached)
    int timed_stage = 0;         // 0: none, 1: the match kernel, 2: the whole step
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
};

// one apply() step; frame is device memory (W*H*C), out_mask optional device W*H
void frame_step(const View& v, const uint8_t* frame, uint8_t* out_mask, double lr, const StepOptions& opt,
                hipStream_t st);
void launch_refresh(const View& v, float frac, int force_fg, uint32_t frame, uint32_t ev, hipStream_t st);
void launch_init_desc(const View& v, const uint8_t* img, hipStream_t st);
void launch_background(const View& v, uint8_t* out_img, uint16_t* out_desc, hipStream_t st);
void launch_lbsp_abs(const uint8_t* img, int W, int H, int C, uint32_t t, uint16_t* out, hipStream_t st);
void launch_rng_probe(const View& v, uint32_t frame, uint32_t phase, uint32_t ev, uint32_t n, uint32_t* out,
                      hipStream_t st);

}  // namespace lvbgs
