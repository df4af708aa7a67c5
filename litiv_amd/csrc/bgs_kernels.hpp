// bgs_kernels.hpp -- host-side launchers of the background-subtraction kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "bgs_view.hpp"

namespace lvbgs {

// largest sequential-match fetch depth; the sample planes are padded by this many samples so a
// scan may load past a pixel's last sample
constexpr int kMaxFetch = 4;
constexpr int kCtxWords = 20;  // words of a cached match context

struct StepOptions {
    int fixpoint_launches = 2;   // fix-point iterations launched as full-grid kernels
    int fixpoint_blocks = 256;   // grid of each fix-point kernel (grid-stride over the worklist)
    int tail_cap = 1 << 16;      // iteration cap of the single-workgroup tail (error flag if reached)
    int seq_limit = 0;           // samples scanned per pixel by the sequential phase of the match
                                 // (0: the single-stream path's automatic choice)
    int fetch_depth = 4;         // sample loads in flight per lane in the sequential phase (2 or 4)
    int cont_blocks = 1024;      // grid of the cooperative phase (grid-stride over unfinished pixels)
    int fix_px_per_block = 0;    // > 0: fix-point grid of a stream = ceil(npx / this), at most fixpoint_blocks
    int cont_px_per_block = 0;   // > 0: cooperative grid of a stream = ceil(npx / this), at most cont_blocks
    int timed_stage = 0;         // 0: none, 1: the match kernel (phase A), 2: the whole step, 3: the update kernel
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
};

// grid kinds of the frame-step kernels (block counts per stream)
enum GridKind : int {
    kGridTile = 0,   // 64x4 pixel tiles (describe, match, update)
    kGridPost,       // 64x16 tiles (median, dilate)
    kGridZone,       // 64x64 hole-filling zones, four per block
    kGridPix,        // 256 pixels per block
    kGridCell,       // 256 frame-level cells per block (0 when learning-rate scaling is off)
    kGridOne,        // one block per stream
    kGridFix,        // fix-point iteration grid
    kGridCont,       // cooperative match grid
    kGridSweep,      // grid-stride sweep over the pixels (pending model refresh; mostly idle)
    kNumGrids
};
void kernel_grids(const View& v, const StepOptions& opt, uint32_t g[kNumGrids]);
// entries of one stripe of the match continuation list of a W x H stream: the most that the tiles
// appending to one stripe can hand over
uint32_t cont_stripe_slots(int W, int H);

// one apply() step; v.frame is device memory (W*H*C), v.out_mask optional device W*H
void frame_step(const View& v, const StepOptions& opt, hipStream_t st);

// one SuBSENSE step of every stream of a batch (see bgs_kernels.hip)
void batch_step(const View* d_views, int ns, const uint32_t* const prefix[kNumGrids], const uint32_t total[kNumGrids],
                const StepOptions& opt, hipStream_t st);
void launch_refresh(const View& v, float frac, int force_fg, uint32_t frame, uint32_t ev, hipStream_t st);
void launch_init_desc(const View& v, const uint8_t* img, hipStream_t st);
void launch_background(const View& v, uint8_t* out_img, uint16_t* out_desc, hipStream_t st);
void launch_lbsp_abs(const uint8_t* img, int W, int H, int C, uint32_t t, uint16_t* out, bool packed, hipStream_t st);
void launch_postprocess_probe(const View& v, hipStream_t st);
// out[i] = the sticky error bits (Scalars::err_flags) of d_views[i], i < n
void launch_gather_errors(const View* d_views, int n, uint32_t* d_out, hipStream_t st);
void launch_rng_probe(const View& v, uint32_t frame, uint32_t phase, uint32_t ev, uint32_t n, uint32_t* out,
                      hipStream_t st);

}  // namespace lvbgs
