// bgs_ctx.hpp -- the per-stream context behind lvbgs_ctx (shared by the C ABI and the server).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/litiv_bgs.h"
#include "bgs_kernels.hpp"

struct lvbgs_ctx {
    int algo = 0;
    lvbgs_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // geometry / state
    bool initialized = false;
    int W = 0, H = 0, C = 0;
    size_t npx = 0, n_roi = 0, orig_roi = 0;
    std::vector<uint8_t> roi;  // validated ROI (host copy)
    int roi_w = 0, roi_h = 0;
    // device
    uint8_t* arena = nullptr;
    size_t arena_size = 0;
    lvbgs::View v{};
    uint8_t* d_frame = nullptr;  // staging for host-pointer apply / initialize
    uint8_t* d_mask = nullptr;
    int ring_frames = 0;         // serving: device slots for frames submitted from host memory
    uint8_t* d_ring = nullptr;   // ring_frames frames of W*H*C bytes (in the arena)
    lvbgs::StepOptions opt{};
    uint32_t apply_count = 0;
    // checkpoint load: the sample planes repacked into the device record layout on the host, held
    // between the geometry set-up and the upload of the scalars
    std::vector<uint32_t> restore_stage;
    // timing
    int timed_stage = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pairs;
    size_t ev_used = 0;
};


// device bytes of the model and per-step buffers of a stream of this geometry (one arena), with
// ring_frames extra frame slots for host-memory ingest
size_t lvbgs_arena_bytes(int algo, const lvbgs_params& p, int width, int height, int channels, int ring_frames = 0);

// initialize from a device frame with the previous full-frame ROI of the same size (or a full frame),
// left queued on the context's stream; `done` is recorded after it
int lvbgs_initialize_async(lvbgs_ctx* c, const void* d_img, int width, int height, int channels, hipEvent_t done);
// the same from a frame in host memory (copied to the context's staging frame on its stream)
int lvbgs_initialize_async_host(lvbgs_ctx* c, const void* img, int width, int height, int channels, size_t stride,
                                hipEvent_t done);
