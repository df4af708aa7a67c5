// bgs_api.hip -- C ABI (include/litiv_bgs.h): per-stream context, device memory layout,
// initialize / apply / getters. Host-side logic mirrors the reference's IIBackgroundSubtractor
// (BackgroundSubtractionUtils.cpp) and the constructors/initialize of SuBSENSE and LOBSTER.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/litiv_bgs.h"
#include "bgs_ctx.hpp"
#include "bgs_kernels.hpp"

using namespace lvbgs;

namespace {

__global__ void k_fill_f32(float* p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct Arena {
    uint8_t* base = nullptr;
    size_t size = 0, off = 0;
    template <typename T>
    T* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;  // sizing pass: offsets only
        off += count * sizeof(T);
        return p;
    }
};

uint8_t sat_u8_host(float v) {
    const long r = lrintf(v);  // cvRound: round half to even
    return (uint8_t)(r < 0 ? 0 : r > 255 ? 255 : r);
}

}  // namespace

static int set_err(lvbgs_ctx* c, int code, const char* msg) {
    if (c) c->err = msg;
    return code;
}

static int hip_check(lvbgs_ctx* c, hipError_t e, const char* where) {
    if (e == hipSuccess) return LVBGS_OK;
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
    return set_err(c, LVBGS_ERR_HIP, buf);
}

extern "C" int lvbgs_default_params(int algo, lvbgs_params* out) {
    if (!out || (algo != LVBGS_SUBSENSE && algo != LVBGS_LOBSTER)) return LVBGS_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    if (algo == LVBGS_SUBSENSE) {  // BackgroundSubtractorSuBSENSE.hpp:22-31
        out->desc_dist_threshold = 3;
        out->color_dist_threshold = 30;
        out->bg_samples = 50;
        out->required_bg_samples = 2;
        out->samples_for_mv_avgs = 100;
    } else {  // BackgroundSubtractorLOBSTER.hpp:22-31
        out->desc_dist_threshold = 4;
        out->color_dist_threshold = 30;
        out->bg_samples = 35;
        out->required_bg_samples = 2;
    }
    out->lbsp_threshold_offset = 0;
    out->rel_lbsp_threshold = 0.333f;
    out->median_blur_kernel = 9;
    out->rng_seed = 0xB65Eu;
    out->stream_id = 0;
    out->device = 0;
    return LVBGS_OK;
}

extern "C" int lvbgs_create(int algo, const lvbgs_params* params, lvbgs_ctx** out) {
    if (!out) return LVBGS_ERR_ARG;
    *out = nullptr;
    lvbgs_params p;
    if (lvbgs_default_params(algo, &p) != LVBGS_OK) return LVBGS_ERR_ARG;
    if (params) p = *params;
    // SUB:76-77, LOB:28-29, plus this build's limits (8-bit slot indices, median halo of 6)
    if (p.bg_samples == 0 || p.bg_samples > 255 || p.required_bg_samples > p.bg_samples ||
        (p.color_dist_threshold == 0 && p.desc_dist_threshold == 0) || !(p.rel_lbsp_threshold >= 0.f) ||
        (algo == LVBGS_SUBSENSE && p.samples_for_mv_avgs < 4) || p.median_blur_kernel < 1 ||
        p.median_blur_kernel > 13 || (p.median_blur_kernel % 2) == 0 || p.color_dist_threshold > 0xFFFFFF ||
        p.desc_dist_threshold > 0xFFFFFF || p.lbsp_threshold_offset > 255)
        return LVBGS_ERR_ARG;
    lvbgs_ctx* c = new lvbgs_ctx();
    c->algo = algo;
    c->p = p;
    c->device = p.device;
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return LVBGS_ERR_HIP;
    }
    *out = c;
    return LVBGS_OK;
}

static void release_device(lvbgs_ctx* c) {
    if (c->arena) (void)hipFree(c->arena);
    c->arena = nullptr;
    c->arena_size = 0;
}

extern "C" void lvbgs_destroy(lvbgs_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& e : c->ev_pairs) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    release_device(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* lvbgs_last_error(const lvbgs_ctx* c) { return c ? c->err.c_str() : "null context"; }
extern "C" void* lvbgs_stream(lvbgs_ctx* c) { return c ? (void*)c->stream : nullptr; }
extern "C" double lvbgs_default_learning_rate(const lvbgs_ctx* c) {
    return (c && c->algo == LVBGS_LOBSTER) ? 16.0 : 0.0;
}

extern "C" int lvbgs_validate_roi(int w, int h, uint8_t* roi, size_t stride) {
    if (!roi || w <= 0 || h <= 0 || stride < (size_t)w) return LVBGS_ERR_ARG;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            if (x < 2 || y < 2 || x >= w - 2 || y >= h - 2) roi[(size_t)y * stride + x] = 0;
    return LVBGS_OK;
}

// carves the arena starting at a.base into the View (a.base == nullptr: sizing dry run)
static void carve(lvbgs_ctx* c, Arena& ar) {
    const size_t npx = c->npx, C = (size_t)c->C, N = c->p.bg_samples;
    const bool sub = c->algo == LVBGS_SUBSENSE;
    const int dsW = c->W / 8, dsH = c->H / 8;
    const size_t nds = (size_t)(dsW > 0 ? dsW : 1) * (size_t)(dsH > 0 ? dsH : 1) * C;
    View& v = c->v;
    v.roi = ar.take<uint8_t>(npx);
    v.last_color = ar.take<uint8_t>(npx * C);
    v.last_desc = ar.take<uint16_t>(npx * C);
    // tile-major, 64-pixel groups, padded for loads past the last sample of the last group
    v.sA = ar.take<uint8_t>((N * ((npx + 63) / 64) + kMaxFetch) * 64 * (C == 3 ? 12 : 4));
    float** maps[] = {&v.T, &v.R, &v.V, &v.MLD, &v.DminLT, &v.DminST, &v.RawLT, &v.RawST, &v.FinLT, &v.FinST};
    for (auto m : maps) *m = sub ? ar.take<float>(npx) : nullptr;
    uint8_t** u8s[] = {&v.unstable, &v.blinks, &v.last_raw, &v.last_raw_blink, &v.last_fg, &v.last_dil_inv,
                       &v.raw, &v.dec};
    for (auto m : u8s) *m = ar.take<uint8_t>(npx);
    const size_t nz = (size_t)((c->W + 63) / 64) * (size_t)((c->H + 63) / 64);
    v.zbits = ar.take<uint64_t>(nz * 64);
    v.hbits = ar.take<uint64_t>(nz * 64);
    v.zedge = ar.take<uint8_t>(nz * 256);
    v.zpar = ar.take<uint32_t>(nz * 128);
    v.zcount = ar.take<uint32_t>(nz);
    v.desc_cur = ar.take<uint16_t>(npx * C);
    v.rec = ar.take<uint32_t>(npx);
    v.mres = ar.take<uint32_t>(npx);
    v.wl0 = ar.take<uint32_t>(npx);
    v.wl1 = ar.take<uint32_t>(npx);
    // continuation stripes sized so no stripe can overflow; cached contexts for up to an eighth
    // of the pixels, spread over the stripes; entries beyond rebuild theirs
    const size_t ctx_cap = std::min(npx, std::max(npx / 8, (size_t)1024));
    v.cont_region = cont_stripe_slots(c->W, c->H);
    v.cont_ctx_region = (uint32_t)((ctx_cap + kContStripes - 1) / kContStripes);
    v.cont = ar.take<uint32_t>((size_t)kContStripes * v.cont_region);
    v.cont_ctx = ar.take<uint32_t>((size_t)kContStripes * v.cont_ctx_region * kCtxWords);
    v.cont2 = ar.take<uint32_t>(npx);
    v.stamp = ar.take<unsigned long long>(npx);
    v.dsLT = ar.take<float>(nds);
    v.dsST = ar.take<float>(nds);
    v.sc = ar.take<Scalars>(1);
    c->d_frame = ar.take<uint8_t>(npx * C);
    c->d_mask = ar.take<uint8_t>(npx);
    c->d_ring = c->ring_frames > 0 ? ar.take<uint8_t>((size_t)c->ring_frames * npx * C) : nullptr;
}

size_t lvbgs_arena_bytes(int algo, const lvbgs_params& p, int width, int height, int channels, int ring_frames) {
    lvbgs_ctx c;
    c.algo = algo;
    c.p = p;
    c.ring_frames = ring_frames;
    c.W = width;
    c.H = height;
    c.C = channels;
    c.npx = (size_t)width * height;
    Arena a;
    carve(&c, a);
    return a.off + 256;
}

extern "C" size_t lvbgs_model_bytes(int algo, const lvbgs_params* p, int width, int height, int channels) {
    lvbgs_params d;
    if (lvbgs_default_params(algo, &d) != LVBGS_OK || width <= 0 || height <= 0 || (channels != 1 && channels != 3))
        return 0;
    return lvbgs_arena_bytes(algo, p ? *p : d, width, height, channels);
}

// allocates (or re-uses) the device arena for the current geometry and carves the View
static int allocate(lvbgs_ctx* c) {
    Arena a;
    carve(c, a);
    const size_t need = a.off + 256;
    if (need != c->arena_size) {
        release_device(c);
        if (hipMalloc(&c->arena, need) != hipSuccess) {
            c->arena = nullptr;
            return set_err(c, LVBGS_ERR_ALLOC, "device allocation failed for the background model");
        }
        c->arena_size = need;
    }
    Arena b;
    b.base = c->arena;
    carve(c, b);
    return hipMemsetAsync(c->arena, 0, c->arena_size, c->stream) == hipSuccess ? LVBGS_OK
                                                                               : set_err(c, LVBGS_ERR_HIP, "memset");
}

static int upload_scalars(lvbgs_ctx* c, const Scalars& s) {
    return hip_check(c, hipMemcpyAsync(c->v.sc, &s, sizeof(Scalars), hipMemcpyHostToDevice, c->stream), "scalars");
}

static int download_scalars(lvbgs_ctx* c, Scalars& s) {
    int rc = hip_check(c, hipMemcpyAsync(&s, c->v.sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream), "scalars");
    if (rc) return rc;
    return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

// Geometry, arena and View of a context initialized on a w x h x ch frame whose ROI (already in
// c->roi) counts orig pixels before and fin after border trimming; fills s with the initial
// frame-level scalars (LUT, caps, automatic reset) of BackgroundSubtractorLBSP.cpp:29-43 and
// BackgroundSubtractorSuBSENSE.cpp:112-128.
static int setup_geometry(lvbgs_ctx* c, int w, int h, int ch, size_t orig, size_t fin, Scalars& s) {
    const size_t npx = (size_t)w * h;
    c->W = w;
    c->H = h;
    c->C = ch;
    c->npx = npx;
    c->orig_roi = orig;
    c->n_roi = fin;
    c->apply_count = 0;
    int rc = allocate(c);
    if (rc) return rc;
    View& v = c->v;
    v.W = w;
    v.H = h;
    v.C = ch;
    v.N = (int)c->p.bg_samples;
    v.REQ = (int)c->p.required_bg_samples;
    v.algo = c->algo;
    v.npx = (uint32_t)npx;
    v.n_roi = (uint32_t)fin;
    v.color_thr = (uint32_t)c->p.color_dist_threshold;
    v.desc_thr = (uint32_t)c->p.desc_dist_threshold;
    v.mv_avgs = (uint32_t)c->p.samples_for_mv_avgs;
    v.lbsp_offset = (uint32_t)c->p.lbsp_threshold_offset;
    v.rel_lbsp = c->p.rel_lbsp_threshold;
    v.seed_lo = (uint32_t)c->p.rng_seed;
    v.seed_hi = (uint32_t)(c->p.rng_seed >> 32);
    v.stream_id = c->p.stream_id;
    v.dsW = w / 8;
    v.dsH = h / 8;

    std::memset(&s, 0, sizeof(s));
    // LBSP LUT (BackgroundSubtractorLBSP.cpp:29-30, :42-43)
    for (int t = 0; t <= 255; ++t) {
        const float base = (float)t * c->p.rel_lbsp_threshold + (float)c->p.lbsp_threshold_offset;
        s.lut[t] = sat_u8_host(ch == 1 ? base / 3.0f : base);
    }
    float lower = 2.0f, upper = 256.0f;
    if (c->algo == LVBGS_SUBSENSE) {
        // SuBSENSE .cpp:112-128
        if (orig >= npx / 2 && (int)npx >= 320 * 240) {
            v.lr_scaling = 1;
            s.auto_reset = 1;
            v.use3x3 = !((int)npx > 320 * 240 * 2);
            const int raw = (int)std::floor((float)npx / (float)(320 * 240) + 0.5f) + c->p.median_blur_kernel;
            const int rk = raw < 14 ? raw : 14;
            v.median_k = (rk % 2) ? rk : rk - 1;
        } else {
            v.lr_scaling = 0;
            s.auto_reset = 0;
            v.use3x3 = 1;
            v.median_k = c->p.median_blur_kernel;
            lower = 4.0f;
            upper = 512.0f;
        }
    } else {
        v.lr_scaling = 0;
        v.use3x3 = 1;
        v.median_k = c->p.median_blur_kernel;
    }
    s.lower_cap = lower;
    s.upper_cap = upper;
    return LVBGS_OK;
}

// done == nullptr: synchronous (the reference's contract); otherwise the work stays queued on the
// context's stream and `done` is recorded after it
static int initialize(lvbgs_ctx* c, const uint8_t* img, bool device_img, int w, int h, int ch, size_t stride,
                      const uint8_t* roi, size_t roi_stride, hipEvent_t done);

extern "C" int lvbgs_initialize(lvbgs_ctx* c, const uint8_t* img, int w, int h, int ch, size_t stride,
                                const uint8_t* roi, size_t roi_stride) {
    return initialize(c, img, false, w, h, ch, stride, roi, roi_stride, nullptr);
}

extern "C" int lvbgs_initialize_device(lvbgs_ctx* c, const void* d_img, int w, int h, int ch, const uint8_t* roi,
                                       size_t roi_stride) {
    return initialize(c, (const uint8_t*)d_img, true, w, h, ch, (size_t)w * (size_t)(ch > 0 ? ch : 1), roi,
                      roi_stride, nullptr);
}

int lvbgs_initialize_async(lvbgs_ctx* c, const void* d_img, int w, int h, int ch, hipEvent_t done) {
    return initialize(c, (const uint8_t*)d_img, true, w, h, ch, (size_t)w * (size_t)(ch > 0 ? ch : 1), nullptr, 0,
                      done);
}

int lvbgs_initialize_async_host(lvbgs_ctx* c, const void* img, int w, int h, int ch, size_t stride, hipEvent_t done) {
    return initialize(c, (const uint8_t*)img, false, w, h, ch, stride, nullptr, 0, done);
}

static int initialize(lvbgs_ctx* c, const uint8_t* img, bool device_img, int w, int h, int ch, size_t stride,
                      const uint8_t* roi, size_t roi_stride, hipEvent_t done) {
    if (!c) return LVBGS_ERR_ARG;
    c->err.clear();
    if (!img || w <= 0 || h <= 0) return set_err(c, LVBGS_ERR_ARG, "provided image for initialization must be non-empty");
    if (ch != 1 && ch != 3)
        return set_err(c, LVBGS_ERR_ARG, "provided image for initialization must be of type 8UC1/8UC3");
    if (stride == 0) stride = (size_t)w * ch;
    if (stride < (size_t)w * ch) return set_err(c, LVBGS_ERR_ARG, "image stride too small");
    if ((size_t)w * (size_t)h > 0x7FFFFFFFull / 8) return set_err(c, LVBGS_ERR_ARG, "image too large");
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, LVBGS_ERR_HIP, "hipSetDevice");
    const size_t npx = (size_t)w * h;
    // ROI (BackgroundSubtractionUtils.cpp:82-101); the last ROI of the same size is reused as
    // stored (:87-88), already border-trimmed, and recounted (:97-99): both counts are its non-zero
    // pixels
    const bool reuse_roi = !roi && !c->roi.empty() && c->roi_w == w && c->roi_h == h;
    std::vector<uint8_t> nr;
    size_t orig = 0, fin = 0;
    if (!reuse_roi) nr.resize(npx);
    if (reuse_roi) {
        for (uint8_t r : c->roi) orig += r ? 1 : 0;
        fin = orig;
        if (!fin) return set_err(c, LVBGS_ERR_ARG, "provided ROI mat contains no useful pixels away from borders");
    } else if (!roi) {
        std::memset(nr.data(), 255, npx);
    } else {
        if (roi_stride == 0) roi_stride = (size_t)w;
        if (roi_stride < (size_t)w) return set_err(c, LVBGS_ERR_ARG, "ROI stride too small");
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const uint8_t r = roi[(size_t)y * roi_stride + x];
                if (r != 0 && r != 255)
                    return set_err(c, LVBGS_ERR_ARG, "provided ROI mat values must be 0 or 255 only");
                nr[(size_t)y * w + x] = r;
            }
        // roi | dilate(roi, 2 iterations)/2 (:92-95): the 5x5 dilation marks nearby pixels (value 128)
        std::vector<uint8_t> dil(npx, 0);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                uint8_t m = 0;
                for (int yy = y - 2; yy <= y + 2 && !m; ++yy)
                    for (int xx = x - 2; xx <= x + 2; ++xx)
                        if (yy >= 0 && yy < h && xx >= 0 && xx < w && nr[(size_t)yy * w + xx]) {
                            m = 128;
                            break;
                        }
                dil[(size_t)y * w + x] = m;
            }
        for (size_t i = 0; i < npx; ++i) nr[i] = nr[i] | dil[i];
    }
    if (!reuse_roi) {
        orig = 0;
        for (size_t i = 0; i < npx; ++i) orig += nr[i] ? 1 : 0;
        if (!orig) return set_err(c, LVBGS_ERR_ARG, "provided ROI mat contains no useful pixels");
        lvbgs_validate_roi(w, h, nr.data(), (size_t)w);
        fin = 0;
        for (size_t i = 0; i < npx; ++i) fin += nr[i] ? 1 : 0;
        if (!fin) return set_err(c, LVBGS_ERR_ARG, "provided ROI mat contains no useful pixels away from borders");
    }

    c->initialized = false;
    c->roi_w = w;
    c->roi_h = h;
    if (!reuse_roi) c->roi = std::move(nr);
    Scalars s;
    int rc = setup_geometry(c, w, h, ch, orig, fin, s);
    if (rc) return rc;
    View& v = c->v;
    if ((rc = upload_scalars(c, s))) return rc;
    if ((rc = hip_check(c, hipMemcpyAsync((void*)v.roi, c->roi.data(), npx, hipMemcpyHostToDevice, c->stream), "roi")))
        return rc;
    const uint8_t* d_init = img;
    if (!device_img) {
        if ((rc = hip_check(c, hipMemcpy2DAsync(c->d_frame, (size_t)w * ch, img, stride, (size_t)w * ch, (size_t)h,
                                                hipMemcpyHostToDevice, c->stream), "frame")))
            return rc;
        d_init = c->d_frame;
    }
    if (c->algo == LVBGS_SUBSENSE) {
        const unsigned b = (unsigned)((npx + 255) / 256);
        hipLaunchKernelGGL(k_fill_f32, dim3(b), dim3(256), 0, c->stream, v.T, npx, s.lower_cap);
        hipLaunchKernelGGL(k_fill_f32, dim3(b), dim3(256), 0, c->stream, v.R, npx, 1.0f);
        hipLaunchKernelGGL(k_fill_f32, dim3(b), dim3(256), 0, c->stream, v.V, npx, 10.0f);
    }
    launch_init_desc(v, d_init, c->stream);
    // refreshModel(1.0f) for SuBSENSE (.cpp:184), refreshModel(1.0f, true) for LOBSTER (.cpp:455)
    launch_refresh(v, 1.0f, c->algo == LVBGS_LOBSTER ? 1 : 0, 0u, 0u, c->stream);
    s.refresh_seq = 1;
    // the refresh consumed event 0; store the counter
    if ((rc = hip_check(c, hipMemcpyAsync(&v.sc->refresh_seq, &s.refresh_seq, sizeof(uint32_t), hipMemcpyHostToDevice,
                                          c->stream), "scalars")))
        return rc;
    if ((rc = hip_check(c, hipGetLastError(), "initialize launch"))) return rc;
    if (done) {
        if ((rc = hip_check(c, hipEventRecord(done, c->stream), "initialize event"))) return rc;
    } else if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "initialize"))) {
        return rc;
    }
    c->initialized = true;
    return LVBGS_OK;
}

static int check_lr(lvbgs_ctx* c, double lr) {
    if (c->algo == LVBGS_LOBSTER && !(lr > 0))
        return set_err(c, LVBGS_ERR_ARG, "learning rate must be a positive value; faster learning is achieved with smaller values");
    return LVBGS_OK;
}

// Automatic sequential-phase limit of the single-stream path, from measurements on MI355X:
// SuBSENSE RGB frames of 1280x720 and larger cut their scans at 8 samples and continue the cut
// pixels densely (1080p 11% and 4K 16% faster per frame than one full pass); smaller frames,
// one-channel SuBSENSE and LOBSTER scan every sample in one pass.
static uint32_t auto_seq_limit(const lvbgs_ctx* c) {
    return (c->algo == LVBGS_SUBSENSE && c->C == 3 && c->npx >= 900000u) ? 8u : 255u;
}

static int step(lvbgs_ctx* c, const uint8_t* d_img, uint8_t* d_mask, double lr) {
    StepOptions o = c->opt;
    o.timed_stage = 0;
    if (c->timed_stage) {
        if (c->ev_used == c->ev_pairs.size()) {
            hipEvent_t a = nullptr, b = nullptr;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
                if (a) (void)hipEventDestroy(a);
                return set_err(c, LVBGS_ERR_HIP, "timing events");
            }
            c->ev_pairs.emplace_back(a, b);
        }
        o.timed_stage = c->timed_stage;
        o.ev_begin = c->ev_pairs[c->ev_used].first;
        o.ev_end = c->ev_pairs[c->ev_used].second;
        ++c->ev_used;
    }
    View v = c->v;
    v.frame = d_img;
    v.out_mask = d_mask;
    v.lr = lr;
    v.seq_limit = o.seq_limit > 0 ? (uint32_t)o.seq_limit : auto_seq_limit(c);
    frame_step(v, o, c->stream);
    ++c->apply_count;
    return LVBGS_OK;
}

extern "C" int lvbgs_apply(lvbgs_ctx* c, const uint8_t* img, size_t stride, uint8_t* fgmask, size_t mask_stride,
                           double lr) {
    if (!c) return LVBGS_ERR_ARG;
    if (!c->initialized) return set_err(c, LVBGS_ERR_STATE, "algo & model must be initialized first");
    if (!img) return set_err(c, LVBGS_ERR_ARG, "input image type/size mismatch with initialization type/size");
    int rc;
    if ((rc = check_lr(c, lr))) return rc;
    (void)hipSetDevice(c->device);
    const size_t row = (size_t)c->W * c->C;
    if (stride == 0) stride = row;
    if (mask_stride == 0) mask_stride = (size_t)c->W;
    if (stride < row || (fgmask && mask_stride < (size_t)c->W)) return set_err(c, LVBGS_ERR_ARG, "stride too small");
    if ((rc = hip_check(c, hipMemcpy2DAsync(c->d_frame, row, img, stride, row, (size_t)c->H, hipMemcpyHostToDevice,
                                            c->stream), "input copy")))
        return rc;
    if ((rc = step(c, c->d_frame, c->d_mask, lr))) return rc;
    if ((rc = hip_check(c, hipGetLastError(), "apply launch"))) return rc;
    if (fgmask &&
        (rc = hip_check(c, hipMemcpy2DAsync(fgmask, mask_stride, c->d_mask, (size_t)c->W, (size_t)c->W, (size_t)c->H,
                                            hipMemcpyDeviceToHost, c->stream), "mask copy")))
        return rc;
    return lvbgs_synchronize(c);
}

extern "C" int lvbgs_apply_device(lvbgs_ctx* c, const void* d_img, void* d_mask, double lr) {
    if (!c) return LVBGS_ERR_ARG;
    if (!c->initialized) return set_err(c, LVBGS_ERR_STATE, "algo & model must be initialized first");
    if (!d_img) return set_err(c, LVBGS_ERR_ARG, "input image must be a device pointer");
    int rc;
    if ((rc = check_lr(c, lr))) return rc;
    (void)hipSetDevice(c->device);
    if ((rc = step(c, (const uint8_t*)d_img, (uint8_t*)d_mask, lr))) return rc;
    return hip_check(c, hipGetLastError(), "apply launch");
}

extern "C" int lvbgs_synchronize(lvbgs_ctx* c) {
    if (!c) return LVBGS_ERR_ARG;
    (void)hipSetDevice(c->device);
    int rc = hip_check(c, hipStreamSynchronize(c->stream), "synchronize");
    if (rc || !c->initialized) return rc;
    uint32_t flags = 0;
    rc = hip_check(c, hipMemcpy(&flags, &c->v.sc->err_flags, sizeof(flags), hipMemcpyDeviceToHost), "error flags");
    if (rc) return rc;
    if (flags & kErrFixpointCap) return set_err(c, LVBGS_ERR_DEVICE, "fix-point iteration cap reached");
    return LVBGS_OK;
}

extern "C" int lvbgs_get_background_image(lvbgs_ctx* c, uint8_t* out, size_t stride) {
    if (!c || !out) return LVBGS_ERR_ARG;
    if (!c->initialized) return set_err(c, LVBGS_ERR_STATE, "algo must be initialized first");
    (void)hipSetDevice(c->device);
    const size_t row = (size_t)c->W * c->C;
    if (stride == 0) stride = row;
    if (stride < row) return set_err(c, LVBGS_ERR_ARG, "stride too small");
    launch_background(c->v, c->d_frame, nullptr, c->stream);
    int rc = hip_check(c, hipMemcpy2DAsync(out, stride, c->d_frame, row, row, (size_t)c->H, hipMemcpyDeviceToHost,
                                           c->stream), "background copy");
    if (rc) return rc;
    return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

extern "C" int lvbgs_get_background_descriptors_image(lvbgs_ctx* c, uint16_t* out, size_t stride) {
    if (!c || !out) return LVBGS_ERR_ARG;
    if (!c->initialized) return set_err(c, LVBGS_ERR_STATE, "algo must be initialized first");
    (void)hipSetDevice(c->device);
    const size_t row = (size_t)c->W * c->C * 2;
    if (stride == 0) stride = row;
    if (stride < row) return set_err(c, LVBGS_ERR_ARG, "stride too small");
    // staging: the per-step descriptor scratch (same size, only live inside a step)
    uint16_t* tmp = c->v.desc_cur;
    launch_background(c->v, nullptr, tmp, c->stream);
    int rc = hip_check(c, hipMemcpy2DAsync(out, stride, tmp, row, row, (size_t)c->H, hipMemcpyDeviceToHost, c->stream),
                       "descriptor copy");
    if (rc) return rc;
    return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

extern "C" int lvbgs_set_automatic_model_reset(lvbgs_ctx* c, int enabled) {
    if (!c) return LVBGS_ERR_ARG;
    if (!c->initialized) return LVBGS_OK;  // initialize() sets the flag itself (SUB:113-123)
    (void)hipSetDevice(c->device);
    const int32_t val = enabled ? 1 : 0;
    int rc = hip_check(c, hipMemcpyAsync(&c->v.sc->auto_reset, &val, sizeof(val), hipMemcpyHostToDevice, c->stream),
                       "auto reset");
    if (rc) return rc;
    return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

extern "C" int lvbgs_refresh_model(lvbgs_ctx* c, float frac, int force_fg) {
    if (!c) return LVBGS_ERR_ARG;
    if (!c->initialized) return set_err(c, LVBGS_ERR_STATE, "algo must be initialized first");
    if (!(frac > 0.0f && frac <= 1.0f)) return set_err(c, LVBGS_ERR_ARG, "model refresh must be given as a non-null fraction");
    (void)hipSetDevice(c->device);
    Scalars s;
    int rc = download_scalars(c, s);
    if (rc) return rc;
    const uint32_t ev = s.refresh_seq;
    launch_refresh(c->v, frac, force_fg ? 1 : 0, s.frame_idx, ev, c->stream);
    const uint32_t next = ev + 1;
    if ((rc = hip_check(c, hipMemcpyAsync(&c->v.sc->refresh_seq, &next, sizeof(next), hipMemcpyHostToDevice, c->stream),
                        "refresh counter")))
        return rc;
    return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

extern "C" int lvbgs_get_roi(lvbgs_ctx* c, uint8_t* out, size_t stride) {
    if (!c || !out || c->roi.empty()) return LVBGS_ERR_ARG;
    if (stride == 0) stride = (size_t)c->roi_w;
    for (int y = 0; y < c->roi_h; ++y) std::memcpy(out + (size_t)y * stride, c->roi.data() + (size_t)y * c->roi_w, c->roi_w);
    return LVBGS_OK;
}

extern "C" int lvbgs_set_roi(lvbgs_ctx* c, uint8_t* roi, size_t stride) {
    if (!c || !roi) return LVBGS_ERR_ARG;
    const int w = c->initialized ? c->W : c->roi_w, h = c->initialized ? c->H : c->roi_h;
    if (w <= 0 || h <= 0) return set_err(c, LVBGS_ERR_ARG, "setROI needs a known frame size (initialize first)");
    if (stride == 0) stride = (size_t)w;
    lvbgs_validate_roi(w, h, roi, stride);
    size_t nz = 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) nz += roi[(size_t)y * stride + x] ? 1 : 0;
    if (!nz) return set_err(c, LVBGS_ERR_ARG, "provided ROI must have at least one valid pixel");
    if (c->initialized) {
        std::vector<uint8_t> bg((size_t)c->W * c->H * c->C);
        int rc = lvbgs_get_background_image(c, bg.data(), 0);
        if (rc) return rc;
        return lvbgs_initialize(c, bg.data(), c->W, c->H, c->C, 0, roi, stride);
    }
    c->roi.assign((size_t)w * h, 0);
    for (int y = 0; y < h; ++y) std::memcpy(c->roi.data() + (size_t)y * w, roi + (size_t)y * stride, w);
    return LVBGS_OK;
}

extern "C" int lvbgs_lbsp_compute_abs(const void* d_img, int w, int h, int ch, int threshold, void* d_out, void* stream) {
    if (!d_img || !d_out || w <= 0 || h <= 0 || (ch != 1 && ch != 3)) return LVBGS_ERR_ARG;
    const uint32_t t = (uint32_t)(threshold < 0 ? 0 : threshold > 255 ? 255 : threshold);  // saturate_cast<uchar>
    launch_lbsp_abs((const uint8_t*)d_img, w, h, ch, t, (uint16_t*)d_out, false, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? LVBGS_OK : LVBGS_ERR_HIP;
}

extern "C" int lvbgs_debug_lbsp_packed(const void* d_img, int w, int h, int ch, int threshold, void* d_out,
                                       void* stream) {
    if (!d_img || !d_out || w <= 0 || h <= 0 || (ch != 1 && ch != 3)) return LVBGS_ERR_ARG;
    const uint32_t t = (uint32_t)(threshold < 0 ? 0 : threshold > 255 ? 255 : threshold);
    launch_lbsp_abs((const uint8_t*)d_img, w, h, ch, t, (uint16_t*)d_out, true, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? LVBGS_OK : LVBGS_ERR_HIP;
}

extern "C" int lvbgs_get_stats(lvbgs_ctx* c, lvbgs_stats* out) {
    if (!c || !out) return LVBGS_ERR_ARG;
    if (!c->initialized) return set_err(c, LVBGS_ERR_STATE, "algo must be initialized first");
    Scalars s;
    int rc = download_scalars(c, s);
    if (rc) return rc;
    auto sum = [](const unsigned long long* a) {
        unsigned long long t = 0;
        for (int i = 0; i < kStripes; ++i) t += a[i];
        return t;
    };
    out->examined = sum(s.ex_part);
    out->writes = sum(s.wr_part);
    out->examined_speculative = sum(s.spec_part);
    out->examined_cooperative = sum(s.coop_part);
    for (int i = 0; i < 8; ++i) out->fixpoint_work[i] = s.wl_work[i];
    uint32_t handed = 0;
    for (int k = 0; k < kContStripes; ++k) handed += s.cont_cnt[k * kContStride];
    out->fixpoint_work[0] = handed;
    out->fixpoint_iters = s.iters;
    out->max_fixpoint_iter = s.max_iter_seen;
    out->width = (uint32_t)c->W;
    out->height = (uint32_t)c->H;
    out->channels = (uint32_t)c->C;
    out->n_roi = (uint32_t)c->n_roi;
    return LVBGS_OK;
}

extern "C" int lvbgs_get_scalars(lvbgs_ctx* c, double* out, int cap) {
    if (!c || !out) return LVBGS_ERR_ARG;
    if (!c->initialized) return set_err(c, LVBGS_ERR_STATE, "algo must be initialized first");
    Scalars s;
    int rc = download_scalars(c, s);
    if (rc) return rc;
    const double vals[] = {(double)s.frame_idx, (double)s.frames_since_reset, (double)s.cooldown, (double)s.auto_reset,
                           (double)c->v.lr_scaling, (double)c->v.use3x3, (double)c->v.median_k, (double)c->n_roi,
                           (double)s.lower_cap, (double)s.upper_cap, (double)s.last_nz_ratio, (double)s.refresh_seq,
                           (double)c->apply_count};
    const int n = (int)(sizeof(vals) / sizeof(vals[0]));
    for (int i = 0; i < n && i < cap; ++i) out[i] = vals[i];
    return n;
}

extern "C" long long lvbgs_get_state(lvbgs_ctx* c, const char* name, void* out, size_t cap) {
    if (!c || !name || !out) return LVBGS_ERR_ARG;
    if (!c->initialized) return LVBGS_ERR_STATE;
    (void)hipSetDevice(c->device);
    const View& v = c->v;
    const size_t npx = c->npx, C = (size_t)c->C, N = c->p.bg_samples;
    const std::string n(name);
    const void* src = nullptr;
    size_t bytes = 0;
    std::vector<uint8_t> unpack;
    if (n == "samples_color" || n == "samples_desc") {
        // unpack the sample planes to the oracle layout: N x (H x W x C)
        const size_t rec = C == 3 ? 12 : 4;
        std::vector<uint8_t> a(N * ((npx + 63) / 64 * 64) * rec);
        // the context's stream does not synchronize with the null stream: wait for its steps
        if (hipStreamSynchronize(c->stream) != hipSuccess) return LVBGS_ERR_HIP;
        if (hipMemcpy(a.data(), v.sA, a.size(), hipMemcpyDeviceToHost) != hipSuccess) return LVBGS_ERR_HIP;
        const bool col = n == "samples_color";
        bytes = N * npx * C * (col ? 1 : 2);
        if (bytes > cap) return LVBGS_ERR_ARG;
        for (size_t i = 0; i < N * npx; ++i) {
            const size_t s = i / npx, p = i % npx;  // output order: sample-major, then pixel
            const size_t src = ((p >> 6) * N + s) * 64 + (p & 63);
            uint32_t w[3] = {0, 0, 0};
            std::memcpy(w, a.data() + src * rec, rec);
            if (C == 3) {
                if (col) {
                    uint8_t* o = (uint8_t*)out + i * 3;
                    o[0] = (uint8_t)(w[1] >> 16);
                    o[1] = (uint8_t)(w[1] >> 24);
                    o[2] = (uint8_t)w[2];
                } else {
                    uint16_t* o = (uint16_t*)out + i * 3;
                    o[0] = (uint16_t)w[0];
                    o[1] = (uint16_t)(w[0] >> 16);
                    o[2] = (uint16_t)w[1];
                }
            } else {
                if (col) ((uint8_t*)out)[i] = (uint8_t)(w[0] >> 16);
                else ((uint16_t*)out)[i] = (uint16_t)w[0];
            }
        }
        return (long long)bytes;
    }
    struct Item {
        const char* name;
        const void* ptr;
        size_t bytes;
    };
    const size_t fb = npx * 4, nds = (size_t)v.dsW * v.dsH * C * 4;
    const Item items[] = {{"last_color", v.last_color, npx * C},  {"last_desc", v.last_desc, npx * C * 2},
                          {"last_fg", v.last_fg, npx},             {"roi", v.roi, npx},
                          {"lut", v.sc->lut, 256},                 {"T", v.T, fb},
                          {"R", v.R, fb},                          {"V", v.V, fb},
                          {"MLD", v.MLD, fb},                      {"DminLT", v.DminLT, fb},
                          {"DminST", v.DminST, fb},                {"RawLT", v.RawLT, fb},
                          {"RawST", v.RawST, fb},                  {"FinLT", v.FinLT, fb},
                          {"FinST", v.FinST, fb},                  {"unstable", v.unstable, npx},
                          {"blinks", v.blinks, npx},               {"last_raw", v.last_raw, npx},
                          {"last_raw_blink", v.last_raw_blink, npx}, {"last_dil_inv", v.last_dil_inv, npx},
                          {"dsLT", v.dsLT, nds},                   {"dsST", v.dsST, nds},
                          {"raw", v.raw, npx},                     {"dec", v.dec, npx},
                          {"mres", v.mres, npx * 4},               {"rec", v.rec, npx * 4},
                          {"desc_cur", v.desc_cur, npx * C * 2}};
    for (const auto& it : items)
        if (n == it.name) {
            src = it.ptr;
            bytes = it.bytes;
        }
    if (!src) return LVBGS_ERR_ARG;
    if (bytes > cap) return LVBGS_ERR_ARG;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return LVBGS_ERR_HIP;
    if (hipMemcpy(out, src, bytes, hipMemcpyDeviceToHost) != hipSuccess) return LVBGS_ERR_HIP;
    return (long long)bytes;
}

// ---- checkpoint restore (bgs_archive.hip): geometry first, then the saved buffers, then the scalars ----
int lvbgs_restore_begin(lvbgs_ctx* c, int w, int h, int ch, const uint8_t* roi, size_t n_roi, size_t orig_roi) {
    (void)hipSetDevice(c->device);
    const size_t npx = (size_t)w * (size_t)h;
    size_t fin = 0;
    for (size_t i = 0; i < npx; ++i) fin += roi[i] ? 1 : 0;
    if (fin != n_roi || !fin) return set_err(c, LVBGS_ERR_ARG, "checkpoint ROI does not match its pixel count");
    c->initialized = false;
    c->roi.assign(roi, roi + npx);
    c->roi_w = w;
    c->roi_h = h;
    Scalars s;
    int rc = setup_geometry(c, w, h, ch, orig_roi, fin, s);
    if (rc) return rc;
    if ((rc = hip_check(c, hipMemcpyAsync((void*)c->v.roi, c->roi.data(), npx, hipMemcpyHostToDevice, c->stream),
                        "roi")))
        return rc;
    const size_t rec = c->C == 3 ? 3 : 1, N = c->p.bg_samples;
    std::vector<uint32_t>& r = c->restore_stage;
    r.assign((N * ((npx + 63) / 64) + kMaxFetch) * 64 * rec, 0u);
    return hip_check(c, hipStreamSynchronize(c->stream), "restore");
}

int lvbgs_restore_buffer(lvbgs_ctx* c, const char* name, const void* data, size_t bytes) {
    const View& v = c->v;
    const size_t npx = c->npx, C = (size_t)c->C, N = c->p.bg_samples;
    const std::string n(name);
    if (n == "samples_color" || n == "samples_desc") {
        // the oracle layout (N x H x W x C) into the tile-major records (sample s of 64-pixel group g
        // at record (g*N + s)*64 + lane); RGB {d0 | d1<<16, d2 | c0<<16 | c1<<24, c2}, grey {d0 | c0<<16}
        std::vector<uint32_t>& r = c->restore_stage;
        const bool col = n == "samples_color";
        if (bytes != N * npx * C * (col ? 1 : 2) || r.empty()) return set_err(c, LVBGS_ERR_ARG, "sample planes");
        const size_t rw = C == 3 ? 3 : 1;
        for (size_t i = 0; i < N * npx; ++i) {
            const size_t s = i / npx, p = i % npx;
            uint32_t* w = &r[(((p >> 6) * N + s) * 64 + (p & 63)) * rw];
            if (C == 3) {
                if (col) {
                    const uint8_t* o = (const uint8_t*)data + i * 3;
                    w[1] = (w[1] & 0xFFFFu) | ((uint32_t)o[0] << 16) | ((uint32_t)o[1] << 24);
                    w[2] = o[2];
                } else {
                    const uint16_t* o = (const uint16_t*)data + i * 3;
                    w[0] = (uint32_t)o[0] | ((uint32_t)o[1] << 16);
                    w[1] = (w[1] & 0xFFFF0000u) | o[2];
                }
            } else if (col) {
                w[0] = (w[0] & 0xFFFFu) | ((uint32_t)((const uint8_t*)data)[i] << 16);
            } else {
                w[0] = (w[0] & 0xFFFF0000u) | ((const uint16_t*)data)[i];
            }
        }
        return LVBGS_OK;
    }
    struct Item {
        const char* name;
        void* ptr;
        size_t bytes;
    };
    const size_t fb = npx * 4, nds = (size_t)v.dsW * v.dsH * C * 4;
    const Item items[] = {{"last_color", v.last_color, npx * C},  {"last_desc", v.last_desc, npx * C * 2},
                          {"last_fg", v.last_fg, npx},             {"T", v.T, fb},
                          {"R", v.R, fb},                          {"V", v.V, fb},
                          {"MLD", v.MLD, fb},                      {"DminLT", v.DminLT, fb},
                          {"DminST", v.DminST, fb},                {"RawLT", v.RawLT, fb},
                          {"RawST", v.RawST, fb},                  {"FinLT", v.FinLT, fb},
                          {"FinST", v.FinST, fb},                  {"unstable", v.unstable, npx},
                          {"blinks", v.blinks, npx},               {"last_raw", v.last_raw, npx},
                          {"last_raw_blink", v.last_raw_blink, npx}, {"last_dil_inv", v.last_dil_inv, npx},
                          {"dsLT", v.dsLT, nds},                   {"dsST", v.dsST, nds}};
    for (const auto& it : items)
        if (n == it.name) {
            if (!it.ptr || it.bytes != bytes) return set_err(c, LVBGS_ERR_ARG, (std::string("buffer size mismatch: ") + name).c_str());
            return hip_check(c, hipMemcpy(it.ptr, data, bytes, hipMemcpyHostToDevice), name);
        }
    return set_err(c, LVBGS_ERR_ARG, (std::string("unknown buffer ") + name).c_str());
}

int lvbgs_restore_scalars(lvbgs_ctx* c, const Scalars& s, uint32_t apply_count) {
    std::vector<uint32_t>& r = c->restore_stage;
    if (r.empty()) return set_err(c, LVBGS_ERR_STATE, "checkpoint restore not begun");
    int rc = hip_check(c, hipMemcpy(c->v.sA, r.data(), r.size() * 4, hipMemcpyHostToDevice), "samples");
    std::vector<uint32_t>().swap(r);
    if (rc) return rc;
    if ((rc = hip_check(c, hipMemcpy(c->v.sc, &s, sizeof(Scalars), hipMemcpyHostToDevice), "scalars"))) return rc;
    c->apply_count = apply_count;
    c->initialized = true;
    return LVBGS_OK;
}

extern "C" int lvbgs_set_timing(lvbgs_ctx* c, const char* stage) {
    if (!c) return LVBGS_ERR_ARG;
    const std::string s(stage ? stage : "");
    if (s.empty()) c->timed_stage = 0;
    else if (s == "match") c->timed_stage = 1;
    else if (s == "step") c->timed_stage = 2;
    else if (s == "update") c->timed_stage = 3;
    else return set_err(c, LVBGS_ERR_ARG, "unknown timing stage");
    c->ev_used = 0;
    return LVBGS_OK;
}

extern "C" int lvbgs_take_timing(lvbgs_ctx* c, double* total_ms, int* count) {
    if (!c || !total_ms || !count) return LVBGS_ERR_ARG;
    int rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
    if (rc) return rc;
    double tot = 0.0;
    for (size_t i = 0; i < c->ev_used; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev_pairs[i].first, c->ev_pairs[i].second) != hipSuccess)
            return set_err(c, LVBGS_ERR_HIP, "event timing");
        tot += ms;
    }
    *total_ms = tot;
    *count = (int)c->ev_used;
    c->ev_used = 0;
    return LVBGS_OK;
}

extern "C" int lvbgs_set_fixpoint_schedule(lvbgs_ctx* c, int launches, int blocks) {
    if (!c || launches < 0 || launches > 64 || blocks < 1 || blocks > 65535) return LVBGS_ERR_ARG;
    c->opt.fixpoint_launches = launches;
    c->opt.fixpoint_blocks = blocks;
    return LVBGS_OK;
}

extern "C" int lvbgs_set_match_schedule(lvbgs_ctx* c, int seq_limit, int cont_blocks) {
    if (!c || seq_limit < 0 || seq_limit > 255 || cont_blocks < 1 || cont_blocks > 65535) return LVBGS_ERR_ARG;
    c->opt.seq_limit = seq_limit;
    c->opt.cont_blocks = cont_blocks;
    return LVBGS_OK;
}

// RNG self-check for tests: draws of group 0 for pixels [0,n) of a (frame, phase, event) counter
extern "C" int lvbgs_debug_rng(lvbgs_ctx* c, uint32_t frame, uint32_t phase, uint32_t ev, uint32_t n, uint32_t* host_out) {
    if (!c || !host_out || !c->initialized || n == 0) return LVBGS_ERR_ARG;
    (void)hipSetDevice(c->device);
    uint32_t* d = nullptr;
    if (hipMalloc(&d, (size_t)n * 16) != hipSuccess) return LVBGS_ERR_ALLOC;
    launch_rng_probe(c->v, frame, phase, ev, n, d, c->stream);
    int rc = hip_check(c, hipMemcpyAsync(host_out, d, (size_t)n * 16, hipMemcpyDeviceToHost, c->stream), "rng copy");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
    (void)hipFree(d);
    return rc;
}

extern "C" int lvbgs_debug_postprocess(lvbgs_ctx* c, const uint8_t* host_raw, uint8_t* host_last_fg) {
    if (!c || !host_raw || !host_last_fg) return LVBGS_ERR_ARG;
    if (!c->initialized) return set_err(c, LVBGS_ERR_STATE, "algo & model must be initialized first");
    if (c->algo != LVBGS_SUBSENSE) return set_err(c, LVBGS_ERR_ARG, "post-processing check is for SuBSENSE");
    (void)hipSetDevice(c->device);
    int rc = hip_check(c, hipMemcpyAsync(c->v.raw, host_raw, c->npx, hipMemcpyHostToDevice, c->stream), "raw copy");
    if (rc) return rc;
    launch_postprocess_probe(c->v, c->stream);
    rc = hip_check(c, hipMemcpyAsync(host_last_fg, c->v.last_fg, c->npx, hipMemcpyDeviceToHost, c->stream), "mask copy");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
    return rc;
}

// ------------------------------------------------------------------------------------------------
// batches: several initialized SuBSENSE contexts stepped together, one launch per kernel
// ------------------------------------------------------------------------------------------------
struct lvbgs_batch {
    std::vector<lvbgs_ctx*> ctx;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    View* d_views = nullptr;
    View* h_views[2] = {nullptr, nullptr};  // pinned staging, double-buffered
    hipEvent_t ev_copy[2] = {nullptr, nullptr};
    bool ev_live[2] = {false, false};
    int flip = 0;
    uint32_t* d_prefix = nullptr;
    StepOptions opt{};
    // Members are split into `ngroups` groups (largest first, dealt round-robin) stepped on their
    // own HIP streams between a fork and a join on the batch stream, so that one group's memory-bound
    // kernels can overlap another's VALU-bound ones. order[k] = member staged at View slot k; a group
    // is a contiguous range of slots.
    static constexpr int kMaxGroups = 4;
    struct Group {
        int first = 0, count = 0;
        const uint32_t* prefix[kNumGrids] = {};
        uint32_t total[kNumGrids] = {};
        hipGraphExec_t graph = nullptr;
    };
    int ngroups = 1, planned_groups = 1;
    std::vector<int> order;
    Group grp[kMaxGroups];
    hipStream_t gstream[kMaxGroups] = {};  // gstream[0] is the batch stream
    hipEvent_t ev_fork = nullptr, ev_join[kMaxGroups] = {};
    // per-stream overrides of the batch options (0 = the batch's value)
    struct StreamOverride {
        int seq_limit = 0, cont_px_per_block = 0, fix_px_per_block = 0;
    };
    std::vector<StreamOverride> sopt;
    bool dirty = false;  // options changed since the grids were planned: re-plan at the next step
    bool use_graph = false;
    int timed = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pairs;
    size_t ev_used = 0;
};

static int bset_err(lvbgs_batch* b, int code, const char* msg) {
    if (b) b->err = msg;
    return code;
}

static void batch_release(lvbgs_batch* b) {
    for (auto& g : b->grp) {
        if (g.graph) (void)hipGraphExecDestroy(g.graph);
        g.graph = nullptr;
    }
    for (int g = 1; g < lvbgs_batch::kMaxGroups; ++g)
        if (b->gstream[g]) (void)hipStreamDestroy(b->gstream[g]);
    for (auto e : b->ev_join)
        if (e) (void)hipEventDestroy(e);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->d_views) (void)hipFree(b->d_views);
    if (b->d_prefix) (void)hipFree(b->d_prefix);
    for (int i = 0; i < 2; ++i) {
        if (b->h_views[i]) (void)hipHostFree(b->h_views[i]);
        if (b->ev_copy[i]) (void)hipEventDestroy(b->ev_copy[i]);
    }
    for (auto& e : b->ev_pairs) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

// the options of stream i: the batch's, with the stream's overrides
static StepOptions stream_options(const lvbgs_batch* b, int i) {
    StepOptions o = b->opt;
    const auto& so = b->sopt[(size_t)i];
    if (so.seq_limit) o.seq_limit = so.seq_limit;
    if (so.cont_px_per_block) o.cont_px_per_block = so.cont_px_per_block;
    if (so.fix_px_per_block) o.fix_px_per_block = so.fix_px_per_block;
    return o;
}

static int batch_plan(lvbgs_batch* b) {
    const int ns = (int)b->ctx.size();
    const int G = std::max(1, std::min(b->ngroups, ns));
    // largest members first, dealt round-robin: every group gets a share of each resolution
    std::vector<int> by_size(ns);
    for (int i = 0; i < ns; ++i) by_size[i] = i;
    std::stable_sort(by_size.begin(), by_size.end(),
                     [&](int x, int y) { return b->ctx[x]->npx > b->ctx[y]->npx; });
    b->order.clear();
    for (int g = 0; g < lvbgs_batch::kMaxGroups; ++g) {
        auto& gr = b->grp[g];
        if (gr.graph) (void)hipGraphExecDestroy(gr.graph);
        gr.graph = nullptr;
        gr.first = (int)b->order.size();
        if (g < G)
            for (int k = g; k < ns; k += G) b->order.push_back(by_size[k]);
        gr.count = (int)b->order.size() - gr.first;
    }
    // grid tables: group g, kind k at row (g * kNumGrids + k), ns + 1 entries per row
    const size_t row = (size_t)ns + 1;
    std::vector<uint32_t> pre((size_t)lvbgs_batch::kMaxGroups * kNumGrids * row, 0);
    for (int g = 0; g < G; ++g) {
        const auto& gr = b->grp[g];
        for (int j = 0; j < gr.count; ++j) {
            const int i = b->order[(size_t)gr.first + j];
            uint32_t gk[kNumGrids];
            kernel_grids(b->ctx[i]->v, stream_options(b, i), gk);
            for (int k = 0; k < kNumGrids; ++k) {
                uint32_t* r = &pre[((size_t)g * kNumGrids + k) * row];
                r[j + 1] = r[j] + gk[k];
            }
        }
    }
    if (!b->d_prefix && hipMalloc(&b->d_prefix, pre.size() * sizeof(uint32_t)) != hipSuccess)
        return bset_err(b, LVBGS_ERR_ALLOC, "batch tables");
    if (hipMemcpy(b->d_prefix, pre.data(), pre.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
        return bset_err(b, LVBGS_ERR_HIP, "batch tables");
    for (int g = 0; g < G; ++g)
        for (int k = 0; k < kNumGrids; ++k) {
            b->grp[g].prefix[k] = b->d_prefix + ((size_t)g * kNumGrids + k) * row;
            b->grp[g].total[k] = pre[((size_t)g * kNumGrids + k) * row + b->grp[g].count];
        }
    // group streams and fork/join events, created on first use
    for (int g = 1; g < G; ++g) {
        if (!b->gstream[g] && hipStreamCreateWithFlags(&b->gstream[g], hipStreamNonBlocking) != hipSuccess)
            return bset_err(b, LVBGS_ERR_HIP, "group stream");
        if (!b->ev_join[g] && hipEventCreateWithFlags(&b->ev_join[g], hipEventDisableTiming) != hipSuccess)
            return bset_err(b, LVBGS_ERR_HIP, "group event");
    }
    if (G > 1 && !b->ev_fork && hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess)
        return bset_err(b, LVBGS_ERR_HIP, "group event");
    b->gstream[0] = b->stream;
    b->planned_groups = G;
    return LVBGS_OK;
}

extern "C" int lvbgs_batch_create(lvbgs_ctx* const* ctxs, int n, lvbgs_batch** out) {
    if (!ctxs || n <= 0 || !out) return LVBGS_ERR_ARG;
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i] || !ctxs[i]->initialized || ctxs[i]->algo != LVBGS_SUBSENSE || ctxs[i]->device != ctxs[0]->device)
            return LVBGS_ERR_ARG;
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return LVBGS_ERR_ARG;
    }
    lvbgs_batch* b = new lvbgs_batch();
    b->ctx.assign(ctxs, ctxs + n);
    b->sopt.assign((size_t)n, {});
    b->device = ctxs[0]->device;
    {
        lvbgs_batch_options d;
        lvbgs_batch_default_options(&d);
        b->use_graph = d.use_graph != 0;
        b->opt.seq_limit = d.seq_limit;
        b->opt.cont_blocks = d.cont_blocks;
        b->opt.cont_px_per_block = d.cont_px_per_block;
        b->opt.fixpoint_launches = d.fixpoint_launches;
        b->opt.fixpoint_blocks = d.fixpoint_blocks;
        b->opt.fix_px_per_block = d.fix_px_per_block;
        b->opt.fetch_depth = d.fetch_depth;
        b->ngroups = d.groups;
    }
    (void)hipSetDevice(b->device);
    int rc = LVBGS_OK;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&b->d_views, sizeof(View) * n) != hipSuccess ||
        hipHostMalloc((void**)&b->h_views[0], sizeof(View) * n) != hipSuccess ||
        hipHostMalloc((void**)&b->h_views[1], sizeof(View) * n) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_copy[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_copy[1], hipEventDisableTiming) != hipSuccess)
        rc = LVBGS_ERR_HIP;
    if (!rc) rc = batch_plan(b);
    if (rc) {
        batch_release(b);
        delete b;
        return rc;
    }
    // member streams must be idle before the batch takes over their state
    for (auto* c : b->ctx) (void)hipStreamSynchronize(c->stream);
    *out = b;
    return LVBGS_OK;
}

extern "C" void lvbgs_batch_destroy(lvbgs_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_release(b);
    delete b;
}

extern "C" const char* lvbgs_batch_last_error(const lvbgs_batch* b) { return b ? b->err.c_str() : "null batch"; }
extern "C" void* lvbgs_batch_stream(lvbgs_batch* b) { return b ? (void*)b->stream : nullptr; }

extern "C" int lvbgs_batch_default_options(lvbgs_batch_options* o) {
    if (!o) return LVBGS_ERR_ARG;
    o->use_graph = 1;
    // what the online tuner settles on for settled models of mixed resolutions
    o->seq_limit = 12;
    o->cont_blocks = 1024;
    o->cont_px_per_block = 2048;
    o->fixpoint_launches = 2;
    o->fixpoint_blocks = 256;
    o->fix_px_per_block = 8192;
    o->fetch_depth = 4;
    o->groups = 1;
    return LVBGS_OK;
}

extern "C" int lvbgs_batch_get_options(const lvbgs_batch* b, lvbgs_batch_options* o) {
    if (!b || !o) return LVBGS_ERR_ARG;
    o->use_graph = b->use_graph ? 1 : 0;
    o->seq_limit = b->opt.seq_limit;
    o->cont_blocks = b->opt.cont_blocks;
    o->cont_px_per_block = b->opt.cont_px_per_block;
    o->fixpoint_launches = b->opt.fixpoint_launches;
    o->fixpoint_blocks = b->opt.fixpoint_blocks;
    o->fix_px_per_block = b->opt.fix_px_per_block;
    o->fetch_depth = b->opt.fetch_depth;
    o->groups = b->ngroups;
    return LVBGS_OK;
}

extern "C" int lvbgs_batch_set_options(lvbgs_batch* b, const lvbgs_batch_options* o) {
    if (!b || !o) return LVBGS_ERR_ARG;
    if (o->seq_limit < 1 || o->seq_limit > 255 || o->cont_blocks < 1 || o->cont_blocks > 65535 ||
        o->fixpoint_launches < 0 || o->fixpoint_launches > 64 || o->fixpoint_blocks < 1 || o->fixpoint_blocks > 65535 ||
        o->cont_px_per_block < 0 || o->fix_px_per_block < 0 || (o->fetch_depth != 2 && o->fetch_depth != 4) ||
        o->groups < 1 || o->groups > lvbgs_batch::kMaxGroups)
        return bset_err(b, LVBGS_ERR_ARG, "invalid batch options");
    b->use_graph = o->use_graph != 0;
    b->opt.seq_limit = o->seq_limit;
    b->opt.cont_blocks = o->cont_blocks;
    b->opt.cont_px_per_block = o->cont_px_per_block;
    b->opt.fixpoint_launches = o->fixpoint_launches;
    b->opt.fixpoint_blocks = o->fixpoint_blocks;
    b->opt.fix_px_per_block = o->fix_px_per_block;
    b->opt.fetch_depth = o->fetch_depth;
    b->ngroups = o->groups;
    b->dirty = true;
    return LVBGS_OK;
}

extern "C" int lvbgs_batch_set_stream_options(lvbgs_batch* b, int index, int seq_limit, int cont_px_per_block,
                                              int fix_px_per_block) {
    if (!b) return LVBGS_ERR_ARG;
    if (index < 0 || index >= (int)b->ctx.size() || seq_limit < 0 || seq_limit > 255 || cont_px_per_block < 0 ||
        fix_px_per_block < 0)
        return bset_err(b, LVBGS_ERR_ARG, "invalid stream options");
    b->sopt[(size_t)index] = {seq_limit, cont_px_per_block, fix_px_per_block};
    b->dirty = true;
    return LVBGS_OK;
}

extern "C" int lvbgs_batch_get_stream_options(const lvbgs_batch* b, int index, int* seq_limit, int* cont_px_per_block,
                                              int* fix_px_per_block) {
    if (!b || index < 0 || index >= (int)b->ctx.size() || !seq_limit || !cont_px_per_block || !fix_px_per_block)
        return LVBGS_ERR_ARG;
    const auto& so = b->sopt[(size_t)index];
    *seq_limit = so.seq_limit;
    *cont_px_per_block = so.cont_px_per_block;
    *fix_px_per_block = so.fix_px_per_block;
    return LVBGS_OK;
}

extern "C" int lvbgs_batch_apply_device(lvbgs_batch* b, const void* const* d_imgs, void* const* d_masks,
                                        const double* lrs) {
    if (!b || !d_imgs) return LVBGS_ERR_ARG;
    (void)hipSetDevice(b->device);
    if (b->dirty) {
        // the grid tables are rewritten: steps still running read them
        if (hipStreamSynchronize(b->stream) != hipSuccess) return bset_err(b, LVBGS_ERR_HIP, "sync");
        const int rc = batch_plan(b);
        if (rc) return rc;
        b->dirty = false;
    }
    const int ns = (int)b->ctx.size(), i = b->flip;
    b->flip ^= 1;
    // the staging buffer's previous copy (two steps ago) must have been consumed
    if (b->ev_live[i] && hipEventSynchronize(b->ev_copy[i]) != hipSuccess) return bset_err(b, LVBGS_ERR_HIP, "sync");
    for (int s = 0; s < ns; ++s)
        if (!d_imgs[s]) return bset_err(b, LVBGS_ERR_ARG, "input image must be a device pointer");
    if ((b->timed == 1 || b->timed == 3) && b->planned_groups > 1)
        return bset_err(b, LVBGS_ERR_ARG, "kernel-stage timing needs groups = 1");
    for (int k = 0; k < ns; ++k) {
        const int s = b->order[(size_t)k];
        View v = b->ctx[s]->v;
        v.frame = (const uint8_t*)d_imgs[s];
        v.out_mask = d_masks ? (uint8_t*)d_masks[s] : nullptr;
        v.lr = lrs ? lrs[s] : 0.0;
        v.seq_limit = (uint32_t)stream_options(b, s).seq_limit;
        b->h_views[i][k] = v;
        ++b->ctx[s]->apply_count;
    }
    if (hipMemcpyAsync(b->d_views, b->h_views[i], sizeof(View) * ns, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipEventRecord(b->ev_copy[i], b->stream) != hipSuccess)
        return bset_err(b, LVBGS_ERR_HIP, "view upload");
    b->ev_live[i] = true;
    StepOptions o = b->opt;
    o.timed_stage = 0;
    if (b->timed) {
        if (b->ev_used == b->ev_pairs.size()) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
                if (e0) (void)hipEventDestroy(e0);
                return bset_err(b, LVBGS_ERR_HIP, "timing events");
            }
            b->ev_pairs.emplace_back(e0, e1);
        }
        o.ev_begin = b->ev_pairs[b->ev_used].first;
        o.ev_end = b->ev_pairs[b->ev_used].second;
        ++b->ev_used;
    }
    const int G = b->planned_groups;
    if (b->timed == 1 || b->timed == 3) {
        // kernel-stage timing (one group): events between the launches, so the step is issued directly
        o.timed_stage = b->timed;
        batch_step(b->d_views, ns, b->grp[0].prefix, b->grp[0].total, o, b->stream);
        return hipGetLastError() == hipSuccess ? LVBGS_OK : bset_err(b, LVBGS_ERR_HIP, "batch launch");
    }
    if (b->timed == 2) (void)hipEventRecord(o.ev_begin, b->stream);
    if (G > 1) {
        // fork: the group streams start after the View upload
        if (hipEventRecord(b->ev_fork, b->stream) != hipSuccess) return bset_err(b, LVBGS_ERR_HIP, "fork");
        for (int g = 1; g < G; ++g)
            if (hipStreamWaitEvent(b->gstream[g], b->ev_fork, 0) != hipSuccess) return bset_err(b, LVBGS_ERR_HIP, "fork");
    }
    StepOptions og = b->opt;
    og.timed_stage = 0;
    for (int g = 0; g < G; ++g) {
        auto& gr = b->grp[g];
        hipStream_t st = b->gstream[g];
        const View* dv = b->d_views + gr.first;
        if (!b->use_graph) {
            batch_step(dv, gr.count, gr.prefix, gr.total, og, st);
            continue;
        }
        if (!gr.graph) {
            hipGraph_t gh = nullptr;
            if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return bset_err(b, LVBGS_ERR_HIP, "graph capture");
            batch_step(dv, gr.count, gr.prefix, gr.total, og, st);
            if (hipStreamEndCapture(st, &gh) != hipSuccess || !gh) return bset_err(b, LVBGS_ERR_HIP, "graph capture");
            const hipError_t e = hipGraphInstantiate(&gr.graph, gh, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gh);
            if (e != hipSuccess) {
                gr.graph = nullptr;
                return bset_err(b, LVBGS_ERR_HIP, "graph instantiate");
            }
        }
        if (hipGraphLaunch(gr.graph, st) != hipSuccess) return bset_err(b, LVBGS_ERR_HIP, "graph launch");
    }
    // join: the batch stream waits for every group
    for (int g = 1; g < G; ++g)
        if (hipEventRecord(b->ev_join[g], b->gstream[g]) != hipSuccess ||
            hipStreamWaitEvent(b->stream, b->ev_join[g], 0) != hipSuccess)
            return bset_err(b, LVBGS_ERR_HIP, "join");
    if (b->timed == 2) (void)hipEventRecord(o.ev_end, b->stream);
    return hipGetLastError() == hipSuccess ? LVBGS_OK : bset_err(b, LVBGS_ERR_HIP, "batch launch");
}

extern "C" int lvbgs_batch_synchronize(lvbgs_batch* b) {
    if (!b) return LVBGS_ERR_ARG;
    (void)hipSetDevice(b->device);
    if (hipStreamSynchronize(b->stream) != hipSuccess) return bset_err(b, LVBGS_ERR_HIP, "synchronize");
    for (auto* c : b->ctx) {
        uint32_t flags = 0;
        if (hipMemcpy(&flags, &c->v.sc->err_flags, sizeof(flags), hipMemcpyDeviceToHost) != hipSuccess)
            return bset_err(b, LVBGS_ERR_HIP, "error flags");
        if (flags & kErrFixpointCap) return bset_err(b, LVBGS_ERR_DEVICE, "fix-point iteration cap reached");
    }
    return LVBGS_OK;
}

extern "C" int lvbgs_batch_set_timing(lvbgs_batch* b, const char* stage) {
    if (!b) return LVBGS_ERR_ARG;
    const std::string s = stage ? stage : "";
    if (s.empty()) b->timed = 0;
    else if (s == "match") b->timed = 1;
    else if (s == "step") b->timed = 2;
    else if (s == "update") b->timed = 3;
    else return bset_err(b, LVBGS_ERR_ARG, "timing stage must be \"match\", \"update\", \"step\" or \"\"");
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    b->ev_used = 0;
    return LVBGS_OK;
}

extern "C" int lvbgs_batch_take_timing(lvbgs_batch* b, double* total_ms, int* count) {
    if (!b || !total_ms || !count) return LVBGS_ERR_ARG;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return bset_err(b, LVBGS_ERR_HIP, "sync");
    double tot = 0.0;
    for (size_t i = 0; i < b->ev_used; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, b->ev_pairs[i].first, b->ev_pairs[i].second) != hipSuccess)
            return bset_err(b, LVBGS_ERR_HIP, "event timing");
        tot += ms;
    }
    *total_ms = tot;
    *count = (int)b->ev_used;
    b->ev_used = 0;
    return LVBGS_OK;
}
