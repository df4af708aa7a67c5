// bgs_archive.hip -- the reference's binary matrix archive (lv::write / lv::read with
// MatArchive_BINARY, modules/utils/src/opencv.cpp:514-531 and :608-625) and checkpoint / resume of a
// context's whole model in that format (include/litiv_bgs.h, lvbgs_archive_*, lvbgs_checkpoint_*).
// Host code only.
//
// Archive layout (little-endian, as the reference writes it on x86): int32 OpenCV type
// (depth | (channels - 1) << 3), uint64 element size in bytes, uint64 element count, int32 number of
// dimensions, int32 size of each dimension, then the elements, densely packed.
//
// A checkpoint is a directory of such archives, one per state buffer, in the layout the state
// getters use (samples as N x H x W matrices of C-channel elements, maps as H x W), plus meta.bin
// (CV_64F): geometry, model parameters, RNG key and the frame-level scalars. Per-step scratch is not
// saved: every step rebuilds it. Loading a checkpoint into a context created with the same model
// parameters restores the state exactly, so the resumed sequence continues bit for bit.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/litiv_bgs.h"
#include "bgs_ctx.hpp"
#include "bgs_kernels.hpp"

using namespace lvbgs;

namespace {

constexpr int kCV8U = 0, kCV16U = 2, kCV32F = 5, kCV64F = 6;
constexpr int kCheckpointVersion = 1;

int cv_type(int depth, int channels) { return depth + ((channels - 1) << 3); }

size_t depth_bytes(int depth) {
    static const size_t b[8] = {1, 1, 2, 2, 4, 4, 8, 2};
    return b[depth & 7];
}

struct ArrayDesc {
    int type = 0, dims = 0;
    int32_t sizes[8] = {};
    uint64_t elem_size = 0, elem_count = 0;
};

bool write_archive(const std::string& path, int type, int dims, const int32_t* sizes, const void* data) {
    if (dims < 1 || dims > 32) return false;
    const uint64_t esz = (uint64_t)depth_bytes(type & 7) * (uint64_t)((type >> 3) + 1);
    uint64_t cnt = 1;
    for (int d = 0; d < dims; ++d) {
        if (sizes[d] < 0) return false;
        cnt *= (uint64_t)sizes[d];
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const int32_t t = type, nd = dims;
    bool ok = std::fwrite(&t, 4, 1, f) == 1 && std::fwrite(&esz, 8, 1, f) == 1 && std::fwrite(&cnt, 8, 1, f) == 1 &&
              std::fwrite(&nd, 4, 1, f) == 1 && std::fwrite(sizes, 4, (size_t)dims, f) == (size_t)dims;
    if (ok && esz * cnt > 0) ok = std::fwrite(data, 1, (size_t)(esz * cnt), f) == (size_t)(esz * cnt);
    return std::fclose(f) == 0 && ok;
}

// reads the header; with data != nullptr also the elements (at most cap bytes); returns the element
// bytes, or -1 (cannot open / truncated / inconsistent header) or -2 (data does not fit cap)
long long read_archive(const std::string& path, ArrayDesc& d, void* data, size_t cap) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return -1;
    int32_t t = 0, nd = 0;
    bool ok = std::fread(&t, 4, 1, f) == 1 && std::fread(&d.elem_size, 8, 1, f) == 1 &&
              std::fread(&d.elem_count, 8, 1, f) == 1 && std::fread(&nd, 4, 1, f) == 1 && nd >= 1 && nd <= 32;
    int32_t sizes[32] = {};
    if (ok) ok = std::fread(sizes, 4, (size_t)nd, f) == (size_t)nd;
    uint64_t cnt = 1;
    for (int i = 0; ok && i < nd; ++i) {
        ok = sizes[i] >= 0;
        cnt *= (uint64_t)sizes[i];
    }
    ok = ok && cnt == d.elem_count && d.elem_size == (uint64_t)depth_bytes(t & 7) * (uint64_t)((t >> 3) + 1);
    if (!ok) {
        std::fclose(f);
        return -1;
    }
    d.type = t;
    d.dims = nd;
    for (int i = 0; i < nd && i < 8; ++i) d.sizes[i] = sizes[i];
    const uint64_t bytes = d.elem_size * d.elem_count;
    if (data) {
        if (bytes > cap) {
            std::fclose(f);
            return -2;
        }
        if (bytes && std::fread(data, 1, (size_t)bytes, f) != (size_t)bytes) {
            std::fclose(f);
            return -1;
        }
    }
    std::fclose(f);
    return (long long)bytes;
}

// the model buffers of a checkpoint: name, depth, channels (0: the context's), dimensions
struct Item {
    const char* name;
    int depth, channels;
    bool samples, subsense_only, downsampled;
};
const Item kItems[] = {
    {"samples_color", kCV8U, 0, true, false, false},  {"samples_desc", kCV16U, 0, true, false, false},
    {"last_color", kCV8U, 0, false, false, false},    {"last_desc", kCV16U, 0, false, false, false},
    {"last_fg", kCV8U, 1, false, false, false},       {"roi", kCV8U, 1, false, false, false},
    {"T", kCV32F, 1, false, true, false},             {"R", kCV32F, 1, false, true, false},
    {"V", kCV32F, 1, false, true, false},             {"MLD", kCV32F, 1, false, true, false},
    {"DminLT", kCV32F, 1, false, true, false},        {"DminST", kCV32F, 1, false, true, false},
    {"RawLT", kCV32F, 1, false, true, false},         {"RawST", kCV32F, 1, false, true, false},
    {"FinLT", kCV32F, 1, false, true, false},         {"FinST", kCV32F, 1, false, true, false},
    {"unstable", kCV8U, 1, false, true, false},       {"blinks", kCV8U, 1, false, true, false},
    {"last_raw", kCV8U, 1, false, true, false},       {"last_raw_blink", kCV8U, 1, false, true, false},
    {"last_dil_inv", kCV8U, 1, false, true, false},   {"dsLT", kCV32F, 0, false, true, true},
    {"dsST", kCV32F, 0, false, true, true},
};

// meta.bin fields (CV_64F vector)
enum Meta : int {
    kMetaVersion, kMetaAlgo, kMetaW, kMetaH, kMetaC, kMetaN, kMetaREQ, kMetaDesc, kMetaColor, kMetaMv, kMetaLbspOff,
    kMetaRel, kMetaMedianParam, kMetaSeedLo, kMetaSeedHi, kMetaStream, kMetaApplyCount, kMetaFrameIdx,
    kMetaSinceReset, kMetaCooldown, kMetaAutoReset, kMetaLowerCap, kMetaUpperCap, kMetaLastNz, kMetaRefreshSeq,
    kMetaEpoch, kMetaNRoi, kMetaOrigRoi, kMetaCount
};

std::string join(const char* dir, const char* name) { return std::string(dir) + "/" + name + ".bin"; }

int set_err(lvbgs_ctx* c, int code, const std::string& msg) {
    c->err = msg;
    return code;
}

}  // namespace

extern "C" int lvbgs_archive_write(const char* path, int cv_type_code, int dims, const int32_t* sizes,
                                   const void* data) {
    if (!path || !sizes || (!data && dims > 0)) return LVBGS_ERR_ARG;
    if ((cv_type_code & 7) == 7 || cv_type_code < 0 || (cv_type_code >> 3) >= 512) return LVBGS_ERR_ARG;
    return write_archive(path, cv_type_code, dims, sizes, data) ? LVBGS_OK : LVBGS_ERR_ARG;
}

extern "C" long long lvbgs_archive_read(const char* path, int* cv_type_code, int* dims, int32_t* sizes, int max_dims,
                                        void* data, size_t cap) {
    if (!path) return LVBGS_ERR_ARG;
    ArrayDesc d;
    const long long r = read_archive(path, d, data, cap);
    if (r < 0) return LVBGS_ERR_ARG;
    if (cv_type_code) *cv_type_code = d.type;
    if (dims) *dims = d.dims;
    if (sizes)
        for (int i = 0; i < d.dims && i < max_dims && i < 8; ++i) sizes[i] = d.sizes[i];
    return r;
}

extern "C" int lvbgs_checkpoint_save(lvbgs_ctx* c, const char* dir) {
    if (!c || !dir) return LVBGS_ERR_ARG;
    if (!c->initialized) return set_err(c, LVBGS_ERR_STATE, "algo must be initialized first");
    const bool sub = c->algo == LVBGS_SUBSENSE;
    std::vector<double> meta(kMetaCount, 0.0);
    meta[kMetaVersion] = kCheckpointVersion;
    meta[kMetaAlgo] = c->algo;
    meta[kMetaW] = c->W, meta[kMetaH] = c->H, meta[kMetaC] = c->C;
    meta[kMetaN] = (double)c->p.bg_samples, meta[kMetaREQ] = (double)c->p.required_bg_samples;
    meta[kMetaDesc] = (double)c->p.desc_dist_threshold, meta[kMetaColor] = (double)c->p.color_dist_threshold;
    meta[kMetaMv] = (double)c->p.samples_for_mv_avgs, meta[kMetaLbspOff] = (double)c->p.lbsp_threshold_offset;
    meta[kMetaRel] = (double)c->p.rel_lbsp_threshold, meta[kMetaMedianParam] = c->p.median_blur_kernel;
    meta[kMetaSeedLo] = (double)(uint32_t)c->p.rng_seed, meta[kMetaSeedHi] = (double)(uint32_t)(c->p.rng_seed >> 32);
    meta[kMetaStream] = c->p.stream_id;
    meta[kMetaApplyCount] = c->apply_count;
    meta[kMetaNRoi] = (double)c->n_roi, meta[kMetaOrigRoi] = (double)c->orig_roi;
    {
        // the raw scalars (epoch and the exact float caps included)
        (void)hipSetDevice(c->device);
        Scalars s;
        if (hipStreamSynchronize(c->stream) != hipSuccess ||
            hipMemcpy(&s, c->v.sc, sizeof(Scalars), hipMemcpyDeviceToHost) != hipSuccess)
            return set_err(c, LVBGS_ERR_HIP, "scalars");
        meta[kMetaFrameIdx] = s.frame_idx, meta[kMetaSinceReset] = s.frames_since_reset;
        meta[kMetaCooldown] = s.cooldown, meta[kMetaAutoReset] = s.auto_reset;
        meta[kMetaLowerCap] = s.lower_cap, meta[kMetaUpperCap] = s.upper_cap;
        meta[kMetaLastNz] = s.last_nz_ratio, meta[kMetaRefreshSeq] = s.refresh_seq, meta[kMetaEpoch] = s.epoch;
        const int32_t n = 256;
        if (!write_archive(join(dir, "lut"), cv_type(kCV8U, 1), 1, &n, s.lut))
            return set_err(c, LVBGS_ERR_ARG, std::string("cannot write ") + join(dir, "lut"));
    }
    const int32_t nm = kMetaCount;
    if (!write_archive(join(dir, "meta"), cv_type(kCV64F, 1), 1, &nm, meta.data()))
        return set_err(c, LVBGS_ERR_ARG, std::string("cannot write ") + join(dir, "meta"));
    const size_t N = c->p.bg_samples;
    std::vector<uint8_t> buf;
    for (const Item& it : kItems) {
        if (it.subsense_only && !sub) continue;
        const int ch = it.channels ? it.channels : c->C;
        int32_t sizes[3];
        int dims;
        if (it.samples) sizes[0] = (int32_t)N, sizes[1] = c->H, sizes[2] = c->W, dims = 3;
        else if (it.downsampled) sizes[0] = c->H / 8, sizes[1] = c->W / 8, dims = 2;
        else sizes[0] = c->H, sizes[1] = c->W, dims = 2;
        size_t bytes = depth_bytes(it.depth) * (size_t)ch;
        for (int d = 0; d < dims; ++d) bytes *= (size_t)sizes[d];
        buf.resize(bytes ? bytes : 1);
        const long long r = lvbgs_get_state(c, it.name, buf.data(), buf.size());
        if (r != (long long)bytes) return set_err(c, LVBGS_ERR_HIP, std::string("state ") + it.name);
        if (!write_archive(join(dir, it.name), cv_type(it.depth, ch), dims, sizes, buf.data()))
            return set_err(c, LVBGS_ERR_ARG, std::string("cannot write ") + join(dir, it.name));
    }
    return LVBGS_OK;
}

// sets up the context for the checkpoint's geometry and model (bgs_api.hip); the caller then uploads
// the saved buffers over the fresh state
int lvbgs_restore_begin(lvbgs_ctx* c, int w, int h, int ch, const uint8_t* roi, size_t n_roi, size_t orig_roi);
int lvbgs_restore_buffer(lvbgs_ctx* c, const char* name, const void* data, size_t bytes);
int lvbgs_restore_scalars(lvbgs_ctx* c, const Scalars& s, uint32_t apply_count);

static int load_buffers(lvbgs_ctx* c, const char* dir, int W, int H, int C, const std::vector<double>& meta);

extern "C" int lvbgs_checkpoint_load(lvbgs_ctx* c, const char* dir) {
    if (!c || !dir) return LVBGS_ERR_ARG;
    c->err.clear();
    std::vector<double> meta(kMetaCount, 0.0);
    ArrayDesc d;
    if (read_archive(join(dir, "meta"), d, meta.data(), meta.size() * sizeof(double)) < 0 || d.type != kCV64F ||
        d.elem_count < (uint64_t)kMetaCount)
        return set_err(c, LVBGS_ERR_ARG, std::string("cannot read ") + join(dir, "meta"));
    if ((int)meta[kMetaVersion] != kCheckpointVersion) return set_err(c, LVBGS_ERR_ARG, "unknown checkpoint version");
    // the checkpoint belongs to a model of the same algorithm and parameters
    const lvbgs_params& p = c->p;
    const uint64_t seed = (uint64_t)(uint32_t)meta[kMetaSeedLo] | ((uint64_t)(uint32_t)meta[kMetaSeedHi] << 32);
    if ((int)meta[kMetaAlgo] != c->algo || (size_t)meta[kMetaN] != p.bg_samples ||
        (size_t)meta[kMetaREQ] != p.required_bg_samples || (size_t)meta[kMetaDesc] != p.desc_dist_threshold ||
        (size_t)meta[kMetaColor] != p.color_dist_threshold || (size_t)meta[kMetaMv] != p.samples_for_mv_avgs ||
        (size_t)meta[kMetaLbspOff] != p.lbsp_threshold_offset || (float)meta[kMetaRel] != p.rel_lbsp_threshold ||
        (int)meta[kMetaMedianParam] != p.median_blur_kernel || seed != p.rng_seed ||
        (uint32_t)meta[kMetaStream] != p.stream_id)
        return set_err(c, LVBGS_ERR_ARG, "checkpoint of a model with other parameters or RNG key");
    const int W = (int)meta[kMetaW], H = (int)meta[kMetaH], C = (int)meta[kMetaC];
    if (W <= 0 || H <= 0 || (C != 1 && C != 3)) return set_err(c, LVBGS_ERR_ARG, "bad checkpoint geometry");
    const size_t npx = (size_t)W * (size_t)H;
    std::vector<uint8_t> roi(npx);
    if (read_archive(join(dir, "roi"), d, roi.data(), roi.size()) != (long long)npx)
        return set_err(c, LVBGS_ERR_ARG, std::string("cannot read ") + join(dir, "roi"));
    int rc = lvbgs_restore_begin(c, W, H, C, roi.data(), (size_t)meta[kMetaNRoi], (size_t)meta[kMetaOrigRoi]);
    if (rc) return rc;
    rc = load_buffers(c, dir, W, H, C, meta);
    if (rc) std::vector<uint32_t>().swap(c->restore_stage);  // the context stays uninitialized
    return rc;
}

// the saved buffers and scalars of a checkpoint whose geometry lvbgs_restore_begin has set up
static int load_buffers(lvbgs_ctx* c, const char* dir, int W, int H, int C, const std::vector<double>& meta) {
    const size_t npx = (size_t)W * (size_t)H, N = c->p.bg_samples;
    const bool sub = c->algo == LVBGS_SUBSENSE;
    ArrayDesc d;
    int rc;
    std::vector<uint8_t> buf;
    for (const Item& it : kItems) {
        if (it.subsense_only && !sub) continue;
        if (!std::strcmp(it.name, "roi")) continue;
        const int ch = it.channels ? it.channels : C;
        size_t bytes = depth_bytes(it.depth) * (size_t)ch *
                       (it.samples ? N * npx : it.downsampled ? (size_t)(W / 8) * (size_t)(H / 8) : npx);
        buf.resize(bytes ? bytes : 1);
        const long long r = read_archive(join(dir, it.name), d, buf.data(), buf.size());
        if (r != (long long)bytes || d.type != cv_type(it.depth, ch))
            return set_err(c, LVBGS_ERR_ARG, std::string("cannot read ") + join(dir, it.name));
        if ((rc = lvbgs_restore_buffer(c, it.name, buf.data(), bytes))) return rc;
    }
    Scalars s;
    std::memset(&s, 0, sizeof(s));
    if (read_archive(join(dir, "lut"), d, s.lut, sizeof(s.lut)) != 256)
        return set_err(c, LVBGS_ERR_ARG, std::string("cannot read ") + join(dir, "lut"));
    s.frame_idx = (uint32_t)meta[kMetaFrameIdx];
    s.frames_since_reset = (uint32_t)meta[kMetaSinceReset];
    s.cooldown = (uint32_t)meta[kMetaCooldown];
    s.auto_reset = (int32_t)meta[kMetaAutoReset];
    s.lower_cap = (float)meta[kMetaLowerCap];
    s.upper_cap = (float)meta[kMetaUpperCap];
    s.last_nz_ratio = (float)meta[kMetaLastNz];
    s.refresh_seq = (uint32_t)meta[kMetaRefreshSeq];
    s.epoch = (uint32_t)meta[kMetaEpoch];
    return lvbgs_restore_scalars(c, s, (uint32_t)meta[kMetaApplyCount]);
}
