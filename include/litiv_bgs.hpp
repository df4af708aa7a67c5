/*
 * litiv_bgs.hpp -- header-only C++ mirror of litiv's IIBackgroundSubtractor over the C ABI of
 * litiv_bgs.h, method for method, with plain image views in place of cv::Mat / cv::InputArray.
 *
 * Reference interface: modules/video/include/litiv/video/BackgroundSubtractionUtils.hpp:24-55
 * (IIBackgroundSubtractor), BackgroundSubtractorLBSP.hpp:41 (getBackgroundDescriptorsImage),
 * BackgroundSubtractorSuBSENSE.hpp:50-68 and BackgroundSubtractorLOBSTER.hpp:50-57 (constructors
 * with their default parameters, refreshModel, getDefaultLearningRate). Errors throw
 * litiv_amd::Exception carrying the library's message, as the reference's lvAssert_ throws
 * lv::Exception. Link with liblitiv_bgs.so.
 */
#ifndef LITIV_BGS_HPP
#define LITIV_BGS_HPP
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "litiv_bgs.h"

namespace litiv_amd {

/* 8-bit interleaved image (CV_8UC1 / CV_8UC3 equivalent); stride in bytes, 0 = tightly packed */
struct ImageView {
    const uint8_t* data = nullptr;
    int width = 0, height = 0, channels = 1;
    size_t stride = 0;
};

/* owning image returned by the getters (u8 or u16 elements) */
template <typename T>
struct Image {
    std::vector<T> data;
    int width = 0, height = 0, channels = 1;
    ImageView view() const {
        static_assert(sizeof(T) == 1, "views are 8-bit");
        return ImageView{reinterpret_cast<const uint8_t*>(data.data()), width, height, channels, 0};
    }
};

class Exception : public std::runtime_error {
public:
    Exception(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

/* IIBackgroundSubtractor + IBackgroundSubtractorLBSP surface */
class BackgroundSubtractor {
public:
    BackgroundSubtractor(const BackgroundSubtractor&) = delete;
    BackgroundSubtractor& operator=(const BackgroundSubtractor&) = delete;
    virtual ~BackgroundSubtractor() { lvbgs_destroy(ctx_); }

    /* initialize(oInitImg) / initialize(oInitImg, oROI) (BackgroundSubtractionUtils.hpp:29-31) */
    void initialize(const ImageView& img) { initialize(img, nullptr); }
    void initialize(const ImageView& img, const ImageView& roi) { initialize(img, &roi); }

    /* apply(image, fgmask, learningRate) (BackgroundSubtractionUtils.hpp:35); fgmask is resized to W x H */
    void apply(const ImageView& img, Image<uint8_t>& fgmask, double learningRate) {
        fgmask.width = w_, fgmask.height = h_, fgmask.channels = 1;
        fgmask.data.resize((size_t)w_ * h_);
        check(lvbgs_apply(ctx_, img.data, img.stride, fgmask.data.data(), (size_t)w_, learningRate));
    }
    void apply(const ImageView& img, Image<uint8_t>& fgmask) { apply(img, fgmask, getDefaultLearningRate()); }

    /* device-resident variant: asynchronous on stream() */
    void applyDevice(const void* d_img, void* d_mask, double learningRate) {
        check(lvbgs_apply_device(ctx_, d_img, d_mask, learningRate));
    }
    void synchronize() { check(lvbgs_synchronize(ctx_)); }

    /* getBackgroundImage (BackgroundSubtractionUtils.hpp:37) */
    void getBackgroundImage(Image<uint8_t>& out) const {
        alloc(out, c_);
        check(lvbgs_get_background_image(ctx_, out.data.data(), (size_t)w_ * c_));
    }
    /* getBackgroundDescriptorsImage (BackgroundSubtractorLBSP.hpp:41) */
    void getBackgroundDescriptorsImage(Image<uint16_t>& out) const {
        alloc(out, c_);
        check(lvbgs_get_background_descriptors_image(ctx_, out.data.data(), (size_t)w_ * c_ * 2));
    }
    virtual double getDefaultLearningRate() const { return lvbgs_default_learning_rate(ctx_); }
    /* setAutomaticModelReset / validateROI / setROI / getROICopy (BackgroundSubtractionUtils.hpp:39-45) */
    void setAutomaticModelReset(bool enabled) { check(lvbgs_set_automatic_model_reset(ctx_, enabled ? 1 : 0)); }
    static void validateROI(Image<uint8_t>& roi) {
        if (lvbgs_validate_roi(roi.width, roi.height, roi.data.data(), (size_t)roi.width) != LVBGS_OK)
            throw Exception(LVBGS_ERR_ARG, "provided ROI must be non-empty and of type 8UC1");
    }
    void setROI(Image<uint8_t>& roi) {
        check(lvbgs_set_roi(ctx_, roi.data.data(), (size_t)roi.width));
        w_ = roi.width, h_ = roi.height;
    }
    Image<uint8_t> getROICopy() const {
        Image<uint8_t> out;
        alloc(out, 1);
        check(lvbgs_get_roi(ctx_, out.data.data(), (size_t)w_));
        return out;
    }

    /* Checkpoint / resume of the whole model as binary matrix archives (lvbgs_checkpoint_save / _load);
     * no counterpart in the reference, which keeps its model in memory only. save needs an existing
     * directory; load needs a subtractor constructed with the same parameters and stream id. */
    void saveCheckpoint(const std::string& dir) const { check(lvbgs_checkpoint_save(ctx_, dir.c_str())); }
    void loadCheckpoint(const std::string& dir) {
        check(lvbgs_checkpoint_load(ctx_, dir.c_str()));
        lvbgs_stats s;
        check(lvbgs_get_stats(ctx_, &s));
        w_ = (int)s.width, h_ = (int)s.height, c_ = (int)s.channels;
    }

    lvbgs_ctx* handle() const { return ctx_; }
    void* stream() const { return lvbgs_stream(ctx_); }

protected:
    BackgroundSubtractor(int algo, const lvbgs_params& p) {
        const int rc = lvbgs_create(algo, &p, &ctx_);
        if (rc != LVBGS_OK) throw Exception(rc, "lvbgs_create failed");
    }
    void check(int rc) const {
        if (rc < 0) throw Exception(rc, lvbgs_last_error(ctx_));
    }
    template <typename T>
    void alloc(Image<T>& out, int channels) const {
        out.width = w_, out.height = h_, out.channels = channels;
        out.data.resize((size_t)w_ * h_ * channels);
    }
    void initialize(const ImageView& img, const ImageView* roi) {
        check(lvbgs_initialize(ctx_, img.data, img.width, img.height, img.channels, img.stride,
                               roi ? roi->data : nullptr, roi ? roi->stride : 0));
        w_ = img.width, h_ = img.height, c_ = img.channels;
    }
    lvbgs_ctx* ctx_ = nullptr;
    int w_ = 0, h_ = 0, c_ = 0;
};

/* BackgroundSubtractorSuBSENSE (BackgroundSubtractorSuBSENSE.hpp:50-68) */
class BackgroundSubtractorSuBSENSE : public BackgroundSubtractor {
public:
    explicit BackgroundSubtractorSuBSENSE(size_t nDescDistThresholdOffset = 3, size_t nMinColorDistThreshold = 30,
                                          size_t nBGSamples = 50, size_t nRequiredBGSamples = 2,
                                          size_t nSamplesForMovingAvgs = 100, float fRelLBSPThreshold = 0.333f,
                                          uint32_t streamId = 0, int device = 0)
        : BackgroundSubtractor(LVBGS_SUBSENSE, params(nDescDistThresholdOffset, nMinColorDistThreshold, nBGSamples,
                                                      nRequiredBGSamples, nSamplesForMovingAvgs, fRelLBSPThreshold,
                                                      streamId, device)) {}
    /* refreshModel(fSamplesRefreshFrac, bForceFGUpdate) (.hpp:59) */
    void refreshModel(float fSamplesRefreshFrac, bool bForceFGUpdate = false) {
        check(lvbgs_refresh_model(ctx_, fSamplesRefreshFrac, bForceFGUpdate ? 1 : 0));
    }
    double getDefaultLearningRate() const override { return 0; }

private:
    static lvbgs_params params(size_t desc, size_t color, size_t n, size_t req, size_t mv, float rel, uint32_t sid,
                               int dev) {
        lvbgs_params p;
        lvbgs_default_params(LVBGS_SUBSENSE, &p);
        p.desc_dist_threshold = desc, p.color_dist_threshold = color, p.bg_samples = n;
        p.required_bg_samples = req, p.samples_for_mv_avgs = mv, p.rel_lbsp_threshold = rel;
        p.stream_id = sid, p.device = dev;
        return p;
    }
};

/* BackgroundSubtractorLOBSTER (BackgroundSubtractorLOBSTER.hpp:50-57) */
class BackgroundSubtractorLOBSTER : public BackgroundSubtractor {
public:
    explicit BackgroundSubtractorLOBSTER(size_t nDescDistThreshold = 4, size_t nColorDistThreshold = 30,
                                         size_t nBGSamples = 35, size_t nRequiredBGSamples = 2,
                                         size_t nLBSPThresholdOffset = 0, float fRelLBSPThreshold = 0.333f,
                                         uint32_t streamId = 0, int device = 0)
        : BackgroundSubtractor(LVBGS_LOBSTER, params(nDescDistThreshold, nColorDistThreshold, nBGSamples,
                                                     nRequiredBGSamples, nLBSPThresholdOffset, fRelLBSPThreshold,
                                                     streamId, device)) {}
    void refreshModel(float fSamplesRefreshFrac, bool bForceFGUpdate = false) {
        check(lvbgs_refresh_model(ctx_, fSamplesRefreshFrac, bForceFGUpdate ? 1 : 0));
    }
    double getDefaultLearningRate() const override { return 16; }

private:
    static lvbgs_params params(size_t desc, size_t color, size_t n, size_t req, size_t off, float rel, uint32_t sid,
                               int dev) {
        lvbgs_params p;
        lvbgs_default_params(LVBGS_LOBSTER, &p);
        p.desc_dist_threshold = desc, p.color_dist_threshold = color, p.bg_samples = n;
        p.required_bg_samples = req, p.lbsp_threshold_offset = off, p.rel_lbsp_threshold = rel;
        p.stream_id = sid, p.device = dev;
        return p;
    }
};

}  // namespace litiv_amd
#endif
