// mirror_run.cpp -- runs a sequence through the header-only C++ mirror of IIBackgroundSubtractor
// (include/litiv_bgs.hpp) so that the GPU tests execute the mirror itself, not only the C ABI.
//
// usage: mirror_run <algo: 1 SuBSENSE | 2 LOBSTER> <width> <height> <channels> <frames> <in.raw> <out.raw>
//                   <refresh_at> [checkpoint_dir]
// in.raw: the frames, tightly packed. out.raw: the mask of every apply (width x height each), then
// getBackgroundImage (width x height x channels), getBackgroundDescriptorsImage (16-bit), then the
// int32 error code the mirror threw for an apply on an image of the wrong size (0 if none).
// refreshModel(0.1) is called after apply number refresh_at (SuBSENSE), refreshModel(0.5) for LOBSTER.
// With checkpoint_dir (an existing directory), the model is saved there after apply number
// refresh_at + 2 and the rest of the sequence runs on a second subtractor that loaded it.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "litiv_bgs.hpp"

int main(int argc, char** argv) {
    if (argc != 9 && argc != 10) {
        std::fprintf(stderr, "usage: %s algo width height channels frames in.raw out.raw refresh_at [checkpoint_dir]\n",
                     argv[0]);
        return 2;
    }
    const int algo = std::atoi(argv[1]), w = std::atoi(argv[2]), h = std::atoi(argv[3]), c = std::atoi(argv[4]);
    const int n = std::atoi(argv[5]), refresh_at = std::atoi(argv[8]);
    const size_t fb = (size_t)w * h * c;
    std::vector<uint8_t> frames(fb * (size_t)n);
    FILE* f = std::fopen(argv[6], "rb");
    if (!f || std::fread(frames.data(), 1, frames.size(), f) != frames.size()) return 3;
    std::fclose(f);
    auto make = [algo]() -> std::unique_ptr<litiv_amd::BackgroundSubtractor> {
        if (algo == 1) return std::unique_ptr<litiv_amd::BackgroundSubtractor>(new litiv_amd::BackgroundSubtractorSuBSENSE());
        return std::unique_ptr<litiv_amd::BackgroundSubtractor>(new litiv_amd::BackgroundSubtractorLOBSTER());
    };
    std::unique_ptr<litiv_amd::BackgroundSubtractor> bgs = make();
    std::vector<uint8_t> out;
    try {
        litiv_amd::ImageView v{frames.data(), w, h, c, 0};
        bgs->initialize(v);
        litiv_amd::Image<uint8_t> mask;
        for (int t = 0; t < n; ++t) {
            v.data = frames.data() + fb * (size_t)t;
            bgs->apply(v, mask);
            out.insert(out.end(), mask.data.begin(), mask.data.end());
            if (t == refresh_at) {
                if (algo == 1) static_cast<litiv_amd::BackgroundSubtractorSuBSENSE*>(bgs.get())->refreshModel(0.1f);
                else static_cast<litiv_amd::BackgroundSubtractorLOBSTER*>(bgs.get())->refreshModel(0.5f);
            }
            if (argc == 10 && t == refresh_at + 2) {
                bgs->saveCheckpoint(argv[9]);
                std::unique_ptr<litiv_amd::BackgroundSubtractor> resumed = make();
                resumed->loadCheckpoint(argv[9]);
                bgs = std::move(resumed);
            }
        }
        litiv_amd::Image<uint8_t> bg;
        litiv_amd::Image<uint16_t> desc;
        bgs->getBackgroundImage(bg);
        bgs->getBackgroundDescriptorsImage(desc);
        out.insert(out.end(), bg.data.begin(), bg.data.end());
        const uint8_t* d = reinterpret_cast<const uint8_t*>(desc.data.data());
        out.insert(out.end(), d, d + desc.data.size() * 2);
    } catch (const litiv_amd::Exception& e) {
        std::fprintf(stderr, "mirror_run: %s (code %d)\n", e.what(), e.code());
        return 4;
    }
    int32_t code = 0;
    try {
        litiv_amd::Image<uint8_t> mask;
        bgs->apply(litiv_amd::ImageView{nullptr, w, h, c, 0}, mask);  // the reference asserts on a bad input
    } catch (const litiv_amd::Exception& e) {
        code = e.code();
    }
    const uint8_t* cb = reinterpret_cast<const uint8_t*>(&code);
    out.insert(out.end(), cb, cb + 4);
    FILE* o = std::fopen(argv[7], "wb");
    if (!o || std::fwrite(out.data(), 1, out.size(), o) != out.size()) return 5;
    std::fclose(o);
    return 0;
}
