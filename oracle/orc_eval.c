/*
 * orc_eval.c -- binary classification counts (lv::BinClassif::accumulate,
 * modules/datasets/src/metrics.cpp:21-50; labels of modules/datasets/include/litiv/datasets/metrics.hpp:23-28).
 *
 * TEST INFRASTRUCTURE (oracle): the checker of the device evaluator. Restated pixel by pixel in the
 * reference's order of tests.
 */
#include "orc.h"

#define ORC_POSITIVE 255u
#define ORC_NEGATIVE 0u
#define ORC_OUTOFSCOPE 85u
#define ORC_UNKNOWN 170u
#define ORC_SHADOW 50u

void orc_binclassif(const uint8_t* classif, const uint8_t* gt, const uint8_t* roi, int w, int h, size_t stride,
                    uint64_t out[6]) {
    /* out: TP, TN, FP, FN, SE, DC (accumulated) */
    if (!gt) { /* metrics.cpp:26-29 */
        out[5] += (uint64_t)w * (uint64_t)h;
        return;
    }
    for (int y = 0; y < h; ++y) {
        const uint8_t* in = classif + (size_t)y * stride;
        const uint8_t* g = gt + (size_t)y * stride;
        const uint8_t* r = roi ? roi + (size_t)y * stride : 0;
        for (int x = 0; x < w; ++x) {
            if (g[x] != ORC_OUTOFSCOPE && g[x] != ORC_UNKNOWN && (!r || r[x] != ORC_NEGATIVE)) {
                if (in[x] == ORC_POSITIVE) {
                    if (g[x] == ORC_POSITIVE) ++out[0];
                    else ++out[2];
                } else {
                    if (g[x] == ORC_POSITIVE) ++out[3];
                    else ++out[1];
                }
                if (g[x] == ORC_SHADOW && in[x] == ORC_POSITIVE) ++out[4];
            } else {
                ++out[5];
            }
        }
    }
}
