/* Runs the CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer (built by
 * tests/test_oracle_sanitize.py): SuBSENSE and LOBSTER, 1 and 3 channels, with and without a ROI,
 * through initialize / apply / refreshModel / getters / state copies, on moving-block scenes long
 * enough to trigger a model reset, plus the primitives used on their own by the tests. Any invalid
 * access or undefined operation aborts with a non-zero status. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

static void scene(uint8_t* img, int w, int h, int c, int t) {
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int k = 0; k < c; ++k) {
                int v = 40 + ((x * 7 + y * 3 + k * 11) % 90) + ((t * 13 + x) % 5);
                const int bx = (t * 3) % w, by = (t * 2) % h;
                if (x >= bx && x < bx + w / 5 && y >= by && y < by + h / 4) v = 220 - k * 30;
                if (t >= 40 && t < 44) v = v / 3;  /* global change: the SuBSENSE controller resets */
                img[((size_t)y * w + x) * c + k] = (uint8_t)v;
            }
}

static int run(int algo, int w, int h, int c, int with_roi, int frames) {
    orc_params p;
    orc_default_params(algo, &p);
    p.stream_id = (uint32_t)(algo * 10 + c);
    orc_bgs* o = orc_create(algo, &p);
    if (!o) return 1;
    const size_t npx = (size_t)w * h;
    uint8_t* img = malloc(npx * c);
    uint8_t* mask = malloc(npx);
    uint8_t* roi = NULL;
    if (with_roi) {
        roi = calloc(npx, 1);
        for (int y = h / 6; y < h - h / 5; ++y)
            for (int x = w / 8; x < w - w / 7; ++x) roi[(size_t)y * w + x] = 255;
    }
    scene(img, w, h, c, 0);
    if (orc_initialize(o, img, w, h, c, roi)) return 2;
    const double lr = algo == 1 ? 0.0 : 16.0;
    for (int t = 0; t < frames; ++t) {
        scene(img, w, h, c, t);
        if (orc_apply(o, img, mask, lr)) return 3;
        if (t == frames / 2 && orc_refresh_model(o, 0.3f, 1)) return 4;
    }
    uint8_t* bg = malloc(npx * c);
    uint16_t* bgd = malloc(npx * c * 2);
    if (orc_get_background_image(o, bg) || orc_get_background_descriptors_image(o, bgd)) return 5;
    static const char* names[] = {"samples_color", "samples_desc", "last_color", "last_desc", "last_fg", "roi", "lut",
                                  "T", "R", "V", "MLD", "DminLT", "DminST", "RawLT", "RawST", "FinLT", "FinST",
                                  "unstable", "blinks", "last_raw", "last_raw_blink", "last_dil_inv", "dsLT", "dsST"};
    const size_t cap = (size_t)p.bg_samples * npx * c * 2 + 4096;
    void* buf = malloc(cap);
    for (size_t i = 0; i < sizeof(names) / sizeof(names[0]); ++i) (void)orc_state(o, names[i], buf, cap);
    double sc[32];
    (void)orc_scalars(o, sc, 32);
    uint64_t ex = 0, wr = 0;
    orc_counters(o, &ex, &wr);
    /* re-initialize on the same geometry without a ROI: the stored ROI is reused */
    scene(img, w, h, c, 1);
    if (orc_initialize(o, img, w, h, c, NULL) || orc_apply(o, img, mask, lr)) return 6;
    orc_destroy(o);
    free(img), free(mask), free(roi), free(bg), free(bgd), free(buf);
    printf("algo %d %dx%dx%d roi %d: %d frames, examined %llu writes %llu\n", algo, w, h, c, with_roi, frames,
           (unsigned long long)ex, (unsigned long long)wr);
    return 0;
}

static int primitives(void) {
    const int w = 37, h = 29;
    uint8_t in[37 * 29], out[37 * 29], holes[37 * 29], img[37 * 29 * 3], small[9 * 7 * 3];
    uint16_t desc[37 * 29 * 3];
    for (int i = 0; i < w * h; ++i) in[i] = ((i * 2654435761u) >> 13) & 1 ? 255 : 0;
    for (int i = 0; i < w * h * 3; ++i) img[i] = (uint8_t)((i * 40503u) >> 7);
    orc_morph(in, out, w, h, 3, 1);
    orc_morph(in, out, w, h, 1, 0);
    orc_median_binary(in, out, w, h, 9);
    orc_flood_holes(in, holes, w, h);
    orc_lbsp_compute_abs(img, w, h, 3, 20, desc);
    orc_lbsp_compute_abs(img, w, h, 1, 20, desc);
    orc_resize_area(img, 36, 28, 3, small, 9, 7);
    uint64_t counts[6];
    orc_binclassif(in, out, holes, w, h, (size_t)w, counts);
    orc_binclassif(in, NULL, NULL, w, h, (size_t)w, counts);
    for (int r = 0; r < 512; ++r) {
        int ox, oy;
        orc_sample_pos_7x7(r, 2, 2, 2, w, h, &ox, &oy);
        orc_neighbor_pos_3x3(r, 0, 0, 0, w, h, &ox, &oy);
        orc_neighbor_pos_5x5(r, w - 1, h - 1, 0, w, h, &ox, &oy);
    }
    return 0;
}

int main(void) {
    int rc = primitives();
    if (!rc) rc = run(1, 96, 72, 3, 0, 60);
    if (!rc) rc = run(1, 96, 72, 3, 1, 30);
    if (!rc) rc = run(1, 80, 60, 1, 1, 30);
    if (!rc) rc = run(2, 96, 72, 3, 0, 30);
    if (!rc) rc = run(2, 80, 60, 1, 1, 30);
    if (rc) fprintf(stderr, "oracle run failed at step %d\n", rc);
    return rc;
}
