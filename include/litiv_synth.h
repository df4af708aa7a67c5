/*
 * litiv_synth.h -- deterministic synthetic video sequences for the background-subtraction
 * workloads (benchmarks and tests). Not part of the subtractor ABI; the reference has no
 * counterpart (its benchmarks read datasets from disk).
 *
 * Every frame is a pure integer function of (config, stream seed, frame index, pixel), so the
 * host generator and the GPU generator produce bit-identical frames.
 */
#ifndef LITIV_SYNTH_H
#define LITIV_SYNTH_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LVSYNTH_MAX_OBJECTS 32

typedef struct lvsynth_config {
    int width, height, channels, frames;
    uint32_t seed;
    int n_objects, obj_min, obj_max;  /* object side range (px) */
    int speed_min, speed_max;         /* per-axis speed range (px/frame), random sign */
    int noise_sigma;                  /* additive noise std-dev (grey levels) */
    int bg_lo, bg_hi;                 /* value-noise background range */
    int bg_cell, octaves;             /* coarsest value-noise cell (px), number of octaves */
    int drift;                        /* global illumination ramp from -drift to +drift over the sequence */
    int flicker_size, flicker_amp, flicker_pmin, flicker_pmax; /* dynamic-background patch */
    int shape;                        /* 0 = axis-aligned rectangles, 1 = ellipses */
    int obj_lo, obj_hi;               /* object value range */
} lvsynth_config;

typedef struct lvsynth_object {
    int x0, y0, w, h, vx, vy;
    int color[3];
} lvsynth_object;

typedef struct lvsynth_scene {
    lvsynth_config cfg;
    int n_obj;
    lvsynth_object obj[LVSYNTH_MAX_OBJECTS];
    int fx0, fy0; /* flicker patch origin */
} lvsynth_scene;

/* Workload presets. stream selects the per-stream seed (and, for the batched and serving
 * presets, the resolution). Returns 0, or -1 for an unknown preset. */
enum {
    LVSYNTH_SUBSENSE_QVGA = 0,    /* 320x240x3, 300 frames, 3 rectangles 20-60 px, sigma 3 */
    LVSYNTH_SUBSENSE_1080P = 1,   /* 1920x1080x3, 1000 frames, 12 rectangles 40-200 px, sigma 4 */
    LVSYNTH_SUBSENSE_4K = 2,      /* 3840x2160x3, 500 frames, 20 rectangles 40-200 px, sigma 4 */
    LVSYNTH_LOBSTER_THERMAL = 3,  /* 640x512x1, 1000 frames, 5 warm ellipses, sigma 2 */
    LVSYNTH_BATCH64 = 4,          /* stream 0-15 QVGA, 16-31 VGA, 32-47 640x512 grey, 48-63 1080p; 600 frames */
    LVSYNTH_SERVE = 5             /* resolution drawn per stream from {QVGA .3, VGA .3, 720p .2, 1080p .2} */
};
int lvsynth_preset(int preset, int stream, lvsynth_config* out);
int lvsynth_build_scene(const lvsynth_config* cfg, lvsynth_scene* out);
/* host generator: writes frame t (HWC, tightly packed) */
int lvsynth_frame_host(const lvsynth_scene* sc, int t, uint8_t* out);
/* GPU generator: writes frames t0..t0+count-1 to d_out + i*frame_stride (device memory),
 * asynchronously on the given hipStream_t (NULL = default stream). */
int lvsynth_frames_device(const lvsynth_scene* sc, int t0, int count, void* d_out, size_t frame_stride, void* stream);

#ifdef __cplusplus
}
#endif
#endif
