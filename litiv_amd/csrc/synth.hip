// synth.hip -- deterministic synthetic video generator (host and gfx950 device, bit-identical).
//
// Scene: a seeded 4-octave value-noise background, moving objects (rectangles or ellipses) that
// wrap around the frame, additive noise with a 4-term Irwin-Hall approximation of a Gaussian,
// a global illumination ramp and a flickering dynamic-background patch. Everything is integer
// arithmetic on 32-bit hashes, so the CPU and GPU produce the same bytes.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "../../include/litiv_synth.h"

#define LV_HD __host__ __device__ __forceinline__

namespace {

// round(1024*sin(2*pi*i/256)); one copy per compilation side (device constant / host const)
#define LVSYNTH_SIN1024 {0, 25, 50, 75, 100, 125, 150, 175, 200, 224, 249, 273, 297, 321, 345, 369, 392, 415, 438, 460, 483, 505, 526, 548, 569, 590, 610, 630, 650, 669, 688, 706, 724, 742, 759, 775, 792, 807, 822, 837, 851, 865, 878, 891, 903, 915, 926, 936, 946, 955, 964, 972, 980, 987, 993, 999, 1004, 1009, 1013, 1016, 1019, 1021, 1023, 1024, 1024, 1024, 1023, 1021, 1019, 1016, 1013, 1009, 1004, 999, 993, 987, 980, 972, 964, 955, 946, 936, 926, 915, 903, 891, 878, 865, 851, 837, 822, 807, 792, 775, 759, 742, 724, 706, 688, 669, 650, 630, 610, 590, 569, 548, 526, 505, 483, 460, 438, 415, 392, 369, 345, 321, 297, 273, 249, 224, 200, 175, 150, 125, 100, 75, 50, 25, 0, -25, -50, -75, -100, -125, -150, -175, -200, -224, -249, -273, -297, -321, -345, -369, -392, -415, -438, -460, -483, -505, -526, -548, -569, -590, -610, -630, -650, -669, -688, -706, -724, -742, -759, -775, -792, -807, -822, -837, -851, -865, -878, -891, -903, -915, -926, -936, -946, -955, -964, -972, -980, -987, -993, -999, -1004, -1009, -1013, -1016, -1019, -1021, -1023, -1024, -1024, -1024, -1023, -1021, -1019, -1016, -1013, -1009, -1004, -999, -993, -987, -980, -972, -964, -955, -946, -936, -926, -915, -903, -891, -878, -865, -851, -837, -822, -807, -792, -775, -759, -742, -724, -706, -688, -669, -650, -630, -610, -590, -569, -548, -526, -505, -483, -460, -438, -415, -392, -369, -345, -321, -297, -273, -249, -224, -200, -175, -150, -125, -100, -75, -50, -25}
#if defined(__HIP_DEVICE_COMPILE__)
__constant__ short kSin1024[256] = LVSYNTH_SIN1024;
#else
const short kSin1024[256] = LVSYNTH_SIN1024;
#endif

LV_HD uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
LV_HD uint32_t hash4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return mix32(a ^ mix32(b ^ mix32(c ^ mix32(d + 0x9E3779B9u))));
}
LV_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
LV_HD int pmod(int a, int m) {
    const int r = a % m;
    return r < 0 ? r + m : r;
}

LV_HD int value_noise(const lvsynth_config& c, int x, int y, int ch) {
    int acc = 0, wsum = 0;
    for (int o = 0; o < c.octaves; ++o) {
        const int s = (c.bg_cell >> o) > 0 ? (c.bg_cell >> o) : 1;
        const int wgt = 1 << (c.octaves - 1 - o);
        const int ix = x / s, iy = y / s;
        const int fx = ((x % s) << 8) / s, fy = ((y % s) << 8) / s;
        const uint32_t base = c.seed * 0x10001u + (uint32_t)o * 0x51u + (uint32_t)ch * 0x3777u;
        const int a = (int)(hash4(base, (uint32_t)ix, (uint32_t)iy, 0xA1u) & 255u);
        const int b = (int)(hash4(base, (uint32_t)(ix + 1), (uint32_t)iy, 0xA1u) & 255u);
        const int cc = (int)(hash4(base, (uint32_t)ix, (uint32_t)(iy + 1), 0xA1u) & 255u);
        const int d = (int)(hash4(base, (uint32_t)(ix + 1), (uint32_t)(iy + 1), 0xA1u) & 255u);
        const int top = a * (256 - fx) + b * fx;
        const int bot = cc * (256 - fx) + d * fx;
        const int v = (int)(((long long)top * (256 - fy) + (long long)bot * fy) >> 16);
        acc += wgt * v;
        wsum += wgt;
    }
    const int n = wsum ? acc / wsum : 0;
    return c.bg_lo + n * (c.bg_hi - c.bg_lo) / 255;
}

LV_HD int object_at(const lvsynth_scene& sc, int x, int y, int t) {
    const lvsynth_config& c = sc.cfg;
    int hit = -1;
    for (int o = 0; o < sc.n_obj; ++o) {
        const lvsynth_object& ob = sc.obj[o];
        const int ox = pmod(ob.x0 + ob.vx * t, c.width), oy = pmod(ob.y0 + ob.vy * t, c.height);
        const int dx = pmod(x - ox, c.width), dy = pmod(y - oy, c.height);
        if (dx >= ob.w || dy >= ob.h) continue;
        if (c.shape == 1) {
            const long long ex = 2LL * dx - (ob.w - 1), ey = 2LL * dy - (ob.h - 1);
            const long long W2 = (long long)ob.w * ob.w, H2 = (long long)ob.h * ob.h;
            if (ex * ex * H2 + ey * ey * W2 > W2 * H2) continue;
        }
        hit = o;
    }
    return hit;
}

LV_HD int drift_at(const lvsynth_config& c, int t) {
    if (c.frames <= 1 || c.drift == 0) return 0;
    const long long num = 2LL * c.drift * 2 * t + (c.frames - 1);
    return (int)(num / (2LL * (c.frames - 1))) - c.drift;
}

LV_HD uint8_t pixel_value(const lvsynth_scene& sc, int x, int y, int t, int ch, int obj) {
    const lvsynth_config& c = sc.cfg;
    int v;
    if (obj >= 0) {
        v = sc.obj[obj].color[ch];
    } else {
        v = value_noise(c, x, y, ch);
        if (c.flicker_size > 0 && x >= sc.fx0 && x < sc.fx0 + c.flicker_size && y >= sc.fy0 &&
            y < sc.fy0 + c.flicker_size) {
            const uint32_t hp = hash4(c.seed, (uint32_t)x, (uint32_t)y, 0xF11Cu);
            const int period = c.flicker_pmin + (int)(hp % (uint32_t)(c.flicker_pmax - c.flicker_pmin + 1));
            const int phase = (int)((hp >> 12) & 255u);
            const int idx = ((t * 256) / period + phase) & 255;
            v += (c.flicker_amp * kSin1024[idx] + 512) >> 10;
        }
    }
    v += drift_at(c, t);
    if (c.noise_sigma > 0) {
        const uint32_t h = hash4(c.seed ^ 0x5EEDu, (uint32_t)(x + y * c.width), (uint32_t)t, (uint32_t)ch);
        const int s = (int)(h & 255u) + (int)((h >> 8) & 255u) + (int)((h >> 16) & 255u) + (int)(h >> 24) - 510;
        v += (s * c.noise_sigma * 443 + 32768) >> 16;
    }
    return (uint8_t)clampi(v, 0, 255);
}

__global__ void __launch_bounds__(256) synth_kernel(lvsynth_scene sc, int t0, uint8_t* out, size_t stride) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int t = t0 + (int)blockIdx.z;
    const lvsynth_config& c = sc.cfg;
    if (x >= c.width || y >= c.height) return;
    const int obj = object_at(sc, x, y, t);
    uint8_t* dst = out + (size_t)blockIdx.z * stride + ((size_t)y * c.width + x) * c.channels;
    for (int ch = 0; ch < c.channels; ++ch) dst[ch] = pixel_value(sc, x, y, t, ch, obj);
}

uint32_t draw_u(uint32_t seed, uint32_t a, uint32_t b) { return hash4(seed, 0x0B1EC7u, a, b); }

}  // namespace

extern "C" int lvsynth_preset(int preset, int stream, lvsynth_config* out) {
    if (!out) return -1;
    lvsynth_config c;
    std::memset(&c, 0, sizeof(c));
    c.channels = 3;
    c.octaves = 4;
    c.bg_lo = 40;
    c.bg_hi = 215;
    c.speed_min = 1;
    c.speed_max = 4;
    c.drift = 10;
    c.flicker_size = 40;
    c.flicker_amp = 15;
    c.flicker_pmin = 5;
    c.flicker_pmax = 20;
    c.obj_lo = 0;
    c.obj_hi = 255;
    c.shape = 0;
    int cfg_id = preset;
    switch (preset) {
        case LVSYNTH_SUBSENSE_QVGA:
            c.width = 320, c.height = 240, c.frames = 300, c.n_objects = 3, c.obj_min = 20, c.obj_max = 60;
            c.noise_sigma = 3, c.bg_cell = 64;
            break;
        case LVSYNTH_SUBSENSE_1080P:
            c.width = 1920, c.height = 1080, c.frames = 1000, c.n_objects = 12, c.obj_min = 40, c.obj_max = 200;
            c.noise_sigma = 4, c.bg_cell = 256;
            break;
        case LVSYNTH_SUBSENSE_4K:
            c.width = 3840, c.height = 2160, c.frames = 500, c.n_objects = 20, c.obj_min = 40, c.obj_max = 200;
            c.noise_sigma = 4, c.bg_cell = 512;
            break;
        case LVSYNTH_LOBSTER_THERMAL:
            c.width = 640, c.height = 512, c.channels = 1, c.frames = 1000, c.n_objects = 5, c.obj_min = 30;
            c.obj_max = 80, c.speed_min = 1, c.speed_max = 3, c.noise_sigma = 2, c.bg_lo = 60, c.bg_hi = 120;
            c.bg_cell = 256, c.octaves = 2, c.drift = 0, c.flicker_size = 0, c.shape = 1, c.obj_lo = 170;
            c.obj_hi = 230;
            break;
        case LVSYNTH_BATCH64:
        case LVSYNTH_SERVE: {
            static const int RW[4] = {320, 640, 640, 1920}, RH[4] = {240, 480, 512, 1080}, RC[4] = {3, 3, 1, 3};
            static const int SW[4] = {320, 640, 1280, 1920}, SH[4] = {240, 480, 720, 1080};
            int g;
            if (preset == LVSYNTH_BATCH64) {
                if (stream < 0 || stream >= 64) return -1;
                g = stream / 16;
                c.width = RW[g], c.height = RH[g], c.channels = RC[g], c.frames = 600;
            } else {
                const uint32_t u = hash4(0x5E12Eu, (uint32_t)stream, 7u, 11u) % 10u;
                g = u < 3 ? 0 : u < 6 ? 1 : u < 8 ? 2 : 3;
                c.width = SW[g], c.height = SH[g], c.channels = 3, c.frames = 2000;
            }
            const double A = (double)c.width * c.height;
            const double f = (A - 76800.0) / (2073600.0 - 76800.0);
            c.n_objects = (int)std::lround(3.0 + 9.0 * f);
            c.obj_min = (int)std::lround(20.0 + 20.0 * f);
            c.obj_max = (int)std::lround(60.0 + 140.0 * f);
            c.noise_sigma = (int)std::lround(3.0 + f);
            c.bg_cell = c.width >= 1280 ? 256 : c.width >= 640 ? 128 : 64;
            break;
        }
        default:
            return -1;
    }
    c.seed = 0x117170u + 1000u * (uint32_t)cfg_id + (uint32_t)stream;
    *out = c;
    return 0;
}

extern "C" int lvsynth_build_scene(const lvsynth_config* cfg, lvsynth_scene* out) {
    if (!cfg || !out || cfg->width <= 0 || cfg->height <= 0 || (cfg->channels != 1 && cfg->channels != 3) ||
        cfg->n_objects < 0 || cfg->n_objects > LVSYNTH_MAX_OBJECTS || cfg->obj_min <= 0 || cfg->obj_max < cfg->obj_min ||
        cfg->speed_max < cfg->speed_min || (cfg->flicker_size > 0 && cfg->flicker_pmax < cfg->flicker_pmin) ||
        (cfg->flicker_size > 0 && cfg->flicker_pmin <= 0) || cfg->bg_cell <= 0)
        return -1;
    std::memset(out, 0, sizeof(*out));
    out->cfg = *cfg;
    out->n_obj = cfg->n_objects;
    const uint32_t s = cfg->seed;
    for (int o = 0; o < cfg->n_objects; ++o) {
        lvsynth_object& ob = out->obj[o];
        const int span = cfg->obj_max - cfg->obj_min + 1, vs = cfg->speed_max - cfg->speed_min + 1;
        ob.w = cfg->obj_min + (int)(draw_u(s, (uint32_t)o, 0) % (uint32_t)span);
        ob.h = cfg->obj_min + (int)(draw_u(s, (uint32_t)o, 1) % (uint32_t)span);
        ob.x0 = (int)(draw_u(s, (uint32_t)o, 2) % (uint32_t)cfg->width);
        ob.y0 = (int)(draw_u(s, (uint32_t)o, 3) % (uint32_t)cfg->height);
        ob.vx = (cfg->speed_min + (int)(draw_u(s, (uint32_t)o, 4) % (uint32_t)vs)) * ((draw_u(s, (uint32_t)o, 5) & 1) ? 1 : -1);
        ob.vy = (cfg->speed_min + (int)(draw_u(s, (uint32_t)o, 6) % (uint32_t)vs)) * ((draw_u(s, (uint32_t)o, 7) & 1) ? 1 : -1);
        const int cspan = cfg->obj_hi - cfg->obj_lo + 1;
        for (int ch = 0; ch < 3; ++ch) ob.color[ch] = cfg->obj_lo + (int)(draw_u(s, (uint32_t)o, 8u + (uint32_t)ch) % (uint32_t)cspan);
    }
    if (cfg->flicker_size > 0) {
        const int fw = cfg->width - cfg->flicker_size, fh = cfg->height - cfg->flicker_size;
        out->fx0 = fw > 0 ? (int)(draw_u(s, 0xF1u, 0) % (uint32_t)fw) : 0;
        out->fy0 = fh > 0 ? (int)(draw_u(s, 0xF1u, 1) % (uint32_t)fh) : 0;
    }
    return 0;
}

extern "C" int lvsynth_frame_host(const lvsynth_scene* sc, int t, uint8_t* out) {
    if (!sc || !out) return -1;
    const lvsynth_config& c = sc->cfg;
    for (int y = 0; y < c.height; ++y)
        for (int x = 0; x < c.width; ++x) {
            const int obj = object_at(*sc, x, y, t);
            for (int ch = 0; ch < c.channels; ++ch)
                out[((size_t)y * c.width + x) * c.channels + ch] = pixel_value(*sc, x, y, t, ch, obj);
        }
    return 0;
}

extern "C" int lvsynth_frames_device(const lvsynth_scene* sc, int t0, int count, void* d_out, size_t frame_stride,
                                     void* stream) {
    if (!sc || !d_out || count <= 0 || count > 65535) return -1;
    const lvsynth_config& c = sc->cfg;
    if (frame_stride < (size_t)c.width * c.height * c.channels) return -1;
    dim3 grid((unsigned)((c.width + 63) / 64), (unsigned)((c.height + 3) / 4), (unsigned)count);
    hipLaunchKernelGGL(synth_kernel, grid, dim3(256), 0, (hipStream_t)stream, *sc, t0, (uint8_t*)d_out, frame_stride);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
