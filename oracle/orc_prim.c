/*
 * orc_prim.c -- oracle primitives: LBSP, threshold LUT, sampling tables, distances,
 * and the closed forms of the OpenCV whole-image operations on the path.
 *
 * TEST INFRASTRUCTURE (oracle). Each function cites the reference it restates.
 */
#include "orc.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* 16-bit double-cross LBSP pattern: bit n compares the centre against the pixel at
 * (x+dx[n], y+dy[n]). Offsets are LBSP.hpp:297-298 (s_oIdxLUT_16bitdbcross_x/_y);
 * the table has to be identical to keep descriptor bits identical. */
static const int LBSP_DX[16] = {-2, 2, 0, 0, -2, 2, 2, -2, 0, -1, 0, 1, -1, 1, 1, -1};
static const int LBSP_DY[16] = { 0, 0, -2, 2, 2, -2, 2, -2, 1, 0, -1, 0, -1, 1, -1, 1};

/* LBSP.hpp:158-168 + :300-319 (lookup of the 16 taps of channel ch at (x,y), HWC image) */
void orc_lbsp_lookup(const uint8_t* img, int w, int c, int x, int y, int ch, uint8_t vals[16]) {
    for (int n = 0; n < 16; ++n)
        vals[n] = img[((size_t)(y + LBSP_DY[n]) * (size_t)w + (size_t)(x + LBSP_DX[n])) * (size_t)c + (size_t)ch];
}

/* LBSP.hpp:193-202: bit n = (|vals[n]-ref| > t) (the SSE paths compute the same predicate) */
uint16_t orc_lbsp_threshold(const uint8_t vals[16], uint8_t ref, uint8_t t) {
    uint16_t d = 0;
    for (int n = 0; n < 16; ++n) {
        const int diff = (int)vals[n] - (int)ref;
        const unsigned ad = (unsigned)(diff < 0 ? -diff : diff);
        d |= (uint16_t)((ad > (unsigned)t) ? 1u : 0u) << n;
    }
    return d;
}

/* LBSP.cpp:102-123 (lbsp_computeImpl, absolute threshold, no reference image): every pixel with
 * a full 5x5 patch is described, each channel against its own centre; border pixels are left 0
 * (the reference leaves them uninitialized; its own test only reads the valid zone). */
void orc_lbsp_compute_abs(const uint8_t* img, int w, int h, int c, int threshold, uint16_t* out) {
    const uint8_t t = (uint8_t)(threshold < 0 ? 0 : threshold > 255 ? 255 : threshold);
    memset(out, 0, sizeof(uint16_t) * (size_t)w * (size_t)h * (size_t)c);
    uint8_t vals[16];
    for (int y = 2; y < h - 2; ++y)
        for (int x = 2; x < w - 2; ++x)
            for (int ch = 0; ch < c; ++ch) {
                orc_lbsp_lookup(img, w, c, x, y, ch, vals);
                out[((size_t)y * w + x) * c + ch] = orc_lbsp_threshold(vals, img[((size_t)y * w + x) * c + ch], t);
            }
}

/* cv::saturate_cast<uchar>(float): cvRound (round half to even under the default FP
 * environment) then clamp to [0,255]. */
uint8_t orc_sat_u8_f(float v) {
    const long r = lrintf(v);
    return (uint8_t)(r < 0 ? 0 : r > 255 ? 255 : r);
}

/* BackgroundSubtractorLBSP.cpp:29-30 (1ch: (t*rel+offset)/3) and :42-43 (3ch: t*rel+offset),
 * all in float, then saturate_cast<uchar>. */
void orc_build_lut(int channels, float rel, size_t offset, uint8_t lut[256]) {
    for (size_t t = 0; t <= 255; ++t) {
        const float base = (float)t * rel + (float)offset;
        lut[t] = orc_sat_u8_f(channels == 1 ? base / 3.0f : base);
    }
}

/* opencv.hpp:859-870 clampImageCoords */
static void clamp_coords(int* x, int* y, int border, int w, int h) {
    if (*x < border) *x = border;
    else if (*x >= w - border) *x = w - border - 1;
    if (*y < border) *y = border;
    else if (*y >= h - border) *y = h - border - 1;
}

/* opencv.hpp:909-925: 'floor(fspecial('gaussian',7,2)*512)' cumulative pattern; table must be identical. */
static const int SAMPLE_7X7[7][7] = {
    {2, 4, 6, 7, 6, 4, 2}, {4, 8, 12, 14, 12, 8, 4}, {6, 12, 21, 25, 21, 12, 6}, {7, 14, 25, 28, 25, 14, 7},
    {6, 12, 21, 25, 21, 12, 6}, {4, 8, 12, 14, 12, 8, 4}, {2, 4, 6, 7, 6, 4, 2}};

/* opencv.hpp:873-891 getSamplePosition<7,7> with total 512 */
void orc_sample_pos_7x7(int rnd, int x, int y, int border, int w, int h, int* ox, int* oy) {
    int r = 1 + (rnd % 512);
    int sx = 0, sy = 0;
    for (sy = 0; sy < 7; ++sy) {
        for (sx = 0; sx < 7; ++sx) {
            r -= SAMPLE_7X7[sy][sx];
            if (r <= 0) goto found;
        }
    }
found:
    sx += x - 3;
    sy += y - 3;
    clamp_coords(&sx, &sy, border, w, h);
    *ox = sx; *oy = sy;
}

/* opencv.hpp:941-950 neighbour table (x,y) pairs, order kept */
static const int NB3[8][2] = {{-1, 1}, {0, 1}, {1, 1}, {-1, 0}, {1, 0}, {-1, -1}, {0, -1}, {1, -1}};
/* opencv.hpp:953-966 */
static const int NB5[24][2] = {{-2, 2}, {-1, 2}, {0, 2}, {1, 2}, {2, 2},   {-2, 1}, {-1, 1}, {0, 1},
                               {1, 1},  {2, 1},  {-2, 0}, {-1, 0}, {1, 0}, {2, 0},  {-2, -1}, {-1, -1},
                               {0, -1}, {1, -1}, {2, -1}, {-2, -2}, {-1, -2}, {0, -2}, {1, -2}, {2, -2}};

void orc_neighbor_pos_3x3(int rnd, int x, int y, int border, int w, int h, int* ox, int* oy) {
    const int r = rnd % 8;
    int nx = x + NB3[r][0], ny = y + NB3[r][1];
    clamp_coords(&nx, &ny, border, w, h);
    *ox = nx; *oy = ny;
}

void orc_neighbor_pos_5x5(int rnd, int x, int y, int border, int w, int h, int* ox, int* oy) {
    const int r = rnd % 24;
    int nx = x + NB5[r][0], ny = y + NB5[r][1];
    clamp_coords(&nx, &ny, border, w, h);
    *ox = nx; *oy = ny;
}

/* math.hpp:228-234 L1dist<3> on uchar: accumulated in the (unsigned char) return type, so the
 * sum wraps modulo 256 (e.g. 200+200+200 -> 88). */
uint8_t orc_l1dist3_u8(const uint8_t a[3], const uint8_t b[3]) {
    uint8_t acc = 0;
    for (int c = 0; c < 3; ++c) {
        const int d = (int)a[c] - (int)b[c];
        acc = (uint8_t)(acc + (uint8_t)(d < 0 ? -d : d));
    }
    return acc;
}

/* cv::dilate / cv::erode with a (2r+1)^2 rectangle and the default border value
 * (morphologyDefaultBorderValue: neutral), i.e. max/min over the window clipped to the image.
 * Iterated 3x3 operations with neutral borders compose into the larger clipped window.
 * Computed separably (rows, then columns), which is exact for a rectangle. */
void orc_morph(const uint8_t* in, uint8_t* out, int w, int h, int radius, int is_max) {
    uint8_t* tmp = (uint8_t*)malloc((size_t)w * h);
    for (int y = 0; y < h; ++y) {
        const uint8_t* row = in + (size_t)y * w;
        for (int x = 0; x < w; ++x) {
            const int x0 = x - radius < 0 ? 0 : x - radius, x1 = x + radius >= w ? w - 1 : x + radius;
            int acc = row[x0];
            for (int xx = x0 + 1; xx <= x1; ++xx) acc = is_max ? (row[xx] > acc ? row[xx] : acc) : (row[xx] < acc ? row[xx] : acc);
            tmp[(size_t)y * w + x] = (uint8_t)acc;
        }
    }
    for (int y = 0; y < h; ++y) {
        const int y0 = y - radius < 0 ? 0 : y - radius, y1 = y + radius >= h ? h - 1 : y + radius;
        for (int x = 0; x < w; ++x) {
            int acc = tmp[(size_t)y0 * w + x];
            for (int yy = y0 + 1; yy <= y1; ++yy) {
                const int v = tmp[(size_t)yy * w + x];
                acc = is_max ? (v > acc ? v : acc) : (v < acc ? v : acc);
            }
            out[(size_t)y * w + x] = (uint8_t)acc;
        }
    }
    free(tmp);
}

/* cv::medianBlur(k) on a 0/255 mask whose border ring of width >= 1 is 0, so that
 * BORDER_REPLICATE equals zero padding: out = 255 iff #255 in the k x k window >= (k*k+1)/2.
 * (Invariant pinned by modules/imgproc/test/global.cpp:67-90; brute-force sorting median in tests.)
 * Counted with an integral image. */
void orc_median_binary(const uint8_t* in, uint8_t* out, int w, int h, int k) {
    const int r = k / 2, need = (k * k + 1) / 2;
    const size_t iw = (size_t)w + 1;
    int* ii = (int*)calloc(iw * ((size_t)h + 1), sizeof(int));
    for (int y = 0; y < h; ++y) {
        int rs = 0;
        for (int x = 0; x < w; ++x) {
            rs += in[(size_t)y * w + x] ? 1 : 0;
            ii[(size_t)(y + 1) * iw + (x + 1)] = ii[(size_t)y * iw + (x + 1)] + rs;
        }
    }
    for (int y = 0; y < h; ++y) {
        const int y0 = y - r < 0 ? 0 : y - r, y1 = y + r >= h ? h - 1 : y + r;
        for (int x = 0; x < w; ++x) {
            const int x0 = x - r < 0 ? 0 : x - r, x1 = x + r >= w ? w - 1 : x + r;
            const int cnt = ii[(size_t)(y1 + 1) * iw + (x1 + 1)] - ii[(size_t)y0 * iw + (x1 + 1)] -
                            ii[(size_t)(y1 + 1) * iw + x0] + ii[(size_t)y0 * iw + x0];
            out[(size_t)y * w + x] = cnt >= need ? 255 : 0;
        }
    }
    free(ii);
}

/* cv::floodFill(m,(0,0),255) with default 4-connectivity and zero tolerance, then bitwise_not:
 * holes = 0-pixels not 4-connected to pixel (0,0) through 0-pixels (BackgroundSubtractorSuBSENSE.cpp:541-543).
 * Explicit-stack flood (no recursion). */
void orc_flood_holes(const uint8_t* in, uint8_t* holes, int w, int h) {
    const size_t n = (size_t)w * h;
    uint8_t* reached = (uint8_t*)calloc(n, 1);
    size_t* stack = (size_t*)malloc(sizeof(size_t) * (n + 1));
    size_t sp = 0;
    if (in[0] == 0) { reached[0] = 1; stack[sp++] = 0; }
    while (sp) {
        const size_t i = stack[--sp];
        const int x = (int)(i % (size_t)w), y = (int)(i / (size_t)w);
        const int nx[4] = {x - 1, x + 1, x, x}, ny[4] = {y, y, y - 1, y + 1};
        for (int d = 0; d < 4; ++d) {
            if (nx[d] < 0 || nx[d] >= w || ny[d] < 0 || ny[d] >= h) continue;
            const size_t j = (size_t)ny[d] * w + nx[d];
            if (!reached[j] && in[j] == 0) { reached[j] = 1; stack[sp++] = j; }
        }
    }
    for (size_t i = 0; i < n; ++i) holes[i] = (in[i] == 0 && !reached[i]) ? 255 : 0;
    free(stack);
    free(reached);
}

/* cv::resize(in, out, Size(dw,dh), 0, 0, INTER_AREA) as used at BackgroundSubtractorSuBSENSE.cpp:568.
 * Integer scale (OpenCV's resizeAreaFast path): out = saturate_cast<uchar>((float)sum * (1.f/area)).
 * Non-integer scale: defined here as the exact area-weighted mean computed in double in a fixed
 * order, then rounded like saturate_cast (parity vs OpenCV's generic INTER_AREA is unpinned). */
void orc_resize_area(const uint8_t* in, int w, int h, int c, uint8_t* out, int dw, int dh) {
    if (dw <= 0 || dh <= 0) return;
    if (w % dw == 0 && h % dh == 0) {
        const int sx = w / dw, sy = h / dh;
        const float scale = 1.f / (float)(sx * sy);
        for (int dy = 0; dy < dh; ++dy)
            for (int dx = 0; dx < dw; ++dx)
                for (int ch = 0; ch < c; ++ch) {
                    int sum = 0;
                    for (int yy = 0; yy < sy; ++yy)
                        for (int xx = 0; xx < sx; ++xx)
                            sum += in[((size_t)(dy * sy + yy) * w + (size_t)(dx * sx + xx)) * c + ch];
                    out[((size_t)dy * dw + dx) * c + ch] = orc_sat_u8_f((float)sum * scale);
                }
        return;
    }
    const double sx = (double)w / dw, sy = (double)h / dh;
    for (int dy = 0; dy < dh; ++dy) {
        const double y0 = dy * sy, y1 = (dy + 1) * sy;
        for (int dx = 0; dx < dw; ++dx) {
            const double x0 = dx * sx, x1 = (dx + 1) * sx;
            for (int ch = 0; ch < c; ++ch) {
                double acc = 0.0;
                for (int yy = (int)y0; yy < h && (double)yy < y1; ++yy) {
                    const double wy = fmin((double)(yy + 1), y1) - fmax((double)yy, y0);
                    for (int xx = (int)x0; xx < w && (double)xx < x1; ++xx) {
                        const double wx = fmin((double)(xx + 1), x1) - fmax((double)xx, x0);
                        acc += (double)in[((size_t)yy * w + xx) * c + ch] * (wy * wx);
                    }
                }
                out[((size_t)dy * dw + dx) * c + ch] = orc_sat_u8_f((float)(acc / (sx * sy)));
            }
        }
    }
}
