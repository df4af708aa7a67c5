// bgs_common.hpp -- device-side building blocks shared by the SuBSENSE / LOBSTER kernels
// (gfx950): counter-based RNG, LBSP pattern, sampling tables, per-stream state layout.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LV_HD __host__ __device__ __forceinline__
#define LV_D __device__ __forceinline__

namespace lvbgs {

// ------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11). Replaces every libc rand() of the reference; one call
// yields draws 4g..4g+3 of a (pixel, frame, phase, event) counter. A draw is the output word
// shifted right by one: a 31-bit value like glibc rand(), so `rand()%n` keeps its meaning.
//   key = {seed_lo, seed_hi}; ctr = {pixel, frame, phase<<24 | (event&0xFFFF)<<8 | group, stream}
// ------------------------------------------------------------------------------------------
enum : uint32_t { PH_APPLY = 1, PH_REFRESH = 2 };

LV_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                         uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

LV_HD void draw4(uint32_t seed_lo, uint32_t seed_hi, uint32_t stream, uint32_t frame, uint32_t phase,
                 uint32_t event, uint32_t pixel, uint32_t group, uint32_t out[4]) {
    philox4x32_10(pixel, frame, (phase << 24) | ((event & 0xFFFFu) << 8) | group, stream, seed_lo, seed_hi, out);
    out[0] >>= 1;
    out[1] >>= 1;
    out[2] >>= 1;
    out[3] >>= 1;
}

// `draw % m` for the reference's size_t moduli: m can be SIZE_MAX (infinite learning rate),
// and a 31-bit draw is then smaller than m, so the result is the draw itself.
LV_HD uint32_t mod64(uint32_t d, uint64_t m) {
    if (m > 0xFFFFFFFFull) return d;
    return d % (uint32_t)m;
}

// ------------------------------------------------------------------------------------------
// LBSP 16-bit double cross (LBSP.hpp:275-298): bit n compares the centre with (x+DX[n], y+DY[n]).
// ------------------------------------------------------------------------------------------
constexpr int kLbspDX[16] = {-2, 2, 0, 0, -2, 2, 2, -2, 0, -1, 0, 1, -1, 1, 1, -1};
constexpr int kLbspDY[16] = {0, 0, -2, 2, 2, -2, 2, -2, 1, 0, -1, 0, -1, 1, -1, 1};

// bit n = (|v[n]-ref| > t), LBSP.hpp:193-224. Two forms with the same bits; taps are passed packed
// (word w holds taps 4w..4w+3, byte k = tap 4w+k) to both.
//
// lbsp_threshold_sad, for callers whose taps come byte by byte (the frame tile in LDS): one v_sad_u8
// per tap. v_sad_u8(v, ref, ~t) = |v-ref| + 2^32-1-t wraps past 2^32 (bit 31 clears) exactly when
// |v-ref| > t (both at most 255); v_alignbit shifts that bit 31 into the result, tap 15 first, so
// bit n ends up holding tap n's "|v-ref| <= t"; the result is its complement. Two VALU per tap, one
// VGPR per tap once the compiler keeps the bytes apart.
LV_D uint32_t lbsp_threshold_sad(const uint32_t taps[4], uint32_t ref, uint32_t t) {
    const uint32_t acc = ~t;
    uint32_t d = 0;
#pragma unroll
    for (int n = 15; n >= 0; --n) {
        const uint32_t v = (taps[n >> 2] >> (8 * (n & 3))) & 0xFFu;
        const uint32_t x = __builtin_amdgcn_sad_u8(v, ref, acc);
        d = __builtin_amdgcn_alignbit(d, x, 31u);  // (d << 1) | (x >> 31)
    }
    return ~d & 0xFFFFu;
}

// lbsp_threshold_packed, for the sample match, whose taps stay in registers across the whole scan:
// they stay packed four per VGPR (a quarter of the registers). One v_mqsad_u32_u8 per word: window
// k of the quad SAD starts at byte k of S0, so it sees tap 4w+k in its byte 0; S1 holds the
// reference in byte 0 and zeros elsewhere, and the masked SAD skips the bytes where S1 is zero, so
// result k = |tap - ref| + ~t, which wraps past 2^32 (bit 31 clears) exactly when |tap - ref| > t
// (both at most 255). v_alignbit shifts those bits in, tap 15 first, so bit n ends up holding tap
// n's "|v-ref| <= t"; the result is its complement. The mask would also skip byte 0 when ref == 0:
// that case compares the complemented taps with 255 instead (255 - v against 255 - 0 has the same
// absolute difference).
LV_D uint32_t lbsp_threshold_packed(const uint32_t taps[4], uint32_t ref, uint32_t t) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const uint32_t acc = ~t;
    const u32x4 a = {acc, acc, acc, acc};
    const uint32_t flip = ref == 0u ? 0xFFFFFFFFu : 0u;
    const uint32_t r = ref ^ (flip & 0xFFu);
    uint32_t d = 0;
#pragma unroll
    for (int w = 3; w >= 0; --w) {
        const u32x4 q = __builtin_amdgcn_mqsad_u32_u8((uint64_t)(taps[w] ^ flip), r, a);
        d = __builtin_amdgcn_alignbit(d, q[3], 31u);
        d = __builtin_amdgcn_alignbit(d, q[2], 31u);
        d = __builtin_amdgcn_alignbit(d, q[1], 31u);
        d = __builtin_amdgcn_alignbit(d, q[0], 31u);
    }
    return ~d & 0xFFFFu;
}

// ------------------------------------------------------------------------------------------
// Sampling tables (opencv.hpp:859-966). Same tables and order as the reference.
// ------------------------------------------------------------------------------------------
constexpr int kSample7x7[49] = {2, 4, 6,  7,  6,  4,  2,  4, 8,  12, 14, 12, 8, 4, 6, 12, 21,
                                                   25, 21, 12, 6, 7, 14, 25, 28, 25, 14, 7, 6, 12, 21, 25, 21, 12,
                                                   6, 4, 8,  12, 14, 12, 8, 4, 2,  4,  6,  7,  6,  4,  2};
constexpr int kNb3[8][2] = {{-1, 1}, {0, 1}, {1, 1}, {-1, 0}, {1, 0}, {-1, -1}, {0, -1}, {1, -1}};
constexpr int kNb5[24][2] = {{-2, 2}, {-1, 2}, {0, 2},  {1, 2},  {2, 2},   {-2, 1},
                                                {-1, 1}, {0, 1},  {1, 1},  {2, 1},  {-2, 0},  {-1, 0},
                                                {1, 0},  {2, 0},  {-2, -1}, {-1, -1}, {0, -1}, {1, -1},
                                                {2, -1}, {-2, -2}, {-1, -2}, {0, -2}, {1, -2}, {2, -2}};

LV_HD void clamp_coords(int& x, int& y, int border, int w, int h) {
    if (x < border) x = border;
    else if (x >= w - border) x = w - border - 1;
    if (y < border) y = border;
    else if (y >= h - border) y = h - border - 1;
}

// getSamplePosition_7x7_std2 (opencv.hpp:873-925), border 2
LV_HD void sample_pos_7x7(uint32_t rnd, int x, int y, int w, int h, int& ox, int& oy) {
    int r = 1 + (int)(rnd % 512u);
    int i = 0;
    for (; i < 48; ++i) {
        r -= kSample7x7[i];
        if (r <= 0) break;
    }
    ox = x + (i % 7) - 3;
    oy = y + (i / 7) - 3;
    clamp_coords(ox, oy, 2, w, h);
}

// getNeighborPosition_3x3 / _5x5 (opencv.hpp:928-966), border 2
LV_HD void neighbor_pos(uint32_t rnd, bool use3, int x, int y, int w, int h, int& ox, int& oy) {
    if (use3) {
        const int r = (int)(rnd % 8u);
        ox = x + kNb3[r][0];
        oy = y + kNb3[r][1];
    } else {
        const int r = (int)(rnd % 24u);
        ox = x + kNb5[r][0];
        oy = y + kNb5[r][1];
    }
    clamp_coords(ox, oy, 2, w, h);
}

// ------------------------------------------------------------------------------------------
// Per-stream device scalars (the reference's frame-level members, kept on the GPU so a frame
// step needs no host round trip).
// ------------------------------------------------------------------------------------------
constexpr int kWlRing = 4;
constexpr int kStripes = 64;
// The match continuation list is split into kContStripes stripes: tile b appends to stripe
// b % kContStripes, each stripe's counter on its own 128-byte line, so appends from different
// tiles do not serialize on one address.
constexpr int kContStripes = 64;
constexpr int kContStride = 32;  // u32 words between stripe counters
struct alignas(128) Scalars {
    uint32_t frame_idx;          // frames applied since initialize (SuBSENSE m_nFrameIdx)
    uint32_t frames_since_reset; // m_nFramesSinceLastReset
    uint32_t cooldown;           // m_nModelResetCooldown
    int32_t auto_reset;          // m_bAutoModelResetEnabled
    float lower_cap, upper_cap;  // m_fCurrLearningRate{Lower,Upper}Cap
    float last_nz_ratio;         // m_fLastNonZeroDescRatio
    uint32_t refresh_seq;        // refresh events so far (RNG event id)
    uint32_t nz_count_unused;
    uint32_t err_flags;          // sticky error bits (see kErr*)
    unsigned long long cd_sum_unused;
    uint32_t wl_count[kWlRing];  // fix-point worklist counts (ring)
    int32_t refresh_pending;     // refresh request consumed by the refresh kernel
    float refresh_frac;
    int32_t refresh_force_fg;
    uint32_t refresh_frame;
    uint32_t refresh_ev;
    uint32_t refresh_start;
    // per-frame sums, striped over kStripes counters (one atomic per workgroup into stripe
    // blockIdx % kStripes) and summed by the consumer
    unsigned long long nz_part[64];   // pixels with >= 4 (1ch: 2) descriptor bits (LUT adaptation)
    unsigned long long cd_part[64];   // frame-level colour difference sum (model reset, caps)
    unsigned long long ex_part[64];   // samples examined by the final match (k_p)
    unsigned long long wr_part[64];   // sample writes (w_p)
    unsigned long long spec_part[64]; // samples examined by the speculative match kernel
    unsigned long long coop_part[64]; // samples examined by the cooperative match phase
    uint32_t iters;              // per-frame: fix-point iterations that had work
    uint32_t max_iter_seen;      // largest fix-point iteration index that had work (sticky)
    uint32_t epoch;              // steps since allocation (worklist de-duplication tags)
    uint32_t wl_work[8];         // per-frame: worklist size of fix-point iterations 1..7
    uint32_t cont2_count;        // per-frame: pixels handed to the cooperative match phase
    uint8_t lut[256];            // LBSP threshold LUT (m_anLBSPThreshold_8bitLUT)
    // per-frame: pixels handed to the match continuation, stripe k at cont_cnt[k * kContStride]
    alignas(128) uint32_t cont_cnt[kContStripes * kContStride];
};
constexpr uint32_t kErrFixpointCap = 1u;

// record layout (one u32 per pixel, written by the describe kernel)
constexpr uint32_t kRecNone = 31u;  // tgt field value for non-ROI pixels
LV_HD uint32_t rec_tgt(uint32_t r) { return r & 31u; }
LV_HD bool rec_gate_a(uint32_t r) { return (r >> 5) & 1u; }
LV_HD bool rec_gate_b(uint32_t r) { return (r >> 6) & 1u; }
LV_HD bool rec_ghost_fwd(uint32_t r) { return (r >> 7) & 1u; }
LV_HD bool rec_self_bg(uint32_t r) { return (r >> 8) & 1u; }
LV_HD bool rec_self_fg(uint32_t r) { return (r >> 9) & 1u; }
LV_HD bool rec_unstable(uint32_t r) { return (r >> 10) & 1u; }
LV_HD uint32_t rec_slot_self(uint32_t r) { return (r >> 16) & 255u; }
LV_HD uint32_t rec_slot_diff(uint32_t r) { return (r >> 24) & 255u; }
// target offset index (dy+2)*5+(dx+2) of a 5x5 window; 12 = self
LV_HD int tgt_dx(uint32_t t) { return (int)(t % 5u) - 2; }
LV_HD int tgt_dy(uint32_t t) { return (int)(t / 5u) - 2; }

// match result (one u32 per pixel): good | min_desc<<8 | min_sum<<14 | examined<<24
LV_HD uint32_t mres_pack(uint32_t good, uint32_t md, uint32_t ms, uint32_t k) {
    return good | (md << 8) | (ms << 14) | (k << 24);
}

}  // namespace lvbgs
