// bgs_kernels.hip -- gfx950 kernels of the SuBSENSE / LOBSTER frame step.
//
// The reference apply() is a raster-order loop (BackgroundSubtractorSuBSENSE.cpp:333-481,
// BackgroundSubtractorLOBSTER.cpp:459-576) in which a background pixel may write a sample of
// a neighbour; a neighbour later in raster order sees that sample in the same frame. The frame
// step below reproduces that order exactly without a sequential sweep:
//   describe : per pixel, everything that does not depend on other pixels' decisions
//              (LBSP, thresholds, RNG draws, diffusion target/slot/gates)       -> rec, desc_cur
//   match    : speculative ordered sample scan; samples diffused in by raster-earlier
//              neighbours are applied under a guess of their decision (previous raw mask)
//   fixpoint : pixels whose writers' decisions differ from the guess are re-matched, and the
//              change propagates through a de-duplicated worklist until nothing changes. The
//              writer graph is acyclic (raster order), so the fix point is unique and equals the
//              sequential result.
//   feedback : FP32 maps (D_last, D_min, raw segm, T, v, R), raw mask, blink bookkeeping
//   update   : each pixel gathers, in raster order, the writes aimed at its samples from the 5x5
//              window (self-update included); the last writer of a slot wins. No atomics.
//   post     : close 3x3, hole fill (union-find CCL of the background), erode 7x7, binary
//              median k x k, dilate 7x7, blink masks, final-segmentation averages
//   control  : frame-level downsampled analysis, LBSP LUT adaptation, model reset, caps
#include <hip/hip_runtime.h>

#include <cmath>

#include "bgs_kernels.hpp"

namespace lvbgs {

constexpr int TW = 64, TH = 4, TPB = 256;  // per-pixel kernels: 64x4 tile, one pixel per thread

// Blocks are dispatched round-robin over the 8 XCDs; remap so each XCD gets a contiguous
// range of tiles (neighbouring tiles share halo lines in that XCD's L2).
LV_D int xcd_tile(int bid, int nblocks) {
    constexpr int X = 8;
    if (nblocks < X) return bid;
    const int xcd = bid % X, idx = bid / X, per = nblocks / X, rem = nblocks % X;
    return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

LV_D uint32_t popc(uint32_t v) { return (uint32_t)__builtin_popcount(v); }
LV_D uint32_t absdiff(uint32_t a, uint32_t b) { return a > b ? a - b : b - a; }
LV_D float fmin_ref(float a, float b) { return (b < a) ? b : a; }  // std::min
LV_D float fmax_ref(float a, float b) { return (a < b) ? b : a; }  // std::max

// Workgroup sum of one value per thread, added to stripe blockIdx % kStripes of a counter array.
// Every thread of the workgroup must call it.
template <int NT>
LV_D void block_add(unsigned long long value, unsigned long long* stripes) {
    __shared__ unsigned long long part[NT / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) value += __shfl_down(value, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = value;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) s += part[i];
        if (s) atomicAdd(&stripes[blockIdx.x % kStripes], s);
    }
}

LV_HD unsigned long long sum_stripes(const unsigned long long* s) {
    unsigned long long t = 0;
    for (int i = 0; i < kStripes; ++i) t += s[i];
    return t;
}

// frame number of the step in flight (m_nFrameIdx after its increment at the top of apply)
LV_D uint32_t cur_frame(const View& v) { return v.sc->frame_idx + 1u; }

// ----------------------------------------------------------------------------------------------
// sample planes
// ----------------------------------------------------------------------------------------------
template <int C>
LV_D void load_sample(const View& v, uint32_t s, uint32_t p, uint32_t col[C], uint32_t desc[C]) {
    const size_t idx = (size_t)s * v.npx + p;
    if constexpr (C == 3) {
        const uint64_t a = ((const uint64_t*)v.sA)[idx];
        desc[0] = (uint32_t)(a & 0xFFFFu);
        desc[1] = (uint32_t)((a >> 16) & 0xFFFFu);
        desc[2] = (uint32_t)((a >> 32) & 0xFFFFu);
        col[0] = (uint32_t)((a >> 48) & 0xFFu);
        col[1] = (uint32_t)(a >> 56);
        col[2] = v.sB[idx];
    } else {
        desc[0] = ((const uint16_t*)v.sA)[idx];
        col[0] = v.sB[idx];
    }
}

template <int C>
LV_D void store_sample(const View& v, uint32_t s, uint32_t p, const uint32_t col[C], const uint32_t desc[C]) {
    const size_t idx = (size_t)s * v.npx + p;
    if constexpr (C == 3) {
        const uint64_t a = (uint64_t)desc[0] | ((uint64_t)desc[1] << 16) | ((uint64_t)desc[2] << 32) |
                           ((uint64_t)col[0] << 48) | ((uint64_t)col[1] << 56);
        ((uint64_t*)v.sA)[idx] = a;
        v.sB[idx] = (uint8_t)col[2];
    } else {
        ((uint16_t*)v.sA)[idx] = (uint16_t)desc[0];
        v.sB[idx] = (uint8_t)col[0];
    }
}

// current colour and intra descriptor of pixel q (the value it writes into samples)
template <int C>
LV_D void current_value(const View& v, const uint8_t* frame, uint32_t q, uint32_t col[C], uint32_t desc[C]) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
        col[c] = frame[(size_t)q * C + c];
        desc[c] = v.desc_cur[(size_t)q * C + c];
    }
}

// ----------------------------------------------------------------------------------------------
// LDS frame tile with a 2-pixel halo (the 5x5 LBSP patch)
// ----------------------------------------------------------------------------------------------
template <int C>
struct FrameTile {
    static constexpr int LW = TW + 4, LH = TH + 4;
    uint8_t px[LW * LH * C];
};

template <int C>
LV_D void load_frame_tile(FrameTile<C>& t, const uint8_t* frame, int W, int H, int x0, int y0) {
    constexpr int LW = FrameTile<C>::LW, LH = FrameTile<C>::LH;
    for (int i = threadIdx.x; i < LW * LH; i += TPB) {
        const int ly = i / LW, lx = i % LW, gx = x0 + lx - 2, gy = y0 + ly - 2;
        const bool in = gx >= 0 && gx < W && gy >= 0 && gy < H;
#pragma unroll
        for (int c = 0; c < C; ++c) t.px[i * C + c] = in ? frame[((size_t)gy * W + gx) * C + c] : (uint8_t)0;
    }
}

// packed 16 taps of channel c around tile-local (lx,ly) (centre coordinates, halo excluded)
template <int C>
LV_D void taps_from_tile(const FrameTile<C>& t, int lx, int ly, int c, uint32_t taps[4]) {
    constexpr int LW = FrameTile<C>::LW;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint32_t acc = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int n = 4 * w + k;
            acc |= (uint32_t)t.px[((ly + 2 + kLbspDY[n]) * LW + (lx + 2 + kLbspDX[n])) * C + c] << (8 * k);
        }
        taps[w] = acc;
    }
}

template <int C>
LV_D void taps_from_global(const uint8_t* frame, int W, int x, int y, int c, uint32_t taps[4]) {
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint32_t acc = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int n = 4 * w + k;
            acc |= (uint32_t)frame[((size_t)(y + kLbspDY[n]) * W + (x + kLbspDX[n])) * C + c] << (8 * k);
        }
        taps[w] = acc;
    }
}

LV_D void load_lut(uint8_t* lut_lds, const Scalars* sc) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) lut_lds[i] = sc->lut[i];
}

// ----------------------------------------------------------------------------------------------
// per-pixel match context
// ----------------------------------------------------------------------------------------------
template <int C>
struct PxCtx {
    uint32_t cur[C], intra[C];
    uint32_t taps[C][4];
    uint32_t t_sc, t_tot_col, t_tot_desc;  // SuBSENSE 3ch / LOBSTER 3ch thresholds
    uint32_t t_col, t_desc;                // 1ch thresholds
};

LV_D uint32_t clamp_u32(uint64_t v) { return v > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)v; }

// SuBSENSE per-pixel thresholds from the previous unstable flag (SUB:355-359, 1ch :222-223)
template <int C>
LV_D void subsense_thresholds(const View& v, float R, uint32_t unst_old, PxCtx<C>& px) {
    const float colf = R * (float)v.color_thr - (float)((uint64_t)(unst_old ? 0u : 1u) * (uint64_t)(v.color_thr / 5u));
    uint64_t col = (uint64_t)colf;
    const uint64_t shift = (uint64_t)floorf(R + 0.5f);
    const uint64_t desc = (1ull << (shift & 63u)) + (uint64_t)v.desc_thr + (uint64_t)unst_old * (uint64_t)v.desc_thr;
    if constexpr (C == 1) {
        col /= 2u;
        px.t_col = clamp_u32(col);
        px.t_desc = clamp_u32(desc);
    } else {
        const uint64_t tot_col = col * 3u, tot_desc = desc * 3u;
        px.t_sc = clamp_u32(tot_col / 2u);
        px.t_tot_col = clamp_u32(tot_col);
        px.t_tot_desc = clamp_u32(tot_desc);
    }
}

// LOBSTER fixed thresholds (LOB:485-488 1ch; :515-518 3ch)
template <int C>
LV_D void lobster_thresholds(const View& v, PxCtx<C>& px) {
    if constexpr (C == 1) {
        px.t_col = v.color_thr / 2u;
        px.t_desc = v.desc_thr;
    } else {
        px.t_tot_desc = v.desc_thr * 3u;
        px.t_tot_col = v.color_thr * 3u;
        px.t_sc = px.t_tot_col / 2u;   // colour, per channel
        px.t_desc = px.t_tot_desc / 2u;  // descriptor, per channel
    }
}

// one sample test; returns true on a match and accumulates the SuBSENSE distances
template <int C, int ALGO>
LV_D bool test_sample(const PxCtx<C>& px, const uint8_t* lut, const uint32_t col[C], const uint32_t desc[C],
                      uint32_t& tot_desc, uint32_t& tot_sum) {
    if constexpr (ALGO == kSuBSENSE && C == 3) {
        uint32_t td = 0, ts = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint32_t cd = absdiff(px.cur[c], col[c]);
            if (cd > px.t_sc) return false;
            const uint32_t inter = lbsp_threshold_packed(px.taps[c], col[c], lut[col[c]]);
            const uint32_t dd = (popc(px.intra[c] ^ desc[c]) + popc(inter ^ desc[c])) >> 1;
            uint32_t sum = (dd >> 1) * 15u + cd;
            sum = sum > 255u ? 255u : sum;
            if (sum > px.t_sc) return false;
            td += dd;
            ts += sum;
        }
        if (td > px.t_tot_desc || ts > px.t_tot_col) return false;
        tot_desc = td;
        tot_sum = ts;
        return true;
    } else if constexpr (ALGO == kSuBSENSE && C == 1) {
        const uint32_t cd = absdiff(px.cur[0], col[0]);
        if (cd > px.t_col) return false;
        const uint32_t inter = lbsp_threshold_packed(px.taps[0], col[0], lut[col[0]]);
        const uint32_t dd = (popc(px.intra[0] ^ desc[0]) + popc(inter ^ desc[0])) >> 1;
        if (dd > px.t_desc) return false;
        uint32_t sum = (dd >> 2) * 15u + cd;
        sum = sum > 255u ? 255u : sum;
        if (sum > px.t_col) return false;
        tot_desc = dd;
        tot_sum = sum;
        return true;
    } else if constexpr (ALGO == kLOBSTER && C == 3) {
        uint32_t tc = 0, td = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint32_t cd = absdiff(px.cur[c], col[c]);
            if (cd > px.t_sc) return false;
            const uint32_t inter = lbsp_threshold_packed(px.taps[c], col[c], lut[col[c]]);
            const uint32_t dd = popc(inter ^ desc[c]);
            if (dd > px.t_desc) return false;
            tc += cd;
            td += dd;
        }
        tot_desc = td;
        tot_sum = tc;
        return td <= px.t_tot_desc && tc <= px.t_tot_col;
    } else {
        const uint32_t cd = absdiff(px.cur[0], col[0]);
        if (cd > px.t_col) return false;
        const uint32_t inter = lbsp_threshold_packed(px.taps[0], col[0], lut[col[0]]);
        tot_desc = 0;
        tot_sum = 0;
        return popc(inter ^ desc[0]) <= px.t_desc;
    }
}

// raster-earlier neighbours that can diffuse into p (5x5 window): offsets in raster order
constexpr int kPredDX[12] = {-2, -1, 0, 1, 2, -2, -1, 0, 1, 2, -2, -1};
constexpr int kPredDY[12] = {-2, -2, -2, -2, -2, -1, -1, -1, -1, -1, 0, 0};

// static forward gate of a record whose target lies after the writer in raster order
LV_D bool fwd_gate(uint32_t r) { return rec_gate_a(r) || (rec_gate_b(r) && rec_ghost_fwd(r)); }

// Candidate writers of p: bit i set when predecessor i targets p with an open forward gate.
LV_D uint32_t candidate_writers(const View& v, int x, int y) {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        const uint32_t q = (uint32_t)((y + kPredDY[i]) * v.W + (x + kPredDX[i]));
        const uint32_t r = v.rec[q];
        const uint32_t code = (uint32_t)((2 - kPredDY[i]) * 5 + (2 - kPredDX[i]));
        if (rec_tgt(r) == code && fwd_gate(r)) m |= 1u << i;
    }
    return m;
}

// Ordered scan s = 0..N-1 until REQ matches (SUB:366-395, LOB:480-495). Samples overwritten by
// applied writers (mask `applied`) are taken from the writer; the last writer in raster order wins.
template <int C, int ALGO>
LV_D uint32_t match_scan(const View& v, const uint8_t* frame, const uint8_t* lut, const PxCtx<C>& px, int x, int y,
                         uint32_t applied) {
    const uint32_t p = (uint32_t)(y * v.W + x);
    uint64_t sbm = 0;  // slots (mod 64) touched by applied writers
    for (uint32_t m = applied; m; m &= m - 1) {
        const int i = __builtin_ctz(m);
        const uint32_t q = (uint32_t)((y + kPredDY[i]) * v.W + (x + kPredDX[i]));
        sbm |= 1ull << (rec_slot_diff(v.rec[q]) & 63u);
    }
    const uint32_t N = (uint32_t)v.N, REQ = (uint32_t)v.REQ;
    uint32_t good = 0, s = 0, md = C * 16u, ms = C * 255u;
    while (good < REQ && s < N) {
        uint32_t col[C], desc[C];
        bool ovr = false;
        if ((sbm >> (s & 63u)) & 1u) {
            for (int i = 11; i >= 0 && !ovr; --i) {
                if (!((applied >> i) & 1u)) continue;
                const uint32_t q = (uint32_t)((y + kPredDY[i]) * v.W + (x + kPredDX[i]));
                if (rec_slot_diff(v.rec[q]) == s) {
                    current_value<C>(v, frame, q, col, desc);
                    ovr = true;
                }
            }
        }
        if (!ovr) load_sample<C>(v, s, p, col, desc);
        uint32_t td = 0, ts = 0;
        if (test_sample<C, ALGO>(px, lut, col, desc, td, ts)) {
            md = td < md ? td : md;
            ms = ts < ms ? ts : ms;
            ++good;
        }
        ++s;
    }
    return mres_pack(good, md, ms, s);
}

// worklist push with de-duplication; stamp = (epoch, iteration)
LV_D void wl_push(const View& v, uint32_t t, uint32_t k, uint64_t epoch, uint32_t* wl_dst) {
    const unsigned long long tag = (unsigned long long)((epoch << 20) | (uint64_t)(k & 0xFFFFFu));
    if (atomicExch(&v.stamp[t], tag) != tag) {
        const uint32_t slot = atomicAdd(&v.sc->wl_count[k % kWlRing], 1u);
        wl_dst[slot] = t;
    }
}

// forward target of p's own diffusion (if any): target after p in raster order with open gate
LV_D bool forward_target(const View& v, uint32_t r, int x, int y, uint32_t& t) {
    const uint32_t tg = rec_tgt(r);
    if (tg == kRecNone || !fwd_gate(r)) return false;
    const int dx = tgt_dx(tg), dy = tgt_dy(tg);
    if (!(dy > 0 || (dy == 0 && dx > 0))) return false;
    t = (uint32_t)((y + dy) * v.W + (x + dx));
    return v.roi[t] != 0;
}

// ----------------------------------------------------------------------------------------------
// describe: LBSP intra descriptors, unstable flag, RNG draws, update record (SUB:355-365,
// :406-449 / LOB:499-512). Writes desc_cur, rec; counts non-zero descriptors.
// ----------------------------------------------------------------------------------------------
template <int C, int ALGO>
__global__ void __launch_bounds__(TPB) k_describe(View v, const uint8_t* __restrict__ frame, double lr) {
    __shared__ FrameTile<C> tile;
    __shared__ uint8_t lut[256];
    const int ntx = (v.W + TW - 1) / TW, nty = (v.H + TH - 1) / TH;
    const int tid = xcd_tile((int)blockIdx.x, ntx * nty);
    const int x0 = (tid % ntx) * TW, y0 = (tid / ntx) * TH;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // per-frame counters consumed by later kernels of this step
        for (int i = 0; i < kWlRing; ++i) v.sc->wl_count[i] = 0;
        v.sc->iters = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < kStripes) {
        v.sc->ex_part[threadIdx.x] = 0;
        v.sc->wr_part[threadIdx.x] = 0;
        v.sc->spec_part[threadIdx.x] = 0;
    }
    load_frame_tile<C>(tile, frame, v.W, v.H, x0, y0);
    load_lut(lut, v.sc);
    __syncthreads();
    const int lx = threadIdx.x % TW, ly = threadIdx.x / TW, x = x0 + lx, y = y0 + ly;
    bool nz = false;
    if (x < v.W && y < v.H) {
        const uint32_t p = (uint32_t)(y * v.W + x);
        if (!v.roi[p]) {
            v.rec[p] = kRecNone;
        } else {
            uint32_t pc = 0;
            uint32_t intra[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                uint32_t taps[4];
                taps_from_tile<C>(tile, lx, ly, c, taps);
                const uint32_t ref = tile.px[((ly + 2) * FrameTile<C>::LW + (lx + 2)) * C + c];
                intra[c] = lbsp_threshold_packed(taps, ref, lut[ref]);
                v.desc_cur[(size_t)p * C + c] = (uint16_t)intra[c];
                pc += popc(intra[c]);
            }
            nz = pc >= (C == 1 ? 2u : 4u);
            const uint32_t frame_no = cur_frame(v);
            uint32_t d[4], d4[4];
            draw4(v.seed_lo, v.seed_hi, v.stream_id, frame_no, PH_APPLY, 0, p, 0, d);
            draw4(v.seed_lo, v.seed_hi, v.stream_id, frame_no, PH_APPLY, 0, p, 1, d4);
            const uint32_t N = (uint32_t)v.N;
            uint32_t r;
            if constexpr (ALGO == kSuBSENSE) {
                const float Rv = v.R[p];
                const uint32_t unst_new =
                    (Rv > 3.0f || (v.RawLT[p] - v.FinLT[p]) > 0.1f || (v.RawST[p] - v.FinST[p]) > 0.1f) ? 1u : 0u;
                const uint64_t L = isinf(lr) ? ~0ull : (lr > 0 ? (uint64_t)ceil(lr) : (uint64_t)ceilf(v.T[p]));
                const bool self_bg = mod64(d[0], L) == 0u;
                const bool self_fg = v.sc->cooldown != 0u && (d[0] % 2u) == 0u;
                const bool use3 = v.use3x3 && !unst_new;
                int tx, ty;
                neighbor_pos(d[2], use3, x, y, v.W, v.H, tx, ty);
                const uint64_t spread = use3 ? L : L / 2u + 1u;
                const bool gate_a = mod64(d[3], spread) == 0u;
                const uint64_t lcap = (uint64_t)v.sc->lower_cap;
                const bool gate_b = mod64(d[3], lcap) == 0u;
                bool ghost_fwd = false;
                if (ty > y || (ty == y && tx > x)) {
                    const uint32_t t = (uint32_t)(ty * v.W + tx);
                    ghost_fwd = v.RawST[t] > 0.995f && v.MLD[t] < 0.010f;
                }
                const uint32_t tg = (uint32_t)((ty - y + 2) * 5 + (tx - x + 2));
                r = tg | ((uint32_t)gate_a << 5) | ((uint32_t)gate_b << 6) | ((uint32_t)ghost_fwd << 7) |
                    ((uint32_t)self_bg << 8) | ((uint32_t)self_fg << 9) | (unst_new << 10) | ((d[1] % N) << 16) |
                    ((d4[0] % N) << 24);
            } else {
                const uint64_t L = isinf(lr) ? ~0ull : (uint64_t)ceil(lr);
                const bool self_bg = mod64(d[0], L) == 0u;
                const bool gate_a = mod64(d[2], L) == 0u;
                int tx, ty;
                neighbor_pos(d[3], true, x, y, v.W, v.H, tx, ty);
                const uint32_t tg = (uint32_t)((ty - y + 2) * 5 + (tx - x + 2));
                r = tg | ((uint32_t)gate_a << 5) | ((uint32_t)self_bg << 8) | ((d[1] % N) << 16) | ((d4[0] % N) << 24);
            }
            v.rec[p] = r;
        }
    }
    if constexpr (ALGO == kSuBSENSE) block_add<TPB>(nz ? 1ull : 0ull, v.sc->nz_part);
}

// builds the match context of ROI pixel (x,y); taps from the LDS tile or from global memory
template <int C, int ALGO, bool FROM_TILE>
LV_D void build_ctx(const View& v, const uint8_t* frame, const FrameTile<C>* tile, int lx, int ly, int x, int y,
                    PxCtx<C>& px) {
    const uint32_t p = (uint32_t)(y * v.W + x);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        px.cur[c] = frame[(size_t)p * C + c];
        px.intra[c] = v.desc_cur[(size_t)p * C + c];
        if constexpr (FROM_TILE) taps_from_tile<C>(*tile, lx, ly, c, px.taps[c]);
        else taps_from_global<C>(frame, v.W, x, y, c, px.taps[c]);
    }
    if constexpr (ALGO == kSuBSENSE) subsense_thresholds<C>(v, v.R[p], v.unstable[p], px);
    else lobster_thresholds<C>(v, px);
}

// ----------------------------------------------------------------------------------------------
// speculative match: raster-earlier writers assumed to decide as in the previous frame
// ----------------------------------------------------------------------------------------------
template <int C, int ALGO>
LV_D uint32_t match_pixel(const View& v, const uint8_t* frame, const uint8_t* lut, const FrameTile<C>& tile, int lx,
                          int ly, int x, int y, uint32_t p);

template <int C, int ALGO>
__global__ void __launch_bounds__(TPB) k_match(View v, const uint8_t* __restrict__ frame) {
    __shared__ FrameTile<C> tile;
    __shared__ uint8_t lut[256];
    const int ntx = (v.W + TW - 1) / TW, nty = (v.H + TH - 1) / TH;
    const int tid = xcd_tile((int)blockIdx.x, ntx * nty);
    const int x0 = (tid % ntx) * TW, y0 = (tid / ntx) * TH;
    load_frame_tile<C>(tile, frame, v.W, v.H, x0, y0);
    load_lut(lut, v.sc);
    __syncthreads();
    const int lx = threadIdx.x % TW, ly = threadIdx.x / TW, x = x0 + lx, y = y0 + ly;
    const uint32_t p = (uint32_t)(y * v.W + x);
    uint32_t examined = 0;
    if (x < v.W && y < v.H && v.roi[p]) examined = match_pixel<C, ALGO>(v, frame, lut, tile, lx, ly, x, y, p);
    block_add<TPB>(examined, v.sc->spec_part);
}

template <int C, int ALGO>
LV_D uint32_t match_pixel(const View& v, const uint8_t* frame, const uint8_t* lut, const FrameTile<C>& tile, int lx,
                          int ly, int x, int y, uint32_t p) {
    PxCtx<C> px;
    build_ctx<C, ALGO, true>(v, frame, &tile, lx, ly, x, y, px);
    const uint32_t cand = candidate_writers(v, x, y);
    uint32_t applied = 0;
    for (uint32_t m = cand; m; m &= m - 1) {
        const int i = __builtin_ctz(m);
        const uint32_t q = (uint32_t)((y + kPredDY[i]) * v.W + (x + kPredDX[i]));
        if (v.last_raw[q] == 0) applied |= 1u << i;  // guess: background as in the last frame
    }
    const uint32_t mr = match_scan<C, ALGO>(v, frame, lut, px, x, y, applied);
    const uint8_t fg = (mr & 255u) < (uint32_t)v.REQ ? 1 : 0;
    v.mres[p] = mr;
    v.dec[p] = fg;
    const uint8_t guess = v.last_raw[p] ? 1 : 0;
    uint32_t t;
    if (fg != guess && forward_target(v, v.rec[p], x, y, t)) wl_push(v, t, 1u, v.sc->epoch, v.wl1);
    return mr >> 24;
}

// re-match of one worklisted pixel with the current decisions of its writers
template <int C, int ALGO>
LV_D void rematch(const View& v, const uint8_t* frame, const uint8_t* lut, uint32_t p, uint32_t k, uint64_t epoch,
                  uint32_t* wl_next) {
    const int x = (int)(p % (uint32_t)v.W), y = (int)(p / (uint32_t)v.W);
    PxCtx<C> px;
    build_ctx<C, ALGO, false>(v, frame, nullptr, 0, 0, x, y, px);
    const uint32_t cand = candidate_writers(v, x, y);
    uint32_t applied = 0;
    for (uint32_t m = cand; m; m &= m - 1) {
        const int i = __builtin_ctz(m);
        const uint32_t q = (uint32_t)((y + kPredDY[i]) * v.W + (x + kPredDX[i]));
        if (__hip_atomic_load(&v.dec[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) applied |= 1u << i;
    }
    const uint32_t mr = match_scan<C, ALGO>(v, frame, lut, px, x, y, applied);
    const uint8_t fg = (mr & 255u) < (uint32_t)v.REQ ? 1 : 0;
    v.mres[p] = mr;
    const uint8_t old = v.dec[p];
    if (fg != old) {
        __hip_atomic_store(&v.dec[p], fg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t t;
        if (forward_target(v, v.rec[p], x, y, t)) wl_push(v, t, k + 1u, epoch, wl_next);
    }
}

// fix-point iteration k (k >= 1): worklist k -> worklist k+1
template <int C, int ALGO>
__global__ void __launch_bounds__(TPB) k_fixpoint(View v, const uint8_t* __restrict__ frame, uint32_t k) {
    __shared__ uint8_t lut[256];
    // slot k+2 was last read by iteration k-2 (finished); it must be empty before iteration k+1
    // pushes into it, whether or not this iteration has work.
    if (blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(&v.sc->wl_count[(k + 2) % kWlRing], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t n = __hip_atomic_load(&v.sc->wl_count[k % kWlRing], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (n == 0) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&v.sc->iters, 1u);
        atomicMax(&v.sc->max_iter_seen, k);
    }
    load_lut(lut, v.sc);
    __syncthreads();
    const uint32_t* wl = (k & 1u) ? v.wl1 : v.wl0;
    uint32_t* wl_next = (k & 1u) ? v.wl0 : v.wl1;
    const uint64_t epoch = v.sc->epoch;
    for (uint32_t i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB)
        rematch<C, ALGO>(v, frame, lut, wl[i], k, epoch, wl_next);
}

// remaining fix-point iterations in one workgroup (normally no work is left)
template <int C, int ALGO>
__global__ void __launch_bounds__(1024) k_fixpoint_tail(View v, const uint8_t* __restrict__ frame, uint32_t k0,
                                                        uint32_t cap) {
    __shared__ uint8_t lut[256];
    __shared__ uint32_t s_n;
    load_lut(lut, v.sc);
    const uint64_t epoch = v.sc->epoch;
    for (uint32_t k = k0;; ++k) {
        __syncthreads();
        if (threadIdx.x == 0) {
            s_n = __hip_atomic_load(&v.sc->wl_count[k % kWlRing], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (s_n != 0) {
                if (k - k0 >= cap) {
                    atomicOr(&v.sc->err_flags, kErrFixpointCap);
                    s_n = 0;
                } else {
                    __hip_atomic_store(&v.sc->wl_count[(k + 2) % kWlRing], 0u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    v.sc->iters += 1u;
                    v.sc->max_iter_seen = k > v.sc->max_iter_seen ? k : v.sc->max_iter_seen;
                }
            }
        }
        __syncthreads();
        const uint32_t n = s_n;
        if (n == 0) break;
        const uint32_t* wl = (k & 1u) ? v.wl1 : v.wl0;
        uint32_t* wl_next = (k & 1u) ? v.wl0 : v.wl1;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) rematch<C, ALGO>(v, frame, lut, wl[i], k, epoch, wl_next);
        __threadfence();
    }
}

// ----------------------------------------------------------------------------------------------
// SuBSENSE feedback maps (SUB:396-420, :451-480; 1ch :254-329), raw mask and blink bookkeeping
// (SUB:536-539). Whole image: pixels outside the ROI get raw = 0.
// ----------------------------------------------------------------------------------------------
template <int C>
__global__ void __launch_bounds__(TPB) k_feedback(View v, const uint8_t* __restrict__ frame) {
    const uint32_t p = blockIdx.x * TPB + threadIdx.x;
    const bool valid = p < v.npx;
    uint32_t examined = 0;
    uint8_t raw = 0;
    if (valid && v.roi[p]) {
        const uint32_t frame_no = cur_frame(v);
        const uint32_t mvn = v.mv_avgs;
        const float aLT = 1.0f / (float)(frame_no < mvn ? frame_no : mvn);
        const float aST = 1.0f / (float)(frame_no < mvn / 4u ? frame_no : mvn / 4u);
        const uint32_t mr = v.mres[p];
        const uint32_t good = mr & 255u, md = (mr >> 8) & 63u, ms = (mr >> 14) & 1023u;
        examined = mr >> 24;
        const bool fg = v.dec[p] != 0;
        const uint32_t r = v.rec[p];
        const uint32_t unst_new = rec_unstable(r) ? 1u : 0u;
        const float col_max = (float)(C * 255), desc_max = (float)(C * 16);
        // D_last (SUB:396-397): L1dist<3> accumulates in uchar (wraps mod 256)
        uint32_t l1 = 0, hd = 0;
        uint32_t cur[C], intra[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            cur[c] = frame[(size_t)p * C + c];
            intra[c] = v.desc_cur[(size_t)p * C + c];
            l1 += absdiff(v.last_color[(size_t)p * C + c], cur[c]);
            hd += popc((uint32_t)v.last_desc[(size_t)p * C + c] ^ intra[c]);
        }
        l1 &= 255u;
        const float nld = ((float)l1 / col_max + (float)hd / desc_max) / 2.0f;
        const float mld = v.MLD[p] * (1.0f - aST) + nld * aST;
        v.MLD[p] = mld;
        float dlt = v.DminLT[p], dst = v.DminST[p], rlt = v.RawLT[p], rst = v.RawST[p];
        const uint32_t REQ = (uint32_t)v.REQ;
        if (fg) {
            const float nmd = fmin_ref(1.0f, ((float)ms / col_max + (float)md / desc_max) / 2.0f +
                                                 (float)(REQ - good) / (float)REQ);
            dlt = dlt * (1.0f - aLT) + nmd * aLT;
            dst = dst * (1.0f - aST) + nmd * aST;
            rlt = rlt * (1.0f - aLT) + aLT;
            rst = rst * (1.0f - aST) + aST;
            raw = 255;
        } else {
            const float nmd = ((float)ms / col_max + (float)md / desc_max) / 2.0f;
            dlt = dlt * (1.0f - aLT) + nmd * aLT;
            dst = dst * (1.0f - aST) + nmd * aST;
            rlt = rlt * (1.0f - aLT);
            rst = rst * (1.0f - aST);
        }
        v.DminLT[p] = dlt;
        v.DminST[p] = dst;
        v.RawLT[p] = rlt;
        v.RawST[p] = rst;
        const float lcap = v.sc->lower_cap, ucap = v.sc->upper_cap;
        const float mn = fmin_ref(dlt, dst), mx = fmax_ref(dlt, dst);
        const bool last_fg = v.last_fg[p] != 0;
        float T = v.T[p], V = v.V[p], R = v.R[p];
        if (last_fg || (mn < 0.1f && raw)) {
            if (T < ucap) T += 0.5f / (mx * V);
        } else if (T > lcap) {
            T -= 0.25f * V / mx;
        }
        if (T < lcap) T = lcap;
        else if (T > ucap) T = ucap;
        if (mx > 0.1f && v.blinks[p]) {
            V += 1.0f;
        } else if (V > 0.1f) {
            V -= last_fg ? 0.1f / 4 : unst_new ? 0.1f / 2 : 0.1f;
            if (V < 0.1f) V = 0.1f;
        }
        const float base = 1.0f + mn * 2.0f;
        if ((double)R < (double)base * (double)base) {  // std::pow(float,2) in double: an exact square
            R += 0.01f * (V - 0.1f);
        } else {
            R -= 0.01f / V;
            if (R < 1.0f) R = 1.0f;
        }
        v.T[p] = T;
        v.V[p] = V;
        v.R[p] = R;
        v.unstable[p] = (uint8_t)unst_new;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            v.last_color[(size_t)p * C + c] = (uint8_t)cur[c];
            v.last_desc[(size_t)p * C + c] = (uint16_t)intra[c];
        }
    }
    if (valid) {
        const uint8_t cb = raw ^ v.last_raw[p];
        v.blinks[p] = cb | v.last_raw_blink[p];
        v.last_raw_blink[p] = cb;
        v.last_raw[p] = raw;
        v.raw[p] = raw;
    }
    block_add<TPB>(examined, v.sc->ex_part);  // workload counter (sum of k_p)
}

// LOBSTER: raw mask (LOB:496-497, :554-555) and previous-raw bookkeeping for the next speculation
__global__ void __launch_bounds__(TPB) k_lobster_raw(View v) {
    const uint32_t p = blockIdx.x * TPB + threadIdx.x;
    uint32_t examined = 0;
    if (p < v.npx) {
        uint8_t raw = 0;
        if (v.roi[p]) {
            raw = v.dec[p] ? 255 : 0;
            examined = v.mres[p] >> 24;
        }
        v.raw[p] = raw;
        v.last_raw[p] = raw;
    }
    block_add<TPB>(examined, v.sc->ex_part);
}

// ----------------------------------------------------------------------------------------------
// model update: every pixel of the inner rectangle gathers the writes aimed at its samples, in
// the raster order of the writers; the last write of each slot wins (SUB:406-449, LOB:498-512).
// ----------------------------------------------------------------------------------------------
template <int C, int ALGO>
__global__ void __launch_bounds__(TPB) k_update(View v, const uint8_t* __restrict__ frame) {
    const uint32_t iw = (uint32_t)(v.W - 4), ih = (uint32_t)(v.H - 4);
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    uint32_t nw = 0;
    if (i < iw * ih) {
        const int x = 2 + (int)(i % iw), y = 2 + (int)(i / iw);
        const uint32_t p = (uint32_t)(y * v.W + x);
        // ghost test of writers at or after p reads p's updated maps (SUB:438-441)
        bool ghost_new = false;
        if constexpr (ALGO == kSuBSENSE) ghost_new = v.RawST[p] > 0.995f && v.MLD[p] < 0.010f;
        uint8_t slot[26];
        uint8_t src[26];
        int n = 0;
#pragma unroll 1
        for (int j = 0; j < 25; ++j) {
            const int dy = j / 5 - 2, dx = j % 5 - 2;
            const uint32_t q = (uint32_t)((y + dy) * v.W + (x + dx));
            const uint32_t r = v.rec[q];
            const uint32_t tg = rec_tgt(r);
            if (tg == kRecNone) continue;
            const bool bg = v.dec[q] == 0;
            if (j == 12) {  // the pixel itself: self update (or foreground replacement) first
                if (bg ? rec_self_bg(r) : (ALGO == kSuBSENSE && rec_self_fg(r))) {
                    slot[n] = (uint8_t)rec_slot_self(r);
                    src[n++] = 12;
                }
            }
            if (!bg || tg != (uint32_t)((2 - dy) * 5 + (2 - dx))) continue;
            const bool gate = j < 12 ? fwd_gate(r) : (rec_gate_a(r) || (rec_gate_b(r) && ghost_new));
            if (gate) {
                slot[n] = (uint8_t)rec_slot_diff(r);
                src[n++] = (uint8_t)j;
            }
        }
        for (int a = n - 1; a >= 0; --a) {
            bool later = false;
            for (int b = a + 1; b < n; ++b) later |= slot[b] == slot[a];
            if (later) continue;
            const int j = src[a];
            const uint32_t q = (uint32_t)((y + j / 5 - 2) * v.W + (x + j % 5 - 2));
            uint32_t col[C], desc[C];
            current_value<C>(v, frame, q, col, desc);
            store_sample<C>(v, slot[a], p, col, desc);
        }
        nw = (uint32_t)n;
    }
    block_add<TPB>(nw, v.sc->wr_part);
}

// LOBSTER: last colour frame := input (LOB:580)
template <int C>
__global__ void __launch_bounds__(TPB) k_copy_last_color(View v, const uint8_t* __restrict__ frame) {
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i < v.npx * (uint32_t)C) v.last_color[i] = frame[i];
}

// ----------------------------------------------------------------------------------------------
// post-processing (SUB:540-554, LOB:578). Masks are 0/255 bytes.
// ----------------------------------------------------------------------------------------------
constexpr int PW = 64, PH = 16;  // post-processing tile; 256 threads x 4 pixels

// morphologyEx CLOSE with a 3x3 rectangle, neutral border: erode3(dilate3(raw))
__global__ void __launch_bounds__(TPB) k_close(View v) {
    constexpr int LW = PW + 4, LH = PH + 4;
    __shared__ uint8_t s_raw[LH][LW];
    __shared__ uint8_t s_dil[LH - 2][LW - 2];
    const int ntx = (v.W + PW - 1) / PW, nty = (v.H + PH - 1) / PH;
    const int tid = xcd_tile((int)blockIdx.x, ntx * nty);
    const int x0 = (tid % ntx) * PW, y0 = (tid / ntx) * PH;
    for (int i = threadIdx.x; i < LW * LH; i += TPB) {
        const int ly = i / LW, lx = i % LW, gx = x0 + lx - 2, gy = y0 + ly - 2;
        s_raw[ly][lx] = (gx >= 0 && gx < v.W && gy >= 0 && gy < v.H) ? v.raw[(size_t)gy * v.W + gx] : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < (LW - 2) * (LH - 2); i += TPB) {
        const int ly = i / (LW - 2), lx = i % (LW - 2), gx = x0 + lx - 1, gy = y0 + ly - 1;
        uint8_t m = 0;
        if (gx >= 0 && gx < v.W && gy >= 0 && gy < v.H) {
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) m = s_raw[ly + a][lx + b] > m ? s_raw[ly + a][lx + b] : m;
        } else {
            m = 255;  // outside the image: neutral for the erosion
        }
        s_dil[ly][lx] = m;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < PW * PH; i += TPB) {
        const int ly = i / PW, lx = i % PW, gx = x0 + lx, gy = y0 + ly;
        if (gx >= v.W || gy >= v.H) continue;
        uint8_t m = 255;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) m = s_dil[ly + a][lx + b] < m ? s_dil[ly + a][lx + b] : m;
        v.pre[(size_t)gy * v.W + gx] = m;
    }
}

// --- hole fill: union-find labelling of the 0-pixels of PreFlood (4-connectivity) ---
constexpr int CT = 32;  // CCL tile (32x32)

__global__ void __launch_bounds__(TPB) k_ccl_local(View v) {
    __shared__ uint32_t lab[CT * CT];
    const int ntx = (v.W + CT - 1) / CT;
    const int x0 = ((int)blockIdx.x % ntx) * CT, y0 = ((int)blockIdx.x / ntx) * CT;
    for (int i = threadIdx.x; i < CT * CT; i += TPB) {
        const int gx = x0 + i % CT, gy = y0 + i / CT;
        const bool bgpx = gx < v.W && gy < v.H && v.pre[(size_t)gy * v.W + gx] == 0;
        lab[i] = bgpx ? (uint32_t)i : 0xFFFFFFFFu;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < CT * CT; i += TPB) {
        if (lab[i] == 0xFFFFFFFFu) continue;
        const int lx = i % CT, ly = i / CT;
        const int nb[2] = {lx + 1 < CT ? i + 1 : -1, ly + 1 < CT ? i + CT : -1};
        for (int k = 0; k < 2; ++k) {
            if (nb[k] < 0) continue;
            uint32_t a = (uint32_t)i, b = (uint32_t)nb[k];
            if (__hip_atomic_load(&lab[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0xFFFFFFFFu) continue;
            bool done = false;
            while (!done) {
                while (true) {
                    const uint32_t pa = __hip_atomic_load(&lab[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (pa == a) break;
                    a = pa;
                }
                while (true) {
                    const uint32_t pb = __hip_atomic_load(&lab[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (pb == b) break;
                    b = pb;
                }
                if (a < b) {
                    const uint32_t old = atomicMin(&lab[b], a);
                    done = old == b;
                    b = old;
                } else if (b < a) {
                    const uint32_t old = atomicMin(&lab[a], b);
                    done = old == a;
                    a = old;
                } else {
                    done = true;
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < CT * CT; i += TPB) {
        const int gx = x0 + i % CT, gy = y0 + i / CT;
        if (gx >= v.W || gy >= v.H) continue;
        uint32_t a = lab[i];
        if (a != 0xFFFFFFFFu) {
            while (lab[a] != a) a = lab[a];
            a = (uint32_t)((y0 + (int)a / CT) * v.W + (x0 + (int)a % CT));
        }
        v.label[(size_t)gy * v.W + gx] = a;
    }
}

LV_D uint32_t g_find(uint32_t* lab, uint32_t a) {
    while (true) {
        const uint32_t pa = __hip_atomic_load(&lab[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (pa == a) return a;
        a = pa;
    }
}

LV_D void g_union(uint32_t* lab, uint32_t a, uint32_t b) {
    bool done = false;
    while (!done) {
        a = g_find(lab, a);
        b = g_find(lab, b);
        if (a < b) {
            const uint32_t old = atomicMin(&lab[b], a);
            done = old == b;
            b = old;
        } else if (b < a) {
            const uint32_t old = atomicMin(&lab[a], b);
            done = old == a;
            a = old;
        } else {
            done = true;
        }
    }
}

// unions across CCL tile borders (right column and bottom row of every tile)
__global__ void __launch_bounds__(TPB) k_ccl_border(View v) {
    const int ntx = (v.W + CT - 1) / CT, nty = (v.H + CT - 1) / CT;
    const uint32_t per_tile = 2 * CT;
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= (uint32_t)(ntx * nty) * per_tile) return;
    const int t = (int)(i / per_tile), k = (int)(i % per_tile);
    const int x0 = (t % ntx) * CT, y0 = (t / ntx) * CT;
    int ax, ay, bx, by;
    if (k < CT) { ax = x0 + CT - 1; ay = y0 + k; bx = ax + 1; by = ay; }
    else { ax = x0 + (k - CT); ay = y0 + CT - 1; bx = ax; by = ay + 1; }
    if (ax >= v.W || ay >= v.H || bx >= v.W || by >= v.H) return;
    const uint32_t a = (uint32_t)(ay * v.W + ax), b = (uint32_t)(by * v.W + bx);
    if (v.pre[a] != 0 || v.pre[b] != 0) return;
    // labels still hold tile-local roots; the previous pixel along the edge usually carries the
    // same pair of local roots, in which case its union already covers this one
    const uint32_t la = v.label[a], lb = v.label[b];
    if (k % CT != 0) {
        const uint32_t pa = k < CT ? a - (uint32_t)v.W : a - 1u, pb = k < CT ? b - (uint32_t)v.W : b - 1u;
        if (v.pre[pa] == 0 && v.pre[pb] == 0 && v.label[pa] == la && v.label[pb] == lb) return;
    }
    g_union(v.label, la, lb);
}

// holes = 0-pixels of PreFlood not connected to pixel (0,0) (floodFill + bitwise_not, SUB:541-543);
// mh = raw | holes
__global__ void __launch_bounds__(TPB) k_holes(View v) {
    const uint32_t p = blockIdx.x * TPB + threadIdx.x;
    if (p >= v.npx) return;
    uint8_t hole = 0;
    if (v.pre[p] == 0) {
        const uint32_t r = g_find(v.label, p);
        v.label[p] = r;
        hole = r != 0u ? 255 : 0;
    }
    v.mh[p] = v.raw[p] | hole;
}

// mask = raw | holes | erode7(PreFlood); last_fg = medianBlur(mask, k) (SUB:544-547; LOB:578 uses
// the raw mask directly). Binary median with zero padding = majority count over the k x k window.
template <bool SUBSENSE>
__global__ void __launch_bounds__(TPB) k_median(View v) {
    constexpr int HM = 6, HE = 3;                  // median halo (k <= 13), erosion halo
    constexpr int MW = PW + 2 * HM, MH = PH + 2 * HM;  // mask region
    constexpr int EW = MW + 2 * HE, EH = MH + 2 * HE;  // PreFlood region
    __shared__ uint8_t s_pre[EH][EW];
    __shared__ uint8_t s_eh[EH][MW];
    __shared__ uint8_t s_mask[MH][MW];
    __shared__ uint8_t s_col[PH][MW];
    const int ntx = (v.W + PW - 1) / PW, nty = (v.H + PH - 1) / PH;
    const int tid = xcd_tile((int)blockIdx.x, ntx * nty);
    const int x0 = (tid % ntx) * PW, y0 = (tid / ntx) * PH;
    const int k = v.median_k, r = k / 2, need = (k * k + 1) / 2;
    if constexpr (SUBSENSE) {
        for (int i = threadIdx.x; i < EW * EH; i += TPB) {
            const int ly = i / EW, lx = i % EW, gx = x0 + lx - HM - HE, gy = y0 + ly - HM - HE;
            s_pre[ly][lx] = (gx >= 0 && gx < v.W && gy >= 0 && gy < v.H) ? v.pre[(size_t)gy * v.W + gx] : 255;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < EH * MW; i += TPB) {
            const int ly = i / MW, lx = i % MW;
            uint8_t m = 255;
            for (int b = 0; b < 2 * HE + 1; ++b) m = s_pre[ly][lx + b] < m ? s_pre[ly][lx + b] : m;
            s_eh[ly][lx] = m;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < MW * MH; i += TPB) {
        const int ly = i / MW, lx = i % MW, gx = x0 + lx - HM, gy = y0 + ly - HM;
        uint8_t m = 0;
        if (gx >= 0 && gx < v.W && gy >= 0 && gy < v.H) {
            if constexpr (SUBSENSE) {
                uint8_t e = 255;
                for (int a = 0; a < 2 * HE + 1; ++a) e = s_eh[ly + a][lx] < e ? s_eh[ly + a][lx] : e;
                m = v.mh[(size_t)gy * v.W + gx] | e;
            } else {
                m = v.raw[(size_t)gy * v.W + gx];
            }
        }
        s_mask[ly][lx] = m ? 1 : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < PH * MW; i += TPB) {
        const int ly = i / MW, lx = i % MW;
        int c = 0;
        for (int a = -r; a <= r; ++a) c += s_mask[ly + HM + a][lx];
        s_col[ly][lx] = (uint8_t)c;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < PW * PH; i += TPB) {
        const int ly = i / PW, lx = i % PW, gx = x0 + lx, gy = y0 + ly;
        if (gx >= v.W || gy >= v.H) continue;
        int c = 0;
        for (int b = -r; b <= r; ++b) c += s_col[ly][lx + HM + b];
        v.last_fg[(size_t)gy * v.W + gx] = c >= need ? 255 : 0;
    }
}

// dilate7(last_fg), blink masks (SUB:548-551), final-segmentation averages (SUB:553-554) and the
// output mask. addWeighted(.., CV_32F) is evaluated as the double-precision scalar form.
__global__ void __launch_bounds__(TPB) k_dilate_final(View v, uint8_t* __restrict__ out_mask) {
    constexpr int H3 = 3, LW = PW + 2 * H3, LH = PH + 2 * H3;
    __shared__ uint8_t s_fg[LH][LW];
    __shared__ uint8_t s_h[LH][PW];
    const int ntx = (v.W + PW - 1) / PW, nty = (v.H + PH - 1) / PH;
    const int tid = xcd_tile((int)blockIdx.x, ntx * nty);
    const int x0 = (tid % ntx) * PW, y0 = (tid / ntx) * PH;
    for (int i = threadIdx.x; i < LW * LH; i += TPB) {
        const int ly = i / LW, lx = i % LW, gx = x0 + lx - H3, gy = y0 + ly - H3;
        s_fg[ly][lx] = (gx >= 0 && gx < v.W && gy >= 0 && gy < v.H) ? v.last_fg[(size_t)gy * v.W + gx] : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < LH * PW; i += TPB) {
        const int ly = i / PW, lx = i % PW;
        uint8_t m = 0;
        for (int b = 0; b < 2 * H3 + 1; ++b) m = s_fg[ly][lx + b] > m ? s_fg[ly][lx + b] : m;
        s_h[ly][lx] = m;
    }
    __syncthreads();
    const uint32_t frame_no = cur_frame(v), mvn = v.mv_avgs;
    const float aLT = 1.0f / (float)(frame_no < mvn ? frame_no : mvn);
    const float aST = 1.0f / (float)(frame_no < mvn / 4u ? frame_no : mvn / 4u);
    const double alt = (double)(1.0f - aLT), blt = (1.0 / 255) * (double)aLT;
    const double ast = (double)(1.0f - aST), bst = (1.0 / 255) * (double)aST;
    for (int i = threadIdx.x; i < PW * PH; i += TPB) {
        const int ly = i / PW, lx = i % PW, gx = x0 + lx, gy = y0 + ly;
        if (gx >= v.W || gy >= v.H) continue;
        uint8_t d = 0;
        for (int a = 0; a < 2 * H3 + 1; ++a) d = s_h[ly + a][lx] > d ? s_h[ly + a][lx] : d;
        const size_t p = (size_t)gy * v.W + gx;
        const uint8_t inv = (uint8_t)~d;
        v.blinks[p] = (uint8_t)(v.blinks[p] & v.last_dil_inv[p] & inv);
        v.last_dil_inv[p] = inv;
        const uint8_t fg = s_fg[ly + H3][lx + H3];
        const double f = (double)fg;
        v.FinLT[p] = (float)((double)v.FinLT[p] * alt + f * blt + 0.0);
        v.FinST[p] = (float)((double)v.FinST[p] * ast + f * bst + 0.0);
        if (out_mask) out_mask[p] = fg;
    }
}

__global__ void __launch_bounds__(TPB) k_copy_mask(View v, uint8_t* __restrict__ out_mask) {
    const uint32_t p = blockIdx.x * TPB + threadIdx.x;
    if (p < v.npx) out_mask[p] = v.last_fg[p];
}

// ----------------------------------------------------------------------------------------------
// frame-level controller (SUB:555-611)
// ----------------------------------------------------------------------------------------------
LV_D uint8_t sat_u8(float f) {
    const float r = rintf(f);  // round half to even, like cvRound
    return (uint8_t)(r < 0.f ? 0.f : r > 255.f ? 255.f : r);
}

// resize(INTER_AREA, W/8 x H/8) + accumulateWeighted LT/ST + per-cell |ST-LT| (SUB:568-583)
template <int C>
__global__ void __launch_bounds__(TPB) k_downsample(View v, const uint8_t* __restrict__ frame) {
    const uint32_t cell = blockIdx.x * TPB + threadIdx.x;
    const uint32_t ncell = (uint32_t)(v.dsW * v.dsH);
    uint32_t contrib = 0;
    if (cell < ncell) {
        const int dx = (int)(cell % (uint32_t)v.dsW), dy = (int)(cell / (uint32_t)v.dsW);
        const uint32_t frame_no = cur_frame(v), mvn = v.mv_avgs;
        const float aL = 1.0f / (float)(frame_no < mvn ? frame_no : mvn);
        const float aS = 1.0f / (float)(frame_no < mvn / 4u ? frame_no : mvn / 4u);
        const float bL = 1.0f - aL, bS = 1.0f - aS;
        uint32_t best = 0;
        for (int c = 0; c < C; ++c) {
            uint8_t m;
            if (v.W % v.dsW == 0 && v.H % v.dsH == 0) {
                const int sx = v.W / v.dsW, sy = v.H / v.dsH;
                int sum = 0;
                for (int yy = 0; yy < sy; ++yy)
                    for (int xx = 0; xx < sx; ++xx) sum += frame[((size_t)(dy * sy + yy) * v.W + (dx * sx + xx)) * C + c];
                m = sat_u8((float)sum * (1.f / (float)(sx * sy)));
            } else {
                const double sx = (double)v.W / v.dsW, sy = (double)v.H / v.dsH;
                const double yb = dy * sy, ye = (dy + 1) * sy, xb = dx * sx, xe = (dx + 1) * sx;
                double acc = 0.0;
                for (int yy = (int)yb; yy < v.H && (double)yy < ye; ++yy) {
                    const double wy = fmin((double)(yy + 1), ye) - fmax((double)yy, yb);
                    for (int xx = (int)xb; xx < v.W && (double)xx < xe; ++xx) {
                        const double wx = fmin((double)(xx + 1), xe) - fmax((double)xx, xb);
                        acc += (double)frame[((size_t)yy * v.W + xx) * C + c] * (wy * wx);
                    }
                }
                m = sat_u8((float)(acc / (sx * sy)));
            }
            const size_t di = (size_t)cell * C + c;
            const float lt = (float)m * aL + v.dsLT[di] * bL;
            const float st = (float)m * aS + v.dsST[di] * bS;
            v.dsLT[di] = lt;
            v.dsST[di] = st;
            const uint32_t d = (uint32_t)fabsf(st - lt);
            best = d > best ? d : best;
        }
        contrib = C == 1 ? best / 2u : best;
    }
    block_add<TPB>(contrib, v.sc->cd_part);
}

// single-workgroup decision: LUT adaptation, model reset, learning-rate caps, counters
__global__ void __launch_bounds__(256) k_decide(View v) {
    Scalars* sc = v.sc;
    __shared__ float s_ratio, s_last;
    __shared__ unsigned long long s_cd;
    if (threadIdx.x == 0) {
        s_ratio = (float)sum_stripes(sc->nz_part) / (float)v.n_roi;
        s_last = sc->last_nz_ratio;
        s_cd = sum_stripes(sc->cd_part);
    }
    __syncthreads();
    if (threadIdx.x < kStripes) {
        sc->nz_part[threadIdx.x] = 0;
        sc->cd_part[threadIdx.x] = 0;
    }
    const float ratio = s_ratio, last = s_last;
    const uint32_t t = threadIdx.x;
    uint8_t lut = sc->lut[t];
    if (ratio < 0.1f && last < 0.1f) {
        const float lb = (float)v.lbsp_offset + ceilf((float)t * v.rel_lbsp / 4.0f);
        if (lut > sat_u8(lb)) --lut;
    } else if (ratio > 0.5f && last > 0.5f) {
        const uint8_t ub = sat_u8((float)v.lbsp_offset + 255.0f * v.rel_lbsp);
        if (lut < ub) ++lut;
    }
    __syncthreads();
    sc->lut[t] = lut;
    if (threadIdx.x != 0) return;
    sc->last_nz_ratio = ratio;
    sc->refresh_pending = 0;
    if (v.lr_scaling) {
        const float cratio = (float)s_cd / (float)(v.dsW * v.dsH);
        const uint32_t thr = v.color_thr / 2u;  // FRAMELEVEL_MIN_COLOR_DIFF_THRESHOLD
        if (sc->auto_reset) {
            if (sc->frames_since_reset > 1000u) {
                sc->auto_reset = 0;
            } else if (cratio >= (float)thr && sc->cooldown == 0u) {
                sc->frames_since_reset = 0;
                sc->refresh_pending = 1;  // refreshModel(0.1f) + T := 1 (SUB:589-592)
                sc->refresh_frac = 0.1f;
                sc->refresh_force_fg = 0;
                sc->refresh_frame = cur_frame(v);
                sc->refresh_ev = sc->refresh_seq++;
                sc->cooldown = v.mv_avgs / 4u;
            } else {
                ++sc->frames_since_reset;
            }
        } else if (cratio >= (float)(thr * 2u)) {
            sc->frames_since_reset = 0;
            sc->auto_reset = 1;
        }
        if (cratio >= (float)(thr / 2u)) {
            const int k = (int)(cratio / 2.0f);  // x86 shifts by the count mod 32
            const int lo = 2 >> (k & 31), hi = 256 >> (k & 31);
            sc->lower_cap = (float)(lo > 1 ? lo : 1);
            sc->upper_cap = (float)(hi > 1 ? hi : 1);
        } else {
            sc->lower_cap = 2.0f;
            sc->upper_cap = 256.0f;
        }
        if (sc->cooldown > 0u) --sc->cooldown;
    }
    sc->frame_idx += 1u;
    sc->epoch += 1u;
}

// LOBSTER end of step: frame counter and worklist epoch
__global__ void k_lobster_end(View v) {
    v.sc->frame_idx += 1u;
    v.sc->epoch += 1u;
}

// refreshModel (SUB:80-105; LOB:410-440). pending = 1: auto-reset request from k_decide (also
// T := 1 everywhere, SUB:592); otherwise an explicit call with the given parameters.
template <int C, int ALGO>
__global__ void __launch_bounds__(TPB) k_refresh(View v, int use_pending, float frac, int force_fg, uint32_t frame,
                                               uint32_t ev) {
    if (use_pending) {
        if (!v.sc->refresh_pending) return;
        frac = v.sc->refresh_frac;
        force_fg = v.sc->refresh_force_fg;
        frame = v.sc->refresh_frame;
        ev = v.sc->refresh_ev;
    }
    const uint32_t p = blockIdx.x * TPB + threadIdx.x;
    if (p >= v.npx) return;
    if (use_pending && ALGO == kSuBSENSE) v.T[p] = 1.0f;
    if (!v.roi[p]) return;
    if (!(force_fg || !v.last_fg[p])) return;
    const uint32_t N = (uint32_t)v.N;
    const uint32_t nref = frac < 1.0f ? (uint32_t)(frac * (float)N) : N;
    uint32_t start = 0;
    if (frac < 1.0f) {
        uint32_t d[4];
        draw4(v.seed_lo, v.seed_hi, v.stream_id, frame, PH_REFRESH, ev, 0xFFFFFFFFu, 0, d);
        start = d[0] % N;
    }
    const int x = (int)(p % (uint32_t)v.W), y = (int)(p / (uint32_t)v.W);
    uint32_t d[4];
    for (uint32_t j = 0; j < nref; ++j) {
        if ((j & 3u) == 0) draw4(v.seed_lo, v.seed_hi, v.stream_id, frame, PH_REFRESH, ev, p, j >> 2, d);
        int sx, sy;
        sample_pos_7x7(d[j & 3u], x, y, v.W, v.H, sx, sy);
        const uint32_t sp = (uint32_t)(sy * v.W + sx);
        if (!(force_fg || !v.last_fg[sp])) continue;
        uint32_t col[C], desc[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            col[c] = v.last_color[(size_t)sp * C + c];
            if constexpr (ALGO == kLOBSTER) {
                uint32_t taps[4];
                taps_from_global<C>(v.last_color, v.W, sx, sy, c, taps);
                desc[c] = lbsp_threshold_packed(taps, col[c], v.sc->lut[col[c]]);
            } else {
                desc[c] = v.last_desc[(size_t)sp * C + c];
            }
        }
        store_sample<C>(v, (start + j) % N, p, col, desc);
    }
}

// initial descriptors of the LBSP base (LBS:31-63): strict bounds x>2 && y>2 && x<W-2 && y<H-2
template <int C>
__global__ void __launch_bounds__(TPB) k_init_desc(View v, const uint8_t* __restrict__ img) {
    const uint32_t p = blockIdx.x * TPB + threadIdx.x;
    if (p >= v.npx) return;
    const int x = (int)(p % (uint32_t)v.W), y = (int)(p / (uint32_t)v.W);
    const bool in = v.roi[p] && x > 2 && y > 2 && x < v.W - 2 && y < v.H - 2;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        uint16_t d = 0;
        if (in) {
            uint32_t taps[4];
            taps_from_global<C>(img, v.W, x, y, c, taps);
            const uint32_t ref = img[(size_t)p * C + c];
            d = (uint16_t)lbsp_threshold_packed(taps, ref, v.sc->lut[ref]);
        }
        v.last_desc[(size_t)p * C + c] = d;
        v.last_color[(size_t)p * C + c] = v.roi[p] ? img[(size_t)p * C + c] : (uint8_t)0;
    }
}

// getBackgroundImage / getBackgroundDescriptorsImage (SUB:614-649): float sum of v/N in sample
// order, then saturate_cast (round half to even)
template <int C>
__global__ void __launch_bounds__(TPB) k_background(View v, uint8_t* __restrict__ out_img, uint16_t* __restrict__ out_desc) {
    const uint32_t p = blockIdx.x * TPB + threadIdx.x;
    if (p >= v.npx) return;
    float ac[C], ad[C];
#pragma unroll
    for (int c = 0; c < C; ++c) ac[c] = ad[c] = 0.0f;
    const float fN = (float)v.N;
    for (int s = 0; s < v.N; ++s) {
        uint32_t col[C], desc[C];
        load_sample<C>(v, (uint32_t)s, p, col, desc);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            ac[c] += (float)col[c] / fN;
            ad[c] += (float)desc[c] / fN;
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (out_img) out_img[(size_t)p * C + c] = sat_u8(ac[c]);
        if (out_desc) {
            const float r = rintf(ad[c]);
            out_desc[(size_t)p * C + c] = (uint16_t)(r < 0.f ? 0.f : r > 65535.f ? 65535.f : r);
        }
    }
}

// stand-alone LBSP map (LBSP::compute2, LBSP.cpp:102-123): absolute threshold t
template <int C>
__global__ void __launch_bounds__(TPB) k_lbsp_abs(const uint8_t* __restrict__ img, int W, int H, uint32_t t,
                                                uint16_t* __restrict__ out) {
    __shared__ FrameTile<C> tile;
    const int ntx = (W + TW - 1) / TW, nty = (H + TH - 1) / TH;
    const int tid = xcd_tile((int)blockIdx.x, ntx * nty);
    const int x0 = (tid % ntx) * TW, y0 = (tid / ntx) * TH;
    load_frame_tile<C>(tile, img, W, H, x0, y0);
    __syncthreads();
    const int lx = threadIdx.x % TW, ly = threadIdx.x / TW, x = x0 + lx, y = y0 + ly;
    if (x >= W || y >= H) return;
    const bool in = x >= 2 && y >= 2 && x < W - 2 && y < H - 2;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        uint16_t d = 0;
        if (in) {
            uint32_t taps[4];
            taps_from_tile<C>(tile, lx, ly, c, taps);
            d = (uint16_t)lbsp_threshold_packed(taps, tile.px[((ly + 2) * FrameTile<C>::LW + (lx + 2)) * C + c], t);
        }
        out[((size_t)y * W + x) * C + c] = d;
    }
}

// RNG self-check: draws of the apply stream for pixels [0, n)
__global__ void k_rng_probe(View v, uint32_t frame, uint32_t phase, uint32_t ev, uint32_t n, uint32_t* out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t d[4];
    draw4(v.seed_lo, v.seed_hi, v.stream_id, frame, phase, ev, p, 0, d);
    for (int i = 0; i < 4; ++i) out[(size_t)p * 4 + i] = d[i];
}

// ==============================================================================================
// host-side launchers
// ==============================================================================================
static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

template <int C, int ALGO>
static void frame_step_t(const View& v, const uint8_t* frame, uint8_t* out_mask, double lr, const StepOptions& opt,
                         hipStream_t st) {
    const unsigned tiles = cdiv(v.W, TW) * cdiv(v.H, TH);
    const unsigned ptiles = cdiv(v.W, PW) * cdiv(v.H, PH);
    const unsigned pix_blocks = cdiv(v.npx, TPB);
    const unsigned inner_blocks = cdiv((size_t)(v.W - 4) * (v.H - 4), TPB);
    if (opt.timed_stage == 2) (void)hipEventRecord(opt.ev_begin, st);
    hipLaunchKernelGGL((k_describe<C, ALGO>), dim3(tiles), dim3(TPB), 0, st, v, frame, lr);
    if (opt.timed_stage == 1) (void)hipEventRecord(opt.ev_begin, st);
    hipLaunchKernelGGL((k_match<C, ALGO>), dim3(tiles), dim3(TPB), 0, st, v, frame);
    if (opt.timed_stage == 1) (void)hipEventRecord(opt.ev_end, st);
    for (uint32_t k = 1; k <= (uint32_t)opt.fixpoint_launches; ++k)
        hipLaunchKernelGGL((k_fixpoint<C, ALGO>), dim3(opt.fixpoint_blocks), dim3(TPB), 0, st, v, frame, k);
    hipLaunchKernelGGL((k_fixpoint_tail<C, ALGO>), dim3(1), dim3(1024), 0, st, v, frame,
                       (uint32_t)opt.fixpoint_launches + 1u, (uint32_t)opt.tail_cap);
    if constexpr (ALGO == kSuBSENSE) {
        hipLaunchKernelGGL((k_feedback<C>), dim3(pix_blocks), dim3(TPB), 0, st, v, frame);
        hipLaunchKernelGGL((k_update<C, ALGO>), dim3(inner_blocks), dim3(TPB), 0, st, v, frame);
        hipLaunchKernelGGL(k_close, dim3(ptiles), dim3(TPB), 0, st, v);
        const unsigned ct = cdiv(v.W, CT) * cdiv(v.H, CT);
        hipLaunchKernelGGL(k_ccl_local, dim3(ct), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL(k_ccl_border, dim3(cdiv((size_t)ct * 2 * CT, TPB)), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL(k_holes, dim3(pix_blocks), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL((k_median<true>), dim3(ptiles), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL(k_dilate_final, dim3(ptiles), dim3(TPB), 0, st, v, out_mask);
        if (v.lr_scaling) hipLaunchKernelGGL((k_downsample<C>), dim3(cdiv((size_t)v.dsW * v.dsH, TPB)), dim3(TPB), 0, st, v, frame);
        hipLaunchKernelGGL(k_decide, dim3(1), dim3(256), 0, st, v);
        hipLaunchKernelGGL((k_refresh<C, ALGO>), dim3(pix_blocks), dim3(TPB), 0, st, v, 1, 0.f, 0, 0u, 0u);
    } else {
        hipLaunchKernelGGL(k_lobster_raw, dim3(pix_blocks), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL((k_update<C, ALGO>), dim3(inner_blocks), dim3(TPB), 0, st, v, frame);
        hipLaunchKernelGGL((k_median<false>), dim3(ptiles), dim3(TPB), 0, st, v);
        if (out_mask) hipLaunchKernelGGL(k_copy_mask, dim3(pix_blocks), dim3(TPB), 0, st, v, out_mask);
        hipLaunchKernelGGL((k_copy_last_color<C>), dim3(cdiv((size_t)v.npx * C, TPB)), dim3(TPB), 0, st, v, frame);
        hipLaunchKernelGGL(k_lobster_end, dim3(1), dim3(1), 0, st, v);
    }
    if (opt.timed_stage == 2) (void)hipEventRecord(opt.ev_end, st);
}

void frame_step(const View& v, const uint8_t* frame, uint8_t* out_mask, double lr, const StepOptions& opt,
                hipStream_t st) {
    if (v.algo == kSuBSENSE) {
        if (v.C == 3) frame_step_t<3, kSuBSENSE>(v, frame, out_mask, lr, opt, st);
        else frame_step_t<1, kSuBSENSE>(v, frame, out_mask, lr, opt, st);
    } else {
        if (v.C == 3) frame_step_t<3, kLOBSTER>(v, frame, out_mask, lr, opt, st);
        else frame_step_t<1, kLOBSTER>(v, frame, out_mask, lr, opt, st);
    }
}

void launch_refresh(const View& v, float frac, int force_fg, uint32_t frame, uint32_t ev, hipStream_t st) {
    const unsigned b = cdiv(v.npx, TPB);
    if (v.algo == kSuBSENSE) {
        if (v.C == 3) hipLaunchKernelGGL((k_refresh<3, kSuBSENSE>), dim3(b), dim3(TPB), 0, st, v, 0, frac, force_fg, frame, ev);
        else hipLaunchKernelGGL((k_refresh<1, kSuBSENSE>), dim3(b), dim3(TPB), 0, st, v, 0, frac, force_fg, frame, ev);
    } else {
        if (v.C == 3) hipLaunchKernelGGL((k_refresh<3, kLOBSTER>), dim3(b), dim3(TPB), 0, st, v, 0, frac, force_fg, frame, ev);
        else hipLaunchKernelGGL((k_refresh<1, kLOBSTER>), dim3(b), dim3(TPB), 0, st, v, 0, frac, force_fg, frame, ev);
    }
}

void launch_init_desc(const View& v, const uint8_t* img, hipStream_t st) {
    const unsigned b = cdiv(v.npx, TPB);
    if (v.C == 3) hipLaunchKernelGGL(k_init_desc<3>, dim3(b), dim3(TPB), 0, st, v, img);
    else hipLaunchKernelGGL(k_init_desc<1>, dim3(b), dim3(TPB), 0, st, v, img);
}

void launch_background(const View& v, uint8_t* out_img, uint16_t* out_desc, hipStream_t st) {
    const unsigned b = cdiv(v.npx, TPB);
    if (v.C == 3) hipLaunchKernelGGL(k_background<3>, dim3(b), dim3(TPB), 0, st, v, out_img, out_desc);
    else hipLaunchKernelGGL(k_background<1>, dim3(b), dim3(TPB), 0, st, v, out_img, out_desc);
}

void launch_lbsp_abs(const uint8_t* img, int W, int H, int C, uint32_t t, uint16_t* out, hipStream_t st) {
    const unsigned tiles = cdiv(W, TW) * cdiv(H, TH);
    if (C == 3) hipLaunchKernelGGL(k_lbsp_abs<3>, dim3(tiles), dim3(TPB), 0, st, img, W, H, t, out);
    else hipLaunchKernelGGL(k_lbsp_abs<1>, dim3(tiles), dim3(TPB), 0, st, img, W, H, t, out);
}

void launch_rng_probe(const View& v, uint32_t frame, uint32_t phase, uint32_t ev, uint32_t n, uint32_t* out,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_rng_probe, dim3(cdiv(n, 256)), dim3(256), 0, st, v, frame, phase, ev, n, out);
}

}  // namespace lvbgs
