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

#include <algorithm>

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
// Every thread of the workgroup must call it. The closing barrier keeps the wavefront partials in
// LDS until thread 0 has summed them, so a second call right after (the model update adds two
// counters) cannot overwrite them early.
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
    __syncthreads();
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
// one sample record: 12 bytes (RGB) or 4 bytes (grey), a single load or store per sample
struct alignas(4) Rec12 {
    uint32_t w0, w1, w2;
};

// Tile-major sample layout: the N samples of each group of 64 consecutive pixels are contiguous
// (group g, sample s, lane p % 64), so a wavefront's ordered scan walks one compact region instead
// of N planes a frame apart.
LV_D size_t sample_index(const View& v, uint32_t s, uint32_t p) {
    return ((size_t)(p >> 6) * (uint32_t)v.N + s) * 64u + (p & 63u);
}

// A device pointer seen through the global address space. The pointers of a batch member's View are
// read from a table in device memory, so the compiler sees them as generic (flat) pointers; loads
// through address space 1 are global loads instead (an SGPR base with a 32-bit lane offset, and
// LDS waits that do not also wait for them).
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
LV_D gptr<T> G(T* p) {
    return (gptr<T>)p;
}

// wavefront-local LDS exchange: the lanes' stores above are visible to the loads below
LV_D void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// raw sample in registers: ra = record words 0-1 (RGB) or word 0 (grey), rb = word 2 (RGB)
template <int C>
LV_D void fetch_raw(const View& v, uint32_t s, uint32_t p, uint64_t& ra, uint32_t& rb) {
    const size_t idx = sample_index(v, s, p);
    if constexpr (C == 3) {
        const gptr<const uint32_t> q = G((const uint32_t*)v.sA) + idx * 3u;
        ra = (uint64_t)q[0] | ((uint64_t)q[1] << 32);
        rb = q[2];
    } else {
        ra = G((const uint32_t*)v.sA)[idx];
        rb = 0;
    }
}

// records of one pixel: sample s at rec_base(v, p) + s * 64 (tile-major layout)
LV_D const uint32_t* rec_base(const View& v, uint32_t p) {
    return (const uint32_t*)v.sA + sample_index(v, 0u, p) * (v.C == 3 ? 3u : 1u);
}

template <int C>
LV_D void fetch_at(const uint32_t* rec, uint64_t& ra, uint32_t& rb) {
    const gptr<const uint32_t> q = G(rec);
    if constexpr (C == 3) {
        ra = (uint64_t)q[0] | ((uint64_t)q[1] << 32);
        rb = q[2];
    } else {
        ra = q[0];
        rb = 0;
    }
}

template <int C>
LV_D void decode_raw(uint64_t ra, uint32_t rb, uint32_t col[C], uint32_t desc[C]) {
    if constexpr (C == 3) {
        desc[0] = (uint32_t)(ra & 0xFFFFu);
        desc[1] = (uint32_t)((ra >> 16) & 0xFFFFu);
        desc[2] = (uint32_t)((ra >> 32) & 0xFFFFu);
        col[0] = (uint32_t)((ra >> 48) & 0xFFu);
        col[1] = (uint32_t)(ra >> 56);
        col[2] = rb & 0xFFu;
    } else {
        desc[0] = (uint32_t)(ra & 0xFFFFu);
        col[0] = (uint32_t)((ra >> 16) & 0xFFu);
    }
}

template <int C>
LV_D void encode_raw(const uint32_t col[C], const uint32_t desc[C], uint64_t& ra, uint32_t& rb) {
    if constexpr (C == 3) {
        ra = (uint64_t)desc[0] | ((uint64_t)desc[1] << 16) | ((uint64_t)desc[2] << 32) | ((uint64_t)col[0] << 48) |
             ((uint64_t)col[1] << 56);
        rb = col[2];
    } else {
        ra = (uint64_t)desc[0] | ((uint64_t)col[0] << 16);
        rb = 0;
    }
}

template <int C>
LV_D void load_sample(const View& v, uint32_t s, uint32_t p, uint32_t col[C], uint32_t desc[C]) {
    uint64_t ra;
    uint32_t rb;
    fetch_raw<C>(v, s, p, ra, rb);
    decode_raw<C>(ra, rb, col, desc);
}

template <int C>
LV_D void store_sample(const View& v, uint32_t s, uint32_t p, const uint32_t col[C], const uint32_t desc[C]) {
    const size_t idx = sample_index(v, s, p);
    uint64_t ra;
    uint32_t rb;
    encode_raw<C>(col, desc, ra, rb);
    // non-temporal: scattered single-record writes gain nothing from being kept in L2
    typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));
    if constexpr (C == 3) {
        u32x3 w;
        w.x = (uint32_t)ra;
        w.y = (uint32_t)(ra >> 32);
        w.z = rb;
        __builtin_nontemporal_store(w, (gptr<u32x3>)((Rec12*)v.sA + idx));
    } else {
        __builtin_nontemporal_store((uint32_t)ra, G((uint32_t*)v.sA + idx));
    }
}

// current colour and intra descriptor of pixel q (the value it writes into samples)
template <int C>
LV_D void current_value(const View& v, const uint8_t* frame, uint32_t q, uint32_t col[C], uint32_t desc[C]) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
        col[c] = G(frame)[(size_t)q * C + c];
        desc[c] = G(v.desc_cur)[(size_t)q * C + c];
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
        for (int c = 0; c < C; ++c) t.px[i * C + c] = in ? G(frame)[((size_t)gy * W + gx) * C + c] : (uint8_t)0;
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
            acc |= (uint32_t)G(frame)[((size_t)(y + kLbspDY[n]) * W + (x + kLbspDX[n])) * C + c] << (8 * k);
        }
        taps[w] = acc;
    }
}

LV_D void load_lut(uint8_t* lut_lds, const Scalars* sc) {
    const gptr<const uint8_t> src = G(sc->lut);
    for (int i = threadIdx.x; i < 256; i += blockDim.x) lut_lds[i] = src[i];
}

// ----------------------------------------------------------------------------------------------
// per-pixel match context
// ----------------------------------------------------------------------------------------------
template <int C>
struct PxCtx {
    uint32_t cur[C], intra[C];
    uint32_t taps[C][4];                   // packed taps (byte k of word w = tap 4w+k)
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
        px.t_sc = px.t_tot_col = px.t_tot_desc = 0u;  // unused by the 1-channel test
    } else {
        const uint64_t tot_col = col * 3u, tot_desc = desc * 3u;
        px.t_sc = clamp_u32(tot_col / 2u);
        px.t_tot_col = clamp_u32(tot_col);
        px.t_tot_desc = clamp_u32(tot_desc);
        px.t_col = px.t_desc = 0u;  // unused by the 3-channel test
    }
}

// LOBSTER fixed thresholds (LOB:485-488 1ch; :515-518 3ch)
template <int C>
LV_D void lobster_thresholds(const View& v, PxCtx<C>& px) {
    if constexpr (C == 1) {
        px.t_col = v.color_thr / 2u;
        px.t_desc = v.desc_thr;
        px.t_sc = px.t_tot_col = px.t_tot_desc = 0u;  // unused by the 1-channel test
    } else {
        px.t_tot_desc = v.desc_thr * 3u;
        px.t_tot_col = v.color_thr * 3u;
        px.t_sc = px.t_tot_col / 2u;   // colour, per channel
        px.t_desc = px.t_tot_desc / 2u;  // descriptor, per channel
    }
}

// one sample test; returns true on a match and accumulates the SuBSENSE distances
template <int C, int ALGO, int RS>
LV_D bool test_sample(const PxCtx<C>& px, const uint8_t* lut, const uint32_t col[C], const uint32_t desc[C],
                      uint32_t& tot_desc, uint32_t& tot_sum) {
    if constexpr (ALGO == kSuBSENSE && C == 3) {
        // all colour tests first: a sample failing any of them fails whatever the descriptors say
        uint32_t cdv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cdv[c] = absdiff(px.cur[c], col[c]);
            if (cdv[c] > px.t_sc) return false;
        }
        uint32_t td = 0, ts = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint32_t cd = cdv[c];
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
        uint32_t cdv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cdv[c] = absdiff(px.cur[c], col[c]);
            if (cdv[c] > px.t_sc) return false;
        }
        uint32_t tc = 0, td = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint32_t cd = cdv[c];
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
        const uint32_t r = G(v.rec)[q];
        const uint32_t code = (uint32_t)((2 - kPredDY[i]) * 5 + (2 - kPredDX[i]));
        if (rec_tgt(r) == code && fwd_gate(r)) m |= 1u << i;
    }
    return m;
}

// Ordered scan s = 0..N-1 until REQ matches (SUB:366-395, LOB:480-495). Samples overwritten by
// applied writers (mask `applied`, slot bitmap `sbm`) are taken from the writer; the last writer
// in raster order wins. Two evaluation orders give the same result as the sequential loop:
//   scan_seq  : one lane per pixel, samples in order, KF loads in flight, up to `limit`, with
//               at most kRepl overwritten slots substituted from the thread's LDS entries;
//   scan_coop : kGroup lanes per pixel evaluate kGroup consecutive samples at once; a ballot finds
//               the REQ-th match in sample order, distances are reduced only up to it.
constexpr int kGroup = 8;
constexpr uint32_t kBig = 0xFFFFu;

struct ScanState {
    uint32_t good, s, md, ms;
};

template <int C>
LV_D ScanState scan_init() {
    return ScanState{0u, 0u, C * 16u, C * 255u};
}

LV_D uint64_t applied_slots(const View& v, int x, int y, uint32_t applied) {
    uint64_t sbm = 0;  // slots (mod 64) touched by applied writers
    for (uint32_t m = applied; m; m &= m - 1) {
        const int i = __builtin_ctz(m);
        const uint32_t q = (uint32_t)((y + kPredDY[i]) * v.W + (x + kPredDX[i]));
        sbm |= 1ull << (rec_slot_diff(G(v.rec)[q]) & 63u);
    }
    return sbm;
}

// decodes a fetched sample, replacing it by the last applied writer's value when one wrote slot s
template <int C>
LV_D void sample_value(const View& v, const uint8_t* frame, int x, int y, uint32_t s, uint32_t applied, uint64_t sbm,
                       uint64_t ra, uint32_t rb, uint32_t col[C], uint32_t desc[C]) {
    decode_raw<C>(ra, rb, col, desc);
    if ((sbm >> (s & 63u)) & 1u) {
        for (int i = 11; i >= 0; --i) {
            if (!((applied >> i) & 1u)) continue;
            const uint32_t q = (uint32_t)((y + kPredDY[i]) * v.W + (x + kPredDX[i]));
            if (rec_slot_diff(G(v.rec)[q]) == s) {
                current_value<C>(v, frame, q, col, desc);
                break;
            }
        }
    }
}

// Overwritten slots of a pixel with their replacement values (the last applied writer's current
// value), kept per thread in LDS so the sequential scan's loop carries only a slot bitmap: a sample
// whose slot bit is set looks its slot up in the thread's entries. At most kRepl distinct slots;
// pixels with more (very rare) go to the cooperative phase.
constexpr int kRepl = 4;
struct ReplLds {
    uint32_t slot[kRepl][TPB];
    uint32_t w0[kRepl][TPB], w1[kRepl][TPB], w2[kRepl][TPB];
};

// one LDS table per kernel, whichever channel counts the kernel instantiates
LV_D ReplLds& repl_lds() {
    __shared__ ReplLds r;
    return r;
}

template <int C>
LV_D void raw_current(const View& v, const uint8_t* frame, uint32_t q, uint64_t& ra, uint32_t& rb) {
    uint32_t col[C], desc[C];
    current_value<C>(v, frame, q, col, desc);
    encode_raw<C>(col, desc, ra, rb);
}

// Fills the thread's entries from the applied writers, the latest in raster order first (it owns
// its slot). Returns the slot bitmap (mod 64) and the entry count, or false beyond kRepl slots.
// rec_of(i) is the update record of predecessor i.
template <int C, typename RecOf>
LV_D bool build_repl(const View& v, const uint8_t* frame, int x, int y, uint32_t applied, uint64_t& sbm, int& n,
                     RecOf rec_of) {
    ReplLds& R = repl_lds();
    const int t = (int)threadIdx.x;
    sbm = 0;
    n = 0;
    for (uint32_t m = applied; m;) {
        const int i = 31 - __builtin_clz(m);
        m &= ~(1u << i);
        const uint32_t q = (uint32_t)((y + kPredDY[i]) * v.W + (x + kPredDX[i]));
        const uint32_t sl = rec_slot_diff(rec_of(i));
        bool seen = false;
#pragma unroll
        for (int k = 0; k < kRepl; ++k) seen = seen || (k < n && R.slot[k][t] == sl);
        if (seen) continue;
        if (n == kRepl) return false;
        uint64_t ra;
        uint32_t rb;
        raw_current<C>(v, frame, q, ra, rb);
        R.slot[n][t] = sl;
        R.w0[n][t] = (uint32_t)ra;
        R.w1[n][t] = (uint32_t)(ra >> 32);
        R.w2[n][t] = rb;
        sbm |= 1ull << (sl & 63u);
        ++n;
    }
    return true;
}

// One batch of KF loaded samples of a sequential scan (stops at the limit or at REQ matches).
template <int C, int ALGO, int RS, int KF>
LV_D void scan_batch(const uint8_t* lut, const PxCtx<C>& px, uint64_t sbm, int nrep, ScanState& st, uint32_t limit,
                     uint32_t REQ, const uint64_t (&ra)[KF], const uint32_t (&rb)[KF]) {
    const ReplLds& R = repl_lds();
    const int t = (int)threadIdx.x;
#pragma unroll
    for (int j = 0; j < KF; ++j) {
        if (st.s >= limit || st.good >= REQ) break;
        uint64_t a = ra[j];
        uint32_t b = rb[j];
        if ((sbm >> (st.s & 63u)) & 1u) {
            for (int k = 0; k < nrep; ++k)
                if (R.slot[k][t] == st.s) {
                    a = (uint64_t)R.w0[k][t] | ((uint64_t)R.w1[k][t] << 32);
                    b = R.w2[k][t];
                    break;
                }
        }
        uint32_t col[C], desc[C];
        decode_raw<C>(a, b, col, desc);
        uint32_t td = 0, ts = 0;
        if (test_sample<C, ALGO, RS>(px, lut, col, desc, td, ts)) {
            st.md = td < st.md ? td : st.md;
            st.ms = ts < st.ms ? ts : st.ms;
            ++st.good;
        }
        ++st.s;
    }
}

// speculative scan: a wavefront whose active lanes drop below kWaveCutLanes after at least
// kWaveCutMinSamples samples hands its remaining pixels to the continuation (measured on the
// 64-stream batch: phase A 2544 -> 2403 us, the whole step 8124 -> 8071 us)
constexpr int kWaveCutLanes = 16;
constexpr uint32_t kWaveCutMinSamples = 4;

// words between the records of two consecutive samples of a pixel
template <int C>
constexpr uint32_t kSampleStride = 64u * (C == 3 ? 3u : 1u);

// Sequential scan from st.s up to the limit. Unconditional loads (the sample planes are padded by
// kMaxFetch samples) at constant offsets from a running pointer keep all KF loads in flight together
// without address arithmetic.
template <int C, int ALGO, int RS, int KF>
LV_D void scan_seq(const View& v, const uint8_t* lut, const PxCtx<C>& px, int x, int y, uint64_t sbm, int nrep,
                   ScanState& st, uint32_t limit) {
    const uint32_t p = (uint32_t)(y * v.W + x);
    const uint32_t REQ = (uint32_t)v.REQ;
    constexpr uint32_t kStride = kSampleStride<C>;
    const uint32_t* cur = rec_base(v, p) + st.s * kStride;
    while (st.s < limit && st.good < REQ) {
        // speculative phase: once few lanes of the wavefront still scan, their pixels go on in the
        // dense continuation instead of holding the whole wavefront (the cut never changes a result)
        if constexpr (RS != 0)
            if (st.s >= kWaveCutMinSamples && __popcll(__ballot(1)) < kWaveCutLanes) break;
        uint64_t ra[KF];
        uint32_t rb[KF];
#pragma unroll
        for (int j = 0; j < KF; ++j) fetch_at<C>(cur + j * kStride, ra[j], rb[j]);
        cur += KF * kStride;
        scan_batch<C, ALGO, RS, KF>(lut, px, sbm, nrep, st, limit, REQ, ra, rb);
    }
}

// Called by every lane of the wavefront; lanes 8g..8g+7 share one pixel and one state (active and
// px/x/y/applied/sbm/st identical within the group). Returns with st final.
template <int C, int ALGO, int RS>
LV_D void scan_coop(const View& v, const uint8_t* frame, const uint8_t* lut, const PxCtx<C>& px, int x, int y,
                    uint32_t applied, uint64_t sbm, ScanState& st, bool active) {
    const uint32_t N = (uint32_t)v.N, REQ = (uint32_t)v.REQ;
    const int lane = (int)(threadIdx.x & 63u), sub = lane & (kGroup - 1), gbase = lane & ~(kGroup - 1);
    const uint32_t p = active ? (uint32_t)(y * v.W + x) : 0u;
    bool done = !active || st.good >= REQ || st.s >= N;
    while (__ballot(!done) != 0ull) {
        const uint32_t s = st.s + (uint32_t)sub;
        bool m = false;
        uint32_t td = kBig, ts = kBig;
        if (!done && s < N) {
            uint64_t ra;
            uint32_t rb;
            fetch_raw<C>(v, s, p, ra, rb);
            uint32_t col[C], desc[C];
            sample_value<C>(v, frame, x, y, s, applied, sbm, ra, rb, col, desc);
            m = test_sample<C, ALGO, RS>(px, lut, col, desc, td, ts);
        }
        const uint32_t bits = (uint32_t)(__ballot(m) >> gbase) & 0xFFu;
        uint32_t lim = kGroup;  // lanes whose sample precedes the stopping point
        bool stop = false;
        if (!done) {
            const uint32_t need = REQ - st.good;
            if ((uint32_t)__popc(bits) >= need) {
                uint32_t c = 0, pos = 0;
                for (uint32_t k = 0; k < (uint32_t)kGroup; ++k)
                    if ((bits >> k) & 1u) {
                        if (++c == need) {
                            pos = k;
                            break;
                        }
                    }
                lim = pos + 1u;
                stop = true;
            }
        }
        uint32_t tdv = (m && (uint32_t)sub < lim) ? td : kBig, tsv = (m && (uint32_t)sub < lim) ? ts : kBig;
#pragma unroll
        for (int o = 1; o < kGroup; o <<= 1) {
            const uint32_t a = (uint32_t)__shfl_xor((int)tdv, o, 64), b = (uint32_t)__shfl_xor((int)tsv, o, 64);
            tdv = a < tdv ? a : tdv;
            tsv = b < tsv ? b : tsv;
        }
        if (!done) {
            st.md = tdv < st.md ? tdv : st.md;
            st.ms = tsv < st.ms ? tsv : st.ms;
            if (stop) {
                st.good = REQ;
                st.s += lim;
                done = true;
            } else {
                st.good += (uint32_t)__popc(bits);
                st.s = st.s + kGroup < N ? st.s + kGroup : N;
                done = st.s >= N;
            }
        }
    }
}

LV_D uint32_t pack_state(const ScanState& st) { return mres_pack(st.good, st.md, st.ms, st.s); }

LV_D ScanState unpack_state(uint32_t mr) {
    return ScanState{mr & 255u, mr >> 24, (mr >> 8) & 63u, (mr >> 14) & 1023u};
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
    return G(v.roi)[t] != 0;
}

// ----------------------------------------------------------------------------------------------
// describe: LBSP intra descriptors, unstable flag, RNG draws, update record (SUB:355-365,
// :406-449 / LOB:499-512). Writes desc_cur, rec; counts non-zero descriptors.
// ----------------------------------------------------------------------------------------------
template <int C, int ALGO>
LV_D void describe_body(const View& v, uint32_t bid, uint32_t nblk) {
    const uint8_t* frame = v.frame;
    const double lr = v.lr;
    __shared__ FrameTile<C> tile;
    __shared__ uint8_t lut[256];
    const int ntx = (v.W + TW - 1) / TW;
    const int tid = (int)bid;
    const int x0 = (tid % ntx) * TW, y0 = (tid / ntx) * TH;
    if (bid == 0 && threadIdx.x == 0) {
        // per-frame counters consumed by later kernels of this step
        for (int i = 0; i < kWlRing; ++i) v.sc->wl_count[i] = 0;
        v.sc->cont2_count = 0;
        for (int i = 0; i < 8; ++i) v.sc->wl_work[i] = 0;
        v.sc->iters = 0;
    }
    if (bid == 0 && threadIdx.x < kContStripes) v.sc->cont_cnt[threadIdx.x * kContStride] = 0;
    if (bid == 0 && threadIdx.x < kStripes) {
        v.sc->ex_part[threadIdx.x] = 0;
        v.sc->wr_part[threadIdx.x] = 0;
        v.sc->spec_part[threadIdx.x] = 0;
        v.sc->coop_part[threadIdx.x] = 0;
    }
    const int lx = threadIdx.x % TW, ly = threadIdx.x / TW, x = x0 + lx, y = y0 + ly;
    const bool inside = x < v.W && y < v.H;
    // the pixel's own state is loaded first, so these loads run alongside the tile's
    const uint32_t pp = inside ? (uint32_t)(y * v.W + x) : 0u;
    const uint8_t roi_p = G(v.roi)[pp];
    float Rv = 0.f, Tv = 0.f, d_lt = 0.f, d_st = 0.f;
    if constexpr (ALGO == kSuBSENSE) {
        Rv = G(v.R)[pp];
        Tv = G(v.T)[pp];
        d_lt = G(v.RawLT)[pp] - G(v.FinLT)[pp];
        d_st = G(v.RawST)[pp] - G(v.FinST)[pp];
    }
    load_frame_tile<C>(tile, frame, v.W, v.H, x0, y0);
    load_lut(lut, v.sc);
    __syncthreads();
    bool nz = false;
    if (inside) {
        const uint32_t p = pp;
        if (!roi_p) {
            G(v.rec)[p] = kRecNone;
        } else {
            uint32_t pc = 0;
            uint32_t intra[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                uint32_t taps[4];
                taps_from_tile<C>(tile, lx, ly, c, taps);
                const uint32_t ref = tile.px[((ly + 2) * FrameTile<C>::LW + (lx + 2)) * C + c];
                intra[c] = lbsp_threshold_sad(taps, ref, lut[ref]);
                G(v.desc_cur)[(size_t)p * C + c] = (uint16_t)intra[c];
                pc += popc(intra[c]);
            }
            nz = pc >= (C == 1 ? 2u : 4u);
            const uint32_t frame_no = cur_frame(v);
            uint32_t d[4];
            draw4(v.seed_lo, v.seed_hi, v.stream_id, frame_no, PH_APPLY, 0, p, 0, d);
            const uint32_t N = (uint32_t)v.N;
            // draw 4 (group 1) is the diffusion slot: drawn only when a diffusion gate is open
            auto diff_slot = [&]() {
                uint32_t d4[4];
                draw4(v.seed_lo, v.seed_hi, v.stream_id, frame_no, PH_APPLY, 0, p, 1, d4);
                return mod64(d4[0], N);
            };
            uint32_t r;
            if constexpr (ALGO == kSuBSENSE) {
                const uint32_t unst_new = (Rv > 3.0f || d_lt > 0.1f || d_st > 0.1f) ? 1u : 0u;
                const uint64_t L = isinf(lr) ? ~0ull : (lr > 0 ? (uint64_t)ceil(lr) : (uint64_t)ceilf(Tv));
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
                    ghost_fwd = G(v.RawST)[t] > 0.995f && G(v.MLD)[t] < 0.010f;
                }
                const uint32_t tg = (uint32_t)((ty - y + 2) * 5 + (tx - x + 2));
                const uint32_t sd = (gate_a || gate_b) ? diff_slot() : 0u;
                r = tg | ((uint32_t)gate_a << 5) | ((uint32_t)gate_b << 6) | ((uint32_t)ghost_fwd << 7) |
                    ((uint32_t)self_bg << 8) | ((uint32_t)self_fg << 9) | (unst_new << 10) | (mod64(d[1], N) << 16) |
                    (sd << 24);
            } else {
                const uint64_t L = isinf(lr) ? ~0ull : (uint64_t)ceil(lr);
                const bool self_bg = mod64(d[0], L) == 0u;
                const bool gate_a = mod64(d[2], L) == 0u;
                int tx, ty;
                neighbor_pos(d[3], true, x, y, v.W, v.H, tx, ty);
                const uint32_t tg = (uint32_t)((ty - y + 2) * 5 + (tx - x + 2));
                const uint32_t sd = gate_a ? diff_slot() : 0u;
                r = tg | ((uint32_t)gate_a << 5) | ((uint32_t)self_bg << 8) | (mod64(d[1], N) << 16) | (sd << 24);
            }
            G(v.rec)[p] = r;
        }
    }
    if constexpr (ALGO == kSuBSENSE) block_add<TPB>(nz ? 1ull : 0ull, v.sc->nz_part);
}

// builds the match context of ROI pixel (x,y); taps from the LDS tile or from global memory
template <int C, int ALGO, int RS>
LV_D void build_ctx(const View& v, const uint8_t* frame, const uint8_t* win, int x, int y, PxCtx<C>& px) {
    const uint32_t p = (uint32_t)(y * v.W + x);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        px.cur[c] = G(frame)[(size_t)p * C + c];
        px.intra[c] = G(v.desc_cur)[(size_t)p * C + c];
    }
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int n = 4 * w + k;
                acc |= (uint32_t)win[(kLbspDY[n] + 2) * RS + (kLbspDX[n] + 2) * C + c] << (8 * k);
            }
            px.taps[c][w] = acc;
        }
    if constexpr (ALGO == kSuBSENSE) subsense_thresholds<C>(v, G(v.R)[p], G(v.unstable)[p], px);
    else lobster_thresholds<C>(v, px);
}

// a tile pixel's per-frame match inputs, loaded together with the tile and staged in LDS (the
// regrouping of a tile's pixels moves them between threads after the tile barrier)
template <int C>
struct PxIn {
    uint16_t desc[C];  // intra descriptors of the current frame
    uint8_t roi, unst;
    float R;
};

template <int C, int ALGO>
LV_D PxIn<C> load_px_in(const View& v, uint32_t q, bool in) {
    PxIn<C> r;
    r.roi = in ? G(v.roi)[q] : (uint8_t)0;
#pragma unroll
    for (int c = 0; c < C; ++c) r.desc[c] = G(v.desc_cur)[(size_t)q * C + c];
    if constexpr (ALGO == kSuBSENSE) {
        r.R = G(v.R)[q];
        r.unst = G(v.unstable)[q];
    } else {
        r.R = 0.f;
        r.unst = 0;
    }
    return r;
}

// the match context from the LDS frame tile (5x5 window at win; the tile is complete behind a
// workgroup barrier, so plain loads) and the staged inputs
template <int C, int ALGO, int RS>
LV_D void build_ctx_staged(const View& v, const uint8_t* win, const PxIn<C>& in, PxCtx<C>& px) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
        px.cur[c] = win[2 * RS + 2 * C + c];
        px.intra[c] = in.desc[c];
    }
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int n = 4 * w + k;
                acc |= (uint32_t)win[(kLbspDY[n] + 2) * RS + (kLbspDX[n] + 2) * C + c] << (8 * k);
            }
            px.taps[c][w] = acc;
        }
    if constexpr (ALGO == kSuBSENSE) subsense_thresholds<C>(v, in.R, in.unst, px);
    else lobster_thresholds<C>(v, px);
}

// A match context in kCtxWords words (16-byte aligned): taps (12 words), packed centre colours and
// intra descriptors (3), thresholds (5). Written by the speculative match for the pixels it leaves
// to the continuation, which then reads it instead of rebuilding it from the frame.
template <int C>
LV_D void store_ctx(uint32_t* d, const PxCtx<C>& px) {
    uint32_t w[kCtxWords] = {};
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k) w[c * 4 + k] = px.taps[c][k];
    uint32_t cur = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) cur |= px.cur[c] << (8 * c);
    w[12] = cur;
    w[13] = px.intra[0] | (C == 3 ? px.intra[1 % C] << 16 : 0u);
    w[14] = C == 3 ? px.intra[2 % C] : 0u;
    w[15] = px.t_sc;
    w[16] = px.t_tot_col;
    w[17] = px.t_tot_desc;
    w[18] = px.t_col;
    w[19] = px.t_desc;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const gptr<u32x4> q = (gptr<u32x4>)d;
#pragma unroll
    for (int k = 0; k < kCtxWords / 4; ++k) q[k] = u32x4{w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]};
}

template <int C>
LV_D void load_ctx(const uint32_t* d, PxCtx<C>& px) {
    uint32_t w[kCtxWords];
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const gptr<const u32x4> q = (gptr<const u32x4>)d;
#pragma unroll
    for (int k = 0; k < kCtxWords / 4; ++k) {
        const u32x4 u = q[k];
        w[4 * k] = u.x;
        w[4 * k + 1] = u.y;
        w[4 * k + 2] = u.z;
        w[4 * k + 3] = u.w;
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
#pragma unroll
        for (int k = 0; k < 4; ++k) px.taps[c][k] = w[c * 4 + k];
        px.cur[c] = (w[12] >> (8 * c)) & 255u;
    }
    px.intra[0] = w[13] & 0xFFFFu;
    if constexpr (C == 3) {
        px.intra[1 % C] = w[13] >> 16;
        px.intra[2 % C] = w[14];
    }
    px.t_sc = w[15];
    px.t_tot_col = w[16];
    px.t_tot_desc = w[17];
    px.t_col = w[18];
    px.t_desc = w[19];
}

// the same context with the taps read from the frame in global memory
template <int C, int ALGO>
LV_D void build_ctx_global(const View& v, const uint8_t* frame, int x, int y, PxCtx<C>& px) {
    const uint32_t p = (uint32_t)(y * v.W + x);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        px.cur[c] = G(frame)[(size_t)p * C + c];
        px.intra[c] = G(v.desc_cur)[(size_t)p * C + c];
    }
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int n = 4 * w + k;
                acc |= (uint32_t)G(frame)[((size_t)(y + kLbspDY[n]) * v.W + (x + kLbspDX[n])) * C + c] << (8 * k);
            }
            px.taps[c][w] = acc;
        }
    if constexpr (ALGO == kSuBSENSE) subsense_thresholds<C>(v, G(v.R)[p], G(v.unstable)[p], px);
    else lobster_thresholds<C>(v, px);
}

// 5x5 patch of one pixel staged in LDS by the kGroup lanes that share it (row stride 5*C)
template <int C>
LV_D void stage_patch(const uint8_t* frame, int W, int x, int y, uint8_t* patch, int sub) {
    for (int j = sub; j < 25 * C; j += kGroup) {
        const int r = j / (5 * C), rem = j % (5 * C);
        patch[j] = G(frame)[((size_t)(y - 2 + r) * W + (x - 2)) * C + rem];
    }
}

// ----------------------------------------------------------------------------------------------
// speculative match: raster-earlier writers assumed to decide as in the previous frame
// ----------------------------------------------------------------------------------------------
struct PredTile;
template <int C, int ALGO, int KF>
LV_D bool match_pixel(const View& v, const uint8_t* frame, const uint8_t* lut, const FrameTile<C>& tile,
                      const PredTile& pred, const PxIn<C>& in, int lx, int ly, int x, int y, uint32_t p,
                      uint32_t seq_limit, uint32_t& examined, PxCtx<C>& px);

// records and previous raw decisions of the raster-earlier part of a tile's 5x5 neighbourhood:
// rows y0-2 .. y0+TH-1, columns x0-2 .. x0+TW+1
struct PredTile {
    static constexpr int LW = TW + 4, LH = TH + 2;
    uint32_t rec[LH][LW];
    uint8_t guess[LH][LW];
};

LV_D void load_pred_tile(PredTile& t, const View& v, int x0, int y0) {
    for (int i = threadIdx.x; i < PredTile::LW * PredTile::LH; i += TPB) {
        const int ly = i / PredTile::LW, lx = i % PredTile::LW, gx = x0 + lx - 2, gy = y0 + ly - 2;
        const bool in = gx >= 0 && gx < v.W && gy >= 0 && gy < v.H;
        const size_t q = (size_t)gy * v.W + gx;
        t.rec[ly][lx] = in ? G(v.rec)[q] : kRecNone;
        t.guess[ly][lx] = in ? (G(v.last_raw)[q] ? 1 : 0) : 1;
    }
}

// candidate writers of tile-local pixel (lx,ly) whose guessed decision is background
LV_D uint32_t applied_writers_tile(const PredTile& t, int lx, int ly) {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        const int ty = ly + 2 + kPredDY[i], tx = lx + 2 + kPredDX[i];
        const uint32_t r = t.rec[ty][tx];
        const uint32_t code = (uint32_t)((2 - kPredDY[i]) * 5 + (2 - kPredDX[i]));
        if (rec_tgt(r) == code && fwd_gate(r) && t.guess[ty][tx] == 0) m |= 1u << i;
    }
    return m;
}

// LDS of the phase-A match for frames of up to CMAX channels. A batch kernel instantiates the match
// for both channel counts and calls both with CMAX = 3, so they share one allocation (separate
// arrays per instantiation would cost ~5 KB of LDS and an occupancy level).
template <int CMAX>
struct MatchLds {
    union {
        FrameTile<CMAX> big;
        FrameTile<1> c1;
    } tile;
    PredTile pred;
    uint8_t lut[256];
    uint16_t perm[TPB];
    uint32_t nfg[TPB / 64];
    union {
        PxIn<CMAX> big[TPB];
        PxIn<1> c1[TPB];
    } in;
};

template <int CMAX>
LV_D MatchLds<CMAX>& match_lds() {
    __shared__ MatchLds<CMAX> m;
    return m;
}

template <int C, int CMAX>
LV_D FrameTile<C>& tile_of(MatchLds<CMAX>& m) {
    if constexpr (C == CMAX) return m.tile.big;
    else return m.tile.c1;
}

template <int C, int CMAX>
LV_D PxIn<C>* px_in_of(MatchLds<CMAX>& m) {
    if constexpr (C == CMAX) return m.in.big;
    else return m.in.c1;
}

template <int C, int ALGO, int KF, int CMAX = C>
LV_D void match_body(const View& v, uint32_t bid, uint32_t nblk) {
    static_assert(C == CMAX || C == 1, "a match instantiation shares the LDS of CMAX channels");
    const uint8_t* frame = v.frame;
    const uint32_t seq_limit = v.seq_limit;
    MatchLds<CMAX>& L = match_lds<CMAX>();
    FrameTile<C>& tile = tile_of<C>(L);
    PredTile& pred = L.pred;
    uint8_t* lut = L.lut;
    const int ntx = (v.W + TW - 1) / TW;
    const int tid = (int)bid;
    const int x0 = (tid % ntx) * TW, y0 = (tid / ntx) * TH;
    uint16_t* s_perm = L.perm;
    uint32_t* s_nfg = L.nfg;
    PxIn<C>* s_in = px_in_of<C>(L);
    // this thread's tile pixel before the regrouping: its inputs are loaded alongside the tile
    const int ox = x0 + (int)threadIdx.x % TW, oy = y0 + (int)threadIdx.x / TW;
    const bool oin = ox < v.W && oy < v.H;
    const PxIn<C> pin = load_px_in<C, ALGO>(v, oin ? (uint32_t)(oy * v.W + ox) : 0u, oin);
    load_frame_tile<C>(tile, frame, v.W, v.H, x0, y0);
    load_pred_tile(pred, v, x0, y0);
    load_lut(lut, v.sc);
    s_in[threadIdx.x] = pin;
    __syncthreads();
    // A wavefront scans as long as its slowest lane, and foreground pixels scan every sample: the
    // tile's pixels are regrouped so those that were foreground in the previous frame share
    // wavefronts (stable order within each group). Each pixel's match depends on the tile data only.
    {
        const int t = (int)threadIdx.x;
        const bool likely_fg = oin && pred.guess[t / TW + 2][t % TW + 2] != 0 && pin.roi;
        const unsigned long long bal = __ballot(likely_fg);
        const int lane = t & 63, wave = t >> 6;
        if (lane == 0) s_nfg[wave] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t fg_before = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull)), nfg = 0;
        for (int w = 0; w < TPB / 64; ++w) {
            fg_before += w < wave ? s_nfg[w] : 0u;
            nfg += s_nfg[w];
        }
        const uint32_t pos = likely_fg ? (uint32_t)TPB - nfg + fg_before : (uint32_t)t - fg_before;
        s_perm[pos] = (uint16_t)t;
        __syncthreads();
    }
    const int src = s_perm[threadIdx.x];
    const int lx = src % TW, ly = src / TW, x = x0 + lx, y = y0 + ly;
    const uint32_t p = (uint32_t)(y * v.W + x);
    uint32_t examined = 0;
    bool cont = false;
    const PxIn<C> in = s_in[src];
    PxCtx<C> px;
    if (x < v.W && y < v.H && in.roi)
        cont = !match_pixel<C, ALGO, KF>(v, frame, lut, tile, pred, in, lx, ly, x, y, p, seq_limit, examined, px);
    // unfinished pixels go to the continuation with their context: one atomic per wavefront, on
    // this tile's stripe of the list (at most cont_region entries reach one stripe)
    const unsigned long long bal = __ballot(cont);
    if (bal) {
        const int lane = (int)(threadIdx.x & 63u), leader = __ffsll((long long)bal) - 1;
        const uint32_t k = bid % (uint32_t)kContStripes;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&v.sc->cont_cnt[k * kContStride], (uint32_t)__popcll(bal));
        base = (uint32_t)__shfl((int)base, leader, 64);
        const uint32_t j = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        if (cont) {
            G(v.cont)[(size_t)k * v.cont_region + j] = p;
            if (j < v.cont_ctx_region)
                store_ctx<C>(v.cont_ctx + ((size_t)k * v.cont_ctx_region + j) * kCtxWords, px);
        }
    }
    block_add<TPB>(examined, v.sc->spec_part);
}

// Phase A: sequential scan of at most v_seq_limit samples. Returns true when the pixel is final;
// otherwise its partial state is left in mres for the cooperative phase.
template <int C, int ALGO, int KF>
LV_D bool match_pixel(const View& v, const uint8_t* frame, const uint8_t* lut, const FrameTile<C>& tile,
                      const PredTile& pred, const PxIn<C>& in, int lx, int ly, int x, int y, uint32_t p,
                      uint32_t seq_limit, uint32_t& examined, PxCtx<C>& px) {
    // guess: raster-earlier writers decide as in the previous frame
    const uint32_t applied = applied_writers_tile(pred, lx, ly);
    uint64_t sbm;
    int nrep;
    ScanState st = scan_init<C>();
    const uint32_t N = (uint32_t)v.N;
    // more than kRepl overwritten slots (very rare): the whole scan goes to the cooperative phase.
    // The replacements are gathered before the context is built, so the two are not live together.
    auto rec_of = [&](int i) { return pred.rec[ly + 2 + kPredDY[i]][lx + 2 + kPredDX[i]]; };
    const uint32_t limit =
        build_repl<C>(v, frame, x, y, applied, sbm, nrep, rec_of) ? (seq_limit < N ? seq_limit : N) : 0u;
    build_ctx_staged<C, ALGO, FrameTile<C>::LW * C>(v, &tile.px[(ly * FrameTile<C>::LW + lx) * C], in, px);
    scan_seq<C, ALGO, FrameTile<C>::LW * C, KF>(v, lut, px, x, y, sbm, nrep, st, limit);
    examined = st.s;
    G(v.mres)[p] = pack_state(st);
    if (st.good < (uint32_t)v.REQ && st.s < N) return false;
    const uint8_t fg = st.good < (uint32_t)v.REQ ? 1 : 0;
    G(v.dec)[p] = fg;
    const uint8_t guess = G(v.last_raw)[p] ? 1 : 0;
    uint32_t t;
    if (fg != guess && forward_target(v, G(v.rec)[p], x, y, t)) wl_push(v, t, 1u, v.sc->epoch, v.wl1);
    return true;
}

// candidate writers of p whose decision is background: guessed from the previous raw mask, or the
// current decisions during the fix point
template <bool CURRENT>
LV_D uint32_t applied_writers_global(const View& v, int x, int y) {
    const uint32_t cand = candidate_writers(v, x, y);
    uint32_t applied = 0;
    for (uint32_t m = cand; m; m &= m - 1) {
        const int i = __builtin_ctz(m);
        const uint32_t q = (uint32_t)((y + kPredDY[i]) * v.W + (x + kPredDX[i]));
        const uint8_t d = CURRENT ? __hip_atomic_load(&v.dec[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : (G(v.last_raw)[q] ? 1 : 0);
        if (d == 0) applied |= 1u << i;
    }
    return applied;
}

// decision of a finished scan: match result, raw decision, and the forward target of a pixel whose
// decision differs from the speculative guess goes to the first fix-point worklist
template <int C>
LV_D void finish_pixel(const View& v, uint32_t p, int x, int y, const ScanState& st, uint64_t epoch) {
    G(v.mres)[p] = pack_state(st);
    const uint8_t fg = st.good < (uint32_t)v.REQ ? 1 : 0;
    G(v.dec)[p] = fg;
    const uint8_t guess = G(v.last_raw)[p] ? 1 : 0;
    uint32_t t;
    if (fg != guess && forward_target(v, G(v.rec)[p], x, y, t)) wl_push(v, t, 1u, epoch, v.wl1);
}

// Phase B: the scans phase A cut at its sample limit, continued with one lane per pixel over the
// dense list of unfinished pixels (its stripes walked in order), so long scans (mostly foreground
// pixels, which examine every sample) fill whole wavefronts. The context is the one phase A cached
// (rebuilt from global memory past a stripe's cache); a pixel with more than kRepl overwritten
// slots goes on to phase C.
template <int C, int ALGO, int KF>
LV_D void match_seq_body(const View& v, uint32_t bid, uint32_t nblk) {
    static_assert(kContStripes <= 64 && (kContStripes & (kContStripes - 1)) == 0, "one wavefront scans the stripe counts");
    const uint8_t* frame = v.frame;
    __shared__ uint8_t lut[256];
    __shared__ uint32_t s_pre[kContStripes + 1];  // exclusive prefix of the stripe counts
    if (threadIdx.x < 64u) {
        const int lane = (int)threadIdx.x;
        uint32_t c = lane < kContStripes ? v.sc->cont_cnt[lane * kContStride] : 0u;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = (uint32_t)__shfl_up((int)c, (unsigned)o, 64);
            if (lane >= o) c += t;
        }
        if (lane < kContStripes) s_pre[lane + 1] = c;
        if (lane == 0) s_pre[0] = 0u;
    }
    load_lut(lut, v.sc);
    __syncthreads();
    const uint32_t n = s_pre[kContStripes];
    if (bid * TPB >= n) return;  // uniform: no item reaches this block
    const uint64_t epoch = v.sc->epoch;
    uint32_t examined = 0;
    for (uint32_t i = bid * TPB + threadIdx.x; i < n; i += nblk * TPB) {
        // item i = entry j of stripe k: the last stripe whose prefix is at most i
        uint32_t k = 0;
        for (uint32_t half = kContStripes / 2; half > 0; half >>= 1)
            if (s_pre[k + half] <= i) k += half;
        const uint32_t j = i - s_pre[k];
        const uint32_t p = G(v.cont)[(size_t)k * v.cont_region + j];
        const int x = (int)(p % (uint32_t)v.W), y = (int)(p / (uint32_t)v.W);
        PxCtx<C> px;
        if (j < v.cont_ctx_region) load_ctx<C>(v.cont_ctx + ((size_t)k * v.cont_ctx_region + j) * kCtxWords, px);
        else build_ctx_global<C, ALGO>(v, frame, x, y, px);
        const uint32_t applied = applied_writers_global<false>(v, x, y);
        uint64_t sbm;
        int nrep;
        auto rec_of = [&](int i) { return G(v.rec)[(size_t)(y + kPredDY[i]) * v.W + (x + kPredDX[i])]; };
        if (!build_repl<C>(v, frame, x, y, applied, sbm, nrep, rec_of)) {
            v.cont2[atomicAdd(&v.sc->cont2_count, 1u)] = p;
            continue;
        }
        ScanState st = unpack_state(G(v.mres)[p]);
        const uint32_t s0 = st.s;
        scan_seq<C, ALGO, 0, KF>(v, lut, px, x, y, sbm, nrep, st, (uint32_t)v.N);
        examined += st.s - s0;
        finish_pixel<C>(v, p, x, y, st, epoch);
    }
    block_add<TPB>(examined, v.sc->coop_part);
}

// Phase C: pixels left by phase B (more than kRepl overwritten slots), kGroup lanes per pixel
template <int C, int ALGO>
LV_D void match_cont_body(const View& v, uint32_t bid, uint32_t nblk) {
    const uint8_t* frame = v.frame;
    __shared__ uint8_t lut[256];
    __shared__ uint8_t patches[TPB / kGroup][25 * C];
    const uint32_t n = v.sc->cont2_count;
    if (bid * (TPB / kGroup) >= n) return;  // uniform: no item reaches this block
    load_lut(lut, v.sc);
    __syncthreads();
    uint8_t* patch = patches[threadIdx.x / kGroup];
    const uint32_t ngroups = nblk * (TPB / kGroup);
    const uint32_t sub = threadIdx.x & (kGroup - 1);
    const uint32_t g0 = (bid * TPB + threadIdx.x) / kGroup;
    const uint32_t rounds = (n + ngroups - 1) / ngroups;  // uniform trip count: every lane takes part
    const uint64_t epoch = v.sc->epoch;
    uint64_t examined = 0;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t i = g0 + r * ngroups;
        const bool active = i < n;
        const uint32_t p = active ? G(v.cont2)[i] : 0u;
        const int x = (int)(p % (uint32_t)v.W), y = (int)(p / (uint32_t)v.W);
        PxCtx<C> px;
        uint32_t applied = 0;
        uint64_t sbm = 0;
        ScanState st = scan_init<C>();
        __syncthreads();  // previous round's readers of the patch are done
        if (active) {
            stage_patch<C>(frame, v.W, x, y, patch, (int)sub);
            wave_sync();  // the group's patch bytes, written by its eight lanes
            build_ctx<C, ALGO, 5 * C>(v, frame, patch, x, y, px);
            applied = applied_writers_global<false>(v, x, y);
            sbm = applied_slots(v, x, y, applied);
            st = unpack_state(G(v.mres)[p]);
        }
        __syncthreads();
        const uint32_t s_before = st.s;
        scan_coop<C, ALGO, 5 * C>(v, frame, lut, px, x, y, applied, sbm, st, active);
        if (active && sub == 0) {
            examined += st.s - s_before;
            finish_pixel<C>(v, p, x, y, st, epoch);
        }
    }
    block_add<TPB>(examined, v.sc->coop_part);
}

// fix-point re-match of worklisted pixels (all lanes call; kGroup lanes per item)
template <int C, int ALGO>
LV_D void rematch_round(const View& v, const uint8_t* frame, const uint8_t* lut, uint8_t* patch, const uint32_t* wl,
                        uint32_t n, uint32_t i, uint32_t k, uint64_t epoch, uint32_t* wl_next) {
    const bool active = i < n;
    const uint32_t p = active ? wl[i] : 0u;
    const int x = (int)(p % (uint32_t)v.W), y = (int)(p / (uint32_t)v.W);
    PxCtx<C> px;
    uint32_t applied = 0;
    uint64_t sbm = 0;
    __syncthreads();  // previous round's readers of the patch are done
    if (active) {
        stage_patch<C>(frame, v.W, x, y, patch, (int)(threadIdx.x & (kGroup - 1)));
        wave_sync();  // the group's patch bytes, written by its eight lanes
        build_ctx<C, ALGO, 5 * C>(v, frame, patch, x, y, px);
        applied = applied_writers_global<true>(v, x, y);
        sbm = applied_slots(v, x, y, applied);
    }
    __syncthreads();
    ScanState st = scan_init<C>();
    scan_coop<C, ALGO, 5 * C>(v, frame, lut, px, x, y, applied, sbm, st, active);
    if (active && (threadIdx.x & (kGroup - 1)) == 0) {
        G(v.mres)[p] = pack_state(st);
        const uint8_t fg = st.good < (uint32_t)v.REQ ? 1 : 0;
        const uint8_t old = G(v.dec)[p];
        if (fg != old) {
            __hip_atomic_store(&v.dec[p], fg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t t;
            if (forward_target(v, G(v.rec)[p], x, y, t)) wl_push(v, t, k + 1u, epoch, wl_next);
        }
    }
}

// fix-point iteration k (k >= 1): worklist k -> worklist k+1
template <int C, int ALGO>
LV_D void fixpoint_body(const View& v, uint32_t bid, uint32_t nblk, uint32_t k) {
    const uint8_t* frame = v.frame;
    __shared__ uint8_t lut[256];
    __shared__ uint8_t patches[TPB / kGroup][25 * C];
    // slot k+2 was last read by iteration k-2 (finished); it must be empty before iteration k+1
    // pushes into it, whether or not this iteration has work.
    if (bid == 0 && threadIdx.x == 0)
        __hip_atomic_store(&v.sc->wl_count[(k + 2) % kWlRing], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t n = __hip_atomic_load(&v.sc->wl_count[k % kWlRing], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (n == 0) return;
    if (bid == 0 && threadIdx.x == 0) {
        atomicAdd(&v.sc->iters, 1u);
        atomicMax(&v.sc->max_iter_seen, k);
        if (k < 8) v.sc->wl_work[k] = n;
    }
    if (bid * (TPB / kGroup) >= n) return;  // uniform: no item reaches this block
    load_lut(lut, v.sc);
    __syncthreads();
    const uint32_t* wl = (k & 1u) ? v.wl1 : v.wl0;
    uint32_t* wl_next = (k & 1u) ? v.wl0 : v.wl1;
    const uint64_t epoch = v.sc->epoch;
    const uint32_t ngroups = nblk * (TPB / kGroup);
    const uint32_t g0 = (bid * TPB + threadIdx.x) / kGroup;
    const uint32_t rounds = (n + ngroups - 1) / ngroups;
    for (uint32_t r = 0; r < rounds; ++r)
        rematch_round<C, ALGO>(v, frame, lut, patches[threadIdx.x / kGroup], wl, n, g0 + r * ngroups, k, epoch, wl_next);
}

// remaining fix-point iterations in one workgroup (normally no work is left)
template <int C, int ALGO>
LV_D void fixpoint_tail_body(const View& v, uint32_t bid, uint32_t nblk, uint32_t k0, uint32_t cap) {
    const uint8_t* frame = v.frame;
    __shared__ uint8_t lut[256];
    __shared__ uint8_t patches[256 / kGroup][25 * C];
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
                    if (k < 8) v.sc->wl_work[k] = s_n;
                }
            }
        }
        __syncthreads();
        const uint32_t n = s_n;
        if (n == 0) break;
        const uint32_t* wl = (k & 1u) ? v.wl1 : v.wl0;
        uint32_t* wl_next = (k & 1u) ? v.wl0 : v.wl1;
        const uint32_t ngroups = blockDim.x / kGroup, g0 = threadIdx.x / kGroup;
        const uint32_t rounds = (n + ngroups - 1) / ngroups;
        for (uint32_t r = 0; r < rounds; ++r)
            rematch_round<C, ALGO>(v, frame, lut, patches[threadIdx.x / kGroup], wl, n, g0 + r * ngroups, k, epoch, wl_next);
        __threadfence();
    }
}

// ----------------------------------------------------------------------------------------------
// SuBSENSE feedback maps (SUB:396-420, :451-480; 1ch :254-329), raw mask and blink bookkeeping
// (SUB:536-539). Whole image: pixels outside the ROI get raw = 0.
// ----------------------------------------------------------------------------------------------
// One pixel (p < npx). Returns p's samples examined (k_p) and its RawST / MLD after the update
// (the values the model update's ghost test reads).
template <int C>
LV_D uint32_t feedback_pixel(const View& v, uint32_t p, float& rst_out, float& mld_out) {
    const uint8_t* frame = v.frame;
    uint32_t examined = 0;
    uint8_t raw = 0;
    if (!G(v.roi)[p]) {
        rst_out = G(v.RawST)[p];
        mld_out = G(v.MLD)[p];
    } else {
        const uint32_t frame_no = cur_frame(v);
        const uint32_t mvn = v.mv_avgs;
        const float aLT = 1.0f / (float)(frame_no < mvn ? frame_no : mvn);
        const float aST = 1.0f / (float)(frame_no < mvn / 4u ? frame_no : mvn / 4u);
        const uint32_t mr = G(v.mres)[p];
        const uint32_t good = mr & 255u, md = (mr >> 8) & 63u, ms = (mr >> 14) & 1023u;
        examined = mr >> 24;
        const bool fg = G(v.dec)[p] != 0;
        const uint32_t r = G(v.rec)[p];
        const uint32_t unst_new = rec_unstable(r) ? 1u : 0u;
        const float col_max = (float)(C * 255), desc_max = (float)(C * 16);
        // D_last (SUB:396-397): L1dist<3> accumulates in uchar (wraps mod 256)
        uint32_t l1 = 0, hd = 0;
        uint32_t cur[C], intra[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            cur[c] = G(frame)[(size_t)p * C + c];
            intra[c] = G(v.desc_cur)[(size_t)p * C + c];
            l1 += absdiff(G(v.last_color)[(size_t)p * C + c], cur[c]);
            hd += popc((uint32_t)G(v.last_desc)[(size_t)p * C + c] ^ intra[c]);
        }
        l1 &= 255u;
        const float nld = ((float)l1 / col_max + (float)hd / desc_max) / 2.0f;
        const float mld = G(v.MLD)[p] * (1.0f - aST) + nld * aST;
        G(v.MLD)[p] = mld;
        float dlt = G(v.DminLT)[p], dst = G(v.DminST)[p], rlt = G(v.RawLT)[p], rst = G(v.RawST)[p];
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
        G(v.DminLT)[p] = dlt;
        G(v.DminST)[p] = dst;
        G(v.RawLT)[p] = rlt;
        G(v.RawST)[p] = rst;
        rst_out = rst;
        mld_out = mld;
        const float lcap = v.sc->lower_cap, ucap = v.sc->upper_cap;
        const float mn = fmin_ref(dlt, dst), mx = fmax_ref(dlt, dst);
        const bool last_fg = G(v.last_fg)[p] != 0;
        float T = G(v.T)[p], V = G(v.V)[p], R = G(v.R)[p];
        if (last_fg || (mn < 0.1f && raw)) {
            if (T < ucap) T += 0.5f / (mx * V);
        } else if (T > lcap) {
            T -= 0.25f * V / mx;
        }
        if (T < lcap) T = lcap;
        else if (T > ucap) T = ucap;
        if (mx > 0.1f && G(v.blinks)[p]) {
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
        G(v.T)[p] = T;
        G(v.V)[p] = V;
        G(v.R)[p] = R;
        G(v.unstable)[p] = (uint8_t)unst_new;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            G(v.last_color)[(size_t)p * C + c] = (uint8_t)cur[c];
            G(v.last_desc)[(size_t)p * C + c] = (uint16_t)intra[c];
        }
    }
    const uint8_t cb = raw ^ G(v.last_raw)[p];
    G(v.blinks)[p] = cb | G(v.last_raw_blink)[p];
    G(v.last_raw_blink)[p] = cb;
    G(v.last_raw)[p] = raw;
    G(v.raw)[p] = raw;
    return examined;
}

// LOBSTER: raw mask (LOB:496-497, :554-555) and previous-raw bookkeeping for the next speculation
LV_D void lobster_raw_body(const View& v, uint32_t bid, uint32_t nblk) {
    const uint32_t p = bid * TPB + threadIdx.x;
    uint32_t examined = 0;
    if (p < v.npx) {
        uint8_t raw = 0;
        if (G(v.roi)[p]) {
            raw = G(v.dec)[p] ? 255 : 0;
            examined = G(v.mres)[p] >> 24;
        }
        G(v.raw)[p] = raw;
        G(v.last_raw)[p] = raw;
    }
    block_add<TPB>(examined, v.sc->ex_part);
}

// ----------------------------------------------------------------------------------------------
// model update: every pixel of the inner rectangle gathers the writes aimed at its samples, in
// the raster order of the writers; the last write of each slot wins (SUB:406-449, LOB:498-512).
// ----------------------------------------------------------------------------------------------
// FUSED (SuBSENSE): the feedback maps of the thread's pixel are updated first, in the same pass
// (they depend on the pixel alone), and the ghost test uses the new values from registers.
template <int C, int ALGO, bool FUSED>
LV_D void update_body(const View& v, uint32_t bid, uint32_t nblk) {
    const uint8_t* frame = v.frame;
    constexpr int LW = TW + 4, LH = TH + 4, NT = (LW * LH + TPB - 1) / TPB;
    // the tile's 5x5 neighbourhood: records (tgt = kRecNone outside the image), "writer decided
    // background", and the writers' current colour and intra descriptor (the values they write)
    __shared__ uint32_t s_rec[LH][LW];
    __shared__ uint8_t s_bg[LH][LW];
    __shared__ uint8_t s_col[LH][LW * C];
    __shared__ uint16_t s_desc[LH][LW * C];
    const int ntx = (v.W + TW - 1) / TW;
    const int tid = (int)bid;
    const int x0 = (tid % ntx) * TW, y0 = (tid / ntx) * TH;
    // neighbourhood loads first, into registers: their latency overlaps the feedback pass below
    uint32_t t_rec[NT], t_col[NT][C], t_desc[NT][C];
    uint8_t t_dec[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const int i = (int)threadIdx.x + k * TPB;
        const int ly = i / LW, lx = i % LW, gx = x0 + lx - 2, gy = y0 + ly - 2;
        const bool in = i < LW * LH && gx >= 0 && gx < v.W && gy >= 0 && gy < v.H;
        const size_t q = in ? (size_t)gy * v.W + gx : 0;
        t_rec[k] = in ? G(v.rec)[q] : kRecNone;
        t_dec[k] = G(v.dec)[q];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            t_col[k][c] = G(frame)[q * C + c];
            t_desc[k][c] = G(v.desc_cur)[q * C + c];
        }
    }
    const int lx = threadIdx.x % TW, ly = threadIdx.x / TW, x = x0 + lx, y = y0 + ly;
    uint32_t nw = 0, examined = 0;
    float rst = 0.f, mld = 1.f;
    if constexpr (FUSED)
        if (x < v.W && y < v.H) examined = feedback_pixel<C>(v, (uint32_t)(y * v.W + x), rst, mld);
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const int i = (int)threadIdx.x + k * TPB;
        if (i >= LW * LH) break;
        const int ly2 = i / LW, lx2 = i % LW;
        s_rec[ly2][lx2] = t_rec[k];
        s_bg[ly2][lx2] = (rec_tgt(t_rec[k]) != kRecNone && t_dec[k] == 0) ? 1 : 0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            s_col[ly2][lx2 * C + c] = (uint8_t)t_col[k][c];
            s_desc[ly2][lx2 * C + c] = (uint16_t)t_desc[k][c];
        }
    }
    __syncthreads();
    if (x >= 2 && y >= 2 && x < v.W - 2 && y < v.H - 2) {
        const uint32_t p = (uint32_t)(y * v.W + x);
        // ghost test of writers at or after p reads p's updated maps (SUB:438-441)
        bool ghost_new = false;
        if constexpr (FUSED) ghost_new = rst > 0.995f && mld < 0.010f;
        else if constexpr (ALGO == kSuBSENSE) ghost_new = G(v.RawST)[p] > 0.995f && G(v.MLD)[p] < 0.010f;
        // Writes landing on p, as bits in sequential order: writer j of the 5x5 window (raster
        // order) owns bit 2j (its self update, j == 12 only) and bit 2j+1 (its diffusion into p).
        uint64_t wm = 0;
#pragma unroll
        for (int j = 0; j < 25; ++j) {
            const int dy = j / 5 - 2, dx = j % 5 - 2;
            const uint32_t r = s_rec[ly + 2 + dy][lx + 2 + dx];
            const bool bg = s_bg[ly + 2 + dy][lx + 2 + dx] != 0;
            if (j == 12 && rec_tgt(r) != kRecNone && (bg ? rec_self_bg(r) : (ALGO == kSuBSENSE && rec_self_fg(r))))
                wm |= 1ull << (2 * j);
            if (bg && rec_tgt(r) == (uint32_t)((2 - dy) * 5 + (2 - dx)) &&
                (j < 12 ? fwd_gate(r) : (rec_gate_a(r) || (rec_gate_b(r) && ghost_new))))
                wm |= 1ull << (2 * j + 1);
        }
        // Visited from the last write back, so the first write seen for a slot is the one the
        // sequential loop leaves in place; earlier writes to a written slot are skipped.
        uint64_t done0 = 0, done1 = 0, done2 = 0, done3 = 0;  // slots already final (N <= 255)
        nw = (uint32_t)__popcll(wm);
        while (wm) {
            const int b = 63 - __clzll((long long)wm);
            wm &= ~(1ull << b);
            const int j = b >> 1, dy = j / 5 - 2, dx = j % 5 - 2;
            const uint32_t r = s_rec[ly + 2 + dy][lx + 2 + dx];
            const uint32_t sl = (b & 1) ? rec_slot_diff(r) : rec_slot_self(r);
            const uint64_t bit = 1ull << (sl & 63u);
            const uint32_t word = sl >> 6;
            const uint64_t dw = word == 0 ? done0 : word == 1 ? done1 : word == 2 ? done2 : done3;
            if (dw & bit) continue;
            if (word == 0) done0 |= bit;
            else if (word == 1) done1 |= bit;
            else if (word == 2) done2 |= bit;
            else done3 |= bit;
            uint32_t col[C], desc[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                col[c] = s_col[ly + 2 + dy][(lx + 2 + dx) * C + c];
                desc[c] = s_desc[ly + 2 + dy][(lx + 2 + dx) * C + c];
            }
            store_sample<C>(v, sl, p, col, desc);
        }
    }
    block_add<TPB>(nw, v.sc->wr_part);
    if constexpr (FUSED) block_add<TPB>(examined, v.sc->ex_part);  // workload counter (sum of k_p)
}

// LOBSTER: last colour frame := input (LOB:580)
template <int C>
LV_D void copy_last_color_body(const View& v, uint32_t bid, uint32_t nblk) {
    const uint8_t* frame = v.frame;
    const uint32_t i = bid * TPB + threadIdx.x;
    if (i < v.npx * (uint32_t)C) G(v.last_color)[i] = G(frame)[i];
}

// ----------------------------------------------------------------------------------------------
// post-processing (SUB:540-554, LOB:578). Masks are 0/255 bytes.
// ----------------------------------------------------------------------------------------------
constexpr int PW = 64, PH = 16;  // post-processing tile; 256 threads x 4 pixels

// --- hole fill (SUB:540-543) on 64x64 zones, one wavefront per zone --------------------------
// PreFlood = morphologyEx(raw, CLOSE, 3x3) = erode3(dilate3(raw)) with neutral borders (SUB:540);
// holes = its 0-pixels not 4-connected through 0-pixels to pixel (0,0) (floodFill + bitwise_not,
// SUB:541-543). Lane y of a zone's wavefront holds row y of the zone as a 64-bit mask (bit x =
// column x). A zone's 0-pixels split into components that touch the zone border, numbered per
// zone, and components inside it, which are holes. The border components of all zones are joined
// across zone edges by union-find; a border component is outside iff its root is the root of the
// component holding pixel (0,0).
constexpr int ZS = 64;                  // zone side
constexpr int ZPB = TPB / 64;           // zones per block, one per wavefront
constexpr uint32_t kZoneComps = 128;    // border components per zone (at most 126 of 252 pixels)

LV_D uint64_t brev64(uint64_t x) {
    return ((uint64_t)__builtin_bitreverse32((uint32_t)x) << 32) | __builtin_bitreverse32((uint32_t)(x >> 32));
}
LV_D uint64_t lane_bcast(uint64_t x, int l) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(x >> 32), l) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
}
LV_D uint64_t from_lane_above(uint64_t x, int d, int lane) {  // lane - d, 0 for the first d lanes
    const uint64_t r = __shfl_up(x, (unsigned)d, 64);
    return lane >= d ? r : 0ull;
}
LV_D uint64_t from_lane_below(uint64_t x, int d, int lane) {  // lane + d, 0 for the last d lanes
    const uint64_t r = __shfl_down(x, (unsigned)d, 64);
    return lane + d < 64 ? r : 0ull;
}

// the runs of z along the row that hold a bit of s (s within z). Adding s to z carries through
// the run above each seed bit and z & ~(z + s) keeps the bits the carry cleared; the same on the
// bit-reversed row covers the run below.
LV_D uint64_t fill_row(uint64_t z, uint64_t zr, uint64_t s) {
    const uint64_t up = (z & ~(z + s)) | s;
    const uint64_t ur = brev64(up);
    return brev64((zr & ~(zr + ur)) | ur);
}
// the runs of z down and up the columns that hold a bit of s: segmented OR scans over the lanes,
// g = reached at this row from the rows above within the window, p = window all in z
LV_D uint64_t fill_cols(uint64_t z, uint64_t s, int lane) {
    uint64_t g = s, p = z;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t gu = from_lane_above(g, d, lane), pu = from_lane_above(p, d, lane);
        g |= p & gu;
        p &= pu;
    }
    p = z;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t gd = from_lane_below(g, d, lane), pd = from_lane_below(p, d, lane);
        g |= p & gd;
        p &= pd;
    }
    return g;
}
// the 4-connected components of z holding a bit of s; each round adds at least one bit or ends
LV_D uint64_t flood(uint64_t z, uint64_t zr, uint64_t s, int lane) {
    uint64_t f = s & z;
    while (true) {
        const uint64_t g = fill_cols(z, fill_row(z, zr, f), lane);
        if (__ballot(g != f) == 0ull) return g;
        f = g;
    }
}

// raw bits of columns x0-2, x0-1, x0+64, x0+65 of row gy (bits 0..3)
LV_D uint32_t raw_halo(const uint8_t* raw, int W, int H, int x0, int gy) {
    if (gy < 0 || gy >= H) return 0u;
    const uint8_t* row = raw + (size_t)gy * (size_t)W;
    const int xs[4] = {x0 - 2, x0 - 1, x0 + ZS, x0 + ZS + 1};
    uint32_t h = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (xs[i] >= 0 && xs[i] < W && row[xs[i]] != 0) h |= 1u << i;
    return h;
}
// 3-wide dilation along a row: core columns x0..x0+63 and the halo bits above -> dilated core and
// dilated halo (bit 0: column x0-1, bit 1: column x0+64)
LV_D void dilate3_row(uint64_t c, uint32_t h, uint64_t& d, uint32_t& dh) {
    d = c | (c << 1) | (uint64_t)((h >> 1) & 1u) | (c >> 1) | ((uint64_t)((h >> 2) & 1u) << 63);
    const uint32_t left = (h & 1u) | ((h >> 1) & 1u) | (uint32_t)(c & 1ull);
    const uint32_t right = (uint32_t)(c >> 63) | ((h >> 2) & 1u) | ((h >> 3) & 1u);
    dh = left | (right << 1);
}

// closing of the zone, its 0-pixels, and their border components
LV_D void zone_close_body(const View& v, uint32_t bid, uint32_t nblk) {
    __shared__ uint64_t s_d[ZPB][ZS + 4];   // row-dilated raw rows y0-2 .. y0+65
    __shared__ uint32_t s_dx[ZPB][ZS + 4];
    const int W = v.W, H = v.H;
    const int lane = (int)(threadIdx.x & 63u), wv = (int)(threadIdx.x >> 6);
    const int nzx = (W + ZS - 1) / ZS, nz = nzx * ((H + ZS - 1) / ZS);
    const int z = (int)bid * ZPB + wv;
    if (z >= nz) return;  // whole wavefronts only; nothing below synchronizes the block
    const int x0 = (z % nzx) * ZS, y0 = (z / nzx) * ZS;
    const uint8_t* raw = v.raw;
    // raw rows y0-2+r as 64-bit masks: lane = column for the loads, one ballot per row; lane y
    // keeps row y, lanes 0..3 also rows 64..67
    const int gx = x0 + lane;
    uint64_t c0 = 0, c1 = 0;
#pragma unroll 17
    for (int r = 0; r < ZS + 4; ++r) {
        const int gy = y0 - 2 + r;
        const uint8_t b = (gx < W && gy >= 0 && gy < H) ? raw[(size_t)gy * (size_t)W + gx] : (uint8_t)0;
        const uint64_t m = __ballot(b != 0);
        if (r < ZS) c0 = lane == r ? m : c0;
        else c1 = lane == r - ZS ? m : c1;
    }
    const uint32_t h0 = raw_halo(raw, W, H, x0, y0 - 2 + lane);
    const uint32_t h1 = lane < 4 ? raw_halo(raw, W, H, x0, y0 - 2 + ZS + lane) : 0u;
    {
        uint64_t d;
        uint32_t dh;
        dilate3_row(c0, h0, d, dh);
        s_d[wv][lane] = d;
        s_dx[wv][lane] = dh;
        dilate3_row(c1, h1, d, dh);
        if (lane < 4) {
            s_d[wv][ZS + lane] = d;
            s_dx[wv][ZS + lane] = dh;
        }
    }
    wave_sync();
    // lane y: zone row y = dilated-row index y+2; erosion over dilated rows y+1..y+3, each the OR
    // of three row-dilated rows; outside the image the dilation is 255 (neutral for the erosion)
    const uint64_t in_cols = W - x0 >= ZS ? ~0ull : ((1ull << (W - x0)) - 1ull);
    const uint32_t out_halo = (x0 == 0 ? 1u : 0u) | (x0 + ZS >= W ? 2u : 0u);
    uint64_t e = ~0ull;
    uint32_t ex = 3u;
#pragma unroll
    for (int a = 1; a <= 3; ++a) {
        const int r = lane + a, gy = y0 - 2 + r;
        uint64_t d = s_d[wv][r - 1] | s_d[wv][r] | s_d[wv][r + 1];
        uint32_t dx = s_dx[wv][r - 1] | s_dx[wv][r] | s_dx[wv][r + 1];
        if (gy < 0 || gy >= H) {
            d = ~0ull;
            dx = 3u;
        } else {
            d |= ~in_cols;
            dx |= out_halo;
        }
        e &= d;
        ex &= dx;
    }
    const uint64_t pre = e & ((e << 1) | (uint64_t)(ex & 1u)) & ((e >> 1) | ((uint64_t)(ex >> 1) << 63));
    const uint64_t zm = y0 + lane < H ? (~pre & in_cols) : 0ull;
    v.zbits[(size_t)z * ZS + lane] = zm;
    // border components in the order top row, left column, right column, bottom row
    const uint64_t zr = brev64(zm);
    uint64_t remT = lane_bcast(zm, 0), remB = lane_bcast(zm, 63);
    uint64_t remL = __ballot((zm & 1ull) != 0), remR = __ballot((zm >> 63) != 0);
    uint32_t cT = 255u, cB = 255u, cL = 255u, cR = 255u, k = 0;
    while ((remT | remB | remL | remR) != 0ull) {
        uint64_t s;
        if (remT) s = lane == 0 ? 1ull << __builtin_ctzll(remT) : 0ull;
        else if (remL) s = lane == __builtin_ctzll(remL) ? 1ull : 0ull;
        else if (remR) s = lane == __builtin_ctzll(remR) ? 1ull << 63 : 0ull;
        else s = lane == 63 ? 1ull << __builtin_ctzll(remB) : 0ull;
        const uint64_t f = flood(zm, zr, s, lane);
        const uint64_t fT = lane_bcast(f, 0), fB = lane_bcast(f, 63);
        const uint64_t fL = __ballot((f & 1ull) != 0), fR = __ballot((f >> 63) != 0);
        if ((fT >> lane) & 1ull) cT = k;
        if ((fB >> lane) & 1ull) cB = k;
        if ((fL >> lane) & 1ull) cL = k;
        if ((fR >> lane) & 1ull) cR = k;
        remT &= ~fT;
        remB &= ~fB;
        remL &= ~fL;
        remR &= ~fR;
        ++k;
    }
    uint8_t* edge = v.zedge + (size_t)z * 256;
    edge[lane] = (uint8_t)cT;
    edge[64 + lane] = (uint8_t)cB;
    edge[128 + lane] = (uint8_t)cL;
    edge[192 + lane] = (uint8_t)cR;
    const uint32_t base = (uint32_t)z * kZoneComps;
    if ((uint32_t)lane < k) v.zpar[base + lane] = base + lane;
    if ((uint32_t)lane + 64u < k) v.zpar[base + lane + 64] = base + lane + 64;
    if (lane == 0) v.zcount[z] = k;
}

LV_D uint32_t g_find(uint32_t* lab, uint32_t a) {
    while (true) {
        const uint32_t pa = __hip_atomic_load(&lab[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (pa == a) return a;
        a = pa;
    }
}

// the same without concurrent unions: path halving (every store is an ancestor of the node)
LV_D uint32_t g_find_halving(uint32_t* lab, uint32_t a) {
    while (true) {
        const uint32_t pa = __hip_atomic_load(&lab[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (pa == a) return a;
        const uint32_t ga = __hip_atomic_load(&lab[pa], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (ga != pa) __hip_atomic_store(&lab[a], ga, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a = ga;
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

// joins the components facing each other across one zone edge (lane = position along the edge);
// a lane whose pair repeats the previous lane's leaves the union to that lane
LV_D void zone_link(uint32_t* par, uint32_t za, uint32_t zb, uint32_t a, uint32_t b, int lane) {
    const uint32_t pa = (uint32_t)__shfl_up((int)a, 1u, 64), pb = (uint32_t)__shfl_up((int)b, 1u, 64);
    const bool dup = lane > 0 && pa == a && pb == b;
    if (a != 255u && b != 255u && !dup) g_union(par, za * kZoneComps + a, zb * kZoneComps + b);
}

// unions across the right and bottom edges of every zone
LV_D void zone_link_body(const View& v, uint32_t bid, uint32_t nblk) {
    const int lane = (int)(threadIdx.x & 63u), wv = (int)(threadIdx.x >> 6);
    const int nzx = (v.W + ZS - 1) / ZS, nzy = (v.H + ZS - 1) / ZS;
    const int z = (int)bid * ZPB + wv;
    if (z >= nzx * nzy) return;
    const int zx = z % nzx, zy = z / nzx;
    const uint8_t* e = v.zedge;
    if (zx + 1 < nzx)
        zone_link(v.zpar, (uint32_t)z, (uint32_t)z + 1u, e[(size_t)z * 256 + 192 + lane],
                  e[(size_t)(z + 1) * 256 + 128 + lane], lane);
    if (zy + 1 < nzy)
        zone_link(v.zpar, (uint32_t)z, (uint32_t)(z + nzx), e[(size_t)z * 256 + 64 + lane],
                  e[(size_t)(z + nzx) * 256 + lane], lane);
}

// holes of a zone: its 0-pixels minus the flood from the border pixels of outside components
LV_D void zone_holes_body(const View& v, uint32_t bid, uint32_t nblk) {
    const int lane = (int)(threadIdx.x & 63u), wv = (int)(threadIdx.x >> 6);
    const int nzx = (v.W + ZS - 1) / ZS, nz = nzx * ((v.H + ZS - 1) / ZS);
    const int z = (int)bid * ZPB + wv;
    if (z >= nz) return;
    const uint64_t zm = v.zbits[(size_t)z * ZS + lane];
    const uint8_t* e = v.zedge + (size_t)z * 256;
    const uint32_t cT = e[lane], cB = e[64 + lane], cL = e[128 + lane], cR = e[192 + lane];
    const uint32_t k = v.zcount[z];
    const uint32_t c00 = v.zedge[0];  // component of pixel (0,0) in zone 0, 255 if it is foreground
    const uint32_t root = c00 == 255u ? 0xFFFFFFFFu : g_find_halving(v.zpar, c00);
    const uint32_t base = (uint32_t)z * kZoneComps;
    const bool o0 = (uint32_t)lane < k && g_find_halving(v.zpar, base + lane) == root;
    const bool o1 = (uint32_t)lane + 64u < k && g_find_halving(v.zpar, base + lane + 64) == root;
    const uint64_t m0 = __ballot(o0), m1 = __ballot(o1);
    auto outside = [&](uint32_t c) {
        return c < 64u ? ((m0 >> c) & 1ull) != 0 : (c < 128u && ((m1 >> (c - 64u)) & 1ull) != 0);
    };
    const uint64_t sT = __ballot(outside(cT)), sB = __ballot(outside(cB));
    uint64_t s = (outside(cL) ? 1ull : 0ull) | (outside(cR) ? 1ull << 63 : 0ull);
    if (lane == 0) s |= sT;
    if (lane == 63) s |= sB;
    const uint64_t o = flood(zm, brev64(zm), s, lane);
    v.hbits[(size_t)z * ZS + lane] = zm & ~o;
}

// Post-processing masks are 0/255 bytes; the median and dilation kernels below work on four pixels
// per 32-bit word. Words start at columns that are multiples of 4 (tile origins are), so when the
// row pitch W is too a word is one aligned load or store; other widths assemble bytes.
// in_image: 0xFF for each byte of the word at (gx..gx+3, gy) that lies inside the image
LV_D uint32_t in_image_bytes(int W, int H, int gx, int gy) {
    if (gy < 0 || gy >= H || gx + 3 < 0 || gx >= W) return 0u;
    if (gx >= 0 && gx + 3 < W) return 0xFFFFFFFFu;
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (gx + i >= 0 && gx + i < W) m |= 0xFFu << (8 * i);
    return m;
}

// bytes (gx..gx+3, gy) of a W x H byte image, `pad` for those outside it; gx % 4 == 0
LV_D uint32_t load_mask_word(const uint8_t* m, int W, int H, int gx, int gy, uint32_t pad) {
    const uint32_t fill = pad * 0x01010101u;
    if (gy < 0 || gy >= H || gx + 3 < 0 || gx >= W) return fill;
    const uint8_t* row = m + (size_t)gy * (size_t)W;
    if ((W & 3) == 0) return *(const uint32_t*)(row + gx);  // gx >= 0 and gx + 3 < W here
    uint32_t w = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int x = gx + i;
        w |= (x >= 0 && x < W ? (uint32_t)row[x] : pad) << (8 * i);
    }
    return w;
}

LV_D void store_mask_word(uint8_t* m, int W, int H, int gx, int gy, uint32_t w) {
    if (gy >= H || gx >= W) return;
    uint8_t* row = m + (size_t)gy * (size_t)W;
    if ((W & 3) == 0) {
        *(uint32_t*)(row + gx) = w;
        return;
    }
    for (int i = 0; i < 4 && gx + i < W; ++i) row[gx + i] = (uint8_t)(w >> (8 * i));
}

// bits (gx..gx+3, gy) of a zone bitmap (gx % 4 == 0), zero outside the image
LV_D uint32_t zone_bits4(const uint64_t* bits, int W, int H, int gx, int gy) {
    if (gx < 0 || gy < 0 || gx >= W || gy >= H) return 0u;
    const size_t zi = (size_t)(gy / ZS) * (size_t)((W + ZS - 1) / ZS) + (size_t)(gx / ZS);
    return (uint32_t)(bits[zi * ZS + (size_t)(gy % ZS)] >> (gx % ZS)) & 15u;
}
// four bits -> four bytes of 0/255 (the products of the bits with 1, 2^7, 2^14, 2^21 never
// overlap, so bit i lands alone at bit 8i)
LV_D uint32_t expand4(uint32_t b) { return ((b * 0x00204081u) & 0x01010101u) * 255u; }

// byte window of two adjacent words: bytes s..s+3 of (lo, hi)
LV_D uint32_t bytes_at(uint32_t hi, uint32_t lo, uint32_t s) { return __builtin_amdgcn_alignbyte(hi, lo, s); }

// AND (OR) of the word whose bytes start 4 bytes into a, b, c and its shifts by up to 3 bytes either
// way: the 7-wide erosion (dilation) of the 0/255 bytes of b along the row
LV_D uint32_t erode7_word(uint32_t a, uint32_t b, uint32_t c) {
    return b & bytes_at(b, a, 1) & bytes_at(b, a, 2) & bytes_at(b, a, 3) & bytes_at(c, b, 1) & bytes_at(c, b, 2) &
           bytes_at(c, b, 3);
}
LV_D uint32_t dilate7_word(uint32_t a, uint32_t b, uint32_t c) {
    return b | bytes_at(b, a, 1) | bytes_at(b, a, 2) | bytes_at(b, a, 3) | bytes_at(c, b, 1) | bytes_at(c, b, 2) |
           bytes_at(c, b, 3);
}

// mask = raw | holes | erode7(PreFlood); last_fg = medianBlur(mask, k) (SUB:544-547; LOB:578 uses
// the raw mask directly). Binary median with zero padding = majority count over the k x k window.
// Counts are kept one per byte (at most 13 per column and 169 per window, so no byte carries into
// the next) and the majority test adds 128 - need to every byte and keeps bit 7.
template <bool SUBSENSE>
LV_D void median_body(const View& v, uint32_t bid, uint32_t nblk) {
    constexpr int HM = 6, HE = 3;           // median halo (k <= 13), erosion halo
    constexpr int MWW = (PW + 16) / 4;      // mask words: columns x0-8 .. x0+PW+7
    constexpr int MH = PH + 2 * HM;         // mask rows: y0-6 .. y0+PH+5
    constexpr int EWW = (PW + 24) / 4;      // PreFlood words: columns x0-12 .. x0+PW+11
    constexpr int EH = MH + 2 * HE;         // PreFlood rows: y0-9 .. y0+PH+8
    constexpr int OWW = PW / 4;             // output words per row
    __shared__ uint32_t s_pre[EH][EWW];
    __shared__ uint32_t s_eh[EH][MWW];
    __shared__ uint32_t s_mask[MH][MWW];
    __shared__ uint32_t s_col[PH][MWW];
    const int W = v.W, H = v.H;
    const int ntx = (W + PW - 1) / PW;
    const int tid = (int)bid;
    const int x0 = (tid % ntx) * PW, y0 = (tid / ntx) * PH;
    const int k = v.median_k, r = k / 2, need = (k * k + 1) / 2;
    if constexpr (SUBSENSE) {
        // erosion treats the outside of the image as 255
        for (int i = threadIdx.x; i < EH * EWW; i += TPB) {
            const int ly = i / EWW, wx = i % EWW;
            // PreFlood = 255 except at its 0-pixels, 255 outside the image
            s_pre[ly][wx] = expand4(~zone_bits4(v.zbits, W, H, x0 - 12 + 4 * wx, y0 - HM - HE + ly) & 15u);
        }
        __syncthreads();
        // horizontal 7-erosion; mask word j (columns x0-8+4j..) is PreFlood word j+1
        for (int i = threadIdx.x; i < EH * MWW; i += TPB) {
            const int ly = i / MWW, j = i % MWW;
            s_eh[ly][j] = erode7_word(s_pre[ly][j], s_pre[ly][j + 1], s_pre[ly][j + 2]);
        }
        __syncthreads();
    }
    // the mask as 0/1 bytes, zero outside the image; mask row ly is PreFlood row ly+3
    for (int i = threadIdx.x; i < MH * MWW; i += TPB) {
        const int ly = i / MWW, j = i % MWW, gx = x0 - 8 + 4 * j, gy = y0 - HM + ly;
        uint32_t m;
        if constexpr (SUBSENSE) {
            uint32_t e = s_eh[ly][j];
#pragma unroll
            for (int a = 1; a <= 2 * HE; ++a) e &= s_eh[ly + a][j];
            m = (load_mask_word(v.raw, W, H, gx, gy, 0u) | expand4(zone_bits4(v.hbits, W, H, gx, gy)) | e) &
                in_image_bytes(W, H, gx, gy);
        } else {
            m = load_mask_word(v.raw, W, H, gx, gy, 0u);
        }
        s_mask[ly][j] = m & 0x01010101u;
    }
    __syncthreads();
    // column counts over k rows
    for (int i = threadIdx.x; i < PH * MWW; i += TPB) {
        const int ly = i / MWW, j = i % MWW;
        uint32_t c = 0;
        for (int a = HM + ly - r; a <= HM + ly + r; ++a) c += s_mask[a][j];
        s_col[ly][j] = c;
    }
    __syncthreads();
    // row counts over k columns: output word q (columns x0+4q..) starts at byte 4q+8 of the counts
    const uint32_t bias = (uint32_t)(128 - need) * 0x01010101u;
    for (int i = threadIdx.x; i < PH * OWW; i += TPB) {
        const int ly = i / OWW, q = i % OWW;
        uint32_t w[5];
#pragma unroll
        for (int u = 0; u < 5; ++u) w[u] = s_col[ly][q + u];
        uint32_t s = 0;
#pragma unroll
        for (int o = -HM; o <= HM; ++o) {
            const int b = 8 + o, wi = b >> 2, sh = b & 3;
            const uint32_t t = sh ? bytes_at(w[wi + 1], w[wi], (uint32_t)sh) : w[wi];
            s += (o >= -r && o <= r) ? t : 0u;
        }
        const uint32_t bits = ((s + bias) >> 7) & 0x01010101u;
        store_mask_word(v.last_fg, W, H, x0 + 4 * q, y0 + ly, bits * 255u);
    }
}

// dilate7(last_fg), blink masks (SUB:548-551), final-segmentation averages (SUB:553-554) and the
// output mask. addWeighted(.., CV_32F) is evaluated as the double-precision scalar form.
LV_D void dilate_final_body(const View& v, uint32_t bid, uint32_t nblk) {
    uint8_t* out_mask = v.out_mask;
    constexpr int H3 = 3;
    constexpr int FWW = (PW + 8) / 4;  // last_fg words: columns x0-4 .. x0+PW+3
    constexpr int LH = PH + 2 * H3;    // rows y0-3 .. y0+PH+2
    constexpr int OWW = PW / 4;
    __shared__ uint32_t s_fg[LH][FWW];
    __shared__ uint32_t s_h[LH][OWW];
    const int W = v.W, H = v.H;
    const int ntx = (W + PW - 1) / PW;
    const int tid = (int)bid;
    const int x0 = (tid % ntx) * PW, y0 = (tid / ntx) * PH;
    for (int i = threadIdx.x; i < LH * FWW; i += TPB) {
        const int ly = i / FWW, wx = i % FWW;
        s_fg[ly][wx] = load_mask_word(v.last_fg, W, H, x0 - 4 + 4 * wx, y0 - H3 + ly, 0u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < LH * OWW; i += TPB) {
        const int ly = i / OWW, q = i % OWW;
        s_h[ly][q] = dilate7_word(s_fg[ly][q], s_fg[ly][q + 1], s_fg[ly][q + 2]);
    }
    __syncthreads();
    const uint32_t frame_no = cur_frame(v), mvn = v.mv_avgs;
    const float aLT = 1.0f / (float)(frame_no < mvn ? frame_no : mvn);
    const float aST = 1.0f / (float)(frame_no < mvn / 4u ? frame_no : mvn / 4u);
    const double alt = (double)(1.0f - aLT), blt = (1.0 / 255) * (double)aLT;
    const double ast = (double)(1.0f - aST), bst = (1.0 / 255) * (double)aST;
    for (int i = threadIdx.x; i < PH * OWW; i += TPB) {
        const int ly = i / OWW, q = i % OWW, gx = x0 + 4 * q, gy = y0 + ly;
        if (gx >= W || gy >= H) continue;
        uint32_t d = s_h[ly][q];
#pragma unroll
        for (int a = 1; a <= 2 * H3; ++a) d |= s_h[ly + a][q];
        const uint32_t inv = ~d, fg = s_fg[ly + H3][q + 1];
        const size_t p = (size_t)gy * W + gx;
        if ((W & 3) == 0) {
            uint32_t* bl = (uint32_t*)(v.blinks + p);
            uint32_t* ldi = (uint32_t*)(v.last_dil_inv + p);
            *bl = *bl & *ldi & inv;
            *ldi = inv;
            float4* lt = (float4*)(v.FinLT + p);
            float4* st = (float4*)(v.FinST + p);
            float4 a = *lt, b = *st;
            const double f0 = (double)(fg & 255u), f1 = (double)((fg >> 8) & 255u);
            const double f2 = (double)((fg >> 16) & 255u), f3 = (double)(fg >> 24);
            a.x = (float)((double)a.x * alt + f0 * blt + 0.0);
            a.y = (float)((double)a.y * alt + f1 * blt + 0.0);
            a.z = (float)((double)a.z * alt + f2 * blt + 0.0);
            a.w = (float)((double)a.w * alt + f3 * blt + 0.0);
            b.x = (float)((double)b.x * ast + f0 * bst + 0.0);
            b.y = (float)((double)b.y * ast + f1 * bst + 0.0);
            b.z = (float)((double)b.z * ast + f2 * bst + 0.0);
            b.w = (float)((double)b.w * ast + f3 * bst + 0.0);
            *lt = a;
            *st = b;
            if (out_mask) {
                if (((uintptr_t)out_mask & 3u) == 0) *(uint32_t*)(out_mask + p) = fg;  // caller's buffer
                else
                    for (int u = 0; u < 4; ++u) out_mask[p + u] = (uint8_t)(fg >> (8 * u));
            }
        } else {
            for (int u = 0; u < 4 && gx + u < W; ++u) {
                const size_t pu = p + u;
                const uint8_t iu = (uint8_t)(inv >> (8 * u)), fu = (uint8_t)(fg >> (8 * u));
                G(v.blinks)[pu] = (uint8_t)(G(v.blinks)[pu] & G(v.last_dil_inv)[pu] & iu);
                G(v.last_dil_inv)[pu] = iu;
                const double f = (double)fu;
                G(v.FinLT)[pu] = (float)((double)G(v.FinLT)[pu] * alt + f * blt + 0.0);
                G(v.FinST)[pu] = (float)((double)G(v.FinST)[pu] * ast + f * bst + 0.0);
                if (out_mask) out_mask[pu] = fu;
            }
        }
    }
}

LV_D void copy_mask_body(const View& v, uint32_t bid, uint32_t nblk) {
    uint8_t* out_mask = v.out_mask;
    if (!out_mask) return;
    const uint32_t p = bid * TPB + threadIdx.x;
    if (p < v.npx) out_mask[p] = G(v.last_fg)[p];
}

// ----------------------------------------------------------------------------------------------
// frame-level controller (SUB:555-611)
// ----------------------------------------------------------------------------------------------
LV_D uint8_t sat_u8(float f) {
    const float r = rintf(f);  // round half to even, like cvRound
    return (uint8_t)(r < 0.f ? 0.f : r > 255.f ? 255.f : r);
}

LV_D bool ds_exact8(const View& v) { return v.W == 8 * v.dsW && v.H == 8 * v.dsH; }

template <int C>
LV_D void ds_update_cell(const View& v, uint32_t cell, const uint8_t m[C], uint32_t& contrib) {
    const uint32_t frame_no = cur_frame(v), mvn = v.mv_avgs;
    const float aL = 1.0f / (float)(frame_no < mvn ? frame_no : mvn);
    const float aS = 1.0f / (float)(frame_no < mvn / 4u ? frame_no : mvn / 4u);
    const float bL = 1.0f - aL, bS = 1.0f - aS;
    uint32_t best = 0;
    for (int c = 0; c < C; ++c) {
        const size_t di = (size_t)cell * C + c;
        const float lt = (float)m[c] * aL + v.dsLT[di] * bL;
        const float st = (float)m[c] * aS + v.dsST[di] * bS;
        v.dsLT[di] = lt;
        v.dsST[di] = st;
        const uint32_t d = (uint32_t)fabsf(st - lt);
        best = d > best ? d : best;
    }
    contrib = C == 1 ? best / 2u : best;
}

// channel sums of one cell row: the 8 pixels (8*C bytes) at `row`, read as C aligned 8-byte words
template <int C>
LV_D void cell_row_sum(const uint8_t* row, uint32_t sum[C]) {
    const uint2* q = (const uint2*)row;
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const uint2 w = q[j];
#pragma unroll
        for (int b = 0; b < 8; ++b) sum[(8 * j + b) % C] += ((b < 4 ? w.x : w.y) >> (8 * (b & 3))) & 0xFFu;
    }
}

// resize(INTER_AREA, W/8 x H/8) + accumulateWeighted LT/ST + per-cell |ST-LT| (SUB:568-583), one
// thread per cell in raster order. Frame sizes that are multiples of 8 (an integer factor: the
// rounded 8x8 block mean) read each cell row as C 8-byte words (bytes when the frame is not 8-byte
// aligned), so a wavefront's loads cover contiguous runs of 64 cells; other sizes take the general
// area weights.
template <int C>
LV_D void downsample_body(const View& v, uint32_t bid, uint32_t nblk) {
    const uint8_t* frame = v.frame;
    uint32_t contrib = 0;
    const uint32_t ncell = (uint32_t)(v.dsW * v.dsH);
    const uint32_t cell = bid * TPB + threadIdx.x;
    if (cell < ncell && ds_exact8(v)) {
        const int dx = (int)(cell % (uint32_t)v.dsW), dy = (int)(cell / (uint32_t)v.dsW);
        uint32_t sum[C];
#pragma unroll
        for (int c = 0; c < C; ++c) sum[c] = 0;
        const uint8_t* cell0 = frame + ((size_t)dy * 8 * v.W + (size_t)dx * 8) * C;
        if (((uintptr_t)frame & 7u) == 0) {
#pragma unroll
            for (int r = 0; r < 8; ++r) cell_row_sum<C>(cell0 + (size_t)r * v.W * C, sum);
        } else {
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int b = 0; b < 8 * C; ++b) sum[b % C] += cell0[(size_t)r * v.W * C + b];
        }
        uint8_t mc[C];
#pragma unroll
        for (int c = 0; c < C; ++c) mc[c] = sat_u8((float)sum[c] * (1.f / 64.f));
        ds_update_cell<C>(v, cell, mc, contrib);
    } else if (cell < ncell) {
        const int dx = (int)(cell % (uint32_t)v.dsW), dy = (int)(cell / (uint32_t)v.dsW);
        uint8_t mc[C];
        for (int c = 0; c < C; ++c) {
            uint8_t m;
            if (v.W % v.dsW == 0 && v.H % v.dsH == 0) {
                const int sx = v.W / v.dsW, sy = v.H / v.dsH;
                int sum = 0;
                for (int yy = 0; yy < sy; ++yy)
                    for (int xx = 0; xx < sx; ++xx)
                        sum += G(frame)[((size_t)(dy * sy + yy) * v.W + (dx * sx + xx)) * C + c];
                m = sat_u8((float)sum * (1.f / (float)(sx * sy)));
            } else {
                const double sx = (double)v.W / v.dsW, sy = (double)v.H / v.dsH;
                const double yb = dy * sy, ye = (dy + 1) * sy, xb = dx * sx, xe = (dx + 1) * sx;
                double acc = 0.0;
                for (int yy = (int)yb; yy < v.H && (double)yy < ye; ++yy) {
                    const double wy = fmin((double)(yy + 1), ye) - fmax((double)yy, yb);
                    for (int xx = (int)xb; xx < v.W && (double)xx < xe; ++xx) {
                        const double wx = fmin((double)(xx + 1), xe) - fmax((double)xx, xb);
                        acc += (double)G(frame)[((size_t)yy * v.W + xx) * C + c] * (wy * wx);
                    }
                }
                m = sat_u8((float)(acc / (sx * sy)));
            }
            mc[c] = m;
        }
        ds_update_cell<C>(v, cell, mc, contrib);
    }
    block_add<TPB>(contrib, v.sc->cd_part);
}

// single-workgroup decision: LUT adaptation, model reset, learning-rate caps, counters
LV_D void decide_body(const View& v, uint32_t bid, uint32_t nblk) {
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
LV_D void lobster_end_body(const View& v, uint32_t bid, uint32_t nblk) {
    if (threadIdx.x != 0) return;
    v.sc->frame_idx += 1u;
    v.sc->epoch += 1u;
}

// refreshModel (SUB:80-105; LOB:410-440). pending = 1: auto-reset request from k_decide (also
// T := 1 everywhere, SUB:592); otherwise an explicit call with the given parameters.
template <int C, int ALGO>
LV_D void refresh_pixel(const View& v, uint32_t p, int use_pending, int force_fg, uint32_t frame, uint32_t ev,
                        uint32_t nref, uint32_t start);

// A grid-stride sweep, so the per-step launch (normally nothing pending) stays small.
template <int C, int ALGO>
LV_D void refresh_body(const View& v, uint32_t bid, uint32_t nblk, int use_pending, float frac, int force_fg, uint32_t frame, uint32_t ev) {
    if (use_pending) {
        if (!v.sc->refresh_pending) return;
        frac = v.sc->refresh_frac;
        force_fg = v.sc->refresh_force_fg;
        frame = v.sc->refresh_frame;
        ev = v.sc->refresh_ev;
    }
    const uint32_t N = (uint32_t)v.N;
    const uint32_t nref = frac < 1.0f ? (uint32_t)(frac * (float)N) : N;
    uint32_t start = 0;
    if (frac < 1.0f) {
        uint32_t d[4];
        draw4(v.seed_lo, v.seed_hi, v.stream_id, frame, PH_REFRESH, ev, 0xFFFFFFFFu, 0, d);
        start = d[0] % N;
    }
    for (uint32_t p = bid * TPB + threadIdx.x; p < v.npx; p += nblk * TPB)
        refresh_pixel<C, ALGO>(v, p, use_pending, force_fg, frame, ev, nref, start);
}

template <int C, int ALGO>
LV_D void refresh_pixel(const View& v, uint32_t p, int use_pending, int force_fg, uint32_t frame, uint32_t ev,
                        uint32_t nref, uint32_t start) {
    if (use_pending && ALGO == kSuBSENSE) G(v.T)[p] = 1.0f;
    if (!G(v.roi)[p]) return;
    if (!(force_fg || !G(v.last_fg)[p])) return;
    const uint32_t N = (uint32_t)v.N;
    const int x = (int)(p % (uint32_t)v.W), y = (int)(p / (uint32_t)v.W);
    uint32_t d[4];
    for (uint32_t j = 0; j < nref; ++j) {
        if ((j & 3u) == 0) draw4(v.seed_lo, v.seed_hi, v.stream_id, frame, PH_REFRESH, ev, p, j >> 2, d);
        int sx, sy;
        sample_pos_7x7(d[j & 3u], x, y, v.W, v.H, sx, sy);
        const uint32_t sp = (uint32_t)(sy * v.W + sx);
        if (!(force_fg || !G(v.last_fg)[sp])) continue;
        uint32_t col[C], desc[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            col[c] = G(v.last_color)[(size_t)sp * C + c];
            if constexpr (ALGO == kLOBSTER) {
                uint32_t taps[4];
                taps_from_global<C>(v.last_color, v.W, sx, sy, c, taps);
                desc[c] = lbsp_threshold_sad(taps, col[c], v.sc->lut[col[c]]);
            } else {
                desc[c] = G(v.last_desc)[(size_t)sp * C + c];
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
    const bool in = G(v.roi)[p] && x > 2 && y > 2 && x < v.W - 2 && y < v.H - 2;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        uint16_t d = 0;
        if (in) {
            uint32_t taps[4];
            taps_from_global<C>(img, v.W, x, y, c, taps);
            const uint32_t ref = img[(size_t)p * C + c];
            d = (uint16_t)lbsp_threshold_sad(taps, ref, v.sc->lut[ref]);
        }
        G(v.last_desc)[(size_t)p * C + c] = d;
        G(v.last_color)[(size_t)p * C + c] = G(v.roi)[p] ? img[(size_t)p * C + c] : (uint8_t)0;
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
// LBSP with an absolute threshold over a whole image, through either form of the threshold test
template <int C, bool PACKED>
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
            const uint32_t ref = tile.px[((ly + 2) * FrameTile<C>::LW + (lx + 2)) * C + c];
            d = (uint16_t)(PACKED ? lbsp_threshold_packed(taps, ref, t) : lbsp_threshold_sad(taps, ref, t));
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

// sticky device error bits of each stream of a step, gathered for one copy to the host
__global__ void k_gather_errors(const View* __restrict__ views, int n, uint32_t* __restrict__ out) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) out[i] = __hip_atomic_load(&views[i].sc->err_flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ==============================================================================================
// kernel entry points. Single stream: the View by value, block id remapped per XCD. Batched: every
// stream of the batch in one launch; a workgroup finds its stream in a per-kernel prefix table of
// block counts (after the XCD remap, so each stream's tiles stay together) and dispatches on the
// stream's channel count.
// ==============================================================================================
struct BatchArgs {
    const View* views;
    const uint32_t* prefix;  // ns+1 block offsets
    int ns;
};

template <class F>
LV_D void for_batch_block(const BatchArgs& b, F&& f) {
    const uint32_t lb = (uint32_t)xcd_tile((int)blockIdx.x, (int)gridDim.x);
    int lo = 0, hi = b.ns;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (b.prefix[mid] <= lb) lo = mid;
        else hi = mid;
    }
    // private copy: the fields become uniform registers that stores to the model cannot alias
    const View v = b.views[lo];
    f(v, lb - b.prefix[lo], b.prefix[lo + 1] - b.prefix[lo]);
}

#define LV_SINGLE_BID ((uint32_t)xcd_tile((int)blockIdx.x, (int)gridDim.x)), gridDim.x

template <int C, int ALGO>
__global__ void __launch_bounds__(TPB) k_describe(View v) { describe_body<C, ALGO>(v, LV_SINGLE_BID); }
// occupancy 6 (at most 80 VGPRs, 26 KB of LDS per 256-thread block) fits both fetch depths without spilling
#define LV_MATCH_WAVES(KF) __attribute__((amdgpu_waves_per_eu(6)))
template <int C, int ALGO, int KF>
__global__ void __launch_bounds__(TPB) LV_MATCH_WAVES(KF) k_match(View v) { match_body<C, ALGO, KF>(v, LV_SINGLE_BID); }
// phase B holds 80 VGPRs at occupancy 6 as well
template <int C, int ALGO, int KF>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(6))) k_match_seq(View v) {
    match_seq_body<C, ALGO, KF>(v, LV_SINGLE_BID);
}
template <int C, int ALGO>
__global__ void __launch_bounds__(TPB) k_match_cont(View v) { match_cont_body<C, ALGO>(v, LV_SINGLE_BID); }
template <int C, int ALGO>
__global__ void __launch_bounds__(TPB) k_fixpoint(View v, uint32_t k) { fixpoint_body<C, ALGO>(v, LV_SINGLE_BID, k); }
template <int C, int ALGO>
__global__ void __launch_bounds__(256) k_fixpoint_tail(View v, uint32_t k0, uint32_t cap) {
    fixpoint_tail_body<C, ALGO>(v, 0u, 1u, k0, cap);
}

__global__ void __launch_bounds__(TPB) k_lobster_raw(View v) { lobster_raw_body(v, LV_SINGLE_BID); }
template <int C, int ALGO, bool FUSED>
__global__ void __launch_bounds__(TPB) k_update(View v) { update_body<C, ALGO, FUSED>(v, LV_SINGLE_BID); }
template <int C>
__global__ void __launch_bounds__(TPB) k_copy_last_color(View v) { copy_last_color_body<C>(v, LV_SINGLE_BID); }
__global__ void __launch_bounds__(TPB) k_zone_close(View v) { zone_close_body(v, LV_SINGLE_BID); }
__global__ void __launch_bounds__(TPB) k_zone_link(View v) { zone_link_body(v, LV_SINGLE_BID); }
__global__ void __launch_bounds__(TPB) k_zone_holes(View v) { zone_holes_body(v, LV_SINGLE_BID); }
template <bool SUB>
__global__ void __launch_bounds__(TPB) k_median(View v) { median_body<SUB>(v, LV_SINGLE_BID); }
__global__ void __launch_bounds__(TPB) k_dilate_final(View v) { dilate_final_body(v, LV_SINGLE_BID); }
__global__ void __launch_bounds__(TPB) k_copy_mask(View v) { copy_mask_body(v, LV_SINGLE_BID); }
template <int C>
__global__ void __launch_bounds__(TPB) k_downsample(View v) { downsample_body<C>(v, LV_SINGLE_BID); }
__global__ void __launch_bounds__(256) k_decide(View v) { decide_body(v, 0u, 1u); }
__global__ void __launch_bounds__(64) k_lobster_end(View v) { lobster_end_body(v, 0u, 1u); }
template <int C, int ALGO>
__global__ void __launch_bounds__(TPB) k_refresh(View v, int use_pending, float frac, int force_fg, uint32_t frame,
                                               uint32_t ev) {
    refresh_body<C, ALGO>(v, LV_SINGLE_BID, use_pending, frac, force_fg, frame, ev);
}

// batched entry points (SuBSENSE; streams of 1 and 3 channels may be mixed)
#define LV_BATCH_C(body, ...)                                                                \
    for_batch_block(b, [&](const View& v, uint32_t bid, uint32_t nblk) {                     \
        if (v.C == 3) body<3 __VA_OPT__(, ) __VA_ARGS__>(v, bid, nblk);                       \
        else body<1 __VA_OPT__(, ) __VA_ARGS__>(v, bid, nblk);                                \
    })
__global__ void __launch_bounds__(TPB) kb_describe(BatchArgs b) { LV_BATCH_C(describe_body, kSuBSENSE); }
template <int KF>
__global__ void __launch_bounds__(TPB) LV_MATCH_WAVES(KF) kb_match(BatchArgs b) {
    LV_BATCH_C(match_body, kSuBSENSE, KF, 3);
}
template <int KF>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(6))) kb_match_seq(BatchArgs b) {
    LV_BATCH_C(match_seq_body, kSuBSENSE, KF);
}
__global__ void __launch_bounds__(TPB) kb_match_cont(BatchArgs b) { LV_BATCH_C(match_cont_body, kSuBSENSE); }
__global__ void __launch_bounds__(TPB) kb_fixpoint(BatchArgs b, uint32_t k) {
    for_batch_block(b, [&](const View& v, uint32_t bid, uint32_t nblk) {
        if (v.C == 3) fixpoint_body<3, kSuBSENSE>(v, bid, nblk, k);
        else fixpoint_body<1, kSuBSENSE>(v, bid, nblk, k);
    });
}
__global__ void __launch_bounds__(256) kb_fixpoint_tail(BatchArgs b, uint32_t k0, uint32_t cap) {
    for_batch_block(b, [&](const View& v, uint32_t bid, uint32_t nblk) {
        if (v.C == 3) fixpoint_tail_body<3, kSuBSENSE>(v, bid, nblk, k0, cap);
        else fixpoint_tail_body<1, kSuBSENSE>(v, bid, nblk, k0, cap);
    });
}
__global__ void __launch_bounds__(TPB) kb_update(BatchArgs b) { LV_BATCH_C(update_body, kSuBSENSE, true); }
__global__ void __launch_bounds__(TPB) kb_zone_close(BatchArgs b) { for_batch_block(b, zone_close_body); }
__global__ void __launch_bounds__(TPB) kb_zone_link(BatchArgs b) { for_batch_block(b, zone_link_body); }
__global__ void __launch_bounds__(TPB) kb_zone_holes(BatchArgs b) { for_batch_block(b, zone_holes_body); }
__global__ void __launch_bounds__(TPB) kb_median(BatchArgs b) { for_batch_block(b, median_body<true>); }
__global__ void __launch_bounds__(TPB) kb_dilate_final(BatchArgs b) { for_batch_block(b, dilate_final_body); }
__global__ void __launch_bounds__(TPB) kb_downsample(BatchArgs b) { LV_BATCH_C(downsample_body); }
__global__ void __launch_bounds__(256) kb_decide(BatchArgs b) { for_batch_block(b, decide_body); }
__global__ void __launch_bounds__(TPB) kb_refresh(BatchArgs b) {
    for_batch_block(b, [&](const View& v, uint32_t bid, uint32_t nblk) {
        if (v.C == 3) refresh_body<3, kSuBSENSE>(v, bid, nblk, 1, 0.f, 0, 0u, 0u);
        else refresh_body<1, kSuBSENSE>(v, bid, nblk, 1, 0.f, 0, 0u, 0u);
    });
}

// ==============================================================================================
// host-side launchers
// ==============================================================================================
static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

uint32_t cont_stripe_slots(int W, int H) {
    const size_t tiles = (size_t)cdiv(W, TW) * cdiv(H, TH);
    return (uint32_t)((tiles + kContStripes - 1) / kContStripes * TPB);
}

void kernel_grids(const View& v, const StepOptions& opt, uint32_t g[kNumGrids]) {
    g[kGridTile] = cdiv(v.W, TW) * cdiv(v.H, TH);
    g[kGridPost] = cdiv(v.W, PW) * cdiv(v.H, PH);
    g[kGridZone] = cdiv((size_t)cdiv(v.W, ZS) * cdiv(v.H, ZS), ZPB);
    g[kGridPix] = cdiv(v.npx, TPB);
    g[kGridCell] = !v.lr_scaling ? 0u : cdiv((size_t)v.dsW * v.dsH, TPB);
    g[kGridOne] = 1u;
    g[kGridFix] = (uint32_t)opt.fixpoint_blocks;
    if (opt.fix_px_per_block > 0) g[kGridFix] = std::min(g[kGridFix], cdiv(v.npx, (size_t)opt.fix_px_per_block));
    g[kGridSweep] = std::min(cdiv(v.npx, TPB), 128u);
    g[kGridCont] = (uint32_t)opt.cont_blocks;
    if (opt.cont_px_per_block > 0) g[kGridCont] = std::min(g[kGridCont], cdiv(v.npx, (size_t)opt.cont_px_per_block));
}

template <int C, int ALGO>
static void frame_step_t(const View& v, const StepOptions& opt, hipStream_t st) {
    uint32_t g[kNumGrids];
    kernel_grids(v, opt, g);
    if (opt.timed_stage == 2) (void)hipEventRecord(opt.ev_begin, st);
    hipLaunchKernelGGL((k_describe<C, ALGO>), dim3(g[kGridTile]), dim3(TPB), 0, st, v);
    if (opt.timed_stage == 1) (void)hipEventRecord(opt.ev_begin, st);
    if (opt.fetch_depth == 4) hipLaunchKernelGGL((k_match<C, ALGO, 4>), dim3(g[kGridTile]), dim3(TPB), 0, st, v);
    else hipLaunchKernelGGL((k_match<C, ALGO, 2>), dim3(g[kGridTile]), dim3(TPB), 0, st, v);
    if (opt.timed_stage == 1) (void)hipEventRecord(opt.ev_end, st);
    // phases B and C finish what phase A left (scans cut at seq_limit, pixels with more than kRepl
    // overwritten slots); they exit at once when nothing was left
    if (opt.fetch_depth == 4) hipLaunchKernelGGL((k_match_seq<C, ALGO, 4>), dim3(g[kGridCont]), dim3(TPB), 0, st, v);
    else hipLaunchKernelGGL((k_match_seq<C, ALGO, 2>), dim3(g[kGridCont]), dim3(TPB), 0, st, v);
    hipLaunchKernelGGL((k_match_cont<C, ALGO>), dim3(g[kGridCont]), dim3(TPB), 0, st, v);
    for (uint32_t k = 1; k <= (uint32_t)opt.fixpoint_launches; ++k)
        hipLaunchKernelGGL((k_fixpoint<C, ALGO>), dim3(g[kGridFix]), dim3(TPB), 0, st, v, k);
    hipLaunchKernelGGL((k_fixpoint_tail<C, ALGO>), dim3(1), dim3(256), 0, st, v, (uint32_t)opt.fixpoint_launches + 1u,
                       (uint32_t)opt.tail_cap);
    if constexpr (ALGO == kSuBSENSE) {
        if (opt.timed_stage == 3) (void)hipEventRecord(opt.ev_begin, st);
        hipLaunchKernelGGL((k_update<C, ALGO, true>), dim3(g[kGridTile]), dim3(TPB), 0, st, v);
        if (opt.timed_stage == 3) (void)hipEventRecord(opt.ev_end, st);
        hipLaunchKernelGGL(k_zone_close, dim3(g[kGridZone]), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL(k_zone_link, dim3(g[kGridZone]), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL(k_zone_holes, dim3(g[kGridZone]), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL((k_median<true>), dim3(g[kGridPost]), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL(k_dilate_final, dim3(g[kGridPost]), dim3(TPB), 0, st, v);
        if (g[kGridCell]) hipLaunchKernelGGL((k_downsample<C>), dim3(g[kGridCell]), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL(k_decide, dim3(1), dim3(256), 0, st, v);
        hipLaunchKernelGGL((k_refresh<C, ALGO>), dim3(g[kGridSweep]), dim3(TPB), 0, st, v, 1, 0.f, 0, 0u, 0u);
    } else {
        hipLaunchKernelGGL(k_lobster_raw, dim3(g[kGridPix]), dim3(TPB), 0, st, v);
        if (opt.timed_stage == 3) (void)hipEventRecord(opt.ev_begin, st);
        hipLaunchKernelGGL((k_update<C, ALGO, false>), dim3(g[kGridTile]), dim3(TPB), 0, st, v);
        if (opt.timed_stage == 3) (void)hipEventRecord(opt.ev_end, st);
        hipLaunchKernelGGL((k_median<false>), dim3(g[kGridPost]), dim3(TPB), 0, st, v);
        if (v.out_mask) hipLaunchKernelGGL(k_copy_mask, dim3(g[kGridPix]), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL((k_copy_last_color<C>), dim3(cdiv((size_t)v.npx * C, TPB)), dim3(TPB), 0, st, v);
        hipLaunchKernelGGL(k_lobster_end, dim3(1), dim3(64), 0, st, v);
    }
    if (opt.timed_stage == 2) (void)hipEventRecord(opt.ev_end, st);
}

void frame_step(const View& v, const StepOptions& opt, hipStream_t st) {
    if (v.algo == kSuBSENSE) {
        if (v.C == 3) frame_step_t<3, kSuBSENSE>(v, opt, st);
        else frame_step_t<1, kSuBSENSE>(v, opt, st);
    } else {
        if (v.C == 3) frame_step_t<3, kLOBSTER>(v, opt, st);
        else frame_step_t<1, kLOBSTER>(v, opt, st);
    }
}

// one SuBSENSE step of every stream of a batch; d_views holds the streams' Views (with their
// per-frame inputs), prefix[k] the (ns+1) block offsets of grid kind k, total[k] their sums
void batch_step(const View* d_views, int ns, const uint32_t* const prefix[kNumGrids], const uint32_t total[kNumGrids],
                const StepOptions& opt, hipStream_t st) {
    auto args = [&](int kind) { return BatchArgs{d_views, prefix[kind], ns}; };
    if (opt.timed_stage == 2) (void)hipEventRecord(opt.ev_begin, st);
    hipLaunchKernelGGL(kb_describe, dim3(total[kGridTile]), dim3(TPB), 0, st, args(kGridTile));
    if (opt.timed_stage == 1) (void)hipEventRecord(opt.ev_begin, st);
    if (opt.fetch_depth == 4) hipLaunchKernelGGL(kb_match<4>, dim3(total[kGridTile]), dim3(TPB), 0, st, args(kGridTile));
    else hipLaunchKernelGGL(kb_match<2>, dim3(total[kGridTile]), dim3(TPB), 0, st, args(kGridTile));
    if (opt.timed_stage == 1) (void)hipEventRecord(opt.ev_end, st);
    if (opt.fetch_depth == 4) hipLaunchKernelGGL(kb_match_seq<4>, dim3(total[kGridCont]), dim3(TPB), 0, st, args(kGridCont));
    else hipLaunchKernelGGL(kb_match_seq<2>, dim3(total[kGridCont]), dim3(TPB), 0, st, args(kGridCont));
    hipLaunchKernelGGL(kb_match_cont, dim3(total[kGridCont]), dim3(TPB), 0, st, args(kGridCont));
    for (uint32_t k = 1; k <= (uint32_t)opt.fixpoint_launches; ++k)
        hipLaunchKernelGGL(kb_fixpoint, dim3(total[kGridFix]), dim3(TPB), 0, st, args(kGridFix), k);
    hipLaunchKernelGGL(kb_fixpoint_tail, dim3(total[kGridOne]), dim3(256), 0, st, args(kGridOne),
                       (uint32_t)opt.fixpoint_launches + 1u, (uint32_t)opt.tail_cap);
    if (opt.timed_stage == 3) (void)hipEventRecord(opt.ev_begin, st);
    hipLaunchKernelGGL(kb_update, dim3(total[kGridTile]), dim3(TPB), 0, st, args(kGridTile));
    if (opt.timed_stage == 3) (void)hipEventRecord(opt.ev_end, st);
    hipLaunchKernelGGL(kb_zone_close, dim3(total[kGridZone]), dim3(TPB), 0, st, args(kGridZone));
    hipLaunchKernelGGL(kb_zone_link, dim3(total[kGridZone]), dim3(TPB), 0, st, args(kGridZone));
    hipLaunchKernelGGL(kb_zone_holes, dim3(total[kGridZone]), dim3(TPB), 0, st, args(kGridZone));
    hipLaunchKernelGGL(kb_median, dim3(total[kGridPost]), dim3(TPB), 0, st, args(kGridPost));
    hipLaunchKernelGGL(kb_dilate_final, dim3(total[kGridPost]), dim3(TPB), 0, st, args(kGridPost));
    if (total[kGridCell]) hipLaunchKernelGGL(kb_downsample, dim3(total[kGridCell]), dim3(TPB), 0, st, args(kGridCell));
    hipLaunchKernelGGL(kb_decide, dim3(total[kGridOne]), dim3(256), 0, st, args(kGridOne));
    hipLaunchKernelGGL(kb_refresh, dim3(total[kGridSweep]), dim3(TPB), 0, st, args(kGridSweep));
    if (opt.timed_stage == 2) (void)hipEventRecord(opt.ev_end, st);
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

void launch_lbsp_abs(const uint8_t* img, int W, int H, int C, uint32_t t, uint16_t* out, bool packed, hipStream_t st) {
    const unsigned tiles = cdiv(W, TW) * cdiv(H, TH);
    if (packed) {
        if (C == 3) hipLaunchKernelGGL((k_lbsp_abs<3, true>), dim3(tiles), dim3(TPB), 0, st, img, W, H, t, out);
        else hipLaunchKernelGGL((k_lbsp_abs<1, true>), dim3(tiles), dim3(TPB), 0, st, img, W, H, t, out);
    } else {
        if (C == 3) hipLaunchKernelGGL((k_lbsp_abs<3, false>), dim3(tiles), dim3(TPB), 0, st, img, W, H, t, out);
        else hipLaunchKernelGGL((k_lbsp_abs<1, false>), dim3(tiles), dim3(TPB), 0, st, img, W, H, t, out);
    }
}

// the SuBSENSE hole-filling and median kernels alone, on whatever v.raw holds
void launch_postprocess_probe(const View& v, hipStream_t st) {
    uint32_t g[kNumGrids];
    kernel_grids(v, StepOptions{}, g);
    hipLaunchKernelGGL(k_zone_close, dim3(g[kGridZone]), dim3(TPB), 0, st, v);
    hipLaunchKernelGGL(k_zone_link, dim3(g[kGridZone]), dim3(TPB), 0, st, v);
    hipLaunchKernelGGL(k_zone_holes, dim3(g[kGridZone]), dim3(TPB), 0, st, v);
    hipLaunchKernelGGL((k_median<true>), dim3(g[kGridPost]), dim3(TPB), 0, st, v);
}

void launch_gather_errors(const View* d_views, int n, uint32_t* d_out, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_gather_errors, dim3(cdiv((size_t)n, 256)), dim3(256), 0, st, d_views, n, d_out);
}

void launch_rng_probe(const View& v, uint32_t frame, uint32_t phase, uint32_t ev, uint32_t n, uint32_t* out,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_rng_probe, dim3(cdiv(n, 256)), dim3(256), 0, st, v, frame, phase, ev, n, out);
}

}  // namespace lvbgs
