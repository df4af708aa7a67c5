// bgs_eval.hip -- binary classification counts of foreground masks against ground truth, on the
// device (include/litiv_bgs.h, lvbgs_binclassif_*): lv::BinClassif::accumulate
// (modules/datasets/src/metrics.cpp:21-50, labels of modules/datasets/include/litiv/datasets/metrics.hpp:23-28)
// without copying the masks to the host.
//
// A pixel whose ground truth is out of scope (85) or unknown (170), or that lies outside a given ROI
// (ROI value 0), is a don't-care. Otherwise the result is positive iff it is 255 and the ground truth
// is positive iff it is 255 (every other value, shadow 50 included, is negative): TP / FP / FN / TN;
// a positive result on shadow ground truth also counts as a special error. Without ground truth every
// pixel is a don't-care.
//
// Two kernels: each block counts a chunk of kChunk pixels (16 per thread, 16-byte loads when the
// rows are packed and the pointers aligned) into one partial record; one block then sums the
// partials in a fixed order into six 64-bit totals. The partials and totals live in a scratch
// buffer: allocated per call by lvbgs_binclassif_accumulate, or the file-scope `scratch` array (one
// copy per device) for lvbgs_binclassif_accumulate_noalloc. Every partial record and total is
// written before it is read, so the scratch needs no clearing between calls.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>

#include "../../include/litiv_bgs.h"

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 16;
constexpr uint32_t kChunk = kThreads * kPerThread;  // pixels per block
constexpr int kCounters = 6;                         // TP, TN, FP, FN, SE, DC
constexpr size_t kPartialBytes = 32;                 // one block's partial record (6 u32, padded)
constexpr size_t kTotalsBytes = 64;                  // the final totals (6 u64, padded)

// the no-alloc entry point's scratch: partial records of up to kMaxBlocks blocks, then the totals
constexpr size_t kScratchBytes = (size_t)512 << 10;
constexpr uint32_t kMaxBlocks = (uint32_t)((kScratchBytes - kTotalsBytes) / kPartialBytes);
static_assert((uint64_t)kMaxBlocks * kChunk == LVBGS_BINCLASSIF_NOALLOC_MAX_PIXELS,
              "the header's pixel limit is the scratch's capacity");

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

struct Geometry {
    const uint8_t* classif;
    const uint8_t* gt;
    const uint8_t* roi;  // may be null
    uint32_t width, npx;
    uint64_t stride;     // row stride of all three masks, bytes
};

// category counts of one pixel into c[TP, TN, FP, FN, SE, DC] (metrics.cpp:38-60)
__device__ __forceinline__ void count_px(uint32_t in, uint32_t gt, uint32_t roi, uint32_t c[kCounters]) {
    const bool valid = gt != 85u && gt != 170u && roi != 0u;
    const bool pin = in == 255u, pgt = gt == 255u;
    c[0] += (valid && pin && pgt) ? 1u : 0u;
    c[1] += (valid && !pin && !pgt) ? 1u : 0u;
    c[2] += (valid && pin && !pgt) ? 1u : 0u;
    c[3] += (valid && !pin && pgt) ? 1u : 0u;
    c[4] += (valid && pin && gt == 50u) ? 1u : 0u;
    c[5] += valid ? 0u : 1u;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// PACKED: rows are contiguous (stride == width) and the three pointers 16-byte aligned
template <bool PACKED>
__global__ void __launch_bounds__(kThreads) k_binclassif_partial(Geometry g, uint32_t* __restrict__ partials) {
    uint32_t c[kCounters] = {0, 0, 0, 0, 0, 0};
    const uint32_t first = blockIdx.x * kChunk + threadIdx.x * kPerThread;
    if (PACKED && first + kPerThread <= g.npx) {
        const uint4 a = *reinterpret_cast<const uint4*>(g.classif + first);
        const uint4 b = *reinterpret_cast<const uint4*>(g.gt + first);
        const uint4 r = g.roi ? *reinterpret_cast<const uint4*>(g.roi + first) : make_uint4(~0u, ~0u, ~0u, ~0u);
        const uint32_t aw[4] = {a.x, a.y, a.z, a.w}, bw[4] = {b.x, b.y, b.z, b.w}, rw[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int w = k >> 2, sh = (k & 3) * 8;
            count_px((aw[w] >> sh) & 255u, (bw[w] >> sh) & 255u, (rw[w] >> sh) & 255u, c);
        }
    } else {
        for (uint32_t i = first; i < first + kPerThread && i < g.npx; ++i) {
            const uint32_t y = i / g.width, x = i - y * g.width;
            const size_t off = (size_t)y * g.stride + x;
            count_px(g.classif[off], g.gt[off], g.roi ? g.roi[off] : 255u, c);
        }
    }
    __shared__ uint32_t s[kThreads / 64][kCounters];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
    for (int k = 0; k < kCounters; ++k) {
        const uint32_t t = wave_sum(c[k]);
        if (lane == 0) s[wave][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < kCounters) {
        uint32_t t = 0;
        for (int w = 0; w < kThreads / 64; ++w) t += s[w][threadIdx.x];
        partials[(size_t)blockIdx.x * (kPartialBytes / 4) + threadIdx.x] = t;
    }
}

// one block: totals[k] = sum of the partial records' counter k, in block order
__global__ void __launch_bounds__(kThreads) k_binclassif_total(const uint32_t* __restrict__ partials, uint32_t nblk,
                                                               unsigned long long* __restrict__ totals) {
    unsigned long long c[kCounters] = {0, 0, 0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < nblk; b += kThreads)
#pragma unroll
        for (int k = 0; k < kCounters; ++k) c[k] += partials[(size_t)b * (kPartialBytes / 4) + k];
    __shared__ unsigned long long s[kThreads][kCounters];
#pragma unroll
    for (int k = 0; k < kCounters; ++k) s[threadIdx.x][k] = c[k];
    __syncthreads();
    if (threadIdx.x < kCounters) {
        unsigned long long t = 0;
        for (int i = 0; i < kThreads; ++i) t += s[i][threadIdx.x];
        totals[threadIdx.x] = t;
    }
}

// argument checks shared by both entry points; returns LVBGS_OK or an error, and the block count
int check_args(const void* d_classif, int width, int height, size_t stride, lvbgs_binclassif* counts,
               uint32_t* nblk) {
    if (!d_classif || !counts || width <= 0 || height <= 0) return LVBGS_ERR_ARG;
    if (stride == 0) stride = (size_t)width;
    if (stride < (size_t)width) return LVBGS_ERR_ARG;
    const uint64_t npx = (uint64_t)width * (uint64_t)height;
    if (npx > 0xFFFFFFFFull - kChunk) return LVBGS_ERR_ARG;
    *nblk = (uint32_t)((npx + kChunk - 1) / kChunk);
    return LVBGS_OK;
}

// launches both kernels on `st` with the scratch at `buf` and copies the totals into `counts`
int run(const void* d_classif, const void* d_gt, const void* d_roi, int width, int height, size_t stride,
        lvbgs_binclassif* counts, hipStream_t st, uint8_t* buf, uint32_t nblk) {
    if (stride == 0) stride = (size_t)width;
    Geometry g;
    g.classif = (const uint8_t*)d_classif;
    g.gt = (const uint8_t*)d_gt;
    g.roi = (const uint8_t*)d_roi;
    g.width = (uint32_t)width;
    g.npx = (uint32_t)((uint64_t)width * (uint64_t)height);
    g.stride = stride;
    const bool packed = stride == (size_t)width && ((uintptr_t)g.classif % 16) == 0 && ((uintptr_t)g.gt % 16) == 0 &&
                        ((uintptr_t)g.roi % 16) == 0;
    uint32_t* partials = reinterpret_cast<uint32_t*>(buf);
    unsigned long long* totals = reinterpret_cast<unsigned long long*>(buf + (size_t)nblk * kPartialBytes);
    if (packed) hipLaunchKernelGGL(k_binclassif_partial<true>, dim3(nblk), dim3(kThreads), 0, st, g, partials);
    else hipLaunchKernelGGL(k_binclassif_partial<false>, dim3(nblk), dim3(kThreads), 0, st, g, partials);
    hipLaunchKernelGGL(k_binclassif_total, dim3(1), dim3(kThreads), 0, st, partials, nblk, totals);
    unsigned long long h[kCounters];
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(h, totals, sizeof(h), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return LVBGS_ERR_HIP;
    counts->tp += h[0];
    counts->tn += h[1];
    counts->fp += h[2];
    counts->fn += h[3];
    counts->se += h[4];
    counts->dc += h[5];
    return LVBGS_OK;
}

// per device: the device address of `scratch`, taken once with that device current (in host code
// the array's name is not a device address)
constexpr int kMaxDevices = 64;
std::atomic<void*> g_scratch_addr[kMaxDevices];

void* scratch_address(int dev) {
    void* p = g_scratch_addr[dev].load(std::memory_order_acquire);
    if (p) return p;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_addr[dev].store(p, std::memory_order_release);  // a racing first call stores the same value
    return p;
}

}  // namespace

extern "C" int lvbgs_binclassif_accumulate(const void* d_classif, const void* d_gt, const void* d_roi, int width,
                                           int height, size_t stride, lvbgs_binclassif* counts, void* stream) {
    uint32_t nblk = 0;
    int rc = check_args(d_classif, width, height, stride, counts, &nblk);
    if (rc) return rc;
    if (!d_gt) {  // metrics.cpp:26-29: no ground truth, every pixel is a don't-care
        counts->dc += (uint64_t)width * (uint64_t)height;
        return LVBGS_OK;
    }
    uint8_t* buf = nullptr;
    if (hipMalloc(&buf, (size_t)nblk * kPartialBytes + kTotalsBytes) != hipSuccess) return LVBGS_ERR_ALLOC;
    rc = run(d_classif, d_gt, d_roi, width, height, stride, counts, (hipStream_t)stream, buf, nblk);
    if (hipFree(buf) != hipSuccess && rc == LVBGS_OK) rc = LVBGS_ERR_HIP;
    return rc;
}

extern "C" int lvbgs_binclassif_accumulate_noalloc(const void* d_classif, const void* d_gt, const void* d_roi,
                                                   int width, int height, size_t stride, lvbgs_binclassif* counts,
                                                   void* stream) {
    uint32_t nblk = 0;
    int rc = check_args(d_classif, width, height, stride, counts, &nblk);
    if (rc) return rc;
    if ((uint64_t)width * (uint64_t)height > LVBGS_BINCLASSIF_NOALLOC_MAX_PIXELS) return LVBGS_ERR_TOO_LARGE;
    if (!d_gt) {
        counts->dc += (uint64_t)width * (uint64_t)height;
        return LVBGS_OK;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return LVBGS_ERR_HIP;
    uint8_t* buf = static_cast<uint8_t*>(scratch_address(dev));
    if (!buf) return LVBGS_ERR_HIP;
    return run(d_classif, d_gt, d_roi, width, height, stride, counts, (hipStream_t)stream, buf, nblk);
}
