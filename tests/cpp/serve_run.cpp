// serve_run.cpp -- drives a batch and a server through the C ABI (include/litiv_bgs.h), so that the
// host code of both runs under the sanitizers when built by tests/cpp/Makefile.
//
// usage: serve_run <frames> <in0.raw> <in1.raw> <in2.raw> <out.raw>
// Streams: 0 = 96x72 RGB, 1 = 64x48 RGB, 2 = 80x60 grey, SuBSENSE defaults, stream_id = index.
// out.raw: the masks of the batch (stream-major, frame by frame), then of the server fed from pinned
// host memory (two ingest slots per stream), then three frames of stream 0 replayed on a stream
// re-opened after every stream was closed (pooled context, stream_id 7).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "litiv_bgs.h"

namespace {

const int kW[3] = {96, 64, 80}, kH[3] = {72, 48, 60}, kC[3] = {3, 3, 1};

int fail(const char* what, int rc) {
    std::fprintf(stderr, "serve_run: %s failed (%d)\n", what, rc);
    return 10;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s frames in0.raw in1.raw in2.raw out.raw\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]);
    std::vector<std::vector<uint8_t>> frames(3);
    for (int k = 0; k < 3; ++k) {
        frames[k].resize((size_t)kW[k] * kH[k] * kC[k] * n);
        FILE* f = std::fopen(argv[2 + k], "rb");
        if (!f || std::fread(frames[k].data(), 1, frames[k].size(), f) != frames[k].size()) return 3;
        std::fclose(f);
    }
    std::vector<uint8_t> out;
    int rc;
    // ---- batch over three contexts, two concurrent groups, graph replay ----
    {
        lvbgs_ctx* ctx[3] = {};
        uint8_t *d_frames[3] = {}, *d_mask[3] = {};
        for (int k = 0; k < 3; ++k) {
            lvbgs_params p;
            lvbgs_default_params(LVBGS_SUBSENSE, &p);
            p.stream_id = (uint32_t)k;
            if ((rc = lvbgs_create(LVBGS_SUBSENSE, &p, &ctx[k]))) return fail("create", rc);
            if ((rc = lvbgs_initialize(ctx[k], frames[k].data(), kW[k], kH[k], kC[k], 0, nullptr, 0)))
                return fail("initialize", rc);
            if (hipMalloc((void**)&d_frames[k], frames[k].size()) != hipSuccess ||
                hipMalloc((void**)&d_mask[k], (size_t)kW[k] * kH[k]) != hipSuccess ||
                hipMemcpy(d_frames[k], frames[k].data(), frames[k].size(), hipMemcpyHostToDevice) != hipSuccess)
                return fail("device buffers", -3);
        }
        lvbgs_batch* b = nullptr;
        if ((rc = lvbgs_batch_create(ctx, 3, &b))) return fail("batch_create", rc);
        lvbgs_batch_options o;
        lvbgs_batch_default_options(&o);
        o.use_graph = 1;
        o.groups = 2;
        if ((rc = lvbgs_batch_set_options(b, &o))) return fail("batch_set_options", rc);
        if ((rc = lvbgs_batch_set_stream_options(b, 1, 4, 0, 0))) return fail("batch_set_stream_options", rc);
        std::vector<std::vector<uint8_t>> masks(3);
        for (int t = 0; t < n; ++t) {
            const void* imgs[3];
            void* ms[3];
            for (int k = 0; k < 3; ++k) {
                imgs[k] = d_frames[k] + (size_t)kW[k] * kH[k] * kC[k] * t;
                ms[k] = d_mask[k];
            }
            if ((rc = lvbgs_batch_apply_device(b, imgs, ms, nullptr))) return fail("batch_apply", rc);
            if ((rc = lvbgs_batch_synchronize(b))) return fail("batch_synchronize", rc);
            for (int k = 0; k < 3; ++k) {
                const size_t np = (size_t)kW[k] * kH[k];
                masks[k].resize(masks[k].size() + np);
                if (hipMemcpy(masks[k].data() + masks[k].size() - np, d_mask[k], np, hipMemcpyDeviceToHost) != hipSuccess)
                    return fail("mask copy", -3);
            }
        }
        for (int k = 0; k < 3; ++k) out.insert(out.end(), masks[k].begin(), masks[k].end());
        lvbgs_batch_destroy(b);
        for (int k = 0; k < 3; ++k) {
            lvbgs_destroy(ctx[k]);
            (void)hipFree(d_frames[k]);
            (void)hipFree(d_mask[k]);
        }
    }
    // ---- server fed from pinned host memory ----
    {
        lvbgs_server_options so;
        lvbgs_server_default_options(&so);
        so.host_ring = 2;
        so.max_inflight = 2;
        lvbgs_server* s = nullptr;
        if ((rc = lvbgs_server_create(0, &so, &s))) return fail("server_create", rc);
        int id[3];
        uint8_t* pinned[3];
        uint8_t* d_masks[3];
        for (int k = 0; k < 3; ++k) {
            lvbgs_params p;
            lvbgs_default_params(LVBGS_SUBSENSE, &p);
            p.stream_id = (uint32_t)k;
            if ((rc = lvbgs_server_open_host(s, &p, frames[k].data(), 0, kW[k], kH[k], kC[k], &id[k])))
                return fail("server_open_host", rc);
            pinned[k] = (uint8_t*)lvbgs_host_alloc(frames[k].size());
            if (!pinned[k]) return fail("host_alloc", -4);
            std::memcpy(pinned[k], frames[k].data(), frames[k].size());
            if (hipMalloc((void**)&d_masks[k], (size_t)kW[k] * kH[k] * n) != hipSuccess) return fail("masks", -3);
        }
        std::vector<lvbgs_done> done(64);
        int completed = 0;
        for (int t = 0; t < n; ++t)
            for (int k = 0; k < 3; ++k) {
                const size_t np = (size_t)kW[k] * kH[k];
                for (;;) {
                    rc = lvbgs_server_submit_host(s, id[k], pinned[k] + np * kC[k] * t, 0, d_masks[k] + np * t, 0.0,
                                                  (uint64_t)(k * 1000 + t));
                    if (rc != LVBGS_ERR_BUSY) break;
                    if ((rc = lvbgs_server_pump(s)) < 0) return fail("pump", rc);
                    const int got = lvbgs_server_poll(s, done.data(), (int)done.size());
                    if (got < 0) return fail("poll", got);
                    for (int i = 0; i < got; ++i)
                        if (done[(size_t)i].status != LVBGS_OK) return fail("frame status", done[(size_t)i].status);
                    completed += got;
                }
                if (rc) return fail("submit_host", rc);
                if ((rc = lvbgs_server_pump(s)) < 0) return fail("pump", rc);
            }
        if ((rc = lvbgs_server_drain(s))) return fail("drain", rc);
        for (;;) {
            const int got = lvbgs_server_poll(s, done.data(), (int)done.size());
            if (got < 0) return fail("poll", got);
            if (got == 0) break;
            for (int i = 0; i < got; ++i)
                if (done[(size_t)i].status != LVBGS_OK) return fail("frame status", done[(size_t)i].status);
            completed += got;
        }
        if (completed != 3 * n) return fail("completion count", completed);
        for (int k = 0; k < 3; ++k) {
            const size_t np = (size_t)kW[k] * kH[k] * n;
            const size_t at = out.size();
            out.resize(at + np);
            if (hipMemcpy(out.data() + at, d_masks[k], np, hipMemcpyDeviceToHost) != hipSuccess) return fail("copy", -3);
        }
        // every stream closed, then stream 0's geometry re-opened: its context comes from the pool
        for (int k = 0; k < 3; ++k)
            if ((rc = lvbgs_server_close(s, id[k]))) return fail("server_close", rc);
        if ((rc = lvbgs_server_drain(s))) return fail("drain", rc);
        lvbgs_params p;
        lvbgs_default_params(LVBGS_SUBSENSE, &p);
        p.stream_id = 7;
        int rid = -1;
        if ((rc = lvbgs_server_open_host(s, &p, frames[0].data(), 0, kW[0], kH[0], kC[0], &rid)))
            return fail("reopen", rc);
        const size_t np0 = (size_t)kW[0] * kH[0];
        for (int t = 0; t < 3; ++t) {
            while ((rc = lvbgs_server_submit_host(s, rid, pinned[0] + np0 * kC[0] * t, 0, d_masks[0] + np0 * t, 0.0,
                                                  (uint64_t)(9000 + t))) == LVBGS_ERR_BUSY) {
                if (lvbgs_server_pump(s) < 0) return fail("pump", -1);
                (void)lvbgs_server_poll(s, done.data(), (int)done.size());
            }
            if (rc) return fail("submit_host after reopen", rc);
        }
        if ((rc = lvbgs_server_drain(s))) return fail("drain", rc);
        const size_t at = out.size();
        out.resize(at + np0 * 3);
        if (hipMemcpy(out.data() + at, d_masks[0], np0 * 3, hipMemcpyDeviceToHost) != hipSuccess) return fail("copy", -3);
        lvbgs_server_stats st;
        if ((rc = lvbgs_server_get_stats(s, &st))) return fail("stats", rc);
        std::printf("server: %llu steps, %llu frames done, %u pooled contexts\n", (unsigned long long)st.steps,
                    (unsigned long long)st.frames_done, st.pooled_contexts);
        lvbgs_server_destroy(s);
        for (int k = 0; k < 3; ++k) {
            lvbgs_host_free(pinned[k]);
            (void)hipFree(d_masks[k]);
        }
    }
    FILE* o = std::fopen(argv[5], "wb");
    if (!o || std::fwrite(out.data(), 1, out.size(), o) != out.size()) return 5;
    std::fclose(o);
    return 0;
}
