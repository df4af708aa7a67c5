// bgs_serve.hip -- serving loop (include/litiv_bgs.h, lvbgs_server_*): SuBSENSE models resident on
// one device, continuous batching of whatever frames are waiting, admission against a device-memory
// budget. Host code only; the steps reuse the batched kernels of bgs_kernels.hip.
//
// Scheduling: every pump harvests finished steps (event queries) and, while fewer than
// max_inflight steps are queued on the device, launches one more step made of the oldest waiting
// frame of every stream that has one (round-robin start, at most max_batch streams). A stream may
// have frames in several queued steps: the steps run in order on one HIP stream, so each frame
// sees the model its predecessor left. Frames arriving while steps run join the next step.
//
// Failure isolation: after each step the sticky device error bits of its streams (Scalars::
// err_flags, e.g. the fix-point iteration cap) are gathered into one pinned copy. At harvest a
// stream whose bits are set has that frame, its later queued frames and its waiting frames
// completed with LVBGS_ERR_DEVICE and is closed by the server; its id stays reserved until
// lvbgs_server_close acknowledges the failure, then the context goes to the pool (re-initialization
// clears the bits). The other streams of the step are unaffected.
//
// Host ingest (options.host_ring > 0): every stream's arena holds host_ring frame slots. A frame
// submitted from host memory is copied at once into the stream's next free slot on the server's
// copy stream, with an event after it, so the copy overlaps the steps already running. A step
// takes a host frame only once its copy has completed (its stream sits that step out otherwise),
// so the copies queued last never hold up the frames already in HBM; the step still waits for the
// event on the device. Slots are freed in order as the stream's frames complete; a stream with
// every slot busy refuses more (LVBGS_ERR_BUSY).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "../../include/litiv_bgs.h"
#include "bgs_ctx.hpp"
#include "bgs_kernels.hpp"

using namespace lvbgs;

namespace {

struct Request {
    const uint8_t* img;
    uint8_t* mask;
    double lr;
    uint64_t tag;
    int ring_slot;  // host-ingest slot holding img (-1: a caller's device frame)
};

struct Slot {
    lvbgs_ctx* ctx = nullptr;
    std::deque<Request> pending;
    int inflight = 0;  // frames of this stream in queued steps
    bool open = false;
    bool closing = false;
    bool failed = false;  // a step reported a device error for this stream: closed by the server
    bool ready = false;  // its initialization (queued on the context's stream) has completed
    int ring_next = 0;   // next host-ingest slot to fill
    int ring_busy = 0;   // host-ingest slots holding frames not completed yet
};

struct StepFrame {
    int stream;
    uint64_t tag;
    int pos;        // the stream's View slot in the step (index into the gathered error bits)
    int ring_slot;  // host-ingest slot of the frame (-1: none)
};

struct QueuedStep {
    hipEvent_t done;
    uint32_t* h_err;  // pinned: error bits of the step's View slots, written after the step
    std::vector<StepFrame> frames;
};

}  // namespace

struct lvbgs_server {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    lvbgs_server_options o{};
    StepOptions opt{};
    size_t used = 0;  // arena bytes of open and pooled contexts
    std::vector<Slot> slots;
    std::vector<hipEvent_t> init_ev;  // per slot id: end of the stream's initialization
    std::vector<std::vector<hipEvent_t>> ring_ev;  // per slot id: end of each host-ingest copy
    // host-ingest copies (one stream: four, dealt by stream id, measured far slower at 256 streams,
    // p50 1.8 s against 4.1 ms, with the runtime's four hardware queues shared by every stream)
    static constexpr int kCopyStreams = 1;
    hipStream_t cstream[kCopyStreams] = {};
    std::vector<int> free_ids;
    std::vector<lvbgs_ctx*> pool;  // closed contexts whose arenas are kept for reuse
    View* d_views = nullptr;
    uint32_t* d_prefix = nullptr;
    std::vector<uint8_t*> stage;  // pinned staging: max_batch Views + prefix tables
    std::vector<hipEvent_t> stage_ev;
    std::vector<bool> stage_live;
    int stage_next = 0;
    std::vector<hipEvent_t> ev_free;
    std::vector<uint32_t*> err_free;  // pinned error-bit buffers (max_batch words), one per queued step
    uint32_t* d_err = nullptr;        // device error bits of the last step (max_batch words)
    std::deque<QueuedStep> queued;
    std::vector<lvbgs_done> done;
    int rr = 0;
    lvbgs_server_stats st{};
    // a step's frames in `groups` groups stepped concurrently (fork/join on `stream`)
    static constexpr int kMaxGroups = 4;
    int groups = 1;
    hipStream_t gstream[kMaxGroups] = {};  // gstream[0] is `stream`
    hipEvent_t ev_fork = nullptr, ev_join[kMaxGroups] = {};
};

static int srv_err(lvbgs_server* s, int code, const char* msg) {
    if (s) s->err = msg;
    return code;
}

// grid tables of a step: group g, kind k at row g * kNumGrids + k, max_batch + 1 entries per row
static size_t table_words(int max_batch) {
    return (size_t)lvbgs_server::kMaxGroups * kNumGrids * (size_t)(max_batch + 1);
}

static size_t stage_bytes(int max_batch) {
    return sizeof(View) * (size_t)max_batch + sizeof(uint32_t) * table_words(max_batch);
}

extern "C" int lvbgs_server_default_options(lvbgs_server_options* o) {
    if (!o) return LVBGS_ERR_ARG;
    std::memset(o, 0, sizeof(*o));
    o->hbm_budget = (size_t)250 << 30;
    o->max_inflight = 2;
    o->max_batch = 256;
    o->host_ring = 0;
    lvbgs_batch_default_options(&o->launch);
    o->launch.use_graph = 0;
    return LVBGS_OK;
}

static void server_release(lvbgs_server* s) {
    for (auto& sl : s->slots)
        if (sl.ctx) lvbgs_destroy(sl.ctx);
    for (auto* c : s->pool) lvbgs_destroy(c);
    if (s->d_views) (void)hipFree(s->d_views);
    if (s->d_prefix) (void)hipFree(s->d_prefix);
    if (s->d_err) (void)hipFree(s->d_err);
    for (auto* p : s->err_free) (void)hipHostFree(p);
    for (auto& q : s->queued) (void)hipHostFree(q.h_err);
    for (auto* p : s->stage) (void)hipHostFree(p);
    for (auto e : s->stage_ev) (void)hipEventDestroy(e);
    for (auto e : s->ev_free) (void)hipEventDestroy(e);
    for (auto& q : s->queued) (void)hipEventDestroy(q.done);
    for (auto e : s->init_ev) (void)hipEventDestroy(e);
    for (auto& v : s->ring_ev)
        for (auto e : v) (void)hipEventDestroy(e);
    for (auto cs : s->cstream)
        if (cs) (void)hipStreamDestroy(cs);
    for (int g = 1; g < lvbgs_server::kMaxGroups; ++g)
        if (s->gstream[g]) (void)hipStreamDestroy(s->gstream[g]);
    for (auto e : s->ev_join)
        if (e) (void)hipEventDestroy(e);
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

extern "C" int lvbgs_server_create(int device, const lvbgs_server_options* opts, lvbgs_server** out) {
    if (!out) return LVBGS_ERR_ARG;
    *out = nullptr;
    lvbgs_server_options o;
    lvbgs_server_default_options(&o);
    if (opts) o = *opts;
    const lvbgs_batch_options& L = o.launch;
    if (o.max_inflight < 1 || o.max_inflight > 16 || o.max_batch < 1 || o.max_batch > 4096 || L.seq_limit < 1 ||
        L.seq_limit > 255 || L.cont_blocks < 1 || L.fixpoint_launches < 0 || L.fixpoint_launches > 64 ||
        L.fixpoint_blocks < 1 || L.cont_px_per_block < 0 || L.fix_px_per_block < 0 ||
        (L.fetch_depth != 2 && L.fetch_depth != 4) || L.groups < 1 || L.groups > lvbgs_server::kMaxGroups ||
        o.host_ring < 0 || o.host_ring > 64)
        return LVBGS_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return LVBGS_ERR_HIP;
    lvbgs_server* s = new lvbgs_server();
    s->device = device;
    s->o = o;
    s->opt.seq_limit = L.seq_limit;
    s->opt.cont_blocks = L.cont_blocks;
    s->opt.cont_px_per_block = L.cont_px_per_block;
    s->opt.fixpoint_launches = L.fixpoint_launches;
    s->opt.fixpoint_blocks = L.fixpoint_blocks;
    s->opt.fix_px_per_block = L.fix_px_per_block;
    s->opt.fetch_depth = L.fetch_depth;
    s->groups = L.groups;
    bool ok = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&s->d_views, sizeof(View) * o.max_batch) == hipSuccess &&
              hipMalloc(&s->d_prefix, sizeof(uint32_t) * table_words(o.max_batch)) == hipSuccess &&
              hipMalloc(&s->d_err, sizeof(uint32_t) * (size_t)o.max_batch) == hipSuccess;
    s->gstream[0] = s->stream;
    for (int g = 1; ok && g < s->groups; ++g)
        ok = hipStreamCreateWithFlags(&s->gstream[g], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&s->ev_join[g], hipEventDisableTiming) == hipSuccess;
    if (ok && s->groups > 1) ok = hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && o.host_ring > 0 && k < lvbgs_server::kCopyStreams; ++k)
        ok = hipStreamCreateWithFlags(&s->cstream[k], hipStreamNonBlocking) == hipSuccess;
    const int nstage = o.max_inflight + 1;
    for (int i = 0; ok && i < nstage; ++i) {
        uint8_t* p = nullptr;
        hipEvent_t e = nullptr;
        ok = hipHostMalloc((void**)&p, stage_bytes(o.max_batch)) == hipSuccess &&
             hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
        if (p) s->stage.push_back(p);
        if (e) s->stage_ev.push_back(e);
        s->stage_live.push_back(false);
    }
    for (int i = 0; ok && i < o.max_inflight; ++i) {
        hipEvent_t e = nullptr;
        uint32_t* h = nullptr;
        ok = hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess &&
             hipHostMalloc((void**)&h, sizeof(uint32_t) * (size_t)o.max_batch) == hipSuccess;
        if (e) s->ev_free.push_back(e);
        if (h) s->err_free.push_back(h);
    }
    if (!ok) {
        server_release(s);
        delete s;
        return LVBGS_ERR_HIP;
    }
    *out = s;
    return LVBGS_OK;
}

extern "C" const char* lvbgs_server_last_error(const lvbgs_server* s) { return s ? s->err.c_str() : "null server"; }
extern "C" void* lvbgs_server_stream(lvbgs_server* s) { return s ? (void*)s->stream : nullptr; }

static bool same_model_params(const lvbgs_params& a, const lvbgs_params& b) {
    return a.desc_dist_threshold == b.desc_dist_threshold && a.color_dist_threshold == b.color_dist_threshold &&
           a.bg_samples == b.bg_samples && a.required_bg_samples == b.required_bg_samples &&
           a.samples_for_mv_avgs == b.samples_for_mv_avgs && a.lbsp_threshold_offset == b.lbsp_threshold_offset &&
           a.rel_lbsp_threshold == b.rel_lbsp_threshold && a.median_blur_kernel == b.median_blur_kernel;
}

// a stream whose step reported a device error: its waiting frames complete with the error and the
// server closes it
static void fail_stream(lvbgs_server* s, int id) {
    Slot& sl = s->slots[(size_t)id];
    if (sl.failed) return;
    sl.failed = true;
    sl.closing = true;
    for (const Request& r : sl.pending) {
        s->done.push_back(lvbgs_done{r.tag, id, LVBGS_ERR_DEVICE});
        if (r.ring_slot >= 0) --sl.ring_busy;
    }
    s->st.frames_done += sl.pending.size();
    sl.pending.clear();
}

// finished steps -> completions; closing streams whose work is done go to the pool
static int harvest(lvbgs_server* s, bool wait_one) {
    while (!s->queued.empty()) {
        QueuedStep& q = s->queued.front();
        const hipError_t e = wait_one ? hipEventSynchronize(q.done) : hipEventQuery(q.done);
        wait_one = false;
        if (e == hipErrorNotReady) break;
        if (e != hipSuccess) return srv_err(s, LVBGS_ERR_HIP, "step failed");
        for (auto& f : q.frames) {
            Slot& sl = s->slots[(size_t)f.stream];
            const bool bad = (q.h_err[f.pos] & kErrFixpointCap) != 0 || sl.failed;
            s->done.push_back(lvbgs_done{f.tag, f.stream, bad ? LVBGS_ERR_DEVICE : LVBGS_OK});
            --sl.inflight;
            if (f.ring_slot >= 0) --sl.ring_busy;
            if (bad) fail_stream(s, f.stream);
        }
        s->st.frames_done += q.frames.size();
        s->ev_free.push_back(q.done);
        s->err_free.push_back(q.h_err);
        s->queued.pop_front();
    }
    for (size_t i = 0; i < s->slots.size(); ++i) {
        Slot& sl = s->slots[i];
        if (sl.open && sl.closing && !sl.failed && sl.pending.empty() && sl.inflight == 0) {
            s->pool.push_back(sl.ctx);
            sl = Slot();
            s->free_ids.push_back((int)i);
        }
    }
    return LVBGS_OK;
}

static int launch_step(lvbgs_server* s) {
    const int nslots = (int)s->slots.size();
    std::vector<int> members;
    for (int k = 0; k < nslots && (int)members.size() < s->o.max_batch; ++k) {
        const int i = (s->rr + k) % nslots;
        Slot& sl = s->slots[(size_t)i];
        if (!sl.open || sl.pending.empty()) continue;
        if (!sl.ready) {
            const hipError_t e = hipEventQuery(s->init_ev[(size_t)i]);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) return srv_err(s, LVBGS_ERR_HIP, "stream initialization failed");
            sl.ready = true;
        }
        // a frame still being copied from host memory waits for a later step, so that the copies
        // queued last do not hold up every frame of this one
        const int rs = sl.pending.front().ring_slot;
        if (rs >= 0) {
            const hipError_t e = hipEventQuery(s->ring_ev[(size_t)i][(size_t)rs]);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) return srv_err(s, LVBGS_ERR_HIP, "ingest copy failed");
        }
        members.push_back(i);
    }
    if (members.empty()) return LVBGS_OK;
    s->rr = (members.back() + 1) % nslots;
    const int n = (int)members.size();
    const int si = s->stage_next;
    s->stage_next = (si + 1) % (int)s->stage.size();
    if (s->stage_live[(size_t)si] && hipEventSynchronize(s->stage_ev[(size_t)si]) != hipSuccess)
        return srv_err(s, LVBGS_ERR_HIP, "staging");
    View* hv = (View*)s->stage[(size_t)si];
    uint32_t* hp = (uint32_t*)(s->stage[(size_t)si] + sizeof(View) * (size_t)s->o.max_batch);
    QueuedStep q;
    q.frames.reserve((size_t)n);
    std::vector<Request> req((size_t)n);
    for (int m = 0; m < n; ++m) {
        Slot& sl = s->slots[(size_t)members[(size_t)m]];
        req[(size_t)m] = sl.pending.front();
        sl.pending.pop_front();
        ++sl.inflight;
        ++sl.ctx->apply_count;
        q.frames.push_back(StepFrame{members[(size_t)m], req[(size_t)m].tag, 0, req[(size_t)m].ring_slot});
        // a frame copied from host memory: the step waits for its copy
        if (req[(size_t)m].ring_slot >= 0 &&
            hipStreamWaitEvent(s->stream, s->ring_ev[(size_t)members[(size_t)m]][(size_t)req[(size_t)m].ring_slot],
                               0) != hipSuccess)
            return srv_err(s, LVBGS_ERR_HIP, "ingest wait");
    }
    // groups: largest frames first, dealt round-robin; a group is a contiguous range of View slots
    const int G = std::min(s->groups, n);
    std::vector<int> by_size((size_t)n);
    for (int m = 0; m < n; ++m) by_size[(size_t)m] = m;
    std::stable_sort(by_size.begin(), by_size.end(), [&](int x, int y) {
        return s->slots[(size_t)members[(size_t)x]].ctx->npx > s->slots[(size_t)members[(size_t)y]].ctx->npx;
    });
    const size_t row = (size_t)n + 1;
    int first[lvbgs_server::kMaxGroups] = {}, count[lvbgs_server::kMaxGroups] = {};
    int pos = 0;
    for (int g = 0; g < G; ++g) {
        first[g] = pos;
        for (int k = 0; k < kNumGrids; ++k) hp[((size_t)g * kNumGrids + k) * row] = 0;
        for (int j = g; j < n; j += G, ++pos) {
            const int m = by_size[(size_t)j];
            const Request& r = req[(size_t)m];
            q.frames[(size_t)m].pos = pos;
            lvbgs_ctx* c = s->slots[(size_t)members[(size_t)m]].ctx;
            View v = c->v;
            v.frame = r.img;
            v.out_mask = r.mask;
            v.lr = r.lr;
            v.seq_limit = (uint32_t)s->opt.seq_limit;
            hv[pos] = v;
            uint32_t gk[kNumGrids];
            kernel_grids(c->v, s->opt, gk);
            const int jj = pos - first[g];
            for (int k = 0; k < kNumGrids; ++k) {
                uint32_t* t = &hp[((size_t)g * kNumGrids + k) * row];
                t[jj + 1] = t[jj] + gk[k];
            }
        }
        count[g] = pos - first[g];
    }
    if (hipMemcpyAsync(s->d_views, hv, sizeof(View) * (size_t)n, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(s->d_prefix, hp, sizeof(uint32_t) * (size_t)G * kNumGrids * row, hipMemcpyHostToDevice,
                       s->stream) != hipSuccess ||
        hipEventRecord(s->stage_ev[(size_t)si], s->stream) != hipSuccess)
        return srv_err(s, LVBGS_ERR_HIP, "step upload");
    s->stage_live[(size_t)si] = true;
    if (G > 1) {
        if (hipEventRecord(s->ev_fork, s->stream) != hipSuccess) return srv_err(s, LVBGS_ERR_HIP, "fork");
        for (int g = 1; g < G; ++g)
            if (hipStreamWaitEvent(s->gstream[g], s->ev_fork, 0) != hipSuccess) return srv_err(s, LVBGS_ERR_HIP, "fork");
    }
    for (int g = 0; g < G; ++g) {
        const uint32_t* prefix[kNumGrids];
        uint32_t total[kNumGrids];
        for (int k = 0; k < kNumGrids; ++k) {
            prefix[k] = s->d_prefix + ((size_t)g * kNumGrids + k) * row;
            total[k] = hp[((size_t)g * kNumGrids + k) * row + (size_t)count[g]];
        }
        batch_step(s->d_views + first[g], count[g], prefix, total, s->opt, s->gstream[g]);
    }
    for (int g = 1; g < G; ++g)
        if (hipEventRecord(s->ev_join[g], s->gstream[g]) != hipSuccess ||
            hipStreamWaitEvent(s->stream, s->ev_join[g], 0) != hipSuccess)
            return srv_err(s, LVBGS_ERR_HIP, "join");
    // the step's error bits, copied out before its completion event
    q.done = s->ev_free.back();
    s->ev_free.pop_back();
    q.h_err = s->err_free.back();
    s->err_free.pop_back();
    launch_gather_errors(s->d_views, n, s->d_err, s->stream);
    if (hipMemcpyAsync(q.h_err, s->d_err, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess ||
        hipEventRecord(q.done, s->stream) != hipSuccess || hipGetLastError() != hipSuccess) {
        s->ev_free.push_back(q.done);
        s->err_free.push_back(q.h_err);
        return srv_err(s, LVBGS_ERR_HIP, "step launch");
    }
    s->queued.push_back(std::move(q));
    ++s->st.steps;
    s->st.frames_launched += (uint64_t)n;
    return LVBGS_OK;
}

extern "C" int lvbgs_server_pump(lvbgs_server* s) {
    if (!s) return LVBGS_ERR_ARG;
    (void)hipSetDevice(s->device);
    int rc = harvest(s, false);
    if (rc) return rc;
    int launched = 0;
    while ((int)s->queued.size() < s->o.max_inflight) {
        const uint64_t before = s->st.steps;
        if ((rc = launch_step(s))) return rc;
        if (s->st.steps == before) break;
        ++launched;
    }
    return launched;
}

extern "C" int lvbgs_server_poll(lvbgs_server* s, lvbgs_done* out, int cap) {
    if (!s || (!out && cap > 0) || cap < 0) return LVBGS_ERR_ARG;
    const int n = std::min(cap, (int)s->done.size());
    if (n > 0) {
        std::memcpy(out, s->done.data(), sizeof(lvbgs_done) * (size_t)n);
        s->done.erase(s->done.begin(), s->done.begin() + n);
    }
    return n;
}

extern "C" int lvbgs_server_drain(lvbgs_server* s) {
    if (!s) return LVBGS_ERR_ARG;
    (void)hipSetDevice(s->device);
    while (true) {
        int rc = lvbgs_server_pump(s);
        if (rc < 0) return rc;
        if (s->queued.empty()) {
            bool waiting = false;
            for (auto& sl : s->slots) waiting = waiting || (sl.open && !sl.pending.empty());
            if (!waiting) break;
            // nothing could launch: the waiting frames' copies (or initializations) are still running
            for (auto cs : s->cstream)
                if (cs && hipStreamSynchronize(cs) != hipSuccess) return srv_err(s, LVBGS_ERR_HIP, "ingest copy");
            continue;
        }
        if ((rc = harvest(s, true))) return rc;
    }
    return harvest(s, false);
}

// admits a stream initialized from img0: a device frame, or host memory with row stride host_stride
static int server_open(lvbgs_server* s, const lvbgs_params* params, const void* img0, bool host, size_t host_stride,
                       int width, int height, int channels, int* stream_out) {
    if (!s || !params || !img0 || !stream_out || width <= 0 || height <= 0 || (channels != 1 && channels != 3))
        return s ? srv_err(s, LVBGS_ERR_ARG, "invalid stream parameters") : LVBGS_ERR_ARG;
    (void)hipSetDevice(s->device);
    lvbgs_params p = *params;
    p.device = s->device;
    const size_t need = lvbgs_arena_bytes(LVBGS_SUBSENSE, p, width, height, channels, s->o.host_ring);
    // a pooled context of the same geometry and model parameters keeps its arena (no allocation)
    lvbgs_ctx* c = nullptr;
    for (size_t i = 0; i < s->pool.size(); ++i)
        if (s->pool[i]->arena_size == need && s->pool[i]->W == width && s->pool[i]->H == height &&
            s->pool[i]->C == channels && same_model_params(s->pool[i]->p, p)) {
            c = s->pool[i];
            s->pool.erase(s->pool.begin() + (long)i);
            break;
        }
    if (!c) {
        while (s->used + need > s->o.hbm_budget && !s->pool.empty()) {
            s->used -= s->pool.back()->arena_size;
            lvbgs_destroy(s->pool.back());
            s->pool.pop_back();
        }
        if (s->used + need > s->o.hbm_budget) return srv_err(s, LVBGS_ERR_BUDGET, "model does not fit the budget");
        const int rc = lvbgs_create(LVBGS_SUBSENSE, &p, &c);
        if (rc) return srv_err(s, rc, "invalid stream parameters");
        c->ring_frames = s->o.host_ring;
    } else {
        c->p.rng_seed = p.rng_seed;
        c->p.stream_id = p.stream_id;
    }
    int id;
    if (!s->free_ids.empty()) {
        id = s->free_ids.back();
        s->free_ids.pop_back();
    } else {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            if (c->arena_size) s->pool.push_back(c);
            else lvbgs_destroy(c);
            return srv_err(s, LVBGS_ERR_HIP, "event");
        }
        std::vector<hipEvent_t> rev((size_t)s->o.host_ring, nullptr);
        bool ok = true;
        for (auto& x : rev) ok = ok && hipEventCreateWithFlags(&x, hipEventDisableTiming) == hipSuccess;
        if (!ok) {
            for (auto x : rev)
                if (x) (void)hipEventDestroy(x);
            (void)hipEventDestroy(e);
            if (c->arena_size) s->pool.push_back(c);
            else lvbgs_destroy(c);
            return srv_err(s, LVBGS_ERR_HIP, "event");
        }
        id = (int)s->slots.size();
        s->slots.emplace_back();
        s->init_ev.push_back(e);
        s->ring_ev.push_back(std::move(rev));
    }
    // initialization is queued on the context's own stream (no host wait): the stream joins steps
    // once its event has completed
    const size_t before = c->arena_size;
    const int rc = host ? lvbgs_initialize_async_host(c, img0, width, height, channels, host_stride,
                                                      s->init_ev[(size_t)id])
                        : lvbgs_initialize_async(c, img0, width, height, channels, s->init_ev[(size_t)id]);
    if (rc) {
        s->err = lvbgs_last_error(c);
        s->used -= before;
        lvbgs_destroy(c);
        s->free_ids.push_back(id);
        return rc;
    }
    s->used += c->arena_size - before;
    Slot& sl = s->slots[(size_t)id];
    sl = Slot();
    sl.ctx = c;
    sl.open = true;
    *stream_out = id;
    return LVBGS_OK;
}

extern "C" int lvbgs_server_open(lvbgs_server* s, const lvbgs_params* params, const void* d_img0, int width,
                                 int height, int channels, int* stream_out) {
    return server_open(s, params, d_img0, false, 0, width, height, channels, stream_out);
}

extern "C" int lvbgs_server_open_host(lvbgs_server* s, const lvbgs_params* params, const void* img0, size_t stride,
                                      int width, int height, int channels, int* stream_out) {
    return server_open(s, params, img0, true, stride, width, height, channels, stream_out);
}

extern "C" int lvbgs_server_close(lvbgs_server* s, int stream) {
    if (s && stream >= 0 && stream < (int)s->slots.size() && s->slots[(size_t)stream].failed) {
        // closed by the server after a device error: the call releases the id
        s->slots[(size_t)stream].failed = false;
        return harvest(s, false);
    }
    if (!s || stream < 0 || stream >= (int)s->slots.size() || !s->slots[(size_t)stream].open ||
        s->slots[(size_t)stream].closing)
        return s ? srv_err(s, LVBGS_ERR_ARG, "unknown stream") : LVBGS_ERR_ARG;
    s->slots[(size_t)stream].closing = true;
    return harvest(s, false);
}

extern "C" int lvbgs_server_submit(lvbgs_server* s, int stream, const void* d_img, void* d_mask, double lr,
                                   uint64_t tag) {
    if (s && stream >= 0 && stream < (int)s->slots.size() && s->slots[(size_t)stream].failed)
        return srv_err(s, LVBGS_ERR_DEVICE, "stream closed after a device error");
    if (!s || stream < 0 || stream >= (int)s->slots.size() || !s->slots[(size_t)stream].open ||
        s->slots[(size_t)stream].closing || !d_img)
        return s ? srv_err(s, LVBGS_ERR_ARG, "unknown or closing stream, or null frame") : LVBGS_ERR_ARG;
    s->slots[(size_t)stream].pending.push_back(Request{(const uint8_t*)d_img, (uint8_t*)d_mask, lr, tag, -1});
    ++s->st.frames_submitted;
    return LVBGS_OK;
}

extern "C" int lvbgs_server_submit_host(lvbgs_server* s, int stream, const void* host_img, size_t stride,
                                        void* d_mask, double lr, uint64_t tag) {
    if (s && stream >= 0 && stream < (int)s->slots.size() && s->slots[(size_t)stream].failed)
        return srv_err(s, LVBGS_ERR_DEVICE, "stream closed after a device error");
    if (!s || stream < 0 || stream >= (int)s->slots.size() || !s->slots[(size_t)stream].open ||
        s->slots[(size_t)stream].closing || !host_img)
        return s ? srv_err(s, LVBGS_ERR_ARG, "unknown or closing stream, or null frame") : LVBGS_ERR_ARG;
    if (s->o.host_ring <= 0) return srv_err(s, LVBGS_ERR_ARG, "the server has no host-ingest slots (host_ring 0)");
    Slot& sl = s->slots[(size_t)stream];
    if (sl.ring_busy >= s->o.host_ring) return srv_err(s, LVBGS_ERR_BUSY, "every host-ingest slot of the stream is busy");
    lvbgs_ctx* c = sl.ctx;
    const size_t row = (size_t)c->W * (size_t)c->C;
    if (stride == 0) stride = row;
    if (stride < row) return srv_err(s, LVBGS_ERR_ARG, "stride too small");
    (void)hipSetDevice(s->device);
    const int k = sl.ring_next;
    uint8_t* d = c->d_ring + (size_t)k * row * (size_t)c->H;
    // the slot's previous frame has completed (slots are freed in order as frames complete)
    hipStream_t cs = s->cstream[stream % lvbgs_server::kCopyStreams];
    if (hipMemcpy2DAsync(d, row, host_img, stride, row, (size_t)c->H, hipMemcpyHostToDevice, cs) != hipSuccess ||
        hipEventRecord(s->ring_ev[(size_t)stream][(size_t)k], cs) != hipSuccess)
        return srv_err(s, LVBGS_ERR_HIP, "ingest copy");
    sl.ring_next = (k + 1) % s->o.host_ring;
    ++sl.ring_busy;
    sl.pending.push_back(Request{d, (uint8_t*)d_mask, lr, tag, k});
    ++s->st.frames_submitted;
    return LVBGS_OK;
}

extern "C" size_t lvbgs_server_stream_bytes(const lvbgs_server* s, const lvbgs_params* params, int width, int height,
                                            int channels) {
    lvbgs_params d;
    if (!s || lvbgs_default_params(LVBGS_SUBSENSE, &d) != LVBGS_OK || width <= 0 || height <= 0 ||
        (channels != 1 && channels != 3))
        return 0;
    return lvbgs_arena_bytes(LVBGS_SUBSENSE, params ? *params : d, width, height, channels, s->o.host_ring);
}

extern "C" void* lvbgs_host_alloc(size_t bytes) {
    void* p = nullptr;
    return hipHostMalloc(&p, bytes) == hipSuccess ? p : nullptr;
}

extern "C" void lvbgs_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

extern "C" lvbgs_ctx* lvbgs_server_context(lvbgs_server* s, int stream) {
    if (!s || stream < 0 || stream >= (int)s->slots.size() || !s->slots[(size_t)stream].open) return nullptr;
    return s->slots[(size_t)stream].ctx;
}

extern "C" int lvbgs_server_debug_fixpoint_cap(lvbgs_server* s, int tail_cap) {
    if (!s || tail_cap < 0) return LVBGS_ERR_ARG;
    s->opt.tail_cap = tail_cap;
    return LVBGS_OK;
}

extern "C" int lvbgs_server_get_stats(lvbgs_server* s, lvbgs_server_stats* out) {
    if (!s || !out) return LVBGS_ERR_ARG;
    lvbgs_server_stats st = s->st;
    st.bytes_used = s->used;
    st.open_streams = 0;
    st.pending_frames = 0;
    for (auto& sl : s->slots) {
        st.open_streams += sl.open ? 1u : 0u;
        st.pending_frames += (uint64_t)sl.pending.size();
    }
    st.pooled_contexts = (uint32_t)s->pool.size();
    st.queued_steps = (uint32_t)s->queued.size();
    *out = st;
    return LVBGS_OK;
}

extern "C" void lvbgs_server_destroy(lvbgs_server* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    server_release(s);
    delete s;
}
