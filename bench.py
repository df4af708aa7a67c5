"""Benchmark of the gfx950 SuBSENSE / LOBSTER frame step on synthetic sequences resident in HBM.

Contract: `python bench.py --gpus N --steps K --warmup W` prints ONE JSON line. A step is one
apply() of one frame of every stream of the workload; the W warm-up steps are applied first
(untimed), then exactly K steps are timed between two device synchronizations. Frames are generated
on the GPU before the timed region (identical to the host generator's frames).

Default workload: 64 SuBSENSE streams (16x 320x240 RGB, 16x 640x480 RGB, 16x 640x512 grey,
16x 1920x1080 RGB; independent generator seeds; 600 frames each) stepped as one batch (one launch
per kernel for all streams, replayed from a hipGraph). The launch options come from a tuning pass
of the online autotuner over the same streams' first frames, run before and apart from the timed
region; the models are then re-initialized and the timed steps replay the frozen choices. The same
64 streams are also run serially through the plain one-stream drop-in (apply + synchronize per
stream) over the same frames, and the line carries both rates and their ratio.

Extra fields: roofline of the dominant kernel (the longer of the match's phase A and the model
update, by HIP-event time; algorithmic bytes from the device's exact examined-sample and
sample-write counters) with both kernels' figures, the CPU baseline (the sequential oracle on a
bounded sample of the same streams, one stream per host thread), per-step GPU latency p50/p99,
fix-point statistics, serving runs. Plain single-stream workloads (1080p, QVGA, 4K, LOBSTER
thermal) are selectable.
"""
import argparse
import collections
import json
import os
import platform
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SuBSENSE Mpix/s: 64-stream batched vs 1-stream plain; p99 frame latency"

WORKLOADS = {
    # name: (kind, algo, synth preset, streams, frames in the reference configuration)
    "subsense_batch64": ("batch", "subsense", 4, 64, 600),
    "subsense_1080p_plain": ("plain", "subsense", 1, 1, 1000),
    "subsense_qvga_plain": ("plain", "subsense", 0, 1, 300),
    "subsense_4k_plain": ("plain", "subsense", 2, 1, 500),
    "lobster_640x512_gray_plain": ("plain", "lobster", 3, 1, 1000),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(workload):
    """HBM bytes per launch of the roofline kernel from the committed hardware-counter summary of
    this workload (profiles/<workload>_traffic.json: FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a device copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", workload + "_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    return int(t["traffic_bytes_per_launch"]), "profiles/%s_traffic.json" % workload


def traffic_fields(workload, algorithmic_bytes):
    traffic, source = measured_traffic(workload)
    out = {"traffic": traffic}
    if traffic is not None:
        out["traffic_source"] = source
        out["traffic_over_algorithmic"] = round(traffic / algorithmic_bytes, 3)
    return out


def calibration_copy(torch, dev, mb):
    """One device copy of known size (profiling runs scale hardware byte counters by it)."""
    if mb:
        a = torch.empty(mb << 20, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b


def match_bytes(C, n_roi, examined):
    """Algorithmic bytes of one speculative-match pass over a stream: per ROI pixel the input frame (C),
    intra descriptors (2C), R (4), previous unstable flag (1), update record (4), previous raw decision
    (1), match result (4) and decision (1) written, plus 3C bytes per examined sample (colour C +
    descriptor 2C)."""
    return n_roi * (3 * C + 15) + 3 * C * examined


def update_bytes(C, n_roi, writes):
    """Algorithmic bytes of the update kernel over a stream (SuBSENSE, feedback maps fused in): per ROI
    pixel it reads roi (1), match result (4), decision (1), record (4), frame (C), intra descriptors
    (2C), last colour and descriptor (3C), eight f32 maps (32: MLD, D_min LT/ST, raw LT/ST, T, V, R),
    last_fg, blinks, last raw and last raw blink (4), and writes the eight maps (32), unstable (1),
    last colour and descriptor (3C), blinks, last raw blink, last raw and raw (4): 79 + 9C bytes;
    plus 3C bytes (colour + descriptor) per sample write."""
    return n_roi * (79 + 9 * C) + 3 * C * writes


def update_bytes_lobster(C, n_roi, writes):
    """Algorithmic bytes of LOBSTER's update kernel over a stream: per ROI pixel the record (4), the
    decision (1), the frame (C) and intra descriptors (2C); plus 3C bytes per sample write."""
    return n_roi * (5 + 3 * C) + 3 * C * writes


def kernel_figures(name, avg_s, bytes_launch):
    """Roofline figures of one kernel: algorithmic bytes per launch over its mean launch time."""
    achieved = bytes_launch / avg_s / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_launch": int(bytes_launch),
            "avg_launch_us": round(avg_s * 1e6, 2)}


def step_bytes(C, n_roi, examined, writes):
    """Algorithmic bytes of a whole frame step (per-pixel figure 114 + 9(k+w) for RGB, 100 + 3(k+w) grey)."""
    return n_roi * (114 if C == 3 else 100) + 3 * C * (examined + writes)


# ------------------------------------------------------------------------------------------------
# CPU baseline: the sequential oracle, one stream per host thread (ctypes releases the GIL)
# ------------------------------------------------------------------------------------------------
def _oracle_rate(algo, classes, lr, budget_s):
    """classes: list of (weight = pixels of the class per workload step, [(stream id, synth sequence,
    frames to prepare)] with one stream per thread). Per class, every thread builds its oracle,
    initializes it and generates its frames (untimed, in parallel), then all threads apply frames
    concurrently for about budget_s/len(classes) seconds. Returns the mix rate (pixels of one workload
    step / CPU wall time of one step) and a description per class."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import LOBSTER, SUBSENSE, Oracle
    per_class = budget_s / len(classes)
    rates, descr = [], []
    for weight, streams in classes:
        n = len(streams)
        orcs, frames = [None] * n, [None] * n
        done = [0] * n
        ready = threading.Barrier(n + 1)
        go = threading.Event()
        stop = [0.0]

        def work(i):
            sid, seq, n_frames = streams[i]
            frames[i] = [seq.frame(t) for t in range(n_frames)]
            o = Oracle(SUBSENSE if algo == "subsense" else LOBSTER, stream_id=sid)
            o.initialize(frames[i][0])
            orcs[i] = o
            ready.wait()
            go.wait()
            while time.perf_counter() < stop[0] and done[i] < n_frames - 1:
                o.apply(frames[i][1 + done[i]], lr)
                done[i] += 1

        ths = [threading.Thread(target=work, args=(i,)) for i in range(n)]
        for t in ths:
            t.start()
        ready.wait()
        t0 = time.perf_counter()
        stop[0] = t0 + per_class
        go.set()
        for t in ths:
            t.join()
        wall = time.perf_counter() - t0
        f = frames[0][0]
        rate = sum(f.shape[0] * f.shape[1] * d for d in done) / wall
        rates.append((weight, rate))
        descr.append("%dx%d: %d frames over %d thread(s) in %.1f s" % (f.shape[1], f.shape[0], sum(done), n, wall))
        log("cpu baseline: " + descr[-1])
        del orcs, frames
    total_px = sum(w for w, _ in rates)
    step_s = sum(w / r for w, r in rates)
    return total_px / step_s / 1e6, descr


def host_cpu():
    """CPU model name (from /proc/cpuinfo), logical CPUs of the host, CPUs this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model or platform.machine(), os.cpu_count(), usable


def oracle_flags():
    """The C compiler flags the oracle library was built with (oracle/Makefile)."""
    try:
        with open(os.path.join(ROOT, "oracle", "Makefile")) as f:
            for line in f:
                if line.startswith("CFLAGS"):
                    return line.split("=", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(algo, classes, lr, budget_s):
    model, nproc, usable = host_cpu()
    threads = max(len(streams) for _, streams in classes)
    v, descr = _oracle_rate(algo, classes, lr, budget_s)
    return {"value": round(v, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "threads": threads, "host_cpu": model, "host_nproc": nproc, "usable_cpus": usable,
            "compiler": "gcc " + (oracle_flags() or "?"),
            "sample": "sequential CPU oracle (oracle/, the C restatement of the reference's apply; gcc %s), one "
                      "stream per thread (%d thread(s) on %s, %s logical CPUs, %s usable), on the first frames of "
                      "streams of each resolution class; per-class rates weighted by the workload's pixel mix: %s"
                      % (oracle_flags(), threads, model, nproc, usable, "; ".join(descr))}


# ------------------------------------------------------------------------------------------------
# batched workload
# ------------------------------------------------------------------------------------------------
def run_batch(args, preset, n_streams, ref_frames):
    import torch
    from litiv_amd import BackgroundSubtractorSuBSENSE, Batch, synth
    from litiv_amd.autotune import OnlineTuner

    steps = args.steps if args.steps is not None else ref_frames - args.warmup
    tune_steps = 0 if args.no_tune else args.tune_steps
    total = max(args.warmup + steps, tune_steps)
    dev = torch.device("cuda:0")
    cfgs = [synth.preset(preset, s) for s in range(n_streams)]
    seqs = []
    for c in cfgs:
        c.frames = max(c.frames, total)
        seqs.append(synth.Sequence(c))
    px_step = sum(c.width * c.height for c in cfgs)
    frames, masks = [], []
    for s in seqs:
        f = torch.empty(total * s.frame_bytes, dtype=torch.uint8, device=dev)
        frames.append(f)
        masks.append(torch.empty((2, s.cfg.width * s.cfg.height), dtype=torch.uint8, device=dev))
    torch.cuda.synchronize()
    for s, f in zip(seqs, frames):
        s.frames_device(0, total, f.data_ptr())
    torch.cuda.synchronize()
    log("workload subsense_batch64: %d streams, %.2f Mpix per step, %d warm-up + %d timed steps, %.1f GB of frames"
        % (n_streams, px_step / 1e6, args.warmup, steps, sum(f.numel() for f in frames) / 1e9))
    calibration_copy(torch, dev, args.calib_copy_mb)
    f0 = [s.frame(0) for s in seqs]
    subs = [BackgroundSubtractorSuBSENSE(stream_id=i, device=0) for i in range(n_streams)]

    def init_all():
        for b, f in zip(subs, f0):
            b.initialize(f)

    def ptrs(t):
        return ([f.data_ptr() + t * s.frame_bytes for f, s in zip(frames, seqs)],
                [m[t % 2].data_ptr() for m in masks])

    # (1) the same streams serially through the plain drop-in: apply + synchronize per stream
    plain = None
    if not args.no_plain:
        init_all()
        for t in range(args.warmup):
            im, mk = ptrs(t)
            for i, b in enumerate(subs):
                b.apply_device(im[i], mk[i])
                b.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            im, mk = ptrs(args.warmup + k)
            for i, b in enumerate(subs):
                b.apply_device(im[i], mk[i])
                b.synchronize()
        torch.cuda.synchronize()
        wall_plain = time.perf_counter() - t0
        plain = px_step * steps / wall_plain / 1e6
        log("plain serial: %.3f ms/step, %.1f Mpix/s" % (wall_plain / steps * 1e3, plain))

    # (2) tuning pass, outside the timed region: the online tuner steps the batch over the first
    # tune_steps frames with a trial round every tune_period steps (a round times each candidate of the
    # shared options, or a resolution class's launch shape, on live frames; options never change
    # results). Its choices are then frozen.
    tuner, tune_info = None, None
    if tune_steps:
        init_all()
        batch = Batch(subs)
        tuner = OnlineTuner(batch, period=args.tune_period, max_period=args.tune_period, log=log)
        t0 = time.perf_counter()
        for t in range(tune_steps):
            tuner.apply(*ptrs(t))
        batch.synchronize()
        log("tuning pass: %d steps, %d rounds, %.1f s" % (tune_steps, tuner.rounds, time.perf_counter() - t0))
        tuned = batch.options()
        tune_info = {"pass_steps": tune_steps, "rounds": tuner.rounds, "period": args.tune_period,
                     "shared_candidates": tuner.shared,
                     "shared_median_step_ms": [round(m, 4) if m else m for m in tuner.shared_medians()],
                     "per_resolution": tuner.class_choices()}
        batch.close()

    # (3) the timed run: models re-initialized, frozen options, the step captured once as a hipGraph
    # (per group) during the warm-up and replayed
    init_all()
    batch = Batch(subs)
    if tuner is not None:
        tuner.apply_choices(batch)
        batch.set_options(use_graph=1)
    tuned = batch.options()
    for t in range(args.warmup):
        batch.apply_device(*ptrs(t))
    batch.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        batch.apply_device(*ptrs(args.warmup + k))
    batch.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    mpix = px_step * steps / wall / 1e6
    log("batched: %.3f ms/step, %.1f Mpix/s (options %s)" % (wall / steps * 1e3, mpix, tuned))
    batch.close()

    # (4) measurement passes over the same frames (re-initialized, frozen options): HIP-event time of
    # the match's phase A and of the update kernel (one group, so one launch of each covers every
    # stream) with the exact per-stream counters behind their algorithmic bytes, then whole-step
    # latency
    n_meas = min(steps, args.meas_steps)
    kern = {}
    sbytes, iters, work = 0.0, [], np.zeros(8)
    for stage in ("match", "update"):
        log("measurement pass: %s-kernel timing and counters over %d steps" % (stage, n_meas))
        init_all()
        batch = Batch(subs, **tuned)
        if tuner is not None:
            tuner.apply_choices(batch)
        batch.set_options(groups=1)
        for t in range(args.warmup):
            batch.apply_device(*ptrs(t))
        batch.set_timing(stage)
        ms_list, nbytes = [], 0.0
        for k in range(n_meas):
            batch.apply_device(*ptrs(args.warmup + k))
            ms, _ = batch.take_timing()
            ms_list.append(ms)
            for i, b in enumerate(subs):
                st = b.stats()
                C = st["channels"]
                if stage == "match":
                    nbytes += match_bytes(C, st["n_roi"], st["examined_speculative"])
                    sbytes += step_bytes(C, st["n_roi"], st["examined"], st["writes"])
                    iters.append(st["fixpoint_iters"])
                    work += np.array(st["fixpoint_work"], dtype=np.float64)
                else:
                    nbytes += update_bytes(C, st["n_roi"], st["writes"])
        batch.set_timing("")
        batch.close()
        kern[stage] = (float(np.mean(ms_list)) / 1e3, nbytes / n_meas)
    log("measurement pass: step latency over %d steps" % n_meas)
    init_all()
    batch = Batch(subs, **tuned)
    if tuner is not None:
        tuner.apply_choices(batch)
    for t in range(args.warmup):
        batch.apply_device(*ptrs(t))
    batch.set_timing("step")
    step_ms = []
    for k in range(n_meas):
        batch.apply_device(*ptrs(args.warmup + k))
        ms, _ = batch.take_timing()
        step_ms.append(ms)
    batch.set_timing("")
    batch.close()
    log("step latency p50 %.3f ms, p99 %.3f ms" % (np.percentile(step_ms, 50), np.percentile(step_ms, 99)))
    kernels = {
        "match": kernel_figures("kb_match (phase A of the ordered sample match, all streams)", *kern["match"]),
        "update": kernel_figures("kb_update (model update, diffusion and feedback maps, all streams)",
                                 *kern["update"]),
    }
    dominant = max(kernels, key=lambda k: kernels[k]["avg_launch_us"])
    result = {
        "metric": METRIC,
        "value": round(mpix, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/u16 integer model, f32 feedback maps",
        "data": "synthetic (deterministic generator, per-stream seeds, frames resident in HBM)",
        "config": {"workload": "subsense_batch64", "streams": n_streams,
                   "mix": "16x320x240x3, 16x640x480x3, 16x640x512x1, 16x1920x1080x3",
                   "pixels_per_step": px_step, "frames": total, "algo": "subsense", "learning_rate": 0.0,
                   "launch": "one launch per kernel for all streams, hipGraph replay", "options": tuned,
                   "autotune": tune_info,
                   "parallelism": "64 streams batched on one GPU"},
        "plain_serial": None if plain is None else {
            "value": round(plain, 2), "unit": "Mpix/s",
            "what": "same 64 streams and frames through the one-stream drop-in, apply + synchronize per stream"},
        "batched_over_plain": None if plain is None else round(mpix / plain, 3),
        "roofline": dict(kernels[dominant], **traffic_fields("subsense_batch64_" + dominant,
                                                              kernels[dominant]["bytes_per_launch"])),
        "kernels": {k: dict(v, **traffic_fields("subsense_batch64_" + k, v["bytes_per_launch"]))
                    for k, v in kernels.items()},
        "frame_latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                             "p99": round(float(np.percentile(step_ms, 99)), 4),
                             "what": "GPU time of one batched step (every stream's frame), HIP events, %d steps"
                                     % n_meas},
        "step_algorithmic_gbs": round(sbytes / n_meas / (float(np.mean(step_ms)) / 1e3) / 1e9, 1),
        "fixpoint": {"mean_iters": round(float(np.mean(iters)), 3),
                     "mean_worklist_per_stream": [round(float(x), 1) for x in (work / len(iters))[1:6]]},
    }
    if args.serve_seconds > 0:
        # the batch's frames (tens of GB) and models leave HBM before the serving runs, whose budgets
        # assume the device to themselves
        for b in subs:
            b.close()
        del frames, masks, subs
        torch.cuda.empty_cache()
        result["serving"] = {}
        runs = [(int(x), 250.0, "device") for x in str(args.serve_streams).split(",") if x]
        runs += [(int(x.split(":")[0]), float(x.split(":")[1]), "device") for x in str(args.serve_bound).split(",")
                 if x]
        runs += [(int(x), 250.0, "host") for x in str(args.serve_host).split(",") if x]
        for n, gb, ingest in runs:
            log("serving: %d stream slots, %.0f GB budget, %s frames, %.0f s warm-up + %.0f s measured" % (
                n, gb, ingest, args.serve_warmup, args.serve_seconds))
            key = "%d_streams" % n + ("" if gb == 250.0 else "_%dGB" % gb) + ("_host_ingest" if ingest == "host" else "")
            result["serving"][key] = run_serve(args, n, gb, ingest)
            torch.cuda.empty_cache()
        log("serving: %s" % json.dumps(result["serving"]))
    if not args.no_cpu:
        # streams 16g .. 16g+threads-1 of each resolution class g, one per thread, first frames
        threads = max(1, min(args.cpu_threads, 16))
        classes = []
        for g in range(4):
            px = cfgs[16 * g].width * cfgs[16 * g].height
            n_frames = max(4, min(40, int(2.5e6 * args.cpu_budget / 4 / px) + 2))
            streams = [(i, seqs[i], n_frames) for i in range(16 * g, 16 * g + threads)]
            weight = sum(cfgs[i].width * cfgs[i].height for i in range(16 * g, 16 * g + 16))
            classes.append((weight, streams))
        result["cpu_baseline"] = cpu_baseline("subsense", classes, 0.0, args.cpu_budget)
        result["cpu_baseline"]["ratio_gpu_over_cpu"] = round(mpix / result["cpu_baseline"]["value"], 1)
    return result


# ------------------------------------------------------------------------------------------------
# serving: streams arrive and leave, frames arrive as Poisson processes (wall clock), the server
# batches whatever is waiting
# ------------------------------------------------------------------------------------------------
def run_serve(args, n_slots, budget_gb=250.0, ingest="device", log_prefix="serve"):
    """n_slots stream slots; resolutions from the SERVE generator preset ({320x240, 640x480, 1280x720,
    1920x1080} RGB with p = .3/.3/.2/.2); frames arrive at 30 fps (Poisson) per stream; lifetimes are
    exponential with a mean of 2000 frames and an ended stream is replaced by a new one; models are
    admitted against a device-memory budget (250 GB by default; streams that do not fit wait).
    Latency = frame arrival -> completion seen by the host. Each stream plays its first `ring`
    generated frames forwards and backwards (0, 1, .., ring-1, ring-2, .., 1, 0, 1, ..), so its
    objects move continuously for the whole lifetime with no jump back. ingest = "device": the frames
    are in HBM and handed over as device pointers; "host": they are in pinned host memory and go
    through the server's ingest path (copied on its copy stream into the stream's device slots as
    they arrive), so the latency includes the host-to-device copy. Replacement streams are prepared
    (frames generated, and for host ingest copied to pinned memory) on a background thread, so the
    serving loop never stalls on their setup."""
    import heapq

    import torch
    from litiv_amd import Server, synth

    from litiv_amd.bgs import PinnedBuffer

    host = ingest == "host"
    fps, mean_life = 30.0, 2000.0
    ring = args.serve_host_ring if host else args.serve_ring
    period = 2 * ring - 2
    rng = np.random.default_rng(0x5E12E)      # arrivals (serving loop)
    life_rng = np.random.default_rng(0x11FE)  # lifetimes (stream preparation)
    srv = Server(hbm_budget=int(budget_gb * (1 << 30)), max_inflight=2, max_batch=n_slots, fix_px_per_block=8192,
                 cont_px_per_block=2048, seq_limit=8, groups=2, host_ring=4 if host else 0)
    dev = torch.device("cuda:0")
    incarnation = [0]
    slots = {}      # server stream id -> dict
    waiting = []    # prepared streams that did not fit the budget yet
    gen_stream = torch.cuda.Stream()

    def prepare():
        """A new stream with its frames generated on the GPU (on a stream of its own) and, for host
        ingest, copied into pinned host memory. Runs on the preparation thread during the window, so
        the serving loop never waits for it."""
        k = incarnation[0]
        incarnation[0] += 1
        cfg = synth.preset(synth.SERVE, k)
        cfg.frames = max(cfg.frames, ring)
        seq = synth.Sequence(cfg)
        st = {"k": k, "seq": seq, "w": cfg.width, "h": cfg.height, "life": max(1, int(life_rng.exponential(mean_life)))}
        with torch.cuda.stream(gen_stream):
            frames = torch.empty((ring, seq.frame_bytes), dtype=torch.uint8, device=dev)
            seq.frames_device(0, ring, frames.data_ptr(), stream=gen_stream.cuda_stream)
            if host:
                pinned = PinnedBuffer(ring * seq.frame_bytes)
                torch.from_numpy(pinned.array).copy_(frames.view(-1), non_blocking=True)
        gen_stream.synchronize()
        st["frames"] = pinned if host else frames
        return st

    ready = collections.deque()   # prepared streams for replacements
    prep_stop = threading.Event()

    def prep_loop():
        while not prep_stop.is_set():
            if len(ready) < 8:
                ready.append(prepare())
            else:
                time.sleep(0.002)

    def admit(st, now):
        seq = st["seq"]
        frames = st["frames"]
        if host:
            sid = srv.open_host(frames.array[:seq.frame_bytes].reshape(st["h"], st["w"], 3), stream_id=st["k"])
        else:
            sid = srv.open(frames.data_ptr(), st["w"], st["h"], 3, stream_id=st["k"])
        if sid is None:
            return None
        st.update(sid=sid, pend=[], ending=False,
                  masks=torch.empty((ring, st["w"] * st["h"]), dtype=torch.uint8, device=dev),
                  t=0, done=0, closing=False, next=now + rng.exponential(1.0 / fps))
        slots[sid] = st
        return sid

    t0 = time.perf_counter()
    for _ in range(n_slots):
        st = prepare()
        if admit(st, t0) is None:
            waiting.append(st)
    torch.cuda.synchronize()
    log("%s: %d streams admitted, %d waiting, %.1f GB of models, setup %.1f s" % (
        log_prefix, len(slots), len(waiting), srv.stats()["bytes_used"] / 1e9, time.perf_counter() - t0))
    prep = threading.Thread(target=prep_loop, daemon=True)
    prep.start()
    start = time.perf_counter()
    # the arrival clock starts now
    for sid, st in slots.items():
        st["next"] = start + rng.exponential(1.0 / fps)
    heap = [(st["next"], sid) for sid, st in slots.items()]
    heapq.heapify(heap)
    t_warm, t_end = start + args.serve_warmup, start + args.serve_warmup + args.serve_seconds
    arrival = {}
    lat, px_done, n_done_window = [], 0, 0
    npx = {}
    replaced = 0
    to_replace = 0   # ended streams whose replacement is not prepared yet
    tag_seq = 0
    last_log = start
    while True:
        now = time.perf_counter()
        if now >= t_end:
            break
        while heap and heap[0][0] <= now:
            ta, sid = heapq.heappop(heap)
            st = slots.get(sid)
            if st is None or st["closing"] or st.get("next") != ta:
                continue
            i = st["t"] % period
            i = i if i < ring else period - i
            tag_seq += 1
            arrival[tag_seq] = ta
            npx[tag_seq] = st["w"] * st["h"]
            if host:
                st["pend"].append((i, tag_seq))
            else:
                srv.submit(sid, st["frames"][i].data_ptr(), st["masks"][i].data_ptr(), 0.0, tag=tag_seq)
            st["t"] += 1
            if st["t"] >= st["life"]:
                st["ending"] = True
            else:
                st["next"] = ta + rng.exponential(1.0 / fps)
                heapq.heappush(heap, (st["next"], sid))
        if host:
            # host frames go to the ingest path in arrival order; a stream whose device slots are all
            # busy keeps the rest waiting (their latency clock runs)
            for sid, st in slots.items():
                pend = st["pend"]
                while pend:
                    i, tag = pend[0]
                    fb = st["seq"].frame_bytes
                    if not srv.submit_host(sid, st["frames"].ptr + i * fb, st["masks"][i].data_ptr(), 0.0, tag=tag):
                        break
                    pend.pop(0)
        for sid, st in slots.items():
            if st.get("ending") and not st["closing"] and not st["pend"]:
                st["closing"] = True
                srv.close(sid)
        srv.pump()
        done = srv.poll()
        if len(done):
            tnow = time.perf_counter()
            for tag in done["tag"]:
                tag = int(tag)
                ta = arrival.pop(tag)
                p = npx.pop(tag)
                if ta >= t_warm:
                    lat.append(tnow - ta)
                if t_warm <= tnow:
                    px_done += p
                    n_done_window += 1
        # streams that ended and finished: replace them (the new stream may have to wait)
        for sid in [s for s, st in slots.items() if st["closing"]]:
            if srv.retired(sid):
                st = slots.pop(sid)
                if host:
                    st["frames"].free()
                to_replace += 1
                replaced += 1
        while to_replace and ready:
            waiting.append(ready.popleft())
            to_replace -= 1
        while waiting:
            sid = admit(waiting[0], time.perf_counter())
            if sid is None:
                break
            waiting.pop(0)
            heapq.heappush(heap, (slots[sid]["next"], sid))
        if now - last_log > 5.0:
            last_log = now
            sst = srv.stats()
            log("%s: t=%.0fs open %d pending %d done %d steps %d" % (log_prefix, now - start, sst["open_streams"],
                                                                   sst["pending_frames"], sst["frames_done"],
                                                                   sst["steps"]))
    window = t_end - t_warm
    sst = srv.stats()
    prep_stop.set()
    prep.join()
    srv.drain()
    srv.close_server()
    lat_ms = np.array(lat) * 1e3
    offered = sum(st["w"] * st["h"] for st in slots.values()) * fps / 1e6
    return {
        "streams": n_slots, "budget_gb": budget_gb, "frames": "%d-frame ping-pong per stream" % ring,
        "ingest": "pinned host memory through the server's copy stream (latency includes the copy)" if host
                  else "device frames",
        "open_at_end": sst["open_streams"], "waiting_at_end": len(waiting) + to_replace,
        "replaced": replaced, "window_s": round(window, 1),
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 3) if len(lat_ms) else None,
                       "p99": round(float(np.percentile(lat_ms, 99)), 3) if len(lat_ms) else None,
                       "max": round(float(lat_ms.max()), 3) if len(lat_ms) else None,
                       "frames": int(len(lat_ms))},
        "mpix_s": round(px_done / window / 1e6, 1),
        "offered_mpix_s": round(offered, 1),
        "mean_batch": round(sst["frames_launched"] / max(1, sst["steps"]), 1),
        "pending_at_end": sst["pending_frames"],
        "model_gb": round(sst["bytes_used"] / 1e9, 1),
    }


# ------------------------------------------------------------------------------------------------
# plain single-stream workloads
# ------------------------------------------------------------------------------------------------
def run_plain(args, name, algo, preset, ref_frames):
    import torch
    from litiv_amd import BackgroundSubtractorLOBSTER, BackgroundSubtractorSuBSENSE, synth

    steps = args.steps if args.steps is not None else ref_frames - args.warmup
    total = args.warmup + steps
    cfg = synth.preset(preset, 0)
    cfg.frames = max(cfg.frames, total)
    seq = synth.Sequence(cfg)
    fb = seq.frame_bytes
    W, H, C = cfg.width, cfg.height, cfg.channels
    lr = 0.0 if algo == "subsense" else 16.0
    dev = torch.device("cuda:0")
    log("workload %s: %dx%dx%d, %d warm-up + %d timed frames, %.2f GB of frames in HBM" % (
        name, W, H, C, args.warmup, steps, total * fb / 1e9))
    frames = torch.empty(total * fb, dtype=torch.uint8, device=dev)
    masks = torch.empty((8, H * W), dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    seq.frames_device(0, total, frames.data_ptr())
    torch.cuda.synchronize()
    base = frames.data_ptr()
    calibration_copy(torch, dev, args.calib_copy_mb)
    bgs = (BackgroundSubtractorSuBSENSE if algo == "subsense" else BackgroundSubtractorLOBSTER)(device=0)
    f0 = seq.frame(0)

    def warm():
        bgs.initialize(f0)
        for t in range(args.warmup):
            bgs.apply_device(base + t * fb, masks[t % 8].data_ptr(), lr)
        bgs.synchronize()

    warm()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        t = args.warmup + i
        bgs.apply_device(base + t * fb, masks[t % 8].data_ptr(), lr)
    bgs.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    mpix = W * H * steps / wall / 1e6
    log("timed: %.3f s, %.3f ms/frame, %.1f Mpix/s" % (wall, wall / steps * 1e3, mpix))

    n_meas = min(steps, args.meas_steps)
    warm()
    ex_spec = ex_fin = wr = 0
    iters, max_iter, work, match_ms = [], 0, np.zeros(8), []
    bgs.set_timing("match")
    for i in range(n_meas):
        t = args.warmup + i
        bgs.apply_device(base + t * fb, masks[t % 8].data_ptr(), lr)
        ms, _ = bgs.take_timing()
        match_ms.append(ms)
        st = bgs.stats()
        ex_spec += st["examined_speculative"]
        ex_fin += st["examined"]
        wr += st["writes"]
        iters.append(st["fixpoint_iters"])
        work += np.array(st["fixpoint_work"], dtype=np.float64)
        max_iter = max(max_iter, st["max_fixpoint_iter"])
    upd_ms, upd_wr = [], 0
    bgs.set_timing("update")
    warm()
    for i in range(n_meas):
        t = args.warmup + i
        bgs.apply_device(base + t * fb, masks[t % 8].data_ptr(), lr)
        ms, _ = bgs.take_timing()
        upd_ms.append(ms)
        upd_wr += bgs.stats()["writes"]
    bgs.set_timing("step")
    warm()
    step_ms = []
    for i in range(n_meas):
        t = args.warmup + i
        bgs.apply_device(base + t * fb, masks[t % 8].data_ptr(), lr)
        ms, _ = bgs.take_timing()
        step_ms.append(ms)
    bgs.set_timing("")
    n_roi = st["n_roi"]
    ub = (update_bytes if algo == "subsense" else update_bytes_lobster)(C, n_roi, upd_wr / n_meas)
    kernels = {
        "match": dict(kernel_figures("k_match (phase A of the ordered sample match)", float(np.mean(match_ms)) / 1e3,
                                     match_bytes(C, n_roi, ex_spec / n_meas)),
                      examined_per_px=round(ex_spec / n_meas / n_roi, 3)),
        "update": kernel_figures("k_update (model update and diffusion%s)" % (
            ", feedback maps fused" if algo == "subsense" else ""), float(np.mean(upd_ms)) / 1e3, ub),
    }
    dominant = max(kernels, key=lambda k: kernels[k]["avg_launch_us"])
    sb = step_bytes(C, n_roi, ex_fin / n_meas, wr / n_meas)
    result = {
        "metric": METRIC,
        "value": round(mpix, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/u16 integer model, f32 feedback maps",
        "data": "synthetic (deterministic generator, frames resident in HBM)",
        "config": {"workload": name, "width": W, "height": H, "channels": C, "frames": total,
                   "algo": algo, "learning_rate": lr, "streams": 1, "parallelism": "single stream, one GPU"},
        "roofline": dict(kernels[dominant], **traffic_fields(name + "_" + dominant,
                                                              kernels[dominant]["bytes_per_launch"])),
        "kernels": {k: dict(v, **traffic_fields(name + "_" + k, v["bytes_per_launch"])) for k, v in kernels.items()},
        "frame_latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                             "p99": round(float(np.percentile(step_ms, 99)), 4),
                             "what": "GPU time of one frame step (HIP events), %d frames" % n_meas},
        "step_algorithmic_gbs": round(sb / (float(np.mean(step_ms)) / 1e3) / 1e9, 1),
        "fixpoint": {"mean_iters": round(float(np.mean(iters)), 3), "max_iter": int(max_iter),
                     "mean_worklist": [round(float(x), 1) for x in (work / n_meas)[1:6]]},
    }
    if not args.no_cpu:
        n_frames = max(4, min(60, int(3e6 * args.cpu_budget / (W * H)) + 2))
        result["cpu_baseline"] = cpu_baseline(algo, [(W * H, [(0, seq, n_frames)])], lr, args.cpu_budget)
        result["cpu_baseline"]["ratio_gpu_over_cpu"] = round(mpix / result["cpu_baseline"]["value"], 1)
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="subsense_batch64", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=24.0, help="seconds of oracle time for the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="batched workload: host threads of the CPU baseline (one stream each)")
    ap.add_argument("--meas-steps", type=int, default=60, help="steps of each measurement pass")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="batched workload: skip the serial plain comparison")
    ap.add_argument("--no-tune", action="store_true", help="batched workload: default launch options")
    ap.add_argument("--tune-steps", type=int, default=144,
                    help="batched workload: steps of the tuning pass run before the timed region")
    ap.add_argument("--tune-period", type=int, default=8, help="tuning pass: steps between trial rounds")
    ap.add_argument("--calib-copy-mb", type=int, default=0, help="profiling runs: one device copy of this size first")
    ap.add_argument("--serve-seconds", type=float, default=12.0,
                    help="batched workload: measured seconds of the serving run (0 = skip)")
    ap.add_argument("--serve-warmup", type=float, default=3.0)
    ap.add_argument("--serve-streams", default="256,200", help="comma-separated stream-slot counts, one run each")
    ap.add_argument("--serve-bound", default="256:100",
                    help="comma-separated slots:GB serving runs with a smaller model budget (admission binds)")
    ap.add_argument("--serve-ring", type=int, default=32, help="generated frames per served stream (ping-pong)")
    ap.add_argument("--serve-host", default="256",
                    help="comma-separated stream-slot counts of serving runs fed from pinned host memory")
    ap.add_argument("--serve-host-ring", type=int, default=8,
                    help="host-fed serving: pinned frames per stream (ping-pong)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs (--gpus %d) are not part of it. Streams "
              "shard across GPUs with no exchange between them." % args.gpus, file=sys.stderr)
        sys.exit(2)
    kind, algo, preset, n_streams, ref_frames = WORKLOADS[args.workload]
    if kind == "batch":
        result = run_batch(args, preset, n_streams, ref_frames)
    else:
        result = run_plain(args, args.workload, algo, preset, ref_frames)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
