"""GPU parity at the workloads' full frame sizes, against the sequential CPU oracle on the same frames:
1920x1080 and 3840x2160 RGB SuBSENSE through the plain drop-in (the 4K model holds more than 4 GiB
of samples, so sample indices pass 2^32 bytes), 640x512 grey LOBSTER, and one stream of each
resolution class of the 64-stream mix stepped as one batch. Masks are compared on every frame and
the whole model state at the end. Frames are generated on the device (bit-identical to the host
generator) and handed over as device pointers, several of them at addresses that are not 8-byte
aligned. Sequences are short so the oracle finishes in well under a minute."""
import numpy as np
import pytest

from litiv_amd import BackgroundSubtractorLOBSTER, BackgroundSubtractorSuBSENSE, Batch, synth
from oracle_lib import LOBSTER, SUBSENSE, Oracle

pytestmark = pytest.mark.gpu

SUB_MAPS = ["last_color", "last_desc", "last_fg", "lut", "T", "R", "V", "MLD", "DminLT", "DminST", "RawLT", "RawST",
            "FinLT", "FinST", "unstable", "blinks", "last_raw", "last_raw_blink", "last_dil_inv", "dsLT", "dsST"]
LOB_MAPS = ["last_color", "last_fg", "lut"]
SAMPLES = ["samples_color", "samples_desc"]


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _device_frames(torch, cfg, n):
    """n frames of the configured sequence generated on the device, and their host copies."""
    s = synth.Sequence(cfg)
    buf = torch.empty(n * s.frame_bytes, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    s.frames_device(0, n, buf.data_ptr())
    torch.cuda.synchronize()
    host = buf.cpu().numpy()
    return [host[t * s.frame_bytes:(t + 1) * s.frame_bytes].reshape(s.shape) for t in range(n)]


def _diff(a, b):
    return "%d values differ" % int((np.asarray(a) != np.asarray(b)).sum())


def _check_state(gpu, orc, names, tag):
    for name in names:
        sg, so = gpu.state(name), orc.state(name)
        assert np.array_equal(sg, so), "%s state %s: %s" % (tag, name, _diff(sg, so))


def _run_plain(gpu, orc, frames, offsets, names):
    torch = _torch()
    h, w = frames[0].shape[:2]
    n = frames[0].size
    buf = torch.empty(n + 64, dtype=torch.uint8, device="cuda")
    dmask = torch.empty((h, w), dtype=torch.uint8, device="cuda")
    gpu.initialize(frames[0])
    orc.initialize(frames[0], None)
    for t, f in enumerate(frames):
        off = offsets[t % len(offsets)]
        buf[off:off + n].copy_(torch.from_numpy(f.reshape(-1)).cuda())
        torch.cuda.synchronize()
        print("%dx%d frame %d at byte offset %d" % (w, h, t, off), flush=True)
        gpu.apply_device(buf.data_ptr() + off, dmask.data_ptr())
        gpu.synchronize()
        mo = orc.apply(f, None)
        mg = dmask.cpu().numpy()
        assert np.array_equal(mg, mo), "%dx%d frame %d mask: %s" % (w, h, t, _diff(mg, mo))
        eg, wg = gpu.stats()["examined"], gpu.stats()["writes"]
        assert (eg, wg) == orc.counters(), "%dx%d frame %d counters" % (w, h, t)
    _check_state(gpu, orc, names, "%dx%d" % (w, h))


def test_subsense_1080p_rgb_full_size():
    torch = _torch()
    frames = _device_frames(torch, synth.preset(synth.HD1080), 5)
    _run_plain(BackgroundSubtractorSuBSENSE(), Oracle(SUBSENSE), frames, (0, 3, 8, 5, 1), SUB_MAPS + SAMPLES)


def test_subsense_4k_rgb_full_size():
    torch = _torch()
    frames = _device_frames(torch, synth.preset(synth.UHD4K), 3)
    _run_plain(BackgroundSubtractorSuBSENSE(), Oracle(SUBSENSE), frames, (0, 7, 16), SUB_MAPS + SAMPLES)


def test_lobster_640x512_gray_full_size():
    torch = _torch()
    frames = _device_frames(torch, synth.preset(synth.THERMAL), 12)
    _run_plain(BackgroundSubtractorLOBSTER(), Oracle(LOBSTER), frames, (0, 1, 2, 3), LOB_MAPS + SAMPLES)


def test_batch_one_stream_per_class_full_size():
    """Streams 0, 16, 32 and 48 of the 64-stream mix (320x240 RGB, 640x480 RGB, 640x512 grey,
    1920x1080 RGB, each with its own generator seed and RNG key) in one batch, their frames packed
    into one device buffer at offsets that leave some of them unaligned."""
    torch = _torch()
    n_frames = 4
    ids = (0, 16, 32, 48)
    seqs = [_device_frames(torch, synth.preset(synth.BATCH64, sid), n_frames) for sid in ids]
    gpus, orcs = [], []
    for sid, frames in zip(ids, seqs):
        params = {"stream_id": sid, "rng_seed": 0x5EED0000 + sid}
        g, o = BackgroundSubtractorSuBSENSE(**params), Oracle(SUBSENSE, **params)
        g.initialize(frames[0])
        o.initialize(frames[0], None)
        gpus.append(g)
        orcs.append(o)
    sizes = [f[0].size for f in seqs]
    offsets = [sum(sizes[:i]) + 5 * i for i in range(len(ids))]
    assert any(off % 8 for off in offsets[1:])
    buf = torch.empty(offsets[-1] + sizes[-1], dtype=torch.uint8, device="cuda")
    dmasks = [torch.empty(f[0].shape[:2], dtype=torch.uint8, device="cuda") for f in seqs]
    torch.cuda.synchronize()
    batch = Batch(gpus)
    for t in range(n_frames):
        for i, frames in enumerate(seqs):
            buf[offsets[i]:offsets[i] + sizes[i]].copy_(torch.from_numpy(frames[t].reshape(-1)).cuda())
        torch.cuda.synchronize()
        print("batch frame %d" % t, flush=True)
        batch.apply_device([buf.data_ptr() + off for off in offsets], [m.data_ptr() for m in dmasks])
        batch.synchronize()
        for i, frames in enumerate(seqs):
            mo = orcs[i].apply(frames[t], None)
            mg = dmasks[i].cpu().numpy()
            assert np.array_equal(mg, mo), "stream %d frame %d mask: %s" % (ids[i], t, _diff(mg, mo))
            assert (gpus[i].stats()["examined"], gpus[i].stats()["writes"]) == orcs[i].counters()
    for i in range(len(ids)):
        _check_state(gpus[i], orcs[i], SUB_MAPS + SAMPLES, "stream %d" % ids[i])
    batch.close()


def _oracle_stream(sid, n_frames, keep_samples):
    """Stream `sid` of the 64-stream mix through the oracle: per-frame masks and counters, end state."""
    seq = synth.Sequence(synth.preset(synth.BATCH64, sid))
    o = Oracle(SUBSENSE, stream_id=sid)
    o.initialize(seq.frame(0), None)
    masks, counters = [], []
    for t in range(n_frames):
        masks.append(o.apply(seq.frame(t), None))
        counters.append(o.counters())
    names = SUB_MAPS + (SAMPLES if keep_samples else [])
    return masks, counters, {k: o.state(k) for k in names}, o.scalars()


def test_batch64_all_streams_with_the_tuned_options():
    """All 64 streams of the 64-stream mix in one batch with the options the online tuner settles on
    (three concurrent groups, graph replay, fetch depth 4, two full-grid fix-point launches,
    sequential limit 12 and 8 for the 1920x1080 class): every stream against its own oracle (run in
    host threads), masks and counters on every frame, maps and scalars at the end (sample planes for
    one stream per class)."""
    from concurrent.futures import ThreadPoolExecutor
    torch = _torch()
    n_frames = 4
    sids = list(range(64))
    keep = {0, 16, 32, 48}
    pool = ThreadPoolExecutor(max_workers=16)
    futs = {sid: pool.submit(_oracle_stream, sid, n_frames, sid in keep) for sid in sids}
    seqs = [synth.Sequence(synth.preset(synth.BATCH64, sid)) for sid in sids]
    bufs = [torch.empty(n_frames * s.frame_bytes, dtype=torch.uint8, device="cuda") for s in seqs]
    masks = [torch.empty((s.cfg.height, s.cfg.width), dtype=torch.uint8, device="cuda") for s in seqs]
    torch.cuda.synchronize()
    for s, b in zip(seqs, bufs):
        s.frames_device(0, n_frames, b.data_ptr())
    torch.cuda.synchronize()
    gpus = [BackgroundSubtractorSuBSENSE(stream_id=sid) for sid in sids]
    for g, s in zip(gpus, seqs):
        g.initialize(s.frame(0))
    batch = Batch(gpus, use_graph=1, groups=3, seq_limit=12, fetch_depth=4, fixpoint_launches=2,
                  cont_px_per_block=2048)
    for i in range(48, 64):
        batch.set_stream_options(i, seq_limit=8)
    got_masks = {sid: [] for sid in sids}
    got_counters = {sid: [] for sid in sids}
    for t in range(n_frames):
        print("batch64 frame %d" % t, flush=True)
        batch.apply_device([b.data_ptr() + t * s.frame_bytes for b, s in zip(bufs, seqs)],
                           [m.data_ptr() for m in masks])
        batch.synchronize()
        for sid in sids:
            got_masks[sid].append(masks[sid].cpu().numpy())
            st = gpus[sid].stats()
            got_counters[sid].append((st["examined"], st["writes"]))
    batch.close()
    for sid in sids:
        om, oc, ostate, osc = futs[sid].result()
        for t in range(n_frames):
            assert np.array_equal(got_masks[sid][t], om[t]), "stream %d frame %d mask: %s" % (
                sid, t, _diff(got_masks[sid][t], om[t]))
            assert got_counters[sid][t] == oc[t], "stream %d frame %d counters" % (sid, t)
        for name, so in ostate.items():
            sg = gpus[sid].state(name)
            assert np.array_equal(sg, so), "stream %d state %s: %s" % (sid, name, _diff(sg, so))
        sc = gpus[sid].scalars()
        for k in ("frame_idx", "frames_since_reset", "cooldown", "auto_reset", "lower_cap", "upper_cap",
                  "refresh_seq", "median_k", "use3x3"):
            assert sc[k] == osc[k], "stream %d scalar %s" % (sid, k)
        futs[sid] = None
    pool.shutdown()
