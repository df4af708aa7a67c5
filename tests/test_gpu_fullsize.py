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
