"""GPU parity of the batched SuBSENSE step: several streams of different sizes, channel counts, ROIs,
RNG keys and learning rates stepped by one launch per kernel (optionally replayed from a hipGraph)
give, stream by stream, exactly the sequential oracle's masks, counters and model state."""
import numpy as np
import pytest

from litiv_amd import BackgroundSubtractorSuBSENSE, Batch, synth
from oracle_lib import SUBSENSE, Oracle

pytestmark = pytest.mark.gpu

STATE = ["samples_color", "samples_desc", "last_color", "last_desc", "last_fg", "T", "R", "V", "MLD", "DminLT",
         "DminST", "RawLT", "RawST", "FinLT", "FinST", "unstable", "blinks", "last_raw", "last_raw_blink",
         "last_dil_inv", "dsLT", "dsST"]


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _streams():
    """(frames, roi, lr, params) per stream: mixed sizes / channels, one custom ROI, per-stream keys."""
    specs = []
    for sid, (w, h, c, n_obj) in enumerate([(96, 72, 3, 2), (160, 120, 1, 3), (128, 96, 3, 3), (64, 48, 1, 1),
                                             (352, 240, 3, 4)]):
        cfg = synth.preset(synth.QVGA, sid, width=w, height=h, channels=c, n_objects=n_obj, obj_min=6,
                           obj_max=max(8, w // 6))
        s = synth.Sequence(cfg)
        frames = [s.frame(t) for t in range(22)]
        roi = None
        if sid == 2:
            roi = np.zeros((h, w), np.uint8)
            roi[8:h - 6, 10:w - 12] = 255
        lr = [0.0, 0.0, 3.0, float("inf"), 0.0][sid]
        specs.append((frames, roi, lr, {"stream_id": sid, "rng_seed": 0xB65E + 17 * sid}))
    return specs


def _upload(torch, frames):
    return [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in frames]


@pytest.mark.parametrize("use_graph,vary,groups", [(False, False, 1), (True, False, 1), (True, True, 2),
                                                   (True, False, 3), (False, True, 4)])
def test_batch_matches_oracle_per_stream(use_graph, vary, groups):
    """vary: every frame gives every stream different launch-shape overrides (sequential scan depth,
    pixels per block of the continuation and fix-point phases), as the autotuner's trials do.
    groups: members stepped in that many concurrent groups."""
    torch = _torch()
    specs = _streams()
    gpus, orcs, dframes, dmasks = [], [], [], []
    for frames, roi, lr, params in specs:
        g, o = BackgroundSubtractorSuBSENSE(**params), Oracle(SUBSENSE, **params)
        g.initialize(frames[0], roi)
        o.initialize(frames[0], roi)
        gpus.append(g)
        orcs.append(o)
        dframes.append(_upload(torch, frames))
        h, w = frames[0].shape[:2]
        dmasks.append(torch.empty((h, w), dtype=torch.uint8, device="cuda"))
    torch.cuda.synchronize()
    batch = Batch(gpus, use_graph=use_graph, groups=groups)
    n_frames = len(specs[0][0])
    seq_limits, cont_ppb, fix_ppb = [1, 3, 8, 0, 255], [256, 4096, 0], [512, 0, 65536]
    for t in range(n_frames):
        if vary:
            for i in range(len(specs)):
                batch.set_stream_options(i, seq_limit=seq_limits[(t + i) % 5], cont_px_per_block=cont_ppb[(t + 2 * i) % 3],
                                         fix_px_per_block=fix_ppb[(t + i) % 3])
        batch.apply_device([d[t].data_ptr() for d in dframes], [m.data_ptr() for m in dmasks],
                           [s[2] for s in specs])
        batch.synchronize()
        for i, (frames, roi, lr, params) in enumerate(specs):
            mo = orcs[i].apply(frames[t], lr)
            mg = dmasks[i].cpu().numpy()
            assert np.array_equal(mg, mo), "stream %d frame %d: %d mask pixels differ" % (i, t, int((mg != mo).sum()))
            eg, wg = gpus[i].stats()["examined"], gpus[i].stats()["writes"]
            assert (eg, wg) == orcs[i].counters(), "stream %d frame %d counters" % (i, t)
    for i in range(len(specs)):
        for name in STATE:
            assert np.array_equal(gpus[i].state(name), orcs[i].state(name)), "stream %d state %s" % (i, name)
        sg, so = gpus[i].scalars(), orcs[i].scalars()
        for k in ("frame_idx", "frames_since_reset", "cooldown", "lower_cap", "upper_cap", "refresh_seq"):
            assert sg[k] == so[k], "stream %d scalar %s" % (i, k)
    batch.close()


def test_batch_then_single_steps_continue_the_sequence():
    # a member leaves the batch (batch closed) and continues alone: same result as never batching
    torch = _torch()
    frames = [synth.Sequence(synth.preset(synth.QVGA, 3, width=128, height=96)).frame(t) for t in range(16)]
    ref = BackgroundSubtractorSuBSENSE(stream_id=3)
    ref.initialize(frames[0])
    ref_masks = [ref.apply(f) for f in frames]
    a, b = BackgroundSubtractorSuBSENSE(stream_id=3), BackgroundSubtractorSuBSENSE(stream_id=4)
    a.initialize(frames[0])
    b.initialize(frames[0])
    d = _upload(torch, frames)
    m = torch.empty((96, 128), dtype=torch.uint8, device="cuda")
    m2 = torch.empty((96, 128), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    batch = Batch([a, b], use_graph=1, seq_limit=4, cont_blocks=64, cont_px_per_block=512, fix_px_per_block=1024)
    for t in range(8):
        batch.apply_device([d[t].data_ptr(), d[t].data_ptr()], [m.data_ptr(), m2.data_ptr()])
        batch.synchronize()
        assert np.array_equal(m.cpu().numpy(), ref_masks[t]), "batched frame %d" % t
    batch.close()
    for t in range(8, 16):
        assert np.array_equal(a.apply(frames[t]), ref_masks[t]), "single frame %d" % t


def test_batch_rejects_bad_members():
    from litiv_amd import BackgroundSubtractorLOBSTER
    from litiv_amd.bgs import LvbgsError
    f = synth.Sequence(synth.preset(synth.QVGA, 0, width=64, height=48)).frame(0)
    s = BackgroundSubtractorSuBSENSE()
    with pytest.raises(LvbgsError):
        Batch([s])  # not initialized
    s.initialize(f)
    lo = BackgroundSubtractorLOBSTER()
    lo.initialize(f)
    with pytest.raises(LvbgsError):
        Batch([s, lo])  # LOBSTER is stepped alone
    with pytest.raises(LvbgsError):
        Batch([s, s])  # duplicate member
    b = Batch([s])
    with pytest.raises(LvbgsError):
        b.set_stream_options(1, seq_limit=8)  # no such member
    with pytest.raises(LvbgsError):
        b.set_stream_options(0, seq_limit=256)
    b.set_stream_options(0, seq_limit=8, cont_px_per_block=1024)
    assert b.stream_options(0) == {"seq_limit": 8, "cont_px_per_block": 1024, "fix_px_per_block": 0}
    b.close()
