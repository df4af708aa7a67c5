"""GPU parity of the device BinClassif counts (lvbgs_binclassif_accumulate, lv::BinClassif::accumulate
of modules/datasets/src/metrics.cpp:21-50) against the oracle's restatement: random masks holding every
label (0, 255, shadow 50, out of scope 85, unknown 170, other values), with and without ROI and
ground truth, at ragged and full frame sizes, packed and strided rows, aligned and unaligned device
addresses. Counts must be equal."""
import numpy as np
import pytest

from litiv_amd import BinClassif
from oracle_lib import binclassif

pytestmark = pytest.mark.gpu

NAMES = ("tp", "tn", "fp", "fn", "se", "dc")


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _masks(rng, h, w):
    res = rng.choice(np.array([0, 255, 255, 0, 7], np.uint8), size=(h, w))
    gt = rng.choice(np.array([0, 255, 50, 85, 170, 0, 255, 3], np.uint8), size=(h, w))
    roi = rng.choice(np.array([0, 255, 255, 1], np.uint8), size=(h, w))
    return res, gt, roi


def _device(torch, arrays, stride, offset):
    """each (h, w) array in its own device buffer at byte `offset`, rows `stride` bytes apart"""
    out, keep = [], []
    for a in arrays:
        if a is None:
            out.append(None)
            continue
        h, w = a.shape
        buf = torch.zeros(offset + h * stride + 16, dtype=torch.uint8, device="cuda")
        rows = np.zeros((h, stride), np.uint8)
        rows[:, :w] = a
        buf[offset:offset + h * stride].copy_(torch.from_numpy(rows.reshape(-1)))
        keep.append(buf)
        out.append(buf.data_ptr() + offset)
    torch.cuda.synchronize()
    return out, keep


@pytest.mark.parametrize("h,w,stride_pad,offset,with_roi,with_gt", [
    (1, 1, 0, 0, True, True), (3, 17, 0, 0, True, True), (217, 333, 0, 0, True, True),
    (240, 320, 0, 0, False, True), (240, 320, 0, 0, True, False), (240, 320, 5, 0, True, True),
    (241, 319, 0, 3, True, True), (1080, 1920, 0, 0, True, True), (2160, 3840, 0, 0, True, True)])
def test_binclassif_matches_the_oracle(h, w, stride_pad, offset, with_roi, with_gt):
    torch = _torch()
    rng = np.random.default_rng(h * 7919 + w)
    res, gt, roi = _masks(rng, h, w)
    gt = gt if with_gt else None
    roi = roi if with_roi else None
    stride = w + stride_pad
    (dres, dgt, droi), keep = _device(torch, [res, gt, roi], stride, offset)
    print("binclassif %dx%d stride %d offset %d" % (w, h, stride, offset), flush=True)
    b = BinClassif().accumulate_device(dres, dgt, droi, w, h, stride if stride_pad else 0)
    exp = binclassif(res, gt, roi)
    assert tuple(b.counts()[n] for n in NAMES) == exp
    # counts accumulate over calls
    b.accumulate_device(dres, dgt, droi, w, h, stride if stride_pad else 0)
    assert tuple(b.counts()[n] for n in NAMES) == tuple(2 * x for x in exp)
    del keep


def test_binclassif_scores_subsense_masks_without_host_copies():
    # a SuBSENSE mask scored on the device against ground truth made from the same mask with some
    # labels changed; the counts follow from the changes
    torch = _torch()
    from litiv_amd import BackgroundSubtractorSuBSENSE, synth
    seq = synth.Sequence(synth.preset(synth.QVGA))
    bgs = BackgroundSubtractorSuBSENSE()
    bgs.initialize(seq.frame(0))
    dframe = torch.empty(seq.frame_bytes, dtype=torch.uint8, device="cuda")
    dmask = torch.empty((240, 320), dtype=torch.uint8, device="cuda")
    for t in range(12):
        dframe.copy_(torch.from_numpy(seq.frame(t).reshape(-1)))
        torch.cuda.synchronize()
        bgs.apply_device(dframe.data_ptr(), dmask.data_ptr())
    bgs.synchronize()
    gt = dmask.clone()
    gt[:40] = 85        # out of scope rows
    gt[200:, :100] = 170
    torch.cuda.synchronize()
    b = BinClassif().accumulate_device(dmask.data_ptr(), gt.data_ptr(), None, 320, 240, 0,
                                       bgs.stream())
    m = dmask.cpu().numpy()
    assert tuple(b.counts()[n] for n in NAMES) == binclassif(m, gt.cpu().numpy())
    assert b.fp == 0 and b.fn == 0 and b.dc == 40 * 320 + 40 * 100
    assert b.metrics()["pbc"] == 0.0
