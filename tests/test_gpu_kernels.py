"""GPU checks of the stand-alone kernels: LBSP known-answer test and random parity, synthetic
frames generated on the device equal the host generator's."""
import os

import numpy as np
import pytest

from litiv_amd import lbsp_compute_abs, synth
from oracle_lib import lib as orc

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _lbsp_gpu(img, t, packed=False):
    torch = _torch()
    h, w = img.shape[:2]
    c = 1 if img.ndim == 2 else img.shape[2]
    d_img = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    d_out = torch.empty((h, w, c), dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    lbsp_compute_abs(d_img.data_ptr(), w, h, c, t, d_out.data_ptr(), packed=packed)
    torch.cuda.synchronize()
    return d_out.cpu().numpy().view(np.uint16).reshape(img.shape)


def test_lbsp_known_answer():
    # modules/features2d/test/lbsp.cpp:22-46: LBSP(20) on the img2.png crop vs test_lbsp.bin, valid zone
    inp = np.load(os.path.join(GOLDEN, "lbsp_kat_input.npy"))
    exp = np.load(os.path.join(GOLDEN, "lbsp_kat_expected.npy"))
    for packed in (False, True):
        out = _lbsp_gpu(inp, 20, packed)
        assert np.array_equal(out[2:63, 2:63], exp[2:63, 2:63]), packed


@pytest.mark.parametrize("shape,t", [((37, 53, 3), 0), ((64, 65, 1), 7), ((130, 200, 3), 35), ((5, 5, 1), 255)])
def test_lbsp_random_vs_oracle(shape, t):
    rng = np.random.default_rng(sum(shape) + t)
    img = rng.integers(0, 256, size=shape, dtype=np.uint8)
    ref = np.zeros(shape, np.uint16)
    orc().orc_lbsp_compute_abs(img, shape[1], shape[0], shape[2], t, ref)
    for packed in (False, True):
        assert np.array_equal(_lbsp_gpu(img, t, packed), ref), packed


@pytest.mark.parametrize("t", [0, 1, 2, 127, 253, 254, 255, 256])
def test_lbsp_extreme_values_vs_oracle(t):
    # centres and taps at the ends of the byte range: a zero centre (which the masked quad SAD of
    # the packed form treats through the complemented taps), differences of exactly t and t + 1, and
    # 255 - 0, through both forms of the threshold test
    rng = np.random.default_rng(1000 + t)
    vals = np.array([0, 1, 2, 126, 127, 128, 253, 254, 255], np.uint8)
    for shape in ((48, 70, 3), (33, 29, 1)):
        img = vals[rng.integers(0, len(vals), size=shape)]
        ref = np.zeros(shape, np.uint16)
        orc().orc_lbsp_compute_abs(img, shape[1], shape[0], shape[2], t, ref)
        for packed in (False, True):
            assert np.array_equal(_lbsp_gpu(img, t, packed), ref), (shape, packed)


@pytest.mark.parametrize("kind,over", [(synth.QVGA, {}), (synth.THERMAL, {}), (synth.BATCH64, {"stream": 50})])
def test_synth_device_equals_host(kind, over):
    torch = _torch()
    stream = over.pop("stream", 0)
    s = synth.Sequence(synth.preset(kind, stream))
    n = 3
    buf = torch.empty(n * s.frame_bytes, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    s.frames_device(10, n, buf.data_ptr())
    torch.cuda.synchronize()
    dev = buf.cpu().numpy()
    for i in range(n):
        host = s.frame(10 + i)
        assert np.array_equal(dev[i * s.frame_bytes:(i + 1) * s.frame_bytes], host.reshape(-1)), i


def _oracle_post(raw, k):
    """close -> flood holes -> 7x7 erosion -> median, composed from the oracle's primitives in the
    order of its SuBSENSE step (orc_bgs.c, the post-processing block after the raw mask)."""
    h, w = raw.shape
    tmp, pre, holes, er, out = (np.empty_like(raw) for _ in range(5))
    orc().orc_morph(raw, tmp, w, h, 1, 1)
    orc().orc_morph(tmp, pre, w, h, 1, 0)
    orc().orc_flood_holes(pre, holes, w, h)
    orc().orc_morph(pre, er, w, h, 3, 0)
    mask = raw | holes | er
    orc().orc_median_binary(mask, out, w, h, k)
    return out, int((holes > 0).sum())


def _annuli(rng, h, w, n, rmax):
    m = np.zeros((h, w), np.uint8)
    yy, xx = np.ogrid[:h, :w]
    for _ in range(n):
        cx, cy, r = rng.integers(0, w), rng.integers(0, h), rng.integers(5, rmax)
        d = (xx - cx) ** 2 + (yy - cy) ** 2
        m[(d < r * r) & (d >= (r - 2) ** 2)] = 255
    return m


def _maze(h, w, sealed):
    """concentric square walls 4 px apart (corridors of 3 survive the 3x3 closing) with a 3-px gap
    per wall on alternating sides: the core reaches the outside only along the whole zigzag"""
    m = np.zeros((h, w), np.uint8)
    cy, cx = h // 2, w // 2
    for i, r in enumerate(range(6, min(h, w) // 2 - 4, 4)):
        y0, y1, x0, x1 = cy - r, cy + r, cx - r, cx + r
        m[y0, x0:x1 + 1] = m[y1, x0:x1 + 1] = 255
        m[y0:y1 + 1, x0] = m[y0:y1 + 1, x1] = 255
        if not (sealed and i == 0):
            if i % 2: m[cy - 1:cy + 2, x0] = 0
            else: m[cy - 1:cy + 2, x1] = 0
    return m


def _stripes(h, w):
    m = np.zeros((h, w), np.uint8)
    m[:, 3::4] = 255
    m[::37, :] = 0
    return m


def _postprocess_cases():
    rng = np.random.default_rng(11)
    cases = []
    for (w, h) in [(130, 70), (64, 64), (333, 190), (50, 40), (320, 240)]:
        noise = np.where(rng.random((h, w)) < 0.3, 255, 0).astype(np.uint8)
        cases.append(("noise+annuli %dx%d" % (w, h), noise | _annuli(rng, h, w, 8, 40)))
    cases.append(("annuli across zones", _annuli(rng, 300, 400, 6, 120)))
    cases.append(("maze open", _maze(256, 300, False)))
    cases.append(("maze sealed", _maze(256, 300, True)))
    cases.append(("stripes", _stripes(200, 260)))
    fg00 = _annuli(rng, 150, 170, 5, 50)
    fg00[:3, :3] = 255
    cases.append(("pixel (0,0) foreground", fg00))
    return cases


@pytest.mark.parametrize("name,raw", _postprocess_cases(), ids=lambda x: x if isinstance(x, str) else "")
def test_hole_fill_and_median_vs_oracle(name, raw):
    from litiv_amd import BackgroundSubtractorSuBSENSE
    _torch()
    h, w = raw.shape
    sub = BackgroundSubtractorSuBSENSE()
    sub.initialize(np.zeros((h, w, 3), np.uint8))
    k = int(sub.scalars()["median_k"])
    got = sub.debug_postprocess(raw)
    exp, n_holes = _oracle_post(raw, k)
    assert np.array_equal(got, exp), "%s: %d px differ (k=%d, %d hole px)" % (name, int((got != exp).sum()), k, n_holes)
