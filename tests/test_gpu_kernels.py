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


def _lbsp_gpu(img, t):
    torch = _torch()
    h, w = img.shape[:2]
    c = 1 if img.ndim == 2 else img.shape[2]
    d_img = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    d_out = torch.empty((h, w, c), dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    lbsp_compute_abs(d_img.data_ptr(), w, h, c, t, d_out.data_ptr())
    torch.cuda.synchronize()
    return d_out.cpu().numpy().view(np.uint16).reshape(img.shape)


def test_lbsp_known_answer():
    # modules/features2d/test/lbsp.cpp:22-46: LBSP(20) on the img2.png crop vs test_lbsp.bin, valid zone
    inp = np.load(os.path.join(GOLDEN, "lbsp_kat_input.npy"))
    exp = np.load(os.path.join(GOLDEN, "lbsp_kat_expected.npy"))
    out = _lbsp_gpu(inp, 20)
    assert np.array_equal(out[2:63, 2:63], exp[2:63, 2:63])


@pytest.mark.parametrize("shape,t", [((37, 53, 3), 0), ((64, 65, 1), 7), ((130, 200, 3), 35), ((5, 5, 1), 255)])
def test_lbsp_random_vs_oracle(shape, t):
    rng = np.random.default_rng(sum(shape) + t)
    img = rng.integers(0, 256, size=shape, dtype=np.uint8)
    out = _lbsp_gpu(img, t)
    ref = np.zeros(shape, np.uint16)
    orc().orc_lbsp_compute_abs(img, shape[1], shape[0], shape[2], t, ref)
    assert np.array_equal(out, ref)


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
