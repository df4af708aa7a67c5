"""The evaluator's entry point that allocates nothing (lvbgs_binclassif_accumulate_noalloc): its
partial counts live in the library's static device array instead of a per-call device buffer.
Called through the C ABI, its counts must equal lvbgs_binclassif_accumulate's on the same masks on
two calls in a row (the second on a smaller input, so the first call's partial records sit past the
second's in the shared array), and it must refuse a frame larger than its array holds before any
launch."""
import ctypes as C

import numpy as np
import pytest

from litiv_amd.bgs import BINCLASSIF_NOALLOC_MAX_PIXELS, ERR_TOO_LARGE, OK, BinClassifCounts, lib

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _counts(c):
    return tuple(getattr(c, k) for k in ("tp", "tn", "fp", "fn", "se", "dc"))


def test_noalloc_counts_equal_the_allocating_entry_point():
    torch = _torch()
    L = lib()
    rng = np.random.default_rng(11)
    # largest first: the second call's partial records overwrite only a prefix of the first's
    for (h, w, with_roi, stride_pad) in [(1080, 1920, True, 0), (217, 333, False, 7)]:
        stride = w + stride_pad
        res = torch.from_numpy(rng.choice(np.array([0, 255, 7], np.uint8), size=(h, stride))).cuda()
        gt = torch.from_numpy(rng.choice(np.array([0, 255, 50, 85, 170], np.uint8), size=(h, stride))).cuda()
        roi = torch.from_numpy(rng.choice(np.array([0, 255], np.uint8), size=(h, stride))).cuda() if with_roi else None
        torch.cuda.synchronize()
        print("noalloc binclassif %dx%d stride %d" % (w, h, stride), flush=True)
        args = (C.c_void_p(res.data_ptr()), C.c_void_p(gt.data_ptr()), C.c_void_p(roi.data_ptr() if with_roi else 0),
                w, h, stride)
        a, b = BinClassifCounts(), BinClassifCounts()
        assert L.lvbgs_binclassif_accumulate(*args, C.byref(a), None) == OK
        assert L.lvbgs_binclassif_accumulate_noalloc(*args, C.byref(b), None) == OK
        assert _counts(a) == _counts(b)
        assert sum(_counts(b)) - b.se == w * h


def test_noalloc_refuses_frames_beyond_its_array():
    torch = _torch()
    L = lib()
    res = torch.zeros(16, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    c = BinClassifCounts()
    # height 1, width just past the limit: refused before any launch or read of the buffers
    w = BINCLASSIF_NOALLOC_MAX_PIXELS + 1
    assert L.lvbgs_binclassif_accumulate_noalloc(C.c_void_p(res.data_ptr()), C.c_void_p(res.data_ptr()), None, w, 1, 0,
                                                 C.byref(c), None) == ERR_TOO_LARGE
    assert _counts(c) == (0,) * 6
