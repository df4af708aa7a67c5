"""Checkpoint / resume on the GPU: a model saved after k frames (lvbgs_checkpoint_save, one binary
archive per state buffer) and loaded into a fresh context (lvbgs_checkpoint_load) continues the
sequence exactly. Every mask of the resumed run equals the uninterrupted run's and the CPU oracle's
(which never stops), and the final model states are identical. Contexts whose arena was sized for
another frame, ROIs and one-channel frames are covered, mismatched parameters are refused, and a load
that fails part-way leaves the context uninitialized."""
import os

import numpy as np
import pytest

from litiv_amd import BackgroundSubtractorLOBSTER, BackgroundSubtractorSuBSENSE, synth
from litiv_amd.bgs import ERR_STATE, LvbgsError, read_archive
from oracle_lib import LOBSTER, SUBSENSE, Oracle

pytestmark = pytest.mark.gpu

SUB_STATE = ["samples_color", "samples_desc", "last_color", "last_desc", "last_fg", "lut", "T", "R", "V", "MLD",
             "DminLT", "DminST", "RawLT", "RawST", "FinLT", "FinST", "unstable", "blinks", "last_raw",
             "last_raw_blink", "last_dil_inv", "dsLT", "dsST", "roi"]
LOB_STATE = ["samples_color", "samples_desc", "last_color", "last_fg", "lut", "roi"]

CASES = {
    # name: (algorithm, synthetic config, frames, checkpoint after, ROI)
    "subsense_qvga_rgb": (SUBSENSE, dict(kind=synth.QVGA), 60, 23, False),
    "subsense_qvga_rgb_roi": (SUBSENSE, dict(kind=synth.QVGA, seed=5), 40, 17, True),
    "subsense_grey": (SUBSENSE, dict(kind=synth.QVGA, width=200, height=150, channels=1), 40, 12, False),
    "lobster_rgb": (LOBSTER, dict(kind=synth.QVGA, width=256, height=192), 40, 15, False),
}


def _roi(h, w):
    r = np.zeros((h, w), np.uint8)
    r[h // 6:h - h // 5, w // 8:w - w // 7] = 255
    r[h // 3:h // 2, w // 3:w // 2] = 0
    return r


def _parse_archive(path):
    """Header and element bytes of a MatArchive_BINARY file: int32 type, uint64 element size, uint64
    element count, int32 dims, int32 sizes, data (modules/utils/src/opencv.cpp:514-531)."""
    import struct
    raw = open(path, "rb").read()
    typ, esz, cnt, nd = struct.unpack_from("<iQQi", raw, 0)
    sizes = struct.unpack_from("<%di" % nd, raw, 24)
    body = raw[24 + 4 * nd:]
    assert len(body) == esz * cnt and cnt == int(np.prod(sizes))
    return typ, sizes, body


def _cls(algo):
    return BackgroundSubtractorSuBSENSE if algo == SUBSENSE else BackgroundSubtractorLOBSTER


def _state(b, names):
    return {n: b.state(n) for n in names}


@pytest.mark.parametrize("name", sorted(CASES))
def test_checkpoint_resume_is_exact(name, tmp_path):
    algo, spec, n, k, with_roi = CASES[name]
    spec = dict(spec)
    seq = synth.Sequence(synth.preset(spec.pop("kind"), **spec))
    frames = [seq.frame(t) for t in range(n)]
    h, w = frames[0].shape[:2]
    roi = _roi(h, w) if with_roi else None
    lr = 0.0 if algo == SUBSENSE else 16.0
    print("case", name, frames[0].shape, "checkpoint after", k, flush=True)

    orc = Oracle(algo, stream_id=3)
    orc.initialize(frames[0], roi)
    expected = [orc.apply(f, lr) for f in frames]

    a = _cls(algo)(stream_id=3)
    a.initialize(frames[0], roi)
    for t in range(k):
        assert np.array_equal(a.apply(frames[t], learningRate=lr), expected[t]), "frame %d before the checkpoint" % t
    ck = str(tmp_path / "ck")
    a.save_checkpoint(ck)
    meta, _ = read_archive(os.path.join(ck, "meta.bin"))
    assert int(meta[2]) == w and int(meta[3]) == h
    # every file is a plain binary matrix archive: an independent parse of the bytes gives the state
    # getters' arrays (the sample planes as N x H x W matrices of C-channel elements)
    names = SUB_STATE if algo == SUBSENSE else LOB_STATE
    for nme in names:
        typ, shape, body = _parse_archive(os.path.join(ck, nme + ".bin"))
        ref = a.state(nme)
        depth = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 2, np.dtype(np.float32): 5}[ref.dtype]
        assert typ & 7 == depth, nme
        assert body == ref.tobytes(), nme
        assert int(np.prod(shape)) * ((typ >> 3) + 1) == ref.size, nme

    b = _cls(algo)(stream_id=3)
    other = frames[0][: h // 2, : w // 2]
    b.initialize(np.ascontiguousarray(other))      # an arena sized for another frame, then replaced
    b.load_checkpoint(ck)
    assert b.shape == a.shape
    for t in range(k, n):
        ma, mb = a.apply(frames[t], learningRate=lr), b.apply(frames[t], learningRate=lr)
        assert np.array_equal(mb, ma), "frame %d resumed vs uninterrupted" % t
        assert np.array_equal(mb, expected[t]), "frame %d resumed vs oracle" % t
    names = SUB_STATE if algo == SUBSENSE else LOB_STATE
    sa, sb = _state(a, names), _state(b, names)
    for nme in names:
        assert np.array_equal(sa[nme], sb[nme]), nme
    assert a.scalars() == b.scalars()


def test_checkpoint_of_a_model_with_other_parameters_is_refused(tmp_path):
    seq = synth.Sequence(synth.preset(synth.QVGA, width=128, height=96))
    f0 = seq.frame(0)
    a = BackgroundSubtractorSuBSENSE(stream_id=1)
    with pytest.raises(LvbgsError):
        a.save_checkpoint(str(tmp_path / "none"))    # not initialized
    a.initialize(f0)
    a.apply(seq.frame(1))
    ck = str(tmp_path / "ck")
    a.save_checkpoint(ck)
    for kw in (dict(stream_id=2), dict(stream_id=1, rng_seed=99), dict(stream_id=1, bg_samples=40)):
        b = BackgroundSubtractorSuBSENSE(**kw)
        with pytest.raises(LvbgsError):
            b.load_checkpoint(ck)
    with pytest.raises(LvbgsError):
        BackgroundSubtractorLOBSTER(stream_id=1).load_checkpoint(ck)
    # a checkpoint missing one buffer: the load fails after the geometry, the context is left
    # uninitialized, and a complete checkpoint loads into it afterwards
    import shutil
    broken = str(tmp_path / "broken")
    shutil.copytree(ck, broken)
    os.remove(os.path.join(broken, "T.bin"))
    b = BackgroundSubtractorSuBSENSE(stream_id=1)
    b.initialize(f0)
    with pytest.raises(LvbgsError):
        b.load_checkpoint(broken)
    with pytest.raises(LvbgsError) as e:
        b.apply(seq.frame(2))   # reaches the library, which reports the uninitialized model
    assert e.value.code == ERR_STATE
    b.load_checkpoint(ck)
    assert np.array_equal(b.apply(seq.frame(2)), a.apply(seq.frame(2)))
