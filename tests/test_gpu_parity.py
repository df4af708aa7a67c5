"""GPU parity: the HIP SuBSENSE / LOBSTER against the sequential CPU oracle on the same inputs.

Bar: masks bit-exact on every frame; the whole model state (samples, FP32 feedback maps, u8 maps,
LUT, frame-level scalars) bit-exact after the sequence. Sizes are chosen so the oracle finishes in
seconds; the full-size cases are in test_gpu_fullsize.py.
"""
import numpy as np
import pytest

from litiv_amd import BackgroundSubtractorLOBSTER, BackgroundSubtractorSuBSENSE, synth
from oracle_lib import LOBSTER, SUBSENSE, Oracle

pytestmark = pytest.mark.gpu

SUB_STATE = ["samples_color", "samples_desc", "last_color", "last_desc", "last_fg", "lut", "T", "R", "V", "MLD",
             "DminLT", "DminST", "RawLT", "RawST", "FinLT", "FinST", "unstable", "blinks", "last_raw",
             "last_raw_blink", "last_dil_inv", "dsLT", "dsST"]
LOB_STATE = ["samples_color", "samples_desc", "last_color", "last_fg", "lut"]


def _pair(algo, **params):
    if algo == SUBSENSE:
        return BackgroundSubtractorSuBSENSE(**params), Oracle(SUBSENSE, **params)
    return BackgroundSubtractorLOBSTER(**params), Oracle(LOBSTER, **params)


def _describe_diff(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype.kind == "f":
        ai = a.view(np.int32).astype(np.int64)
        bi = b.view(np.int32).astype(np.int64)
        return "%d values differ, max ulp %d" % (int((a != b).sum()), int(np.abs(ai - bi).max()))
    return "%d values differ" % int((a != b).sum())


def run_sequence(algo, frames, lr=None, roi=None, state_names=None, check_scalars=True, schedule=None, **params):
    gpu, orc = _pair(algo, **params)
    if schedule is not None:
        gpu.set_match_schedule(*schedule[:2])
        gpu.set_fixpoint_schedule(*schedule[2:])
    gpu.initialize(frames[0], roi)
    orc.initialize(frames[0], roi)
    for t, f in enumerate(frames):
        kw = {} if lr is None else {"learningRate": lr}
        mg = gpu.apply(f, **kw)
        mo = orc.apply(f, lr)
        assert np.array_equal(mg, mo), "frame %d mask: %s" % (t, _describe_diff(mg, mo))
        eg, wg = gpu.stats()["examined"], gpu.stats()["writes"]
        eo, wo = orc.counters()
        assert (eg, wg) == (eo, wo), "frame %d counters gpu %r oracle %r" % (t, (eg, wg), (eo, wo))
    for name in state_names or (SUB_STATE if algo == SUBSENSE else LOB_STATE):
        sg, so = gpu.state(name), orc.state(name)
        assert np.array_equal(sg, so), "state %s: %s" % (name, _describe_diff(sg, so))
    if check_scalars and algo == LOBSTER:
        assert gpu.scalars()["apply_count"] == orc.scalars()["apply_count"]
    elif check_scalars:
        sg, so = gpu.scalars(), orc.scalars()
        for k in ("frame_idx", "frames_since_reset", "cooldown", "auto_reset", "lr_scaling", "use3x3", "median_k",
                  "n_roi", "lower_cap", "upper_cap", "refresh_seq"):
            assert sg[k] == so[k], "scalar %s gpu %r oracle %r" % (k, sg[k], so[k])
        assert np.float32(sg["last_nz_ratio"]) == np.float32(so["last_nz_ratio"])
    return gpu, orc


def seq(preset=synth.QVGA, n=30, **over):
    s = synth.Sequence(synth.preset(preset, **over))
    return [s.frame(t) for t in range(n)]


def test_rng_stream_matches_oracle():
    import oracle_lib
    frames = seq(n=1, width=32, height=24, n_objects=1, obj_min=4, obj_max=8)
    g = BackgroundSubtractorSuBSENSE(rng_seed=0x1234567890ABCDEF, stream_id=7)
    g.initialize(frames[0])
    for frame, phase, ev in ((1, 1, 0), (5, 2, 3), (0xFFFF1, 1, 0)):
        d = g.debug_rng(frame, phase, ev, 300)
        for p in (0, 1, 17, 299):
            for k in range(4):
                assert d[p, k] == oracle_lib.lib().orc_draw(0x1234567890ABCDEF, 7, frame, phase, ev, p, k)


def test_subsense_rgb_small_unscaled():
    # 96x72 < 320x240: learning-rate scaling off, 3x3 spread, median 9, caps 4/512
    run_sequence(SUBSENSE, seq(n=40, width=96, height=72, n_objects=2, obj_min=10, obj_max=24))


def test_subsense_rgb_qvga():
    run_sequence(SUBSENSE, seq(n=40))


def test_subsense_gray_qvga():
    run_sequence(SUBSENSE, seq(n=40, channels=1))


def test_subsense_rgb_ragged_with_area_downsample():
    # 330x246: width not a multiple of 4 (byte paths of the word-wise median/dilation), frame large
    # enough for learning-rate scaling, so the 8x8 area downsample runs with fractional cell weights
    run_sequence(SUBSENSE, seq(n=14, width=330, height=246, n_objects=3, obj_min=12, obj_max=50))


def test_subsense_gray_ragged_small():
    run_sequence(SUBSENSE, seq(n=24, width=101, height=77, channels=1, n_objects=2, obj_min=6, obj_max=20))


def test_subsense_rgb_vga_5x5_spread():
    # 640x480 > 2*QVGA: 5x5 diffusion and median 13
    run_sequence(SUBSENSE, seq(n=15, width=640, height=480, n_objects=4, obj_min=30, obj_max=80))


@pytest.mark.parametrize("schedule", [(1, 7, 0, 1), (255, 1024, 6, 256), (3, 1, 1, 1)])
def test_subsense_schedule_independent(schedule):
    # the match / fix-point schedule (sequential depth, cooperative grid, full-grid fix-point launches)
    # changes the order of evaluation, never the result
    run_sequence(SUBSENSE, seq(n=25, width=160, height=120), schedule=schedule)


@pytest.mark.parametrize("lr", [float("inf"), 3.0, 1.0])
def test_subsense_learning_rate_override(lr):
    run_sequence(SUBSENSE, seq(n=20, width=160, height=120), lr=lr)


def test_subsense_global_change_triggers_model_reset():
    # LT and ST frame-level averages are identical while frame_idx <= 25 (1/min(t,100) == 1/min(t,25)),
    # so the jump comes later; the ST/LT gap then reaches the reset threshold (15) within ~15 frames.
    frames = seq(n=52)
    for t in range(28, 52):
        frames[t] = np.clip(frames[t].astype(np.int16) + 120, 0, 255).astype(np.uint8)
    gpu, orc = run_sequence(SUBSENSE, frames)
    assert orc.scalars()["refresh_seq"] >= 2, "the brightness jump should trigger refreshModel(0.1)"


def test_subsense_custom_roi_and_getters():
    frames = seq(n=20, width=128, height=96)
    roi = np.zeros((96, 128), np.uint8)
    roi[10:80, 20:110] = 255
    roi[40:50, 50:60] = 0
    gpu, orc = run_sequence(SUBSENSE, frames, roi=roi)
    assert np.array_equal(gpu.getBackgroundImage(), orc.background_image())
    assert np.array_equal(gpu.getBackgroundDescriptorsImage(), orc.background_descriptors())
    assert np.array_equal(gpu.getROICopy(), orc.state("roi"))


def test_subsense_refresh_model_api():
    frames = seq(n=10, width=96, height=72)
    gpu, orc = _pair(SUBSENSE)
    gpu.initialize(frames[0])
    orc.initialize(frames[0])
    for t in range(10):
        assert np.array_equal(gpu.apply(frames[t]), orc.apply(frames[t]))
        if t in (3, 6):
            gpu.refreshModel(0.3, t == 6)
            orc.refresh_model(0.3, t == 6)
    for name in ("samples_color", "samples_desc"):
        assert np.array_equal(gpu.state(name), orc.state(name)), name


def test_lobster_gray_thermal():
    run_sequence(LOBSTER, seq(synth.THERMAL, n=30, width=160, height=128))


def test_lobster_rgb():
    run_sequence(LOBSTER, seq(n=30, width=128, height=96), lr=4.0)


def test_lobster_refresh_and_background():
    frames = seq(n=12, width=96, height=72)
    gpu, orc = _pair(LOBSTER)
    gpu.initialize(frames[0])
    orc.initialize(frames[0])
    for t in range(12):
        assert np.array_equal(gpu.apply(frames[t]), orc.apply(frames[t]))
        if t == 5:
            gpu.refreshModel(0.5)
            orc.refresh_model(0.5)
    assert np.array_equal(gpu.getBackgroundImage(), orc.background_image())
    assert np.array_equal(gpu.getBackgroundDescriptorsImage(), orc.background_descriptors())


def _continue(gpu, orc, frames, names=SUB_STATE):
    for t, f in enumerate(frames):
        mg, mo = gpu.apply(f), orc.apply(f)
        assert np.array_equal(mg, mo), "frame %d mask: %s" % (t, _describe_diff(mg, mo))
    for name in names:
        sg, so = gpu.state(name), orc.state(name)
        assert np.array_equal(sg, so), "state %s: %s" % (name, _describe_diff(sg, so))
    sg, so = gpu.scalars(), orc.scalars()
    for k in ("frame_idx", "frames_since_reset", "cooldown", "auto_reset", "lr_scaling", "use3x3", "median_k",
              "n_roi", "lower_cap", "upper_cap", "refresh_seq"):
        assert sg[k] == so[k], "scalar %s gpu %r oracle %r" % (k, sg[k], so[k])


def test_reinitialize_without_roi_recounts_the_stored_roi():
    # A left band of 158 of 320 columns counts 160 columns with its dilated ring: half the frame, so
    # learning-rate scaling is on (BackgroundSubtractorSuBSENSE.cpp:112). Re-initializing without a
    # ROI reuses the stored, border-trimmed ROI (BackgroundSubtractionUtils.cpp:87-88) and recounts it
    # (:97): 236 x 158 pixels, under half, so the second initialize turns scaling off.
    frames = seq(n=12)
    roi = np.zeros((240, 320), np.uint8)
    roi[:, :158] = 255
    gpu, orc = _pair(SUBSENSE)
    gpu.initialize(frames[0], roi)
    orc.initialize(frames[0], roi)
    assert gpu.scalars()["lr_scaling"] == 1 == orc.scalars()["lr_scaling"]
    _continue(gpu, orc, frames[:6])
    gpu.initialize(frames[6])
    orc.initialize(frames[6])
    assert gpu.scalars()["lr_scaling"] == 0 == orc.scalars()["lr_scaling"]
    assert gpu.scalars()["n_roi"] == 236 * 158
    _continue(gpu, orc, frames[6:])


def test_set_roi_on_an_initialized_context():
    # setROI after initialize re-initializes from the current background image with the new ROI
    # (BackgroundSubtractionUtils.cpp:38-48); the oracle does the same two calls
    frames = seq(n=16, width=160, height=120)
    gpu, orc = _pair(SUBSENSE)
    gpu.initialize(frames[0])
    orc.initialize(frames[0])
    _continue(gpu, orc, frames[:6])
    roi = np.zeros((120, 160), np.uint8)
    roi[20:100, 30:140] = 255
    gpu.setROI(roi)
    bg = orc.background_image()
    orc.initialize(bg, gpu.validateROI(roi))
    assert np.array_equal(gpu.getROICopy(), orc.state("roi"))
    _continue(gpu, orc, frames[6:])
