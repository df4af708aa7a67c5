"""Pins of the oracle's whole apply() (CPU only): closed forms derived by hand from the reference's
lines, on scenes simple enough that the answer follows from the constants without running anything.

Each case states its derivation. Together they fix the match thresholds and their flip points
(SuBSENSE `BackgroundSubtractorSuBSENSE.cpp:355-395` 3ch and `:222-253` 1ch, LOBSTER
`BackgroundSubtractorLOBSTER.cpp:480-495, 532-553`), the feedback maps' constants and branch order
(SUB:396-474), the post-processing's blink bookkeeping (SUB:536-549), the frame-level controller
(SUB:555-611), refreshModel (SUB:80-105) and initialize's scaling table (SUB:112-128,
BackgroundSubtractionUtils.cpp:80-103), so that a one-constant slip in the oracle (a wrong offset,
divisor, cap, increment, compare or branch order) fails at least one of them.

Frames are flat unless stated: on a flat frame every LBSP descriptor is 0 (no tap differs from the
centre), every model sample holds the init colour with descriptor 0, and the inter descriptor of a
flat frame of colour b+d against a sample of colour b is 0xFFFF on that channel iff d > LUT[b]
(`LBSP.hpp:193-224`), else 0. The LUT is `sat(t*0.333f)` (3ch) / `sat(t*0.333f/3)` (1ch)
(`BackgroundSubtractorLBSP.cpp:29-30, 42-43`), itself pinned in test_oracle_pins.py.
"""
import numpy as np
import pytest

from oracle_lib import LOBSTER, SUBSENSE, Oracle

f32 = np.float32
H, W = 40, 48                 # 1920 px < 320x240: no learning-rate scaling, caps 4/512, 3x3 spread
CY, CX = 20, 24               # a pixel far from the borders


def _lut(c):
    t = np.arange(256, dtype=np.float32)
    v = t * f32(0.333) if c == 3 else (t * f32(0.333)) / f32(3)
    return np.clip(np.rint(v), 0, 255).astype(int)


LUT3, LUT1 = _lut(3), _lut(1)


def _flat(c, v, h=H, w=W):
    if c == 1:
        return np.full((h, w), v, np.uint8)
    img = np.empty((h, w, 3), np.uint8)
    img[...] = v
    return img


def _fg_after_one_step(algo, c, b, offs, **kw):
    """initialize on flat b, apply one flat frame b+offs (per channel); True iff the interior is
    foreground. The decision is the same for every interior pixel (all see the same taps and the
    same samples), which is asserted."""
    o = Oracle(algo, **kw)
    o.initialize(_flat(c, b))
    cur = np.array([b, b, b]) + np.array(offs if c == 3 else [offs[0], 0, 0])
    m = o.apply(_flat(c, tuple(int(v) for v in cur) if c == 3 else int(cur[0])))
    if algo == SUBSENSE:
        raw = o.state("last_raw")[2:-2, 2:-2]
        assert raw.min() == raw.max()
        return bool(raw[0, 0])
    # LOBSTER keeps no raw mask: the median of a uniform interior decision is that decision far
    # from the border
    return bool(m[CY, CX])


def _flip(algo, c, b, offs_of_d, lo, hi, **kw):
    """smallest d in [lo, hi] for which the step is foreground; asserts monotonicity around it"""
    flip = None
    for d in range(lo, hi + 1):
        fg = _fg_after_one_step(algo, c, b, offs_of_d(d), **kw)
        if flip is None and fg:
            flip = d
        assert fg == (flip is not None), "decision not monotone in d"
    return flip


def one(d):
    return (d, 0, 0)


def two(d):
    return (d, d, 0)


def three(d):
    return (d, d, d)


# ---- (a) match thresholds --------------------------------------------------------------------------
# SuBSENSE 3ch at the first apply: R = 1, unstable = 0, so
#   colour threshold = (size_t)(1*30 - 30/5) = 24, total 72, per channel 72/2 = 36   (SUB:355, :357-359)
#   descriptor threshold = (1 << floor(1+0.5)) + 3 = 5, total 15                     (SUB:356, :358)
#   per channel: fail if L1 > 36; desc = (intra + inter)/2; sum = min((desc/2)*(255/16) + L1, 255),
#   fail if sum > 36 (SUB:370-383); totals compared after the three channels (SUB:384-385).
# One channel offset by d: if d <= LUT[b], desc = 0 and sum = d, so foreground from d = 37;
# if d > LUT[b] the inter descriptor is 0xFFFF, desc = 8, sum = 60 + d > 36: foreground from LUT[b]+1.
@pytest.mark.parametrize("b", [150, 60])
def test_subsense_rgb_single_channel_flip(b):
    expected = 37 if LUT3[b] >= 36 else LUT3[b] + 1
    assert (LUT3[150], LUT3[60]) == (50, 20)
    assert _flip(SUBSENSE, 3, b, one, expected - 3, expected + 2) == expected


def test_subsense_rgb_total_colour_flip():
    # three channels offset by d <= 36 < LUT[150]: per channel sum = d, total 3d <= 72 -> d <= 24
    assert _flip(SUBSENSE, 3, 150, three, 21, 27) == 25


def test_subsense_rgb_descriptor_total_and_sum_scale():
    # color_dist_threshold 100: colour threshold (size_t)(100 - 20) = 80, total 240, per channel 120.
    # b = 60, d = 30 > LUT[60] = 20: each offset channel has desc 8 and sum 4*15 + 30 = 90 <= 120.
    assert not _fg_after_one_step(SUBSENSE, 3, 60, one(30), color_dist_threshold=100)   # total desc 8
    assert _fg_after_one_step(SUBSENSE, 3, 60, two(30), color_dist_threshold=100)       # 16 > 3*5
    # descriptor offset 4: threshold 2 + 4 = 6, total 18 >= 16
    assert not _fg_after_one_step(SUBSENSE, 3, 60, two(30), color_dist_threshold=100, desc_dist_threshold=4)
    # the sum's (desc/2)*(255/16) = 60: one channel flips where 60 + d exceeds 120
    assert _flip(SUBSENSE, 3, 60, one, 57, 63, color_dist_threshold=100) == 61


# SuBSENSE 1ch at the first apply: colour threshold (size_t)(1*30 - 6)/2 = 12, descriptor 5 (SUB:222-223);
# fail if L1 > 12, desc = (intra+inter)/2 > 5, or min((desc/4)*15 + L1, 255) > 12 (SUB:233-245).
@pytest.mark.parametrize("b", [150, 60])
def test_subsense_gray_flip(b):
    assert (LUT1[150], LUT1[60]) == (17, 7)
    expected = 13 if LUT1[b] >= 12 else LUT1[b] + 1   # desc 8 > 5 once d > LUT[b]
    assert _flip(SUBSENSE, 1, b, one, expected - 3, expected + 2) == expected


def test_subsense_gray_descriptor_and_sum_scale():
    # colour threshold (size_t)(100 - 20)/2 = 40; with descriptor offset 6 the threshold is 2 + 6 = 8,
    # so the desc = 8 of d > LUT[60] = 7 passes and the sum (8/4)*15 + d flips past 40 at d = 11
    assert _flip(SUBSENSE, 1, 60, one, 8, 14, color_dist_threshold=100, desc_dist_threshold=6) == 11
    # descriptor offset 5: threshold 7 < 8, foreground as soon as d > LUT[60]
    assert _flip(SUBSENSE, 1, 60, one, 5, 10, color_dist_threshold=100, desc_dist_threshold=5) == 8


# LOBSTER 3ch (LOB:532-553): per channel L1 <= 3*30/2 = 45 and inter-descriptor distance <= 3*4/2 = 6,
# totals <= 90 and <= 12. LOBSTER 1ch (LOB:480-495): L1 <= 30/2 = 15 and distance <= 4.
@pytest.mark.parametrize("b", [150, 20])
def test_lobster_rgb_single_channel_flip(b):
    expected = 46 if LUT3[b] >= 45 else LUT3[b] + 1   # distance 16 > 6 once d > LUT[b]
    assert _flip(LOBSTER, 3, b, one, expected - 3, expected + 2) == expected


def test_lobster_rgb_totals():
    # b = 200 (LUT 67): three channels of d <= 45 -> total 3d <= 90 -> d <= 30
    assert _flip(LOBSTER, 3, 200, three, 28, 33) == 31
    # descriptor distances: per channel 3*t/2, total 3*t; one channel at d = LUT[20]+1 has distance 16
    assert _fg_after_one_step(LOBSTER, 3, 20, one(8), desc_dist_threshold=10)         # 16 > 15
    assert not _fg_after_one_step(LOBSTER, 3, 20, one(8), desc_dist_threshold=11)     # 16 <= 16
    assert _fg_after_one_step(LOBSTER, 3, 20, three(8), desc_dist_threshold=11)       # 48 > 33
    assert not _fg_after_one_step(LOBSTER, 3, 20, three(8), desc_dist_threshold=16)   # 48 <= 48


@pytest.mark.parametrize("b", [150, 60])
def test_lobster_gray_flip(b):
    expected = 16 if LUT1[b] >= 15 else LUT1[b] + 1
    assert _flip(LOBSTER, 1, b, one, expected - 3, expected + 2) == expected


# ---- (b) feedback maps -----------------------------------------------------------------------------
def _v_static(frames, v=f32(10)):
    # SUB:466-470 on a background pixel with last_fg = 0 and unstable = 0: v -= 0.1 while v > 0.1,
    # clamped at 0.1
    out = []
    for _ in range(frames):
        if v > f32(0.1):
            v = f32(v - f32(0.1))
            if v < f32(0.1):
                v = f32(0.1)
        out.append(v)
    return out


@pytest.mark.parametrize("c,size,lower", [(3, (H, W), 4.0), (1, (H, W), 4.0), (3, (240, 320), 2.0)])
def test_static_scene_feedback_trajectory(c, size, lower):
    # Every pixel matches every sample at distance 0: min distances 0, D_min LT/ST and the raw
    # segmentation averages stay 0 (SUB:414-419); T never leaves its lower cap (decrement branch
    # guarded by T > cap, SUB:457-461); R < (1 + 2*0)^2 is false, so R -= 0.01/v and clamps to 1
    # (SUB:471-476); v decreases by 0.1 per frame down to 0.1. Lower cap 4 below 320x240, 2 at it.
    h, w = size
    o = Oracle(SUBSENSE)
    f = _flat(c, 150, h, w)
    o.initialize(f)
    vs = _v_static(110)
    roi = np.zeros((h, w), bool)
    roi[2:-2, 2:-2] = True
    for k in range(110):
        assert not o.apply(f).any()
        if k in (0, 1, 50, 98, 99, 109):
            V = o.state("V")
            assert (V[roi] == vs[k]).all() and (V[~roi] == f32(10)).all()
    assert (o.state("R") == f32(1)).all()
    assert (o.state("T") == f32(lower)).all()
    for name in ("DminLT", "DminST", "RawLT", "RawST", "MLD", "FinLT", "FinST"):
        assert not o.state(name).any(), name
    assert vs[98] == f32(0.1) and vs[97] > f32(0.1)


def test_foreground_patch_feedback_values():
    # A 20x20 patch of (250,250,250) on flat 150 from the first apply; its centre never matches
    # (L1 = 100 > 36). Frame 1 (rolling factors 1): D_min = min(1, (765/765 + 48/48)/2 + 2/2) = 1,
    # raw averages 1 (SUB:399-404); the mean last distance uses the uchar-wrapped L1 of the three
    # channels, (3*100) mod 256 = 44 (math.hpp:228-234, SUB:396-397); T stays at 4 (last_fg = 0 and
    # min D >= 0.1 take the guarded decrement); v -= 0.1 (blinks 0, stable); R grows by
    # 0.01*(v - 0.1) since R = 1 < (1+2)^2.
    # Frame 2 (factors 1/2): last_fg = 1 at the centre, so T += 0.5/(max D * v) with v before its
    # update; blinks are 0 (the stale inverted dilation of SUB:549 is 0 after frame 1), so
    # v -= 0.1/4 (SUB:468). Frame 3: the same with the frame-2 values.
    o = Oracle(SUBSENSE)
    bg = _flat(3, 150)
    o.initialize(bg)
    fr = bg.copy()
    fr[10:30, 14:34] = 250
    nld = (f32(44) / f32(765) + f32(0) / f32(48)) / f32(2)
    v1 = f32(f32(10) - f32(0.1))
    r1 = f32(f32(1) + f32(f32(0.01) * f32(v1 - f32(0.1))))
    o.apply(fr)
    at = lambda name: o.state(name)[CY, CX]
    assert at("DminLT") == 1 and at("DminST") == 1 and at("RawLT") == 1 and at("RawST") == 1
    assert at("MLD") == nld
    assert at("T") == f32(4) and at("V") == v1 and at("R") == r1
    assert at("FinLT") == 1 and at("FinST") == 1 and o.state("last_fg")[CY, CX] == 255
    o.apply(fr)
    t2 = f32(f32(4) + f32(f32(0.5) / f32(f32(1) * v1)))
    v2 = f32(v1 - f32(f32(0.1) / f32(4)))
    r2 = f32(r1 + f32(f32(0.01) * f32(v2 - f32(0.1))))
    assert at("T") == t2 and at("V") == v2 and at("R") == r2
    assert at("MLD") == f32(f32(nld * f32(1 - f32(0.5))) + f32(0))
    assert at("DminLT") == 1 and at("RawLT") == 1
    o.apply(fr)
    a3 = f32(1) / f32(3)
    d3 = f32(f32(f32(1) * f32(f32(1) - a3)) + f32(f32(1) * a3))   # both D_min averages, 1 before and now
    t3 = f32(t2 + f32(f32(0.5) / f32(d3 * v2)))
    v3 = f32(v2 - f32(f32(0.1) / f32(4)))
    r3 = f32(r2 + f32(f32(0.01) * f32(v3 - f32(0.1))))
    assert at("T") == t3 and at("V") == v3 and at("R") == r3
    # the flat background far from the patch is untouched by the feedback
    assert o.state("T")[5, 5] == f32(4) and o.state("R")[5, 5] == f32(1)


@pytest.mark.parametrize("c", [3, 1])
def test_single_pixel_flicker_blinks_and_variation(c):
    # One pixel P alternates 250 (odd frames) and 150 on flat 150. P is foreground when lit and
    # background otherwise; a single raw pixel never survives the 9x9 median, so last_fg = 0.
    # Blinks (SUB:536-549): after frame 1 the stale inverted dilation (still 0) clears them; from
    # frame 2 on blinks[P] = 1. v at P: frame 1 v -= 0.1 (stable); frame 2 P is unstable (raw LT
    # average 1 > final LT average 0 + 0.1, SUB:365) and blinks are still 0: v -= 0.1/2; frames 3
    # and 4 have max D >= 0.5 > 0.1 and blinks: v += 1 (SUB:462-465). R grows by 0.01*(v - 0.1) each
    # frame (min D > 0 keeps R under (1+2 min D)^2).
    # The 16 pixels that have P among their LBSP taps get one set bit per channel: too few for the
    # non-zero descriptor count (>= 4 bits over 3 channels, >= 2 for 1 channel; SUB:477, :327), so
    # the frame's ratio is exactly 1 / (ROI pixels) on lit frames (P itself has 16 bits a channel).
    o = Oracle(SUBSENSE)
    bg = _flat(c, 150)
    o.initialize(bg)
    lit = bg.copy()
    lit[CY, CX] = 250
    n_roi = (H - 4) * (W - 4)
    v = [f32(f32(10) - f32(0.1))]
    v.append(f32(v[0] - f32(f32(0.1) / f32(2))))
    v.append(f32(v[1] + f32(1)))
    v.append(f32(v[2] + f32(1)))
    r = f32(1)
    for k in range(4):
        m = o.apply(lit if k % 2 == 0 else bg)
        assert not m.any()
        r = f32(r + f32(f32(0.01) * f32(v[k] - f32(0.1))))
        assert o.state("V")[CY, CX] == v[k], k
        assert o.state("R")[CY, CX] == r, k
        blinks = o.state("blinks")
        assert blinks[CY, CX] == (0 if k == 0 else 255) and blinks.sum() == blinks[CY, CX]
        nz = o.scalars()["last_nz_ratio"]
        assert f32(nz) == (f32(1) / f32(n_roi) if k % 2 == 0 else f32(0)), k
        assert o.state("T")[CY, CX] == f32(4)


# ---- (c) frame-level controller ------------------------------------------------------------------
def _lut_floor():
    # SUB:555-559: lower bound sat(0 + ceil(t*0.333f/4))
    t = np.arange(256, dtype=np.float32)
    return np.clip(np.ceil((t * f32(0.333)) / f32(4)), 0, 255).astype(int)


def test_lut_decays_on_a_flat_scene():
    # no descriptor is non-zero: the ratio is 0 < 0.1 on every frame and the previous one (initially
    # 0), so each entry above its floor drops by one per frame from the first apply
    o = Oracle(SUBSENSE)
    f = _flat(3, 150)
    o.initialize(f)
    lo = _lut_floor()
    for k in range(1, 41):
        o.apply(f)
        if k in (1, 3, 40):
            assert np.array_equal(o.state("lut").astype(int), np.maximum(LUT3 - k, np.minimum(LUT3, lo)))
    assert (lo[255], lo[13]) == (22, 2)


def test_lut_grows_on_a_textured_scene():
    # uniform noise frames: ratio > 0.5, but the first frame's predecessor is 0, so the LUT moves
    # from the second apply on, by one per frame up to sat(255*0.333f) = 85 (SUB:560-564)
    rng = np.random.default_rng(7)
    o = Oracle(SUBSENSE)
    o.initialize(rng.integers(0, 256, (H, W, 3), dtype=np.uint8))
    for k in range(1, 5):
        o.apply(rng.integers(0, 256, (H, W, 3), dtype=np.uint8))
        assert o.scalars()["last_nz_ratio"] > 0.5
        assert np.array_equal(o.state("lut").astype(int), np.minimum(LUT3 + (k - 1), 85))


def test_brightness_step_resets_the_model_on_the_closed_form_frame():
    # 320x240 RGB (learning-rate scaling and automatic reset on). Frames 1-100 flat 100, then flat
    # 160. The 8x8 area mean of a flat frame is the flat value; the down-sampled averages are
    # EMAs with factors 1/min(t,100) and 1/min(t,25) (SUB:567-570), so n frames after the step
    # ST - LT = 60 (0.99^n - 0.96^n), equal in every cell, and the frame ratio is its integer part
    # (SUB:571-583). Reset when the ratio reaches 30/2 = 15 with no cooldown (SUB:586-591): n = 11.
    # Caps from ratio >= 7: max(2 >> (ratio/2), 1) and max(256 >> (ratio/2), 1) (SUB:601-604).
    def diff(n):
        return 60.0 * (0.99 ** n - 0.96 ** n)

    for n in range(1, 14):  # float32 EMAs of values <= 160 over ~110 frames stay within ~1e-3
        assert abs(diff(n) - round(diff(n))) > 0.05
    o = Oracle(SUBSENSE)
    a, b = _flat(3, 100, 240, 320), _flat(3, 160, 240, 320)
    o.initialize(a)
    for _ in range(100):
        o.apply(a)
    s = o.scalars()
    assert (s["lower_cap"], s["upper_cap"], s["frames_since_reset"], s["auto_reset"]) == (2, 256, 100, 1)
    for n in range(1, 13):
        o.apply(b)
        s = o.scalars()
        ratio = int(diff(n))
        if n < 11:
            assert s["frames_since_reset"] == 100 + n and s["cooldown"] == 0 and s["refresh_seq"] == 1, n
        elif n == 11:
            # reset: T := 1 everywhere, cooldown 100/4 = 25 then decremented once this frame
            assert s["frames_since_reset"] == 0 and s["cooldown"] == 24 and s["refresh_seq"] == 2
            assert (o.state("T") == f32(1)).all()
        else:
            assert s["frames_since_reset"] == 1 and s["cooldown"] == 23 and s["refresh_seq"] == 2
        if ratio >= 7:
            k = ratio // 2
            assert (s["lower_cap"], s["upper_cap"]) == (max(2 >> k, 1), max(256 >> k, 1)), n
        else:
            assert (s["lower_cap"], s["upper_cap"]) == (2, 256), n


def test_automatic_reset_turns_off_after_1000_quiet_frames():
    # SUB:585-587: the counter grows by one per quiet frame; the frame that finds it above 1000
    # disables the automatic reset (320x240 grey keeps it cheap)
    o = Oracle(SUBSENSE)
    f = _flat(1, 120, 240, 320)
    o.initialize(f)
    for _ in range(1001):
        o.apply(f)
    s = o.scalars()
    assert (s["auto_reset"], s["frames_since_reset"]) == (1, 1001)
    o.apply(f)
    s = o.scalars()
    assert (s["auto_reset"], s["frames_since_reset"]) == (0, 1001)


# ---- refreshModel ---------------------------------------------------------------------------------
@pytest.mark.parametrize("frac,count", [(0.1, 5), (0.33, 16), (1.0, 50)])
def test_refresh_model_rewrites_a_common_window_of_slots(frac, count):
    # SUB:80-105: (size_t)(frac*50) consecutive slots from one start (0 for frac = 1) modulo 50, taken
    # from the last colour frame at 7x7 positions (all inside the flat ROI here). The apply at 110
    # with lr = inf is background everywhere (L1 30 <= 72, 10 <= LUT[100] = 33) and writes nothing.
    o = Oracle(SUBSENSE)
    o.initialize(_flat(3, 100))
    assert not o.apply(_flat(3, 110), float("inf")).any()
    assert (o.state("samples_color")[:, 2:-2, 2:-2] == 100).all()   # samples outside the ROI stay 0
    o.refresh_model(frac, False)
    sc = o.state("samples_color")[:, 2:-2, 2:-2, 0]           # (N, h, w)
    new = sc == 110
    assert ((sc == 100) | new).all()
    per_slot = new.reshape(50, -1)
    assert (per_slot.all(axis=1) | ~per_slot.any(axis=1)).all()  # a slot changes at every pixel or none
    slots = np.flatnonzero(per_slot.all(axis=1))
    assert len(slots) == count
    start = 0 if count == 50 else [s for s in slots if (s - 1) % 50 not in slots][0]
    assert sorted(slots) == sorted((start + j) % 50 for j in range(count))


# ---- (d) initialize's scaling table ----------------------------------------------------------------
@pytest.mark.parametrize("w,h,scaled,use3x3,median", [
    (320, 240, 1, 1, 9),    # 1 + 0.5 -> 1, 9 + 1 = 10 -> even -> 9
    (360, 240, 1, 1, 9),    # 1.125 + 0.5 -> 1
    (400, 300, 1, 1, 11),   # 1.5625 + 0.5 -> 2
    (480, 320, 1, 1, 11),   # area 153600 = 2*76800: 3x3 spread kept (strict >)
    (482, 320, 1, 0, 11),
    (560, 420, 1, 0, 11),   # 3.0625 + 0.5 -> 3, 12 -> 11
    (640, 480, 1, 0, 13),   # 4 + 0.5 -> 4, 13
    (1920, 1080, 1, 0, 13),  # 27 + 9 -> min(.., 14) = 14 -> 13
    (319, 240, 0, 1, 9),    # below 320x240: no scaling, caps doubled
])
def test_initialize_scaling_table(w, h, scaled, use3x3, median):
    # SUB:112-128
    o = Oracle(SUBSENSE)
    o.initialize(_flat(1, 90, h, w))
    s = o.scalars()
    caps = (2, 256) if scaled else (4, 512)
    assert (s["lr_scaling"], s["auto_reset"], s["use3x3"], s["median_k"]) == (scaled, scaled, use3x3, median)
    assert (s["lower_cap"], s["upper_cap"]) == caps
    assert (o.state("T") == f32(caps[0])).all() and (o.state("R") == 1).all() and (o.state("V") == 10).all()


@pytest.mark.parametrize("cols,scaled", [(157, 0), (158, 1)])
def test_roi_count_includes_the_dilated_ring(cols, scaled):
    # BackgroundSubtractionUtils.cpp:92-97: the ROI is OR-ed with half its 2-iteration dilation
    # before counting, so a left band of `cols` columns counts cols + 2 columns; scaling needs at
    # least half the 320x240 frame (SUB:112): 159*240 < 38400 <= 160*240
    roi = np.zeros((240, 320), np.uint8)
    roi[:, :cols] = 255
    o = Oracle(SUBSENSE)
    o.initialize(_flat(1, 90, 240, 320), roi)
    assert o.scalars()["lr_scaling"] == scaled
    # re-initializing without a ROI reuses the stored, border-trimmed one (UTL:87-88), recounted:
    # rows 2..237 and columns 2..cols+1, under half the frame either way
    o.initialize(_flat(1, 90, 240, 320))
    s = o.scalars()
    assert s["lr_scaling"] == 0 and s["n_roi"] == 236 * cols


# ---- (b') feedback decrements, unstable regions, partial matches -----------------------------------
def _ema(x, a, target):
    # the per-pixel rolling averages: x*(1-a) + target*a in float32, as SUB:399-419 write them
    return f32(f32(x * f32(f32(1) - a)) + f32(target * a))


def test_t_decrement_and_mean_last_distance_after_a_short_patch():
    # 100 static frames bring v to its floor 0.1 (rolling factors then 1/100 and 1/25). Frame 101
    # lights the 20x20 patch: the centre is raw foreground with D_min LT/ST = 0.01/0.04, so
    # min D < 0.1 with the raw mask set increments T by 0.5/(max D * v) (SUB:451-455); the mean last
    # distance takes the short-term factor (SUB:397). Frame 102 (patch gone) increments again since
    # last_fg is set; frame 103 decrements by 0.25*v/max D (SUB:456-457), staying above the cap 4.
    o = Oracle(SUBSENSE)
    bg = _flat(3, 150)
    o.initialize(bg)
    for _ in range(100):
        o.apply(bg)
    fr = bg.copy()
    fr[10:30, 14:34] = 250
    aLT, aST, v = f32(1) / f32(100), f32(1) / f32(25), f32(0.1)
    at = lambda name: o.state(name)[CY, CX]
    assert at("V") == v
    o.apply(fr)
    dlt, dst = _ema(f32(0), aLT, f32(1)), _ema(f32(0), aST, f32(1))
    t = f32(f32(4) + f32(f32(0.5) / f32(max(dlt, dst) * v)))
    nld = (f32(44) / f32(765) + f32(0) / f32(48)) / f32(2)
    assert at("DminLT") == dlt and at("DminST") == dst and at("T") == t and at("V") == v and at("R") == 1
    assert at("MLD") == _ema(f32(0), aST, nld)
    o.apply(bg)
    dlt, dst = _ema(dlt, aLT, f32(0)), _ema(dst, aST, f32(0))
    t = f32(t + f32(f32(0.5) / f32(max(dlt, dst) * v)))
    assert at("T") == t
    o.apply(bg)
    dlt, dst = _ema(dlt, aLT, f32(0)), _ema(dst, aST, f32(0))
    t = f32(t - f32(f32(f32(0.25) * v) / max(dlt, dst)))
    assert at("T") == t and t > f32(100)


def test_r_decrement_when_the_distance_average_falls():
    # lr = inf freezes the samples at 150. Frame 1 at 155 matches with min sum 3*5 = 15, min desc 0:
    # D_min = (15/765)/2 ~ 0.0098 (factors 1), R = 1 < (1 + 2 D)^2 so R += 0.01*(v - 0.1). Frames 2, 3
    # at 150 match at distance 0: D_min halves then falls to a third, (1 + 2 D)^2 < R and R -= 0.01/v
    # (SUB:471-476); v loses 0.1 a frame (stable: R <= 3, raw averages 0).
    o = Oracle(SUBSENSE)
    o.initialize(_flat(3, 150))
    inf = float("inf")
    o.apply(_flat(3, 155), inf)
    d = (f32(15) / f32(765) + f32(0) / f32(48)) / f32(2)
    v = f32(f32(10) - f32(0.1))
    r = f32(f32(1) + f32(f32(0.01) * f32(v - f32(0.1))))
    at = lambda name: o.state(name)[CY, CX]
    assert at("DminLT") == d and at("R") == r and at("V") == v
    for t in (2, 3):
        o.apply(_flat(3, 150), inf)
        a = f32(1) / f32(t)
        d = _ema(d, a, f32(0))
        v = f32(v - f32(0.1))
        assert float(r) >= (1.0 + 2.0 * float(d)) ** 2
        r = f32(r - f32(f32(0.01) / v))
        assert at("DminLT") == d and at("V") == v and at("R") == r, t
        assert at("T") == f32(4)


def test_one_matching_sample_scales_the_foreground_distance():
    # refreshModel(0.03) puts the last colour 110 into (size_t)(0.03*50) = 1 slot of every pixel
    # (others 100). A frame (140, 100, 100) fails the 100 samples (L1 40 > 36) and matches the 110
    # one (sums 30 + 10 + 10 = 50 <= 72): one good sample of the 2 required, foreground with
    # D = min(1, (50/765 + 0/48)/2 + (2 - 1)/2) (SUB:399).
    o = Oracle(SUBSENSE)
    o.initialize(_flat(3, 100))
    inf = float("inf")
    o.apply(_flat(3, 110), inf)
    d1 = (f32(30) / f32(765) + f32(0) / f32(48)) / f32(2)
    assert o.state("DminLT")[CY, CX] == d1
    o.refresh_model(0.03, False)
    assert ((o.state("samples_color")[:, CY, CX, 0] == 110).sum()) == 1
    o.apply(_flat(3, (140, 100, 100)), inf)
    assert o.state("last_raw")[CY, CX] == 255
    dfg = min(f32(1), f32(f32(f32(f32(50) / f32(765)) + f32(f32(0) / f32(48))) / f32(2)) + f32(f32(1) / f32(2)))
    assert o.state("DminLT")[CY, CX] == _ema(d1, f32(0.5), dfg)


def test_unstable_flag_from_the_long_term_raw_average():
    # After 100 static frames, P is lit for 20 frames (raw foreground alone: never in the final mask)
    # and then dark for 48. The raw averages (factors 0.01 and 0.04) are then
    # LT = (1 - 0.99^20) 0.99^48 ~ 0.112 and ST = (1 - 0.96^20) 0.96^48 ~ 0.078 with final averages 0,
    # so the next frame flags P unstable through the long-term term alone (SUB:365).
    o = Oracle(SUBSENSE)
    bg = _flat(3, 150)
    lit = bg.copy()
    lit[CY, CX] = 250
    o.initialize(bg)
    for _ in range(100):
        o.apply(bg)
    for _ in range(20):
        o.apply(lit)
    for _ in range(48):
        o.apply(bg)
    lt = (1 - 0.99 ** 20) * 0.99 ** 48
    st = (1 - 0.96 ** 20) * 0.96 ** 48
    assert lt > 0.11 and st < 0.08
    assert abs(o.state("RawLT")[CY, CX] - lt) < 1e-4 and abs(o.state("RawST")[CY, CX] - st) < 1e-4
    assert o.state("FinLT")[CY, CX] == 0 and o.state("R")[CY, CX] < 3
    o.apply(bg)
    assert o.state("unstable")[CY, CX] == 1
    assert o.state("unstable").sum() == 1


def test_learning_rate_caps_below_the_reset_threshold():
    # color_dist_threshold 4: reset threshold 4/2 = 2, caps from ratio >= 1 (SUB:584, :600-607).
    # A step of 8 at frame 101 gives ratio floor(8 (0.99^n - 0.96^n)): 0 for n <= 4, 1 for n = 5..10
    # (caps 2 >> 0 = 2 and 256 >> 0 = 256), 2 at n = 11: reset, caps 2 >> 1 -> 1 and 128.
    def diff(n):
        return 8.0 * (0.99 ** n - 0.96 ** n)

    for n in range(1, 13):
        assert abs(diff(n) - round(diff(n))) > 0.05
    o = Oracle(SUBSENSE, color_dist_threshold=4)
    a, b = _flat(3, 100, 240, 320), _flat(3, 108, 240, 320)
    o.initialize(a)
    for _ in range(100):
        o.apply(a)
    for n in range(1, 13):
        o.apply(b)
        s = o.scalars()
        ratio = int(diff(n))
        assert (s["lower_cap"], s["upper_cap"]) == ((2, 256) if ratio < 2 else (1, 128)), n
        assert s["refresh_seq"] == (1 if n < 11 else 2), n
