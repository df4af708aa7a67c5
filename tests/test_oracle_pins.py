"""Pins of the CPU oracle (CPU only): each primitive is checked against something outside the oracle's
own formula -- a published known-answer vector, the reference's golden test data, a closed form, a
library routine (scipy) that computes the same operation, or brute-force enumeration -- chosen so a
plausible slip (wrong bit order, transposed offset, dropped wrap-around, wrong border) fails here.
"""
import ctypes as C
import os

import numpy as np
import pytest
from scipy import ndimage

import oracle_lib
from oracle_lib import lib

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _u8p(a):
    # the oracle's ctypes signatures take contiguous numpy arrays
    return np.ascontiguousarray(a)


# ---- counter-based RNG ---------------------------------------------------------------------------
# Random123 (Salmon et al., SC'11) known-answer vectors for philox4x32 with 10 rounds
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expected", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expected):
    assert tuple(oracle_lib.philox(list(ctr), list(key))) == expected


def test_draw_is_philox_word_of_the_counter():
    # draw d of (seed, stream, frame, phase, event, pixel) = word d%4 of the block at
    # ctr = {pixel, frame, phase<<24 | event<<8 | d/4, stream}, key = seed, shifted to 31 bits
    seed, stream, frame, phase, event, pixel = 0x0123456789ABCDEF, 9, 77, 1, 3, 4242
    for d in range(8):
        ctr = [pixel, frame, (phase << 24) | ((event & 0xFFFF) << 8) | (d // 4), stream]
        word = oracle_lib.philox(ctr, [seed & 0xFFFFFFFF, seed >> 32])[d % 4]
        assert lib().orc_draw(seed, stream, frame, phase, event, pixel, d) == word >> 1


def test_draws_are_uniform_enough():
    # 31-bit draws: bit balance over 20k pixels (a stuck or shifted-out bit fails)
    v = np.array([lib().orc_draw(0xB65E, 0, 5, 1, 0, p, 0) for p in range(20000)], np.uint64)
    assert v.max() < 2 ** 31
    bits = np.array([((v >> b) & 1).mean() for b in range(31)])
    assert np.all(np.abs(bits - 0.5) < 0.02)


# ---- LBSP --------------------------------------------------------------------------------------------
def _lbsp_oracle(img, t):
    h, w = img.shape[:2]
    c = 1 if img.ndim == 2 else img.shape[2]
    out = np.zeros((h, w, c), np.uint16)
    lib().orc_lbsp_compute_abs(_u8p(img), w, h, c, t, out)
    return out.reshape(img.shape)


def test_lbsp_reference_golden_vector():
    # modules/features2d/test/lbsp.cpp:22-46: LBSP with absolute threshold 20 on a 65x65 RGB crop,
    # expected descriptors from the reference's test_lbsp.bin (valid zone, 2-pixel border excluded)
    inp = np.load(os.path.join(GOLDEN, "lbsp_kat_input.npy"))
    exp = np.load(os.path.join(GOLDEN, "lbsp_kat_expected.npy"))
    out = _lbsp_oracle(inp, 20)
    assert np.array_equal(out[2:63, 2:63], exp[2:63, 2:63])


def test_lbsp_special_cases():
    rng = np.random.default_rng(3)
    flat = np.full((9, 9), 77, np.uint8)
    assert not _lbsp_oracle(flat, 0)[2:7, 2:7].any()      # |n - c| > t never holds on a flat patch
    noisy = rng.integers(0, 256, (12, 12), dtype=np.uint8)
    assert not _lbsp_oracle(noisy, 255)[2:10, 2:10].any()  # no difference exceeds 255
    # a single bright centre: every neighbour differs by 200 > t -> all 16 bits
    img = np.zeros((5, 5), np.uint8)
    img[2, 2] = 200
    assert _lbsp_oracle(img, 199)[2, 2] == 0xFFFF
    assert _lbsp_oracle(img, 200)[2, 2] == 0


def test_lut_closed_form():
    lut3 = np.zeros(256, np.uint8)
    lut1 = np.zeros(256, np.uint8)
    lib().orc_build_lut(3, C.c_float(0.333), 0, _u8p(lut3))
    lib().orc_build_lut(1, C.c_float(0.333), 0, _u8p(lut1))
    t = np.arange(256, dtype=np.float32)
    # saturate_cast<uchar>(float) rounds half to even; t*0.333f is exact in float32 arithmetic here
    assert np.array_equal(lut3, np.clip(np.rint(t * np.float32(0.333)), 0, 255).astype(np.uint8))
    assert np.array_equal(lut1, np.clip(np.rint((t * np.float32(0.333)) / np.float32(3)), 0, 255).astype(np.uint8))
    assert (lut3[0], lut3[255], lut1[255]) == (0, 85, 28)


@pytest.mark.parametrize("v,expected", [(0.5, 0), (1.5, 2), (2.5, 2), (-3.0, 0), (254.6, 255), (300.0, 255)])
def test_saturate_round_half_even(v, expected):
    assert lib().orc_sat_u8_f(C.c_float(v)) == expected


# ---- sampling tables ---------------------------------------------------------------------------------
def _pos(fn, r, x, y, border=0, w=64, h=64):
    ox, oy = C.c_int(), C.c_int()
    getattr(lib(), fn)(r, x, y, border, w, h, C.byref(ox), C.byref(oy))
    return ox.value, oy.value


# opencv.hpp:913-923 (getSamplePosition_7x7_std2), the reference's own table: floor(fspecial('gaussian',
# 7, 2) * 512) outside the central 3x3, whose weights are raised so the total is 512
SAMPLE_7X7_REFERENCE = np.array([[2, 4, 6, 7, 6, 4, 2], [4, 8, 12, 14, 12, 8, 4], [6, 12, 21, 25, 21, 12, 6],
                                 [7, 14, 25, 28, 25, 14, 7], [6, 12, 21, 25, 21, 12, 6], [4, 8, 12, 14, 12, 8, 4],
                                 [2, 4, 6, 7, 6, 4, 2]])


def test_sample_7x7_histogram_is_the_reference_table():
    g = np.exp(-(np.arange(-3, 4)[:, None] ** 2 + np.arange(-3, 4)[None, :] ** 2) / 8.0)
    floored = np.floor(g / g.sum() * 512).astype(int)
    outer = np.ones((7, 7), bool)
    outer[2:5, 2:5] = False
    assert np.array_equal(SAMPLE_7X7_REFERENCE[outer], floored[outer]) and SAMPLE_7X7_REFERENCE.sum() == 512
    # enumerating all 512 residues at an interior pixel hits each offset exactly table[offset] times
    hist = np.zeros((7, 7), int)
    for r in range(512):
        x, y = _pos("orc_sample_pos_7x7", r, 20, 30)
        hist[y - 30 + 3, x - 20 + 3] += 1
    assert np.array_equal(hist, SAMPLE_7X7_REFERENCE)
    assert _pos("orc_sample_pos_7x7", 512 + 5, 20, 30) == _pos("orc_sample_pos_7x7", 5, 20, 30)


@pytest.mark.parametrize("fn,n,radius", [("orc_neighbor_pos_3x3", 8, 1), ("orc_neighbor_pos_5x5", 24, 2)])
def test_neighbour_tables_cover_the_ring(fn, n, radius):
    seen = {_pos(fn, r, 20, 30) for r in range(n)}
    want = {(20 + dx, 30 + dy) for dx in range(-radius, radius + 1) for dy in range(-radius, radius + 1)
            if (dx, dy) != (0, 0)}
    assert seen == want
    # clamped into the valid zone (border 2) at the image corner
    for r in range(n):
        x, y = _pos(fn, r, 2, 2, border=2, w=10, h=8)
        assert 2 <= x <= 7 and 2 <= y <= 5


def test_pow_square_compare_is_an_exact_product():
    """The R feedback compare `R < std::pow(1.0f + 2*min(Dlt, Dst), 2)` (BackgroundSubtractorSuBSENSE.cpp:467)
    promotes the float base to double and calls libm pow; the oracle and the kernel compare against
    (double)base * (double)base instead. Every float32 base the compare can see (min of two maps kept in
    [0, 1], so base in [1, 3]) is enumerated and libm's pow(base, 2) -- math.pow calls it directly -- must
    equal the double product bit for bit; the two compares then agree for every R."""
    import itertools
    import math
    lo = int(np.float32(1.0).view(np.uint32))
    hi = int(np.float32(3.0).view(np.uint32))
    step = 1 << 20
    for start in range(lo, hi + 1, step):
        bits = np.arange(start, min(start + step, hi + 1), dtype=np.uint32)
        base = bits.view(np.float32).astype(np.float64)
        prod = base * base
        ref = np.fromiter(map(math.pow, base.tolist(), itertools.repeat(2.0)), np.float64, count=base.size)
        bad = np.flatnonzero(ref != prod)
        assert bad.size == 0, "pow(%r, 2) = %r, product %r" % (base[bad[0]], ref[bad[0]], prod[bad[0]])


def test_l1dist_uchar_wraps():
    a = np.array([200, 200, 200], np.uint8)
    b = np.zeros(3, np.uint8)
    assert lib().orc_l1dist3_u8(_u8p(a), _u8p(b)) == (600 % 256)
    a2 = np.array([10, 250, 3], np.uint8)
    b2 = np.array([20, 240, 9], np.uint8)
    assert lib().orc_l1dist3_u8(_u8p(a2), _u8p(b2)) == 26


# ---- whole-image closed forms vs scipy -------------------------------------------------------------
def _mask(rng, h, w, p=0.35, ring=2):
    m = (rng.random((h, w)) < p).astype(np.uint8) * 255
    if ring:
        m[:ring] = 0
        m[-ring:] = 0
        m[:, :ring] = 0
        m[:, -ring:] = 0
    return m


@pytest.mark.parametrize("radius,is_max", [(1, 1), (1, 0), (3, 1), (3, 0)])
def test_morph_matches_scipy(radius, is_max):
    rng = np.random.default_rng(radius + 10 * is_max)
    m = _mask(rng, 37, 53, ring=0)
    out = np.empty_like(m)
    lib().orc_morph(_u8p(m), _u8p(out), 53, 37, radius, is_max)
    f = ndimage.maximum_filter if is_max else ndimage.minimum_filter
    assert np.array_equal(out, f(m, size=2 * radius + 1, mode="nearest"))


@pytest.mark.parametrize("k", [3, 9, 13])
def test_binary_median_matches_scipy(k):
    rng = np.random.default_rng(k)
    for p in (0.2, 0.5, 0.8):
        m = _mask(rng, 41, 60, p=p)
        out = np.empty_like(m)
        lib().orc_median_binary(_u8p(m), _u8p(out), 60, 41, k)
        assert np.array_equal(out, ndimage.median_filter(m, size=k, mode="nearest")), p


def test_hole_fill_matches_scipy():
    rng = np.random.default_rng(5)
    for p in (0.3, 0.55, 0.7):
        m = _mask(rng, 50, 70, p=p)
        holes = np.empty_like(m)
        lib().orc_flood_holes(_u8p(m), _u8p(holes), 70, 50)
        filled = ndimage.binary_fill_holes(m > 0)
        assert np.array_equal(holes > 0, filled & (m == 0)), p


def test_resize_area_integer_factor_is_rounded_block_mean():
    rng = np.random.default_rng(8)
    img = rng.integers(0, 256, (48, 64, 3), dtype=np.uint8)
    out = np.empty((6, 8, 3), np.uint8)
    lib().orc_resize_area(_u8p(img), 64, 48, 3, _u8p(out), 8, 6)
    # sums / 64 are exact in float32, so the result is the half-to-even rounded block mean
    mean = img.reshape(6, 8, 8, 8, 3).astype(np.int64).sum(axis=(1, 3)) / 64.0
    assert np.array_equal(out, np.rint(mean).astype(np.uint8))


# ---- BinClassif (modules/datasets/src/metrics.cpp:21-50) ------------------------------------------
# One pixel at a time, the category each (result, ground truth, ROI) triple takes by the reference's
# tests: GT 85 / 170 or ROI 0 -> don't care; else result 255 -> TP (GT 255) / FP (any other GT);
# result != 255 -> FN (GT 255) / TN; result 255 on GT 50 -> also SE.
def _expected_category(res, gt, roi):
    if gt in (85, 170) or roi == 0:
        return {"dc"}
    cat = {("tp" if gt == 255 else "fp") if res == 255 else ("fn" if gt == 255 else "tn")}
    if gt == 50 and res == 255:
        cat.add("se")
    return cat


def test_binclassif_known_answers():
    names = ("tp", "tn", "fp", "fn", "se", "dc")
    for res in (0, 7, 255):
        for gt in (0, 3, 50, 85, 170, 255):
            for roi in (None, 0, 1, 255):
                got = oracle_lib.binclassif(np.full((1, 1), res, np.uint8), np.full((1, 1), gt, np.uint8),
                                            None if roi is None else np.full((1, 1), roi, np.uint8))
                assert {n for n, v in zip(names, got) if v} == _expected_category(res, gt, roi), (res, gt, roi)
                assert sum(got) - got[4] == 1
    # no ground truth: every pixel is a don't-care (metrics.cpp:26-29)
    assert oracle_lib.binclassif(np.zeros((3, 5), np.uint8)) == (0, 0, 0, 0, 0, 15)
