"""C ABI checks that need no GPU: every function declared in include/*.h is exported by the built
libraries, the host-only entry points behave as the reference's (default parameters, validateROI),
the synthetic generator is deterministic and matches its configuration, and bench.py refuses
multi-GPU runs. No call here launches device work."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"litiv_bgs.h": "liblitiv_bgs.so", "litiv_synth.h": "liblitiv_synth.so"}


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lv(?:bgs|synth)_\w+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_every_declared_function_is_exported(header):
    lib = C.CDLL(os.path.join(ROOT, "litiv_amd", HEADERS[header]))
    names = _declared(header)
    assert len(names) >= 4
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def _bgs():
    from litiv_amd import bgs
    return bgs


def test_default_parameters_are_the_reference_defaults():
    bgs = _bgs()
    s = bgs.default_params(bgs.SUBSENSE)  # BackgroundSubtractorSuBSENSE.hpp:22-31
    assert (s.desc_dist_threshold, s.color_dist_threshold, s.bg_samples, s.required_bg_samples,
            s.samples_for_mv_avgs, s.lbsp_threshold_offset, s.median_blur_kernel) == (3, 30, 50, 2, 100, 0, 9)
    assert abs(s.rel_lbsp_threshold - 0.333) < 1e-7
    lo = bgs.default_params(bgs.LOBSTER)  # BackgroundSubtractorLOBSTER.hpp:22-31
    assert (lo.desc_dist_threshold, lo.color_dist_threshold, lo.bg_samples, lo.required_bg_samples) == (4, 30, 35, 2)
    with pytest.raises(ValueError):
        bgs.default_params(7)


def test_validate_roi_clears_the_lbsp_border():
    bgs = _bgs()
    roi = np.full((12, 17), 255, np.uint8)
    out = bgs.BackgroundSubtractor.validateROI(roi)
    assert out[2:10, 2:15].all() and not out[:2].any() and not out[10:].any()
    assert not out[:, :2].any() and not out[:, 15:].any()
    assert roi.all()  # the input is not modified (returns a validated copy)


def test_synthetic_generator_is_deterministic_and_configured():
    from litiv_amd import synth
    cfg = synth.preset(synth.QVGA, 0)
    assert (cfg.width, cfg.height, cfg.channels, cfg.frames, cfg.n_objects) == (320, 240, 3, 300, 3)
    a = synth.Sequence(cfg)
    b = synth.Sequence(synth.preset(synth.QVGA, 0))
    for t in (0, 7, 150):
        assert np.array_equal(a.frame(t), b.frame(t))
    assert a.frame(0).shape == (240, 320, 3)
    assert not np.array_equal(a.frame(0), a.frame(5))  # objects move
    other = synth.Sequence(synth.preset(synth.QVGA, 1))
    assert not np.array_equal(a.frame(0), other.frame(0))  # per-stream seed
    # background without noise or objects stays in the value-noise range [40, 215]
    quiet = synth.Sequence(synth.preset(synth.QVGA, 0, noise_sigma=0, n_objects=0))
    f = quiet.frame(0).astype(int)
    assert f.min() >= 40 - 10 and f.max() <= 215 + 10  # +-10 illumination ramp


def test_batch64_mix():
    from litiv_amd import synth
    shapes = [(c.width, c.height, c.channels) for c in (synth.preset(synth.BATCH64, s) for s in range(64))]
    for g, want in enumerate([(320, 240, 3), (640, 480, 3), (640, 512, 1), (1920, 1080, 3)]):
        assert shapes[16 * g:16 * g + 16] == [want] * 16
    with pytest.raises(ValueError):
        synth.preset(synth.BATCH64, 64)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_cpp_mirror_compiles_and_links(tmp_path):
    # include/litiv_bgs.hpp (header-only IIBackgroundSubtractor mirror) against the built library
    src = tmp_path / "use.cpp"
    src.write_text('#include "litiv_bgs.hpp"\n'
                   "int main(int argc, char**) {\n"
                   "  if (argc > 5) {\n"
                   "    litiv_amd::BackgroundSubtractorSuBSENSE s;\n"
                   "    litiv_amd::Image<uint8_t> m; litiv_amd::ImageView v;\n"
                   "    s.initialize(v); s.apply(v, m); s.refreshModel(0.1f);\n"
                   "    litiv_amd::Image<uint16_t> d; s.getBackgroundDescriptorsImage(d);\n"
                   "    litiv_amd::BackgroundSubtractorLOBSTER l; l.getBackgroundImage(m);\n"
                   "  }\n"
                   "  return 0;\n"
                   "}\n")
    exe = tmp_path / "use"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                        "-L", os.path.join(ROOT, "litiv_amd"), "-llitiv_bgs", "-o", str(exe)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]


def test_python_mirror_of_the_noalloc_pixel_limit():
    text = open(os.path.join(ROOT, "include", "litiv_bgs.h")).read()
    m = re.search(r"#define LVBGS_BINCLASSIF_NOALLOC_MAX_PIXELS (\d+)ull", text)
    assert m and int(m.group(1)) == _bgs().BINCLASSIF_NOALLOC_MAX_PIXELS
    m = re.search(r"#define LVBGS_ERR_TOO_LARGE \((-\d+)\)", text)
    assert m and int(m.group(1)) == _bgs().ERR_TOO_LARGE

