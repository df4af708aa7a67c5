"""The header-only C++ mirror of IIBackgroundSubtractor (include/litiv_bgs.hpp), executed: a small
C++ program (tests/cpp/mirror_run.cpp, built by __graft_entry__.build()) runs a sequence through
the mirror's constructor, initialize, apply, refreshModel and getters, and an apply on a bad input
that must throw; its outputs must equal the oracle's for the same calls."""
import os
import subprocess

import numpy as np
import pytest

from litiv_amd import synth
from oracle_lib import LOBSTER, SUBSENSE, Oracle

pytestmark = pytest.mark.gpu

RUNNER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "mirror_run")


@pytest.mark.parametrize("algo,c", [(SUBSENSE, 3), (SUBSENSE, 1), (LOBSTER, 3)])
def test_cpp_mirror_runs_a_sequence(tmp_path, algo, c):
    assert os.path.exists(RUNNER), "tests/cpp/mirror_run is not built: run __graft_entry__.build()"
    w, h, n, refresh_at = 128, 96, 16, 7
    seq = synth.Sequence(synth.preset(synth.QVGA, 3, width=w, height=h, channels=c, n_objects=3, obj_min=8,
                                      obj_max=24))
    frames = [seq.frame(t) for t in range(n)]
    inp, out = tmp_path / "in.raw", tmp_path / "out.raw"
    inp.write_bytes(b"".join(f.tobytes() for f in frames))
    print("mirror_run algo %d %dx%dx%d" % (algo, w, h, c), flush=True)
    r = subprocess.run([RUNNER, str(algo), str(w), str(h), str(c), str(n), str(inp), str(out), str(refresh_at)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    data = np.frombuffer(out.read_bytes(), np.uint8)
    o = Oracle(algo)
    o.initialize(frames[0])
    pos = 0
    for t in range(n):
        mask = data[pos:pos + w * h].reshape(h, w)
        pos += w * h
        exp = o.apply(frames[t])
        assert np.array_equal(mask, exp), "frame %d: %d px differ" % (t, int((mask != exp).sum()))
        if t == refresh_at:
            o.refresh_model(0.1 if algo == SUBSENSE else 0.5, False)
    bg = data[pos:pos + w * h * c]
    pos += w * h * c
    desc = data[pos:pos + w * h * c * 2].view(np.uint16)
    pos += w * h * c * 2
    assert np.array_equal(bg, o.background_image().reshape(-1))
    assert np.array_equal(desc, o.background_descriptors().reshape(-1))
    code = int(data[pos:pos + 4].view(np.int32)[0])
    assert code == -1  # LVBGS_ERR_ARG, thrown as litiv_amd::Exception
