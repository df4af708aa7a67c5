"""The header-only C++ mirror of IIBackgroundSubtractor (include/litiv_bgs.hpp), executed: a small
C++ program (tests/cpp/mirror_run.cpp, built by __graft_entry__.build()) runs a sequence through
the mirror's constructor, initialize, apply, refreshModel and getters, and an apply on a bad input
that must throw; its outputs must equal the oracle's for the same calls. With a checkpoint
directory, the runner saves the model part-way through (saveCheckpoint) and finishes the sequence
on a second subtractor that loaded it (loadCheckpoint): the outputs must still equal the oracle's,
which never stops."""
import os
import subprocess

import numpy as np
import pytest

from litiv_amd import synth
from oracle_lib import LOBSTER, SUBSENSE, Oracle

pytestmark = pytest.mark.gpu

RUNNER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "mirror_run")


@pytest.mark.parametrize("algo,c,resume", [(SUBSENSE, 3, False), (SUBSENSE, 1, False), (LOBSTER, 3, False),
                                           (SUBSENSE, 3, True), (LOBSTER, 1, True)])
def test_cpp_mirror_runs_a_sequence(tmp_path, algo, c, resume):
    run_and_compare(RUNNER, tmp_path, algo, c, resume)


def run_and_compare(runner, tmp_path, algo, c, resume, env=None):
    """Runs `runner` (mirror_run or its sanitized build) on a small sequence and compares every output
    with the oracle's."""
    assert os.path.exists(runner), "%s is not built: run __graft_entry__.build()" % runner
    w, h, n, refresh_at = 128, 96, 16, 7
    seq = synth.Sequence(synth.preset(synth.QVGA, 3, width=w, height=h, channels=c, n_objects=3, obj_min=8,
                                      obj_max=24))
    frames = [seq.frame(t) for t in range(n)]
    inp, out = tmp_path / "in.raw", tmp_path / "out.raw"
    inp.write_bytes(b"".join(f.tobytes() for f in frames))
    print("mirror_run algo %d %dx%dx%d resume %d" % (algo, w, h, c, resume), flush=True)
    cmd = [runner, str(algo), str(w), str(h), str(c), str(n), str(inp), str(out), str(refresh_at)]
    if resume:
        os.makedirs(tmp_path / "ck")
        cmd.append(str(tmp_path / "ck"))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
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
    return r
