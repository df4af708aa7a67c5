"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer: tests/cpp/Makefile
links the library sources with a driver into one executable with the host code instrumented (device
code is not). The C++ mirror's runner drives initialize, apply, refreshModel, the getters,
checkpoint save / load (the host-side packing of the sample planes) and an error path; serve_run
drives a batch (two concurrent groups, graph replay, a per-stream override) and a server fed from
pinned host memory (ingest slots, busy retries, close, pooled re-open) through the C ABI. Every
output must equal the oracle's, and the runs must finish with no sanitizer report."""
import os
import subprocess

import numpy as np

import pytest

from litiv_amd import synth
from oracle_lib import LOBSTER, SUBSENSE, Oracle
from test_gpu_mirror import run_and_compare

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
RUNNER = os.path.join(HERE, "cpp", "mirror_run_asan")
SERVE = os.path.join(HERE, "cpp", "serve_run_asan")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:verify_asan_link_order=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.mark.parametrize("algo,c", [(SUBSENSE, 3), (LOBSTER, 1)])
def test_host_code_runs_clean_under_asan_and_ubsan(tmp_path, algo, c):
    r = run_and_compare(RUNNER, tmp_path, algo, c, True, env=ENV)
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]


def test_batch_and_server_host_code_runs_clean_under_asan_and_ubsan(tmp_path):
    assert os.path.exists(SERVE), "tests/cpp/serve_run_asan is not built: run __graft_entry__.build()"
    n = 8
    shapes = [(96, 72, 3), (64, 48, 3), (80, 60, 1)]
    frames = []
    for k, (w, h, c) in enumerate(shapes):
        seq = synth.Sequence(synth.preset(synth.QVGA, 11 + k, width=w, height=h, channels=c, n_objects=2, obj_min=8,
                                          obj_max=20))
        frames.append([seq.frame(t) for t in range(n)])
        (tmp_path / ("in%d.raw" % k)).write_bytes(b"".join(f.tobytes() for f in frames[k]))
    out = tmp_path / "out.raw"
    print("serve_run: batch and server over 3 streams, %d frames" % n, flush=True)
    r = subprocess.run([SERVE, str(n)] + [str(tmp_path / ("in%d.raw" % k)) for k in range(3)] + [str(out)],
                       capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    data = np.frombuffer(out.read_bytes(), np.uint8)

    def oracle_masks(k, sid, count):
        o = Oracle(SUBSENSE, stream_id=sid)
        o.initialize(frames[k][0])
        return [o.apply(frames[k][t]) for t in range(count)]

    expected = [oracle_masks(k, k, n) for k in range(3)]
    pos = 0
    for part in ("batch", "server"):
        for k, (w, h, c) in enumerate(shapes):
            for t in range(n):
                got = data[pos:pos + w * h].reshape(h, w)
                pos += w * h
                assert np.array_equal(got, expected[k][t]), "%s stream %d frame %d: %d px differ" % (
                    part, k, t, int((got != expected[k][t]).sum()))
    w, h, c = shapes[0]
    for t, exp in enumerate(oracle_masks(0, 7, 3)):
        got = data[pos:pos + w * h].reshape(h, w)
        pos += w * h
        assert np.array_equal(got, exp), "re-opened stream frame %d" % t
    assert pos == data.size

