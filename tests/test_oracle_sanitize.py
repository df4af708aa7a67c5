"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: its sources and a driver
(tests/cpp/oracle_sanitize_main.c) are built with -fsanitize=address,undefined and run over
SuBSENSE and LOBSTER sequences (1 and 3 channels, ROI, refresh, model reset, re-initialization,
state copies) and the primitives. Any invalid access or undefined operation fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SRCS = ["orc_rng.c", "orc_prim.c", "orc_bgs.c", "orc_eval.c"]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_runs_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_sanitize")
    cmd = ["gcc", "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-I", ORACLE,
           os.path.join(ROOT, "tests", "cpp", "oracle_sanitize_main.c")] + [os.path.join(ORACLE, s) for s in SRCS] + [
           "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.stdout.count("frames, examined") == 5
