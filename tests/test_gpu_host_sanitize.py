"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer: tests/cpp/Makefile
links the library sources and the C++ mirror's runner into one executable with the host code
instrumented (device code is not), and the runner drives initialize, apply, refreshModel, the
getters, checkpoint save / load (the host-side packing of the sample planes) and an error path on
the GPU. Every output must equal the oracle's, and the run must finish with no sanitizer report."""
import os

import pytest

from oracle_lib import LOBSTER, SUBSENSE
from test_gpu_mirror import run_and_compare

pytestmark = pytest.mark.gpu

RUNNER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "mirror_run_asan")


@pytest.mark.parametrize("algo,c", [(SUBSENSE, 3), (LOBSTER, 1)])
def test_host_code_runs_clean_under_asan_and_ubsan(tmp_path, algo, c):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = run_and_compare(RUNNER, tmp_path, algo, c, True, env=env)
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
