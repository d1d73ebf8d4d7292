"""The oracle under the host sanitizers: oracle/sanitize_check.c drives every
oracle entry point on a small chain, built here with AddressSanitizer and
UndefinedBehaviorSanitizer (no recovery), so that an out-of-bounds access,
a leak or undefined behaviour in the checker fails this test."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(
    __file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_runs_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_check")
    cmd = ["gcc", "-O1", "-g", "-std=c11", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
           "-Wall", "-Wextra", "-Wno-unused-parameter",
           os.path.join(ORACLE, "sanitize_check.c"),
           os.path.join(ORACLE, "ckks_oracle.c"), "-o", exe, "-lm"]
    build = subprocess.run(cmd, capture_output=True, text=True)
    if build.returncode != 0 and "asan" in build.stderr.lower():
        pytest.skip("sanitizer runtime not available: " + build.stderr[-200:])
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe], capture_output=True, text=True, env=env,
                         timeout=300)
    assert run.returncode == 0, run.stderr[-4000:]
    assert run.stdout.strip() == "ok"
