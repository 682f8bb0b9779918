# SPDX-License-Identifier: BSD-3-Clause
"""The tile server's threads under ThreadSanitizer, on the CPU: its host code built with
-fsanitize=thread (tests/tsan/Makefile) and driven over a CPU stand-in of the kernel library whose
streams are worker threads (tests/tsan/hip_standin.cpp). Submitters on four threads wait for 240
tiles in shuffled order, each tile checked cell by cell against the stand-in's height for its own
tile and bins (tests/tsan/server_driver.cpp); the run must end cleanly with no race reported."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

TSAN = Path(__file__).resolve().parent / "tsan"
CLANG = Path("/opt/rocm/llvm/bin/clang++")


@pytest.mark.skipif(not CLANG.exists() or shutil.which("make") is None,
                    reason="needs ROCm's clang++ (its ThreadSanitizer runtime) and make")
def test_server_threads_are_race_free(tmp_path):
    out = tmp_path / "tsan"
    build = subprocess.run(["make", "-C", str(TSAN), f"OUT={out}", "-j8"], capture_output=True,
                           text=True, timeout=600)
    assert build.returncode == 0, build.stderr[-4000:]
    env = dict(os.environ, TERRAINGEN_HIP_LIBRARY=str(out / "libtg_hip_standin.so"),
               TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    run = subprocess.run([str(out / "server_driver")], capture_output=True, text=True,
                         timeout=300, env=env)
    print(run.stdout)
    assert "ThreadSanitizer" not in run.stderr, run.stderr[-6000:]
    assert run.returncode == 0, run.stderr[-4000:]
    assert "bad 0" in run.stdout
