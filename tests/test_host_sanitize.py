"""The library's host code under sanitizers (tests/sanitize): the driver
calls the C-ABI through argument checks, the k-space generators at three grid
sizes in both precisions, context set-up, steps, the batch executor on
several host threads, checkpoints, tuning, particle sampling and device-built
starts.

* host paths under AddressSanitizer + UndefinedBehaviorSanitizer (no GPU);
* every path on the GPU with UndefinedBehaviorSanitizer on the host side (the
  ROCm runtime does not run under AddressSanitizer's allocator).
"""
import os
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "sanitize"


def _run(exe, mode, env):
    res = subprocess.run([str(exe), mode], cwd=HERE, capture_output=True, text=True, timeout=900,
                         env={**os.environ, **env})
    assert res.returncode == 0, res.stdout[-4000:] + res.stderr[-4000:]
    assert "0 failed checks" in res.stdout, res.stdout
    return res


def test_host_paths_under_asan_and_ubsan():
    subprocess.run(["make", "-C", str(HERE), "-s", "_build/host_driver"], check=True, timeout=1200)
    _run(HERE / "_build" / "host_driver", "cpu",
         {"ASAN_OPTIONS": "detect_leaks=0", "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"})


@pytest.mark.gpu
def test_gpu_paths_under_ubsan():
    exe = HERE / "_build" / "host_driver_ubsan"
    if not exe.exists():
        pytest.skip("tests/sanitize/_build/host_driver_ubsan not built (__graft_entry__.build() builds it)")
    print("host UBSan driver: cpu and gpu paths", flush=True)
    _run(exe, "gpu", {"UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"})
