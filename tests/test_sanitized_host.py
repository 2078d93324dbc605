"""The host code under the sanitizers: the scene packer (rayna_amd/csrc/scene_build.cpp, index-heavy: trees,
flattening, merging) and the oracle, built by g++ with AddressSanitizer + UndefinedBehaviorSanitizer and with
ThreadSanitizer (tests/sanitize/Makefile), driven through every kind of scene (tests/sanitize/drive.py) in a child
process with the sanitizer runtime preloaded; any report fails the run."""
import os
import subprocess
import sys

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


def _runtime(name):
    path = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.parametrize("kind,threads", [("asan", 1), ("tsan", 4)])
def test_host_code_is_clean_under_sanitizers(kind, threads):
    rt = _runtime("libasan.so" if kind == "asan" else "libtsan.so")
    if rt is None:
        pytest.skip("sanitizer runtime not installed")
    subprocess.run(["make", "-s", "-C", HERE, f"_build/libhost_{kind}.so"], check=True)
    env = dict(os.environ, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1")
    lib = os.path.join(HERE, "_build", f"libhost_{kind}.so")
    p = subprocess.run([sys.executable, os.path.join(HERE, "drive.py"), lib, str(threads)], env=env,
                       capture_output=True, text=True, timeout=600)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-4000:]
    assert "host ok" in p.stdout
    for marker in ("AddressSanitizer", "runtime error:", "ThreadSanitizer"):
        assert marker not in out, out[-4000:]
