"""The engine's own integration test (rayna_engine/tests/simple_renders.rs) ported to the C++ Renderer API
(tests/cpp/simple_renders.cpp, built by build()), run on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_sphere_colours_through_the_cpp_api(gpu):
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "simple_renders")
    assert os.path.exists(exe), "build() compiles the C++ API tests"
    print("case cpp simple_renders", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "sphere_colours: ok" in r.stdout
