"""The C-ABI library: it loads without a GPU and exports every entry point include/rayna_gpu.h declares; the
ctypes struct mirrors have the C layout; host-only preparation reports errors through status codes."""
import ctypes as C
import os
import re

import pytest

from rayna_amd import _abi as A
from rayna_amd import renderer

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rayna_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rayna_[a-z0-9_]+)\s*\(", text)))


def test_header_and_python_list_agree():
    assert declared_functions() == sorted(renderer.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(renderer.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_null_and_bad_arguments_fail_cleanly():
    L = renderer.lib()
    out = C.c_void_p()
    st = L.rayna_create(None, None, None, C.byref(out))
    assert st == A.RAYNA_INVALID_ARGUMENT and not out.value
    assert b"null" in L.rayna_last_error(None)
    d = A.SceneDesc()
    d.n_objects = -1
    assert L.rayna_scene_prepare(C.byref(d), None, None, None, 0, None, None) == A.RAYNA_INVALID_ARGUMENT


def test_struct_layouts_match_a_compiled_probe(tmp_path):
    """Compile a probe against the header with the system compiler and compare every offset ctypes uses."""
    import subprocess
    probe = tmp_path / "probe.c"
    fields = {"rayna_material": A.Material, "rayna_mesh": A.Mesh, "rayna_object": A.Object, "rayna_sdf": A.Sdf,
              "rayna_camera": A.CameraDesc, "rayna_render_opts": A.RenderOptsDesc,
              "rayna_render_stats": A.RenderStats, "rayna_scene_desc": A.SceneDesc, "rayna_tuning": A.Tuning,
              "rayna_request": A.Request, "rayna_batch_stats": A.BatchStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in fields.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    probe.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", str(probe), "-o", str(exe)], check=True)
    got = dict((" ".join(l.split()[:2]), int(l.split()[2])) for l in subprocess.check_output([str(exe)]).decode().splitlines())
    for cname, py in fields.items():
        assert got[f"{cname} size"] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[f"{cname} {f}"] == getattr(py, f).offset, (cname, f)


def test_status_names_cover_the_header():
    text = open(HEADER).read()
    codes = dict((int(v), n) for n, v in re.findall(r"RAYNA_([A-Z_]+) = (\d+)", text.split("} rayna_status;")[0]))
    assert codes == A.STATUS_NAMES
    assert A.RAYNA_RENDERER_LOST == 5


def test_fault_injection_flag_is_host_state():
    """rayna_debug_fail_launches only sets a countdown (no HIP call, so it runs without a GPU): it returns the
    previous setting, and any negative value means off."""
    L = renderer.lib()
    assert L.rayna_debug_fail_launches(-1) == -1
    assert renderer.debug_fail_launches(3) == -1
    assert renderer.debug_fail_launches(-7) == 3
    assert renderer.debug_fail_launches(0) == -1
    assert renderer.debug_fail_launches(-1) == 0
    assert renderer.debug_fail_launches(-1) == -1
