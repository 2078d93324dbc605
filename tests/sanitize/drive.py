"""Drives the sanitized host library (tests/sanitize/Makefile) through the packer and the oracle: run in a child
process with the sanitizer runtime preloaded (tests/test_sanitized_host.py), it prints 'host ok' at the end."""
import ctypes as C
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402

oracle._LIB_PATH = sys.argv[1]
threads = int(sys.argv[2]) if len(sys.argv) > 2 else 1

from rayna_amd import _abi as A  # noqa: E402
from rayna_amd import presets  # noqa: E402

L = oracle.lib()
L.san_pack_scene.restype = C.c_int32
L.san_pack_scene.argtypes = [C.POINTER(A.SceneDesc), C.POINTER(C.c_uint64), C.POINTER(C.c_int32),
                             C.POINTER(C.c_int32), C.POINTER(C.c_int32)]

SCENES = [
    ("random_spheres", presets.random_spheres(width=16, height=9, samples=2)),
    ("mixed_primitives", presets.mixed_primitives(width=16, height=12, samples=2)),
    ("icosphere_mesh", presets.triangle_mesh(subdiv=3, width=16, height=9, samples=2)),
    ("sdf_volume", presets.sdf_volume(width=12, height=8, samples=1)),
    ("cornell", presets.cornell(width=12, height=12, samples=2)),
    ("rttnw", presets.rttnw(width=12, height=12, samples=2)),
    ("textured", presets.textured(width=12, height=8, samples=2)),
]

for name, (sc, cam, opts) in SCENES:
    d = sc.desc()
    h, need, pneed, nodes = C.c_uint64(0), C.c_int32(0), C.c_int32(0), C.c_int32(0)
    st = L.san_pack_scene(C.byref(d), C.byref(h), C.byref(need), C.byref(pneed), C.byref(nodes))
    assert st == 0, (name, st)
    tw = oracle.TwinScene(sc)
    assert tw.tree_hash() == h.value, name
    img32, _, _, _ = tw.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT, threads=threads)
    img64, _, _, _ = tw.render(cam, opts, precision=64, box_mode=oracle.BOX_REFERENCE, threads=threads)
    assert np.isfinite(img32).all() and np.isfinite(img64).all(), name
    tw.primary_hits(cam, opts, precision=32)
    tw.path_ranks(cam, opts, 1, 1, precision=32)
    _, counts = tw.render_checked(cam, opts, 32, threads=threads)
    assert not any(counts.values()), (name, counts)
    print(name, "packed", nodes.value, "nodes, stack", need.value, pneed.value, flush=True)
# the literal O(n^2) builder restatement on a small scene
sc, cam, opts = presets.random_spheres(width=8, height=8, samples=1)
assert oracle.TwinScene(sc, builder="literal").tree_hash() == oracle.TwinScene(sc).tree_hash()
print("host ok", flush=True)
