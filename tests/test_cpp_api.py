"""The C++ host API (include/rayna/engine.hpp): it compiles against the C ABI, and its scene flattening gives
the same engine object graph (tree hash, depth-first order) as the Python mirror for a scene with every mesh,
material and object kind.  Host only: rayna_scene_prepare builds the trees without a GPU."""
import os
import subprocess

import numpy as np

from rayna_amd import prepare_scene
from rayna_amd.scene import (AxisBoxMesh, DielectricMaterial, IsotropicMaterial, LambertianMaterial, LightMaterial,
                             MetalMaterial, ParallelogramMesh, RaymarchedIsosurfaceMesh, Scene, SdfMandelbulb,
                             SdfSminSpheres, SimpleObject, SimpleSkybox, SphereMesh, TriangleListMesh, TriangleMesh,
                             VolumetricObject)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r'''
#include <cstdio>
#include "rayna/engine.hpp"
using namespace rayna;
int main() {
    std::vector<Triangle> tris;
    for (int i = 0; i < 40; ++i) {
        const double x = 0.1 * i, y = 0.05 * (i % 7), z = -1.0 - 0.02 * i;
        tris.push_back(Triangle{{{x, y, z}, {x + 0.3, y, z}, {x, y + 0.3, z + 0.1}},
                                {{0, 0, 1}, {0, 0, 1}, {0, 0, 1}}});
    }
    StandardScene s;
    s.objects.push_back(SimpleObject::new_uncorrected(SphereMesh{{0, -1000, 0}, 1000}, LambertianMaterial{{0.5f, 0.5f, 0.5f}}));
    s.objects.push_back(SimpleObject::new_uncorrected(SphereMesh{{1, 1, 0}, 0.5}, MetalMaterial{{0.8f, 0.7f, 0.6f}, 0.1}));
    s.objects.push_back(SimpleObject::new_uncorrected(MeshList{tris}, DielectricMaterial{{1, 1, 1}, 1.5, 0.2}));
    s.objects.push_back(SimpleObject::new_uncorrected(Triangle{{{-2, 0, -2}, {2, 0, -2}, {0, 2, -3}}, {{0, 0, 1}, {0, 0, 1}, {0, 0, 1}}},
                                                      LambertianMaterial{{0.2f, 0.3f, 0.4f}}));
    s.objects.push_back(SimpleObject::new_uncorrected(ParallelogramMesh{{-1, 3, -1}, {2, 0, 0}, {0, 0, 2}}, LightMaterial{{4, 4, 4}}));
    s.objects.push_back(SimpleObject::new_uncorrected(AxisBoxMesh{{1.5, 0, 0.5}, {2.5, 1, 1.5}}, LambertianMaterial{{0.7f, 0.3f, 0.2f}}));
    s.objects.push_back(VolumetricObject{SphereMesh{{-2, 1, 1.5}, 0.8}, IsotropicMaterial{{0.9f, 0.7f, 0.7f}, 0.1}, 1.5});
    s.objects.push_back(SimpleObject::new_uncorrected(
        RaymarchedIsosurfaceMesh{SdfSminSpheres{{{1, 0.5, -2}, {1.5, 0.8, -2.2}}, {0.4, 0.3}, 0.1}, 64, 1e-4},
        LambertianMaterial{{0.3f, 0.3f, 0.8f}}));
    s.objects.push_back(SimpleObject::new_uncorrected(RaymarchedIsosurfaceMesh{SdfMandelbulb{{0, 0, 0}, 8, 10, 2}, 256, 1e-4},
                                                      LambertianMaterial{{0.8f, 0.5f, 0.3f}}));
    s.skybox = SkyboxInstance::SimpleSkybox;
    detail::SceneArrays a(s);
    uint64_t hash = 0;
    int32_t n = 0, need = 0;
    std::vector<int32_t> obj(4096), loc(4096);
    if (rayna_scene_prepare(&a.desc, &hash, obj.data(), loc.data(), 4096, &n, &need) != RAYNA_OK) return 2;
    std::printf("%llu %d", (unsigned long long)hash, n);
    for (int i = 0; i < n; ++i) std::printf(" %d:%d", obj[i], loc[i]);
    std::printf("\n");
    return 0;
}
'''


def _python_scene():
    tv = []
    for i in range(40):
        x, y, z = 0.1 * i, 0.05 * (i % 7), -1.0 - 0.02 * i
        tv.append([[x, y, z], [x + 0.3, y, z], [x, y + 0.3, z + 0.1]])
    tv = np.array(tv)
    tn = np.tile(np.array([0.0, 0.0, 1.0]), (40, 3, 1))
    objs = [
        SimpleObject(SphereMesh((0, -1000, 0), 1000), LambertianMaterial((0.5, 0.5, 0.5))),
        SimpleObject(SphereMesh((1, 1, 0), 0.5), MetalMaterial((0.8, 0.7, 0.6), 0.1)),
        SimpleObject(TriangleListMesh(tv, tn), DielectricMaterial((1, 1, 1), 1.5, 0.2)),
        SimpleObject(TriangleMesh(((-2, 0, -2), (2, 0, -2), (0, 2, -3)), ((0, 0, 1), (0, 0, 1), (0, 0, 1))),
                     LambertianMaterial((0.2, 0.3, 0.4))),
        SimpleObject(ParallelogramMesh((-1, 3, -1), (2, 0, 0), (0, 0, 2)), LightMaterial((4, 4, 4))),
        SimpleObject(AxisBoxMesh((1.5, 0, 0.5), (2.5, 1, 1.5)), LambertianMaterial((0.7, 0.3, 0.2))),
        VolumetricObject(SphereMesh((-2, 1, 1.5), 0.8), IsotropicMaterial((0.9, 0.7, 0.7), 0.1), 1.5),
        SimpleObject(RaymarchedIsosurfaceMesh(SdfSminSpheres(np.array([[1, 0.5, -2], [1.5, 0.8, -2.2]]),
                                                             np.array([0.4, 0.3]), 0.1), 64, 1e-4),
                     LambertianMaterial((0.3, 0.3, 0.8))),
        SimpleObject(RaymarchedIsosurfaceMesh(SdfMandelbulb((0, 0, 0), 8, 10, 2.0), 256, 1e-4),
                     LambertianMaterial((0.8, 0.5, 0.3))),
    ]
    return Scene(objs, SimpleSkybox)


def test_cpp_scene_flattening_matches_python(tmp_path):
    src = tmp_path / "probe.cpp"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    lib_dir = os.path.join(ROOT, "rayna_amd")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", f"-I{os.path.join(ROOT, 'include')}", str(src),
                    f"-L{lib_dir}", "-lrayna_amd", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    h, n = int(out[0]), int(out[1])
    ph, obj, loc, _ = prepare_scene(_python_scene())
    assert h == ph and n == len(obj)
    assert [f"{o}:{l}" for o, l in zip(obj, loc)] == out[2:]


def test_cpp_tests_compile_cleanly(tmp_path):
    for name in os.listdir(os.path.join(ROOT, "tests", "cpp")):
        if name.endswith(".cpp"):
            subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                            f"-I{os.path.join(ROOT, 'include')}", os.path.join(ROOT, "tests", "cpp", name)], check=True)
