"""The C++ host API (include/rayna/engine.hpp): it compiles against the C ABI, and its scene flattening gives
the same engine object graph (tree hash, depth-first order) and the same object transforms as the Python mirror
for a scene with every mesh, material and object kind, corrected and uncorrected transforms, a BVH mesh and an
instanced triangle list.  Host only: rayna_scene_prepare builds the trees without a GPU."""
import os
import subprocess

import numpy as np

from rayna_amd import prepare_scene
from rayna_amd.scene import (AxisBoxMesh, BvhMesh, DielectricMaterial, IsotropicMaterial, LambertianMaterial, LightMaterial,
                             MetalMaterial, ParallelogramMesh, RaymarchedIsosurfaceMesh, Scene, SdfMandelbulb,
                             SdfSminSpheres, SimpleObject, SimpleSkybox, SphereMesh, Transform3, TriangleListMesh,
                             TriangleMesh, VolumetricObject, object_matrix)

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
    s.objects.push_back(VolumetricObject::new_uncorrected(SphereMesh{{-2, 1, 1.5}, 0.8},
                                                          IsotropicMaterial{{0.9f, 0.7f, 0.7f}, 0.1}, 1.5));
    s.objects.push_back(SimpleObject::new_uncorrected(
        RaymarchedIsosurfaceMesh{SdfSminSpheres{{{1, 0.5, -2}, {1.5, 0.8, -2.2}}, {0.4, 0.3}, 0.1}, 64, 1e-4},
        LambertianMaterial{{0.3f, 0.3f, 0.8f}}));
    s.objects.push_back(SimpleObject::new_uncorrected(RaymarchedIsosurfaceMesh{SdfMandelbulb{{0, 0, 0}, 8, 10, 2}, 256, 1e-4},
                                                      LambertianMaterial{{0.8f, 0.5f, 0.3f}}));
    s.objects.push_back(SimpleObject::make(AxisBoxMesh{{0.231, 0, 0.117}, {0.531, 0.595, 0.414}},
                                           LambertianMaterial{{0.85f, 0.74f, 0.55f}},
                                           Transform3::from_axis_angle({0, 1, 0}, Angle::from_degrees(15))));
    std::vector<PrimitiveMesh> members;
    for (int i = 0; i < 30; ++i) {
        const double x = 0.07 * i, y = 2 + 0.03 * (i % 5), z = 0.5 - 0.05 * i;
        if (i % 2) members.push_back(SphereMesh{{x, y, z}, 0.1});
        else members.push_back(AxisBoxMesh{{x, y, z}, {x + 0.1, y + 0.2, z + 0.1}});
    }
    s.objects.push_back(SimpleObject::make(BvhMesh{members}, LambertianMaterial{{0.85f, 0.85f, 0.85f}},
                                           Transform3::from_scale_rotation_translation(
                                               {1, 1, 1}, {0, 1, 0}, Angle::from_degrees(15), {-1, 2.7, 3.95})));
    s.objects.push_back(SimpleObject::new_uncorrected(MeshList{tris}, MetalMaterial{{0.5f, 0.6f, 0.7f}, 0.0},
                                                      Transform3::from_scale({2, 1, 0.5}).then_translate({0, 3, 0})));
    s.objects.push_back(VolumetricObject::make(SphereMesh{{3, 1, 3}, 0.5}, IsotropicMaterial{{0.5f, 0.5f, 0.5f}, 0.1}, 2.0,
                                               Transform3::from_scale({1.5, 1, 1})));
    s.skybox = SkyboxInstance::SimpleSkybox;
    detail::SceneArrays a(s);
    std::printf("%d", a.desc.n_transforms);
    for (double v : a.transforms) std::printf(" %.17g", v);
    std::printf("\n");
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
        SimpleObject(AxisBoxMesh((0.231, 0, 0.117), (0.531, 0.595, 0.414)), LambertianMaterial((0.85, 0.74, 0.55)),
                     Transform3.from_axis_angle((0, 1, 0), 15.0)),
    ]
    members = []
    for i in range(30):
        x, y, z = 0.07 * i, 2 + 0.03 * (i % 5), 0.5 - 0.05 * i
        members.append(SphereMesh((x, y, z), 0.1) if i % 2 else AxisBoxMesh((x, y, z), (x + 0.1, y + 0.2, z + 0.1)))
    objs.append(SimpleObject(BvhMesh(members), LambertianMaterial((0.85, 0.85, 0.85)),
                             Transform3.from_scale_rotation_translation((1, 1, 1), (0, 1, 0), 15.0, (-1, 2.7, 3.95))))
    objs.append(SimpleObject(TriangleListMesh(tv, tn), MetalMaterial((0.5, 0.6, 0.7), 0.0),
                             Transform3.from_scale((2, 1, 0.5)).then_translate((0, 3, 0)), corrected=False))
    objs.append(VolumetricObject(SphereMesh((3, 1, 3), 0.5), IsotropicMaterial((0.5, 0.5, 0.5), 0.1), 2.0,
                                 Transform3.from_scale((1.5, 1, 1))))
    return Scene(objs, SimpleSkybox)


def test_cpp_scene_flattening_matches_python(tmp_path):
    src = tmp_path / "probe.cpp"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    lib_dir = os.path.join(ROOT, "rayna_amd")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", f"-I{os.path.join(ROOT, 'include')}", str(src),
                    f"-L{lib_dir}", "-lrayna_amd", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    xf = lines[0].split()
    sc = _python_scene()
    mats = [object_matrix(o) for o in sc.objects if o.transform is not None]
    assert int(xf[0]) == len(mats) == 4
    got = np.array([float(v) for v in xf[1:]]).reshape(-1, 12)
    for g, m in zip(got, mats):
        assert np.allclose(g, m[:3, :4].reshape(-1), rtol=0, atol=1e-15)
    out = lines[1].split()
    h, n = int(out[0]), int(out[1])
    ph, obj, loc, _ = prepare_scene(sc)
    assert h == ph and n == len(obj)
    assert [f"{o}:{l}" for o, l in zip(obj, loc)] == out[2:]


def test_cpp_tests_compile_cleanly(tmp_path):
    for name in os.listdir(os.path.join(ROOT, "tests", "cpp")):
        if name.endswith(".cpp"):
            subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                            f"-I{os.path.join(ROOT, 'include')}", os.path.join(ROOT, "tests", "cpp", name)], check=True)
