"""Scene, camera and render options: the rayna_engine API surface the GPU renderer accepts.

Names and argument meanings follow the engine:
  materials  rayna_engine/src/material/{lambertian,metal,dielectric,isotropic,light}.rs
  meshes     rayna_engine/src/mesh/primitive/{sphere,triangle,axis_box}.rs, mesh/planar/parallelogram.rs,
             mesh/advanced/list.rs (a list of triangles), mesh/isosurface/raymarched.rs
  objects    rayna_engine/src/object/{simple,volumetric}.rs (identity transform)
  scene      rayna_engine/src/scene/mod.rs:11-24 (StandardScene: objects + skybox)
  camera     rayna_engine/src/scene/camera.rs:10-38 (angles given in degrees, as Angle::from_degrees)
  options    rayna_engine/src/render/render_opts.rs:8-72
A scene becomes the flat arrays of `rayna_scene_desc` (include/rayna_gpu.h) with `Scene.desc()`.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Union

import numpy as np
from enum import IntEnum

from . import _abi as A

# ------------------------------------------------------------------------------------------ materials


@dataclass
class LambertianMaterial:
    albedo: Sequence[float] = (0.5, 0.5, 0.5)

    def _desc(self):
        return A.Material(A.MAT_LAMBERTIAN, (C.c_float * 3)(*self.albedo), 0.0, 0.0, 0.0)


@dataclass
class MetalMaterial:
    albedo: Sequence[float]
    fuzz: float = 0.0

    def _desc(self):
        return A.Material(A.MAT_METAL, (C.c_float * 3)(*self.albedo), float(self.fuzz), 0.0, 0.0)


@dataclass
class DielectricMaterial:
    albedo: Sequence[float] = (1.0, 1.0, 1.0)
    refractive_index: float = 1.5
    density: float = 0.0

    def _desc(self):
        return A.Material(A.MAT_DIELECTRIC, (C.c_float * 3)(*self.albedo), 0.0, float(self.refractive_index),
                          float(self.density))


@dataclass
class IsotropicMaterial:
    albedo: Sequence[float] = (0.5, 0.5, 0.5)
    density: float = 1.0

    def _desc(self):
        return A.Material(A.MAT_ISOTROPIC, (C.c_float * 3)(*self.albedo), 0.0, 0.0, float(self.density))


@dataclass
class LightMaterial:
    emissive: Sequence[float]

    def _desc(self):
        return A.Material(A.MAT_LIGHT, (C.c_float * 3)(*self.emissive), 0.0, 0.0, 0.0)


MaterialT = Union[LambertianMaterial, MetalMaterial, DielectricMaterial, IsotropicMaterial, LightMaterial]

# ------------------------------------------------------------------------------------------ meshes


@dataclass
class SphereMesh:
    pos: Sequence[float]
    radius: float


@dataclass
class TriangleMesh:
    """One `Triangle::new(vertices, normals)`; normals must be unit length (triangle.rs:246-249)."""
    vertices: Sequence[Sequence[float]]
    normals: Sequence[Sequence[float]]


@dataclass
class TriangleListMesh:
    """A `MeshList` of triangles (mesh/advanced/list.rs), i.e. a `BvhMesh` over them.

    vertices and normals are arrays of shape (n, 3, 3)."""
    vertices: np.ndarray
    normals: np.ndarray


@dataclass
class ParallelogramMesh:
    """`ParallelogramMesh::new(Planar::new(p, u, v))` (planar/mod.rs:303-314)."""
    p: Sequence[float]
    u: Sequence[float]
    v: Sequence[float]


@dataclass
class AxisBoxMesh:
    """`AxisBoxMesh::new(a, b)`: the box spanned by two corners."""
    a: Sequence[float]
    b: Sequence[float]

    @staticmethod
    def new_centred(centre, size):
        c, s = np.asarray(centre, float), np.asarray(size, float)
        return AxisBoxMesh(tuple(c + s / 2.0), tuple(c - s / 2.0))


@dataclass
class SdfSphere:
    centre: Sequence[float]
    radius: float


@dataclass
class SdfSminSpheres:
    """Polynomial smooth minimum (parameter k) of sphere distances, folded in order."""
    centres: np.ndarray
    radii: np.ndarray
    k: float = 0.1


@dataclass
class SdfMandelbulb:
    centre: Sequence[float] = (0.0, 0.0, 0.0)
    power: int = 8
    iterations: int = 10
    bailout: float = 2.0


@dataclass
class RaymarchedIsosurfaceMesh:
    """`RaymarchedIsosurfaceMesh::new_custom(sdf, max_iterations, epsilon)` (raymarched.rs:51-57)."""
    sdf: Union[SdfSphere, SdfSminSpheres, SdfMandelbulb]
    max_iterations: int = 150
    epsilon: float = 1e-7


MeshT = Union[SphereMesh, TriangleMesh, TriangleListMesh, ParallelogramMesh, AxisBoxMesh, RaymarchedIsosurfaceMesh]

# ------------------------------------------------------------------------------------------ objects


@dataclass
class SimpleObject:
    mesh: MeshT
    material: MaterialT


@dataclass
class VolumetricObject:
    mesh: MeshT
    material: MaterialT
    density: float


NoSkybox, SimpleSkybox, WhiteSkybox = A.SKY_NONE, A.SKY_SIMPLE, A.SKY_WHITE


@dataclass
class Scene:
    objects: List[Union[SimpleObject, VolumetricObject]]
    skybox: int = SimpleSkybox
    _cache: Optional[tuple] = field(default=None, repr=False, compare=False)

    def desc(self) -> A.SceneDesc:
        """The flat C description (cached; the arrays stay alive with the Scene)."""
        if self._cache is None:
            self._cache = _build_desc(self)
        return self._cache[0]

    def n_triangles(self) -> int:
        self.desc()
        return self._cache[1]["n_triangles"]

    def materials_list(self) -> List[MaterialT]:
        """The material table in the order the C description holds it (first use by object order)."""
        self.desc()
        return self._cache[1]["materials"]

    def colours(self) -> np.ndarray:
        """The material colours (albedo, or emission for lights) as an (n_materials, 3) float32 array."""
        self.desc()
        return np.array([[self._cache[0].materials[i].colour[c] for c in range(3)]
                         for i in range(self._cache[0].n_materials)], np.float32).reshape(-1, 3)


@dataclass
class Camera:
    pos: Sequence[float] = (0.0, 0.0, 0.0)
    fwd: Sequence[float] = (0.0, 0.0, 1.0)
    v_fov: float = 45.0           # degrees
    focus_dist: float = 1.0
    defocus_angle: float = 0.0    # degrees

    def desc(self) -> A.CameraDesc:
        return A.CameraDesc((C.c_double * 3)(*self.pos), (C.c_double * 3)(*self.fwd), math.radians(self.v_fov),
                            float(self.focus_dist), math.radians(self.defocus_angle))


class RenderMode(IntEnum):
    """RenderMode (render/render_opts.rs:32-52), in the engine's order."""
    PBR = A.MODE_PBR
    RayNormal = A.MODE_RAY_NORMAL
    OutwardNormal = A.MODE_OUTWARD_NORMAL
    Scatter = A.MODE_SCATTER
    FrontFace = A.MODE_FRONT_FACE
    Distance = A.MODE_DISTANCE
    Uv = A.MODE_UV
    Side = A.MODE_SIDE


@dataclass
class RenderOpts:
    width: int = 740
    height: int = 480
    samples: int = 1
    ray_depth: int = 5
    ray_branching: int = 1
    mode: int = A.MODE_PBR
    seed: int = 0

    def desc(self) -> A.RenderOptsDesc:
        return A.RenderOptsDesc(self.width, self.height, self.samples, self.ray_depth, self.ray_branching,
                                self.mode, self.seed)


# ------------------------------------------------------------------------------------------ flattening


def _ptr(arr: np.ndarray, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


def _build_desc(scene: Scene):
    materials: List[A.Material] = []
    mat_objs: List[MaterialT] = []
    mat_ids = {}
    meshes: List[A.Mesh] = []
    spheres: List[np.ndarray] = []
    n_spheres = 0
    tri_v: List[np.ndarray] = []
    tri_n: List[np.ndarray] = []
    n_tri = 0
    paras: List[Sequence[float]] = []
    boxes: List[Sequence[float]] = []
    sdfs: List[A.Sdf] = []
    objects: List[A.Object] = []

    def mat_index(m):
        key = id(m)
        if key not in mat_ids:
            mat_ids[key] = len(materials)
            materials.append(m._desc())
            mat_objs.append(m)
        return mat_ids[key]

    def add_mesh(m) -> int:
        nonlocal n_spheres, n_tri
        if isinstance(m, SphereMesh):
            spheres.append(np.array([[*m.pos, m.radius]], float))
            meshes.append(A.Mesh(A.MESH_SPHERE, n_spheres, 1, 0))
            n_spheres += 1
        elif isinstance(m, TriangleMesh):
            tri_v.append(np.asarray(m.vertices, float).reshape(1, 3, 3))
            tri_n.append(np.asarray(m.normals, float).reshape(1, 3, 3))
            meshes.append(A.Mesh(A.MESH_TRIANGLE, n_tri, 1, 0))
            n_tri += 1
        elif isinstance(m, TriangleListMesh):
            v = np.ascontiguousarray(m.vertices, float).reshape(-1, 3, 3)
            nn = np.ascontiguousarray(m.normals, float).reshape(-1, 3, 3)
            tri_v.append(v)
            tri_n.append(nn)
            meshes.append(A.Mesh(A.MESH_TRIANGLE_LIST, n_tri, len(v), 0))
            n_tri += len(v)
        elif isinstance(m, ParallelogramMesh):
            meshes.append(A.Mesh(A.MESH_PARALLELOGRAM, len(paras), 1, 0))
            paras.append([*m.p, *m.u, *m.v])
        elif isinstance(m, AxisBoxMesh):
            meshes.append(A.Mesh(A.MESH_AXIS_BOX, len(boxes), 1, 0))
            boxes.append([*m.a, *m.b])
        elif isinstance(m, RaymarchedIsosurfaceMesh):
            f = m.sdf
            d = A.Sdf()
            d.max_iterations = int(m.max_iterations)
            d.epsilon = float(m.epsilon)
            if isinstance(f, SdfSphere):
                d.kind = A.SDF_SPHERE
                d.first_sphere, d.n_spheres = n_spheres, 1
                spheres.append(np.array([[*f.centre, f.radius]], float))
                n_spheres += 1
            elif isinstance(f, SdfSminSpheres):
                d.kind = A.SDF_SMIN_SPHERES
                cs = np.asarray(f.centres, float).reshape(-1, 3)
                rs = np.asarray(f.radii, float).reshape(-1, 1)
                d.first_sphere, d.n_spheres = n_spheres, len(cs)
                d.k = float(f.k)
                spheres.append(np.hstack([cs, rs]))
                n_spheres += len(cs)
            elif isinstance(f, SdfMandelbulb):
                d.kind = A.SDF_MANDELBULB
                d.power = float(f.power)
                d.bulb_iterations = int(f.iterations)
                d.bailout = float(f.bailout)
                d.centre = (C.c_double * 3)(*f.centre)
            else:
                raise TypeError(f"unsupported SDF {type(f).__name__}")
            meshes.append(A.Mesh(A.MESH_SDF, len(sdfs), 1, 0))
            sdfs.append(d)
        else:
            raise TypeError(f"unsupported mesh {type(m).__name__}")
        return len(meshes) - 1

    for o in scene.objects:
        mi = add_mesh(o.mesh)
        if isinstance(o, VolumetricObject):
            objects.append(A.Object(A.OBJ_VOLUMETRIC, mi, mat_index(o.material), 0, float(o.density)))
        else:
            objects.append(A.Object(A.OBJ_SIMPLE, mi, mat_index(o.material), 0, 0.0))

    keep = {}
    sph = np.ascontiguousarray(np.vstack(spheres) if spheres else np.zeros((0, 4)), float)
    tv = np.ascontiguousarray(np.concatenate(tri_v) if tri_v else np.zeros((0, 3, 3)), float)
    tn = np.ascontiguousarray(np.concatenate(tri_n) if tri_n else np.zeros((0, 3, 3)), float)
    pa = np.ascontiguousarray(np.array(paras, float).reshape(-1, 9))
    bx = np.ascontiguousarray(np.array(boxes, float).reshape(-1, 6))
    obj_arr = (A.Object * max(1, len(objects)))(*objects)
    mesh_arr = (A.Mesh * max(1, len(meshes)))(*meshes)
    mat_arr = (A.Material * max(1, len(materials)))(*materials)
    sdf_arr = (A.Sdf * max(1, len(sdfs)))(*sdfs)
    keep.update(sph=sph, tv=tv, tn=tn, pa=pa, bx=bx, obj=obj_arr, mesh=mesh_arr, mat=mat_arr, sdf=sdf_arr,
                n_triangles=len(tv), materials=mat_objs)
    d = A.SceneDesc()
    d.objects, d.n_objects = C.cast(obj_arr, C.POINTER(A.Object)), len(objects)
    d.meshes, d.n_meshes = C.cast(mesh_arr, C.POINTER(A.Mesh)), len(meshes)
    d.materials, d.n_materials = C.cast(mat_arr, C.POINTER(A.Material)), len(materials)
    d.spheres, d.n_spheres = _ptr(sph, C.c_double), len(sph)
    d.tri_vertices, d.tri_normals, d.n_triangles = _ptr(tv, C.c_double), _ptr(tn, C.c_double), len(tv)
    d.parallelograms, d.n_parallelograms = _ptr(pa, C.c_double), len(pa)
    d.boxes, d.n_boxes = _ptr(bx, C.c_double), len(bx)
    d.sdfs, d.n_sdfs = C.cast(sdf_arr, C.POINTER(A.Sdf)), len(sdfs)
    d.skybox = int(scene.skybox)
    return d, keep
