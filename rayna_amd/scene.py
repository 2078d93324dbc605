"""Scene, camera and render options: the rayna_engine API surface the GPU renderer accepts.

Names and argument meanings follow the engine:
  materials  rayna_engine/src/material/{lambertian,metal,dielectric,isotropic,light}.rs
  meshes     rayna_engine/src/mesh/primitive/{sphere,triangle,axis_box}.rs, mesh/planar/parallelogram.rs,
             mesh/advanced/list.rs (a list of triangles), mesh/advanced/bvh.rs (a BVH over primitive meshes),
             mesh/isosurface/raymarched.rs
  objects    rayna_engine/src/object/{simple,volumetric}.rs, with object/transform.rs (ObjectTransform)
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

# ------------------------------------------------------------------------------------------ textures


@dataclass
class SolidTexture:
    """SolidTexture { albedo } (texture/solid.rs:8-30); a plain colour tuple means the same."""
    albedo: Sequence[float] = (0.0, 0.0, 0.0)


@dataclass
class WorldCheckerTexture:
    """WorldCheckerTexture (texture/checker.rs:12-26): cells of `scale` in world space, shifted by `offset`."""
    even: "TextureT"
    odd: "TextureT"
    scale: float = 1.0
    offset: Sequence[float] = (0.0, 0.0, 0.0)


@dataclass
class UvCheckerTexture:
    """UvCheckerTexture (texture/checker.rs:28-42): cells of `scale` in (u, v), shifted by `offset`."""
    even: "TextureT"
    odd: "TextureT"
    scale: float = 1.0
    offset: Sequence[float] = (0.0, 0.0)


@dataclass
class ImageTexture:
    """ImageTexture (texture/image.rs:7-40): `image` is a (height, width, 3) float array, row-major (pixel (x, y)
    at image[y, x]); uv is scaled by `scale`, offset by `offset`, and sampled bilinearly."""
    image: np.ndarray
    scale: Sequence[float] = (1.0, 1.0)
    offset: Sequence[float] = (0.0, 0.0)


@dataclass
class Perlin:
    """An improved Perlin noise (Perlin 2002) whose permutation is drawn from `seed` -- a restatement, not the
    `noise` crate's Perlin (its permutation differs, so values differ for the same seed)."""
    seed: int = 0


@dataclass
class ScalePoint:
    """ScalePoint::new(source).set_scale(scale): the noise at the point times `scale`."""
    source: Perlin
    scale: float = 1.0


@dataclass
class ColourSource:
    """ColourSource (texture/noise.rs:18-45): Greyscale (one noise) or Rgb (three), mapped from [-1, 1] to
    [0, 1]."""
    noises: List[Union[Perlin, ScalePoint]]

    @staticmethod
    def greyscale(noise):
        return ColourSource([noise])

    @staticmethod
    def rgb(r, g, b):
        return ColourSource([r, g, b])


@dataclass
class UvNoiseTexture:
    source: ColourSource


@dataclass
class WorldNoiseTexture:
    source: ColourSource


@dataclass
class LocalNoiseTexture:
    source: ColourSource


TextureT = Union[Sequence[float], SolidTexture, WorldCheckerTexture, UvCheckerTexture, ImageTexture, UvNoiseTexture,
                 WorldNoiseTexture, LocalNoiseTexture]


def _solid(t) -> Optional[Sequence[float]]:
    """The colour of a solid texture (a colour tuple or SolidTexture), else None."""
    if isinstance(t, SolidTexture):
        return tuple(t.albedo)
    if isinstance(t, (WorldCheckerTexture, UvCheckerTexture, ImageTexture, UvNoiseTexture, WorldNoiseTexture,
                      LocalNoiseTexture)):
        return None
    return tuple(t)


# ------------------------------------------------------------------------------------------ materials


def _material(kind, tex, t, fuzz=0.0, ior=0.0, density=0.0):
    col = _solid(t)
    if col is None and tex is None:
        raise ValueError("a textured material is described within its scene")
    k = 0 if col is not None else tex(t)
    return A.Material(kind=kind, colour=(C.c_float * 3)(*(col if col is not None else (0.0, 0.0, 0.0))), texture=k,
                      fuzz=fuzz, refractive_index=ior, density=density)


@dataclass
class LambertianMaterial:
    albedo: TextureT = (0.5, 0.5, 0.5)

    def _desc(self, tex=None):
        return _material(A.MAT_LAMBERTIAN, tex, self.albedo)


@dataclass
class MetalMaterial:
    albedo: TextureT
    fuzz: float = 0.0

    def _desc(self, tex=None):
        return _material(A.MAT_METAL, tex, self.albedo, fuzz=float(self.fuzz))


@dataclass
class DielectricMaterial:
    albedo: TextureT = (1.0, 1.0, 1.0)
    refractive_index: float = 1.5
    density: float = 0.0

    def _desc(self, tex=None):
        return _material(A.MAT_DIELECTRIC, tex, self.albedo, ior=float(self.refractive_index),
                         density=float(self.density))


@dataclass
class IsotropicMaterial:
    albedo: TextureT = (0.5, 0.5, 0.5)
    density: float = 1.0

    def _desc(self, tex=None):
        return _material(A.MAT_ISOTROPIC, tex, self.albedo, density=float(self.density))


@dataclass
class LightMaterial:
    emissive: TextureT

    def _desc(self, tex=None):
        return _material(A.MAT_LIGHT, tex, self.emissive)


MaterialT = Union[LambertianMaterial, MetalMaterial, DielectricMaterial, IsotropicMaterial, LightMaterial]

# ------------------------------------------------------------------------------------------ meshes


@dataclass
class SphereMesh:
    pos: Sequence[float]
    radius: float


@dataclass
class TriangleMesh:
    """One `Triangle::new(vertices, normals)`; normals must be unit length (triangle.rs:26-29)."""
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
    """`ParallelogramMesh::new(Planar::new(p, u, v))` (planar/mod.rs:79-90)."""
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


@dataclass
class BvhMesh:
    """`BvhMesh::new(meshes)` (mesh/advanced/bvh.rs:28-45): an SAH tree over sphere, triangle, parallelogram and
    axis-box meshes, intersected as one mesh (one material, one transform)."""
    meshes: List[Union[SphereMesh, TriangleMesh, ParallelogramMesh, AxisBoxMesh]]


MeshT = Union[SphereMesh, TriangleMesh, TriangleListMesh, ParallelogramMesh, AxisBoxMesh, RaymarchedIsosurfaceMesh,
              BvhMesh]

# ------------------------------------------------------------------------------------------ transforms


class Transform3:
    """An affine map of 3D space held as a 4x4 matrix (`glamour::Transform3<f64>`, core/types.rs:18), mapping
    object space to world space.  `a.then(b)` applies a first, then b; angles are in degrees (Angle::from_degrees).
    """

    def __init__(self, matrix=None):
        self.matrix = np.eye(4) if matrix is None else np.array(matrix, float).reshape(4, 4)

    @staticmethod
    def from_translation(v) -> "Transform3":
        m = np.eye(4)
        m[:3, 3] = np.asarray(v, float)
        return Transform3(m)

    @staticmethod
    def from_scale(v) -> "Transform3":
        m = np.eye(4)
        m[0, 0], m[1, 1], m[2, 2] = (float(x) for x in np.broadcast_to(np.asarray(v, float), (3,)))
        return Transform3(m)

    @staticmethod
    def from_axis_angle(axis, degrees: float) -> "Transform3":
        """A rotation by `degrees` about the unit vector `axis` (Rodrigues' formula, arranged as glam's
        Mat4::from_axis_angle)."""
        x, y, z = (float(a) for a in axis)
        s, c = math.sin(math.radians(degrees)), math.cos(math.radians(degrees))
        omc = 1.0 - c
        xy, xz, yz = x * y * omc, x * z * omc, y * z * omc
        m = np.eye(4)
        m[:3, 0] = (x * x * omc + c, xy + z * s, xz - y * s)
        m[:3, 1] = (xy - z * s, y * y * omc + c, yz + x * s)
        m[:3, 2] = (xz + y * s, yz - x * s, z * z * omc + c)
        return Transform3(m)

    @staticmethod
    def from_scale_rotation_translation(scale, axis, degrees: float, translation) -> "Transform3":
        return Transform3.from_scale(scale).then(Transform3.from_axis_angle(axis, degrees)).then_translate(translation)

    def then(self, other: "Transform3") -> "Transform3":
        return Transform3(other.matrix @ self.matrix)

    def then_translate(self, v) -> "Transform3":
        return self.then(Transform3.from_translation(v))

    def corrected(self, centre) -> "Transform3":
        """ObjectTransform::new_corrected (object/transform.rs): the map applied about `centre`."""
        c = np.asarray(centre, float)
        return Transform3.from_translation(-c).then(self).then_translate(c)

    def map_point(self, p) -> np.ndarray:
        return self.matrix[:3, :3] @ np.asarray(p, float) + self.matrix[:3, 3]

    def map_vector(self, v) -> np.ndarray:
        return self.matrix[:3, :3] @ np.asarray(v, float)

    def inverse(self) -> "Transform3":
        return Transform3(np.linalg.inv(self.matrix))


Transform3.IDENTITY = Transform3()


def mesh_centre(m) -> np.ndarray:
    """Mesh::centre (mesh/mod.rs:93): the point a corrected transform turns about."""
    if isinstance(m, SphereMesh):
        return np.asarray(m.pos, float)
    if isinstance(m, TriangleMesh):
        v = np.asarray(m.vertices, float).reshape(3, 3)
        return (v[0] + v[1] + v[2]) / 3.0
    if isinstance(m, ParallelogramMesh):
        return np.asarray(m.p, float)
    if isinstance(m, AxisBoxMesh):
        a, b = np.asarray(m.a, float), np.asarray(m.b, float)
        return (np.minimum(a, b) + np.maximum(a, b)) / 2.0
    if isinstance(m, BvhMesh):  # the mean of the meshes' centres (bvh.rs:36-44)
        c = np.zeros(3)
        for sub in m.meshes:
            c = c + mesh_centre(sub)
        return c / len(m.meshes)
    if isinstance(m, TriangleListMesh):  # MeshList::new sums its meshes' centres without dividing (list.rs:33-36)
        v = np.asarray(m.vertices, float).reshape(-1, 3, 3)
        return ((v[:, 0] + v[:, 1] + v[:, 2]) / 3.0).sum(axis=0)
    return np.zeros(3)  # SDFs (raymarched.rs:69)


# ------------------------------------------------------------------------------------------ objects


@dataclass
class SimpleObject:
    """`SimpleObject::new(mesh, material, transform)` (object/simple.rs:67-71): the transform turns about the
    mesh's centre; `corrected=False` is `new_uncorrected` (the transform as given)."""
    mesh: MeshT
    material: MaterialT
    transform: Optional[Transform3] = None
    corrected: bool = True


@dataclass
class VolumetricObject:
    """`VolumetricObject::new(mesh, material, density, transform)` (object/volumetric.rs:42-72)."""
    mesh: MeshT
    material: MaterialT
    density: float
    transform: Optional[Transform3] = None
    corrected: bool = True


def object_matrix(o) -> Optional[np.ndarray]:
    """The object-to-world matrix an object's ObjectTransform holds, or None for no transform."""
    if o.transform is None:
        return None
    t = o.transform.corrected(mesh_centre(o.mesh)) if o.corrected else o.transform
    return t.matrix


NoSkybox, SimpleSkybox, WhiteSkybox = A.SKY_NONE, A.SKY_SIMPLE, A.SKY_WHITE


@dataclass
class HdrImageSkybox:
    """HdrImageSkybox (skybox/hdri.rs:8-27): the image, (height, width, 3) row-major, looked up by the ray
    direction's sphere uv and sampled bilinearly."""
    image: np.ndarray


@dataclass
class Scene:
    objects: List[Union[SimpleObject, VolumetricObject]]
    skybox: Union[int, HdrImageSkybox] = SimpleSkybox
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
    transforms: List[np.ndarray] = []

    textures: List[A.Texture] = []
    tex_ids = {}
    images: List[np.ndarray] = []
    img_ids = {}

    def img_index(arr) -> int:
        key = id(arr)
        if key not in img_ids:
            a = np.ascontiguousarray(arr, np.float32)
            if a.ndim != 3 or a.shape[2] != 3:
                raise ValueError("an image is a (height, width, 3) array")
            img_ids[key] = len(images)
            images.append(a)
        return img_ids[key]

    def noise_params(d, source):
        ns = source.noises
        scales = {float(n.scale) if isinstance(n, ScalePoint) else 1.0 for n in ns}
        if len(ns) not in (1, 3) or len(scales) != 1:
            raise ValueError("a colour source is one noise or three with one scale")
        seeds = [(n.source if isinstance(n, ScalePoint) else n).seed for n in ns]
        d.noise_rgb = 1 if len(ns) == 3 else 0
        d.noise_seed = (C.c_uint32 * 3)(*(seeds + [0, 0])[:3])
        d.scale = scales.pop()

    def tex_index(t) -> int:
        """The 1-based texture index of a non-solid texture (its sub-textures registered first)."""
        key = id(t)
        if key in tex_ids:
            return tex_ids[key]
        d = A.Texture()
        if isinstance(t, (WorldCheckerTexture, UvCheckerTexture)):
            odd, even = tex_ref(t.odd), tex_ref(t.even)
            d.kind = A.TEX_WORLD_CHECKER if isinstance(t, WorldCheckerTexture) else A.TEX_UV_CHECKER
            d.odd, d.even, d.scale = odd, even, float(t.scale)
            off = list(t.offset) + [0.0] * (3 - len(t.offset))
            d.offset = (C.c_double * 3)(*off[:3])
        elif isinstance(t, ImageTexture):
            d.kind = A.TEX_IMAGE
            d.image = img_index(t.image)
            d.image_scale = (C.c_double * 2)(*t.scale)
            d.offset = (C.c_double * 3)(t.offset[0], t.offset[1], 0.0)
        elif isinstance(t, (UvNoiseTexture, WorldNoiseTexture, LocalNoiseTexture)):
            d.kind = {UvNoiseTexture: A.TEX_UV_NOISE, WorldNoiseTexture: A.TEX_WORLD_NOISE,
                      LocalNoiseTexture: A.TEX_LOCAL_NOISE}[type(t)]
            noise_params(d, t.source)
        else:
            raise TypeError(f"unsupported texture {type(t).__name__}")
        textures.append(d)
        tex_ids[key] = len(textures)
        return tex_ids[key]

    def tex_ref(t) -> int:
        """Any texture as a 1-based index (a solid colour becomes a SOLID texture entry)."""
        col = _solid(t)
        if col is None:
            return tex_index(t)
        d = A.Texture()
        d.kind = A.TEX_SOLID
        d.colour = (C.c_float * 3)(*col)
        textures.append(d)
        return len(textures)

    def mat_index(m):
        key = id(m)
        if key not in mat_ids:
            mat_ids[key] = len(materials)
            materials.append(m._desc(tex_index))
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
        elif isinstance(m, BvhMesh):
            subs = [add_mesh(sub) for sub in m.meshes]
            for sub in m.meshes:
                if not isinstance(sub, (SphereMesh, TriangleMesh, ParallelogramMesh, AxisBoxMesh)):
                    raise TypeError(f"a BvhMesh holds primitive meshes, not {type(sub).__name__}")
            meshes.append(A.Mesh(A.MESH_BVH, subs[0] if subs else len(meshes), len(subs), 0))
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
        m = object_matrix(o)
        ti = 0
        if m is not None:
            transforms.append(np.ascontiguousarray(m[:3, :4], float))
            ti = len(transforms)
        if isinstance(o, VolumetricObject):
            objects.append(A.Object(A.OBJ_VOLUMETRIC, mi, mat_index(o.material), ti, float(o.density)))
        else:
            objects.append(A.Object(A.OBJ_SIMPLE, mi, mat_index(o.material), ti, 0.0))

    keep = {}
    sph = np.ascontiguousarray(np.vstack(spheres) if spheres else np.zeros((0, 4)), float)
    tv = np.ascontiguousarray(np.concatenate(tri_v) if tri_v else np.zeros((0, 3, 3)), float)
    tn = np.ascontiguousarray(np.concatenate(tri_n) if tri_n else np.zeros((0, 3, 3)), float)
    pa = np.ascontiguousarray(np.array(paras, float).reshape(-1, 9))
    bx = np.ascontiguousarray(np.array(boxes, float).reshape(-1, 6))
    xf = np.ascontiguousarray(np.array(transforms, float).reshape(-1, 12))
    obj_arr = (A.Object * max(1, len(objects)))(*objects)
    mesh_arr = (A.Mesh * max(1, len(meshes)))(*meshes)
    mat_arr = (A.Material * max(1, len(materials)))(*materials)
    sdf_arr = (A.Sdf * max(1, len(sdfs)))(*sdfs)
    keep.update(sph=sph, tv=tv, tn=tn, pa=pa, bx=bx, xf=xf, obj=obj_arr, mesh=mesh_arr, mat=mat_arr, sdf=sdf_arr,
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
    if isinstance(scene.skybox, HdrImageSkybox):
        d.skybox, d.sky_image = A.SKY_HDRI, img_index(scene.skybox.image)
    else:
        d.skybox, d.sky_image = int(scene.skybox), 0
    d.transforms, d.n_transforms = _ptr(xf, C.c_double), len(xf)
    tex_arr = (A.Texture * max(1, len(textures)))(*textures)
    img_arr = (A.Image * max(1, len(images)))(*[A.Image(a.shape[1], a.shape[0], _ptr(a, C.c_float)) for a in images])
    keep.update(tex=tex_arr, img=img_arr, img_data=images)
    d.textures, d.n_textures = C.cast(tex_arr, C.POINTER(A.Texture)), len(textures)
    d.images, d.n_images = C.cast(img_arr, C.POINTER(A.Image)), len(images)
    return d, keep
