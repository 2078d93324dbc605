"""Seeded workload scenes, in the manner of rayna_engine/src/scene/preset.rs.

Each function returns (Scene, Camera, RenderOpts) for one workload:
  random_spheres   the "Ray Tracing in One Weekend" demo (template: RTIAW_DEMO, preset.rs:157-254) with the
                   textbook grid and probabilities; 400x225, 32 spp, depth 10, SimpleSkybox
  triangle_mesh    a value-noise-displaced icosphere of 20 * 4^subdiv triangles (1,310,720 at subdiv 8), one
                   material per base face, a two-triangle ground and 8 parallelogram lights; 1920x1080,
                   64 spp, depth 8, NoSkybox
  sdf_volume       a power-8 Mandelbulb and 16 smooth-min spheres inside a 4x4x4 constant-density isotropic
                   medium; 1280x720, 128 spp, depth 12, SimpleSkybox
  sphere_colours   the scene of the engine's one integration test (rayna_engine/tests/simple_renders.rs:35-53)
  cornell          the engine's CORNELL preset (preset.rs:606-683): five walls, a light and two boxes rotated
                   about their centres; NoSkybox
  transformed_mix  objects of every transformable kind under rotations, scalings and shears (parity workload)
  rttnw            the engine's RTTNW_DEMO (preset.rs:394-605): a BvhMesh floor of 400 boxes, a BvhMesh cube of
                   1000 balls rotated as an instance, glass, metal and medium-filled spheres, a noise-textured
                   sphere, a light; NoSkybox
  textured         every texture kind (world and uv checkers, image, uv / world / local noise) on every primitive,
                   a textured light and an image sky (parity workload)
All randomness comes from numpy's PCG64 seeded with `seed`, so every scene is a pure function of its arguments.
"""
from __future__ import annotations

import math

import numpy as np

from .scene import (AxisBoxMesh, BvhMesh, Camera, ColourSource, HdrImageSkybox, ImageTexture, LocalNoiseTexture,
                    Perlin, ScalePoint, UvCheckerTexture, UvNoiseTexture, WorldCheckerTexture, WorldNoiseTexture, DielectricMaterial, IsotropicMaterial, LambertianMaterial, LightMaterial,
                    MetalMaterial, NoSkybox, ParallelogramMesh, RaymarchedIsosurfaceMesh, RenderOpts, Scene,
                    SdfMandelbulb, SdfSminSpheres, SimpleObject, SimpleSkybox, SphereMesh, Transform3,
                    TriangleListMesh, TriangleMesh, VolumetricObject, WhiteSkybox)

# ------------------------------------------------------------------------------------------ random spheres


def random_spheres(seed: int = 0, width: int = 400, height: int = 225, samples: int = 32, ray_depth: int = 10):
    rng = np.random.Generator(np.random.PCG64(seed))
    objects = []
    big = np.array([4.0, 0.2, 0.0])
    for a in range(-11, 11):
        for b in range(-11, 11):
            choose = rng.random()
            centre = np.array([a + 0.9 * rng.random(), 0.2, b + 0.9 * rng.random()])
            if np.linalg.norm(centre - big) <= 0.9:  # preset.rs:169
                continue
            if choose < 0.8:
                albedo = rng.random(3) * rng.random(3)  # rng::colour_rgb * rng::colour_rgb (preset.rs:176)
                mat = LambertianMaterial(tuple(float(x) for x in albedo))
            elif choose < 0.95:
                albedo = 0.5 + 0.5 * rng.random(3)
                mat = MetalMaterial(tuple(float(x) for x in albedo), float(0.5 * rng.random()))
            else:
                mat = DielectricMaterial((1.0, 1.0, 1.0), 1.5, 0.0)
            objects.append(SimpleObject(SphereMesh(tuple(centre), 0.2), mat))
    objects.append(SimpleObject(SphereMesh((0.0, 1.0, 0.0), 1.0), DielectricMaterial((1.0, 1.0, 1.0), 1.5, 0.0)))
    objects.append(SimpleObject(SphereMesh((-4.0, 1.0, 0.0), 1.0), LambertianMaterial((0.4, 0.2, 0.1))))
    objects.append(SimpleObject(SphereMesh((4.0, 1.0, 0.0), 1.0), MetalMaterial((0.7, 0.6, 0.5), 0.0)))
    objects.append(SimpleObject(SphereMesh((0.0, -1000.0, 0.0), 1000.0), LambertianMaterial((0.5, 0.5, 0.5))))
    fwd = np.array([-13.0, -2.0, -3.0])
    cam = Camera(pos=(13.0, 2.0, 3.0), fwd=tuple(fwd / np.linalg.norm(fwd)), v_fov=20.0, focus_dist=10.0,
                 defocus_angle=0.6)
    opts = RenderOpts(width=width, height=height, samples=samples, ray_depth=ray_depth, seed=seed)
    return Scene(objects, SimpleSkybox), cam, opts


# ------------------------------------------------------------------------------------------ icosphere

_PHI = (1.0 + math.sqrt(5.0)) / 2.0
_ICO_V = np.array([[-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
                   [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
                   [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1]], float)
_ICO_F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                   [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                   [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                   [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], int)


def _value_noise(points: np.ndarray, seed: int) -> np.ndarray:
    """Lattice value noise in [-1, 1]: seeded random values on the integer lattice (a 256-entry permutation
    hash), trilinearly interpolated with smoothstep weights."""
    rng = np.random.Generator(np.random.PCG64(seed))
    perm = rng.permutation(256)
    vals = rng.random(256) * 2.0 - 1.0
    fl = np.floor(points)
    f = points - fl
    i = fl.astype(np.int64)
    w = f * f * (3.0 - 2.0 * f)

    def lat(dx, dy, dz):
        h = perm[(i[:, 0] + dx) & 255]
        h = perm[(h + i[:, 1] + dy) & 255]
        h = perm[(h + i[:, 2] + dz) & 255]
        return vals[h]

    out = np.zeros(len(points))
    for dx in (0, 1):
        wx = w[:, 0] if dx else 1.0 - w[:, 0]
        for dy in (0, 1):
            wy = w[:, 1] if dy else 1.0 - w[:, 1]
            for dz in (0, 1):
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out += wx * wy * wz * lat(dx, dy, dz)
    return out


def icosphere(subdiv: int, seed: int, amplitude: float = 0.1, frequency: float = 4.0, octaves: int = 3):
    """Vertices (n, 3, 3) and area-weighted vertex normals (n, 3, 3) of a displaced icosphere, and the base
    face of every triangle.  Each base face is cut into the 4^subdiv triangles of a barycentric grid with
    N = 2^subdiv segments per edge; grid points are shared exactly between faces (keyed by their integer
    barycentric weights), projected to the unit sphere and moved radially to
    r = 1 + amplitude * sum_o 2^-o noise(2^o * frequency * p) / sum_o 2^-o."""
    N = 1 << subdiv
    ii, jj = np.meshgrid(np.arange(N + 1), np.arange(N + 1), indexing="ij")
    mask = ii + jj <= N
    gi, gj = ii[mask], jj[mask]
    gk = N - gi - gj
    local = -np.ones((N + 1, N + 1), np.int64)
    local[gi, gj] = np.arange(len(gi))
    # triangles of one face in local grid indices: up (i,j),(i+1,j),(i,j+1); down (i+1,j),(i+1,j+1),(i,j+1)
    ui, uj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    up = ui + uj <= N - 1
    dn = ui + uj <= N - 2
    t_up = np.stack([local[ui[up], uj[up]], local[ui[up] + 1, uj[up]], local[ui[up], uj[up] + 1]], 1)
    t_dn = np.stack([local[ui[dn] + 1, uj[dn]], local[ui[dn] + 1, uj[dn] + 1], local[ui[dn], uj[dn] + 1]], 1)
    face_tris = np.concatenate([t_up, t_dn])  # N^2 per face

    keys_all, tris_all, face_of = [], [], []
    sentinel = 12 * 1024
    for f, (a, b, c) in enumerate(_ICO_F):
        codes = np.stack([np.where(gk > 0, a * 1024 + gk, sentinel),   # weight on a = k
                          np.where(gi > 0, b * 1024 + gi, sentinel),   # weight on b = i
                          np.where(gj > 0, c * 1024 + gj, sentinel)], 1)
        codes.sort(axis=1)
        keys = (codes[:, 0] << 28) | (codes[:, 1] << 14) | codes[:, 2]
        keys_all.append(keys)
        tris_all.append(face_tris + f * len(gi))
        face_of.append(np.full(len(face_tris), f))
    keys_all = np.concatenate(keys_all)
    uniq, inv = np.unique(keys_all, return_inverse=True)
    # position of each unique point from its canonical key (sorted vertex order)
    codes = np.stack([(uniq >> 28) & 0x3fff, (uniq >> 14) & 0x3fff, uniq & 0x3fff], 1)
    pos = np.zeros((len(uniq), 3))
    for m in range(3):
        valid = codes[:, m] != sentinel
        vi = codes[:, m] // 1024
        wt = codes[:, m] % 1024
        pos[valid] += _ICO_V[vi[valid]] * wt[valid, None]
    pos /= N
    unit = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    n = np.zeros(len(unit))
    norm = 0.0
    for o in range(octaves):
        n += 2.0 ** -o * _value_noise(unit * (frequency * 2.0 ** o), seed * 7919 + o)
        norm += 2.0 ** -o
    verts = unit * (1.0 + amplitude * n / norm)[:, None]
    tris = inv[np.concatenate(tris_all)]
    fn = np.cross(verts[tris[:, 1]] - verts[tris[:, 0]], verts[tris[:, 2]] - verts[tris[:, 0]])
    vn = np.zeros_like(verts)
    for m in range(3):
        np.add.at(vn, tris[:, m], fn)
    vn /= np.linalg.norm(vn, axis=1, keepdims=True)
    # Wind clockwise as seen from outside: the engine's Triangle::intersect treats the side that
    # (v1 - v0) x (v2 - v0) faces as the back face (front_face = det < 0, triangle.rs:102), so outward-facing
    # hits need the opposite winding; the vertex normals stay outward.
    tris = tris[:, [0, 2, 1]]
    return verts[tris], vn[tris], np.concatenate(face_of)


def triangle_mesh(seed: int = 1, subdiv: int = 8, width: int = 1920, height: int = 1080, samples: int = 64,
                  ray_depth: int = 8):
    rng = np.random.Generator(np.random.PCG64(seed))
    tv, tn, face = icosphere(subdiv, seed)
    objects = []
    for f in range(20):
        sel = face == f
        if rng.random() < 0.7:
            mat = LambertianMaterial(tuple(float(x) for x in 0.2 + 0.7 * rng.random(3)))
        else:
            mat = MetalMaterial(tuple(float(x) for x in 0.5 + 0.5 * rng.random(3)), float(0.3 * rng.random()))
        objects.append(SimpleObject(TriangleListMesh(tv[sel], tn[sel]), mat))
    # ground: two triangles at y = -1.25 spanning [-8, 8]^2, facing up (wound clockwise seen from above)
    g = -1.25
    quad = np.array([[-8, g, -8], [-8, g, 8], [8, g, 8], [8, g, -8]], float)
    gv = np.array([[quad[0], quad[2], quad[1]], [quad[0], quad[3], quad[2]]])
    gn = np.tile(np.array([0.0, 1.0, 0.0]), (2, 3, 1))
    objects.append(SimpleObject(TriangleListMesh(gv, gn), LambertianMaterial((0.5, 0.5, 0.5))))
    # 8 area lights on a ring of radius 3 at height 2.5, 1x1, facing down (light is two-sided)
    for i in range(8):
        ang = 2.0 * math.pi * i / 8.0
        c = np.array([3.0 * math.cos(ang), 2.5, 3.0 * math.sin(ang)])
        emit = tuple(float(x) for x in 3.0 + 3.0 * rng.random(3))
        objects.append(SimpleObject(ParallelogramMesh(tuple(c - [0.5, 0, 0.5]), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)),
                                    LightMaterial(emit)))
    pos = np.array([0.0, 0.9, 3.6])
    cam = Camera(pos=tuple(pos), fwd=tuple(-pos / np.linalg.norm(pos)), v_fov=38.0,
                 focus_dist=float(np.linalg.norm(pos)), defocus_angle=0.0)
    opts = RenderOpts(width=width, height=height, samples=samples, ray_depth=ray_depth, seed=seed)
    return Scene(objects, NoSkybox), cam, opts


# ------------------------------------------------------------------------------------------ SDF + volume


def sdf_volume(seed: int = 2, width: int = 1280, height: int = 720, samples: int = 128, ray_depth: int = 12,
               max_iterations: int = 256, epsilon: float = 1e-4):
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(-2.0, 2.0, (16, 3))
    radii = rng.uniform(0.2, 0.6, 16)
    bulb = RaymarchedIsosurfaceMesh(SdfMandelbulb((0.0, 0.0, 0.0), 8, 10, 2.0), max_iterations, epsilon)
    blobs = RaymarchedIsosurfaceMesh(SdfSminSpheres(centres, radii, 0.1), max_iterations, epsilon)
    objects = [
        SimpleObject(bulb, LambertianMaterial((0.8, 0.55, 0.35))),
        SimpleObject(blobs, LambertianMaterial((0.35, 0.5, 0.8))),
        # isotropic density 0: the medium's sigma_t lives in the VolumetricObject alone
        VolumetricObject(AxisBoxMesh.new_centred((0.0, 0.0, 0.0), (4.0, 4.0, 4.0)), IsotropicMaterial((0.8, 0.8, 0.8), 0.0),
                         0.5),
    ]
    pos = np.array([0.0, 1.5, 7.0])
    cam = Camera(pos=tuple(pos), fwd=tuple(-pos / np.linalg.norm(pos)), v_fov=40.0,
                 focus_dist=float(np.linalg.norm(pos)), defocus_angle=0.0)
    opts = RenderOpts(width=width, height=height, samples=samples, ray_depth=ray_depth, seed=seed)
    return Scene(objects, SimpleSkybox), cam, opts


# ------------------------------------------------------------------------------------------ engine test scene


def sphere_colours(albedo, cam_z: float, width: int = 320, height: int = 320, samples: int = 10, ray_depth: int = 5,
                   seed: int = 0):
    """The unit Lambertian sphere under a white sky of tests/simple_renders.rs:35-53, camera at (0, 0, cam_z)
    looking along +z with a 45 degree field of view (tests/common/mod.rs:13-22 for the options)."""
    scene = Scene([SimpleObject(SphereMesh((0.0, 0.0, 0.0), 1.0), LambertianMaterial(tuple(albedo)))], WhiteSkybox)
    cam = Camera(pos=(0.0, 0.0, cam_z), fwd=(0.0, 0.0, 1.0), v_fov=45.0, focus_dist=1.0, defocus_angle=0.0)
    return scene, cam, RenderOpts(width=width, height=height, samples=samples, ray_depth=ray_depth, seed=seed)


def cornell(width: int = 600, height: int = 600, samples: int = 64, ray_depth: int = 8, seed: int = 0):
    """CORNELL (preset.rs:606-683): the walls as parallelograms (SimpleObject::new with no transform), the light
    quad, and the two boxes under Transform3::from_axis_angle(Y, 15 deg / -18 deg), corrected about their centres."""
    red, green, grey = (0.65, 0.05, 0.05), (0.12, 0.45, 0.15), (0.85, 0.74, 0.55)
    X, Y, Z = (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)
    grey_m = LambertianMaterial(grey)

    def quad(p, u, v, mat):
        return SimpleObject(ParallelogramMesh(p, u, v), mat)

    objects = [
        quad((0.0, 0.0, 0.0), Y, Z, LambertianMaterial(red)),   # left
        quad((0.0, 0.0, 0.0), X, Y, LambertianMaterial(grey)),  # back
        quad((0.0, 0.0, 0.0), Z, X, LambertianMaterial(grey)),  # floor
        quad((1.0, 0.0, 0.0), Z, Y, LambertianMaterial(green)),  # right
        quad((0.0, 1.0, 0.0), X, Z, LambertianMaterial(grey)),  # ceiling
        quad((0.4, 0.9999, 0.4), (0.2, 0.0, 0.0), (0.0, 0.0, 0.2), LightMaterial((15.0, 15.0, 15.0))),
        SimpleObject(AxisBoxMesh((0.231, 0.0, 0.117), (0.531, 0.595, 0.414)), grey_m,
                     Transform3.from_axis_angle(Y, 15.0)),
        SimpleObject(AxisBoxMesh((0.477, 0.0, 0.531), (0.774, 0.297, 0.829)), LambertianMaterial(grey),
                     Transform3.from_axis_angle(Y, -18.0)),
    ]
    cam = Camera(pos=(0.5, 0.5, 2.3), fwd=(0.0, 0.0, -1.0), v_fov=40.0, focus_dist=1.0, defocus_angle=0.0)
    opts = RenderOpts(width=width, height=height, samples=samples, ray_depth=ray_depth, seed=seed)
    return Scene(objects, NoSkybox), cam, opts


def transformed_mix(seed: int = 6, width: int = 96, height: int = 64, samples: int = 4, ray_depth: int = 6):
    """Every transformable object kind (sphere, triangle, parallelogram, axis box, sphere and box volumes) under
    seeded rotations about random axes, non-uniform scalings and a shear, corrected and uncorrected, beside
    untransformed objects and a light, under a SimpleSkybox."""
    rng = np.random.Generator(np.random.PCG64(seed))

    def rot():
        a = rng.normal(size=3)
        return Transform3.from_axis_angle(a / np.linalg.norm(a), float(rng.uniform(-180.0, 180.0)))

    def scale():
        return Transform3.from_scale(rng.uniform(0.5, 1.8, size=3))

    shear = Transform3(np.array([[1.0, 0.35, 0.0, 0.0], [0.0, 1.0, 0.0, 0.0], [0.2, 0.0, 1.0, 0.0],
                                 [0.0, 0.0, 0.0, 1.0]]))
    mats = [LambertianMaterial((0.7, 0.3, 0.2)), MetalMaterial((0.8, 0.8, 0.7), 0.2),
            DielectricMaterial((0.95, 0.95, 1.0), 1.5, 0.3), LambertianMaterial((0.2, 0.5, 0.7))]
    objects = [
        SimpleObject(SphereMesh((-2.0, 1.0, 0.0), 0.7), mats[1], scale().then(rot())),
        SimpleObject(TriangleMesh(((-1.0, 0.2, -1.0), (1.0, 0.4, -1.2), (0.0, 1.8, -1.5)),
                                  ((0.0, 0.0, 1.0), (0.2, 0.0, 0.98), (0.0, 0.3, 0.954))), mats[0], rot()),
        SimpleObject(ParallelogramMesh((0.5, 0.1, 0.5), (1.0, 0.0, 0.2), (0.0, 1.0, 0.3)), mats[3], shear),
        SimpleObject(AxisBoxMesh((1.5, 0.0, -0.5), (2.5, 1.2, 0.5)), mats[0], scale().then(rot())),
        SimpleObject(AxisBoxMesh((-0.6, 0.0, 1.0), (0.6, 0.8, 1.8)), mats[2], rot()),
        SimpleObject(SphereMesh((0.0, 0.5, 0.0), 0.5), mats[3], Transform3.from_translation((0.3, 0.4, -2.0)),
                     corrected=False),
        VolumetricObject(AxisBoxMesh.new_centred((-1.5, 0.8, 1.5), (1.0, 1.4, 1.0)),
                         IsotropicMaterial((0.9, 0.8, 0.7), 0.05), 1.2, rot()),
        VolumetricObject(SphereMesh((1.2, 0.8, 1.8), 0.6), IsotropicMaterial((0.6, 0.7, 0.9), 0.1), 2.0,
                         scale().then(shear)),
        SimpleObject(ParallelogramMesh((-1.0, 3.0, -1.0), (2.0, 0.0, 0.0), (0.0, 0.0, 2.0)),
                     LightMaterial((4.0, 4.0, 4.0))),
        SimpleObject(SphereMesh((0.0, -1000.0, 0.0), 1000.0), LambertianMaterial((0.5, 0.5, 0.5))),
    ]
    cam = Camera(pos=(0.0, 2.0, 7.0), fwd=(0.0, -0.25, -1.0), v_fov=50.0, focus_dist=6.0, defocus_angle=0.0)
    opts = RenderOpts(width=width, height=height, samples=samples, ray_depth=ray_depth, seed=seed)
    return Scene(objects, SimpleSkybox), cam, opts


def rttnw(seed: int = 0, width: int = 600, height: int = 600, samples: int = 32, ray_depth: int = 10,
          balls: int = 1000):
    """RTTNW_DEMO (preset.rs:394-605) with numpy's PCG64 for the engine's thread_rng: box heights U[0, 1] and ball
    centres U[0, 1)^3 * 0.825.  The noise sphere's WorldNoiseTexture(Greyscale(ScalePoint(Perlin(69), 4))) uses the
    improved-Perlin restatement, whose permutation differs from the noise crate's."""
    rng = np.random.Generator(np.random.PCG64(seed))
    objects = []
    floor = []
    for i in range(20):  # BOXES (FLOOR): COUNT = 20, WIDTH = 1
        for j in range(20):
            low = np.array([-10.0 + i, 0.0, -10.0 + j])
            high = low + np.array([1.0, float(rng.uniform(0.0, 1.0)), 1.0])
            floor.append(AxisBoxMesh(tuple(low), tuple(high)))
    objects.append(SimpleObject(BvhMesh(floor), LambertianMaterial((0.48, 0.83, 0.53))))
    objects.append(SimpleObject(ParallelogramMesh((1.23, 5.54, 1.47), (3.0, 0.0, 0.0), (0.0, 0.0, 2.65)),
                                LightMaterial((7.0, 7.0, 7.0))))
    objects.append(SimpleObject(SphereMesh((4.0, 4.0, 2.0), 0.5), LambertianMaterial((0.7, 0.3, 0.1))))
    objects.append(SimpleObject(SphereMesh((2.6, 1.5, 0.45), 0.5), DielectricMaterial((1.0, 1.0, 1.0), 1.5, 1.0)))
    objects.append(SimpleObject(SphereMesh((0.0, 1.5, 1.45), 0.5), MetalMaterial((0.8, 0.8, 0.9), 1.0)))
    objects.append(SimpleObject(SphereMesh((3.6, 1.5, 1.45), 0.7), DielectricMaterial((1.0, 1.0, 1.0), 1.5, 0.0)))
    objects.append(VolumetricObject(SphereMesh((3.6, 1.5, 1.45), 0.6999), IsotropicMaterial((0.2, 0.4, 0.9), 0.3),
                                    2.0))
    noise = WorldNoiseTexture(ColourSource.greyscale(ScalePoint(Perlin(69), 4.0)))
    objects.append(SimpleObject(SphereMesh((2.2, 2.8, 3.0), 0.8), LambertianMaterial(noise)))
    spread = 0.825  # CUBE OF BALLS
    cube = [SphereMesh(tuple(float(x) for x in rng.random(3) * spread), 0.1) for _ in range(balls)]
    objects.append(SimpleObject(BvhMesh(cube), LambertianMaterial((0.85, 0.85, 0.85)),
                                Transform3.from_scale_rotation_translation(
                                    (1.0, 1.0, 1.0), (0.0, 1.0, 0.0), 15.0,
                                    np.array([-1.0, 2.7, 3.95]) + spread)))
    fwd = np.array([-1.0, 0.0, 3.0])
    cam = Camera(pos=(4.78, 2.78, -6.0), fwd=tuple(fwd / np.linalg.norm(fwd)), v_fov=40.0, focus_dist=1.0,
                 defocus_angle=0.0)
    opts = RenderOpts(width=width, height=height, samples=samples, ray_depth=ray_depth, seed=seed)
    return Scene(objects, NoSkybox), cam, opts


def test_image(width: int, height: int, seed: int) -> np.ndarray:
    """A (height, width, 3) float image: smooth colour ramps crossed by seeded stripes."""
    rng = np.random.Generator(np.random.PCG64(seed))
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    img = np.stack([x / max(1, width - 1), y / max(1, height - 1), 0.5 + 0.5 * np.sin(0.7 * x + 0.3 * y)], axis=-1)
    img[:, rng.integers(0, width, size=max(1, width // 8))] *= 0.25
    return img.astype(np.float32)


def textured(seed: int = 7, width: int = 96, height: int = 64, samples: int = 4, ray_depth: int = 6):
    """Every texture kind: a world-checker ground, an image-textured sphere, uv checkers and uv noise on a box and
    a parallelogram, a world-noise sphere, an rgb local-noise box under a rotation, an image-textured light, and an
    image sky."""
    checker = WorldCheckerTexture(even=(0.9, 0.9, 0.9), odd=(0.2, 0.3, 0.1), scale=0.5, offset=(0.25, 0.0, 0.1))
    nested = UvCheckerTexture(even=checker, odd=(0.8, 0.2, 0.2), scale=0.25, offset=(0.1, 0.0))
    img = ImageTexture(test_image(37, 23, seed), scale=(2.0, 1.0), offset=(0.1, 0.0))
    rgb_noise = LocalNoiseTexture(ColourSource.rgb(ScalePoint(Perlin(seed), 3.0), ScalePoint(Perlin(seed + 1), 3.0),
                                                   ScalePoint(Perlin(seed + 2), 3.0)))
    objects = [
        SimpleObject(ParallelogramMesh((-20.0, 0.0, -20.0), (40.0, 0.0, 0.0), (0.0, 0.0, 40.0)),
                     LambertianMaterial(checker)),
        SimpleObject(SphereMesh((-1.6, 0.8, 0.0), 0.8), LambertianMaterial(img)),
        SimpleObject(AxisBoxMesh((-0.3, 0.0, -0.6), (0.5, 1.2, 0.2)), MetalMaterial(nested, 0.3)),
        SimpleObject(SphereMesh((1.5, 0.7, 0.3), 0.7),
                     LambertianMaterial(WorldNoiseTexture(ColourSource.greyscale(ScalePoint(Perlin(69), 4.0))))),
        SimpleObject(AxisBoxMesh.new_centred((0.2, 0.5, 1.6), (0.8, 1.0, 0.8)), LambertianMaterial(rgb_noise),
                     Transform3.from_axis_angle((0.0, 1.0, 0.0), 30.0)),
        SimpleObject(ParallelogramMesh((2.2, 0.05, -1.5), (1.0, 0.0, 0.2), (0.0, 1.2, 0.0)),
                     DielectricMaterial(UvNoiseTexture(ColourSource.greyscale(ScalePoint(Perlin(5), 6.0))), 1.4, 0.4)),
        SimpleObject(TriangleMesh(((-2.5, 2.5, -2.0), (-0.5, 2.5, -2.2), (-1.5, 3.5, -2.1)),
                                  ((0.0, 0.0, 1.0), (0.0, 0.0, 1.0), (0.0, 0.0, 1.0))),
                     LightMaterial(ImageTexture(test_image(8, 8, seed + 3) * 6.0))),
    ]
    sky = HdrImageSkybox(test_image(64, 32, seed + 5) * 1.5)
    cam = Camera(pos=(0.0, 2.0, 6.0), fwd=(0.0, -0.25, -1.0), v_fov=55.0, focus_dist=6.0, defocus_angle=0.0)
    opts = RenderOpts(width=width, height=height, samples=samples, ray_depth=ray_depth, seed=seed)
    return Scene(objects, sky), cam, opts


def mixed_primitives(seed: int = 5, width: int = 96, height: int = 64, samples: int = 4, ray_depth: int = 6):
    """A small scene with every primitive kind, material kind and a box medium, for parity tests."""
    rng = np.random.Generator(np.random.PCG64(seed))
    objects = []
    for i in range(12):
        c = (float(rng.uniform(-3, 3)), float(rng.uniform(0.2, 1.5)), float(rng.uniform(-3, 1)))
        k = i % 4
        mat = [LambertianMaterial((0.7, 0.3, 0.2)), MetalMaterial((0.8, 0.8, 0.9), 0.2),
               DielectricMaterial((0.9, 1.0, 0.9), 1.5, 0.3), LambertianMaterial((0.2, 0.6, 0.3))][k]
        if i % 3 == 0:
            objects.append(SimpleObject(AxisBoxMesh.new_centred(c, (0.5, 0.7, 0.4)), mat))
        else:
            objects.append(SimpleObject(SphereMesh(c, float(rng.uniform(0.2, 0.6))), mat))
    tv = rng.uniform(-1, 1, (40, 3, 3)) * 0.5 + np.array([0.0, 1.0, -1.0])
    nrm = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    tn = np.repeat(nrm[:, None, :], 3, axis=1)
    objects.append(SimpleObject(TriangleListMesh(tv, tn), LambertianMaterial((0.6, 0.6, 0.2))))
    objects.append(SimpleObject(TriangleMesh(((-4, 0, -4), (4, 0, -4), (0, 3, -5)),
                                             ((0, 0, 1), (0, 0, 1), (0, 0, 1))), MetalMaterial((0.9, 0.9, 0.9), 0.0)))
    objects.append(SimpleObject(ParallelogramMesh((-1, 3, -1), (2, 0, 0), (0, 0, 2)), LightMaterial((4.0, 4.0, 4.0))))
    objects.append(VolumetricObject(AxisBoxMesh.new_centred((2.0, 1.0, 1.0), (1.5, 1.5, 1.5)),
                                    IsotropicMaterial((0.8, 0.8, 0.8), 0.0), 0.8))
    objects.append(VolumetricObject(SphereMesh((-2.0, 1.0, 1.5), 0.8), IsotropicMaterial((0.9, 0.7, 0.7), 0.1), 1.5))
    objects.append(SimpleObject(SphereMesh((0.0, -1000.0, 0.0), 1000.0), LambertianMaterial((0.5, 0.5, 0.5))))
    objects.append(SimpleObject(RaymarchedIsosurfaceMesh(SdfSminSpheres(np.array([[1.0, 0.5, -2.0], [1.5, 0.8, -2.2]]),
                                                                        np.array([0.4, 0.3]), 0.1), 64, 1e-4),
                                LambertianMaterial((0.3, 0.3, 0.8))))
    pos = np.array([0.0, 2.0, 7.0])
    cam = Camera(pos=tuple(pos), fwd=tuple(-(pos - np.array([0.0, 0.8, 0.0])) / np.linalg.norm(pos - [0, 0.8, 0])),
                 v_fov=50.0, focus_dist=7.0, defocus_angle=0.5)
    opts = RenderOpts(width=width, height=height, samples=samples, ray_depth=ray_depth, seed=seed)
    return Scene(objects, SimpleSkybox), cam, opts


# ------------------------------------------------------------------------------------------ inverse rendering


def inverse_spheres(seed: int = 4, width: int = 256, height: int = 256, samples: int = 16, ray_depth: int = 10):
    """The inverse-rendering scene: the random-spheres layout with 512 spheres (509 jittered r=0.2 spheres
    from a 24x24 grid in the textbook's material mix, plus the three r=1 spheres), a spherical isotropic
    medium (radius 1.2, density 0.5, albedo 0.8: the SDF workload's medium; a sphere rather than a box because
    the engine's box media never scatter), 8 parallelogram lights (2x2, a ring of radius 5 at height 6) and
    the r=1000 ground; SimpleSkybox, 256x256, 16 spp, depth 10.

    Returns (scene, camera, opts, hidden): the colours of materials [0, 521) are the hidden parameters,
    in this order: the 512 sphere albedos, the medium albedo, the 8 light emissions.  Their true values are
    the scene's colours; the ground (material 521) is known."""
    rng = np.random.Generator(np.random.PCG64(seed))
    spheres = []
    big = np.array([4.0, 0.2, 0.0])
    for a in range(-12, 12):
        for b in range(-12, 12):
            choose = rng.random()
            centre = np.array([a + 0.9 * rng.random(), 0.2, b + 0.9 * rng.random()])
            if np.linalg.norm(centre - big) <= 0.9:
                continue
            if choose < 0.8:
                mat = LambertianMaterial(tuple(float(x) for x in rng.random(3) * rng.random(3)))
            elif choose < 0.95:
                mat = MetalMaterial(tuple(float(x) for x in 0.5 + 0.5 * rng.random(3)), float(0.5 * rng.random()))
            else:
                mat = DielectricMaterial(tuple(float(x) for x in 0.5 + 0.5 * rng.random(3)), 1.5, 0.0)
            spheres.append(SimpleObject(SphereMesh(tuple(centre), 0.2), mat))
    spheres = spheres[:509]
    spheres.append(SimpleObject(SphereMesh((0.0, 1.0, 0.0), 1.0), DielectricMaterial((0.9, 0.95, 1.0), 1.5, 0.0)))
    spheres.append(SimpleObject(SphereMesh((-4.0, 1.0, 0.0), 1.0), LambertianMaterial((0.4, 0.2, 0.1))))
    spheres.append(SimpleObject(SphereMesh((4.0, 1.0, 0.0), 1.0), MetalMaterial((0.7, 0.6, 0.5), 0.0)))
    assert len(spheres) == 512
    objects = list(spheres)
    objects.append(VolumetricObject(SphereMesh((-2.0, 1.2, 2.0), 1.2), IsotropicMaterial((0.8, 0.8, 0.8), 0.0), 0.5))
    for i in range(8):
        ang = 2.0 * math.pi * i / 8
        c = np.array([5.0 * math.cos(ang), 6.0, 5.0 * math.sin(ang)])
        emission = tuple(float(x) for x in 1.0 + 2.0 * rng.random(3))
        objects.append(SimpleObject(ParallelogramMesh(tuple(c - [1.0, 0.0, 1.0]), (2.0, 0.0, 0.0), (0.0, 0.0, 2.0)),
                                    LightMaterial(emission)))
    objects.append(SimpleObject(SphereMesh((0.0, -1000.0, 0.0), 1000.0), LambertianMaterial((0.5, 0.5, 0.5))))
    fwd = np.array([-13.0, -2.0, -3.0])
    cam = Camera(pos=(13.0, 2.0, 3.0), fwd=tuple(fwd / np.linalg.norm(fwd)), v_fov=20.0, focus_dist=10.0,
                 defocus_angle=0.6)
    opts = RenderOpts(width=width, height=height, samples=samples, ray_depth=ray_depth, seed=seed)
    return Scene(objects, SimpleSkybox), cam, opts, np.arange(521)


# ------------------------------------------------------------------------------------------ batched requests


def batch_requests(seed: int = 3, n: int = 64, mesh_subdiv: int = 8, max_w: int = 1920, max_h: int = 1080,
                   min_side: int = 128, max_spp: int = 256):
    """The batched-rendering workload: scenes [random_spheres, triangle_mesh, sdf_volume] and `n` requests.
    Request i draws t ~ U[0, 1] and renders min_side * (max_w / min_side)^t by min_side * (max_h / min_side)^t
    pixels (log-uniform along the diagonal from min_side^2 to max_w x max_h), spp ~ U{1..max_spp}, a scene ~
    U{0, 1, 2} with that scene's camera and ray depth, and seed 100 + i.  Returns (scenes, requests)."""
    from .batch import Request

    c0 = random_spheres()
    c1 = triangle_mesh(subdiv=mesh_subdiv)
    c2 = sdf_volume()
    presets_ = [c0, c1, c2]
    rng = np.random.Generator(np.random.PCG64(seed))
    reqs = []
    for i in range(n):
        t = rng.random()
        w = int(round(min_side * (max_w / min_side) ** t))
        h = int(round(min_side * (max_h / min_side) ** t))
        spp = int(rng.integers(1, max_spp + 1))
        k = int(rng.integers(0, 3))
        _, cam, opts = presets_[k]
        reqs.append(Request(k, cam, RenderOpts(width=w, height=h, samples=spp, ray_depth=opts.ray_depth,
                                               seed=100 + i)))
    return [c0[0], c1[0], c2[0]], reqs
