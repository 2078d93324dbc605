// rayna/engine.hpp — rayna_engine's scene and Renderer API in C++17, over the C ABI of rayna_gpu.h.
//
// The names, argument meanings and error behaviour follow rayna_engine (Rust):
//   scene      StandardScene { objects, skybox }                      scene/mod.rs:18-24
//   objects    SimpleObject::new_uncorrected(mesh, material, transform) object/simple.rs:67-100
//              SimpleObject::make(mesh, material, transform) -- the engine's SimpleObject::new (`new` is a C++
//              keyword): the transform turned about the mesh's centre (ObjectTransform::new_corrected)
//              VolumetricObject::make / new_uncorrected(mesh, material, density, transform)
//                                                                     object/volumetric.rs:42-72
//   transforms Transform3 (glamour::Transform3<f64>): from_translation, from_scale, from_axis_angle,
//              from_scale_rotation_translation, then, then_translate  core/types.rs:18, object/transform.rs
//   meshes     SphereMesh, Triangle, MeshList (Vec<Triangle>), ParallelogramMesh, AxisBoxMesh,
//              RaymarchedIsosurfaceMesh, BvhMesh (of primitive meshes)  mesh/mod.rs:68-85
//   materials  LambertianMaterial, MetalMaterial, DielectricMaterial, IsotropicMaterial, LightMaterial
//                                                                     material/mod.rs:168-175
//   textures   SolidTexture, WorldCheckerTexture, UvCheckerTexture, ImageTexture, UvNoiseTexture,
//              WorldNoiseTexture, LocalNoiseTexture (over ScalePoint(Perlin), an improved-Perlin restatement)
//                                                                     texture/mod.rs:26-37
//   skyboxes   NoSkybox, SimpleSkybox, WhiteSkybox, HdrImageSkybox(image)
//                                                                     skybox/mod.rs:28-34
//   camera     Camera { pos, fwd, v_fov, focus_dist, defocus_angle } scene/camera.rs:10-38
//   options    RenderOpts { width, height, samples, mode, ray_depth, ray_branching }
//                                                                     render/render_opts.rs:8-29
//   renderer   Renderer::new_from(scene, camera, options, num_threads) -> Renderer or RendererCreateError,
//              render() -> Render { img, stats }, set_scene / set_camera / set_options (each clears
//              accumulation), clear_accumulation, set_thread_count, scene() / camera() / options()
//                                                                     render/renderer.rs:36-264
//   image      Image of ColourRgb, shape (width, height), indexed [x][y] core/image.rs:17-26
//
// The renderer runs on the GPU (librayna_amd.so); num_threads has no meaning there and is kept for the
// signature.  RenderOpts gains `seed`, the key of the counter RNG that replaces the engine's entropy-seeded
// generators.  Link with -lrayna_amd.
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <variant>
#include <vector>

#include "../rayna_gpu.h"

namespace rayna {

// ---------------------------------------------------------------------------------------------- core types
using Number = double;   // core/types.rs:7
using Channel = float;   // core/types.rs:3

struct Vector3 {
    Number x = 0, y = 0, z = 0;
    static constexpr Vector3 ZERO() { return {0, 0, 0}; }
};
using Point3 = Vector3;

struct ColourRgb {
    Channel r = 0, g = 0, b = 0;
    static constexpr ColourRgb WHITE() { return {1, 1, 1}; }
    static constexpr ColourRgb BLACK() { return {0, 0, 0}; }
};

struct Angle {
    Number radians = 0;
    static Angle from_degrees(Number d) { return {d * (3.14159265358979323846 / 180.0)}; }
};

// An affine map, object space to world space: row-major 3x4 [m | t], p -> m p + t.
struct Transform3 {
    Number m[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    static Transform3 IDENTITY() { return {}; }
    static Transform3 from_translation(const Vector3 &v) {
        Transform3 r;
        r.m[3] = v.x;
        r.m[7] = v.y;
        r.m[11] = v.z;
        return r;
    }
    static Transform3 from_scale(const Vector3 &s) {
        Transform3 r;
        r.m[0] = s.x;
        r.m[5] = s.y;
        r.m[10] = s.z;
        return r;
    }
    // A rotation about the unit vector `axis` (Rodrigues' formula, arranged as glam's Mat4::from_axis_angle).
    static Transform3 from_axis_angle(const Vector3 &axis, Angle angle) {
        const Number s = std::sin(angle.radians), c = std::cos(angle.radians), omc = 1.0 - c;
        const Number x = axis.x, y = axis.y, z = axis.z;
        const Number xy = x * y * omc, xz = x * z * omc, yz = y * z * omc;
        Transform3 r;
        const Number rows[3][3] = {{x * x * omc + c, xy - z * s, xz + y * s},
                                   {xy + z * s, y * y * omc + c, yz - x * s},
                                   {xz - y * s, yz + x * s, z * z * omc + c}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.m[4 * i + j] = rows[i][j];
        return r;
    }
    static Transform3 from_scale_rotation_translation(const Vector3 &scale, const Vector3 &axis, Angle angle,
                                                      const Vector3 &translation) {
        return from_scale(scale).then(from_axis_angle(axis, angle)).then_translate(translation);
    }
    // self, then `next`: next.m * self.m
    Transform3 then(const Transform3 &next) const {
        Transform3 r;
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 4; ++j) {
                Number v = next.m[4 * i] * m[j] + next.m[4 * i + 1] * m[4 + j] + next.m[4 * i + 2] * m[8 + j];
                if (j == 3) v += next.m[4 * i + 3];
                r.m[4 * i + j] = v;
            }
        }
        return r;
    }
    Transform3 then_translate(const Vector3 &v) const { return then(from_translation(v)); }
    Point3 map_point(const Point3 &p) const {
        return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
                m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
    }
    // ObjectTransform::new_corrected (object/transform.rs): the map applied about `centre`.
    Transform3 corrected(const Point3 &centre) const {
        return from_translation({-centre.x, -centre.y, -centre.z}).then(*this).then_translate(centre);
    }
};

// ---------------------------------------------------------------------------------------------- images
// core/image.rs: pixel (x, y) at x * height + y (x-major), as the engine indexes Image[(x, y)].
struct Image {
    std::size_t width = 0, height = 0;
    std::vector<ColourRgb> pixels;  // x-major: pixel (x, y) at x * height + y
    const ColourRgb &operator()(std::size_t x, std::size_t y) const { return pixels[x * height + y]; }
    auto begin() const { return pixels.begin(); }
    auto end() const { return pixels.end(); }
};

// ---------------------------------------------------------------------------------------------- textures
struct TextureInstance;
using TexturePtr = std::shared_ptr<const TextureInstance>;
struct Vector2 { Number x = 0, y = 0; };
struct SolidTexture { ColourRgb albedo; };
struct WorldCheckerTexture { Vector3 offset; TexturePtr even, odd; Number scale = 1; };
struct UvCheckerTexture { Vector2 offset; TexturePtr even, odd; Number scale = 1; };
struct ImageTexture { std::shared_ptr<const Image> image; Vector2 scale{1, 1}; Vector2 offset; };
// An improved Perlin noise whose permutation is drawn from `seed` (not the noise crate's permutation).
struct Perlin { std::uint32_t seed = 0; };
struct ScalePoint { Perlin source; Number scale = 1; };
// ColourSource::Greyscale (one noise) or ::Rgb (three noises sharing one scale).
struct ColourSource {
    std::vector<ScalePoint> noises;
    static ColourSource Greyscale(ScalePoint n) { return {{n}}; }
    static ColourSource Rgb(ScalePoint r, ScalePoint g, ScalePoint b) { return {{r, g, b}}; }
};
struct UvNoiseTexture { ColourSource source; };
struct WorldNoiseTexture { ColourSource source; };
struct LocalNoiseTexture { ColourSource source; };
struct TextureInstance : std::variant<SolidTexture, WorldCheckerTexture, UvCheckerTexture, ImageTexture,
                                      UvNoiseTexture, WorldNoiseTexture, LocalNoiseTexture> {
    using variant::variant;
    TextureInstance() : variant(SolidTexture{}) {}
    TextureInstance(ColourRgb c) : variant(SolidTexture{c}) {}  // impl From<Colour> for TextureInstance
    TextureInstance(Channel r, Channel g, Channel b) : variant(SolidTexture{{r, g, b}}) {}
};
template <class T>
inline TexturePtr texture(T t) { return std::make_shared<const TextureInstance>(std::move(t)); }

// ---------------------------------------------------------------------------------------------- materials
struct LambertianMaterial { TextureInstance albedo; };
struct MetalMaterial { TextureInstance albedo; Number fuzz = 0; };
struct DielectricMaterial { TextureInstance albedo = ColourRgb::WHITE(); Number refractive_index = 1.5; Number density = 0; };
struct IsotropicMaterial { TextureInstance albedo; Number density = 0; };
struct LightMaterial { TextureInstance emissive; };
using MaterialInstance =
    std::variant<LambertianMaterial, MetalMaterial, DielectricMaterial, IsotropicMaterial, LightMaterial>;

// ---------------------------------------------------------------------------------------------- meshes
struct SphereMesh { Point3 pos; Number radius = 1; };
struct Triangle { Point3 vertices[3]; Vector3 normals[3]; };
struct MeshList { std::vector<Triangle> triangles; };  // MeshInstance::from(Vec<Triangle>)
struct ParallelogramMesh { Point3 p; Vector3 u, v; };  // ParallelogramMesh::new(Planar::new(p, u, v))
struct AxisBoxMesh { Point3 a, b; };
struct SdfSphere { Point3 centre; Number radius = 1; };
struct SdfSminSpheres { std::vector<Point3> centres; std::vector<Number> radii; Number k = 0.1; };
struct SdfMandelbulb { Point3 centre; int power = 8; int iterations = 10; Number bailout = 2; };
using SdfFunction = std::variant<SdfSphere, SdfSminSpheres, SdfMandelbulb>;
struct RaymarchedIsosurfaceMesh { SdfFunction sdf; int max_iterations = 150; Number epsilon = 1e-7; };
using PrimitiveMesh = std::variant<SphereMesh, Triangle, ParallelogramMesh, AxisBoxMesh>;
struct BvhMesh { std::vector<PrimitiveMesh> meshes; };  // BvhMesh::new(meshes), mesh/advanced/bvh.rs:28-45
using MeshInstance =
    std::variant<SphereMesh, Triangle, MeshList, ParallelogramMesh, AxisBoxMesh, RaymarchedIsosurfaceMesh, BvhMesh>;

// Mesh::centre (mesh/mod.rs:93): the point the engine's corrected transforms turn about.
inline Point3 mesh_centre(const MeshInstance &m) {
    auto prim = [](const auto &x) -> Point3 {
        using T = std::decay_t<decltype(x)>;
        if constexpr (std::is_same_v<T, SphereMesh>) return x.pos;
        if constexpr (std::is_same_v<T, Triangle>)  // triangle.rs:42-45
            return {(x.vertices[0].x + x.vertices[1].x + x.vertices[2].x) / 3.0,
                    (x.vertices[0].y + x.vertices[1].y + x.vertices[2].y) / 3.0,
                    (x.vertices[0].z + x.vertices[1].z + x.vertices[2].z) / 3.0};
        if constexpr (std::is_same_v<T, ParallelogramMesh>) return x.p;
        if constexpr (std::is_same_v<T, AxisBoxMesh>)
            return {(std::fmin(x.a.x, x.b.x) + std::fmax(x.a.x, x.b.x)) / 2.0,
                    (std::fmin(x.a.y, x.b.y) + std::fmax(x.a.y, x.b.y)) / 2.0,
                    (std::fmin(x.a.z, x.b.z) + std::fmax(x.a.z, x.b.z)) / 2.0};
        return Point3{};
    };
    return std::visit([&](const auto &x) -> Point3 {
        using T = std::decay_t<decltype(x)>;
        if constexpr (std::is_same_v<T, MeshList>) {  // the sum of the triangles' centres (list.rs:33-36)
            Point3 c{};
            for (const Triangle &t : x.triangles) {
                const Point3 tc = prim(t);
                c = {c.x + tc.x, c.y + tc.y, c.z + tc.z};
            }
            return c;
        } else if constexpr (std::is_same_v<T, BvhMesh>) {  // the mean of the meshes' centres (bvh.rs:36-44)
            Point3 c{};
            for (const PrimitiveMesh &p : x.meshes) {
                const Point3 pc = std::visit(prim, p);
                c = {c.x + pc.x, c.y + pc.y, c.z + pc.z};
            }
            const Number n = (Number)x.meshes.size();
            return {c.x / n, c.y / n, c.z / n};
        } else if constexpr (std::is_same_v<T, RaymarchedIsosurfaceMesh>) {
            return Point3{};  // raymarched.rs:69
        } else {
            return prim(x);
        }
    }, m);
}

// ---------------------------------------------------------------------------------------------- objects
struct SimpleObject {
    MeshInstance mesh;
    MaterialInstance material;
    Transform3 transform;  // object to world, as given to the renderer (already corrected)
    static SimpleObject new_uncorrected(MeshInstance mesh, MaterialInstance material,
                                        Transform3 transform = Transform3::IDENTITY()) {
        return SimpleObject{std::move(mesh), std::move(material), transform};
    }
    // SimpleObject::new: the transform turned about the mesh's centre.
    static SimpleObject make(MeshInstance mesh, MaterialInstance material,
                             Transform3 transform = Transform3::IDENTITY()) {
        const Transform3 t = transform.corrected(mesh_centre(mesh));
        return SimpleObject{std::move(mesh), std::move(material), t};
    }
};
struct VolumetricObject {
    MeshInstance mesh;
    MaterialInstance material;
    Number density = 1;
    Transform3 transform;
    static VolumetricObject new_uncorrected(MeshInstance mesh, MaterialInstance material, Number density,
                                            Transform3 transform = Transform3::IDENTITY()) {
        return VolumetricObject{std::move(mesh), std::move(material), density, transform};
    }
    static VolumetricObject make(MeshInstance mesh, MaterialInstance material, Number density,
                                 Transform3 transform = Transform3::IDENTITY()) {
        const Transform3 t = transform.corrected(mesh_centre(mesh));
        return VolumetricObject{std::move(mesh), std::move(material), density, t};
    }
};
using ObjectInstance = std::variant<SimpleObject, VolumetricObject>;

// SkyboxInstance (skybox/mod.rs:28-34): a kind (RAYNA_SKY_*) and, for HdrImageSkybox, its image.
struct SkyboxInstance {
    int32_t kind = RAYNA_SKY_SIMPLE;
    std::shared_ptr<const Image> image;
    static const SkyboxInstance NoSkybox, SimpleSkybox, WhiteSkybox;
    static SkyboxInstance HdrImageSkybox(std::shared_ptr<const Image> image) { return {RAYNA_SKY_HDRI, image}; }
};
inline const SkyboxInstance SkyboxInstance::NoSkybox{RAYNA_SKY_NONE, nullptr};
inline const SkyboxInstance SkyboxInstance::SimpleSkybox{RAYNA_SKY_SIMPLE, nullptr};
inline const SkyboxInstance SkyboxInstance::WhiteSkybox{RAYNA_SKY_WHITE, nullptr};

struct StandardScene {
    std::vector<ObjectInstance> objects;
    SkyboxInstance skybox = SkyboxInstance::SimpleSkybox;
};

struct Camera {
    Point3 pos;
    Vector3 fwd{0, 0, 1};
    Angle v_fov = Angle::from_degrees(45);
    Number focus_dist = 1;
    Angle defocus_angle;
};

enum class RenderMode {
    PBR = RAYNA_MODE_PBR,
    RayNormal = RAYNA_MODE_RAY_NORMAL,
    OutwardNormal = RAYNA_MODE_OUTWARD_NORMAL,
    Scatter = RAYNA_MODE_SCATTER,
    FrontFace = RAYNA_MODE_FRONT_FACE,
    Distance = RAYNA_MODE_DISTANCE,
    Uv = RAYNA_MODE_UV,
    Side = RAYNA_MODE_SIDE
};

struct RenderOpts {
    std::size_t width = 740, height = 480, samples = 1;
    RenderMode mode = RenderMode::PBR;
    std::size_t ray_depth = 5, ray_branching = 1;
    std::uint64_t seed = 0;
};

// ---------------------------------------------------------------------------------------------- results

struct RenderStats {
    double duration = 0;           // seconds
    std::uint64_t accum_frames = 0;
    std::uint64_t samples = 0;
    std::uint64_t segments = 0;    // closest-hit queries
    double kernel_ms = 0;
};

struct Render {
    Image img;
    RenderStats stats;
};

// The engine's RendererCreateError (renderer.rs:54-62), carrying the library's status and message.
class RendererCreateError : public std::runtime_error {
public:
    RendererCreateError(rayna_status st, const std::string &msg) : std::runtime_error(msg), status(st) {}
    rayna_status status;
};

// ---------------------------------------------------------------------------------------------- flattening
namespace detail {

struct SceneArrays {
    std::vector<rayna_object> objects;
    std::vector<rayna_mesh> meshes;
    std::vector<rayna_material> materials;
    std::vector<double> spheres, tri_v, tri_n, paras, boxes, transforms;
    std::vector<rayna_sdf> sdfs;
    std::vector<rayna_texture> textures;
    std::vector<std::vector<float>> image_data;  // row-major RGB
    std::vector<rayna_image> images;
    std::vector<const Image *> image_keys;
    rayna_scene_desc desc{};

    int32_t image(const Image *im) {
        for (std::size_t i = 0; i < image_keys.size(); ++i)
            if (image_keys[i] == im) return (int32_t)i;
        std::vector<float> rgb(3 * im->width * im->height);
        for (std::size_t y = 0; y < im->height; ++y)
            for (std::size_t x = 0; x < im->width; ++x) {
                const ColourRgb &c = (*im)(x, y);
                float *o = &rgb[3 * (y * im->width + x)];
                o[0] = c.r;
                o[1] = c.g;
                o[2] = c.b;
            }
        image_data.push_back(std::move(rgb));
        image_keys.push_back(im);
        images.push_back({(uint32_t)im->width, (uint32_t)im->height, nullptr});
        return (int32_t)images.size() - 1;
    }

    // A texture as a 1-based index into `textures` (sub-textures first, so checkers name earlier entries).
    int32_t texture_ref(const TextureInstance &t) {
        rayna_texture r{};
        std::visit([&](const auto &x) {
            using T = std::decay_t<decltype(x)>;
            auto noise = [&](const ColourSource &s) {
                r.noise_rgb = s.noises.size() == 3 ? 1 : 0;
                for (std::size_t k = 0; k < s.noises.size() && k < 3; ++k) r.noise_seed[k] = s.noises[k].source.seed;
                r.scale = s.noises.empty() ? 1.0 : s.noises[0].scale;
            };
            if constexpr (std::is_same_v<T, SolidTexture>) {
                r.kind = RAYNA_TEX_SOLID;
                r.colour[0] = x.albedo.r;
                r.colour[1] = x.albedo.g;
                r.colour[2] = x.albedo.b;
            }
            if constexpr (std::is_same_v<T, WorldCheckerTexture> || std::is_same_v<T, UvCheckerTexture>) {
                r.kind = std::is_same_v<T, WorldCheckerTexture> ? RAYNA_TEX_WORLD_CHECKER : RAYNA_TEX_UV_CHECKER;
                r.odd = texture_ref(x.odd ? *x.odd : TextureInstance{});
                r.even = texture_ref(x.even ? *x.even : TextureInstance{});
                r.scale = x.scale;
                r.offset[0] = x.offset.x;
                r.offset[1] = x.offset.y;
                if constexpr (std::is_same_v<T, WorldCheckerTexture>) r.offset[2] = x.offset.z;
            }
            if constexpr (std::is_same_v<T, ImageTexture>) {
                r.kind = RAYNA_TEX_IMAGE;
                r.image = image(x.image.get());
                r.image_scale[0] = x.scale.x;
                r.image_scale[1] = x.scale.y;
                r.offset[0] = x.offset.x;
                r.offset[1] = x.offset.y;
            }
            if constexpr (std::is_same_v<T, UvNoiseTexture>) { r.kind = RAYNA_TEX_UV_NOISE; noise(x.source); }
            if constexpr (std::is_same_v<T, WorldNoiseTexture>) { r.kind = RAYNA_TEX_WORLD_NOISE; noise(x.source); }
            if constexpr (std::is_same_v<T, LocalNoiseTexture>) { r.kind = RAYNA_TEX_LOCAL_NOISE; noise(x.source); }
        }, static_cast<const TextureInstance::variant &>(t));
        textures.push_back(r);
        return (int32_t)textures.size();
    }

    void put3(std::vector<double> &out, const Vector3 &v) {
        out.push_back(v.x);
        out.push_back(v.y);
        out.push_back(v.z);
    }

    rayna_material material(const MaterialInstance &m) {
        rayna_material r{};
        auto colour = [&](const TextureInstance &t) {  // a solid colour in place, any other texture by index
            if (const SolidTexture *s = std::get_if<SolidTexture>(&t)) {
                r.colour[0] = s->albedo.r;
                r.colour[1] = s->albedo.g;
                r.colour[2] = s->albedo.b;
            } else {
                r.texture = texture_ref(t);
            }
        };
        std::visit([&](const auto &x) {
            using T = std::decay_t<decltype(x)>;
            if constexpr (std::is_same_v<T, LambertianMaterial>) { r.kind = RAYNA_MAT_LAMBERTIAN; colour(x.albedo); }
            if constexpr (std::is_same_v<T, MetalMaterial>) { r.kind = RAYNA_MAT_METAL; colour(x.albedo); r.fuzz = x.fuzz; }
            if constexpr (std::is_same_v<T, DielectricMaterial>) {
                r.kind = RAYNA_MAT_DIELECTRIC;
                colour(x.albedo);
                r.refractive_index = x.refractive_index;
                r.density = x.density;
            }
            if constexpr (std::is_same_v<T, IsotropicMaterial>) { r.kind = RAYNA_MAT_ISOTROPIC; colour(x.albedo); r.density = x.density; }
            if constexpr (std::is_same_v<T, LightMaterial>) { r.kind = RAYNA_MAT_LIGHT; colour(x.emissive); }
        }, m);
        return r;
    }

    int32_t mesh(const MeshInstance &m) {
        rayna_mesh r{};
        std::visit([&](const auto &x) {
            using T = std::decay_t<decltype(x)>;
            if constexpr (std::is_same_v<T, SphereMesh>) {
                r = {RAYNA_MESH_SPHERE, (int32_t)(spheres.size() / 4), 1, 0};
                put3(spheres, x.pos);
                spheres.push_back(x.radius);
            }
            if constexpr (std::is_same_v<T, Triangle>) {
                r = {RAYNA_MESH_TRIANGLE, (int32_t)(tri_v.size() / 9), 1, 0};
                for (int i = 0; i < 3; ++i) { put3(tri_v, x.vertices[i]); put3(tri_n, x.normals[i]); }
            }
            if constexpr (std::is_same_v<T, MeshList>) {
                r = {RAYNA_MESH_TRIANGLE_LIST, (int32_t)(tri_v.size() / 9), (int32_t)x.triangles.size(), 0};
                for (const Triangle &t : x.triangles)
                    for (int i = 0; i < 3; ++i) { put3(tri_v, t.vertices[i]); put3(tri_n, t.normals[i]); }
            }
            if constexpr (std::is_same_v<T, ParallelogramMesh>) {
                r = {RAYNA_MESH_PARALLELOGRAM, (int32_t)(paras.size() / 9), 1, 0};
                put3(paras, x.p);
                put3(paras, x.u);
                put3(paras, x.v);
            }
            if constexpr (std::is_same_v<T, AxisBoxMesh>) {
                r = {RAYNA_MESH_AXIS_BOX, (int32_t)(boxes.size() / 6), 1, 0};
                put3(boxes, x.a);
                put3(boxes, x.b);
            }
            if constexpr (std::is_same_v<T, BvhMesh>) {  // the members take the mesh slots just before it
                const int32_t first = (int32_t)meshes.size();
                for (const PrimitiveMesh &p : x.meshes) std::visit([&](const auto &q) { mesh(MeshInstance{q}); }, p);
                r = {RAYNA_MESH_BVH, first, (int32_t)x.meshes.size(), 0};
            }
            if constexpr (std::is_same_v<T, RaymarchedIsosurfaceMesh>) {
                rayna_sdf s{};
                s.max_iterations = x.max_iterations;
                s.epsilon = x.epsilon;
                std::visit([&](const auto &f) {
                    using F = std::decay_t<decltype(f)>;
                    if constexpr (std::is_same_v<F, SdfSphere>) {
                        s.kind = RAYNA_SDF_SPHERE;
                        s.first_sphere = (int32_t)(spheres.size() / 4);
                        s.n_spheres = 1;
                        put3(spheres, f.centre);
                        spheres.push_back(f.radius);
                    }
                    if constexpr (std::is_same_v<F, SdfSminSpheres>) {
                        s.kind = RAYNA_SDF_SMIN_SPHERES;
                        s.first_sphere = (int32_t)(spheres.size() / 4);
                        s.n_spheres = (int32_t)f.centres.size();
                        s.k = f.k;
                        for (std::size_t i = 0; i < f.centres.size(); ++i) {
                            put3(spheres, f.centres[i]);
                            spheres.push_back(i < f.radii.size() ? f.radii[i] : 0.0);
                        }
                    }
                    if constexpr (std::is_same_v<F, SdfMandelbulb>) {
                        s.kind = RAYNA_SDF_MANDELBULB;
                        s.power = f.power;
                        s.bulb_iterations = f.iterations;
                        s.bailout = f.bailout;
                        s.centre[0] = f.centre.x;
                        s.centre[1] = f.centre.y;
                        s.centre[2] = f.centre.z;
                    }
                }, x.sdf);
                r = {RAYNA_MESH_SDF, (int32_t)sdfs.size(), 1, 0};
                sdfs.push_back(s);
            }
        }, m);
        meshes.push_back(r);
        return (int32_t)meshes.size() - 1;
    }

    explicit SceneArrays(const StandardScene &sc) {
        for (const ObjectInstance &o : sc.objects) {
            std::visit([&](const auto &x) {
                using T = std::decay_t<decltype(x)>;
                const int32_t mi = mesh(x.mesh);
                materials.push_back(material(x.material));
                const int32_t mat = (int32_t)materials.size() - 1;
                int32_t ti = 0;
                bool identity = true;
                const Transform3 I = Transform3::IDENTITY();
                for (int k = 0; k < 12; ++k) identity = identity && x.transform.m[k] == I.m[k];
                if (!identity) {
                    transforms.insert(transforms.end(), x.transform.m, x.transform.m + 12);
                    ti = (int32_t)(transforms.size() / 12);
                }
                if constexpr (std::is_same_v<T, SimpleObject>) objects.push_back({RAYNA_OBJ_SIMPLE, mi, mat, ti, 0.0});
                else objects.push_back({RAYNA_OBJ_VOLUMETRIC, mi, mat, ti, x.density});
            }, o);
        }
        desc.objects = objects.data();
        desc.n_objects = (int32_t)objects.size();
        desc.meshes = meshes.data();
        desc.n_meshes = (int32_t)meshes.size();
        desc.materials = materials.data();
        desc.n_materials = (int32_t)materials.size();
        desc.spheres = spheres.data();
        desc.n_spheres = (int32_t)(spheres.size() / 4);
        desc.tri_vertices = tri_v.data();
        desc.tri_normals = tri_n.data();
        desc.n_triangles = (int32_t)(tri_v.size() / 9);
        desc.parallelograms = paras.data();
        desc.n_parallelograms = (int32_t)(paras.size() / 9);
        desc.boxes = boxes.data();
        desc.n_boxes = (int32_t)(boxes.size() / 6);
        desc.sdfs = sdfs.data();
        desc.n_sdfs = (int32_t)sdfs.size();
        desc.skybox = sc.skybox.kind;
        desc.transforms = transforms.data();
        desc.n_transforms = (int32_t)(transforms.size() / 12);
        if (sc.skybox.kind == RAYNA_SKY_HDRI && sc.skybox.image) desc.sky_image = image(sc.skybox.image.get());
        for (std::size_t i = 0; i < images.size(); ++i) images[i].rgb = image_data[i].data();
        desc.textures = textures.data();
        desc.n_textures = (int32_t)textures.size();
        desc.images = images.data();
        desc.n_images = (int32_t)images.size();
    }
};

inline rayna_camera camera_desc(const Camera &c) {
    return rayna_camera{{c.pos.x, c.pos.y, c.pos.z}, {c.fwd.x, c.fwd.y, c.fwd.z}, c.v_fov.radians, c.focus_dist,
                        c.defocus_angle.radians};
}

inline rayna_render_opts opts_desc(const RenderOpts &o) {
    return rayna_render_opts{(uint32_t)o.width, (uint32_t)o.height, (uint32_t)o.samples, (uint32_t)o.ray_depth,
                             (uint32_t)o.ray_branching, (int32_t)o.mode, o.seed};
}

}  // namespace detail

// ---------------------------------------------------------------------------------------------- renderer
class Renderer {
public:
    // Renderer::new_from (renderer.rs:86-107).  Throws RendererCreateError.
    static Renderer new_from(StandardScene scene, Camera camera, RenderOpts options, std::size_t num_threads) {
        Renderer r;
        r.scene_ = std::move(scene);
        r.camera_ = camera;
        r.options_ = options;
        r.threads_ = num_threads;
        detail::SceneArrays arrays(r.scene_);
        const rayna_camera c = detail::camera_desc(camera);
        const rayna_render_opts o = detail::opts_desc(options);
        const rayna_status st = rayna_create(&arrays.desc, &c, &o, &r.h_);
        if (st != RAYNA_OK) throw RendererCreateError(st, rayna_last_error(nullptr));
        return r;
    }

    Renderer(Renderer &&o) noexcept { *this = std::move(o); }
    Renderer &operator=(Renderer &&o) noexcept {
        if (this != &o) {
            reset();
            h_ = std::exchange(o.h_, nullptr);
            scene_ = std::move(o.scene_);
            camera_ = o.camera_;
            options_ = o.options_;
            threads_ = o.threads_;
        }
        return *this;
    }
    Renderer(const Renderer &) = delete;
    Renderer &operator=(const Renderer &) = delete;
    ~Renderer() { reset(); }

    // Renderer::render (renderer.rs:224-264): never fails for an invalid camera (the engine's checkerboard);
    // a device failure throws std::runtime_error.
    Render render() {
        Render out;
        std::vector<float> rgb((std::size_t)options_.width * options_.height * 3);
        uint32_t dims[2] = {0, 0};
        rayna_render_stats st{};
        check(rayna_render(h_, rgb.data(), RAYNA_LAYOUT_XMAJOR, dims, &st));
        out.img.width = dims[0];
        out.img.height = dims[1];
        out.img.pixels.resize((std::size_t)dims[0] * dims[1]);
        for (std::size_t i = 0; i < out.img.pixels.size(); ++i)
            out.img.pixels[i] = ColourRgb{rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]};
        out.stats = RenderStats{st.duration_s, st.accum_frames, st.samples, st.segments, st.kernel_ms};
        return out;
    }

    // render() then the UI's 8-bit conversion (rayna_ui/src/ext/img_ext.rs:15-45): RGBA, row-major, gamma 1/2.2
    std::vector<std::uint8_t> render_srgb8(std::size_t *width = nullptr, std::size_t *height = nullptr) {
        std::vector<std::uint8_t> rgba((std::size_t)options_.width * options_.height * 4);
        uint32_t dims[2] = {0, 0};
        check(rayna_render_srgb8(h_, rgba.data(), dims, nullptr));
        rgba.resize((std::size_t)dims[0] * dims[1] * 4);
        if (width) *width = dims[0];
        if (height) *height = dims[1];
        return rgba;
    }

    // set_scene / set_camera / set_options (renderer.rs:158-176): each clears accumulation
    void set_scene(StandardScene scene) {
        detail::SceneArrays arrays(scene);
        check(rayna_set_scene(h_, &arrays.desc));
        scene_ = std::move(scene);
    }
    void set_camera(Camera camera) {
        const rayna_camera c = detail::camera_desc(camera);
        check(rayna_set_camera(h_, &c));
        camera_ = camera;
    }
    void set_options(RenderOpts options) {
        const rayna_render_opts o = detail::opts_desc(options);
        check(rayna_set_options(h_, &o));
        options_ = options;
    }
    void clear_accumulation() { check(rayna_clear_accumulation(h_)); }
    // set_thread_count (renderer.rs:179-182): the GPU renderer has no host pool; kept for the signature
    void set_thread_count(std::size_t num_threads) { threads_ = num_threads; }

    const StandardScene &scene() const { return scene_; }
    const Camera &camera() const { return camera_; }
    const RenderOpts &options() const { return options_; }
    rayna_renderer *handle() const { return h_; }

private:
    Renderer() = default;
    void reset() {
        if (h_) rayna_destroy(h_);
        h_ = nullptr;
    }
    void check(rayna_status st) const {
        if (st != RAYNA_OK) throw std::runtime_error(rayna_last_error(h_));
    }

    rayna_renderer *h_ = nullptr;
    StandardScene scene_;
    Camera camera_;
    RenderOpts options_;
    std::size_t threads_ = 0;
};

}  // namespace rayna
