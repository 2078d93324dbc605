// rayna/engine.hpp — rayna_engine's scene and Renderer API in C++17, over the C ABI of rayna_gpu.h.
//
// The names, argument meanings and error behaviour follow rayna_engine (Rust):
//   scene      StandardScene { objects, skybox }                      scene/mod.rs:18-24
//   objects    SimpleObject::new_uncorrected(mesh, material)          object/simple.rs:175-200
//              VolumetricObject::new(mesh, material, density)         object/volumetric.rs:42-72
//   meshes     SphereMesh, Triangle, MeshList (Vec<Triangle>), ParallelogramMesh, AxisBoxMesh,
//              RaymarchedIsosurfaceMesh                               mesh/mod.rs:68-85
//   materials  LambertianMaterial, MetalMaterial, DielectricMaterial, IsotropicMaterial, LightMaterial
//                                                                     material/mod.rs:168-175
//   skyboxes   NoSkybox, SimpleSkybox, WhiteSkybox                     skybox/mod.rs:28-34
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
    static Angle from_degrees(Number d) { return {d * 3.14159265358979323846 / 180.0}; }
};

// ---------------------------------------------------------------------------------------------- materials
struct LambertianMaterial { ColourRgb albedo; };
struct MetalMaterial { ColourRgb albedo; Number fuzz = 0; };
struct DielectricMaterial { ColourRgb albedo = ColourRgb::WHITE(); Number refractive_index = 1.5; Number density = 0; };
struct IsotropicMaterial { ColourRgb albedo; Number density = 0; };
struct LightMaterial { ColourRgb emissive; };
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
using MeshInstance =
    std::variant<SphereMesh, Triangle, MeshList, ParallelogramMesh, AxisBoxMesh, RaymarchedIsosurfaceMesh>;

// ---------------------------------------------------------------------------------------------- objects
struct SimpleObject {
    MeshInstance mesh;
    MaterialInstance material;
    static SimpleObject new_uncorrected(MeshInstance mesh, MaterialInstance material) {
        return SimpleObject{std::move(mesh), std::move(material)};
    }
};
struct VolumetricObject {
    MeshInstance mesh;
    MaterialInstance material;
    Number density = 1;
};
using ObjectInstance = std::variant<SimpleObject, VolumetricObject>;

enum class SkyboxInstance { NoSkybox = RAYNA_SKY_NONE, SimpleSkybox = RAYNA_SKY_SIMPLE, WhiteSkybox = RAYNA_SKY_WHITE };

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
struct Image {
    std::size_t width = 0, height = 0;
    std::vector<ColourRgb> pixels;  // x-major: pixel (x, y) at x * height + y
    const ColourRgb &operator()(std::size_t x, std::size_t y) const { return pixels[x * height + y]; }
    auto begin() const { return pixels.begin(); }
    auto end() const { return pixels.end(); }
};

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
    std::vector<double> spheres, tri_v, tri_n, paras, boxes;
    std::vector<rayna_sdf> sdfs;
    rayna_scene_desc desc{};

    void put3(std::vector<double> &out, const Vector3 &v) {
        out.push_back(v.x);
        out.push_back(v.y);
        out.push_back(v.z);
    }

    rayna_material material(const MaterialInstance &m) {
        rayna_material r{};
        auto colour = [&](const ColourRgb &c) {
            r.colour[0] = c.r;
            r.colour[1] = c.g;
            r.colour[2] = c.b;
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
                if constexpr (std::is_same_v<T, SimpleObject>) objects.push_back({RAYNA_OBJ_SIMPLE, mi, mat, 0, 0.0});
                else objects.push_back({RAYNA_OBJ_VOLUMETRIC, mi, mat, 0, x.density});
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
        desc.skybox = (int32_t)sc.skybox;
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
