"""ctypes mirrors of the C structs declared in include/rayna_gpu.h (kept field-for-field identical)."""
import ctypes as C

RAYNA_OK = 0
RAYNA_INVALID_ARGUMENT = 1
RAYNA_UNSUPPORTED = 2
RAYNA_DEVICE_ERROR = 3
RAYNA_OUT_OF_MEMORY = 4
RAYNA_RENDERER_LOST = 5
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "UNSUPPORTED", 3: "DEVICE_ERROR", 4: "OUT_OF_MEMORY",
                5: "RENDERER_LOST"}

MAT_LAMBERTIAN, MAT_METAL, MAT_DIELECTRIC, MAT_ISOTROPIC, MAT_LIGHT = 0, 1, 2, 3, 4
MESH_SPHERE, MESH_TRIANGLE, MESH_TRIANGLE_LIST, MESH_PARALLELOGRAM, MESH_AXIS_BOX, MESH_SDF, MESH_BVH = 0, 1, 2, 3, 4, 5, 6
SDF_SPHERE, SDF_SMIN_SPHERES, SDF_MANDELBULB = 0, 1, 2
OBJ_SIMPLE, OBJ_VOLUMETRIC = 0, 1
SKY_NONE, SKY_SIMPLE, SKY_WHITE, SKY_HDRI = 0, 1, 2, 3
TEX_SOLID, TEX_WORLD_CHECKER, TEX_UV_CHECKER, TEX_IMAGE, TEX_UV_NOISE, TEX_WORLD_NOISE, TEX_LOCAL_NOISE = range(7)
MODE_PBR = 0
MODE_RAY_NORMAL, MODE_OUTWARD_NORMAL, MODE_SCATTER, MODE_FRONT_FACE, MODE_DISTANCE, MODE_UV, MODE_SIDE = range(1, 8)
LAYOUT_XMAJOR, LAYOUT_ROWMAJOR = 0, 1


class Material(C.Structure):
    _fields_ = [("kind", C.c_int32), ("colour", C.c_float * 3), ("texture", C.c_int32), ("fuzz", C.c_double),
                ("refractive_index", C.c_double), ("density", C.c_double)]


class Texture(C.Structure):
    _fields_ = [("kind", C.c_int32), ("colour", C.c_float * 3), ("odd", C.c_int32), ("even", C.c_int32),
                ("image", C.c_int32), ("noise_rgb", C.c_int32), ("noise_seed", C.c_uint32 * 3),
                ("reserved", C.c_int32), ("scale", C.c_double), ("offset", C.c_double * 3),
                ("image_scale", C.c_double * 2)]


class Image(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.POINTER(C.c_float))]


class Mesh(C.Structure):
    _fields_ = [("kind", C.c_int32), ("index", C.c_int32), ("count", C.c_int32), ("reserved", C.c_int32)]


class Sdf(C.Structure):
    _fields_ = [("kind", C.c_int32), ("first_sphere", C.c_int32), ("n_spheres", C.c_int32),
                ("max_iterations", C.c_int32), ("bulb_iterations", C.c_int32), ("reserved", C.c_int32),
                ("epsilon", C.c_double), ("k", C.c_double), ("power", C.c_double), ("bailout", C.c_double),
                ("centre", C.c_double * 3)]


class Object(C.Structure):
    _fields_ = [("kind", C.c_int32), ("mesh", C.c_int32), ("material", C.c_int32), ("transform", C.c_int32),
                ("density", C.c_double)]


class SceneDesc(C.Structure):
    _fields_ = [("objects", C.POINTER(Object)), ("n_objects", C.c_int32),
                ("meshes", C.POINTER(Mesh)), ("n_meshes", C.c_int32),
                ("materials", C.POINTER(Material)), ("n_materials", C.c_int32),
                ("spheres", C.POINTER(C.c_double)), ("n_spheres", C.c_int32),
                ("tri_vertices", C.POINTER(C.c_double)), ("tri_normals", C.POINTER(C.c_double)),
                ("n_triangles", C.c_int32),
                ("parallelograms", C.POINTER(C.c_double)), ("n_parallelograms", C.c_int32),
                ("boxes", C.POINTER(C.c_double)), ("n_boxes", C.c_int32),
                ("sdfs", C.POINTER(Sdf)), ("n_sdfs", C.c_int32),
                ("skybox", C.c_int32),
                ("transforms", C.POINTER(C.c_double)), ("n_transforms", C.c_int32),
                ("textures", C.POINTER(Texture)), ("n_textures", C.c_int32),
                ("images", C.POINTER(Image)), ("n_images", C.c_int32), ("sky_image", C.c_int32)]


class CameraDesc(C.Structure):
    _fields_ = [("pos", C.c_double * 3), ("fwd", C.c_double * 3), ("v_fov", C.c_double),
                ("focus_dist", C.c_double), ("defocus_angle", C.c_double)]


class RenderOptsDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("samples", C.c_uint32),
                ("ray_depth", C.c_uint32), ("ray_branching", C.c_uint32), ("mode", C.c_int32),
                ("seed", C.c_uint64)]


class RenderStats(C.Structure):
    _fields_ = [("duration_s", C.c_double), ("kernel_ms", C.c_double), ("extend_ms", C.c_double),
                ("shade_ms", C.c_double), ("accum_frames", C.c_uint64), ("samples", C.c_uint64),
                ("segments", C.c_uint64), ("extend_launches", C.c_uint64), ("extend_camera_ms", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("chunk_slots", C.c_uint32), ("refill_lanes", C.c_uint32), ("wave_batch", C.c_uint32),
                ("shade_blocks_per_cu", C.c_uint32), ("use_graph", C.c_int32), ("schedule", C.c_int32),
                ("pipelines", C.c_int32), ("packet_depths", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_int32), ("layout", C.c_int32), ("camera", CameraDesc), ("opts", RenderOptsDesc),
                ("out_rgb", C.POINTER(C.c_float))]


class BatchStats(C.Structure):
    _fields_ = [("duration_s", C.c_double), ("kernel_ms", C.c_double), ("samples", C.c_uint64),
                ("segments", C.c_uint64), ("chunks", C.c_uint32), ("launches", C.c_uint32),
                ("request_ms", C.POINTER(C.c_double))]
