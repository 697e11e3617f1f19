"""ctypes mirror of include/noor_hip.h and include/noor_scene.h.

The structures below are field-for-field copies of the C declarations; the
layout checks in tests/test_abi.py compare their sizes and offsets with the
values the C compiler reports through ``noor_abi_layout``.
"""
import ctypes as C

c_f = C.c_float
c_u32 = C.c_uint32
c_i32 = C.c_int32
c_u16 = C.c_uint16
c_u8 = C.c_uint8
c_u64 = C.c_uint64


class float2(C.Structure):
    _fields_ = [("x", c_f), ("y", c_f)]


class float3(C.Structure):
    _fields_ = [("x", c_f), ("y", c_f), ("z", c_f)]


class float4(C.Structure):
    _fields_ = [("x", c_f), ("y", c_f), ("z", c_f), ("w", c_f)]


class uint4(C.Structure):
    _fields_ = [("x", c_u32), ("y", c_u32), ("z", c_u32), ("w", c_u32)]


class Transform(C.Structure):
    _fields_ = [("row", float3 * 3), ("t", float3)]


class Camera(C.Structure):
    _fields_ = [
        ("camera_to_world", Transform),
        ("raster_to_camera", Transform),
        ("world_to_raster", Transform),
        ("dx_camera", float3), ("dy_camera", float3),
        ("lens_radius", c_f), ("focal_length", c_f), ("image_area", c_f),
        ("w", c_u32), ("h", c_u32), ("center", c_u32), ("type", c_i32),
    ]


class HosekState(C.Structure):
    _fields_ = [
        ("configs", (c_f * 9) * 3),
        ("radiances", c_f * 3),
        ("turbidity", c_f),
        ("solar_radius", c_f),
        ("emission_correction_factor_sky", c_f * 3),
        ("emission_correction_factor_sun", c_f * 3),
        ("albedo", c_f),
        ("elevation", c_f),
    ]


class Hosek(C.Structure):
    _fields_ = [
        ("state_r", HosekState), ("state_g", HosekState), ("state_b", HosekState),
        ("solar_dir", float3), ("ground_albedo", float3),
        ("solar_scale", c_f), ("solar_radius", c_f),
    ]


class Spec(C.Structure):
    _fields_ = [
        ("white", float3), ("black", float3),
        ("reflection_bias", c_f), ("shadow_bias", c_f), ("world_radius", c_f), ("wr2", c_f),
        ("bvh_root_node", c_u32),
        ("lighting_type", c_u16), ("bounces", c_u16), ("rr", c_u16), ("bvh_height", c_u16),
        ("num_gpus", c_u16), ("pad0", c_u16),
        ("skydome_type", c_i32),
        ("debug_sky", c_u8), ("mis", c_u8), ("outofsync", c_u8), ("pad1", c_u8),
    ]


class Material(C.Structure):
    _fields_ = [
        ("diffuse_uvscale", float2), ("specular_uvscale", float2), ("transparency_uvscale", float2),
        ("emittance_uvscale", float2), ("bump_uvscale", float2), ("roughness_uvscale", float2),
        ("metalness_uvscale", float2),
        ("diffuse_tex_idx", c_u32), ("specular_tex_idx", c_u32), ("transmission_tex_idx", c_u32),
        ("emittance_tex_idx", c_u32), ("bump_tex_idx", c_u32), ("transparency_tex_idx", c_u32),
        ("ior_tex_idx", c_u32), ("k_tex_idx", c_u32), ("roughness_tex_idx", c_u32),
        ("metalness_tex_idx", c_u32),
        ("bumpfactor", c_f), ("emitter_scale", c_f), ("coat_weight", c_f), ("coat_roughness", c_f),
        ("coat_thickness", c_f), ("coat_sigma", c_f), ("coat_ior", c_f),
        ("type", c_u32),
    ]


class Shape(C.Structure):
    _fields_ = [
        ("center", float3), ("u", float3), ("v", float3), ("n", float3),
        ("area", c_f), ("radius", c_f), ("radius2", c_f),
        ("light_idx", c_i32), ("type", c_i32),
    ]


class AreaLight(C.Structure):
    _fields_ = [("ke", float3), ("power", float3), ("type", c_i32), ("shape", Shape), ("two_sided", c_i32)]


class PointLight(C.Structure):
    _fields_ = [("ke", float3), ("power", float3), ("type", c_i32), ("position", float3)]


class SpotLight(C.Structure):
    _fields_ = [("ke", float3), ("power", float3), ("type", c_i32), ("position", float3),
                ("direction", float3), ("cosinner", c_f), ("cosouter", c_f)]


class DistantLight(C.Structure):
    _fields_ = [("ke", float3), ("power", float3), ("type", c_i32), ("direction", float3)]


class TextureView(C.Structure):
    _fields_ = [("width", c_u32), ("height", c_u32), ("rgba", C.POINTER(c_f))]


class PayloadView(C.Structure):
    _fields_ = [
        ("num_triangles", c_u32),
        ("attribute_indices", C.POINTER(uint4)),
        ("wald_ax_nu_nv_nd", C.POINTER(uint4)),
        ("wald_bnu_bnv_au_av", C.POINTER(uint4)),
        ("wald_cnu_cnv_deg_mat", C.POINTER(uint4)),
        ("prim_ids", C.POINTER(c_u32)),
        ("num_vertices", c_u32),
        ("vertices", C.POINTER(float4)),
        ("normals", C.POINTER(float4)),
        ("uvs", C.POINTER(float2)),
        ("bvh_root_node", c_u32),
        ("num_bvh_nodes", c_u32),
        ("bvh_min_right_start", C.POINTER(uint4)),
        ("bvh_max_axis_count", C.POINTER(uint4)),
        ("num_instances", c_u32),
        ("world_to_object_row", C.POINTER(float4) * 3),
        ("object_to_world_row", C.POINTER(float4) * 3),
        ("normal_row", C.POINTER(float4) * 3),
        ("num_textures", c_u32),
        ("textures", C.POINTER(TextureView)),
        ("num_materials", c_u32),
        ("materials", C.POINTER(Material)),
        ("num_area_lights", c_u32),
        ("area_lights", C.POINTER(AreaLight)),
        ("num_point_lights", c_u32),
        ("point_lights", C.POINTER(PointLight)),
        ("num_spot_lights", c_u32),
        ("spot_lights", C.POINTER(SpotLight)),
        ("num_distant_lights", c_u32),
        ("distant_lights", C.POINTER(DistantLight)),
        ("num_infinite_lights", c_u32),
    ]


class Options(C.Structure):
    _fields_ = [("rng", c_i32), ("tri_test", c_i32), ("compat", c_u32), ("stats", c_u32)]


class Stats(C.Structure):
    _fields_ = [(n, c_u64) for n in (
        "extension_rays", "shadow_rays", "node_visits", "tri_tests", "tri_second_fetch",
        "instance_entries", "shadow_node_visits", "shadow_tri_tests", "shadow_tri_second_fetch",
        "shadow_instance_entries", "extension_hits", "drain_extension_rays", "drain_shadow_rays")]


class Ray(C.Structure):
    _fields_ = [("o", float3), ("tmax", c_f), ("d", float3), ("pad", c_f)]


class Hit(C.Structure):
    _fields_ = [("t", c_f), ("u", c_f), ("v", c_f), ("prim", c_u32), ("prim_id", c_u32),
                ("inst", c_u32), ("kind", c_u32), ("pad", c_u32)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", c_u64), ("total_ms", C.c_double)]


class View(C.Structure):
    _fields_ = [("camera", Camera), ("frame", c_u32), ("spp", c_u32), ("seed", c_u32), ("pad", c_u32)]


class TuneEntry(C.Structure):
    _fields_ = [("queue_capacity", c_u32), ("block", c_i32), ("stack_lds", c_u32), ("ms", c_f)]


class MeshDesc(C.Structure):
    _fields_ = [
        ("num_vertices", c_u32), ("positions", C.POINTER(c_f)), ("normals", C.POINTER(c_f)),
        ("uvs", C.POINTER(c_f)), ("num_triangles", c_u32), ("indices", C.POINTER(c_u32)),
        ("material", C.POINTER(c_u32)), ("material_all", c_u32),
    ]


class InstanceDesc(C.Structure):
    _fields_ = [("mesh", c_u32), ("object_to_world", c_f * 12)]


class BvhParams(C.Structure):
    _fields_ = [("num_bins", c_u32), ("max_height", c_u32), ("min_leaf_tris", c_u32),
                ("max_leaf_tris", c_u32), ("Ci", c_f), ("Ct", c_f), ("threads", c_i32)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_meshes", c_u32), ("meshes", C.POINTER(MeshDesc)),
        ("num_instances", c_u32), ("instances", C.POINTER(InstanceDesc)),
        ("num_materials", c_u32), ("materials", C.POINTER(Material)),
        ("num_textures", c_u32), ("textures", C.POINTER(TextureView)),
        ("num_area_lights", c_u32), ("area_lights", C.POINTER(AreaLight)),
        ("num_point_lights", c_u32), ("point_lights", C.POINTER(PointLight)),
        ("num_spot_lights", c_u32), ("spot_lights", C.POINTER(SpotLight)),
        ("num_distant_lights", c_u32), ("distant_lights", C.POINTER(DistantLight)),
    ]


class BuildStats(C.Structure):
    _fields_ = [
        ("height", c_u32), ("stack_depth", c_u32), ("num_nodes", c_u32), ("num_inner", c_u32),
        ("num_leaves", c_u32), ("max_leaf_tris", c_u32),
        ("scene_min", c_f * 3), ("scene_max", c_f * 3), ("scene_radius", c_f),
        ("build_seconds", C.c_double),
    ]


# include/noor_loaders.h
PATH_MAX_ = 256


class ObjMaterial(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("type", C.c_uint32), ("diffuse", C.c_float * 3),
                ("specular", C.c_float * 3), ("transmission", C.c_float * 3), ("emittance", C.c_float * 3),
                ("ior", C.c_float), ("roughness", C.c_float), ("transparency", C.c_float),
                ("bumpfactor", C.c_float), ("uvscale", C.c_float * 2),
                ("map_kd", C.c_char * PATH_MAX_), ("map_ks", C.c_char * PATH_MAX_), ("map_ke", C.c_char * PATH_MAX_),
                ("map_d", C.c_char * PATH_MAX_), ("map_bump", C.c_char * PATH_MAX_)]


class ObjView(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)),
                ("num_triangles", C.c_uint32), ("indices", C.POINTER(C.c_uint32)),
                ("material", C.POINTER(C.c_uint32)), ("num_materials", C.c_uint32),
                ("materials", C.POINTER(ObjMaterial)), ("num_lights", C.c_uint32),
                ("lights", C.POINTER(C.c_float)), ("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3),
                ("unitize_scale", C.c_float)]


# enums (include/noor_hip.h)
MAT_DIFFUSE, MAT_TRANSLUCENT, MAT_GLOSSY, MAT_MIRROR = 1, 4, 8, 16
MAT_GLASS, MAT_ROUGHGLASS, MAT_METAL, MAT_EMITTER = 32, 64, 128, 512
MAT_SUBSTRATE, MAT_CLEARCOAT, MAT_SHADOWCATCHER = 256, 1024, 2048
MAT_ALPHA, MAT_BUMP = 1 << 12, 1 << 13
MAT_MESHLIGHT = 1 << 14
LIGHT_POINT, LIGHT_SPOT, LIGHT_DISTANT, LIGHT_AREA, LIGHT_INFINITE = 1, 2, 4, 8, 16
SHAPE_QUAD, SHAPE_SPHERE, SHAPE_DISK = 0, 1, 2
SKY_HDR, SKY_PHYSICAL, SKY_CONSTANT = 0, 1, 2
CAMERA_PERSP, CAMERA_ORTHO, CAMERA_ENV = 0, 1, 2
RNG_TEA_LCG, RNG_PCG32 = 0, 1
TRI_WATERTIGHT, TRI_WALD = 0, 1
COMPAT_QUAD_PDF, COMPAT_CDF_N_PLUS_1, COMPAT_CLAMP, COMPAT_WASTED_DRAWS = 1, 2, 4, 8
COMPAT_MIS_LIGHT_COUNT, COMPAT_MIS_SELF_SHADOW, COMPAT_RESIZE_SUM = 16, 32, 64
COMPAT_ALL = 127
LIGHT_SKY_MASK = 0x20

OK, ERR_INVALID_ARG, ERR_INVALID_PAYLOAD, ERR_HIP, ERR_OUT_OF_MEMORY = 0, 1, 2, 3, 4
ERR_STACK_DEPTH, ERR_NO_DEVICE, ERR_UNSUPPORTED, ERR_QUEUE_OVERFLOW = 5, 6, 7, 8

STATUS = {0: "ok", 1: "invalid argument", 2: "invalid payload", 3: "HIP error", 4: "out of memory",
          5: "stack depth", 6: "no device", 7: "unsupported", 8: "queue overflow"}


class NoorError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"noor status {status} ({STATUS.get(status, '?')}) {msg}")
