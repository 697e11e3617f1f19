/*
 * noor_hip.h -- plain C ABI of the MI355X (gfx950) render path.
 *
 * This header replaces the `extern "C"` block of the reference renderer
 * (noor/cuda/path_tracer.cuh:136-151).  Those eight functions took C++
 * objects (std::unique_ptr<CudaPayload>, glm vectors, GLuint) and called
 * exit() on any CUDA error (noor/cuda/path_tracer.cuh:154-171).  Here every
 * argument is a POD struct or a pointer + count, and every call returns a
 * status code.  The PODs keep the field order and meaning of the reference
 * wire format so that a host which fills CudaPayload / CudaSpec / CudaCamera /
 * CudaHosekSky can fill these instead, field by field.
 *
 * Mapping of the reference entry points:
 *   load_cuda_data      (noor/cuda/path_tracer_data.cuh:53-62)  -> noor_create
 *   update_cuda_camera  (noor/cuda/path_tracer_data.cuh:72)     -> noor_update_camera
 *   update_cuda_hosek   (noor/cuda/path_tracer_data.cuh:76)     -> noor_update_hosek
 *   update_cuda_spec    (noor/cuda/path_tracer_data.cuh:68)     -> noor_update_spec
 *   cuda_path_tracer    (noor/cuda/path_tracer.cu:132-155)      -> noor_render
 *   debug_skydome       (noor/cuda/path_tracer.cu:124-130)      -> noor_debug_skydome
 *   device_free_memory  (noor/cuda/path_tracer_data.cuh:80-82)  -> noor_destroy
 *   get_lookAt          (noor/cuda/path_tracer_data.cuh:64-66)  -> noor_get_lookat
 * plus the film read-back that the reference did through GL interop
 * (noor/cuda/render.cuh:213-232) -> noor_read_film, and test / extension hooks
 * (noor_trace, noor_render_wavefront, noor_render_batch, noor_adjoint_*).
 *
 * Threading: one context owns one HIP stream; calls on one context must not
 * overlap.  Independent contexts may be used from different threads.
 */
#ifndef NOOR_HIP_H
#define NOOR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: noor_stats gained extension_hits and the drain's ray counts */
#define NOOR_ABI_VERSION 2

typedef enum noor_status {
    NOOR_OK = 0,
    NOOR_ERR_INVALID_ARG = 1,      /* null pointer, zero size, bad enum */
    NOOR_ERR_INVALID_PAYLOAD = 2,  /* index out of range, NaN vertex, bad node */
    NOOR_ERR_HIP = 3,              /* a HIP runtime call failed */
    NOOR_ERR_OUT_OF_MEMORY = 4,
    NOOR_ERR_STACK_DEPTH = 5,      /* BVH deeper than the traversal stack */
    NOOR_ERR_NO_DEVICE = 6,
    NOOR_ERR_UNSUPPORTED = 7,      /* feature outside the supported set */
    NOOR_ERR_QUEUE_OVERFLOW = 8    /* wavefront queue capacity exceeded */
} noor_status;

/* ---- vector PODs (CUDA float3/float4/uint4 layouts) ---------------------- */
typedef struct noor_float2 { float x, y; } noor_float2;
typedef struct noor_float3 { float x, y, z; } noor_float3;
typedef struct noor_float4 { float x, y, z, w; } noor_float4;
typedef struct noor_uint4 { uint32_t x, y, z, w; } noor_uint4;

/* ---- enums and masks of the reference (noor/cuda/path_tracer.cuh:48-131) -- */
#define NOOR_TRI_LEAF_MASK      0x80000000u
#define NOOR_MESH_LEAF_MASK     0x40000000u
#define NOOR_MESH_INSTANCE_MASK 0x20000000u
#define NOOR_MESH_INNER_MASK    0x10000000u
#define NOOR_MESH_NODE_MASK     0x70000000u
#define NOOR_COUNT_MASK         0x0FFFFFFFu
#define NOOR_TRI_COUNT_MASK     0x7FFFFFFFu
#define NOOR_AXIS_MASK          0x0FFFFFFFu
#define NOOR_TRANS_MASK         0x0FFFFFFFu
#define NOOR_LIGHT_NODE_MASK    0x08000000u
#define NOOR_LIGHT_MASK         0x07FFFFFFu

/* MaterialType (noor/cuda/path_tracer.cuh:87-103) */
enum {
    NOOR_MAT_DIFFUSE = 1 << 0, NOOR_MAT_ORENNAYAR = 1 << 1, NOOR_MAT_TRANSLUCENT = 1 << 2,
    NOOR_MAT_GLOSSY = 1 << 3, NOOR_MAT_MIRROR = 1 << 4, NOOR_MAT_GLASS = 1 << 5,
    NOOR_MAT_ROUGHGLASS = 1 << 6, NOOR_MAT_METAL = 1 << 7, NOOR_MAT_SUBSTRATE = 1 << 8,
    NOOR_MAT_EMITTER = 1 << 9, NOOR_MAT_CLEARCOAT = 1 << 10, NOOR_MAT_SHADOWCATCHER = 1 << 11,
    NOOR_MAT_ALPHA = 1 << 12, NOOR_MAT_BUMP = 1 << 13, NOOR_MAT_MESHLIGHT = 1 << 14
};
/* LightType (noor/cuda/light.cuh:33-40) */
enum { NOOR_LIGHT_POINT = 1, NOOR_LIGHT_SPOT = 2, NOOR_LIGHT_DISTANT = 4,
       NOOR_LIGHT_AREA = 8, NOOR_LIGHT_INFINITE = 16 };
/* AreaMeshLightType, SkydomeType, CameraType (noor/cuda/path_tracer.cuh:128-131) */
enum { NOOR_SHAPE_QUAD = 0, NOOR_SHAPE_SPHERE = 1, NOOR_SHAPE_DISK = 2 };
enum { NOOR_SKY_HDR = 0, NOOR_SKY_PHYSICAL = 1, NOOR_SKY_CONSTANT = 2 };
enum { NOOR_CAMERA_PERSP = 0, NOOR_CAMERA_ORTHO = 1, NOOR_CAMERA_ENV = 2 };
/* CudaSpec._lighting_type bits (noor/cuda/spec.cuh:26-31) */
#define NOOR_LIGHT_SKY_MASK 0x0020u

/* ---- scene PODs ---------------------------------------------------------- */

/* CudaTransform (noor/cuda/transform.cuh:26-69): 3x3 rows + translation. */
typedef struct noor_transform { noor_float3 row[3]; noor_float3 t; } noor_transform;

/* CudaCamera (noor/cuda/camera.cuh:95-102). */
typedef struct noor_camera {
    noor_transform camera_to_world;
    noor_transform raster_to_camera;
    noor_transform world_to_raster;
    noor_float3 dx_camera, dy_camera;
    float lens_radius, focal_length, image_area;
    uint32_t w, h;
    uint32_t center;
    int32_t type;           /* NOOR_CAMERA_* */
} noor_camera;

/* CudaArHosekSkyModelState (noor/cuda/hosek.cuh:30-40). */
typedef struct noor_hosek_state {
    float configs[3][9];
    float radiances[3];
    float turbidity;
    float solar_radius;
    float emission_correction_factor_sky[3];
    float emission_correction_factor_sun[3];
    float albedo;
    float elevation;
} noor_hosek_state;

/* CudaHosekSky (noor/cuda/hosek.cuh:42-51). */
typedef struct noor_hosek {
    noor_hosek_state state_r, state_g, state_b;
    noor_float3 solar_dir;
    noor_float3 ground_albedo;
    float solar_scale;
    float solar_radius;
} noor_hosek;

/* CudaSpec (noor/cuda/spec.cuh:41-58).  bool fields are widened to uint8. */
typedef struct noor_spec {
    noor_float3 white, black;
    float reflection_bias, shadow_bias, world_radius, wr2;
    uint32_t bvh_root_node;
    uint16_t lighting_type, bounces, rr, bvh_height, num_gpus;
    uint16_t pad0;
    int32_t skydome_type;   /* NOOR_SKY_* */
    uint8_t debug_sky, mis, outofsync, pad1;
} noor_spec;

/* CudaMaterial (noor/cuda/material.cuh:26-58), 128 bytes.  Every colour is a
 * texture index into noor_payload_view.textures; constants are 1x1 textures
 * (noor/source/model.cpp:211-299).  Larger textures are image textures, read
 * mip-mapped at the pixel footprint (transmission, ior and k at uv scale 1,
 * roughness at level 0).  Type bit NOOR_MAT_BUMP reads bump_tex_idx (first
 * channel, times bumpfactor, at bump_uvscale); NOOR_MAT_ALPHA reads
 * transparency_tex_idx (first channel, point-sampled at transparency_uvscale;
 * <= 1e-4 cuts the surface away).  The v coordinate is flipped (1 - v) on
 * read, as the reference's mesh manager does. */
typedef struct noor_material {
    noor_float2 diffuse_uvscale, specular_uvscale, transparency_uvscale,
                emittance_uvscale, bump_uvscale, roughness_uvscale, metalness_uvscale;
    uint32_t diffuse_tex_idx, specular_tex_idx, transmission_tex_idx, emittance_tex_idx,
             bump_tex_idx, transparency_tex_idx, ior_tex_idx, k_tex_idx,
             roughness_tex_idx, metalness_tex_idx;
    float bumpfactor, emitter_scale, coat_weight, coat_roughness, coat_thickness,
          coat_sigma, coat_ior;
    uint32_t type;          /* NOOR_MAT_* */
} noor_material;

/* CudaShape (noor/cuda/shape.cuh:30-40). */
typedef struct noor_shape {
    noor_float3 center, u, v, n;
    float area, radius, radius2;
    int32_t light_idx;
    int32_t type;           /* NOOR_SHAPE_* */
} noor_shape;

/* CudaAreaLight (noor/cuda/light.cuh:42-100): CudaLight base + shape + flag. */
typedef struct noor_area_light {
    noor_float3 ke, power;
    int32_t type;           /* NOOR_LIGHT_AREA */
    noor_shape shape;
    int32_t two_sided;
} noor_area_light;

/* CudaPointLight / CudaSpotLight / CudaDistantLight (noor/cuda/light.cuh:160-249). */
typedef struct noor_point_light { noor_float3 ke, power; int32_t type; noor_float3 position; } noor_point_light;
typedef struct noor_spot_light {
    noor_float3 ke, power; int32_t type; noor_float3 position, direction; float cosinner, cosouter;
} noor_spot_light;
typedef struct noor_distant_light { noor_float3 ke, power; int32_t type; noor_float3 direction; } noor_distant_light;

/* ImageTexture as the device sees it: RGBA32F texels, row-major, row 0 at
 * v = 0.  Material image textures (at most 32768 a side) are resampled to
 * the nearest power of two per side (Lanczos, see NOOR_COMPAT_RESIZE_SUM),
 * then read with wrap addressing from a mip chain as CudaMipMap builds it,
 * bilinear within a level and linear between levels
 * (noor/cuda/texture.cuh:30-400, noor/cuda/image.cuh:28-125).
 * The last texture of the payload is the environment / sky texture (any
 * size up to 65536 a side, resampled the same way; clamp addressing,
 * bilinear at level 0). */
typedef struct noor_texture_view {
    uint32_t width, height;
    const float* rgba;      /* width*height*4 floats */
} noor_texture_view;

/* CudaPayload (noor/header/payload.h:27-127) as pointer + count views.
 * Triangles are stored in BVH leaf order (noor/source/model.cpp:333-360);
 * prim_ids (an addition) maps leaf order back to the caller's triangle ids so
 * that closest-hit ties can be broken by the lowest primitive id. */
typedef struct noor_payload_view {
    uint32_t num_triangles;
    const noor_uint4* attribute_indices;     /* {v0, v1, v2, material index} */
    const noor_uint4* wald_ax_nu_nv_nd;
    const noor_uint4* wald_bnu_bnv_au_av;
    const noor_uint4* wald_cnu_cnv_deg_mat;  /* .w = material type bits */
    const uint32_t* prim_ids;                /* may be NULL: identity */

    uint32_t num_vertices;
    const noor_float4* vertices;
    const noor_float4* normals;
    const noor_float2* uvs;

    uint32_t bvh_root_node;
    uint32_t num_bvh_nodes;
    const noor_uint4* bvh_min_right_start;
    const noor_uint4* bvh_max_axis_count;

    uint32_t num_instances;                  /* transforms, one per mesh/instance */
    const noor_float4* world_to_object_row[3];
    const noor_float4* object_to_world_row[3];
    const noor_float4* normal_row[3];

    uint32_t num_textures;
    const noor_texture_view* textures;
    uint32_t num_materials;
    const noor_material* materials;

    uint32_t num_area_lights;
    const noor_area_light* area_lights;
    uint32_t num_point_lights;
    const noor_point_light* point_lights;
    uint32_t num_spot_lights;
    const noor_spot_light* spot_lights;
    uint32_t num_distant_lights;
    const noor_distant_light* distant_lights;
    uint32_t num_infinite_lights;            /* 0 or 1; the infinite light is last */
} noor_payload_view;

/* ---- render options that the reference fixed at compile time ------------- */
enum { NOOR_RNG_TEA_LCG = 0, NOOR_RNG_PCG32 = 1 };
enum { NOOR_TRI_WATERTIGHT = 0, NOOR_TRI_WALD = 1 };
/* Reference behaviours reproduced when the bit is set (all set by default). */
#define NOOR_COMPAT_QUAD_PDF      0x1u  /* area-light pdf with unnormalised wi (noor/cuda/geometry.cuh:39-41) */
#define NOOR_COMPAT_CDF_N_PLUS_1  0x2u  /* sampleContinuous2D divides by W+1, H+1 (noor/cuda/distribution2d.cuh:331-333) */
#define NOOR_COMPAT_CLAMP         0x4u  /* clamp each sample to [0,100] (noor/cuda/path_tracer.cu:73) */
#define NOOR_COMPAT_WASTED_DRAWS  0x8u  /* two unused draws in scatter (noor/cuda/scatter.cuh:34) */
#define NOOR_COMPAT_MIS_LIGHT_COUNT 0x10u /* MIS: the BSDF-sampled NEE term is not scaled by the light count
                                             (noor/cuda/direct.cuh:69-113; sample_Li is, noor/cuda/lightimp.cuh:415-456) */
#define NOOR_COMPAT_MIS_SELF_SHADOW 0x20u /* MIS: the shadow ray to a BSDF-sampled light point ends on the light
                                             and can be blocked by it (noor/cuda/direct.cuh:98-104,
                                             noor/cuda/intersection.cuh:201-207) */
#define NOOR_COMPAT_RESIZE_SUM    0x40u /* non-power-of-two textures: the horizontal and vertical Lanczos passes
                                         are summed, not composed (noor/cuda/image.cuh:61-123); read at
                                         noor_create only */
#define NOOR_COMPAT_ALL           0x7Fu

typedef struct noor_options {
    int32_t rng;            /* NOOR_RNG_* */
    int32_t tri_test;       /* NOOR_TRI_* */
    uint32_t compat;        /* NOOR_COMPAT_* */
    uint32_t stats;         /* nonzero: count traversal events (slower) */
} noor_options;

/* Traversal / path counters.  Filled when options.stats != 0, except rays
 * which are always counted. */
typedef struct noor_stats {
    uint64_t extension_rays, shadow_rays;
    /* traversal events of closest-hit (extension) queries */
    uint64_t node_visits, tri_tests, tri_second_fetch, instance_entries;
    /* traversal events of any-hit (shadow) queries */
    uint64_t shadow_node_visits, shadow_tri_tests, shadow_tri_second_fetch, shadow_instance_entries;
    /* extension rays that hit a triangle or an area light (the rest escape) */
    uint64_t extension_hits;
    /* of extension_rays and shadow_rays, those the wavefront renderer's drain
     * kernel traced (the last paths of a group, finished in one launch) */
    uint64_t drain_extension_rays, drain_shadow_rays;
} noor_stats;

/* ---- hit queries (test hook; noor/cuda/bvh.cuh:142-234) ------------------ */
typedef struct noor_ray { noor_float3 o; float tmax; noor_float3 d; float pad; } noor_ray;
typedef struct noor_hit {
    float t, u, v;          /* u weights v1, v weights v2 (noor/cuda/intersection.cuh:343-345) */
    uint32_t prim;          /* leaf-order triangle index, 0xFFFFFFFF = miss */
    uint32_t prim_id;       /* caller triangle id (prim_ids[prim]) */
    uint32_t inst;          /* instance / transform index, or light index for a light hit */
    uint32_t kind;          /* 0 miss, 1 triangle, 2 light node */
    uint32_t pad;
} noor_hit;

/* ---- context API --------------------------------------------------------- */
typedef struct noor_ctx noor_ctx;

const char* noor_status_string(noor_status s);
int noor_abi_version(void);
/* Number of HIP devices visible; 0 without a GPU. */
int noor_device_count(void);

noor_status noor_create(const noor_payload_view* payload, const noor_hosek* hosek,
                        const noor_camera* camera, const noor_spec* spec,
                        const noor_options* options, noor_ctx** out);
noor_status noor_destroy(noor_ctx* ctx);
noor_status noor_update_camera(noor_ctx* ctx, const noor_camera* camera);
noor_status noor_update_hosek(noor_ctx* ctx, const noor_hosek* hosek);
noor_status noor_update_spec(noor_ctx* ctx, const noor_spec* spec);
noor_status noor_update_options(noor_ctx* ctx, const noor_options* options);
/* Last error message of this context (static storage owned by ctx). */
const char* noor_last_error(const noor_ctx* ctx);

/* Megakernel: renders `spp` progressive frames starting at *frame (frame 1
 * overwrites the film) and advances *frame by spp, like `spp` calls of
 * cuda_path_tracer. */
noor_status noor_render(noor_ctx* ctx, uint32_t* frame, uint32_t spp);
/* Wavefront renderer with material-sorted queues; same film semantics and
 * identical per-path arithmetic, so its film matches noor_render. */
noor_status noor_render_wavefront(noor_ctx* ctx, uint32_t* frame, uint32_t spp);
noor_status noor_debug_skydome(noor_ctx* ctx, uint32_t* frame);
noor_status noor_get_lookat(noor_ctx* ctx, noor_float4* lookat);
noor_status noor_read_film(noor_ctx* ctx, float* rgba, size_t count_floats);
/* Device address of the film (float4 per pixel); valid until noor_destroy. */
noor_status noor_film_device_ptr(noor_ctx* ctx, void** ptr);
noor_status noor_get_stats(noor_ctx* ctx, noor_stats* stats);
noor_status noor_reset_stats(noor_ctx* ctx);
noor_status noor_synchronize(noor_ctx* ctx);

/* Closest-hit (any_hit = 0) or any-hit (any_hit = 1) queries on host arrays. */
noor_status noor_trace(noor_ctx* ctx, const noor_ray* rays, uint32_t n, int any_hit,
                       noor_hit* hits);
/* noor_trace without allocating: no counterpart in the reference; for callers
 * that cannot allocate device memory (graph capture, serving loops).  Rays and
 * hits are DEVICE arrays of n entries; the call only enqueues the queries on
 * the context's stream (results are ready after noor_synchronize).  The
 * traversal stack beyond its LDS part uses a 16 MiB scratch array fixed in the
 * library, one per GPU: calls that overlap on one GPU (several contexts or
 * threads) race on it.  NOOR_ERR_UNSUPPORTED when one wave's stack does not fit
 * (more than 65,536 device entries per lane); no allocation is made. */
noor_status noor_trace_noalloc(noor_ctx* ctx, const noor_ray* d_rays, uint32_t n, int any_hit,
                               noor_hit* d_hits);

/* Live kernel timing with HIP events on the context's stream (per kernel
 * name: launches and total milliseconds).  Timing makes the wavefront
 * renderer launch kernels one by one instead of replaying its graph. */
typedef struct noor_kernel_time { char name[48]; uint64_t launches; double total_ms; } noor_kernel_time;
noor_status noor_set_timing(noor_ctx* ctx, int enable);
noor_status noor_get_kernel_times(noor_ctx* ctx, noor_kernel_time* out, uint32_t max, uint32_t* count);
noor_status noor_reset_kernel_times(noor_ctx* ctx);
/* Wavefront tunables: queue capacity in paths (0 = automatic: the items of
 * the render, at most 2^25, about 9 GB of path state, halved down to 2^22
 * when the device cannot hold it; otherwise 64 to 2^28) and lanes per block
 * of the traversal kernels (64, 128 or 256). */
noor_status noor_set_wavefront_params(noor_ctx* ctx, uint32_t queue_capacity, int32_t block);

/* Traversal stack entries per lane kept in LDS (the rest of the exact
 * per-BVH bound lives in device memory); 0 restores the default (24). */
noor_status noor_set_stack_lds(noor_ctx* ctx, uint32_t entries);

/* ---- multi-view batches (no counterpart in the reference, which rendered
 * one camera per process) ----------------------------------------------- */
typedef struct noor_view {
    noor_camera camera;
    uint32_t frame;   /* running-mean frame of the view's first sample (>= 1) */
    uint32_t spp;     /* samples to add */
    uint32_t seed;    /* RNG stream offset: sample s uses stream frame + s + seed */
    uint32_t pad;
} noor_view;
/* Renders every view of the batch in one wavefront pass into films the
 * context owns (one float4 film per view, in view order).  The films persist
 * between calls with the same view sizes, so a later call with frame > 1
 * continues each view's running mean; other sizes reallocate them (zeroed). */
noor_status noor_render_batch(noor_ctx* ctx, const noor_view* views, uint32_t n);
noor_status noor_read_batch_film(noor_ctx* ctx, uint32_t view, float* rgba, size_t count_floats);
noor_status noor_batch_film_device_ptr(noor_ctx* ctx, uint32_t view, void** ptr);

/* Wavefront autotuner: times noor_render_batch(views) for every setting of
 * block {64, 128, 256} x queue capacity {2^20 .. 2^25, up to the first that
 * holds every item of the batch} x LDS stack entries
 * {16, 24, 32} (clamped to what the BVH needs), keeps the fastest on the
 * context and returns the measurements (at most `max`).  The batch films are
 * overwritten. */
typedef struct noor_tune_entry {
    uint32_t queue_capacity;
    int32_t block;
    uint32_t stack_lds;
    float ms;
} noor_tune_entry;
noor_status noor_autotune(noor_ctx* ctx, const noor_view* views, uint32_t n, noor_tune_entry* out, uint32_t max,
                          uint32_t* count);

/* ---- differentiable rendering (no counterpart in the reference) -------- *
 * Parameters: the diffuse colour (albedo, RGB) and isotropic roughness of
 * materials [0, num_materials) and the RGB texels of the HDR environment.
 * All pointers below are DEVICE pointers on the context's GPU, row-major
 * float arrays: albedo num_materials x 3, roughness num_materials,
 * env_rgb env_h x env_w x 3, images h x w x 4 (the fourth channel unused).
 * Calls are ordered on the context's stream and return after it drains. */

/* Overwrites the parameters (any pointer may be NULL to keep that group);
 * env_w, env_h must match the environment of the payload.  The environment
 * sampling distribution is rebuilt from the new texels. */
noor_status noor_set_params_device(noor_ctx* ctx, const float* albedo, const float* roughness,
                                   uint32_t num_materials, const float* env_rgb, uint32_t env_w,
                                   uint32_t env_h);
/* One sample per pixel from RNG stream `stream` (>= 1), written to
 * `radiance` instead of the film (the megakernel's estimator). */
noor_status noor_render_sample(noor_ctx* ctx, uint32_t stream, float* radiance);
/* Path-replay adjoint of noor_render_sample(stream): given the forward
 * `radiance` and dLoss/dradiance (`grad_image`), ADDS dLoss/dparameter to
 * grad_albedo, grad_roughness (num_materials) and grad_env (may be NULL).
 * Sampling is detached (the random decisions are those of the forward
 * pass).  Requires MIS off and the per-sample clamp (NOOR_COMPAT_CLAMP) off. */
noor_status noor_adjoint(noor_ctx* ctx, uint32_t stream, const float* radiance, const float* grad_image,
                         float* grad_albedo, float* grad_roughness, uint32_t num_materials, float* grad_env);

/* Host-side check of the four-wide BVH the traversal kernels use (no GPU
 * needed): builds it from a payload and verifies that every triangle of every
 * instanced BLAS is reachable exactly once per BLAS, that every child box
 * equals the box of the binary node it came from, and the stack bound.
 * Reports the node count and the stack entries a traversal can need. */
noor_status noor_wide_bvh_check(const noor_payload_view* payload, uint32_t* num_nodes, uint32_t* stack_need);

/* sizeof() of a named ABI struct as compiled into the library (layout checks). */
size_t noor_abi_struct_size(const char* name);

#ifdef __cplusplus
}
#endif
#endif /* NOOR_HIP_H */
