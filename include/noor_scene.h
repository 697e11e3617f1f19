/*
 * noor_scene.h -- host-side scene assembly: binned-SAH BVH build, Wald
 * triangle precompute, payload flattening and camera / spec set-up.
 *
 * Replaces the host layer that produced the reference's CudaPayload:
 *   BVH::BVH / BVH::build / makeInner / makeLeaf  (noor/source/bvh.cpp:29-258)
 *   Splitter::findBestSplit / split               (noor/source/splitter.cpp:28-161)
 *   WaldTriangle                                  (noor/header/triangle.h:56-120)
 *   Model::buildWaldTriangles / loadCudaPayload   (noor/source/model.cpp:333-458)
 *   BVH::loadCudaPayload (instance transforms)    (noor/source/bvh.cpp:53-86)
 *   Camera ctor / reset / update*                 (noor/header/camera.h:99-334)
 *   Scene::load spec fill                         (noor/source/scene.cpp:51-91)
 * The produced arrays use the reference node layout (two uint4 per node) and
 * triangle records, so a payload built here can be handed to noor_create or
 * read by any consumer of the reference format.
 */
#ifndef NOOR_SCENE_H
#define NOOR_SCENE_H

#include "noor_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One source mesh in object space. */
typedef struct noor_mesh_desc {
    uint32_t num_vertices;
    const float* positions;      /* xyz per vertex */
    const float* normals;        /* xyz per vertex, or NULL: face normals are used */
    const float* uvs;            /* uv per vertex, or NULL: zeros */
    uint32_t num_triangles;
    const uint32_t* indices;     /* 3 per triangle, counter-clockwise */
    const uint32_t* material;    /* per-triangle material index, or NULL: material_all */
    uint32_t material_all;
} noor_mesh_desc;

/* A placement of a source mesh: object-to-world as a 3x4 row-major matrix. */
typedef struct noor_instance_desc {
    uint32_t mesh;
    float object_to_world[12];
} noor_instance_desc;

/* BVHSpec (noor/header/options.h:130-135 defaults). */
typedef struct noor_bvh_params {
    uint32_t num_bins;           /* 16 */
    uint32_t max_height;         /* 32 */
    uint32_t min_leaf_tris;      /* 2 */
    uint32_t max_leaf_tris;      /* 2 */
    float Ci, Ct;                /* 1, 0.125 */
    int32_t threads;             /* 0 = hardware concurrency */
} noor_bvh_params;

typedef struct noor_scene_desc {
    uint32_t num_meshes;
    const noor_mesh_desc* meshes;
    uint32_t num_instances;      /* 0: each mesh placed once with identity */
    const noor_instance_desc* instances;
    uint32_t num_materials;
    const noor_material* materials;
    uint32_t num_textures;       /* the last one is the environment texture */
    const noor_texture_view* textures;
    uint32_t num_area_lights;
    const noor_area_light* area_lights;  /* each becomes a light node in the top level */
    uint32_t num_point_lights;
    const noor_point_light* point_lights;
    uint32_t num_spot_lights;
    const noor_spot_light* spot_lights;
    uint32_t num_distant_lights;
    const noor_distant_light* distant_lights;
} noor_scene_desc;

/* Build statistics (noor/header/stat.h:51-77). */
typedef struct noor_build_stats {
    uint32_t height;             /* Stat._height after the top-level build */
    uint32_t stack_depth;        /* traversal stack entries needed in the worst case */
    uint32_t num_nodes, num_inner, num_leaves, max_leaf_tris;
    float scene_min[3], scene_max[3];
    float scene_radius;          /* BBox::radius of the scene box */
    double build_seconds;
} noor_build_stats;

typedef struct noor_payload noor_payload;

noor_status noor_scene_build(const noor_scene_desc* desc, const noor_bvh_params* params,
                             noor_payload** out);
/* The same build with the bottom-level trees of meshes of 16384 or more
 * triangles built on HIP device `device` (a level-by-level binned SAH that makes
 * the host builder's decisions with the same arithmetic: the payload is the
 * same, byte for byte).  Replaces the single-threaded recursion of
 * noor/source/bvh.cpp:180-258 and noor/source/splitter.cpp:62-161 for large
 * meshes; needs num_bins <= 16 (NOOR_ERR_UNSUPPORTED otherwise). */
noor_status noor_scene_build_device(const noor_scene_desc* desc, const noor_bvh_params* params, int device,
                                    noor_payload** out);
noor_status noor_payload_view_get(const noor_payload* p, noor_payload_view* view);
noor_status noor_payload_stats(const noor_payload* p, noor_build_stats* stats);
void noor_payload_free(noor_payload* p);

/* Camera set-up as the reference host Camera does it: look-at frame,
 * perspective (fov in radians, near 0.01, far 1000) or orthographic
 * projection, screen window from the aspect ratio, raster grid of w x h. */
noor_status noor_camera_look_at(const float eye[3], const float look_at[3], const float up[3],
                                float fov, float ortho_zoom, float lens_radius,
                                uint32_t w, uint32_t h, int32_t type, noor_camera* out);

/* Spec fill of Scene::load: biases 1e-4 * scene radius, world radius, root,
 * height, bounces, rr, sky enabled when fewer than 2 lights exist. */
noor_status noor_spec_fill(const noor_payload* p, uint16_t bounces, uint16_t rr,
                           int32_t skydome_type, uint8_t mis, noor_spec* out);

/* Hosek-Wilkie RGB sky state for turbidity in [1, 10], ground albedo in
 * [0, 1] and solar elevation in [0, pi/2] radians: the nine configuration
 * parameters and the radiance scale per channel, blended from the published
 * coefficient tables.  Replaces arhosek_rgb_skymodelstate_alloc_update
 * (noor/external/ArHosekSkyModel.cpp:147-304,667-695); bit-identical cooks. */
noor_status noor_hosek_cook_state(float turbidity, float albedo, float elevation, noor_hosek_state* out);

/* The whole sky as HosekSky::update builds it (noor/header/hosek.h:61-83):
 * one state per ground-albedo channel at elevation pi/2 - sun_lat, sun
 * direction sphericalDirection(sun_lat, sun_lon) (y up), solar_scale (the
 * reference's default 250) and solar_radius_factor x the terrestrial solar
 * radius (default factor 2).  sun_lat in [0, pi/2].  Hand the result to
 * noor_create or noor_update_hosek. */
noor_status noor_hosek_cook(float turbidity, const float ground_albedo[3], float sun_lat, float sun_lon,
                            float solar_scale, float solar_radius_factor, noor_hosek* out);

#ifdef __cplusplus
}
#endif
#endif
