/*
 * oracle.h -- CPU restatement of the reference render path (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load the
 * library built from this directory, and only as the reference results of
 * parity checks or as the CPU baseline.  The product (noor_amd/) never links or calls it.
 *
 * The restatement follows the reference files named in each function's
 * comment.  It consumes the same wire format as the product's C ABI
 * (include/noor_hip.h): the payload arrays built by the host BVH builder.
 * Two instantiations of the integrator exist: FP32 (the bit-exact parity
 * target for the HIP kernels) and FP64 (the convergence oracle).
 *
 * Deliberate, documented choices where the reference is unspecified or
 * unreproducible (see also the module headers):
 *   - RNG draws inside one expression are taken left to right.
 *   - Transcendentals are the portable ones of omath.h.
 *   - normalize(v) = v * (1 / sqrt(dot(v, v))), IEEE division and sqrt.
 *   - Environment texture lookups are software bilinear with FP32 weights.
 *   - Closest-hit ties are broken by (t, prim_id, instance); a box is entered
 *     when its entry distance is <= the current tmax.
 *   - Material textures are filtered in software (bilinear per level, linear
 *     between levels at log2 of the footprint, point sampling for alpha);
 *     the mip chains and the Lanczos resampling follow the reference's
 *     kernels with binary32 arithmetic and a binary64 sinc.  These are pinned
 *     by closed forms (tests/test_oracle_textures.py); parity with the
 *     reference's hardware texture filtering is unpinned (CUDA does not
 *     specify its filter weights).
 */
#ifndef NOOR_ORACLE_H
#define NOOR_ORACLE_H

#include <stdint.h>
#include "../include/noor_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Material fields that can read an image texture (noor/cuda/materialimp.cuh:120-186). */
enum { ORC_TX_DIFFUSE = 0, ORC_TX_SPECULAR, ORC_TX_TRANSMISSION, ORC_TX_EMITTANCE, ORC_TX_ETA, ORC_TX_K,
       ORC_TX_ROUGH, ORC_TX_BUMP, ORC_TX_ALPHA, ORC_TX_COUNT };

/* Material: every 1x1 texture read once (noor/cuda/materialimp.cuh:134-214);
 * field f with bit f of tex_mask reads texture tex[f] of the store at uv
 * scale uvs[f] instead (image textures, and every bump and transparency map). */
typedef struct orc_material {
    uint32_t type;
    float diffuse[3], specular[3], transmission[3], emittance[3];
    float eta[3], k[3];
    float rough[2];
    float coat_ior;
    float coat_weight, coat_thickness, coat_sigma;
    uint32_t tex_mask;
    uint32_t tex[ORC_TX_COUNT];
    float uvs[ORC_TX_COUNT][2];
    float bumpfactor, emitter_scale;
} orc_material;

/* One texture of the store: level l is max(1, w >> l) x max(1, h >> l) RGBA
 * texels, the levels packed from texel `offset`. */
typedef struct orc_texdesc {
    uint32_t offset, w, h, levels;
} orc_texdesc;

/* 2D piecewise-constant distribution in the reference's single-buffer layout
 * (noor/cuda/distribution2d.cuh:165-200): rows of W+1 floats, the marginal
 * at row H, func[row*(W+1)+W] holds the row integral. */
typedef struct orc_dist2d {
    int w, h;
    float* func;
    float* cdf;
} orc_dist2d;

typedef struct orc_scene {
    noor_payload_view p;
    noor_hosek hosek;
    noor_camera cam;
    noor_spec spec;
    noor_options opt;
    orc_material* mats;
    /* environment texture (the last payload texture) */
    uint32_t env_w, env_h;
    const float* env;            /* RGBA32F */
    float* env_resized;          /* owned copy when the payload's sides were not powers of two */
    float* sky_bake;             /* PHYSICAL: sin(theta) * sky baked texture */
    uint32_t sky_w, sky_h;
    orc_dist2d dist;
    /* material texture store: mip chains (CudaMipMap, noor/cuda/texture.cuh:226-400) */
    float* tex_texels;           /* RGBA */
    orc_texdesc* tex_desc;
    uint32_t num_tex;
    int textured;                /* any material reads a texture of the store */
    int num_lights;              /* all lights incl. the infinite one */
    int sky_enabled;
    int stack_size;
} orc_scene;

/* Traversal events of the reference's binary BVH, counted separately for
   closest-hit (extension) and any-hit (shadow) traversals.  tri_second_fetch
   counts the triangle tests whose Wald projection puts t in [0, tmax], i.e.
   the tests at which the reference reads its second Wald record
   (noor/cuda/triangle.cuh:100-104), whichever test decides the hit. */
typedef struct orc_counters {
    uint64_t extension_rays, shadow_rays;
    uint64_t node_visits, tri_tests, tri_second_fetch, instance_entries;
    uint64_t shd_node_visits, shd_tri_tests, shd_tri_second_fetch, shd_instance_entries;
} orc_counters;

/* scene preparation (oracle_scene.c) */
int oracle_scene_create(const noor_payload_view* p, const noor_hosek* hosek,
                        const noor_camera* cam, const noor_spec* spec,
                        const noor_options* opt, orc_scene** out);
void oracle_scene_destroy(orc_scene* s);
void oracle_scene_set_camera(orc_scene* s, const noor_camera* cam);
void oracle_scene_set_spec(orc_scene* s, const noor_spec* spec);
void oracle_scene_set_options(orc_scene* s, const noor_options* opt);
/* Replace the environment texture (at its stored size, after resampling to
 * powers of two) and rebuild its distribution. */
int oracle_scene_set_env(orc_scene* s, const float* rgba);
/* Replace per-material diffuse rgb and roughness xy (adjoint tests). */
void oracle_scene_set_material_params(orc_scene* s, int mat, const float* diffuse,
                                      const float* rough);

/* the 2D distribution (noor/cuda/distribution2d.cuh:63-228) */
int oracle_dist_build(const float* rgba, int tw, int th, orc_dist2d* d);
void oracle_dist_free(orc_dist2d* d);
const float* oracle_scene_dist_func(const orc_scene* s, int* w, int* h);
const float* oracle_scene_dist_cdf(const orc_scene* s);
const float* oracle_scene_sky_bake(const orc_scene* s, int* w, int* h);

/* shared float helpers (oracle_scene.c) */
float oracle_tex_bilinear(const float* rgba, int w, int h, float u, float v, int c);
void oracle_hosek_query(const noor_hosek* hk, const float dir[3], float rgb[3]);
float oracle_hosek_radiance(const noor_hosek_state* st, float theta, float gamma, int channel);
uint32_t oracle_tea16(uint32_t v0, uint32_t v1);
uint32_t oracle_tea(uint32_t v0, uint32_t v1, const uint32_t key[4], int rounds, uint32_t* v1_out);
void oracle_pcg32_seed(uint64_t* state, uint64_t* inc, uint64_t initstate, uint64_t initseq);
uint32_t oracle_pcg32_next(uint64_t* state, uint64_t inc);

/* integrator instantiations (oracle_core.inc) */
int oracle_render_f32(orc_scene* s, uint32_t frame0, uint32_t spp, int nthreads,
                      uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                      float* film, orc_counters* cnt);
int oracle_render_f64(orc_scene* s, uint32_t frame0, uint32_t spp, int nthreads,
                      uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                      double* film, orc_counters* cnt);
/* the same with each pixel's path signature (hash of its last sample's
   discrete decisions) written to sig, row-major over the rectangle */
int oracle_render_sig_f32(orc_scene* s, uint32_t frame0, uint32_t spp, int nthreads, uint32_t x0, uint32_t y0,
                          uint32_t x1, uint32_t y1, float* film, orc_counters* cnt, uint64_t* sig);
int oracle_render_sig_f64(orc_scene* s, uint32_t frame0, uint32_t spp, int nthreads, uint32_t x0, uint32_t y0,
                          uint32_t x1, uint32_t y1, double* film, orc_counters* cnt, uint64_t* sig);
int oracle_trace_f32(orc_scene* s, const noor_ray* rays, uint32_t n, int any_hit,
                     noor_hit* hits, orc_counters* cnt);
int oracle_trace_f64(orc_scene* s, const noor_ray* rays, uint32_t n, int any_hit,
                     noor_hit* hits, orc_counters* cnt);
/* Brute-force closest hit over every (instance, triangle) pair, same formula. */
int oracle_trace_brute_f32(orc_scene* s, const noor_ray* rays, uint32_t n, noor_hit* hits);
/* Generate the camera ray of pixel (x, y) for frame (test hook). */
void oracle_camera_ray_f32(orc_scene* s, uint32_t x, uint32_t y, uint32_t frame, noor_ray* out);

/* Texture store hooks: the descriptor of store texture i (0 on success),
 * its texels (all levels), and one lookup as the shading code makes it:
 * kind 0 tex2D (level 0), 1 tex2DGrad with the uv gradients g = (dudx, dvdx,
 * dudy, dvdy), 2 point-filtered. */
int oracle_texture_desc(const orc_scene* s, uint32_t i, orc_texdesc* out);
const float* oracle_texture_texels(const orc_scene* s);
void oracle_tex_lookup_f32(const orc_scene* s, uint32_t i, int kind, float u, float v, const float g[4],
                           float out[4]);
/* Camera ray of pixel (x, y) for `frame` with its differentials:
 * out = o, d, ox, oy, dx, dy (18 floats) (test hook). */
void oracle_camera_ray_diff_f32(orc_scene* s, uint32_t x, uint32_t y, uint32_t frame, float out[18]);
/* The camera ray's first surface hit with its differentials: out = u, v,
 * dudx, dvdx, dudy, dvdy, p, dpdx, dpdy, gn; returns 0 when it hits no
 * triangle (test hook). */
int oracle_hit_diff_f32(orc_scene* s, uint32_t x, uint32_t y, uint32_t frame, float out[18]);

/* Detached-sampling render for finite differences (FP64): sampling decisions
 * use `s0` (frozen parameters), evaluation uses `s` (perturbed parameters). */
int oracle_render_detached_f64(orc_scene* s, orc_scene* s0, uint32_t frame0, uint32_t spp,
                               int nthreads, double* film);

#ifdef __cplusplus
}
#endif
#endif
