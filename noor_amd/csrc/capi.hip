// C ABI of the render path (include/noor_hip.h).
//
// Replaces the reference's load_cuda_data / update_cuda_* / cuda_path_tracer /
// device_free_memory / get_lookAt (noor/cuda/path_tracer.cuh:136-151,
// noor/cuda/path_tracer_data.cuh:51-84, noor/cuda/render.cuh:37-265):
//   - the payload is validated and deep-copied to HBM; nothing is retained;
//   - the __constant__ manager singletons become one DevScene passed by value;
//   - CUDA texture objects become raw pointers with software filtering;
//   - every failure returns a status and records a message instead of exit().
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "ctx.h"
#include "texture_store.h"
#include "wide_bvh.h"
#include "noor_loaders.h"
#include "noor_scene.h"

using noor::Context;

struct noor_ctx {
    Context c;
};

// Traversal-stack scratch of noor_trace_noalloc: the stack entries beyond a
// lane's LDS part, lane-interleaved, for callers that cannot allocate device
// memory.  One copy per GPU, held until the process ends.  Host code never
// passes this name to a kernel (on the host it is a stand-in, not a device
// address): each device's address is looked up once, with that device
// current, at the first noor_create on it (scratch_address below).
constexpr size_t kScratchWords = (size_t)4 << 20;  // 16 MiB
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once per
// device), or nullptr with the HIP error in *err.
uint32_t* scratch_address(int device, hipError_t* err) {
    *err = hipSuccess;
    if (device < 0 || device >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*err != hipSuccess) return nullptr;
        g_scratch_ptr[device] = (uint32_t*)p;
    }
    return g_scratch_ptr[device];
}
}  // namespace

namespace {

const char* kStatus[] = {"ok", "invalid argument", "invalid payload", "HIP error", "out of memory",
                         "stack depth exceeded", "no device", "unsupported", "queue overflow"};

noor_status fail(Context* c, noor_status s, const std::string& msg) {
    if (c) c->last_error = msg;
    return s;
}

#define HIP_TRY(c, expr)                                                                           \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            return fail((c), e_ == hipErrorOutOfMemory ? NOOR_ERR_OUT_OF_MEMORY : NOOR_ERR_HIP,    \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                        \
        }                                                                                          \
    } while (0)

template <class T>
noor_status dev_upload(Context* c, T** dst, const T* src, size_t count) {
    *dst = nullptr;
    if (count == 0) count = 1;
    void* p = nullptr;
    HIP_TRY(c, hipMalloc(&p, sizeof(T) * count));
    c->allocs.push_back(p);
    if (src) HIP_TRY(c, hipMemcpyAsync(p, src, sizeof(T) * count, hipMemcpyHostToDevice, c->stream));
    else HIP_TRY(c, hipMemsetAsync(p, 0, sizeof(T) * count, c->stream));
    *dst = (T*)p;
    return NOOR_OK;
}

uint32_t stack_need(const noor_payload_view* P, uint32_t ni, std::vector<int32_t>& memo, int depth, bool* bad) {
    if (ni >= P->num_bvh_nodes || depth > 4096) {
        *bad = true;
        return 0;
    }
    if (memo[ni] >= 0) return (uint32_t)memo[ni];
    const noor_uint4 a = P->bvh_min_right_start[ni];
    const uint32_t w = P->bvh_max_axis_count[ni].w;
    uint32_t need = 0;
    if (w & NOOR_LIGHT_NODE_MASK) {
        need = 0;
    } else if (w & NOOR_MESH_INSTANCE_MASK) {
        need = std::max(1u, stack_need(P, a.w, memo, depth + 1, bad));
    } else if (w & NOOR_MESH_LEAF_MASK) {
        const uint32_t k = w & NOOR_COUNT_MASK;
        need = k;
        for (uint32_t j = 0; j < k && !*bad; ++j)
            need = std::max(need, (k - 1 - j) + stack_need(P, ni + 1 + j, memo, depth + 1, bad));
    } else if (w & NOOR_TRI_LEAF_MASK) {
        need = 0;
    } else {
        const uint32_t l = stack_need(P, ni + 1, memo, depth + 1, bad);
        const uint32_t r = stack_need(P, a.w, memo, depth + 1, bad);
        need = std::max(2u, 1u + std::max(l, r));
    }
    memo[ni] = (int32_t)need;
    return need;
}

const uint32_t kSupportedMaterials = NOOR_MAT_DIFFUSE | NOOR_MAT_TRANSLUCENT | NOOR_MAT_GLOSSY | NOOR_MAT_MIRROR |
                                     NOOR_MAT_GLASS | NOOR_MAT_METAL | NOOR_MAT_EMITTER | NOOR_MAT_ROUGHGLASS |
                                     NOOR_MAT_SUBSTRATE | NOOR_MAT_CLEARCOAT | NOOR_MAT_SHADOWCATCHER |
                                     NOOR_MAT_ALPHA | NOOR_MAT_BUMP;


noor_status validate(Context* c, const noor_payload_view* P) {
    if (P->num_triangles && (!P->attribute_indices || !P->wald_ax_nu_nv_nd || !P->wald_bnu_bnv_au_av ||
                             !P->wald_cnu_cnv_deg_mat))
        return fail(c, NOOR_ERR_INVALID_ARG, "triangle arrays missing");
    if (!P->bvh_min_right_start || !P->bvh_max_axis_count || P->num_bvh_nodes == 0)
        return fail(c, NOOR_ERR_INVALID_ARG, "BVH arrays missing");
    if (P->bvh_root_node >= P->num_bvh_nodes) return fail(c, NOOR_ERR_INVALID_PAYLOAD, "root out of range");
    if (P->num_vertices && (!P->vertices || !P->normals || !P->uvs))
        return fail(c, NOOR_ERR_INVALID_ARG, "vertex arrays missing");
    if (P->num_materials == 0 || !P->materials) return fail(c, NOOR_ERR_INVALID_ARG, "no materials");
    if (P->num_textures == 0 || !P->textures) return fail(c, NOOR_ERR_INVALID_ARG, "no textures");
    for (uint32_t i = 0; i < P->num_vertices; ++i) {
        const noor_float4 v = P->vertices[i];
        if (!std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z))
            return fail(c, NOOR_ERR_INVALID_PAYLOAD, "non-finite vertex " + std::to_string(i));
    }
    for (uint32_t t = 0; t < P->num_triangles; ++t) {
        const noor_uint4 a = P->attribute_indices[t];
        if (a.x >= P->num_vertices || a.y >= P->num_vertices || a.z >= P->num_vertices)
            return fail(c, NOOR_ERR_INVALID_PAYLOAD, "vertex index out of range in triangle " + std::to_string(t));
        if (a.w >= P->num_materials)
            return fail(c, NOOR_ERR_INVALID_PAYLOAD, "material index out of range in triangle " + std::to_string(t));
        if (P->wald_ax_nu_nv_nd[t].x > 2u) return fail(c, NOOR_ERR_INVALID_PAYLOAD, "bad Wald axis");
        // the alpha test reads the transparency map of the triangle's material
        if ((P->wald_cnu_cnv_deg_mat[t].w & NOOR_MAT_ALPHA) && !(P->materials[a.w].type & NOOR_MAT_ALPHA))
            return fail(c, NOOR_ERR_INVALID_PAYLOAD, "alpha-tested triangle of a material without a transparency map");
    }
    for (uint32_t n = 0; n < P->num_bvh_nodes; ++n) {
        const noor_uint4 a = P->bvh_min_right_start[n];
        const uint32_t w = P->bvh_max_axis_count[n].w;
        if (w & NOOR_LIGHT_NODE_MASK) {
            if ((w & NOOR_LIGHT_MASK) >= P->num_area_lights)
                return fail(c, NOOR_ERR_INVALID_PAYLOAD, "light node index out of range");
        } else if (w & NOOR_MESH_INSTANCE_MASK) {
            if ((w & NOOR_TRANS_MASK) >= P->num_instances || a.w >= P->num_bvh_nodes)
                return fail(c, NOOR_ERR_INVALID_PAYLOAD, "instance node out of range");
        } else if (w & NOOR_MESH_LEAF_MASK) {
            if ((uint64_t)n + (w & NOOR_COUNT_MASK) >= P->num_bvh_nodes)
                return fail(c, NOOR_ERR_INVALID_PAYLOAD, "mesh leaf overruns the node array");
        } else if (w & NOOR_TRI_LEAF_MASK) {
            if ((uint64_t)a.w + (w & NOOR_TRI_COUNT_MASK) > P->num_triangles)
                return fail(c, NOOR_ERR_INVALID_PAYLOAD, "triangle leaf out of range");
        } else {
            if ((w & NOOR_AXIS_MASK) > 2u || a.w >= P->num_bvh_nodes || n + 1 >= P->num_bvh_nodes)
                return fail(c, NOOR_ERR_INVALID_PAYLOAD, "inner node out of range");
        }
    }
    for (uint32_t m = 0; m < P->num_materials; ++m) {
        const noor_material& mt = P->materials[m];
        if ((mt.type & ~kSupportedMaterials) != 0 || mt.type == 0)
            return fail(c, NOOR_ERR_UNSUPPORTED, "material type " + std::to_string(mt.type) + " not supported");
        const uint32_t idx[] = {mt.diffuse_tex_idx, mt.specular_tex_idx, mt.transmission_tex_idx,
                                mt.emittance_tex_idx, mt.ior_tex_idx, mt.k_tex_idx, mt.roughness_tex_idx,
                                (mt.type & NOOR_MAT_BUMP) ? mt.bump_tex_idx : mt.diffuse_tex_idx,
                                (mt.type & NOOR_MAT_ALPHA) ? mt.transparency_tex_idx : mt.diffuse_tex_idx};
        for (uint32_t t : idx) {
            if (t >= P->num_textures) return fail(c, NOOR_ERR_INVALID_PAYLOAD, "texture index out of range");
            const noor_texture_view& tv = P->textures[t];
            if (!tv.rgba || tv.width == 0 || tv.height == 0)
                return fail(c, NOOR_ERR_INVALID_PAYLOAD, "empty material texture");
            if (tv.width > 1u << 15 || tv.height > 1u << 15)
                return fail(c, NOOR_ERR_UNSUPPORTED, "material textures are at most 32768 texels a side");
        }
        if ((mt.type & NOOR_MAT_BUMP) && !std::isfinite(mt.bumpfactor))
            return fail(c, NOOR_ERR_INVALID_PAYLOAD, "non-finite bump factor");
    }
    const noor_texture_view& env = P->textures[P->num_textures - 1];
    if (!env.rgba || env.width == 0 || env.height == 0) return fail(c, NOOR_ERR_INVALID_PAYLOAD, "no environment texture");
    if (env.width > 1u << 16 || env.height > 1u << 16)
        return fail(c, NOOR_ERR_UNSUPPORTED, "environment textures are at most 65536 texels a side");
    if (P->num_infinite_lights > 1) return fail(c, NOOR_ERR_UNSUPPORTED, "more than one infinite light");
    return NOOR_OK;
}

int nearest_pow2(int n) {
    int next = 1;
    while (next < n) next <<= 1;
    int prev = next;
    if (prev > n) prev >>= 1;
    if (prev < 1) prev = 1;
    return (std::abs(n - next) < std::abs(n - prev)) ? next : prev;
}

noor_status apply_spec(Context* c, const noor_spec* spec) {
    c->scene_dirty = true;
    c->spec = *spec;
    c->S.sky_enabled = (spec->lighting_type & NOOR_LIGHT_SKY_MASK) ? 1 : 0;
    c->S.skydome_type = spec->skydome_type;
    c->S.bounces = spec->bounces;
    c->S.rr = spec->rr;
    c->S.mis = spec->mis ? 1 : 0;
    c->S.bias = spec->reflection_bias;
    c->S.world_radius = spec->world_radius;
    if (spec->bvh_root_node != c->bvh_root) return fail(c, NOOR_ERR_INVALID_ARG, "spec root differs from payload root");
    return NOOR_OK;
}

noor_status apply_options(Context* c, const noor_options* o) {
    c->scene_dirty = true;
    noor_options def{NOOR_RNG_PCG32, NOOR_TRI_WATERTIGHT, NOOR_COMPAT_ALL, 0};
    c->opt = o ? *o : def;
    if (c->opt.rng != NOOR_RNG_PCG32 && c->opt.rng != NOOR_RNG_TEA_LCG) return fail(c, NOOR_ERR_INVALID_ARG, "rng");
    if (c->opt.tri_test != NOOR_TRI_WATERTIGHT && c->opt.tri_test != NOOR_TRI_WALD)
        return fail(c, NOOR_ERR_INVALID_ARG, "tri_test");
    c->S.rng_kind = c->opt.rng;
    c->S.tri_test = c->opt.tri_test;
    c->S.compat = c->opt.compat;
    return NOOR_OK;
}

noor_status apply_camera(Context* c, const noor_camera* cam) {
    c->scene_dirty = true;
    if (cam->w == 0 || cam->h == 0 || cam->w > 16384 || cam->h > 16384)
        return fail(c, NOOR_ERR_INVALID_ARG, "camera size");
    if (cam->type < 0 || cam->type > 2) return fail(c, NOOR_ERR_INVALID_ARG, "camera type");
    c->cam = *cam;
    c->S.cam = *cam;
    if (cam->w != c->film_w || cam->h != c->film_h) {
        if (c->film) {
            hipFree(c->film);
            c->film = nullptr;
        }
        void* p = nullptr;
        HIP_TRY(c, hipMalloc(&p, sizeof(float4) * (size_t)cam->w * cam->h));
        HIP_TRY(c, hipMemsetAsync(p, 0, sizeof(float4) * (size_t)cam->w * cam->h, c->stream));
        c->film = (float4*)p;
        c->film_w = cam->w;
        c->film_h = cam->h;
    }
    return NOOR_OK;
}

// Splits the traversal stack: `k` entries per lane in LDS, the rest of the
// BVH's exact bound in device memory (reallocated when the split changes).
noor_status set_stack_lds(Context* c, int k) {
    noor::DevScene& S = c->S;
    k = std::max(1, std::min(k, S.stack_depth));
    const size_t need = (size_t)(S.stack_depth - k) * S.gstack_lanes;
    if (c->gstack) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipFree(c->gstack));
        c->gstack = nullptr;
    }
    if (need) {
        void* p = nullptr;
        HIP_TRY(c, hipMalloc(&p, need * sizeof(uint32_t)));
        c->gstack = (uint32_t*)p;
    }
    S.gstack = c->gstack;
    S.stack_lds = k;
    c->scene_dirty = true;
    return NOOR_OK;
}

noor_status build_environment(Context* c) {
    c->scene_dirty = true;
    // PHYSICAL: bake sin(theta) * sky into a 2048^2 texture, then the CDF of it
    // (noor/cuda/skydome.cuh:58-71); HDR: the CDF of the environment texture.
    const float4* tex = c->S.env;
    int tw = (int)c->S.env_w, th = (int)c->S.env_h;
    if (c->S.skydome_type == NOOR_SKY_PHYSICAL) {
        if (!c->sky_bake) {
            void* p = nullptr;
            HIP_TRY(c, hipMalloc(&p, sizeof(float4) * 2048 * 2048));
            c->sky_bake = (float4*)p;
        }
        HIP_TRY(c, noor::launch_sky_bake(*c, c->sky_bake, 2048, 2048));
        tex = c->sky_bake;
        tw = th = 2048;
    }
    const int W = std::min(std::max(nearest_pow2(tw), 32), 1024);
    const int H = std::min(std::max(nearest_pow2(th), 32), 1024);
    const size_t n = (size_t)H * (W + 2) + 1;
    if (c->dfunc) hipFree(c->dfunc);
    if (c->dcdf) hipFree(c->dcdf);
    c->dfunc = c->dcdf = nullptr;
    void* p = nullptr;
    HIP_TRY(c, hipMalloc(&p, sizeof(float) * n));
    c->dfunc = (float*)p;
    HIP_TRY(c, hipMalloc(&p, sizeof(float) * n));
    c->dcdf = (float*)p;
    HIP_TRY(c, hipMemsetAsync(c->dfunc, 0, sizeof(float) * n, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->dcdf, 0, sizeof(float) * n, c->stream));
    HIP_TRY(c, noor::launch_dist_build(*c, tex, tw, th, c->dfunc, c->dcdf, W, H));
    c->S.dfunc = c->dfunc;
    c->S.dcdf = c->dcdf;
    c->S.dw = W;
    c->S.dh = H;
    return NOOR_OK;
}

}  // namespace

extern "C" {

const char* noor_status_string(noor_status s) {
    if ((int)s < 0 || (int)s > 8) return "unknown";
    return kStatus[(int)s];
}

int noor_abi_version(void) { return NOOR_ABI_VERSION; }

int noor_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* noor_last_error(const noor_ctx* ctx) { return ctx ? ctx->c.last_error.c_str() : ""; }

noor_status noor_create(const noor_payload_view* P, const noor_hosek* hosek, const noor_camera* camera,
                        const noor_spec* spec, const noor_options* options, noor_ctx** out) {
    if (!P || !camera || !spec || !out) return NOOR_ERR_INVALID_ARG;
    *out = nullptr;
    noor_ctx* ctx = new (std::nothrow) noor_ctx();
    if (!ctx) return NOOR_ERR_OUT_OF_MEMORY;
    Context* c = &ctx->c;
    auto bail = [&](noor_status s) {
        noor_destroy(ctx);
        return s;
    };
    noor_status st = validate(c, P);
    if (st != NOOR_OK) return bail(st);
    if (spec->skydome_type == NOOR_SKY_PHYSICAL && !hosek) return bail(NOOR_ERR_INVALID_ARG);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return bail(fail(c, NOOR_ERR_NO_DEVICE, "no HIP device"));
    if (hipGetDevice(&c->device) != hipSuccess) return bail(NOOR_ERR_NO_DEVICE);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(c, NOOR_ERR_HIP, "stream creation failed"));
    {
        // this device's copy of the no-allocation scratch (set up once per device)
        hipError_t se = hipSuccess;
        c->scratch = scratch_address(c->device, &se);
        if (!c->scratch) return bail(fail(c, NOOR_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(se)));
    }

    // walk of the binary BVH: detects out-of-range links and cycles
    std::vector<int32_t> memo(P->num_bvh_nodes, -1);
    bool bad = false;
    stack_need(P, P->bvh_root_node, memo, 0, &bad);
    if (bad) return bail(fail(c, NOOR_ERR_INVALID_PAYLOAD, "BVH references out of range or cyclic"));
    noor::DevScene& S = c->S;
    S.num_tris = P->num_triangles;
    c->bvh_root = P->bvh_root_node;

    // four-wide BVH for the traversal kernels
    {
        noor::WideBvh wb;
        const std::string werr = noor::build_wide_bvh(P, wb);
        if (!werr.empty()) return bail(fail(c, NOOR_ERR_UNSUPPORTED, "wide BVH: " + werr));
        c->stack_need = wb.stack_need;
        if (c->stack_need > 4096) return bail(fail(c, NOOR_ERR_STACK_DEPTH, "BVH needs " +
                                                  std::to_string(c->stack_need) + " stack entries, limit 4096"));
        S.stack_depth = (int)c->stack_need;
        S.stack_lds = std::min<int>(S.stack_depth, c->stack_lds);
        S.wroot = wb.root;
        if ((st = dev_upload(c, (noor::QNode**)&S.wbvh, wb.qnodes.data(), wb.qnodes.size())) != NOOR_OK) return bail(st);
        if ((st = dev_upload(c, (uint32_t**)&S.inst_root, wb.inst_root.data(), wb.inst_root.size())) != NOOR_OK)
            return bail(st);
        c->wide_nodes = wb.nodes.size();
        int cus = 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) cus = prop.multiProcessorCount;
        S.gstack_lanes = (uint32_t)cus * 32u * 64u;
        if ((st = set_stack_lds(c, c->stack_lds)) != NOOR_OK) return bail(st);
    }

    // triangles
    const uint32_t nt = P->num_triangles;
    if ((st = dev_upload(c, (noor_uint4**)&S.wald0, P->wald_ax_nu_nv_nd, nt)) != NOOR_OK) return bail(st);
    if ((st = dev_upload(c, (noor_uint4**)&S.wald1, P->wald_bnu_bnv_au_av, nt)) != NOOR_OK) return bail(st);
    if ((st = dev_upload(c, (noor_uint4**)&S.wald2, P->wald_cnu_cnv_deg_mat, nt)) != NOOR_OK) return bail(st);
    if ((st = dev_upload(c, (noor_uint4**)&S.attr, P->attribute_indices, nt)) != NOOR_OK) return bail(st);
    std::vector<uint32_t> ids(nt ? nt : 1);
    for (uint32_t i = 0; i < nt; ++i) ids[i] = P->prim_ids ? P->prim_ids[i] : i;
    if ((st = dev_upload(c, (uint32_t**)&S.prim_ids, ids.data(), ids.size())) != NOOR_OK) return bail(st);
    {
        // watertight records: the three vertices of each triangle in leaf order
        std::vector<noor_float4> tri((size_t)3 * (nt ? nt : 1));
        for (uint32_t i = 0; i < nt; ++i) {
            const noor_uint4 a = P->attribute_indices[i];
            tri[3 * (size_t)i + 0] = P->vertices[a.x];
            tri[3 * (size_t)i + 1] = P->vertices[a.y];
            tri[3 * (size_t)i + 2] = P->vertices[a.z];
        }
        if ((st = dev_upload(c, (noor_float4**)&S.tri, tri.data(), tri.size())) != NOOR_OK) return bail(st);
    }
    if ((st = dev_upload(c, (noor_float4**)&S.vert, P->vertices, P->num_vertices)) != NOOR_OK) return bail(st);
    if ((st = dev_upload(c, (noor_float4**)&S.nrm, P->normals, P->num_vertices)) != NOOR_OK) return bail(st);
    if ((st = dev_upload(c, (noor_float2**)&S.uv, P->uvs, P->num_vertices)) != NOOR_OK) return bail(st);
    {
        std::vector<noor_float4> xf((size_t)9 * (P->num_instances ? P->num_instances : 1));
        for (uint32_t i = 0; i < P->num_instances; ++i)
            for (int r = 0; r < 3; ++r) {
                xf[9 * (size_t)i + r] = P->world_to_object_row[r][i];
                xf[9 * (size_t)i + 3 + r] = P->object_to_world_row[r][i];
                xf[9 * (size_t)i + 6 + r] = P->normal_row[r][i];
            }
        if ((st = dev_upload(c, (noor_float4**)&S.xf, xf.data(), xf.size())) != NOOR_OK) return bail(st);
    }
    {
        c->num_materials = P->num_materials;
        std::vector<noor::DevMat> mats(P->num_materials);
        auto rgb = [&](uint32_t t, float* o) {
            const float* px = P->textures[t].rgba;
            o[0] = px[0];
            o[1] = px[1];
            o[2] = px[2];
        };
        // image textures (and every bump or transparency map) go to the texture
        // store with their mip chains; 1x1 textures of the other fields are
        // constants, read here once
        noor::TextureStore store;
        std::vector<int64_t> store_idx(P->num_textures, -1);
        auto stored = [&](uint32_t t) -> uint32_t {
            if (store_idx[t] < 0) {
                const noor_texture_view& tv = P->textures[t];
                store_idx[t] = noor::store_add_texture(store, tv.rgba, tv.width, tv.height,
                                                       ((options ? options->compat : NOOR_COMPAT_ALL) &
                                                        NOOR_COMPAT_RESIZE_SUM) != 0);
            }
            return (uint32_t)store_idx[t];
        };
        S.textured = 0;
        for (uint32_t m = 0; m < P->num_materials; ++m) {
            const noor_material& mt = P->materials[m];
            noor::DevMat& d = mats[m];
            std::memset(&d, 0, sizeof(d));
            const noor_float2 one{1.0f, 1.0f};
            const struct {
                uint32_t field, tex;
                noor_float2 uvs;
                bool always;
            } fields[] = {
                {noor::TX_DIFFUSE, mt.diffuse_tex_idx, mt.diffuse_uvscale, false},
                {noor::TX_SPECULAR, mt.specular_tex_idx, mt.specular_uvscale, false},
                {noor::TX_TRANSMISSION, mt.transmission_tex_idx, one, false},
                {noor::TX_EMITTANCE, mt.emittance_tex_idx, mt.emittance_uvscale, false},
                {noor::TX_ETA, mt.ior_tex_idx, one, false},
                {noor::TX_K, mt.k_tex_idx, one, false},
                {noor::TX_ROUGH, mt.roughness_tex_idx, mt.roughness_uvscale, false},
                {noor::TX_BUMP, mt.bump_tex_idx, mt.bump_uvscale, true},
                {noor::TX_ALPHA, mt.transparency_tex_idx, mt.transparency_uvscale, true},
            };
            for (const auto& f : fields) {
                if (f.field == noor::TX_BUMP && !(mt.type & NOOR_MAT_BUMP)) continue;
                if (f.field == noor::TX_ALPHA && !(mt.type & NOOR_MAT_ALPHA)) continue;
                const noor_texture_view& tv = P->textures[f.tex];
                if (!f.always && tv.width * tv.height == 1) continue;
                d.tex_mask |= 1u << f.field;
                d.tex[f.field] = stored(f.tex);
                d.uvs[f.field][0] = f.uvs.x;
                d.uvs[f.field][1] = f.uvs.y;
            }
            d.bumpfactor = mt.bumpfactor;
            d.emitter_scale = mt.emitter_scale;
            if (d.tex_mask) S.textured = 1;
            float tmp[3];
            rgb(mt.diffuse_tex_idx, d.diffuse);
            rgb(mt.specular_tex_idx, d.specular);
            rgb(mt.transmission_tex_idx, d.transmission);
            rgb(mt.emittance_tex_idx, tmp);
            for (int k = 0; k < 3; ++k) d.emittance[k] = mt.emitter_scale * tmp[k];
            rgb(mt.ior_tex_idx, d.eta);
            rgb(mt.k_tex_idx, d.k);
            rgb(mt.roughness_tex_idx, tmp);
            d.rough[0] = tmp[0];
            d.rough[1] = tmp[1];
            d.coat_ior = mt.coat_ior;
            d.coat_weight = mt.coat_weight;
            d.coat_thickness = mt.coat_thickness;
            d.coat_sigma = mt.coat_sigma;
            d.type = mt.type;
        }
        if ((st = dev_upload(c, (noor::DevMat**)&S.mats, mats.data(), mats.size())) != NOOR_OK) return bail(st);
        static_assert(sizeof(noor::TexLevels) == sizeof(noor::TexDesc), "texture descriptor layouts differ");
        if ((st = dev_upload(c, (float4**)&S.tex_texels, (const float4*)store.texels.data(), store.texels.size() / 4)) !=
            NOOR_OK)
            return bail(st);
        if ((st = dev_upload(c, (noor::TexDesc**)&S.tex_desc, (const noor::TexDesc*)store.desc.data(),
                             store.desc.size())) != NOOR_OK)
            return bail(st);
        c->texels = store.texels.size() / 4;
    }
    if ((st = dev_upload(c, (noor_area_light**)&S.area, P->area_lights, P->num_area_lights)) != NOOR_OK) return bail(st);
    if ((st = dev_upload(c, (noor_point_light**)&S.point, P->point_lights, P->num_point_lights)) != NOOR_OK) return bail(st);
    if ((st = dev_upload(c, (noor_spot_light**)&S.spot, P->spot_lights, P->num_spot_lights)) != NOOR_OK) return bail(st);
    if ((st = dev_upload(c, (noor_distant_light**)&S.distant, P->distant_lights, P->num_distant_lights)) != NOOR_OK)
        return bail(st);
    S.n_area = P->num_area_lights;
    S.n_point = P->num_point_lights;
    S.n_spot = P->num_spot_lights;
    S.n_distant = P->num_distant_lights;
    S.num_lights = (int)(P->num_area_lights + P->num_point_lights + P->num_spot_lights + P->num_distant_lights +
                         P->num_infinite_lights);
    {
        // the environment is a CudaMipMap too (noor/cuda/skydome.cuh:40-52): sides
        // resampled to the nearest power of two like the material textures
        const noor_texture_view& env = P->textures[P->num_textures - 1];
        const uint32_t ew = noor::nearest_pow2(env.width), eh = noor::nearest_pow2(env.height);
        std::vector<float> resized;
        const float* src = env.rgba;
        if (ew != env.width || eh != env.height) {
            resized = noor::lanczos_resize(env.rgba, env.width, env.height, ew, eh,
                                           ((options ? options->compat : NOOR_COMPAT_ALL) & NOOR_COMPAT_RESIZE_SUM) != 0);
            src = resized.data();
        }
        if ((st = dev_upload(c, (noor_float4**)&S.env, (const noor_float4*)src, (size_t)ew * eh)) != NOOR_OK)
            return bail(st);
        S.env_w = ew;
        S.env_h = eh;
    }
    noor_hosek hk{};
    if (hosek) hk = *hosek;
    c->hosek_host = hk;
    if ((st = dev_upload(c, (noor_hosek**)&c->d_hosek, &hk, 1)) != NOOR_OK) return bail(st);
    S.hosek = c->d_hosek;
    if ((st = dev_upload(c, (noor::DevScene**)&c->d_scene, (const noor::DevScene*)nullptr, 1)) != NOOR_OK) return bail(st);
    if ((st = dev_upload(c, (noor::DevStats**)&c->d_stats, (const noor::DevStats*)nullptr, 1)) != NOOR_OK) return bail(st);
    if ((st = dev_upload(c, (float4**)&c->d_lookat, (const float4*)nullptr, 1)) != NOOR_OK) return bail(st);
    if ((st = apply_options(c, options)) != NOOR_OK) return bail(st);
    if ((st = apply_spec(c, spec)) != NOOR_OK) return bail(st);
    if ((st = apply_camera(c, camera)) != NOOR_OK) return bail(st);
    if ((st = build_environment(c)) != NOOR_OK) return bail(st);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(fail(c, NOOR_ERR_HIP, "upload failed"));
    *out = ctx;
    return NOOR_OK;
}

noor_status noor_destroy(noor_ctx* ctx) {
    if (!ctx) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    if (c->stream) hipStreamSynchronize(c->stream);
    noor::wave_free(*c);
    for (void* p : c->allocs) hipFree(p);
    if (c->film) hipFree(c->film);
    if (c->gstack) hipFree(c->gstack);
    if (c->batch_film) hipFree(c->batch_film);
    if (c->sky_bake) hipFree(c->sky_bake);
    if (c->sample_lo) hipFree(c->sample_lo);
    if (c->dfunc) hipFree(c->dfunc);
    if (c->dcdf) hipFree(c->dcdf);
    if (c->stream) hipStreamDestroy(c->stream);
    delete ctx;
    return NOOR_OK;
}

noor_status noor_update_camera(noor_ctx* ctx, const noor_camera* camera) {
    if (!ctx || !camera) return NOOR_ERR_INVALID_ARG;
    return apply_camera(&ctx->c, camera);
}

noor_status noor_update_hosek(noor_ctx* ctx, const noor_hosek* hosek) {
    if (!ctx || !hosek) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    c->hosek_host = *hosek;
    HIP_TRY(c, hipMemcpyAsync(c->d_hosek, hosek, sizeof(noor_hosek), hipMemcpyHostToDevice, c->stream));
    // a new sky needs a new bake and distribution (noor/cuda/skydome.cuh:58-71)
    if (c->S.skydome_type == NOOR_SKY_PHYSICAL) {
        noor_status st = build_environment(c);
        if (st != NOOR_OK) return st;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NOOR_OK;
}

noor_status noor_update_spec(noor_ctx* ctx, const noor_spec* spec) {
    if (!ctx || !spec) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    const int old_sky = c->S.skydome_type;
    noor_status st = apply_spec(c, spec);
    if (st != NOOR_OK) return st;
    if (old_sky != c->S.skydome_type) {
        st = build_environment(c);
        if (st != NOOR_OK) return st;
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return NOOR_OK;
}

noor_status noor_update_options(noor_ctx* ctx, const noor_options* options) {
    if (!ctx || !options) return NOOR_ERR_INVALID_ARG;
    return apply_options(&ctx->c, options);
}

noor_status noor_render(noor_ctx* ctx, uint32_t* frame, uint32_t spp) {
    if (!ctx || !frame || *frame < 1 || spp == 0) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    HIP_TRY(c, noor::launch_megakernel(*c, *frame, spp));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->timer.collect();
    *frame += spp;
    return NOOR_OK;
}

noor_status noor_set_timing(noor_ctx* ctx, int enable) {
    if (!ctx) return NOOR_ERR_INVALID_ARG;
    ctx->c.timer.enabled = enable != 0;
    return NOOR_OK;
}

noor_status noor_get_kernel_times(noor_ctx* ctx, noor_kernel_time* out, uint32_t max, uint32_t* count) {
    if (!ctx || !count) return NOOR_ERR_INVALID_ARG;
    uint32_t n = 0;
    for (const auto& kv : ctx->c.timer.totals) {
        if (out && n < max) {
            std::memset(&out[n], 0, sizeof(out[n]));
            std::strncpy(out[n].name, kv.first.c_str(), sizeof(out[n].name) - 1);
            out[n].launches = kv.second.first;
            out[n].total_ms = kv.second.second;
        }
        ++n;
    }
    *count = n;
    return NOOR_OK;
}

noor_status noor_reset_kernel_times(noor_ctx* ctx) {
    if (!ctx) return NOOR_ERR_INVALID_ARG;
    ctx->c.timer.totals.clear();
    return NOOR_OK;
}

noor_status noor_set_wavefront_params(noor_ctx* ctx, uint32_t queue_capacity, int32_t block) {
    if (!ctx) return NOOR_ERR_INVALID_ARG;
    if (block != 64 && block != 128 && block != 256) return NOOR_ERR_INVALID_ARG;
    if (queue_capacity != 0 && (queue_capacity < 64 || queue_capacity > (1u << 28))) return NOOR_ERR_INVALID_ARG;
    ctx->c.wave_capacity = queue_capacity;
    ctx->c.wave_block = block;
    return NOOR_OK;
}

noor_status noor_render_wavefront(noor_ctx* ctx, uint32_t* frame, uint32_t spp) {
    if (!ctx || !frame || *frame < 1 || spp == 0) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    HIP_TRY(c, noor::wave_render(*c, *frame, spp));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->timer.collect();
    *frame += spp;
    return NOOR_OK;
}

noor_status noor_set_stack_lds(noor_ctx* ctx, uint32_t entries) {
    if (!ctx) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    c->stack_lds = entries ? (int)std::min<uint32_t>(entries, 4096u) : 24;
    return set_stack_lds(c, c->stack_lds);
}

namespace {
noor_status check_views(Context* c, const noor_view* views, uint32_t n) {
    if (!views || n == 0) return fail(c, NOOR_ERR_INVALID_ARG, "no views");
    uint64_t pixels = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const noor_view& v = views[i];
        if (v.camera.w == 0 || v.camera.h == 0 || v.camera.w > 16384 || v.camera.h > 16384)
            return fail(c, NOOR_ERR_INVALID_ARG, "view " + std::to_string(i) + ": camera size");
        if (v.camera.type < 0 || v.camera.type > 2)
            return fail(c, NOOR_ERR_INVALID_ARG, "view " + std::to_string(i) + ": camera type");
        if (v.frame < 1 || v.spp == 0) return fail(c, NOOR_ERR_INVALID_ARG, "view " + std::to_string(i) + ": frame/spp");
        pixels += (uint64_t)v.camera.w * v.camera.h;
    }
    if (pixels >= (1ull << 31)) return fail(c, NOOR_ERR_INVALID_ARG, "batch larger than 2^31 pixels");
    return NOOR_OK;
}

noor_status batch_films(Context* c, const noor_view* views, uint32_t n) {
    bool same = c->batch_film && c->batch_w.size() == n;
    for (uint32_t i = 0; same && i < n; ++i)
        same = c->batch_w[i] == views[i].camera.w && c->batch_h[i] == views[i].camera.h;
    if (same) return NOOR_OK;
    if (c->batch_film) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipFree(c->batch_film));
        c->batch_film = nullptr;
    }
    c->batch_w.resize(n);
    c->batch_h.resize(n);
    c->batch_off.resize(n);
    size_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
        c->batch_w[i] = views[i].camera.w;
        c->batch_h[i] = views[i].camera.h;
        c->batch_off[i] = off;
        off += (size_t)views[i].camera.w * views[i].camera.h;
    }
    void* p = nullptr;
    HIP_TRY(c, hipMalloc(&p, off * sizeof(float4)));
    HIP_TRY(c, hipMemsetAsync(p, 0, off * sizeof(float4), c->stream));
    c->batch_film = (float4*)p;
    return NOOR_OK;
}
}  // namespace

noor_status noor_render_batch(noor_ctx* ctx, const noor_view* views, uint32_t n) {
    if (!ctx) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    noor_status st = check_views(c, views, n);
    if (st != NOOR_OK) return st;
    if ((st = batch_films(c, views, n)) != NOOR_OK) return st;
    HIP_TRY(c, noor::wave_render_batch(*c, views, n, c->batch_film, c->batch_off.data()));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->timer.collect();
    return NOOR_OK;
}

noor_status noor_read_batch_film(noor_ctx* ctx, uint32_t view, float* rgba, size_t count_floats) {
    if (!ctx || !rgba) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    if (view >= c->batch_w.size()) return fail(c, NOOR_ERR_INVALID_ARG, "no such batch view");
    const size_t need = (size_t)4 * c->batch_w[view] * c->batch_h[view];
    if (count_floats < need) return fail(c, NOOR_ERR_INVALID_ARG, "buffer too small");
    HIP_TRY(c, hipMemcpyAsync(rgba, c->batch_film + c->batch_off[view], need * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NOOR_OK;
}

noor_status noor_batch_film_device_ptr(noor_ctx* ctx, uint32_t view, void** ptr) {
    if (!ctx || !ptr) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    if (view >= c->batch_w.size()) return fail(c, NOOR_ERR_INVALID_ARG, "no such batch view");
    *ptr = c->batch_film + c->batch_off[view];
    return NOOR_OK;
}

noor_status noor_autotune(noor_ctx* ctx, const noor_view* views, uint32_t n, noor_tune_entry* out, uint32_t max,
                          uint32_t* count) {
    if (!ctx) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    noor_status st = check_views(c, views, n);
    if (st != NOOR_OK) return st;
    if ((st = batch_films(c, views, n)) != NOOR_OK) return st;
    const int blocks[3] = {64, 128, 256};
    // pools of 2^20 .. 2^25 paths; larger ones than the batch's items add nothing
    uint64_t items = 0;
    for (uint32_t i = 0; i < n; ++i) items += (uint64_t)views[i].camera.w * views[i].camera.h * views[i].spp;
    std::vector<uint32_t> caps;
    for (uint32_t cap = 1u << 20; cap <= (1u << 25); cap <<= 1) {
        caps.push_back(cap);
        if (cap >= items) break;
    }
    std::vector<int> ks;
    for (int k : {16, 24, 32}) {
        const int kk = std::min(k, c->S.stack_depth);
        if (std::find(ks.begin(), ks.end(), kk) == ks.end()) ks.push_back(kk);
    }
    const bool timing = c->timer.enabled;
    c->timer.enabled = false;
    std::vector<noor_tune_entry> res;
    noor_tune_entry best{0, 64, 24, 1e30f};
    for (int k : ks) {
        if ((st = set_stack_lds(c, k)) != NOOR_OK) break;
        for (int b : blocks)
            for (uint32_t cap : caps) {
                c->wave_block = b;
                c->wave_capacity = cap;
                // first render allocates and captures; the second is timed
                hipError_t e = noor::wave_render_batch(*c, views, n, c->batch_film, c->batch_off.data());
                if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
                const auto t0 = std::chrono::steady_clock::now();
                if (e == hipSuccess) e = noor::wave_render_batch(*c, views, n, c->batch_film, c->batch_off.data());
                if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
                const auto t1 = std::chrono::steady_clock::now();
                if (e != hipSuccess) {
                    // a setting that does not fit (memory) is skipped
                    (void)hipGetLastError();
                    continue;
                }
                const float ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
                const noor_tune_entry r{cap, b, (uint32_t)k, ms};
                res.push_back(r);
                if (ms < best.ms) best = r;
            }
    }
    c->timer.enabled = timing;
    if (res.empty()) return fail(c, NOOR_ERR_HIP, "no wavefront setting could run");
    c->wave_block = best.block;
    c->wave_capacity = best.queue_capacity;
    c->stack_lds = (int)best.stack_lds;
    if ((st = set_stack_lds(c, c->stack_lds)) != NOOR_OK) return st;
    if (count) *count = (uint32_t)std::min<size_t>(res.size(), max);
    if (out)
        for (uint32_t i = 0; i < std::min<size_t>(res.size(), max); ++i) out[i] = res[i];
    return NOOR_OK;
}

noor_status noor_set_params_device(noor_ctx* ctx, const float* albedo, const float* roughness,
                                   uint32_t num_materials, const float* env_rgb, uint32_t env_w, uint32_t env_h) {
    if (!ctx) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    if ((albedo || roughness) && num_materials > c->num_materials)
        return fail(c, NOOR_ERR_INVALID_ARG, "more parameter materials than the payload holds");
    if (env_rgb) {
        if (c->S.skydome_type != NOOR_SKY_HDR || !c->S.env)
            return fail(c, NOOR_ERR_UNSUPPORTED, "environment parameters need an HDR environment");
        if (env_w != c->S.env_w || env_h != c->S.env_h) return fail(c, NOOR_ERR_INVALID_ARG, "environment size");
    }
    HIP_TRY(c, noor::launch_set_params(*c, albedo, roughness, num_materials, env_rgb));
    if (env_rgb) {
        noor_status st = build_environment(c);
        if (st != NOOR_OK) return st;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NOOR_OK;
}

noor_status noor_render_sample(noor_ctx* ctx, uint32_t stream, float* radiance) {
    if (!ctx || !radiance || stream < 1) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    const size_t px = (size_t)c->S.cam.w * c->S.cam.h;
    if (c->sample_lo_px != px) {
        if (c->sample_lo) hipFree(c->sample_lo);
        c->sample_lo = nullptr;
        c->sample_lo_px = 0;
        HIP_TRY(c, hipMalloc((void**)&c->sample_lo, sizeof(float4) * px));
        c->sample_lo_px = px;
    }
    c->sample_lo_of = nullptr;
    HIP_TRY(c, noor::launch_megakernel(*c, stream, 1, (float4*)radiance, c->sample_lo));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->sample_lo_of = radiance;
    c->sample_lo_stream = stream;
    c->timer.collect();
    return NOOR_OK;
}

noor_status noor_adjoint(noor_ctx* ctx, uint32_t stream, const float* radiance, const float* grad_image,
                         float* grad_albedo, float* grad_roughness, uint32_t num_materials, float* grad_env) {
    if (!ctx || !radiance || !grad_image || stream < 1) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    if (num_materials && (!grad_albedo || !grad_roughness)) return NOOR_ERR_INVALID_ARG;
    if (num_materials > c->num_materials) return fail(c, NOOR_ERR_INVALID_ARG, "num_materials");
    if (c->S.mis) return fail(c, NOOR_ERR_UNSUPPORTED, "the adjoint does not implement MIS");
    if (c->S.textured) return fail(c, NOOR_ERR_UNSUPPORTED, "the adjoint does not implement image textures");
    if (c->S.compat & NOOR_COMPAT_CLAMP) return fail(c, NOOR_ERR_UNSUPPORTED, "the adjoint needs the clamp off");
    // the forward's rounding errors when `radiance` is the last sample rendered
    const float4* lo = (c->sample_lo && c->sample_lo_of == radiance && c->sample_lo_stream == stream &&
                        c->sample_lo_px == (size_t)c->S.cam.w * c->S.cam.h)
                           ? c->sample_lo
                           : nullptr;
    HIP_TRY(c, noor::launch_adjoint(*c, stream, radiance, lo, grad_image, grad_albedo, grad_roughness, num_materials,
                                    grad_env));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->timer.collect();
    return NOOR_OK;
}

noor_status noor_wide_bvh_check(const noor_payload_view* P, uint32_t* num_nodes, uint32_t* stack_need_out) {
    if (!P) return NOOR_ERR_INVALID_ARG;
    Context tmp;
    noor_status st = validate(&tmp, P);
    if (st != NOOR_OK) return st;
    std::vector<int32_t> memo(P->num_bvh_nodes, -1);
    bool bad = false;
    stack_need(P, P->bvh_root_node, memo, 0, &bad);
    if (bad) return NOOR_ERR_INVALID_PAYLOAD;
    noor::WideBvh wb;
    if (!noor::build_wide_bvh(P, wb).empty()) return NOOR_ERR_UNSUPPORTED;
    if (!noor::check_wide_bvh(P, wb).empty()) return NOOR_ERR_INVALID_PAYLOAD;
    if (num_nodes) *num_nodes = (uint32_t)wb.nodes.size();
    if (stack_need_out) *stack_need_out = wb.stack_need;
    return NOOR_OK;
}

noor_status noor_debug_skydome(noor_ctx* ctx, uint32_t* frame) {
    if (!ctx || !frame) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    HIP_TRY(c, noor::launch_debug_skydome(*c, *frame));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    ++*frame;
    return NOOR_OK;
}

noor_status noor_get_lookat(noor_ctx* ctx, noor_float4* lookat) {
    if (!ctx || !lookat) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    HIP_TRY(c, hipMemcpyAsync(lookat, c->d_lookat, sizeof(noor_float4), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NOOR_OK;
}

noor_status noor_read_film(noor_ctx* ctx, float* rgba, size_t count_floats) {
    if (!ctx || !rgba) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    const size_t need = (size_t)4 * c->film_w * c->film_h;
    if (count_floats < need) return fail(c, NOOR_ERR_INVALID_ARG, "film buffer too small");
    HIP_TRY(c, hipMemcpyAsync(rgba, c->film, need * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NOOR_OK;
}

noor_status noor_film_device_ptr(noor_ctx* ctx, void** ptr) {
    if (!ctx || !ptr) return NOOR_ERR_INVALID_ARG;
    *ptr = ctx->c.film;
    return NOOR_OK;
}

noor_status noor_get_stats(noor_ctx* ctx, noor_stats* out) {
    if (!ctx || !out) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    noor::DevStats d{};
    HIP_TRY(c, hipMemcpyAsync(&d, c->d_stats, sizeof(d), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::memset(out, 0, sizeof(*out));
    for (const auto& sl : d.slot) {
        out->extension_rays += sl.extension_rays;
        out->shadow_rays += sl.shadow_rays;
        out->drain_extension_rays += sl.drain_extension_rays;
        out->drain_shadow_rays += sl.drain_shadow_rays;
    }
    out->node_visits = d.ext[0];
    out->tri_tests = d.ext[1];
    out->tri_second_fetch = d.ext[2];
    out->instance_entries = d.ext[3];
    out->shadow_node_visits = d.shd[0];
    out->shadow_tri_tests = d.shd[1];
    out->shadow_tri_second_fetch = d.shd[2];
    out->shadow_instance_entries = d.shd[3];
    out->extension_hits = d.ext_hits;
    if (d.overflow) return fail(c, NOOR_ERR_STACK_DEPTH, "traversal stack overflow");
    return NOOR_OK;
}

noor_status noor_reset_stats(noor_ctx* ctx) {
    if (!ctx) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    HIP_TRY(c, hipMemsetAsync(c->d_stats, 0, sizeof(noor::DevStats), c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return NOOR_OK;
}

noor_status noor_synchronize(noor_ctx* ctx) {
    if (!ctx) return NOOR_ERR_INVALID_ARG;
    HIP_TRY(&ctx->c, hipStreamSynchronize(ctx->c.stream));
    return NOOR_OK;
}

noor_status noor_trace(noor_ctx* ctx, const noor_ray* rays, uint32_t n, int any_hit, noor_hit* hits) {
    if (!ctx || !rays || !hits) return NOOR_ERR_INVALID_ARG;
    if (n == 0) return NOOR_OK;
    Context* c = &ctx->c;
    void *dr = nullptr, *dh = nullptr;
    HIP_TRY(c, hipMalloc(&dr, sizeof(noor_ray) * n));
    hipError_t e = hipMalloc(&dh, sizeof(noor_hit) * n);
    if (e != hipSuccess) {
        hipFree(dr);
        return fail(c, NOOR_ERR_OUT_OF_MEMORY, "hit buffer");
    }
    e = hipMemcpyAsync(dr, rays, sizeof(noor_ray) * n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = noor::launch_trace(*c, (const noor_ray*)dr, n, any_hit, (noor_hit*)dh);
    if (e == hipSuccess) e = hipMemcpyAsync(hits, dh, sizeof(noor_hit) * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(dr);
    hipFree(dh);
    if (e != hipSuccess) return fail(c, NOOR_ERR_HIP, std::string("trace: ") + hipGetErrorString(e));
    return NOOR_OK;
}

noor_status noor_trace_noalloc(noor_ctx* ctx, const noor_ray* d_rays, uint32_t n, int any_hit, noor_hit* d_hits) {
    if (!ctx || (n && (!d_rays || !d_hits))) return NOOR_ERR_INVALID_ARG;
    if (n == 0) return NOOR_OK;
    if (n > (1u << 31)) return NOOR_ERR_INVALID_ARG;
    Context* c = &ctx->c;
    if (!c->scratch) return fail(c, NOOR_ERR_HIP, "trace_noalloc: no scratch on this context's device");
    // the lanes whose deep stack entries fit the scratch
    const size_t deep = (size_t)std::max(0, c->S.stack_depth - c->S.stack_lds);
    uint32_t lanes = c->S.gstack_lanes;
    if (deep) lanes = (uint32_t)std::min<size_t>(lanes, kScratchWords / deep);
    lanes &= ~63u;
    if (lanes < 64)
        return fail(c, NOOR_ERR_UNSUPPORTED,
                    "trace_noalloc: the BVH needs " + std::to_string(deep) +
                        " device stack entries per lane, more than the 16 MiB scratch holds for one wave "
                        "(raise the LDS part with noor_set_stack_lds)");
    HIP_TRY(c, noor::launch_trace(*c, d_rays, n, any_hit, d_hits, c->scratch, lanes));
    return NOOR_OK;
}

size_t noor_abi_struct_size(const char* name) {
    if (!name) return 0;
    const std::string n(name);
    if (n == "noor_camera") return sizeof(noor_camera);
    if (n == "noor_hosek") return sizeof(noor_hosek);
    if (n == "noor_spec") return sizeof(noor_spec);
    if (n == "noor_material") return sizeof(noor_material);
    if (n == "noor_shape") return sizeof(noor_shape);
    if (n == "noor_area_light") return sizeof(noor_area_light);
    if (n == "noor_payload_view") return sizeof(noor_payload_view);
    if (n == "noor_options") return sizeof(noor_options);
    if (n == "noor_stats") return sizeof(noor_stats);
    if (n == "noor_ray") return sizeof(noor_ray);
    if (n == "noor_hit") return sizeof(noor_hit);
    if (n == "noor_transform") return sizeof(noor_transform);
    if (n == "noor_hosek_state") return sizeof(noor_hosek_state);
    if (n == "noor_point_light") return sizeof(noor_point_light);
    if (n == "noor_spot_light") return sizeof(noor_spot_light);
    if (n == "noor_distant_light") return sizeof(noor_distant_light);
    if (n == "noor_texture_view") return sizeof(noor_texture_view);
    if (n == "noor_kernel_time") return sizeof(noor_kernel_time);
    if (n == "noor_mesh_desc") return sizeof(noor_mesh_desc);
    if (n == "noor_obj_material") return sizeof(noor_obj_material);
    if (n == "noor_obj_view") return sizeof(noor_obj_view);
    if (n == "noor_instance_desc") return sizeof(noor_instance_desc);
    if (n == "noor_bvh_params") return sizeof(noor_bvh_params);
    if (n == "noor_scene_desc") return sizeof(noor_scene_desc);
    if (n == "noor_build_stats") return sizeof(noor_build_stats);
    if (n == "noor_view") return sizeof(noor_view);
    if (n == "noor_tune_entry") return sizeof(noor_tune_entry);
    return 0;
}

}  // extern "C"
