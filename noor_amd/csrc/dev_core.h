// Device library of the gfx950 render path: scene layout in HBM, RNG,
// two-level BVH traversal, triangle tests, BSDF lobes, lights, sky and
// environment sampling, camera rays and the per-bounce shading step.
//
// The megakernel and the wavefront kernels both call shade_bounce(), so their
// per-path arithmetic is identical; the CPU oracle restates the same
// reference code with the same operation order, which is what the parity
// tests compare.  Reference citations are given per function.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/noor_hip.h"
#include "dev_math.h"
#include "wide_bvh.h"

namespace noor {

// ---------------------------------------------------------------------------------------------
// constants (noor/cuda/path_tracer.cuh:65-75)
#define K_PI 3.141592654f
#define K_2PI (2.0f * K_PI)
#define K_INV_PI (1.0f / K_PI)
#define K_INV_2PI (1.0f / K_2PI)
#define K_PI_OVER_2 (K_PI / 2.0f)
#define K_PI_OVER_4 (K_PI / 4.0f)
#define K_ONE_MINUS_EPS (1.0f - 1.1920928955078125e-07f)
#define K_INF __builtin_huge_valf()

// Address-space qualified pointers: loads through them compile to global
// (or LDS) instructions instead of flat ones.  Device allocations are global.
#define NOOR_GLOBAL __attribute__((address_space(1)))
#define NOOR_LDS __attribute__((address_space(3)))
typedef uint32_t noor_v4u __attribute__((ext_vector_type(4)));
typedef uint32_t noor_v2u __attribute__((ext_vector_type(2)));
typedef float noor_v4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 ldg(const uint4* p) {
    const noor_v4u v = *(const NOOR_GLOBAL noor_v4u*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 ldg(const uint2* p) {
    const noor_v2u v = *(const NOOR_GLOBAL noor_v2u*)p;
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ float4 ldg(const float4* p) {
    const noor_v4f v = *(const NOOR_GLOBAL noor_v4f*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t ldg(const uint32_t* p) { return *(const NOOR_GLOBAL uint32_t*)p; }
__device__ __forceinline__ float ldg(const float* p) { return *(const NOOR_GLOBAL float*)p; }
typedef float noor_v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ldg(const float2* p) {
    const noor_v2f v = *(const NOOR_GLOBAL noor_v2f*)p;
    return make_float2(v.x, v.y);
}

// BxDFType (noor/cuda/path_tracer.cuh:104-117)
enum : int { B_REFLECTION = 1, B_TRANSMISSION = 2, B_DIFFUSE = 4, B_GLOSSY = 8, B_SPECULAR = 16,
             B_CONDUCTOR = 32, B_DIELECTRIC = 64, B_NOOP = 128, B_ALL = 255 };
enum : int { L_SHADOWCATCHER = 0, L_LAMBERT_R = 1, L_SPEC_R_NOOP = 2, L_MICRO_R_DIEL = 4, L_MICRO_R_COND = 5,
             L_LAMBERT_T = 6, L_MICRO_T = 8, L_FRESNEL_BLEND = 9, L_FRESNEL_SPEC = 10, L_CLEARCOAT = 11 };

constexpr uint32_t M_EMITTER_ANY = NOOR_MAT_EMITTER | NOOR_MAT_MESHLIGHT;
constexpr uint32_t M_TRANSPARENT = NOOR_MAT_TRANSLUCENT | NOOR_MAT_GLASS | NOOR_MAT_ROUGHGLASS;
constexpr uint32_t M_SPECULAR = NOOR_MAT_MIRROR | NOOR_MAT_GLASS | NOOR_MAT_CLEARCOAT;
// materials whose spawned rays carry differentials (noor/cuda/path_tracer.cuh:125-126)
constexpr uint32_t M_GLOSSY = M_SPECULAR | NOOR_MAT_ROUGHGLASS | NOOR_MAT_CLEARCOAT | NOOR_MAT_METAL |
                              NOOR_MAT_SUBSTRATE | NOOR_MAT_GLOSSY;

// ---------------------------------------------------------------------------------------------
// vectors (CUDA-samples helper_math semantics, IEEE ops, no contraction)
struct f3 {
    float x, y, z;
};
struct f2 {
    float x, y;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul3(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 scl3(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 sdiv3(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 div3(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ f3 neg3(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ f3 nrm3(f3 a) {
    const float inv = 1.0f / sqrtf(dot3(a, a));
    return scl3(a, inv);
}
__device__ __forceinline__ f3 nrm3s(f3 a) {
    const float l = dot3(a, a);
    if (l == 0.0f) return a;
    const float inv = 1.0f / sqrtf(l);
    return scl3(a, inv);
}
__device__ __forceinline__ float absdot3(f3 a, f3 b) { return fabsf(dot3(a, b)); }
__device__ __forceinline__ bool is_black(f3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
__device__ __forceinline__ float rmax(float a, float b) { return a > b ? a : (b == b ? b : a); }
__device__ __forceinline__ float rmin(float a, float b) { return a < b ? a : (b == b ? b : a); }
__device__ __forceinline__ float clampr(float v, float a, float b) { return rmax(a, rmin(v, b)); }
__device__ __forceinline__ float maxcomp3(f3 f) { return rmax(rmax(f.x, f.y), f.z); }
__device__ __forceinline__ int isign(float v) { return (0.0f < v) - (v < 0.0f); }
__device__ __forceinline__ bool is_inf(float v) { return fabsf(v) == K_INF; }
__device__ __forceinline__ float c3(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
__device__ __forceinline__ f3 ld3(const float* f) { return f3{f[0], f[1], f[2]}; }
__device__ __forceinline__ f3 fromf3(noor_float3 a) { return f3{a.x, a.y, a.z}; }
__device__ __forceinline__ f3 fromf4(float4 a) { return f3{a.x, a.y, a.z}; }

// ---------------------------------------------------------------------------------------------
// scene layout in device memory

// 1x1 material textures resolved at upload (noor/cuda/materialimp.cuh:134-214).
// Material fields that can read an image texture (noor/cuda/materialimp.cuh:120-186)
enum : uint32_t { TX_DIFFUSE = 0, TX_SPECULAR, TX_TRANSMISSION, TX_EMITTANCE, TX_ETA, TX_K, TX_ROUGH, TX_BUMP,
                  TX_ALPHA, TX_COUNT };

struct DevMat {
    float diffuse[3], specular[3], transmission[3], emittance[3], eta[3], k[3];
    float rough[2];
    float coat_ior;
    float coat_weight, coat_thickness, coat_sigma;
    uint32_t type;
    // image textures: bit f of tex_mask set when field f reads texture tex[f]
    // at uvscale uvs[f] (the constants above are then unused)
    uint32_t tex_mask;
    uint32_t tex[TX_COUNT];
    float uvs[TX_COUNT][2];
    float bumpfactor, emitter_scale;
};

// A texture's mip chain in the texel store: level l is max(1, w >> l) by
// max(1, h >> l) RGBA texels, the levels packed from `offset` (w, h powers
// of two).
struct TexDesc {
    uint32_t offset, w, h, levels;
};

struct DevScene {
    // four-wide BVH derived from the payload's binary BVH (wide_bvh.h): 8 uint4 per node
    const uint4* __restrict__ wbvh;
    const uint32_t* __restrict__ inst_root;  // per transform: reference of its BLAS root
    uint32_t* gstack;                        // traversal stack entries beyond the LDS part
    // triangle records in leaf order
    const uint4* __restrict__ wald0;
    const uint4* __restrict__ wald1;
    const uint4* __restrict__ wald2;
    const float4* __restrict__ tri;      // watertight: v0, v1, v2 per triangle
    const uint4* __restrict__ attr;
    const uint32_t* __restrict__ prim_ids;
    const float4* __restrict__ vert;
    const float4* __restrict__ nrm;
    const float2* __restrict__ uv;
    const float4* __restrict__ xf;       // 9 rows per instance: w2o, o2w, normal
    const DevMat* __restrict__ mats;
    const float4* __restrict__ tex_texels;  // mip chains of the material textures
    const TexDesc* __restrict__ tex_desc;   // per texture of the store
    uint32_t textured;                      // any image texture, bump or alpha material
    const noor_area_light* __restrict__ area;
    const noor_point_light* __restrict__ point;
    const noor_spot_light* __restrict__ spot;
    const noor_distant_light* __restrict__ distant;
    const float4* __restrict__ env;
    const float* __restrict__ dfunc;
    const float* __restrict__ dcdf;
    const noor_hosek* __restrict__ hosek;
    uint32_t wroot;          // reference of the root
    uint32_t gstack_lanes;   // lanes the global stack part is laid out for
    uint32_t n_area, n_point, n_spot, n_distant;
    int32_t num_lights;
    uint32_t env_w, env_h;
    int32_t dw, dh;
    int32_t sky_enabled, skydome_type, bounces, rr, mis;
    int32_t rng_kind, tri_test;
    uint32_t compat;
    int32_t stack_depth;     // entries a traversal can need (exact bound for this BVH)
    int32_t stack_lds;       // of which the first stack_lds live in LDS
    float bias, world_radius;
    uint32_t num_tris;
    noor_camera cam;
};

// ---------------------------------------------------------------------------------------------
// RNG: TEA-seeded LCG (noor/cuda/rng.cuh:37-76) or PCG32 (O'Neill, pcg-random.org)
struct Rng {
    uint64_t state;  // PCG32 state, or the LCG state in the low 32 bits
    uint64_t inc;
};

__device__ __forceinline__ uint32_t tea16(uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
#pragma unroll
    for (int n = 0; n < 16; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
__device__ __forceinline__ uint32_t pcg32_step(uint64_t& state, uint64_t inc) {
    const uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
__device__ __forceinline__ void rng_init(Rng& r, int kind, uint32_t gid, uint32_t frame) {
    if (kind == NOOR_RNG_PCG32) {
        r.state = 0;
        r.inc = ((uint64_t)frame << 1u) | 1u;
        pcg32_step(r.state, r.inc);
        r.state += (uint64_t)gid;
        pcg32_step(r.state, r.inc);
    } else {
        r.state = tea16(gid, frame);
        r.inc = 0;
    }
}
__device__ __forceinline__ float rng_next(Rng& r, int kind) {
    if (kind == NOOR_RNG_PCG32) {
        const uint32_t x = pcg32_step(r.state, r.inc);
        return (float)(x >> 8) / (float)0x01000000;
    }
    uint32_t s = (uint32_t)r.state;
    s = 1664525u * s + 1013904223u;
    r.state = s;
    return (float)(s & 0x00FFFFFFu) / (float)0x01000000;
}

// ---------------------------------------------------------------------------------------------
// rays (noor/cuda/ray.cuh:26-207) plus the watertight shear constants
struct Ray {
    f3 o, d, inv;
    float tmax;
    uint32_t pn;  // bit a set when d[a] <= 0 (the reference posneg); bit 3: box4_hit's fast form applies
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

// Box-test terms of a ray: reciprocal direction and octant flags.
__device__ __forceinline__ void ray_update_box(Ray& r) {
    r.inv = f3{1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    r.pn = (r.d.x > 0.0f ? 0u : 1u) | (r.d.y > 0.0f ? 0u : 2u) | (r.d.z > 0.0f ? 0u : 4u);
    // the fast slab form needs finite reciprocals whose products with a
    // node's step cannot overflow (steps stay far below 2^40)
    const float mi = fmaxf(fmaxf(fabsf(r.inv.x), fabsf(r.inv.y)), fabsf(r.inv.z));
    if (mi <= 0x1p80f) r.pn |= 8u;
}

// Triangle-test terms: Woop, Benthin & Wald, "Watertight Ray/Triangle
// Intersection", JCGT 2(1), 2013 (axis permutation and shear).  Needs inv.
__device__ __forceinline__ void ray_update_tri(Ray& r) {
    const float ax = fabsf(r.d.x), ay = fabsf(r.d.y), az = fabsf(r.d.z);
    int kz = 0;
    float m = ax;
    if (ay > m) {
        kz = 1;
        m = ay;
    }
    if (az > m) kz = 2;
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    if (c3(r.d, kz) < 0.0f) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    const float dz = c3(r.d, kz);
    r.Sx = c3(r.d, kx) / dz;
    r.Sy = c3(r.d, ky) / dz;
    r.Sz = c3(r.inv, kz);  // 1 / dz, already rounded once in inv
}

__device__ __forceinline__ void ray_update(Ray& r) {
    ray_update_box(r);
    ray_update_tri(r);
}
__device__ __forceinline__ Ray mkray(f3 o, f3 d, float tmax) {
    Ray r;
    r.o = o;
    r.d = d;
    r.tmax = tmax;
    ray_update(r);
    return r;
}
__device__ __forceinline__ f3 ray_at(const Ray& r, float t) { return add3(r.o, scl3(r.d, t)); }

struct Xf {
    f3 r0, r1, r2, t;
};
__device__ __forceinline__ Xf xf_load(const float4* rows) {
    const float4 a = ldg(rows), b = ldg(rows + 1), c = ldg(rows + 2);
    Xf x;
    x.r0 = f3{a.x, a.y, a.z};
    x.r1 = f3{b.x, b.y, b.z};
    x.r2 = f3{c.x, c.y, c.z};
    x.t = f3{a.w, b.w, c.w};
    return x;
}
__device__ __forceinline__ f3 xf_vec(const Xf& x, f3 v) { return f3{dot3(x.r0, v), dot3(x.r1, v), dot3(x.r2, v)}; }
__device__ __forceinline__ f3 xf_pt(const Xf& x, f3 p) { return add3(x.t, xf_vec(x, p)); }
__device__ __forceinline__ Xf xf_from(const noor_transform& t) {
    Xf x;
    x.r0 = fromf3(t.row[0]);
    x.r1 = fromf3(t.row[1]);
    x.r2 = fromf3(t.row[2]);
    x.t = fromf3(t.t);
    return x;
}

// ---------------------------------------------------------------------------------------------
// traversal

// Conservative slab test of the four child boxes of a wide node
// (CudaBBox::intersect, noor/cuda/bbox.cuh:62-80).  The exit distance and the
// ray's tmax are widened by 2^-20 relative, more than the rounding of this
// test and of the triangle tests, so a box holding a hit that a triangle test
// accepts is always entered.  The decoded boxes of a compressed node contain
// the exact ones, which only adds entered boxes.  Which boxes a traversal
// visits, and in which order, then cannot change the closest hit (ties are
// ordered by primitive id): the wide BVH returns the hits of the binary BVH it
// was built from.  fmaxf/fminf ignore the NaN of an origin on the slab plane
// of a zero direction component.
constexpr float BOX_SLACK = 1.00000095367431640625f;  // 1 + 2^-20

// A compressed node (QNode in wide_bvh.h) as loaded: 56 of its 64 bytes.
struct WideNode {
    uint4 a;  // org.xyz, meta
    uint4 b;  // qlo.xyz, qhi.x
    uint4 c;  // qhi.yz, ref0, ref1
    uint2 d;  // ref2, ref3
    __device__ __forceinline__ uint4 ref() const { return make_uint4(c.z, c.w, d.x, d.y); }
};
__device__ __forceinline__ WideNode load_wnode(const DevScene& S, uint32_t idx) {
    const uint4* p = S.wbvh + (size_t)idx * 4;
    WideNode n;
    n.a = ldg(p);
    n.b = ldg(p + 1);
    n.c = ldg(p + 2);
    n.d = ldg(reinterpret_cast<const uint2*>(p + 3));
    // the references are needed only when a child is entered, but loading
    // them then would add a second dependent fetch: issue them with the boxes
    asm volatile("" ::"v"(n.d.x), "v"(n.d.y), "v"(n.c.z), "v"(n.c.w));
    return n;
}
// Plane c (byte c of q) of an axis with origin org and step s: the decode the
// builder checked (quantize in wide_bvh.cpp).  q * s is exact (8-bit q, s a
// power of two), so the fused multiply-add rounds org + q * s once, exactly
// as the builder's separate product and sum do, in one instruction.
__device__ __forceinline__ float q_plane(float org, uint32_t q, int c, float s) {
    return __builtin_fmaf((float)((q >> (8 * c)) & 0xFFu), s, org);
}

// Slab test of the four children of a compressed node, for rays with a zero
// (or subnormal) direction component: each plane is decoded as the builder
// checked it (q_plane) and measured from the ray origin, the reference's
// arithmetic (noor/cuda/bbox.cuh:62-80) including its infinities.
__device__ __forceinline__ uint32_t box4_hit_exact(const Ray& r, const WideNode& n, float tn[4]) {
    const float ox = __uint_as_float(n.a.x), oy = __uint_as_float(n.a.y), oz = __uint_as_float(n.a.z);
    const uint32_t meta = n.a.w;
    const float sx = __uint_as_float((meta & 0xFFu) << 23);
    const float sy = __uint_as_float(((meta >> 8) & 0xFFu) << 23);
    const float sz = __uint_as_float(((meta >> 16) & 0xFFu) << 23);
    const uint32_t qlx = n.b.x, qly = n.b.y, qlz = n.b.z, qhx = n.b.w, qhy = n.c.x, qhz = n.c.y;
    const uint32_t nx = (r.pn & 1u) ? qhx : qlx, fx = (r.pn & 1u) ? qlx : qhx;
    const uint32_t ny = (r.pn & 2u) ? qhy : qly, fy = (r.pn & 2u) ? qly : qhy;
    const uint32_t nz = (r.pn & 4u) ? qhz : qlz, fz = (r.pn & 4u) ? qlz : qhz;
    const float tmaxs = r.tmax * BOX_SLACK;
    uint32_t mask = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float tx0 = (q_plane(ox, nx, c, sx) - r.o.x) * r.inv.x;
        const float tx1 = (q_plane(ox, fx, c, sx) - r.o.x) * r.inv.x;
        const float ty0 = (q_plane(oy, ny, c, sy) - r.o.y) * r.inv.y;
        const float ty1 = (q_plane(oy, fy, c, sy) - r.o.y) * r.inv.y;
        const float tz0 = (q_plane(oz, nz, c, sz) - r.o.z) * r.inv.z;
        const float tz1 = (q_plane(oz, fz, c, sz) - r.o.z) * r.inv.z;
        const float t0 = fmaxf(fmaxf(tx0, ty0), tz0);
        const float t1 = fminf(fminf(tx1, ty1), tz1) * BOX_SLACK;
        tn[c] = t0;
        if (t0 <= t1 && t0 <= tmaxs && t1 >= 0.0f) mask |= 1u << c;
    }
    return mask & (meta >> 24);
}

// Slab test of the four children of a compressed node.  Bit c of the result
// is set when child c is entered; tn[c] is its entry distance.
//
// Per axis, A = (org - o) inv and B = s inv (exact: s is a power of two), and
// the distance of plane q is one fused multiply-add, fma(q, B, A): two VALU
// operations per plane (byte extract, fma) instead of four.  Against the
// exact distance t* = (org + q s - o) / d of the exact grid value (which the
// builder checked contains the child, q_exact in wide_bvh.cpp) it errs by a
// relative (1 + d0)(1 + d3), from inv and the fma, and an absolute at most
// 2.01 u |A| (u = 2^-24), from the rounding of org - o and of its product
// with inv (q B is exact inside the fma).  The relative part is covered by
// the 2^-20 relative widening of exit distances and tmax; the absolute part by
// W = 2^-22 max|A|, subtracted from entry and added to exit distances.  So a
// box holding a point of the ray between 0 and tmax is always entered.
__device__ __forceinline__ uint32_t box4_hit(const Ray& r, const WideNode& n, float tn[4]) {
    if (!(r.pn & 8u)) return box4_hit_exact(r, n, tn);
    const uint32_t meta = n.a.w;
    const float sx = __uint_as_float((meta & 0xFFu) << 23);
    const float sy = __uint_as_float(((meta >> 8) & 0xFFu) << 23);
    const float sz = __uint_as_float(((meta >> 16) & 0xFFu) << 23);
    const float Ax = (__uint_as_float(n.a.x) - r.o.x) * r.inv.x, Bx = sx * r.inv.x;
    const float Ay = (__uint_as_float(n.a.y) - r.o.y) * r.inv.y, By = sy * r.inv.y;
    const float Az = (__uint_as_float(n.a.z) - r.o.z) * r.inv.z, Bz = sz * r.inv.z;
    const float W = fmaxf(fmaxf(fabsf(Ax), fabsf(Ay)), fabsf(Az)) * 0x1p-22f;
    const uint32_t qlx = n.b.x, qly = n.b.y, qlz = n.b.z, qhx = n.b.w, qhy = n.c.x, qhz = n.c.y;
    const uint32_t nx = (r.pn & 1u) ? qhx : qlx, fx = (r.pn & 1u) ? qlx : qhx;
    const uint32_t ny = (r.pn & 2u) ? qhy : qly, fy = (r.pn & 2u) ? qly : qhy;
    const uint32_t nz = (r.pn & 4u) ? qhz : qlz, fz = (r.pn & 4u) ? qlz : qhz;
    const float tmaxs = r.tmax * BOX_SLACK;
    uint32_t mask = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float tx0 = __builtin_fmaf((float)((nx >> (8 * c)) & 0xFFu), Bx, Ax);
        const float tx1 = __builtin_fmaf((float)((fx >> (8 * c)) & 0xFFu), Bx, Ax);
        const float ty0 = __builtin_fmaf((float)((ny >> (8 * c)) & 0xFFu), By, Ay);
        const float ty1 = __builtin_fmaf((float)((fy >> (8 * c)) & 0xFFu), By, Ay);
        const float tz0 = __builtin_fmaf((float)((nz >> (8 * c)) & 0xFFu), Bz, Az);
        const float tz1 = __builtin_fmaf((float)((fz >> (8 * c)) & 0xFFu), Bz, Az);
        const float t0 = fmaxf(fmaxf(tx0, ty0), tz0) - W;
        const float t1 = fminf(fminf(tx1, ty1), tz1) * BOX_SLACK + W;
        tn[c] = t0;
        if (t0 <= t1 && t0 <= tmaxs && t1 >= 0.0f) mask |= 1u << c;
    }
    return mask & (meta >> 24);
}

struct Counters {
    uint32_t nodes, tris, tri2, inst;
};

// Wald projection test (noor/cuda/triangle.cuh:82-141)
__device__ __forceinline__ bool tri_wald(const DevScene& S, const Ray& r, uint32_t i, float& to, float& uo,
                                         float& vo, Counters& c) {
    const uint4 a = ldg(S.wald0 + i);
    const uint4 cc = ldg(S.wald2 + i);
    if (cc.z == 1u) return false;
    const uint32_t k = a.x;
    const uint32_t ku = (k == 0u) ? 1u : ((k == 1u) ? 2u : 0u);
    const uint32_t kv = (k == 0u) ? 2u : ((k == 1u) ? 0u : 1u);
    const float ro_u = c3(r.o, ku), ro_v = c3(r.o, kv), ro_k = c3(r.o, k);
    const float rd_u = c3(r.d, ku), rd_v = c3(r.d, kv), rd_k = c3(r.d, k);
    const float nd = __uint_as_float(a.w), nu = __uint_as_float(a.y), nv = __uint_as_float(a.z);
    const float denom = rd_k + nu * rd_u + nv * rd_v;
    if (denom == 0.0f) return false;
    const float t = (nd - ro_k - nu * ro_u - nv * ro_v) / denom;
    if (t < 0.0f || t > r.tmax) return false;
    const uint4 b = ldg(S.wald1 + i);
    c.tri2++;
    const float hu = ro_u + t * rd_u - __uint_as_float(b.z);
    const float hv = ro_v + t * rd_v - __uint_as_float(b.w);
    const float u = hv * __uint_as_float(b.x) + hu * __uint_as_float(b.y);
    const float v = hu * __uint_as_float(cc.x) + hv * __uint_as_float(cc.y);
    if (!(u >= 0.0f && v >= 0.0f && u + v <= 1.0f)) return false;
    to = t;
    uo = u;
    vo = v;
    return true;
}

// Watertight test (Woop, Benthin & Wald 2013) with a double-precision
// fallback for zero edge functions.
__device__ __forceinline__ bool tri_watertight(const DevScene& S, const Ray& r, uint32_t i, float& to,
                                               float& uo, float& vo) {
    const float4* T = S.tri + 3 * (size_t)i;
    const float4 p0 = ldg(T), p1 = ldg(T + 1), p2 = ldg(T + 2);
    const f3 A = sub3(f3{p0.x, p0.y, p0.z}, r.o);
    const f3 B = sub3(f3{p1.x, p1.y, p1.z}, r.o);
    const f3 C = sub3(f3{p2.x, p2.y, p2.z}, r.o);
    const float Az0 = c3(A, r.kz), Bz0 = c3(B, r.kz), Cz0 = c3(C, r.kz);
    const float Ax = c3(A, r.kx) - r.Sx * Az0;
    const float Ay = c3(A, r.ky) - r.Sy * Az0;
    const float Bx = c3(B, r.kx) - r.Sx * Bz0;
    const float By = c3(B, r.ky) - r.Sy * Bz0;
    const float Cx = c3(C, r.kx) - r.Sx * Cz0;
    const float Cy = c3(C, r.ky) - r.Sy * Cz0;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * Az0, Bz = r.Sz * Bz0, Cz = r.Sz * Cz0;
    const float Tt = U * Az + V * Bz + W * Cz;
    const float inv = 1.0f / det;
    const float t = Tt * inv;
    if (!(t >= 0.0f) || t > r.tmax) return false;
    to = t;
    uo = V * inv;
    vo = W * inv;
    return true;
}

__device__ __forceinline__ bool alpha_pass(const DevScene& S, uint32_t tri, float u, float v);

// ALPHA: triangles of alpha-tested materials whose transparency texture reads
// <= 1e-4 at the hit are not hit (noor/cuda/triangle.cuh:124-135)
template <bool STATS, bool ALPHA = false>
__device__ __forceinline__ bool tri_test(const DevScene& S, const Ray& r, uint32_t i, float& t, float& u,
                                         float& v, Counters& c) {
    if (STATS) c.tris++;
    const bool hit = (S.tri_test == NOOR_TRI_WALD) ? tri_wald(S, r, i, t, u, v, c) : tri_watertight(S, r, i, t, u, v);
    if (ALPHA && hit && (ldg(S.wald2 + i).w & NOOR_MAT_ALPHA)) return alpha_pass(S, i, u, v);
    return hit;
}

// analytic light shapes (noor/cuda/geometry.cuh:106-211)
__device__ __forceinline__ bool solve_quadratic(float a, float b, float c, float& x0, float& x1) {
    if (a == 0.0f) {
        if (b != 0.0f) {
            x0 = x1 = -c / b;
            return true;
        }
        return false;
    }
    const float discrim = b * b - 4.0f * a * c;
    if (discrim < 0.0f) return false;
    const float sqrtD = sqrtf(discrim);
    const float temp = (b < 0.0f) ? -0.5f * (b - sqrtD) : -0.5f * (b + sqrtD);
    x0 = temp / a;
    x1 = c / temp;
    if (x0 > x1) {
        const float t = x0;
        x0 = x1;
        x1 = t;
    }
    return true;
}

__device__ __forceinline__ bool shape_intersect(const noor_shape& sh, Ray& r) {
    const f3 n = fromf3(sh.n), ctr = fromf3(sh.center);
    if (sh.type == NOOR_SHAPE_QUAD || sh.type == NOOR_SHAPE_DISK) {
        const float denom = dot3(n, r.d);
        if (denom == 0.0f) return false;
        const float t_hit = -(dot3(r.o, n) - dot3(ctr, n)) / denom;
        if (t_hit > r.tmax) return false;
        const f3 inplane = sub3(ray_at(r, t_hit), ctr);
        if (sh.type == NOOR_SHAPE_QUAD) {
            const f3 u = fromf3(sh.u), v = fromf3(sh.v);
            const float t0 = dot3(inplane, u) / dot3(u, u);
            if (t0 < 0.0f || t0 > 1.0f) return false;
            const float t1 = dot3(inplane, v) / dot3(v, v);
            if (t1 < 0.0f || t1 > 1.0f) return false;
        } else {
            if (dot3(inplane, inplane) > sh.radius2) return false;
        }
        r.tmax = t_hit;
        return true;
    }
    const f3 o = sub3(r.o, ctr);
    const float A = dot3(r.d, r.d), B = 2.0f * dot3(o, r.d), C = dot3(o, o) - sh.radius2;
    float nearT, farT;
    if (!solve_quadratic(A, B, C, nearT, farT)) return false;
    if (!(nearT <= r.tmax && farT >= 0.0f)) return false;
    if (nearT < 0.0f) {
        if (farT > r.tmax) return false;
        r.tmax = farT;
    } else {
        r.tmax = nearT;
    }
    return true;
}

__device__ __forceinline__ bool light_node_intersect(const DevScene& S, uint32_t li, Ray& r) {
    if (li < S.n_area) return shape_intersect(S.area[li].shape, r);
    return false;
}

struct Hit {
    float t, u, v;
    uint32_t prim, key, inst, kind;
};

// Per-lane traversal stack.  Entries [0, K) live in LDS, lane-interleaved
// (entry d of a lane at base[d * stride], so the 64 lanes of a wave use 64
// banks); deeper entries, which few rays reach, live in device memory laid
// out the same way over all lanes of the launch.  `depth` is the exact bound
// of the BVH, so the stack cannot overflow.
// The two parts are typed by address space so that entries move with LDS and
// global instructions: through a generic pointer every access would be a flat
// one, and each wait for an LDS entry would also wait for the node loads in
// flight.
struct TravStack {
    NOOR_LDS uint32_t* lds;
    int lstride;
    int K;
    NOOR_GLOBAL uint32_t* glob;
    uint32_t gstride;
    int depth;
    __device__ __forceinline__ void put(int d, uint32_t v) const {
        if (d < K) lds[d * lstride] = v;
        else glob[(size_t)(d - K) * gstride] = v;
    }
    __device__ __forceinline__ uint32_t get(int d) const {
        return d < K ? lds[d * lstride] : glob[(size_t)(d - K) * gstride];
    }
};
// Stack of this lane whose entries beyond the LDS part live in `gstack`,
// lane-interleaved over `glanes` lanes; the launch must use at most `glanes`.
__device__ __forceinline__ TravStack make_stack(const DevScene& S, uint32_t* smem, uint32_t* gstack, uint32_t glanes) {
    TravStack t;
    t.lds = (NOOR_LDS uint32_t*)(smem + threadIdx.x);
    t.lstride = (int)blockDim.x;
    t.K = S.stack_lds;
    t.glob = (NOOR_GLOBAL uint32_t*)(gstack + (blockIdx.x * blockDim.x + threadIdx.x));
    t.gstride = glanes;
    t.depth = S.stack_depth;
    return t;
}
// The context's stack; the launch must use at most S.gstack_lanes lanes.
__device__ __forceinline__ TravStack make_stack(const DevScene& S, uint32_t* smem) {
    return make_stack(S, smem, S.gstack, S.gstack_lanes);
}

__device__ __forceinline__ uint32_t ref_of(const uint4& r, int i) {
    return i == 0 ? r.x : (i == 1 ? r.y : (i == 2 ? r.z : r.w));
}
__device__ __forceinline__ void cswap(float& a, float& b, uint32_t& ra, uint32_t& rb) {
    const bool s = b < a;
    const float ta = s ? b : a, tb = s ? a : b;
    const uint32_t ua = s ? rb : ra, ub = s ? ra : rb;
    a = ta;
    b = tb;
    ra = ua;
    rb = ub;
}

// Two-level closest hit (noor/cuda/bvh.cuh:142-191) over the wide BVH, as a
// resumable state machine: closest_steps() advances a traversal by at most
// `max_steps` node or leaf visits and returns true when it has finished, so a
// persistent kernel can hand a finished lane a new ray while the others go
// on.  Ties in t are broken by (prim_id, instance); light nodes win only when
// strictly closer.  Hit children are followed nearest first.
// The world ray is kept as origin and direction only; lray is the ray of the
// level being traversed (world, or object space inside an instance) with its
// precomputed terms and the current closest distance, and is rebuilt from the
// world ray on the way back to the top level.  Keeping one full Ray instead of
// two saves the registers that decide how many waves a CU can keep in flight.
struct ClosestTrav {
    f3 o, d;
    Ray lray;
    uint32_t cur, ins;
    int sp;
    bool hit, world;
    Hit h;
};

__device__ __forceinline__ void closest_begin(const DevScene& S, const Ray& r, ClosestTrav& t) {
    t.o = r.o;
    t.d = r.d;
    t.lray = r;
    t.world = true;
    t.cur = S.wroot;
    t.ins = 0;
    t.sp = 0;
    t.hit = false;
    t.h.kind = 0;
    t.h.prim = 0xFFFFFFFFu;
    t.h.key = 0xFFFFFFFFu;
    t.h.inst = 0xFFFFFFFFu;
    t.h.t = 0.0f;
    t.h.u = t.h.v = 0.0f;
}

// One step of a traversal visits one inner node, tests ONE triangle of a leaf
// (a leaf with triangles left stays the current reference, its start advanced
// and its count decremented), enters one instance or tests one light.  Lanes
// of a wave step in lockstep, so a wave pays for the union of the kinds its
// lanes visit in a step: testing a whole leaf per step would make every step
// with a leaf lane cost as many triangle tests as the wave's largest leaf.
// The ray of the level is rebuilt in one place for both transitions that need
// it: entering an instance (object space) and popping back to the top level
// (world space); a pop straight into another instance builds only the
// instance's ray, and the top level's ray gets only its box-test terms.
__device__ __forceinline__ uint32_t leaf_next(uint32_t cur) {
    return ((cur >> REF_LEAF_COUNT_SHIFT) & 7u) ? cur + 1u - (1u << REF_LEAF_COUNT_SHIFT) : REF_EMPTY;
}

template <bool STATS, bool ALPHA = false>
__device__ __forceinline__ bool closest_steps(const DevScene& S, ClosestTrav& t, const TravStack& st, Counters& c,
                                              bool& overflow, int max_steps) {
    for (int step = 0; step < max_steps; ++step) {
        const uint32_t cur = t.cur;
        const uint32_t kind = (cur >> REF_KIND_SHIFT) & 3u;
        if (kind == REF_INST || ((cur & REF_TLAS) && !t.world)) {
            f3 o = t.o, d = t.d;
            if (kind == REF_INST) {
                t.ins = cur & REF_INDEX_MASK;
                if (STATS) c.inst++;
                const Xf T = xf_load(S.xf + 9 * (size_t)t.ins);
                o = xf_pt(T, t.o);
                d = xf_vec(T, t.d);
            }
            t.lray.o = o;
            t.lray.d = d;
            ray_update_box(t.lray);
            if (kind == REF_INST) ray_update_tri(t.lray);  // the top level holds no triangles
            t.world = kind != REF_INST;
        }
        if (kind == REF_INNER) {
            if (STATS) c.nodes++;
            const WideNode n = load_wnode(S, cur & REF_INDEX_MASK);
            float tn[4];
            const uint32_t m = box4_hit(t.lray, n, tn);
            if (m != 0u) {
                float d0 = (m & 1u) ? tn[0] : K_INF, d1 = (m & 2u) ? tn[1] : K_INF;
                float d2 = (m & 4u) ? tn[2] : K_INF, d3 = (m & 8u) ? tn[3] : K_INF;
                const uint4 rf = n.ref();
                uint32_t r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
                cswap(d0, d1, r0, r1);
                cswap(d2, d3, r2, r3);
                cswap(d0, d2, r0, r2);
                cswap(d1, d3, r1, r3);
                cswap(d1, d2, r1, r2);
                const int k = __popc(m);
                if (t.sp + k - 1 > st.depth) {
                    overflow = true;
                    return true;
                }
                if (k > 3) st.put(t.sp++, r3);
                if (k > 2) st.put(t.sp++, r2);
                if (k > 1) st.put(t.sp++, r1);
                t.cur = r0;
                continue;
            }
        } else if (kind == REF_LEAF) {
            const uint32_t ti = cur & REF_LEAF_START_MASK;
            float tt, u, v;
            if (tri_test<STATS, ALPHA>(S, t.lray, ti, tt, u, v, c)) {
                const uint32_t key = ldg(S.prim_ids + ti);
                bool take = false;
                if (tt < t.lray.tmax) take = true;
                else if (tt == t.lray.tmax) {
                    if (!t.hit) take = true;
                    else if (key < t.h.key || (key == t.h.key && t.ins < t.h.inst)) take = true;
                }
                if (take) {
                    t.hit = true;
                    t.lray.tmax = tt;
                    t.h.t = tt;
                    t.h.u = u;
                    t.h.v = v;
                    t.h.prim = ti;
                    t.h.key = key;
                    t.h.inst = t.ins;
                    t.h.kind = 1;
                }
            }
            t.cur = leaf_next(cur);
            if (t.cur != REF_EMPTY) continue;
        } else if (kind == REF_INST) {
            t.cur = ldg(S.inst_root + t.ins);
            if (t.cur != REF_EMPTY) continue;
        } else {
            // light nodes are top-level references: lray is the world ray here
            const uint32_t li = cur & REF_INDEX_MASK;
            Ray tr = t.lray;
            if (light_node_intersect(S, li, tr) && tr.tmax < t.lray.tmax) {
                t.lray.tmax = tr.tmax;
                t.hit = true;
                t.h.t = tr.tmax;
                t.h.kind = 2;
                t.h.inst = li;
                t.h.key = 0xFFFFFFFFu;
                t.h.prim = 0xFFFFFFFFu;
                t.h.u = t.h.v = 0.0f;
            }
        }
        if (t.sp == 0) return true;
        t.cur = st.get(--t.sp);
    }
    return false;
}

template <bool STATS, bool ALPHA = false>
__device__ __forceinline__ bool bvh_closest(const DevScene& S, Ray& ray, Hit& h, const TravStack& st, Counters& c,
                                            bool& overflow) {
    ClosestTrav t;
    closest_begin(S, ray, t);
    while (!closest_steps<STATS, ALPHA>(S, t, st, c, overflow, 1 << 30)) {
    }
    ray.tmax = t.lray.tmax;
    h = t.h;
    return t.hit;
}

// Any hit (noor/cuda/bvh.cuh:193-234) over the wide BVH, resumable like the
// closest-hit traversal.
struct AnyTrav {
    f3 o, d;  // world ray; lray as in ClosestTrav (tmax fixed)
    Ray lray;
    uint32_t cur;
    int sp;
    bool occluded, world;
};

__device__ __forceinline__ void any_begin(const DevScene& S, const Ray& r, AnyTrav& t) {
    t.o = r.o;
    t.d = r.d;
    t.lray = r;
    t.world = true;
    t.cur = S.wroot;
    t.sp = 0;
    t.occluded = false;
}

// Steps as in closest_steps (one node, triangle, instance or light per step).
template <bool STATS, bool ALPHA = false>
__device__ __forceinline__ bool any_steps(const DevScene& S, AnyTrav& t, const TravStack& st, Counters& c,
                                          bool& overflow, int max_steps) {
    for (int step = 0; step < max_steps; ++step) {
        const uint32_t cur = t.cur;
        const uint32_t kind = (cur >> REF_KIND_SHIFT) & 3u;
        if (kind == REF_INST || ((cur & REF_TLAS) && !t.world)) {
            f3 o = t.o, d = t.d;
            if (kind == REF_INST) {
                if (STATS) c.inst++;
                const Xf T = xf_load(S.xf + 9 * (size_t)(cur & REF_INDEX_MASK));
                o = xf_pt(T, t.o);
                d = xf_vec(T, t.d);
            }
            t.lray.o = o;
            t.lray.d = d;
            ray_update_box(t.lray);
            if (kind == REF_INST) ray_update_tri(t.lray);  // the top level holds no triangles
            t.world = kind != REF_INST;
        }
        if (kind == REF_INNER) {
            if (STATS) c.nodes++;
            const WideNode n = load_wnode(S, cur & REF_INDEX_MASK);
            float tn[4];
            uint32_t m = box4_hit(t.lray, n, tn);
            if (m != 0u) {
                const int k = __popc(m);
                if (t.sp + k - 1 > st.depth) {
                    overflow = true;
                    return true;
                }
                const uint4 rf = n.ref();
                t.cur = ref_of(rf, __ffs(m) - 1);
                m &= m - 1u;
                while (m) {
                    st.put(t.sp++, ref_of(rf, __ffs(m) - 1));
                    m &= m - 1u;
                }
                continue;
            }
        } else if (kind == REF_LEAF) {
            float tt, u, v;
            if (tri_test<STATS, ALPHA>(S, t.lray, cur & REF_LEAF_START_MASK, tt, u, v, c)) {
                t.occluded = true;
                return true;
            }
            t.cur = leaf_next(cur);
            if (t.cur != REF_EMPTY) continue;
        } else if (kind == REF_INST) {
            t.cur = ldg(S.inst_root + (cur & REF_INDEX_MASK));
            if (t.cur != REF_EMPTY) continue;
        } else {
            Ray tr = t.lray;
            if (light_node_intersect(S, cur & REF_INDEX_MASK, tr)) {
                t.occluded = true;
                return true;
            }
        }
        if (t.sp == 0) return true;
        t.cur = st.get(--t.sp);
    }
    return false;
}

template <bool STATS, bool ALPHA = false>
__device__ __forceinline__ bool bvh_any(const DevScene& S, const Ray& ray, const TravStack& st, Counters& c,
                                        bool& overflow) {
    AnyTrav t;
    any_begin(S, ray, t);
    while (!any_steps<STATS, ALPHA>(S, t, st, c, overflow, 1 << 30)) {
    }
    return t.occluded;
}

// ---------------------------------------------------------------------------------------------
// sampling utilities (noor/cuda/utils.cuh:359-470)
__device__ __forceinline__ f2 concentric_disk(f2 u) {
    const f2 o = f2{2.0f * u.x - 1.0f, 2.0f * u.y - 1.0f};
    if (o.x == 0.0f && o.y == 0.0f) return f2{0.0f, 0.0f};
    float theta, r;
    if (fabsf(o.x) > fabsf(o.y)) {
        r = o.x;
        theta = K_PI_OVER_4 * (o.y / o.x);
    } else {
        r = o.y;
        theta = K_PI_OVER_2 - K_PI_OVER_4 * (o.x / o.y);
    }
    return f2{r * nm_cosf(theta), r * nm_sinf(theta)};
}
__device__ __forceinline__ f3 cosine_hemisphere(f2 u) {
    const f2 d = concentric_disk(u);
    const float z = sqrtf(rmax(0.0f, 1.0f - d.x * d.x - d.y * d.y));
    return f3{d.x, d.y, z};
}
__device__ __forceinline__ float spherical_phi(f3 v) {
    const float p = nm_atan2f(v.z, v.x);
    return (p < 0.0f) ? (p + K_2PI) : p;
}
__device__ __forceinline__ float spherical_theta(f3 v) { return nm_acosf(clampr(v.y, -1.0f, 1.0f)); }
__device__ __forceinline__ float power_heuristic(float f, float g) { return (f * f) / (f * f + g * g); }
__device__ __forceinline__ void coordinate_system(f3 v1, f3& v2, f3& v3o) {
    if (fabsf(v1.x) > fabsf(v1.y))
        v2 = sdiv3(f3{-v1.z, 0.0f, v1.x}, sqrtf(v1.x * v1.x + v1.z * v1.z));
    else
        v2 = sdiv3(f3{0.0f, v1.z, -v1.y}, sqrtf(v1.y * v1.y + v1.z * v1.z));
    v3o = cross3(v1, v2);
}

// ---------------------------------------------------------------------------------------------
// environment: 2D distribution (noor/cuda/distribution2d.cuh:165-337)
__device__ __forceinline__ float d_cdf(const DevScene& S, int row, int col) {
    const long idx = (long)row * (S.dw + 1) + col;
    const long n = (long)S.dh * (S.dw + 2) + 1;
    if (idx < 0 || idx >= n) return 0.0f;
    return ldg(S.dcdf + idx);
}
__device__ __forceinline__ float d_func(const DevScene& S, int row, int col) {
    const long idx = (long)row * (S.dw + 1) + col;
    const long n = (long)S.dh * (S.dw + 2) + 1;
    if (idx < 0 || idx >= n) return 0.0f;
    return ldg(S.dfunc + idx);
}
__device__ __forceinline__ float d_funcint(const DevScene& S, int row) {
    return d_func(S, row, row == S.dh ? S.dh : S.dw);
}
__device__ __forceinline__ int d_find_interval(const DevScene& S, int row, int n, float x) {
    int l = 0, r = n, interval = -1;
    while (l <= r) {
        const int m = l + (r - l) / 2;
        if (d_cdf(S, row, m) <= x) {
            l = m + 1;
            interval = m;
        } else {
            r = m - 1;
        }
    }
    return interval < 0 ? 0 : (interval > n - 1 ? n - 1 : interval);
}
__device__ __forceinline__ float d_sample1d(const DevScene& S, int row, int n, float u, float& pdf, int* off) {
    const int offset = d_find_interval(S, row, n, u);
    if (off) *off = offset;
    const float func = d_func(S, row, offset);
    const float cdf = d_cdf(S, row, offset);
    const float cdf_n = d_cdf(S, row, offset + 1);
    const float funcInt = d_funcint(S, row);
    float du = u - cdf;
    if ((cdf_n - cdf) > 0.0f) du /= (cdf_n - cdf);
    pdf = (funcInt > 0.0f) ? func / funcInt : 0.0f;
    return ((float)offset + du) / (float)n;
}
__device__ __forceinline__ f2 d_sample2d(const DevScene& S, f2 u, float& pdf) {
    const int plus = (S.compat & NOOR_COMPAT_CDF_N_PLUS_1) ? 1 : 0;
    float py, px;
    int v;
    const float y = d_sample1d(S, S.dh, S.dh + plus, u.y, py, &v);
    const float x = d_sample1d(S, v, S.dw + plus, u.x, px, nullptr);
    pdf = px * py;
    return f2{x, y};
}
__device__ __forceinline__ float d_pdf(const DevScene& S, f2 p) {
    int iu = (int)(p.x * (float)S.dw), iv = (int)(p.y * (float)S.dh);
    iu = iu < 0 ? 0 : (iu > S.dw - 1 ? S.dw - 1 : iu);
    iv = iv < 0 ? 0 : (iv > S.dh - 1 ? S.dh - 1 : iv);
    return d_func(S, iv, iu) / d_funcint(S, S.dh);
}

// tex2D, normalized coordinates, linear filter, clamp addressing
// (noor/cuda/texture.cuh:107-113), FP32 weights.
__device__ __forceinline__ f3 tex_bilinear(const float4* tex, int w, int h, float u, float v) {
    if (u != u) u = 0.0f;
    if (v != v) v = 0.0f;
    float x = u * (float)w - 0.5f;
    float y = v * (float)h - 0.5f;
    x = clampr(x, -2.0f, (float)w + 1.0f);
    y = clampr(y, -2.0f, (float)h + 1.0f);
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    const int ix = (int)fx0, iy = (int)fy0;
    const int i0 = ix < 0 ? 0 : (ix > w - 1 ? w - 1 : ix);
    const int i1 = ix + 1 < 0 ? 0 : (ix + 1 > w - 1 ? w - 1 : ix + 1);
    const int j0 = iy < 0 ? 0 : (iy > h - 1 ? h - 1 : iy);
    const int j1 = iy + 1 < 0 ? 0 : (iy + 1 > h - 1 ? h - 1 : iy + 1);
    const float4 t00 = ldg(tex + (size_t)j0 * w + i0), t10 = ldg(tex + (size_t)j0 * w + i1);
    const float4 t01 = ldg(tex + (size_t)j1 * w + i0), t11 = ldg(tex + (size_t)j1 * w + i1);
    const float bx = 1.0f - ax, by = 1.0f - ay;
    f3 r;
    r.x = by * (bx * t00.x + ax * t10.x) + ay * (bx * t01.x + ax * t11.x);
    r.y = by * (bx * t00.y + ax * t10.y) + ay * (bx * t01.y + ax * t11.y);
    r.z = by * (bx * t00.z + ax * t10.z) + ay * (bx * t01.z + ax * t11.z);
    return r;
}

// The fourth channel of tex_bilinear's lookup.
__device__ __forceinline__ float tex_bilinear_w(const float4* tex, int w, int h, float u, float v) {
    if (u != u) u = 0.0f;
    if (v != v) v = 0.0f;
    float x = u * (float)w - 0.5f;
    float y = v * (float)h - 0.5f;
    x = clampr(x, -2.0f, (float)w + 1.0f);
    y = clampr(y, -2.0f, (float)h + 1.0f);
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    const int ix = (int)fx0, iy = (int)fy0;
    const int i0 = ix < 0 ? 0 : (ix > w - 1 ? w - 1 : ix);
    const int i1 = ix + 1 < 0 ? 0 : (ix + 1 > w - 1 ? w - 1 : ix + 1);
    const int j0 = iy < 0 ? 0 : (iy > h - 1 ? h - 1 : iy);
    const int j1 = iy + 1 < 0 ? 0 : (iy + 1 > h - 1 ? h - 1 : iy + 1);
    const float t00 = ldg(tex + (size_t)j0 * w + i0).w, t10 = ldg(tex + (size_t)j0 * w + i1).w;
    const float t01 = ldg(tex + (size_t)j1 * w + i0).w, t11 = ldg(tex + (size_t)j1 * w + i1).w;
    const float bx = 1.0f - ax, by = 1.0f - ay;
    return by * (bx * t00 + ax * t10) + ay * (bx * t01 + ax * t11);
}

// ---- material image textures (noor/cuda/texture.cuh:30-423) -------------------------------
// Software filtering restated identically in the oracle: bilinear within a
// level (texel centres at half-integers, wrap addressing) and linear between
// the two levels around lod = log2(footprint in texels), the footprint being
// the longer of the two screen-space uv gradients.  The reference samples
// through CUDA texture units with 64x anisotropy, whose weights are not
// specified; its images are not reproduced bit for bit here.
__device__ __forceinline__ int tex_wrap(int x, int n) {
    const int r = x % n;
    return r < 0 ? r + n : r;
}
__device__ __forceinline__ float4 tex_level(const DevScene& S, const TexDesc& t, uint32_t l, float s, float q) {
    uint32_t off = t.offset, lw = t.w, lh = t.h;
    for (uint32_t j = 0; j < l; ++j) {
        off += lw * lh;
        lw = lw > 1u ? lw >> 1 : 1u;
        lh = lh > 1u ? lh >> 1 : 1u;
    }
    if (!(fabsf(s) < 1e30f)) s = 0.0f;  // NaN and infinities
    if (!(fabsf(q) < 1e30f)) q = 0.0f;
    s = s - floorf(s);  // wrap into [0, 1)
    q = q - floorf(q);
    const float x = s * (float)lw - 0.5f, y = q * (float)lh - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const int x0 = tex_wrap((int)fx, (int)lw), x1 = tex_wrap((int)fx + 1, (int)lw);
    const int y0 = tex_wrap((int)fy, (int)lh), y1 = tex_wrap((int)fy + 1, (int)lh);
    const float4* T = S.tex_texels + off;
    const float4 t00 = ldg(T + (size_t)y0 * lw + x0), t10 = ldg(T + (size_t)y0 * lw + x1);
    const float4 t01 = ldg(T + (size_t)y1 * lw + x0), t11 = ldg(T + (size_t)y1 * lw + x1);
    const float bx = 1.0f - ax, by = 1.0f - ay;
    float4 r;
    r.x = by * (bx * t00.x + ax * t10.x) + ay * (bx * t01.x + ax * t11.x);
    r.y = by * (bx * t00.y + ax * t10.y) + ay * (bx * t01.y + ax * t11.y);
    r.z = by * (bx * t00.z + ax * t10.z) + ay * (bx * t01.z + ax * t11.z);
    r.w = by * (bx * t00.w + ax * t10.w) + ay * (bx * t01.w + ax * t11.w);
    return r;
}
// tex2D: level 0
__device__ __forceinline__ float4 tex_eval(const DevScene& S, uint32_t ti, float s, float q) {
    const TexDesc t = S.tex_desc[ti];
    return tex_level(S, t, 0, s, q);
}
// tex2DGrad: (gx0, gx1) and (gy0, gy1) the uv changes per pixel step in x and y
__device__ __forceinline__ float4 tex_grad(const DevScene& S, uint32_t ti, float s, float q, float gx0, float gx1,
                                           float gy0, float gy1) {
    const TexDesc t = S.tex_desc[ti];
    const float fw = (float)t.w, fh = (float)t.h;
    const float ax = gx0 * fw, ay = gx1 * fh, bx = gy0 * fw, by = gy1 * fh;
    const float width = rmax(sqrtf(ax * ax + ay * ay), sqrtf(bx * bx + by * by));
    if (!(width > 1.0f) || t.levels <= 1u) return tex_level(S, t, 0, s, q);
    const float lod = nm_logf(width) * 1.44269504088896340736f;  // log2
    if (lod >= (float)(t.levels - 1u)) return tex_level(S, t, t.levels - 1u, s, q);
    const float fl = floorf(lod);
    const uint32_t l0 = (uint32_t)fl;
    const float f = lod - fl;
    const float4 a = tex_level(S, t, l0, s, q), b = tex_level(S, t, l0 + 1u, s, q);
    const float g = 1.0f - f;
    return make_float4(g * a.x + f * b.x, g * a.y + f * b.y, g * a.z + f * b.z, g * a.w + f * b.w);
}

// tex2D with point filtering (the transparency maps, noor/source/model.cpp:286-297):
// the texel under (s, q) at level 0, wrap addressing
__device__ __forceinline__ float4 tex_point(const DevScene& S, uint32_t ti, float s, float q) {
    const TexDesc t = S.tex_desc[ti];
    if (!(fabsf(s) < 1e30f)) s = 0.0f;
    if (!(fabsf(q) < 1e30f)) q = 0.0f;
    s = s - floorf(s);
    q = q - floorf(q);
    const int x = tex_wrap((int)floorf(s * (float)t.w), (int)t.w);
    const int y = tex_wrap((int)floorf(q * (float)t.h), (int)t.h);
    return ldg(S.tex_texels + t.offset + (size_t)y * t.w + x);
}

// MaterialManager::getAlpha at the hit's interpolated uv (noor/cuda/materialimp.cuh:159-163)
__device__ __forceinline__ bool alpha_pass(const DevScene& S, uint32_t tri, float u, float v) {
    const uint4 ai = ldg(S.attr + tri);
    const float2 q0 = ldg(S.uv + ai.x), q1 = ldg(S.uv + ai.y), q2 = ldg(S.uv + ai.z);
    const float w = 1.0f - u - v;
    const float tu = w * q0.x + u * q1.x + v * q2.x;
    const float tv = w * (1.0f - q0.y) + u * (1.0f - q1.y) + v * (1.0f - q2.y);
    const DevMat* m = S.mats + ai.w;
    const float a = tex_point(S, m->tex[TX_ALPHA], m->uvs[TX_ALPHA][0] * tu, m->uvs[TX_ALPHA][1] * tv).x;
    return !(a <= 0.0001f);
}

// Screen-space uv derivatives of a hit (zero for a ray without differentials)
struct UvGrad {
    float dudx, dvdx, dudy, dvdy;
};

// One texture lookup as the material getters make it: tex2D at level 0
// (lod0) or tex2DGrad.  Out of line: the shading kernels call it from
// several sites.
__device__ __noinline__ float4 tex_field(const DevScene& S, uint32_t ti, float s, float q, float gx0, float gx1,
                                         float gy0, float gy1, bool lod0) {
    return lod0 ? tex_eval(S, ti, s, q) : tex_grad(S, ti, s, q, gx0, gx1, gy0, gy1);
}

// MaterialManager getters: the fields of `m` that read image textures looked
// up at (u, v) (noor/cuda/materialimp.cuh:120-186, noor/cuda/texture.cuh:114-138:
// the uv scaled component-wise, the x and y gradients scaled by uvscale.x
// and uvscale.y; roughness reads level 0).
__device__ __forceinline__ void resolve_material(const DevScene& S, float u, float v, UvGrad g, DevMat& m) {
    const uint32_t mask = m.tex_mask;
#pragma unroll
    for (uint32_t f = 0; f < TX_BUMP; ++f) {
        if (!(mask & (1u << f))) continue;
        const float sx = m.uvs[f][0], sy = m.uvs[f][1];
        const float4 c = tex_field(S, m.tex[f], sx * u, sy * v, sx * g.dudx, sx * g.dvdx, sy * g.dudy, sy * g.dvdy,
                                   f == TX_ROUGH);
        switch (f) {
            case TX_DIFFUSE: m.diffuse[0] = c.x; m.diffuse[1] = c.y; m.diffuse[2] = c.z; break;
            case TX_SPECULAR: m.specular[0] = c.x; m.specular[1] = c.y; m.specular[2] = c.z; break;
            case TX_TRANSMISSION: m.transmission[0] = c.x; m.transmission[1] = c.y; m.transmission[2] = c.z; break;
            case TX_EMITTANCE:
                m.emittance[0] = m.emitter_scale * c.x;
                m.emittance[1] = m.emitter_scale * c.y;
                m.emittance[2] = m.emitter_scale * c.z;
                break;
            case TX_ETA: m.eta[0] = c.x; m.eta[1] = c.y; m.eta[2] = c.z; break;
            case TX_K: m.k[0] = c.x; m.k[1] = c.y; m.k[2] = c.z; break;
            case TX_ROUGH: m.rough[0] = c.x; m.rough[1] = c.y; break;
            default: break;
        }
    }
}

// Hosek-Wilkie (noor/cuda/hosek.cuh:55-101)
__device__ __forceinline__ float hosek_rad(const noor_hosek_state& st, float theta, float gamma, int ch) {
    const float* c = st.configs[ch];
    const float cg = nm_cosf(gamma), cg2 = cg * cg, ct = nm_cosf(theta);
    const float expM = nm_expf(c[4] * gamma);
    const float mieM = (1.0f + cg2) / nm_powf((1.0f + c[8] * c[8] - 2.0f * c[8] * cg), 1.5f);
    const float zenith = sqrtf(ct);
    return (1.0f + c[0] * nm_expf(c[1] / (ct + 0.01f))) *
           (c[2] + c[3] * expM + c[5] * cg2 + c[6] * mieM + c[7] * zenith) * st.radiances[ch];
}
__device__ __forceinline__ float inv_gamma(float v) {
    if (v <= 0.04045f) return v * 1.f / 12.92f;
    return nm_powf((v + 0.055f) * 1.f / 1.055f, 2.4f);
}
__device__ __forceinline__ f3 hosek_query(const noor_hosek& hk, f3 v) {
    if (v.y < 0.0f) return f3{0.0f, 0.0f, 0.0f};
    const float theta = clampr(spherical_theta(v), 0.0f, K_PI_OVER_2);
    const float gamma = nm_acosf(clampr(dot3(v, fromf3(hk.solar_dir)), -1.0f, 1.0f));
    const f3 rgb = f3{hosek_rad(hk.state_r, theta, gamma, 0), hosek_rad(hk.state_g, theta, gamma, 1),
                      hosek_rad(hk.state_b, theta, gamma, 2)};
    if (gamma < hk.solar_radius) return scl3(rgb, hk.solar_scale);
    const f3 sc = scl3(rgb, 0.025f);
    return f3{inv_gamma(sc.x), inv_gamma(sc.y), inv_gamma(sc.z)};
}

// CudaSkyDomeManager::evaluate (noor/cuda/skydome.cuh:91-101)
__device__ __forceinline__ f3 sky_eval(const DevScene& S, f3 dir) {
    const float phi = spherical_phi(dir) * K_INV_2PI;
    const float theta = spherical_theta(dir) * K_INV_PI;
    if (S.skydome_type == NOOR_SKY_PHYSICAL) return hosek_query(*S.hosek, dir);
    if (S.skydome_type == NOOR_SKY_CONSTANT) return f3{0.25f, 0.25f, 0.25f};
    return tex_bilinear(S.env, (int)S.env_w, (int)S.env_h, phi, theta);
}

// importance_sample_dir (noor/cuda/skydome.cuh:131-140)
__device__ __forceinline__ f3 env_sample_dir(const DevScene& S, Rng& rng, float& pdf, f2* uv_out = nullptr) {
    f2 u;
    u.x = rng_next(rng, S.rng_kind);
    u.y = rng_next(rng, S.rng_kind);
    float map_pdf = 1.0f;
    u = d_sample2d(S, u, map_pdf);
    if (uv_out) *uv_out = u;
    const float phi = u.x * K_2PI, theta = u.y * K_PI;
    const float cosT = nm_cosf(theta), sinT = nm_sinf(theta);
    const f3 dir = nrm3(f3{sinT * nm_cosf(phi), cosT, sinT * nm_sinf(phi)});
    pdf = (sinT == 0.0f) ? 0.0f : map_pdf / (2.0f * K_PI * K_PI * sinT);
    return dir;
}

// ---------------------------------------------------------------------------------------------
// surface interaction (noor/cuda/intersection.cuh:95-442)
struct Isect {
    f3 p, wo, wi, gn, sn, ss, ts;
    float u, v;
    uint32_t tri, ins, mat, type;
    f3 catcher;  // shadow catcher: the sky behind the surface, sky(-wo)
};

__device__ __forceinline__ bool is_emitter(const Isect& I) { return (I.type & M_EMITTER_ANY) != 0; }
__device__ __forceinline__ bool is_specular(const Isect& I) { return (I.type & M_SPECULAR) != 0; }
__device__ __forceinline__ bool is_transparent(const Isect& I) { return (I.type & M_TRANSPARENT) != 0; }
__device__ __forceinline__ bool is_shadowcatcher(const Isect& I) { return (I.type & NOOR_MAT_SHADOWCATCHER) != 0; }

// Ray differentials (noor/cuda/ray.cuh:26-207): the auxiliary rays one pixel
// to the right (x) and one pixel down (y).  Carried only by camera rays and
// the spawns of glossy materials.
struct RayDiff {
    f3 ox, oy, dx, dy;
    int has;
};
// Derivatives at a hit (noor/cuda/intersection.cuh:370-440); zero when the
// ray had no differentials.
struct HitDiff {
    f3 dpdx, dpdy, dndu, dndv;
    UvGrad g;
};

// NOOR::solve2x2 (noor/cuda/math.cuh:193-207)
__device__ __forceinline__ void solve2x2(float a00, float a01, float a10, float a11, float b0, float b1, float& x0,
                                         float& x1) {
    const float det = a00 * a11 - a10 * a01;
    if (fabsf(det) < 1e-10f) {
        x0 = x1 = 0.0f;
        return;
    }
    const float inv = 1.0f / det;
    x0 = (a11 * b0 - a01 * b1) * inv;
    x1 = (a00 * b1 - a10 * b0) * inv;
    if (x0 != x0 || x1 != x1) x0 = x1 = 0.0f;
}

// CudaIntersection::computeDifferential (noor/cuda/intersection.cuh:399-439)
__device__ __forceinline__ void hit_differential(const Isect& I, const RayDiff& rd, f3 dpdu, f3 dpdv, HitDiff& hd) {
    const float d = dot3(I.gn, I.p);
    const float tx = -(dot3(I.gn, rd.ox) - d) / dot3(I.gn, rd.dx);
    if (isinf(tx) || tx != tx) return;
    const f3 px = add3(rd.ox, scl3(rd.dx, tx));
    const float ty = -(dot3(I.gn, rd.oy) - d) / dot3(I.gn, rd.dy);
    if (isinf(ty) || ty != ty) return;
    const f3 py = add3(rd.oy, scl3(rd.dy, ty));
    hd.dpdx = sub3(px, I.p);
    hd.dpdy = sub3(py, I.p);
    int dx, dy;
    const f3 n = I.gn;
    if (fabsf(n.x) > fabsf(n.y) && fabsf(n.x) > fabsf(n.z)) {
        dx = 1;
        dy = 2;
    } else if (fabsf(n.y) > fabsf(n.z)) {
        dx = 0;
        dy = 2;
    } else {
        dx = 0;
        dy = 1;
    }
    solve2x2(c3(dpdu, dx), c3(dpdv, dx), c3(dpdu, dy), c3(dpdv, dy), c3(hd.dpdx, dx), c3(hd.dpdx, dy), hd.g.dudx,
             hd.g.dvdx);
    solve2x2(c3(dpdu, dx), c3(dpdv, dx), c3(dpdu, dy), c3(dpdv, dy), c3(hd.dpdy, dx), c3(hd.dpdy, dy), hd.g.dudy,
             hd.g.dvdy);
}

// The hit record (CudaIntersection::updateIntersection, noor/cuda/intersection.cuh:317-397).
// TEX: also the derivatives of the ray differentials `rd` at the hit into `hd`
// (zero when the ray has none), and an emitter keeps its barycentrics as uv.
template <bool TEX = false>
__device__ __forceinline__ void make_isect(const DevScene& S, const Ray& ray, const Hit& h, Isect& I,
                                           const RayDiff* rd = nullptr, HitDiff* hd = nullptr) {
    if (TEX) {
        hd->dpdx = hd->dpdy = hd->dndu = hd->dndv = f3{0.0f, 0.0f, 0.0f};
        hd->g = UvGrad{0.0f, 0.0f, 0.0f, 0.0f};
    }
    I.p = ray_at(ray, ray.tmax);
    I.wo = nrm3(scl3(ray.d, -1.0f));
    I.ins = h.inst;
    I.wi = f3{0.0f, 0.0f, 0.0f};
    I.gn = I.sn = I.ss = I.ts = f3{0.0f, 0.0f, 0.0f};
    I.u = I.v = 0.0f;
    if (h.kind == 2) {
        I.type = NOOR_MAT_MESHLIGHT | NOOR_MAT_EMITTER;
        I.tri = 0;
        I.mat = 0;
        return;
    }
    I.tri = h.prim;
    I.type = ldg(S.wald2 + h.prim).w;
    const uint4 ai = ldg(S.attr + I.tri);
    I.mat = ai.w;
    if (I.type & M_EMITTER_ANY) {
        // an emitter keeps the barycentrics as its uv (computeTangentSpace
        // returns before interpolating, noor/cuda/intersection.cuh:326-327)
        if (TEX) {
            I.u = h.u;
            I.v = h.v;
        }
        return;
    }
    const float2 q0 = ldg(S.uv + ai.x), q1 = ldg(S.uv + ai.y), q2 = ldg(S.uv + ai.z);
    const f2 uv0 = f2{q0.x, 1.0f - q0.y}, uv1 = f2{q1.x, 1.0f - q1.y}, uv2 = f2{q2.x, 1.0f - q2.y};
    const f3 p0 = fromf4(ldg(S.vert + ai.x)), p1 = fromf4(ldg(S.vert + ai.y)), p2 = fromf4(ldg(S.vert + ai.z));
    const f3 n0 = fromf4(ldg(S.nrm + ai.x)), n1 = fromf4(ldg(S.nrm + ai.y)), n2 = fromf4(ldg(S.nrm + ai.z));
    const float4* X = S.xf + 9 * (size_t)I.ins;
    Xf T = xf_load(X + 6);
    T.t = f3{0.0f, 0.0f, 0.0f};
    const float bu = h.u, bv = h.v;
    const float w = 1.0f - bu - bv;
    const f3 ni = add3(add3(scl3(n0, w), scl3(n1, bu)), scl3(n2, bv));
    I.sn = nrm3s(nrm3(xf_vec(T, ni)));
    I.u = w * uv0.x + bu * uv1.x + bv * uv2.x;
    I.v = w * uv0.y + bu * uv1.y + bv * uv2.y;
    const Xf Sx = xf_load(X + 3);
    const f3 dp01 = xf_vec(Sx, sub3(p1, p0));
    const f3 dp02 = xf_vec(Sx, sub3(p2, p0));
    I.gn = nrm3s(cross3(dp01, dp02));
    const f2 d01 = f2{uv1.x - uv0.x, uv1.y - uv0.y};
    const f2 d02 = f2{uv2.x - uv0.x, uv2.y - uv0.y};
    const float det = d01.x * d02.y - d02.x * d01.y;
    const bool degenerate = (double)fabsf(det) < 1e-8;
    f3 dpdu, dpdv;
    if (!degenerate) {
        const float invdet = 1.0f / det;
        const float a00 = d02.y * invdet, a01 = -d01.y * invdet;
        const float a10 = -d02.x * invdet, a11 = d01.x * invdet;
        dpdu = add3(scl3(dp01, a00), scl3(dp02, a01));
        dpdv = add3(scl3(dp01, a10), scl3(dp02, a11));
        const f3 cr = cross3(dpdu, dpdv);
        if (dot3(cr, cr) == 0.0f) coordinate_system(I.gn, dpdu, dpdv);
        if (TEX && rd->has) {
            const f3 dn01 = sub3(n1, n0), dn02 = sub3(n2, n0);
            hd->dndu = add3(scl3(dn01, a00), scl3(dn02, a01));
            hd->dndv = add3(scl3(dn01, a10), scl3(dn02, a11));
        }
    } else {
        coordinate_system(I.gn, dpdu, dpdv);
        if (TEX && rd->has) {
            const f3 dn = nrm3s(nrm3(xf_vec(T, cross3(sub3(n1, n0), sub3(n2, n0)))));
            if (dot3(dn, dn) != 0.0f) coordinate_system(dn, hd->dndu, hd->dndv);
        }
    }
    if (TEX && rd->has) hit_differential(I, *rd, dpdu, dpdv, *hd);
    const f3 n = I.sn;
    const f3 ss = nrm3s(sub3(dpdu, scl3(n, dot3(n, dpdu))));
    const float sign = dot3(cross3(n, ss), dpdv) < 0.0f ? -1.0f : 1.0f;
    I.ss = ss;
    I.ts = nrm3s(scl3(cross3(n, ss), sign));
    // CudaShadowCatcher::Eval reads the sky behind the surface (noor/cuda/bxdf.cuh:210-216)
    if (I.type & NOOR_MAT_SHADOWCATCHER) I.catcher = sky_eval(S, scl3(I.wo, -1.0f));
}

// bump (noor/cuda/accumulate.cuh:47-81): finite differences of the
// displacement texture over half the uv footprint of a pixel perturb the
// shading frame
__device__ __forceinline__ float bump_tex(const DevScene& S, const DevMat& m, float u, float v, const UvGrad& g) {
    const float sx = m.uvs[TX_BUMP][0], sy = m.uvs[TX_BUMP][1];
    return tex_grad(S, m.tex[TX_BUMP], sx * u, sy * v, sx * g.dudx, sx * g.dvdx, sy * g.dudy, sy * g.dvdy).x;
}
struct Frame {
    f3 ss, ts, sn;
};
__device__ __noinline__ Frame bump_frame(const DevScene& S, const DevMat& m, float u, float v, UvGrad g, Frame F,
                                         f3 gn) {
    float du = 0.5f * (fabsf(g.dudx) + fabsf(g.dudy));
    float dv = 0.5f * (fabsf(g.dvdx) + fabsf(g.dvdy));
    du = du == 0.0f ? 0.001f : du;
    dv = dv == 0.0f ? 0.001f : dv;
    const float disp = bump_tex(S, m, u, v, g);
    const float ud = m.bumpfactor * (bump_tex(S, m, u + du, v, g) - disp) / du;
    const float vd = m.bumpfactor * (bump_tex(S, m, u, v + dv, g) - disp) / dv;
    const f3 dpdu = add3(F.ss, scl3(F.sn, ud));
    const f3 dpdv = add3(F.ts, scl3(F.sn, vd));
    const f3 n = nrm3s(cross3(dpdu, dpdv));
    Frame o;
    o.ss = nrm3s(sub3(dpdu, scl3(n, dot3(n, dpdu))));
    o.ts = nrm3s(cross3(n, o.ss));
    o.sn = dot3(gn, n) < 0.0f ? scl3(n, -1.0f) : n;
    return o;
}

// spawnRay with differentials (noor/cuda/intersection.cuh:115-182): the
// reflected or refracted auxiliary rays of a glossy material, their spread
// scaled by 1/4.  `rd` holds the incoming ray's differentials and receives
// the spawned ray's.  The refraction branch's dndy pairs dndv with dudy, as
// the reference does.
__device__ __forceinline__ void spawn_diff(const DevScene& S, const Isect& I, const HitDiff& hd, f3 origin,
                                           RayDiff& rd) {
    const f3 wi = I.wi, wo = I.wo, n = I.sn;
    const f3 ox = add3(origin, hd.dpdx), oy = add3(origin, hd.dpdy);
    const f3 dwodx = scl3(add3(rd.dx, wo), -1.0f);
    const f3 dwody = scl3(add3(rd.dy, wo), -1.0f);
    f3 dx, dy;
    if (dot3(wi, I.gn) >= 0.0f) {
        const f3 dndx = add3(scl3(hd.dndu, hd.g.dudx), scl3(hd.dndv, hd.g.dvdx));
        const f3 dndy = add3(scl3(hd.dndu, hd.g.dudy), scl3(hd.dndv, hd.g.dvdy));
        const float dDNdx = dot3(dwodx, n) + dot3(wo, dndx);
        const float dDNdy = dot3(dwody, n) + dot3(wo, dndy);
        const float wn = dot3(wo, n);
        dx = add3(sub3(wi, dwodx), scl3(add3(scl3(dndx, wn), scl3(n, dDNdx)), 2.0f));
        dy = add3(sub3(wi, dwody), scl3(add3(scl3(dndy, wn), scl3(n, dDNdy)), 2.0f));
    } else {
        const f3 w = scl3(wo, -1.0f);
        const float ieta = (I.type & (NOOR_MAT_GLASS | NOOR_MAT_ROUGHGLASS)) ? S.mats[I.mat].coat_ior : 1.0f;
        const float eta = dot3(wo, n) > 0.0f ? 1.0f / ieta : ieta;
        const f3 dndx = add3(scl3(hd.dndu, hd.g.dudx), scl3(hd.dndv, hd.g.dvdx));
        const f3 dndy = add3(scl3(hd.dndv, hd.g.dudy), scl3(hd.dndv, hd.g.dvdy));
        const float dDNdx = dot3(dwodx, n) + dot3(wo, dndx);
        const float dDNdy = dot3(dwody, n) + dot3(wo, dndy);
        const float mu = eta * dot3(w, n) - dot3(wi, n);
        const float k = eta - (eta * eta * dot3(w, n)) / dot3(wi, n);
        const float dmudx = k * dDNdx;
        const float dmudy = k * dDNdy;
        dx = sub3(add3(wi, scl3(dwodx, eta)), add3(scl3(dndx, mu), scl3(n, dmudx)));
        dy = sub3(add3(wi, scl3(dwody, eta)), add3(scl3(dndy, mu), scl3(n, dmudy)));
    }
    // CudaRay::scaleDifferentials(0.25)
    const float s = 0.25f, r = 1.0f - s;
    rd.ox = add3(scl3(origin, r), scl3(ox, s));
    rd.oy = add3(scl3(origin, r), scl3(oy, s));
    rd.dx = add3(scl3(wi, r), scl3(dx, s));
    rd.dy = add3(scl3(wi, r), scl3(dy, s));
    rd.has = 1;
}

__device__ __forceinline__ f3 to_local(const Isect& I, f3 v) { return f3{dot3(v, I.ss), dot3(v, I.ts), dot3(v, I.sn)}; }
__device__ __forceinline__ f3 to_world(const Isect& I, f3 v) {
    return add3(add3(scl3(I.ss, v.x), scl3(I.ts, v.y)), scl3(I.sn, v.z));
}

// ---------------------------------------------------------------------------------------------
// Fresnel (noor/cuda/fresnel.cuh:58-107)
__device__ __forceinline__ float fr_dielectric(float cosI, float etaI, float etaT) {
    cosI = clampr(cosI, -1.0f, 1.0f);
    if (!(cosI > 0.0f)) {
        const float t = etaI;
        etaI = etaT;
        etaT = t;
        cosI = fabsf(cosI);
    }
    const float sinI = sqrtf(rmax(0.0f, 1.0f - cosI * cosI));
    const float sinT = etaI / etaT * sinI;
    if (sinT >= 1.0f) return 1.0f;
    const float cosT = sqrtf(rmax(0.0f, 1.0f - sinT * sinT));
    const float Rparl = ((etaT * cosI) - (etaI * cosT)) / ((etaT * cosI) + (etaI * cosT));
    const float Rperp = ((etaI * cosI) - (etaT * cosT)) / ((etaI * cosI) + (etaT * cosT));
    return (Rparl * Rparl + Rperp * Rperp) / 2.0f;
}
__device__ __forceinline__ f3 fr_conductor(float cosI, f3 etaI, f3 etaT, f3 k) {
    cosI = clampr(cosI, -1.0f, 1.0f);
    const f3 eta = div3(etaT, etaI), etak = div3(k, etaI);
    const float c2 = cosI * cosI, s2 = 1.0f - c2;
    const f3 eta2 = mul3(eta, eta), etak2 = mul3(etak, etak);
    const f3 t0 = sub3(sub3(eta2, etak2), f3{s2, s2, s2});
    const f3 q = add3(mul3(t0, t0), mul3(scl3(eta2, 4.0f), etak2));
    const f3 a2b2 = f3{sqrtf(q.x), sqrtf(q.y), sqrtf(q.z)};
    const f3 t1 = add3(a2b2, f3{c2, c2, c2});
    const f3 hh = scl3(add3(a2b2, t0), 0.5f);
    const f3 a = f3{sqrtf(hh.x), sqrtf(hh.y), sqrtf(hh.z)};
    const f3 t2 = scl3(a, 2.0f * cosI);
    const f3 Rs = div3(sub3(t1, t2), add3(t1, t2));
    const f3 t3 = add3(scl3(a2b2, c2), f3{s2 * s2, s2 * s2, s2 * s2});
    const f3 t4 = scl3(t2, s2);
    const f3 Rp = div3(mul3(Rs, sub3(t3, t4)), add3(t3, t4));
    return scl3(add3(Rp, Rs), 0.5f);
}

// ---------------------------------------------------------------------------------------------
// GGX / Trowbridge-Reitz (noor/cuda/distribution.cuh:58-183)
__device__ __forceinline__ float cos2t(f3 w) { return w.z * w.z; }
__device__ __forceinline__ float sin2t(f3 w) { return rmax(0.0f, 1.0f - cos2t(w)); }
__device__ __forceinline__ float sint(f3 w) { return sqrtf(sin2t(w)); }
__device__ __forceinline__ float tant(f3 w) { return sint(w) / w.z; }
__device__ __forceinline__ float tan2t(f3 w) { return sin2t(w) / cos2t(w); }
__device__ __forceinline__ float cosphi(f3 w) {
    const float s = sint(w);
    return (s == 0.0f) ? 1.0f : clampr(w.x / s, -1.0f, 1.0f);
}
__device__ __forceinline__ float sinphi(f3 w) {
    const float s = sint(w);
    return (s == 0.0f) ? 0.0f : clampr(w.y / s, -1.0f, 1.0f);
}
__device__ __forceinline__ float cos2p(f3 w) { return cosphi(w) * cosphi(w); }
__device__ __forceinline__ float sin2p(f3 w) { return sinphi(w) * sinphi(w); }
__device__ __forceinline__ bool same_hemi(f3 a, f3 b) { return a.z * b.z > 0.0f; }

struct Ggx {
    float ax, ay;
};
__device__ __forceinline__ Ggx ggx_make(const DevMat& m) {
    const float rx = m.rough[0], ry = m.rough[1];
    return Ggx{rmax(rx * rx, 1e-3f), rmax(ry * ry, 1e-3f)};
}
__device__ __forceinline__ float ggx_D(Ggx g, f3 wh) {
    const float t2 = tan2t(wh);
    if (is_inf(t2)) return 0.0f;
    const float c4 = cos2t(wh) * cos2t(wh);
    const float e = (cos2p(wh) / (g.ax * g.ax) + sin2p(wh) / (g.ay * g.ay)) * t2;
    return 1.0f / (K_PI * g.ax * g.ay * c4 * (1.0f + e) * (1.0f + e));
}
__device__ __forceinline__ float ggx_lambda(Ggx g, f3 w) {
    const float at = fabsf(tant(w));
    if (is_inf(at)) return 0.0f;
    const float alpha = sqrtf(cos2p(w) * g.ax * g.ax + sin2p(w) * g.ay * g.ay);
    const float a2t2 = (alpha * at) * (alpha * at);
    return (-1.0f + sqrtf(1.0f + a2t2)) / 2.0f;
}
__device__ __forceinline__ float ggx_G(Ggx g, f3 wo, f3 wi) { return 1.0f / (1.0f + ggx_lambda(g, wo) + ggx_lambda(g, wi)); }
__device__ __forceinline__ float ggx_G1(Ggx g, f3 w) { return 1.0f / (1.0f + ggx_lambda(g, w)); }
__device__ __forceinline__ float ggx_pdf(Ggx g, f3 wo, f3 wh) {
    return ggx_D(g, wh) * ggx_G1(g, wo) * absdot3(wo, wh) / fabsf(wo.z);
}
__device__ __forceinline__ void ggx_sample11(float cosT, float U1, float U2, float& sx, float& sy) {
    if (cosT > .9999f) {
        const float r = sqrtf(U1 / (1.0f - U1));
        const float phi = 6.28318530718f * U2;
        sx = r * nm_cosf(phi);
        sy = r * nm_sinf(phi);
        return;
    }
    const float sinT = sqrtf(rmax(0.0f, 1.0f - cosT * cosT));
    const float tanT = sinT / cosT;
    const float a = 1.0f / tanT;
    const float G1 = 2.0f / (1.0f + sqrtf(1.0f + 1.0f / (a * a)));
    const float A = 2.0f * U1 / G1 - 1.0f;
    float tmp = 1.0f / (A * A - 1.0f);
    if (tmp > 1e10f) tmp = 1e10f;
    const float B = tanT;
    const float D = sqrtf(rmax(B * B * tmp * tmp - (A * A - B * B) * tmp, 0.0f));
    const float sx1 = B * tmp - D, sx2 = B * tmp + D;
    sx = (A < 0.0f || sx2 > 1.0f / tanT) ? sx1 : sx2;
    float Sg;
    if (U2 > 0.5f) {
        Sg = 1.0f;
        U2 = 2.0f * (U2 - .5f);
    } else {
        Sg = -1.0f;
        U2 = 2.0f * (.5f - U2);
    }
    const float z = (U2 * (U2 * (U2 * 0.27385f - 0.73369f) + 0.46341f)) /
                    (U2 * (U2 * (U2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
    sy = Sg * z * sqrtf(1.0f + sx * sx);
}
__device__ __forceinline__ f3 ggx_sample_wh(Ggx g, f3 wo, f2 u) {
    const float sg = (float)isign(wo.z);
    const f3 wi = scl3(wo, sg);
    const f3 st = nrm3(f3{g.ax * wi.x, g.ay * wi.y, wi.z});
    float sx, sy;
    ggx_sample11(st.z, u.x, u.y, sx, sy);
    const float tmp = cosphi(st) * sx - sinphi(st) * sy;
    sy = sinphi(st) * sx + cosphi(st) * sy;
    sx = tmp;
    sx *= g.ax;
    sy *= g.ay;
    return scl3(nrm3(f3{-sx, -sy, 1.0f}), sg);
}

// ---------------------------------------------------------------------------------------------
// lobes and BSDF (noor/cuda/bxdf.cuh, noor/cuda/bsdf.cuh)
struct Lobe {
    int id, type;
};
struct Bsdf {
    Lobe l[2];
    int n;
};
__device__ __forceinline__ bool lobe_matches(int type, int flags) { return (type & flags) == type; }

// factoryBSDF (noor/cuda/bsdf.cuh:425-461)
__device__ __forceinline__ void make_bsdf(uint32_t mtype, Bsdf& b) {
    const uint32_t t = mtype & ~(uint32_t)(NOOR_MAT_BUMP | NOOR_MAT_ALPHA);
    b.n = 0;
    const Lobe lam{L_LAMBERT_R, B_REFLECTION | B_DIFFUSE | B_DIELECTRIC};
    const Lobe lamt{L_LAMBERT_T, B_TRANSMISSION | B_DIFFUSE | B_DIELECTRIC};
    const Lobe mdiel{L_MICRO_R_DIEL, B_DIELECTRIC | B_REFLECTION | B_GLOSSY};
    const Lobe mcond{L_MICRO_R_COND, B_CONDUCTOR | B_REFLECTION | B_GLOSSY};
    const Lobe fspec{L_FRESNEL_SPEC, B_REFLECTION | B_TRANSMISSION | B_SPECULAR | B_DIELECTRIC};
    const Lobe mirror{L_SPEC_R_NOOP, B_NOOP | B_REFLECTION | B_SPECULAR};
    const Lobe catcher{L_SHADOWCATCHER, B_REFLECTION | B_TRANSMISSION | B_DIFFUSE};
    const Lobe mtrans{L_MICRO_T, B_TRANSMISSION | B_GLOSSY | B_DIELECTRIC};
    const Lobe blend{L_FRESNEL_BLEND, B_REFLECTION | B_GLOSSY | B_DIELECTRIC};
    const Lobe coat{L_CLEARCOAT, B_GLOSSY | B_REFLECTION | B_SPECULAR};
    switch (t) {
        case NOOR_MAT_DIFFUSE: b.l[b.n++] = lam; break;
        case NOOR_MAT_TRANSLUCENT: b.l[b.n++] = lam; b.l[b.n++] = lamt; break;
        case NOOR_MAT_GLASS: b.l[b.n++] = fspec; break;
        case NOOR_MAT_GLOSSY: b.l[b.n++] = lam; b.l[b.n++] = mdiel; break;
        case NOOR_MAT_METAL: b.l[b.n++] = mcond; break;
        case NOOR_MAT_MIRROR: b.l[b.n++] = mirror; break;
        case NOOR_MAT_SHADOWCATCHER: b.l[b.n++] = catcher; break;
        case NOOR_MAT_ROUGHGLASS: b.l[b.n++] = mdiel; b.l[b.n++] = mtrans; break;
        case NOOR_MAT_SUBSTRATE: b.l[b.n++] = blend; break;
        case NOOR_MAT_CLEARCOAT: b.l[b.n++] = coat; break;
        default: break;
    }
}

__device__ __forceinline__ f3 fresnel_eval(int ltype, const DevMat& m, float cosI) {
    if (ltype & B_DIELECTRIC) {
        const float F = fr_dielectric(cosI, 1.0f, m.coat_ior);
        return f3{F, F, F};
    }
    if (ltype & B_CONDUCTOR) return fr_conductor(cosI, f3{1.0f, 1.0f, 1.0f}, ld3(m.eta), ld3(m.k));
    return f3{1.0f, 1.0f, 1.0f};
}
__device__ __forceinline__ f3 micro_eval(int ltype, const DevMat& m, f3 wo, f3 wi) {
    if (!same_hemi(wo, wi)) return f3{0.0f, 0.0f, 0.0f};
    const float cO = fabsf(wo.z), cI = fabsf(wi.z);
    if (cI == 0.0f || cO == 0.0f) return f3{0.0f, 0.0f, 0.0f};
    f3 wh = nrm3(add3(wi, wo));
    if (is_black(wh)) return f3{0.0f, 0.0f, 0.0f};
    wh = scl3(wh, (float)isign(wh.z));
    const f3 F = fresnel_eval(ltype, m, dot3(wo, wh));
    const Ggx g = ggx_make(m);
    const float D = ggx_D(g, wh), G = ggx_G(g, wo, wi);
    return sdiv3(scl3(scl3(mul3(F, ld3(m.specular)), D), G), 4.0f * cI * cO);
}
__device__ __forceinline__ float micro_pdf(const DevMat& m, f3 wo, f3 wi) {
    if (!same_hemi(wo, wi)) return 0.0f;
    const f3 wh = nrm3s(add3(wo, wi));
    const Ggx g = ggx_make(m);
    return ggx_pdf(g, wo, wh) / (4.0f * dot3(wo, wh));
}
__device__ __forceinline__ f3 refl_z(f3 wo) { return f3{-wo.x, -wo.y, wo.z}; }
__device__ __forceinline__ bool refract_(f3 wi, f3 n, float eta, f3& wt) {
    const float cosI = dot3(n, wi);
    const float sin2I = rmax(0.0f, 1.0f - cosI * cosI);
    const float sin2T = eta * eta * sin2I;
    if (sin2T >= 1.0f) return false;
    const float cosT = sqrtf(1.0f - sin2T);
    wt = add3(scl3(wi, -eta), scl3(n, eta * cosI - cosT));
    return true;
}
__device__ __forceinline__ float pow5r(float v) {  // NOOR::pow5f
    const float a = v * v;
    return a * a * v;
}

// CudaMicrofacetTransmission (noor/cuda/bxdf.cuh:559-667): eta from coat_ior
__device__ __forceinline__ f3 micro_t_eval(const DevMat& m, f3 wo, f3 wi) {
    if (same_hemi(wo, wi)) return f3{0.0f, 0.0f, 0.0f};
    const float cO = wo.z, cI = wi.z;
    if (cI == 0.0f || cO == 0.0f) return f3{0.0f, 0.0f, 0.0f};
    const float eta = wo.z > 0.0f ? m.coat_ior / 1.0f : 1.0f / m.coat_ior;
    f3 wh = nrm3(add3(wo, scl3(wi, eta)));
    wh = scl3(wh, (float)isign(wh.z));
    const float F = fr_dielectric(dot3(wo, wh), 1.0f, m.coat_ior);
    const float sd = dot3(wo, wh) + eta * dot3(wi, wh);
    const Ggx g = ggx_make(m);
    const float v = ggx_D(g, wh) * ggx_G(g, wo, wi) * eta * eta * absdot3(wi, wh) * absdot3(wo, wh) / (cI * cO * sd * sd);
    return scl3(scl3(ld3(m.transmission), 1.0f - F), fabsf(v));
}
__device__ __forceinline__ float micro_t_pdf(const DevMat& m, f3 wo, f3 wi) {
    if (same_hemi(wo, wi)) return 0.0f;
    const float eta = wo.z > 0.0f ? m.coat_ior / 1.0f : 1.0f / m.coat_ior;
    const f3 wh = nrm3(add3(wo, scl3(wi, eta)));
    const float sd = dot3(wo, wh) + eta * dot3(wi, wh);
    const float dwh = fabsf((eta * eta * dot3(wi, wh)) / (sd * sd));
    return ggx_pdf(ggx_make(m), wo, wh) * dwh;
}

// CudaFresnelBlend (noor/cuda/bxdf.cuh:669-752)
__device__ __forceinline__ f3 blend_eval(const DevMat& m, f3 wo, f3 wi) {
    const f3 S = ld3(m.specular);
    f3 d = scl3(ld3(m.diffuse), 28.0f / (23.0f * K_PI));
    d = mul3(d, sub3(f3{1.0f, 1.0f, 1.0f}, S));
    d = scl3(d, 1.0f - pow5r(1.0f - .5f * fabsf(wi.z)));
    d = scl3(d, 1.0f - pow5r(1.0f - .5f * fabsf(wo.z)));
    f3 wh = add3(wi, wo);
    if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return f3{0.0f, 0.0f, 0.0f};
    wh = nrm3(wh);
    const Ggx g = ggx_make(m);
    const float sp = ggx_D(g, wh) / (4.0f * absdot3(wi, wh) * rmax(fabsf(wi.z), fabsf(wo.z)));
    const f3 schlick = add3(S, scl3(sub3(f3{1.0f, 1.0f, 1.0f}, S), pow5r(1.0f - dot3(wi, wh))));
    return add3(d, scl3(schlick, sp));
}
__device__ __forceinline__ float blend_pdf(const DevMat& m, f3 wo, f3 wi) {
    if (!same_hemi(wo, wi)) return 0.0f;
    const f3 wh = nrm3(add3(wo, wi));
    const float pwh = ggx_pdf(ggx_make(m), wo, wh);
    return .5f * (fabsf(wi.z) * K_INV_PI + pwh / (4.0f * dot3(wo, wh)));
}

// CudaClearCoat (noor/cuda/bxdf.cuh:842-940): a conductor microfacet
// substrate under a dielectric coat of index coat_ior
__device__ __forceinline__ float coat_absorption(const DevMat& m, f3 wo, f3 wi) {
    const float sigmaA = m.coat_sigma * m.coat_thickness;
    if (sigmaA != 0.0f) return nm_expf(-sigmaA * (1.0f / fabsf(wo.z) + 1.0f / fabsf(wi.z)));
    return 1.0f;
}
__device__ __forceinline__ f3 coat_eval(const DevMat& m, f3 wo, f3 wi) {
    const float eta = 1.0f / m.coat_ior;
    const float F0 = fr_dielectric(wo.z, 1.0f, m.coat_ior);
    const float F1 = fr_dielectric(wi.z, 1.0f, m.coat_ior);
    const f3 n = f3{0.0f, 0.0f, 1.0f};
    f3 wo_c = f3{0.0f, 0.0f, 0.0f}, wi_c = f3{0.0f, 0.0f, 0.0f};
    refract_(wo, n, eta, wo_c);
    refract_(wi, n, eta, wi_c);
    wo_c = scl3(wo_c, -1.0f);
    wi_c = scl3(wi_c, -1.0f);
    const f3 r = scl3(scl3(micro_eval(B_CONDUCTOR | B_REFLECTION | B_GLOSSY, m, wo_c, wi_c), 1.0f - F0), 1.0f - F1);
    return scl3(r, coat_absorption(m, wo_c, wi_c));
}
__device__ __forceinline__ float coat_pdf(const DevMat& m, f3 wo, f3 wi) {
    const float eta = 1.0f / m.coat_ior;
    const float F = fr_dielectric(wo.z, 1.0f, m.coat_ior);
    const f3 n = f3{0.0f, 0.0f, 1.0f};
    f3 wo_c = f3{0.0f, 0.0f, 0.0f}, wi_c = f3{0.0f, 0.0f, 0.0f};
    refract_(wo, n, eta, wo_c);
    refract_(wi, n, eta, wi_c);
    wo_c = scl3(wo_c, -1.0f);
    wi_c = scl3(wi_c, -1.0f);
    const float w = m.coat_weight;
    const float cp = F * w / (F * w + (1.0f - F) * (1.0f - w));
    return micro_pdf(m, wo_c, wi_c) * (1.0f - cp) + cp;
}

// The rough-glass transmission, substrate and clear-coat lobes are evaluated
// out of line: inlined at every lobe loop they raise the register budget of
// the shading kernels for all materials (the megakernel from 211 to 256
// VGPRs, one wave per SIMD instead of two); as calls they cost only the lanes
// whose material has them.  Values are passed by value and the material by
// its global-memory address, so the call spills nothing to the stack.
__device__ __forceinline__ f3 rare_eval_inl(int id, const DevMat& m, f3 wo, f3 wi) {
    if (id == L_MICRO_T) return micro_t_eval(m, wo, wi);
    if (id == L_FRESNEL_BLEND) return blend_eval(m, wo, wi);
    return coat_eval(m, wo, wi);
}
__device__ __forceinline__ float rare_pdf_inl(int id, const DevMat& m, f3 wo, f3 wi) {
    if (id == L_MICRO_T) return micro_t_pdf(m, wo, wi);
    if (id == L_FRESNEL_BLEND) return blend_pdf(m, wo, wi);
    return coat_pdf(m, wo, wi);
}
__device__ __noinline__ f3 rare_eval(int id, const DevMat* m, f3 wo, f3 wi) { return rare_eval_inl(id, *m, wo, wi); }
__device__ __noinline__ float rare_pdf(int id, const DevMat* m, f3 wo, f3 wi) { return rare_pdf_inl(id, *m, wo, wi); }
// The same for a material held in registers (textured shading resolves its
// material per hit): the fields the rare lobes read, passed by value, so the
// caller's copy need not live in memory.
struct RareMat {
    float diffuse[3], specular[3], transmission[3], eta[3], k[3], rough[2];
    float coat_ior, coat_weight, coat_thickness, coat_sigma;
};
__device__ __forceinline__ RareMat rare_pack(const DevMat& m) {
    RareMat r;
    for (int c = 0; c < 3; ++c) {
        r.diffuse[c] = m.diffuse[c];
        r.specular[c] = m.specular[c];
        r.transmission[c] = m.transmission[c];
        r.eta[c] = m.eta[c];
        r.k[c] = m.k[c];
    }
    r.rough[0] = m.rough[0];
    r.rough[1] = m.rough[1];
    r.coat_ior = m.coat_ior;
    r.coat_weight = m.coat_weight;
    r.coat_thickness = m.coat_thickness;
    r.coat_sigma = m.coat_sigma;
    return r;
}
__device__ __forceinline__ DevMat rare_unpack(const RareMat& r) {
    DevMat m;
    for (int c = 0; c < 3; ++c) {
        m.diffuse[c] = r.diffuse[c];
        m.specular[c] = r.specular[c];
        m.transmission[c] = r.transmission[c];
        m.eta[c] = r.eta[c];
        m.k[c] = r.k[c];
        m.emittance[c] = 0.0f;
    }
    m.rough[0] = r.rough[0];
    m.rough[1] = r.rough[1];
    m.coat_ior = r.coat_ior;
    m.coat_weight = r.coat_weight;
    m.coat_thickness = r.coat_thickness;
    m.coat_sigma = r.coat_sigma;
    m.type = 0;
    m.tex_mask = 0;
    return m;
}
__device__ __noinline__ f3 rare_eval_v(int id, RareMat r, f3 wo, f3 wi) {
    const DevMat m = rare_unpack(r);
    return rare_eval_inl(id, m, wo, wi);
}
__device__ __noinline__ float rare_pdf_v(int id, RareMat r, f3 wo, f3 wi) {
    const DevMat m = rare_unpack(r);
    return rare_pdf_inl(id, m, wo, wi);
}

// LOCAL: m is a copy in registers (see rare_eval_v)
template <bool LOCAL = false>
__device__ __forceinline__ f3 lobe_eval(const Lobe& L, const Isect& I, const DevMat& m, f3 wo, f3 wi) {
    switch (L.id) {
        case L_LAMBERT_R:
        case L_LAMBERT_T: return scl3(ld3(m.diffuse), K_INV_PI);
        case L_MICRO_R_DIEL:
        case L_MICRO_R_COND: return micro_eval(L.type, m, wo, wi);
        case L_SHADOWCATCHER: return sdiv3(I.catcher, fabsf(wi.z));
        case L_MICRO_T:
        case L_FRESNEL_BLEND:
        case L_CLEARCOAT: return LOCAL ? rare_eval_v(L.id, rare_pack(m), wo, wi) : rare_eval(L.id, &m, wo, wi);
        default: return f3{0.0f, 0.0f, 0.0f};
    }
}
template <bool LOCAL = false>
__device__ __forceinline__ float lobe_pdf(const Lobe& L, const DevMat& m, f3 wo, f3 wi) {
    switch (L.id) {
        case L_LAMBERT_R: return same_hemi(wo, wi) ? fabsf(wi.z) * K_INV_PI : 0.0f;
        case L_LAMBERT_T: return !same_hemi(wo, wi) ? fabsf(wi.z) * K_INV_PI : 0.0f;
        case L_MICRO_R_DIEL:
        case L_MICRO_R_COND: return micro_pdf(m, wo, wi);
        case L_SHADOWCATCHER: return 1.0f;
        case L_MICRO_T:
        case L_FRESNEL_BLEND:
        case L_CLEARCOAT: return LOCAL ? rare_pdf_v(L.id, rare_pack(m), wo, wi) : rare_pdf(L.id, &m, wo, wi);
        default: return 0.0f;
    }
}
// MicrofacetReflection::Sample_f (noor/cuda/bxdf.cuh:519-537)
__device__ __forceinline__ f3 micro_sample(int ltype, const DevMat& m, f3 wo, f3& wi, f2 u, float& pdf) {
    if (wo.z == 0.0f) return f3{0.0f, 0.0f, 0.0f};
    const Ggx g = ggx_make(m);
    const f3 wh = ggx_sample_wh(g, wo, u);
    wi = add3(scl3(wo, -1.0f), scl3(wh, 2.0f * dot3(wo, wh)));
    if (!same_hemi(wo, wi)) {
        pdf = 0.0f;
        return f3{0.0f, 0.0f, 0.0f};
    }
    pdf = ggx_pdf(g, wo, wh) / (4.0f * dot3(wo, wh));
    return micro_eval(ltype, m, wo, wi);
}
template <bool LOCAL = false>
__device__ __forceinline__ f3 lobe_sample(const Lobe& L, const Isect& I, const DevMat& m, f3 wo, f3& wi, f2 u,
                                          float& pdf, int& sampled) {
    switch (L.id) {
        case L_LAMBERT_R:
        case L_LAMBERT_T: {
            sampled = L.type;
            wi = cosine_hemisphere(u);
            if (L.id == L_LAMBERT_R) {
                if (wo.z < 0.0f) wi.z *= -1.0f;
            } else {
                if (wo.z > 0.0f) wi.z *= -1.0f;
            }
            pdf = lobe_pdf<LOCAL>(L, m, wo, wi);
            return lobe_eval<LOCAL>(L, I, m, wo, wi);
        }
        case L_MICRO_R_DIEL:
        case L_MICRO_R_COND:
            sampled = L.type;
            return micro_sample(L.type, m, wo, wi, u, pdf);
        case L_SPEC_R_NOOP: {
            sampled = L.type;
            wi = refl_z(wo);
            pdf = 1.0f;
            const f3 F = fresnel_eval(L.type, m, wi.z);
            return sdiv3(mul3(F, ld3(m.diffuse)), fabsf(wi.z));
        }
        case L_FRESNEL_SPEC: {
            const float F = fr_dielectric(wo.z, 1.0f, m.coat_ior);
            if (u.x < F) {
                wi = refl_z(wo);
                pdf = F;
                sampled = B_REFLECTION | B_SPECULAR;
                return sdiv3(scl3(ld3(m.specular), F), fabsf(wi.z));
            }
            const bool entering = wo.z > 0.0f;
            const float etaT = m.coat_ior;
            const float eta = entering ? 1.0f / etaT : etaT / 1.0f;
            const f3 nf = (dot3(wo, f3{0.0f, 0.0f, 1.0f}) < 0.0f) ? f3{0.0f, 0.0f, -1.0f} : f3{0.0f, 0.0f, 1.0f};
            if (!refract_(wo, nf, eta, wi)) return f3{0.0f, 0.0f, 0.0f};
            pdf = 1.0f - F;
            sampled = B_TRANSMISSION | B_SPECULAR;
            return sdiv3(scl3(scl3(scl3(ld3(m.transmission), eta), eta), 1.0f - F), fabsf(wi.z));
        }
        case L_SHADOWCATCHER: {
            // uniformSampleHemisphere, LEFT_HANDED (noor/cuda/utils.cuh:364-372)
            sampled = L.type;
            pdf = 1.0f;
            const float z = u.x;
            const float r = sqrtf(rmax(0.0f, 1.0f - z * z));
            const float phi = 2.0f * K_PI * u.y;
            wi = f3{r * nm_cosf(phi), r * nm_sinf(phi), z};
            return sdiv3(I.catcher, fabsf(wi.z));
        }
        case L_MICRO_T: {
            sampled = L.type;
            if (wo.z == 0.0f) return f3{0.0f, 0.0f, 0.0f};
            const Ggx g = ggx_make(m);
            f3 wh = ggx_sample_wh(g, wo, u);
            const float eta = wo.z > 0.0f ? 1.0f / m.coat_ior : m.coat_ior / 1.0f;
            if (!refract_(wo, wh, eta, wi)) {
                // total internal reflection: helper_math reflect(i, n) = i - 2 n (n.i)
                wi = sub3(wo, scl3(scl3(wh, 2.0f), dot3(wh, wo)));
                wh = scl3(wh, (float)isign(wh.z));
                const float F = fr_dielectric(dot3(wi, wh), 1.0f, m.coat_ior);
                const float cO = wo.z, cI = wi.z;
                const float dg = ggx_D(g, wh) * ggx_G(g, wo, wi);
                pdf = ggx_pdf(g, wo, wh) / (4.0f * dot3(wo, wh));
                return sdiv3(scl3(scl3(ld3(m.specular), F), dg), 4.0f * cI * cO);
            }
            pdf = micro_t_pdf(m, wo, wi);
            return micro_t_eval(m, wo, wi);
        }
        case L_FRESNEL_BLEND: {
            sampled = L.type;
            f2 lu = u;
            if (lu.x < .5f) {
                lu.x = rmin(2.0f * lu.x, K_ONE_MINUS_EPS);
                wi = cosine_hemisphere(lu);
                wi.z *= (float)isign(wi.z);
            } else {
                lu.x = rmin(2.0f * (lu.x - .5f), K_ONE_MINUS_EPS);
                const f3 wh = ggx_sample_wh(ggx_make(m), wo, lu);
                wi = add3(scl3(wo, -1.0f), scl3(wh, 2.0f * dot3(wo, wh)));
                if (!same_hemi(wo, wi)) return f3{0.0f, 0.0f, 0.0f};
            }
            pdf = blend_pdf(m, wo, wi);
            return blend_eval(m, wo, wi);
        }
        case L_CLEARCOAT: {
            const float eta = 1.0f / m.coat_ior;
            const float F0 = fr_dielectric(wo.z, 1.0f, m.coat_ior);
            const float w = m.coat_weight;
            const float ps = F0 * w / (F0 * w + (1.0f - F0) * (1.0f - w));
            f2 smp = u;
            bool spec;
            if (smp.y < ps) {
                smp.y /= ps;
                spec = true;
            } else {
                smp.y = (smp.y - ps) / (1.0f - ps);
                spec = false;
            }
            if (spec) {
                // the coat: CudaSpecularReflection(BSDF_DIELECTRIC) (noor/cuda/bxdf.cuh:377-391)
                wi = refl_z(wo);
                sampled = B_DIELECTRIC | B_REFLECTION | B_SPECULAR;
                const float F = fr_dielectric(wi.z, 1.0f, m.coat_ior);
                pdf = ps;
                return sdiv3(scl3(ld3(m.diffuse), F), fabsf(wi.z));
            }
            const f3 n = f3{0.0f, 0.0f, 1.0f};
            f3 wo_c = f3{0.0f, 0.0f, 0.0f}, wi_c = f3{0.0f, 0.0f, 0.0f};
            refract_(wo, n, eta, wo_c);
            sampled = B_CONDUCTOR | B_REFLECTION | B_GLOSSY;
            float spdf = 0.0f;
            f3 r = micro_sample(B_CONDUCTOR | B_REFLECTION | B_GLOSSY, m, scl3(wo_c, -1.0f), wi_c, smp, spdf);
            const float F1 = fr_dielectric(-wi_c.z, 1.0f, m.coat_ior);
            refract_(scl3(wi_c, -1.0f), scl3(n, -1.0f), 1.0f / eta, wi);
            if (F1 == 1.0f) {
                pdf = 0.0f;
                return f3{0.0f, 0.0f, 0.0f};
            }
            const float k = coat_absorption(m, wo_c, wi_c) * (1.0f - F0) * (1.0f - F1);
            r = sdiv3(scl3(r, k), fabsf(wi_c.z));
            pdf = spdf * (1.0f - ps);
            return r;
        }
        default: return f3{0.0f, 0.0f, 0.0f};
    }
}

// CudaBSDF::Eval / Pdf / Sample_f (noor/cuda/bsdf.cuh:259-356)
template <bool LOCAL = false>
__device__ __forceinline__ f3 bsdf_eval(const Bsdf& b, const Isect& I, const DevMat& m, f3 woW, f3 wiW, int flags) {
    const f3 wo = to_local(I, woW), wi = to_local(I, wiW);
    const bool reflect = dot3(wiW, I.gn) * dot3(woW, I.gn) > 0.0f;
    f3 f = f3{0.0f, 0.0f, 0.0f};
    for (int i = 0; i < b.n; ++i)
        if (lobe_matches(b.l[i].type, flags) &&
            ((reflect && (b.l[i].type & B_REFLECTION)) || (!reflect && (b.l[i].type & B_TRANSMISSION))))
            f = add3(f, lobe_eval<LOCAL>(b.l[i], I, m, wo, wi));
    return f;
}
template <bool LOCAL = false>
__device__ __forceinline__ float bsdf_pdf(const Bsdf& b, const Isect& I, const DevMat& m, f3 woW, f3 wiW, int flags) {
    if (b.n == 0) return 0.0f;
    const f3 wo = to_local(I, woW), wi = to_local(I, wiW);
    if (wo.z == 0.0f) return 0.0f;
    float pdf = 0.0f;
    int matching = 0;
    for (int i = 0; i < b.n; ++i)
        if (lobe_matches(b.l[i].type, flags)) {
            ++matching;
            pdf += lobe_pdf<LOCAL>(b.l[i], m, wo, wi);
        }
    return matching > 0 ? pdf / (float)matching : 0.0f;
}
template <bool LOCAL = false>
__device__ __forceinline__ f3 bsdf_sample(const Bsdf& b, const Isect& I, const DevMat& m, f3 woW, f3& wiW, float& pdf,
                          int flags, int& sampled, Rng& rng, int rk, int* lobe_out = nullptr) {
    int matching = 0;
    for (int i = 0; i < b.n; ++i)
        if (lobe_matches(b.l[i].type, flags)) ++matching;
    if (matching == 0) {
        pdf = 0.0f;
        sampled = 0;
        return f3{0.0f, 0.0f, 0.0f};
    }
    f2 u;
    u.x = rng_next(rng, rk);
    u.y = rng_next(rng, rk);
    int comp = (int)floorf(u.x * (float)matching);
    if (comp > matching - 1) comp = matching - 1;
    const f2 ur = f2{rmin(u.x * (float)matching - (float)comp, K_ONE_MINUS_EPS), u.y};
    int li = -1;
    int count = comp;
    for (int i = 0; i < b.n; ++i)
        if (lobe_matches(b.l[i].type, flags) && count-- == 0) {
            li = i;
            break;
        }
    const Lobe L = (li == 1) ? b.l[1] : b.l[0];  // a BSDF has at most two lobes
    if (lobe_out) *lobe_out = li;
    const f3 wo = to_local(I, woW);
    if (wo.z == 0.0f) return f3{0.0f, 0.0f, 0.0f};
    pdf = 0.0f;
    sampled = L.type;
    f3 wi = f3{0.0f, 0.0f, 0.0f};
    f3 f = lobe_sample<LOCAL>(L, I, m, wo, wi, ur, pdf, sampled);
    if (pdf == 0.0f) {
        sampled = 0;
        return f3{0.0f, 0.0f, 0.0f};
    }
    wiW = to_world(I, wi);
    const bool spec = (L.type & B_SPECULAR) != 0;
    if (!spec && matching > 1)
        for (int i = 0; i < b.n; ++i)
            if (i != li && lobe_matches(b.l[i].type, flags)) pdf += lobe_pdf<LOCAL>(b.l[i], m, wo, wi);
    if (matching > 1) pdf /= (float)matching;
    if (!spec) {
        const bool reflect = dot3(wiW, I.gn) * dot3(woW, I.gn) > 0.0f;
        for (int i = 0; i < b.n; ++i)
            if (i != li && lobe_matches(b.l[i].type, flags) &&
                ((reflect && (b.l[i].type & B_REFLECTION)) || (!reflect && (b.l[i].type & B_TRANSMISSION))))
                f = add3(f, lobe_eval<LOCAL>(b.l[i], I, m, wo, wi));
    }
    return f;
}

// ---------------------------------------------------------------------------------------------
// lights (noor/cuda/lightimp.cuh, noor/cuda/geometry.cuh)
struct Vis {
    f3 from, to, wi;
    float dist;
};
__device__ __forceinline__ Vis mkvis(f3 from, f3 to) {
    Vis v;
    v.from = from;
    v.to = to;
    v.wi = sub3(to, from);
    v.dist = len3(v.wi);
    v.wi = sdiv3(v.wi, v.dist);
    return v;
}

// light index -> kind in setup_lights order (noor/cuda/lightimp.cuh:244-281)
__device__ __forceinline__ int light_kind(const DevScene& S, int idx, int& local) {
    int i = idx;
    if (i < (int)S.n_area) { local = i; return NOOR_LIGHT_AREA; }
    i -= (int)S.n_area;
    if (i < (int)S.n_point) { local = i; return NOOR_LIGHT_POINT; }
    i -= (int)S.n_point;
    if (i < (int)S.n_spot) { local = i; return NOOR_LIGHT_SPOT; }
    i -= (int)S.n_spot;
    if (i < (int)S.n_distant) { local = i; return NOOR_LIGHT_DISTANT; }
    local = 0;
    return NOOR_LIGHT_INFINITE;
}
__device__ __forceinline__ bool is_delta(const DevScene& S, int idx) {
    int l;
    const int k = light_kind(S, idx, l);
    return k == NOOR_LIGHT_POINT || k == NOOR_LIGHT_SPOT || k == NOOR_LIGHT_DISTANT;
}
__device__ __forceinline__ f3 area_Le(const noor_area_light& L, f3 w) {
    if (L.shape.type != NOOR_SHAPE_SPHERE)
        return (L.two_sided || dot3(w, fromf3(L.shape.n)) < 0.0f) ? fromf3(L.ke) : f3{0.0f, 0.0f, 0.0f};
    return fromf3(L.ke);
}
__device__ __forceinline__ f3 light_Le(const DevScene& S, f3 w, int idx) {
    int l;
    switch (light_kind(S, idx, l)) {
        case NOOR_LIGHT_AREA: return area_Le(S.area[l], w);
        case NOOR_LIGHT_INFINITE: return S.sky_enabled ? sky_eval(S, w) : f3{0.0f, 0.0f, 0.0f};
        default: return f3{0.0f, 0.0f, 0.0f};
    }
}

__device__ __forceinline__ void shape_sample(const DevScene& S, const noor_shape& sh, const Isect& I, f3& p, float& pdf,
                             Rng& rng) {
    const int rk = S.rng_kind;
    const float bias = S.bias;
    const f3 n = fromf3(sh.n);
    if (sh.type == NOOR_SHAPE_QUAD) {
        const float r1 = rng_next(rng, rk);
        const float r2 = rng_next(rng, rk);
        p = add3(add3(add3(fromf3(sh.center), scl3(fromf3(sh.u), r1)), scl3(fromf3(sh.v), r2)), scl3(n, bias));
        const f3 wi = sub3(I.p, p);
        const float dist2 = dot3(wi, wi);
        if (S.compat & NOOR_COMPAT_QUAD_PDF)
            pdf = dist2 / (absdot3(n, scl3(wi, -1.0f)) * sh.area);
        else
            pdf = dist2 / (absdot3(n, scl3(nrm3(wi), -1.0f)) * sh.area);
        if (is_inf(pdf)) pdf = 0.0f;
    } else if (sh.type == NOOR_SHAPE_DISK) {
        f2 uu;
        uu.x = rng_next(rng, rk);
        uu.y = rng_next(rng, rk);
        const f2 pd = concentric_disk(uu);
        const f3 q = f3{pd.x * sh.radius, pd.y * sh.radius, 0.0f};
        f3 u, v;
        coordinate_system(n, u, v);
        p = add3(add3(add3(add3(scl3(u, q.x), scl3(v, q.y)), scl3(n, q.z)), fromf3(sh.center)), scl3(n, bias));
        const f3 wi = sub3(I.p, p);
        const float dist2 = dot3(wi, wi);
        if (S.compat & NOOR_COMPAT_QUAD_PDF)
            pdf = dist2 / (absdot3(n, scl3(wi, -1.0f)) * sh.area);
        else
            pdf = dist2 / (absdot3(n, scl3(nrm3(wi), -1.0f)) * sh.area);
        if (is_inf(pdf)) pdf = 0.0f;
    } else {
        f2 r;
        r.x = rng_next(rng, rk);
        r.y = rng_next(rng, rk);
        const f3 ctr = fromf3(sh.center);
        const f3 w = nrm3(sub3(ctr, I.p));
        f3 u, v;
        coordinate_system(w, u, v);
        const f3 pc = sub3(I.p, ctr);
        const float sinMax2 = sh.radius2 / dot3(pc, pc);
        const float cosMax = sqrtf(rmax(0.0f, 1.0f - sinMax2));
        const float cosT = (1.0f - r.x) + r.x * cosMax;
        const float sinT = sqrtf(rmax(0.0f, 1.0f - cosT * cosT));
        const float phi = r.y * K_2PI;
        const float dc = len3(pc);
        const float rad = sh.radius;
        const float ds = dc * cosT - sqrtf(rmax(0.0f, sh.radius2 - dc * dc * sinT * sinT));
        const float cosA = (dc * dc + sh.radius2 - ds * ds) / (2.0f * dc * rad);
        const float sinA = sqrtf(rmax(0.0f, 1.0f - cosA * cosA));
        const f3 N = add3(add3(scl3(neg3(u), sinA * nm_cosf(phi)), scl3(neg3(v), sinA * nm_sinf(phi))),
                          scl3(neg3(w), cosA));
        f3 pp = add3(ctr, scl3(N, rad));
        pp = scl3(pp, rad / len3(sub3(pp, ctr)));
        p = add3(pp, scl3(N, bias));
        pdf = K_INV_2PI / (1.0f - cosMax);
    }
}

// CudaLightManager::sample_Li (noor/cuda/lightimp.cuh:415-456)
__device__ __forceinline__ f3 light_sample_Li(const DevScene& S, const Isect& I, int& light_idx, Vis& vis, float& pdf, Rng& rng) {
    pdf = 0.0f;
    const int num_lights = S.sky_enabled ? S.num_lights : S.num_lights - 1;
    if (num_lights <= 0) return f3{0.0f, 0.0f, 0.0f};
    const int idx = (int)((float)num_lights * rng_next(rng, S.rng_kind));
    light_idx = idx;
    int l;
    f3 Ld = f3{0.0f, 0.0f, 0.0f};
    const float wr = S.world_radius;
    switch (light_kind(S, idx, l)) {
        case NOOR_LIGHT_AREA: {
            const noor_area_light& L = S.area[l];
            f3 p;
            shape_sample(S, L.shape, I, p, pdf, rng);
            if (pdf == 0.0f) break;
            vis = mkvis(I.p, p);
            Ld = area_Le(L, vis.wi);
            break;
        }
        case NOOR_LIGHT_POINT: {
            const noor_point_light& L = S.point[l];
            pdf = 1.0f;
            vis = mkvis(I.p, fromf3(L.position));
            const f3 d = sub3(vis.to, vis.from);
            const float d2 = dot3(d, d);
            if (d2 == 0.0f) {
                pdf = 0.0f;
                break;
            }
            Ld = sdiv3(fromf3(L.ke), d2);
            break;
        }
        case NOOR_LIGHT_SPOT: {
            const noor_spot_light& L = S.spot[l];
            pdf = 1.0f;
            vis = mkvis(I.p, fromf3(L.position));
            const f3 d = sub3(vis.to, vis.from);
            const float d2 = dot3(d, d);
            if (d2 == 0.0f) {
                pdf = 0.0f;
                break;
            }
            const float ct = dot3(fromf3(L.direction), vis.wi);
            float fo;
            if (ct < L.cosouter) fo = 0.0f;
            else if (ct >= L.cosinner) fo = 1.0f;
            else {
                const float dl = (ct - L.cosouter) / (L.cosinner - L.cosouter);
                fo = (dl * dl) * (dl * dl);
            }
            if (fo == 0.0f) {
                pdf = 0.0f;
                break;
            }
            Ld = sdiv3(scl3(fromf3(L.ke), fo), d2);
            break;
        }
        case NOOR_LIGHT_DISTANT: {
            const noor_distant_light& L = S.distant[l];
            pdf = 1.0f;
            vis = mkvis(I.p, add3(I.p, scl3(fromf3(L.direction), 2.0f * wr)));
            Ld = fromf3(L.ke);
            break;
        }
        default: {
            const f3 wi = env_sample_dir(S, rng, pdf);
            if (is_inf(pdf) || pdf != pdf) pdf = 0.0f;
            vis = mkvis(I.p, add3(I.p, scl3(wi, 2.0f * wr)));
            Ld = S.sky_enabled ? sky_eval(S, vis.wi) : f3{0.0f, 0.0f, 0.0f};
            break;
        }
    }
    if (is_inf(pdf)) pdf = 0.0f;
    return scl3(Ld, (float)num_lights);
}

__device__ __forceinline__ float shape_pdf_solid(const noor_shape& sh, const Isect& I, f3 wi) {
    float pdf = 0.0f;
    if (sh.type == NOOR_SHAPE_QUAD || sh.type == NOOR_SHAPE_DISK) {
        Ray r = mkray(I.p, wi, K_INF);
        if (!shape_intersect(sh, r)) return 0.0f;
        const f3 p = ray_at(r, r.tmax);
        const f3 d = sub3(p, I.p);
        pdf = dot3(d, d) / (absdot3(fromf3(sh.n), scl3(r.d, -1.0f)) * sh.area);
    } else {
        const f3 pc = sub3(I.p, fromf3(sh.center));
        const float sinMax2 = sh.radius2 / dot3(pc, pc);
        const float cosMax = sqrtf(rmax(0.0f, 1.0f - sinMax2));
        pdf = K_INV_2PI / (1.0f - cosMax);
    }
    return is_inf(pdf) ? 0.0f : pdf;
}
// CudaLightManager::pdf_Li (noor/cuda/lightimp.cuh:458-483)
__device__ __forceinline__ float light_pdf_Li(const DevScene& S, const Isect& I, f3 wi, int idx) {
    int l;
    float pdf = 0.0f;
    switch (light_kind(S, idx, l)) {
        case NOOR_LIGHT_AREA: pdf = shape_pdf_solid(S.area[l].shape, I, wi); break;
        case NOOR_LIGHT_INFINITE: {
            if (!S.sky_enabled) return 0.0f;
            const float theta = spherical_theta(wi), phi = spherical_phi(wi);
            const float sinT = nm_sinf(theta);
            if (sinT == 0.0f) return 0.0f;
            pdf = d_pdf(S, f2{phi * K_INV_2PI, theta * K_INV_PI}) / (2.0f * K_PI * K_PI * sinT);
            break;
        }
        default: pdf = 0.0f;
    }
    return is_inf(pdf) ? 0.0f : pdf;
}
// CudaLightManager::intersect (noor/cuda/lightimp.cuh:398-408)
__device__ __forceinline__ bool light_intersect(const DevScene& S, Ray& r, int idx) {
    int l;
    const int k = light_kind(S, idx, l);
    if (k == NOOR_LIGHT_INFINITE && S.sky_enabled) {
        r.tmax = 2.0f * S.world_radius;
        return true;
    }
    if (k != NOOR_LIGHT_AREA) return false;
    return shape_intersect(S.area[l].shape, r);
}

// spawnShadowRay(CudaVisibility) (noor/cuda/intersection.cuh:199-207)
__device__ __forceinline__ Ray shadow_ray_vis(const DevScene& S, const Isect& I, const Vis& v) {
    const f3 p = dot3(v.wi, I.gn) >= 0.0f ? add3(v.from, scl3(I.gn, S.bias)) : sub3(v.from, scl3(I.gn, S.bias));
    const float dist = len3(sub3(p, v.to));
    return mkray(p, v.wi, dist);
}

// ---------------------------------------------------------------------------------------------
// camera rays (noor/cuda/cameraimp.cuh:58-269); with TEX also the ray
// differentials of the neighbouring pixels, which CudaRay::transform carries
// to world space (origins as points, directions as unnormalised vectors,
// noor/cuda/ray.cuh:162-175)
__device__ __forceinline__ f3 env_cam_dir(const noor_camera& cam, float x, float y) {
    const float theta = K_PI * y / (float)cam.h + K_PI;
    const float phi = K_2PI * x / (float)cam.w - K_PI_OVER_2;
    const float sT = nm_sinf(theta), cT = nm_cosf(theta);
    return nrm3(f3{sT * nm_cosf(phi), cT, sT * nm_sinf(phi)});
}

template <bool TEX = false>
__device__ __forceinline__ Ray gen_ray(const noor_camera& cam, uint32_t x, uint32_t y, Rng& rng, int rk,
                                       RayDiff* rd = nullptr) {
    const float rx = (float)x + rng_next(rng, rk) - 0.5f;
    const float ry = (float)y + rng_next(rng, rk) - 0.5f;
    const Xf c2w = xf_from(cam.camera_to_world);
    const Xf r2c = xf_from(cam.raster_to_camera);
    f3 o = f3{0.0f, 0.0f, 0.0f}, d;
    f3 ox, oy, dx, dy;
    const f3 dxc = f3{cam.dx_camera.x, cam.dx_camera.y, cam.dx_camera.z};
    const f3 dyc = f3{cam.dy_camera.x, cam.dy_camera.y, cam.dy_camera.z};
    if (cam.type == NOOR_CAMERA_ORTHO) {
        o = xf_pt(r2c, f3{rx, ry, 0.0f});
        d = f3{0.0f, 0.0f, -1.0f};
        if (TEX) {
            ox = add3(o, dxc);
            oy = add3(o, dyc);
            dx = dy = d;
        }
    } else if (cam.type == NOOR_CAMERA_ENV) {
        d = nrm3(env_cam_dir(cam, rx, ry));
        if (TEX) {
            ox = oy = o;
            dx = env_cam_dir(cam, rx + 1.0f, ry);
            dy = env_cam_dir(cam, rx, ry + 1.0f);
        }
    } else {
        const f3 pc = xf_pt(r2c, f3{rx, ry, 0.0f});
        d = nrm3(pc);
        if (cam.lens_radius > 0.0f) {
            f2 u;
            u.x = rng_next(rng, rk);
            u.y = rng_next(rng, rk);
            const f2 pl = concentric_disk(u);
            const float lr = cam.lens_radius;
            const float ft = -cam.focal_length / d.z;
            const f3 pf = add3(o, scl3(d, ft));
            o = f3{lr * pl.x, lr * pl.y, 0.0f};
            d = nrm3(sub3(pf, o));
            if (TEX) {
                const f3 ddx = nrm3(add3(pc, dxc));
                const f3 pfx = add3(f3{0.0f, 0.0f, 0.0f}, scl3(ddx, -cam.focal_length / ddx.z));
                const f3 ddy = nrm3(add3(pc, dyc));
                const f3 pfy = add3(f3{0.0f, 0.0f, 0.0f}, scl3(ddy, -cam.focal_length / ddy.z));
                ox = oy = o;
                dx = nrm3(sub3(pfx, ox));
                dy = nrm3(sub3(pfy, oy));
            }
        } else if (TEX) {
            ox = oy = o;
            dx = nrm3(add3(pc, dxc));
            dy = nrm3(add3(pc, dyc));
        }
    }
    if (TEX) {
        rd->ox = xf_pt(c2w, ox);
        rd->oy = xf_pt(c2w, oy);
        rd->dx = xf_vec(c2w, dx);
        rd->dy = xf_vec(c2w, dy);
        rd->has = 1;
    }
    return mkray(xf_pt(c2w, o), nrm3(xf_vec(c2w, d)), K_INF);
}

// ---------------------------------------------------------------------------------------------
// one bounce of the path tracer (noor/cuda/path_tracer.cu:36-72,
// noor/cuda/accumulate.cuh:28-99, noor/cuda/direct.cuh:33-132,
// noor/cuda/scatter.cuh:27-41).
//
// Next-event estimation is split in two halves so that the occlusion test can
// run elsewhere (inline in the megakernel, in the shadow kernel of the
// wavefront renderer): shade_bounce() returns up to two candidate
// contributions with their shadow rays, and resolve_nee() adds
// beta_prev * (visible_a + visible_b) to L exactly as `L += beta * direct`.
struct PathState {
    f3 L, beta;
    f3 Lc;  // rounding error of the additions into L (compensated paths only)
    Rng rng;
    f3 o, d;
    uint32_t bounce;
    uint32_t specular;
};

struct Nee {
    f3 a, b, beta;     // candidate contributions (before occlusion) and the throughput they scale
    Ray ra, rb;        // their shadow rays
    int na, nb;        // 1 when the candidate needs its occlusion test
    int has;           // 1 when `L += beta * Ld` runs for this bounce
};

// s = a + b, with the rounding error of the addition added to *err
// (Knuth's TwoSum: exact in binary32 without contraction).
__device__ __forceinline__ float two_sum(float a, float b, float& err) {
    const float s = a + b;
    const float bb = s - a;
    err += (a - (s - bb)) + (b - bb);
    return s;
}

// L += a.  COMP: the rounding errors are kept in Lc, so that L + Lc is the
// sum of the addends to about eps^2 while L itself is bit-identical to the
// plain sum.  The adjoint needs the radiance a path gathers after a bounce as
// (total - gathered so far): in binary32 alone that difference loses all its
// digits when the rest is small against the total.
template <bool COMP>
__device__ __forceinline__ void add_L(PathState& ps, f3 a) {
    if (!COMP) {
        ps.L = add3(ps.L, a);
        return;
    }
    ps.L.x = two_sum(ps.L.x, a.x, ps.Lc.x);
    ps.L.y = two_sum(ps.L.y, a.y, ps.Lc.y);
    ps.L.z = two_sum(ps.L.z, a.z, ps.Lc.z);
}

template <bool COMP = false>
__device__ __forceinline__ void resolve_nee(PathState& ps, const Nee& n, bool occ_a, bool occ_b) {
    if (!n.has) return;
    f3 Ld = (n.na && !occ_a) ? n.a : f3{0.0f, 0.0f, 0.0f};
    if (n.nb) Ld = add3(Ld, occ_b ? f3{0.0f, 0.0f, 0.0f} : n.b);
    add_L<COMP>(ps, mul3(n.beta, Ld));
}

// Returns true when the path continues with the ray stored in ps.
// ---------------------------------------------------------------------------------------------
// parameter derivatives for the adjoint

// d ln(D G) / d alpha of an isotropic GGX lobe (alpha_x = alpha_y), with
// D = 1 / (pi a^2 cos^4 (1 + e)^2), e = tan^2(wh) / a^2 and
// Lambda(w) = (sqrt(1 + a^2 tan^2 w) - 1) / 2:
//   d ln D / da = (2 / a) (e - 1) / (e + 1)
//   d Lambda / da = a tan^2 / (2 sqrt(1 + a^2 tan^2)),  d ln G / da = -(dLo + dLi) G
__device__ __forceinline__ float ggx_dlog_dalpha(float a, f3 wo, f3 wi, f3 wh) {
    float dD = 0.0f;
    const float t2 = tan2t(wh);
    if (!is_inf(t2)) {
        const float e = t2 / (a * a);
        dD = (2.0f / a) * (e - 1.0f) / (e + 1.0f);
    }
    float lam = 0.0f, dlam = 0.0f;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const float tt = fabsf(tant(k == 0 ? wo : wi));
        if (is_inf(tt)) continue;
        const float q = sqrtf(1.0f + a * a * tt * tt);
        lam += (q - 1.0f) * 0.5f;
        dlam += a * tt * tt / (2.0f * q);
    }
    return dD - dlam / (1.0f + lam);
}

// Derivatives of the f that bsdf_eval (lobe < 0, `flags`) or a sample of lobe
// `lobe` (bsdf_sample, flags B_ALL) sums, with respect to the diffuse colour
// (dR, per channel: d f_c / d diffuse_c) and the roughness r of an isotropic
// material (dr, alpha = max(r^2, 1e-3)).
__device__ __forceinline__ void bsdf_grad(const Bsdf& b, const Isect& I, const DevMat& m, f3 woW, f3 wiW, int flags,
                                          int lobe, f3& dR, f3& dr) {
    dR = f3{0.0f, 0.0f, 0.0f};
    dr = f3{0.0f, 0.0f, 0.0f};
    const f3 wo = to_local(I, woW), wi = to_local(I, wiW);
    const bool reflect = dot3(wiW, I.gn) * dot3(woW, I.gn) > 0.0f;
    bool spec = false;
#pragma unroll
    for (int i = 0; i < 2; ++i)
        if (i < b.n && i == lobe && (b.l[i].type & B_SPECULAR)) spec = true;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        if (i >= b.n) break;
        const Lobe& L = b.l[i];
        bool inc;
        if (i == lobe) inc = true;
        else if (spec) inc = false;
        else
            inc = lobe_matches(L.type, flags) &&
                  ((reflect && (L.type & B_REFLECTION)) || (!reflect && (L.type & B_TRANSMISSION)));
        if (!inc) continue;
        switch (L.id) {
            case L_LAMBERT_R:
            case L_LAMBERT_T: dR = add3(dR, f3{K_INV_PI, K_INV_PI, K_INV_PI}); break;
            case L_SPEC_R_NOOP: {
                if (wi.z == 0.0f) break;
                const f3 F = fresnel_eval(L.type, m, wi.z);
                dR = add3(dR, sdiv3(F, fabsf(wi.z)));
                break;
            }
            case L_MICRO_R_DIEL:
            case L_MICRO_R_COND: {
                const float r = m.rough[0];
                if (!(r * r > 1e-3f)) break;
                const f3 fm = micro_eval(L.type, m, wo, wi);
                if (is_black(fm)) break;
                f3 wh = nrm3(add3(wi, wo));
                wh = scl3(wh, (float)isign(wh.z));
                const float dl = ggx_dlog_dalpha(r * r, wo, wi, wh);
                dr = add3(dr, scl3(fm, dl * 2.0f * r));
                break;
            }
            default: break;
        }
    }
}

// Texels and weights of the bilinear environment lookup of sky_eval (HDR).
struct EnvTaps {
    uint32_t idx[4];
    float w[4];
};
__device__ __forceinline__ EnvTaps env_taps(const DevScene& S, f3 dir) {
    EnvTaps t;
    float u = spherical_phi(dir) * K_INV_2PI, v = spherical_theta(dir) * K_INV_PI;
    const int w = (int)S.env_w, h = (int)S.env_h;
    if (u != u) u = 0.0f;
    if (v != v) v = 0.0f;
    float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    x = clampr(x, -2.0f, (float)w + 1.0f);
    y = clampr(y, -2.0f, (float)h + 1.0f);
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    const int ix = (int)fx0, iy = (int)fy0;
    const int i0 = ix < 0 ? 0 : (ix > w - 1 ? w - 1 : ix);
    const int i1 = ix + 1 < 0 ? 0 : (ix + 1 > w - 1 ? w - 1 : ix + 1);
    const int j0 = iy < 0 ? 0 : (iy > h - 1 ? h - 1 : iy);
    const int j1 = iy + 1 < 0 ? 0 : (iy + 1 > h - 1 ? h - 1 : iy + 1);
    const float bx = 1.0f - ax, by = 1.0f - ay;
    t.idx[0] = (uint32_t)(j0 * w + i0);
    t.w[0] = by * bx;
    t.idx[1] = (uint32_t)(j0 * w + i1);
    t.w[1] = by * ax;
    t.idx[2] = (uint32_t)(j1 * w + i0);
    t.w[2] = ay * bx;
    t.idx[3] = (uint32_t)(j1 * w + i1);
    t.w[3] = ay * ax;
    return t;
}

// What the adjoint replay needs from one bounce of shade_bounce(): the sky
// lookup of an escaping path, the light sample of NEE and the BSDF sample
// that continued the path, with the hit record they were evaluated at.
struct Tape {
    int esc;          // the sky was added for an escaping ray
    f3 esc_dir, esc_beta;
    int nee;          // NEE candidate `a` was formed
    int nee_light;    // its light index
    f3 nee_wi, nee_Li;
    float nee_pdf;
    int scat;         // a BSDF sample continued (or ended) the path
    int scat_lobe;
    f3 scat_wi;
    f3 scat_f;
    Isect I;
};

// TEX: image textures, bump and ray differentials; `rd` holds the
// differentials of `ray` and receives those of the next ray.
template <bool TEX = false, bool COMP = false>
__device__ __forceinline__ bool shade_bounce(const DevScene& S, PathState& ps, const Ray& ray, bool hit, const Hit& h, Nee& nee,
                             float4* lookat, Tape* tp = nullptr, RayDiff* rd = nullptr) {
    const int rk = S.rng_kind;
    nee.has = 0;
    nee.na = 0;
    nee.nb = 0;
    if (tp) {
        tp->esc = 0;
        tp->nee = 0;
        tp->scat = 0;
    }
    if (!hit) {
        if (S.sky_enabled && (ps.bounce == 0 || ps.specular)) {
            if (tp) {
                tp->esc = 1;
                tp->esc_dir = nrm3(ray.d);
                tp->esc_beta = ps.beta;
            }
            add_L<COMP>(ps, mul3(ps.beta, sky_eval(S, nrm3(ray.d))));
        }
        return false;
    }
    Isect I;
    HitDiff hd;
    make_isect<TEX>(S, ray, h, I, rd, &hd);
    // autofocus point: the first hit of the sample, w = 1 (noor/cuda/path_tracer.cu:40-42)
    if (ps.bounce == 0 && lookat) *lookat = make_float4(I.p.x, I.p.y, I.p.z, 1.0f);
    // the material with its image textures looked up (TEX), else the constants
    DevMat mt;
    if (TEX) {
        mt = S.mats[I.mat];
        if (!(I.type & NOOR_MAT_MESHLIGHT)) resolve_material(S, I.u, I.v, hd.g, mt);
    }
    if (is_emitter(I) && (ps.bounce == 0 || ps.specular)) {
        const f3 Le = (I.type & NOOR_MAT_MESHLIGHT) ? light_Le(S, ray.d, (int)I.ins)
                                                    : ld3(TEX ? mt.emittance : S.mats[I.mat].emittance);
        add_L<COMP>(ps, mul3(ps.beta, Le));
        return false;
    }
    // accumulate: bump before the BSDF is built (noor/cuda/accumulate.cuh:84-99)
    if (TEX && rd->has && (I.type & NOOR_MAT_BUMP)) {
        const Frame F = bump_frame(S, S.mats[I.mat], I.u, I.v, hd.g, Frame{I.ss, I.ts, I.sn}, I.gn);
        I.ss = F.ss;
        I.ts = F.ts;
        I.sn = F.sn;
    }
    Bsdf b;
    make_bsdf(I.type, b);
    const DevMat& m = TEX ? mt : S.mats[I.mat];
    // sampleLight (noor/cuda/direct.cuh:33-67)
    nee.has = 1;
    nee.beta = ps.beta;
    nee.a = f3{0.0f, 0.0f, 0.0f};
    nee.b = f3{0.0f, 0.0f, 0.0f};
    int light_idx = -1;
    bool light_ok = false;
    if (!is_specular(I)) {
        const int flags = B_ALL & ~B_SPECULAR;
        Vis vis;
        float light_pdf = 0.0f;
        const f3 Li = light_sample_Li(S, I, light_idx, vis, light_pdf, ps.rng);
        if (!(light_pdf == 0.0f || is_black(Li))) {
            light_ok = true;
            f3 f = bsdf_eval<TEX>(b, I, m, I.wo, vis.wi, flags);
            if (!is_black(f)) {
                f = scl3(f, absdot3(vis.wi, I.sn));
                f3 c;
                if (is_shadowcatcher(I)) c = f;
                else if (is_delta(S, light_idx) || !S.mis) c = sdiv3(mul3(Li, f), light_pdf);
                else {
                    const float sp = bsdf_pdf<TEX>(b, I, m, I.wo, vis.wi, flags);
                    const float w = power_heuristic(light_pdf, sp);
                    c = sdiv3(scl3(mul3(Li, f), w), light_pdf);
                }
                nee.a = c;
                nee.ra = shadow_ray_vis(S, I, vis);
                nee.na = 1;
                if (tp) {
                    tp->I = I;
                    tp->nee = 1;
                    tp->nee_light = light_idx;
                    tp->nee_wi = vis.wi;
                    tp->nee_Li = Li;
                    tp->nee_pdf = light_pdf;
                }
            }
        }
    }
    // sampleBSDF (noor/cuda/direct.cuh:69-113), MIS only
    if (S.mis && !is_shadowcatcher(I) && light_ok && !is_delta(S, light_idx)) {
        const int flags = is_specular(I) ? B_ALL : (B_ALL & ~B_SPECULAR);
        if (S.compat & NOOR_COMPAT_WASTED_DRAWS) {
            rng_next(ps.rng, rk);
            rng_next(ps.rng, rk);
        }
        float spdf = 0.0f;
        int sampled = 0;
        f3 wi = f3{0.0f, 0.0f, 0.0f};
        f3 f = bsdf_sample<TEX>(b, I, m, I.wo, wi, spdf, flags, sampled, ps.rng, rk);
        if (!(is_black(f) || spdf == 0.0f)) {
            f = scl3(f, absdot3(wi, I.sn));
            float sw = 1.0f;
            bool ok = true;
            if (!(sampled & B_SPECULAR)) {
                const float lp = light_pdf_Li(S, I, wi, light_idx);
                if (lp == 0.0f) ok = false;
                else sw = power_heuristic(spdf, lp);
            }
            if (ok) {
                const f3 o = dot3(wi, I.gn) >= 0.0f ? add3(I.p, scl3(I.gn, S.bias)) : sub3(I.p, scl3(I.gn, S.bias));
                Ray sr = mkray(o, wi, 2.0f * S.world_radius);
                if (light_intersect(S, sr, light_idx)) {
                    // the reference aims the shadow ray at the light point itself, where the
                    // light can block it; the corrected mode stops one bias short
                    const float te = (S.compat & NOOR_COMPAT_MIS_SELF_SHADOW) ? sr.tmax : sr.tmax - S.bias;
                    const Vis v2 = mkvis(I.p, ray_at(sr, te));
                    nee.b = sdiv3(scl3(mul3(light_Le(S, wi, light_idx), f), sw), spdf);
                    // the light was chosen with probability 1 / (light count); the
                    // reference leaves this half of the estimate unscaled
                    if (!(S.compat & NOOR_COMPAT_MIS_LIGHT_COUNT))
                        nee.b = scl3(nee.b, (float)(S.sky_enabled ? S.num_lights : S.num_lights - 1));
                    nee.rb = shadow_ray_vis(S, I, v2);
                    nee.nb = 1;
                }
            }
        }
    }
    // scatter (noor/cuda/scatter.cuh:27-41)
    if (S.compat & NOOR_COMPAT_WASTED_DRAWS) {
        rng_next(ps.rng, rk);
        rng_next(ps.rng, rk);
    }
    float pdf = 0.0f;
    int sampled = 0;
    int lobe = -1;
    f3 wi = I.wi;
    const f3 f = bsdf_sample<TEX>(b, I, m, I.wo, wi, pdf, B_ALL, sampled, ps.rng, rk, tp ? &lobe : nullptr);
    I.wi = wi;
    if (tp) {
        tp->scat = (pdf != 0.0f) ? 1 : 0;
        tp->scat_lobe = lobe;
        tp->scat_wi = wi;
        tp->scat_f = f;
        tp->I = I;
    }
    const f3 sc = (pdf == 0.0f) ? f3{0.0f, 0.0f, 0.0f} : sdiv3(scl3(f, absdot3(I.wi, I.sn)), pdf);
    ps.beta = mul3(ps.beta, sc);
    // correct_sidedness (noor/cuda/accumulate.cuh:41-45)
    if (!is_transparent(I) && dot3(I.wi, I.gn) < 0.0f) return false;
    // rr_terminate (noor/cuda/accumulate.cuh:28-38)
    if ((int)ps.bounce >= S.rr) {
        const float q = rmax(0.05f, 1.0f - maxcomp3(ps.beta));
        if (rng_next(ps.rng, rk) < q) return false;
        ps.beta = sdiv3(ps.beta, 1.0f - q);
    }
    ps.specular = is_specular(I) ? 1u : 0u;
    // spawnRay (noor/cuda/intersection.cuh:115-189)
    ps.o = dot3(I.wi, I.gn) >= 0.0f ? add3(I.p, scl3(I.gn, S.bias)) : sub3(I.p, scl3(I.gn, S.bias));
    ps.d = I.wi;
    if (TEX) {
        if (rd->has && (I.type & M_GLOSSY)) spawn_diff(S, I, hd, ps.o, *rd);
        else rd->has = 0;
    }
    return true;
}

// clamp (noor/cuda/path_tracer.cu:73) and the running mean of the film
// (noor/cuda/framebuffer.cuh:66-76)
__device__ __forceinline__ f3 final_radiance(const DevScene& S, f3 L) {
    if (S.compat & NOOR_COMPAT_CLAMP) {
        L.x = clampr(L.x, 0.0f, 100.f);
        L.y = clampr(L.y, 0.0f, 100.f);
        L.z = clampr(L.z, 0.0f, 100.f);
    }
    return L;
}
__device__ __forceinline__ float4 film_lerp(float4 old, f3 c, uint32_t frame) {
    if (frame <= 1) old = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    const float inv = 1.0f / (float)frame;
    float4 r;
    r.x = old.x + inv * (c.x - old.x);
    r.y = old.y + inv * (c.y - old.y);
    r.z = old.z + inv * (c.z - old.z);
    r.w = old.w + inv * (1.0f - old.w);
    return r;
}

}  // namespace noor
