// Megakernel renderer (the drop-in for path_tracer_kernel,
// noor/cuda/path_tracer.cu:77-100) and the hit-query test hook.
//
// Launch shape: 1-D blocks of 64 lanes; each wave renders an 8x8 pixel tile
// so that primary rays of a wave are coherent.  A launch renders `spp`
// progressive frames and keeps the pixel's film value in registers between
// them; the per-frame running mean is applied in frame order, so the film is
// identical to `spp` single-frame launches.  The traversal stack lives in LDS,
// interleaved by lane.
#include <algorithm>

#include "ctx.h"

namespace noor {

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void flush_counters(unsigned long long* dst, const Counters& c) {
    const unsigned long long v[4] = {wave_sum(c.nodes), wave_sum(c.tris), wave_sum(c.tri2), wave_sum(c.inst)};
    if ((threadIdx.x & 63) == 0)
        for (int i = 0; i < 4; ++i)
            if (v[i]) atomicAdd(&dst[i], v[i]);
}

// `hits` (closest-hit queries that hit) is counted with STATS only, else 0.
__device__ __forceinline__ void flush_stats(DevStats* st, unsigned long long ext, unsigned long long shd,
                                            unsigned long long hits, const Counters& ce, const Counters& cs,
                                            bool overflow) {
    const unsigned long long e = wave_sum(ext), s = wave_sum(shd), ov = wave_sum(overflow ? 1ull : 0ull);
    const unsigned long long nh = wave_sum(hits);
    if ((threadIdx.x & 63) == 0) {
        DevStats::Slot& sl = st->slot[blockIdx.x % kStatSlots];  // one wave per block
        if (e) atomicAdd(&sl.extension_rays, e);
        if (s) atomicAdd(&sl.shadow_rays, s);
        if (ov) atomicAdd(&st->overflow, ov);
        if (nh) atomicAdd(&st->ext_hits, nh);
    }
    flush_counters(st->ext, ce);
    flush_counters(st->shd, cs);
}

// DIRECT: write each pixel's last sample to `direct` instead of updating the
// film (a separate instantiation keeps the film path's register budget).
// Two waves per SIMD are required: at one the latency of the traversal's
// node loads is exposed and the kernel runs about 1.7x slower.
// TEX: image textures, alpha tests, bump and ray differentials.
template <bool STATS, bool DIRECT, bool TEX = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_megakernel(const DevScene* __restrict__ Sp, float4* __restrict__ film, uint32_t frame0,
                                                   uint32_t spp, DevStats* __restrict__ stats,
                                                   float4* __restrict__ lookat_out, float4* __restrict__ direct,
                                                   float4* __restrict__ direct_lo) {
    const DevScene& S = *Sp;
    extern __shared__ uint32_t smem[];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t W = S.cam.w, H = S.cam.h;
    const uint32_t tiles_x = (W + 7) / 8, tiles = tiles_x * ((H + 7) / 8);
    const TravStack st = make_stack(S, smem);
    Counters cnt{0, 0, 0, 0}, cns{0, 0, 0, 0};
    unsigned long long ext = 0, shd = 0, nhit = 0;
    bool overflow = false;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t x = (tile % tiles_x) * 8 + (lane & 7);
        const uint32_t y = (tile / tiles_x) * 8 + (lane >> 3);
        if (x >= W || y >= H) continue;
        const uint32_t gid = y * W + x;
        float4 acc = DIRECT ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : film[gid];
        // the autofocus point of the centre pixel's last sample: its first hit,
        // or zero when that sample's camera ray escapes
        float4 lookat = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const bool center = (gid == S.cam.center);
        for (uint32_t s = 0; s < spp; ++s) {
            const uint32_t frame = frame0 + s;
            lookat = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            PathState ps;
            ps.L = f3{0.0f, 0.0f, 0.0f};
            ps.Lc = f3{0.0f, 0.0f, 0.0f};
            ps.beta = f3{1.0f, 1.0f, 1.0f};
            ps.bounce = 0;
            ps.specular = 0;
            rng_init(ps.rng, S.rng_kind, gid, frame);
            RayDiff rd;
            Ray ray = gen_ray<TEX>(S.cam, x, y, ps.rng, S.rng_kind, &rd);
            for (int b = 0; b < S.bounces; ++b) {
                Hit h;
                ++ext;
                const bool hit = bvh_closest<STATS, TEX>(S, ray, h, st, cnt, overflow);
                if (STATS && hit) ++nhit;
                Nee nee;
                const bool cont =
                    shade_bounce<TEX, DIRECT>(S, ps, ray, hit, h, nee, center ? &lookat : nullptr, nullptr, &rd);
                bool oa = false, ob = false;
                if (nee.na) {
                    ++shd;
                    oa = bvh_any<STATS, TEX>(S, nee.ra, st, cns, overflow);
                }
                if (nee.nb) {
                    ++shd;
                    ob = bvh_any<STATS, TEX>(S, nee.rb, st, cns, overflow);
                }
                resolve_nee<DIRECT>(ps, nee, oa, ob);
                if (!cont) break;
                ray = mkray(ps.o, ps.d, K_INF);
                ps.bounce++;
            }
            if (DIRECT) {
                const f3 c = final_radiance(S, ps.L);
                acc = make_float4(c.x, c.y, c.z, 1.0f);
                // the rounding error of the sample's sum (unclamped samples only)
                if (direct_lo) direct_lo[gid] = make_float4(ps.Lc.x, ps.Lc.y, ps.Lc.z, 0.0f);
            } else {
                acc = film_lerp(acc, final_radiance(S, ps.L), frame);
            }
        }
        if (DIRECT) direct[gid] = acc;
        else film[gid] = acc;
        if (center && lookat_out) *lookat_out = lookat;
    }
    flush_stats(stats, ext, shd, nhit, cnt, cns, overflow);
}

hipError_t launch_megakernel(const Context& c, uint32_t frame0, uint32_t spp, float4* direct, float4* direct_lo) {
    hipError_t se = sync_scene(c);
    if (se != hipSuccess) return se;
    const uint32_t tiles = ((c.S.cam.w + 7) / 8) * ((c.S.cam.h + 7) / 8);
    // one wave per 8x8 tile, grid-stride over tiles within the lanes the global stack covers
    const uint32_t blocks = std::min<uint32_t>(tiles, c.S.gstack_lanes / 64);
    const size_t lds = (size_t)64 * c.S.stack_lds * sizeof(uint32_t);
    KernelTimer& T = const_cast<KernelTimer&>(c.timer);
    T.begin(c.stream, "megakernel");
    if (direct && c.S.textured) return hipErrorNotSupported;
    if (direct)
        hipLaunchKernelGGL((k_megakernel<false, true>), dim3(blocks), dim3(64), lds, c.stream, c.d_scene, c.film, frame0,
                           spp, c.d_stats, c.d_lookat, direct, direct_lo);
    else if (c.S.textured && c.opt.stats)
        hipLaunchKernelGGL((k_megakernel<true, false, true>), dim3(blocks), dim3(64), lds, c.stream, c.d_scene, c.film,
                           frame0, spp, c.d_stats, c.d_lookat, direct, direct_lo);
    else if (c.S.textured)
        hipLaunchKernelGGL((k_megakernel<false, false, true>), dim3(blocks), dim3(64), lds, c.stream, c.d_scene, c.film,
                           frame0, spp, c.d_stats, c.d_lookat, direct, direct_lo);
    else if (c.opt.stats)
        hipLaunchKernelGGL((k_megakernel<true, false>), dim3(blocks), dim3(64), lds, c.stream, c.d_scene, c.film, frame0,
                           spp, c.d_stats, c.d_lookat, direct, direct_lo);
    else
        hipLaunchKernelGGL((k_megakernel<false, false>), dim3(blocks), dim3(64), lds, c.stream, c.d_scene, c.film, frame0,
                           spp, c.d_stats, c.d_lookat, direct, direct_lo);
    T.end(c.stream);
    return hipGetLastError();
}

// noor_trace: closest or any hit for a batch of rays.
// `gstack`, `glanes`: the device part of the traversal stack and the lanes it
// covers (the context's, or a caller's scratch); the grid uses at most glanes.
template <bool STATS, bool ALPHA = false>
__global__ __launch_bounds__(64) void k_trace(const DevScene* __restrict__ Sp, const noor_ray* __restrict__ rays, uint32_t n,
                                              int any_hit, noor_hit* __restrict__ hits, DevStats* stats,
                                              uint32_t* gstack, uint32_t glanes) {
    const DevScene& S = *Sp;
    extern __shared__ uint32_t smem[];
    const TravStack st = make_stack(S, smem, gstack, glanes);
    Counters cnt{0, 0, 0, 0}, cns{0, 0, 0, 0};
    bool overflow = false;
    unsigned long long ext = 0, shd = 0, nhit = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const noor_ray r = rays[i];
        Ray ray = mkray(f3{r.o.x, r.o.y, r.o.z}, f3{r.d.x, r.d.y, r.d.z}, r.tmax);
        noor_hit out;
        out.t = out.u = out.v = 0.0f;
        out.prim = out.prim_id = out.inst = 0xFFFFFFFFu;
        out.kind = 0;
        out.pad = 0;
        if (any_hit) {
            ++shd;
            out.kind = bvh_any<STATS, ALPHA>(S, ray, st, cns, overflow) ? 1u : 0u;
        } else {
            ++ext;
            Hit h;
            if (bvh_closest<STATS, ALPHA>(S, ray, h, st, cnt, overflow)) {
                if (STATS) ++nhit;
                out.t = h.t;
                out.u = h.u;
                out.v = h.v;
                out.kind = h.kind;
                out.inst = h.inst;
                if (h.kind == 1) {
                    out.prim = h.prim;
                    out.prim_id = h.key;
                }
            }
        }
        hits[i] = out;
    }
    flush_stats(stats, ext, shd, nhit, cnt, cns, overflow);
}

hipError_t launch_trace(const Context& c, const noor_ray* d_rays, uint32_t n, int any_hit, noor_hit* d_hits,
                        uint32_t* gstack, uint32_t glanes) {
    hipError_t se = sync_scene(c);
    if (se != hipSuccess) return se;
    if (!gstack) {
        gstack = c.S.gstack;
        glanes = c.S.gstack_lanes;
    }
    if (glanes < 64) return hipErrorInvalidValue;
    const size_t lds = (size_t)64 * c.S.stack_lds * sizeof(uint32_t);
    const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>((n + 63) / 64, glanes / 64));
    auto k = c.S.textured ? (c.opt.stats ? k_trace<true, true> : k_trace<false, true>)
                          : (c.opt.stats ? k_trace<true> : k_trace<false>);
    hipLaunchKernelGGL(k, dim3(blocks), dim3(64), lds, c.stream, c.d_scene, d_rays, n, any_hit, d_hits, c.d_stats,
                       gstack, glanes);
    return hipGetLastError();
}

// debug_skydome (noor/cuda/path_tracer.cu:102-122): paints the sky texture
// (the sin(theta)-weighted bake of a physical sky, or the HDR environment)
// over the film, read bilinearly at u = (x / w, y / h) with its four channels
// (noor/cuda/skydome.cuh:87-89); importance-sample dots when MIS is on.
// Pixels hit by a dot are written by two threads in no set order, as in the
// reference.
__global__ void k_debug_skydome(const DevScene* __restrict__ Sp, float4* film, uint32_t frame, const float4* tex,
                                int tw, int th) {
    const DevScene& S = *Sp;
    const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15);
    const uint32_t y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= S.cam.w || y >= S.cam.h) return;
    const uint32_t id = y * S.cam.w + x;
    const float u = x / (float)S.cam.w, v = y / (float)S.cam.h;
    float4 c = make_float4(0.25f, 0.25f, 0.25f, 1.0f);
    if (S.skydome_type != NOOR_SKY_CONSTANT) {
        c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (tex) {
            const f3 rgb = tex_bilinear(tex, tw, th, u, v);
            c = make_float4(rgb.x, rgb.y, rgb.z, tex_bilinear_w(tex, tw, th, u, v));
        }
    }
    film[id] = c;
    if (S.mis) {
        Rng rng;
        rng_init(rng, S.rng_kind, id, frame);
        float pdf;
        f2 q;
        q.x = rng_next(rng, S.rng_kind);
        q.y = rng_next(rng, S.rng_kind);
        q = d_sample2d(S, q, pdf);
        const uint32_t px = (uint32_t)(q.x * S.cam.w), py = (uint32_t)(q.y * S.cam.h);
        if (px < S.cam.w && py < S.cam.h) film[py * S.cam.w + px] = make_float4(1.0f, 0.0f, 0.0f, 1.0f);
    }
}

hipError_t launch_debug_skydome(const Context& c, uint32_t frame) {
    hipError_t se = sync_scene(c);
    if (se != hipSuccess) return se;
    const bool phys = c.S.skydome_type == NOOR_SKY_PHYSICAL;
    const float4* tex = phys ? c.sky_bake : c.S.env;
    const int tw = phys ? 2048 : (int)c.S.env_w, th = phys ? 2048 : (int)c.S.env_h;
    hipLaunchKernelGGL(k_debug_skydome, dim3((c.S.cam.w + 15) / 16, (c.S.cam.h + 15) / 16), dim3(256), 0, c.stream,
                       c.d_scene, c.film, frame, tex, tw, th);
    return hipGetLastError();
}

}  // namespace noor
