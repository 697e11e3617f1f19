// Internal host-side declarations shared by the C ABI (capi.hip) and the
// kernel translation units.  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "dev_core.h"

namespace noor {

// Device counters, accumulated with one atomic per wave.  The ray counts,
// which every launch flushes, are spread over kStatSlots counters 256 bytes
// apart (slot: wave index mod kStatSlots): returning atomics on one address
// complete about one per 11 ns on MI355X (measured), so thousands of waves
// flushing into a single counter would serialize behind each other.
constexpr int kStatSlots = 64;
struct DevStats {
    struct Slot {
        unsigned long long extension_rays, shadow_rays, drain_extension_rays, drain_shadow_rays, pad[28];
    } slot[kStatSlots];
    unsigned long long ext[4];   // nodes, triangle tests, second Wald fetches, instance entries
    unsigned long long shd[4];
    unsigned long long overflow;
    unsigned long long ext_hits;  // closest-hit queries that hit (counted with STATS)
};

struct WaveBuffers;  // wavefront queues (kernels_wave.hip)

// Live kernel timing with HIP events on the context's stream.
struct KernelTimer {
    struct Rec {
        const char* name;
        hipEvent_t a, b;
    };
    bool enabled = false;
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<unsigned long long, double>> totals;  // launches, ms
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        hipEventCreate(&e);
        return e;
    }
    void begin(hipStream_t s, const char* name) {
        if (!enabled) return;
        Rec r{name, get(), get()};
        hipEventRecord(r.a, s);
        pending.push_back(r);
    }
    void end(hipStream_t s) {
        if (!enabled || pending.empty()) return;
        hipEventRecord(pending.back().b, s);
    }
    // call after the stream is synchronized
    void collect() {
        for (auto& r : pending) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                auto& t = totals[r.name];
                t.first += 1;
                t.second += ms;
            }
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    ~KernelTimer() {
        for (auto& r : pending) {
            hipEventDestroy(r.a);
            hipEventDestroy(r.b);
        }
        for (auto e : pool) hipEventDestroy(e);
    }
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    DevScene S{};
    noor_hosek hosek_host{};
    noor_camera cam{};
    noor_spec spec{};
    noor_options opt{};
    // device allocations
    std::vector<void*> allocs;
    float4* film = nullptr;
    uint32_t film_w = 0, film_h = 0;
    // the rounding errors of the last noor_render_sample (one float4 per
    // pixel), and which sample they belong to (its output buffer and stream)
    float4* sample_lo = nullptr;
    size_t sample_lo_px = 0;
    const float* sample_lo_of = nullptr;
    uint32_t sample_lo_stream = 0;
    float4* sky_bake = nullptr;
    float* dfunc = nullptr;
    float* dcdf = nullptr;
    noor_hosek* d_hosek = nullptr;
    DevStats* d_stats = nullptr;
    float4* d_lookat = nullptr;
    DevScene* d_scene = nullptr;  // device copy of S read by every kernel
    bool scene_dirty = true;
    WaveBuffers* wave = nullptr;
    uint32_t wave_capacity = 0;   // 0: min(pixels per layer, 2^22)
    int wave_block = 64;          // lanes per block of the persistent traversal kernels
    KernelTimer timer;
    uint32_t stack_need = 0;      // traversal stack entries the wide BVH can need
    int stack_lds = 24;           // of which at most this many per lane live in LDS
    size_t wide_nodes = 0;
    size_t texels = 0;      // texels of the material texture store, all mip levels
    uint32_t bvh_root = 0;        // the payload's binary root (the spec must name it)
    uint32_t num_materials = 0;
    uint32_t* gstack = nullptr;   // device part of the traversal stack
    uint32_t* scratch = nullptr;  // this device's no-allocation scratch (not owned)
    // multi-view batch films
    float4* batch_film = nullptr;
    std::vector<uint32_t> batch_w, batch_h;
    std::vector<size_t> batch_off;  // float4 offset of each view's film
    std::string last_error;
};

// Uploads the host copy of the scene descriptor when it changed.  Called by
// every launch wrapper before its first launch; no kernel is in flight then.
inline hipError_t sync_scene(const Context& cc) {
    Context& c = const_cast<Context&>(cc);
    if (!c.scene_dirty) return hipSuccess;
    hipError_t e = hipMemcpyAsync(c.d_scene, &c.S, sizeof(DevScene), hipMemcpyHostToDevice, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e == hipSuccess) c.scene_dirty = false;
    return e;
}

// kernels_mega.hip
// `direct`: write each pixel's last sample there instead of updating the film
// `direct_lo` (DIRECT only, may be null): the rounding error of each sample's
// radiance sum, which the adjoint uses to recover the radiance after a bounce
hipError_t launch_megakernel(const Context& c, uint32_t frame0, uint32_t spp, float4* direct = nullptr,
                             float4* direct_lo = nullptr);
// `gstack`, `glanes`: the device part of the traversal stack to use instead of
// the context's (lane-interleaved, (stack depth - LDS entries) x glanes words)
hipError_t launch_trace(const Context& c, const noor_ray* d_rays, uint32_t n, int any_hit, noor_hit* d_hits,
                        uint32_t* gstack = nullptr, uint32_t glanes = 0);
hipError_t launch_debug_skydome(const Context& c, uint32_t frame);
// kernels_env.hip
hipError_t launch_sky_bake(const Context& c, float4* out, int w, int h);
hipError_t launch_dist_build(const Context& c, const float4* tex, int tw, int th, float* func, float* cdf, int W,
                             int H);
// kernels_adjoint.hip
hipError_t launch_set_params(const Context& c, const float* albedo, const float* rough, uint32_t n, const float* env_rgb);
// `radiance_lo`: the forward's direct_lo for the same sample, or null
hipError_t launch_adjoint(const Context& c, uint32_t stream, const float* radiance, const float4* radiance_lo,
                          const float* grad, float* g_albedo, float* g_rough, uint32_t num_mats, float* g_env);
// kernels_wave.hip
hipError_t wave_render(Context& c, uint32_t frame0, uint32_t spp);
hipError_t wave_render_batch(Context& c, const noor_view* views, uint32_t n, float4* films, const size_t* offsets);
void wave_free(Context& c);

}  // namespace noor
