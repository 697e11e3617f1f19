// Internal host-side declarations shared by the C ABI (capi.hip) and the
// kernel translation units.  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "dev_core.h"

namespace noor {

// Device counters, accumulated with one atomic per wave.
struct DevStats {
    unsigned long long extension_rays, shadow_rays, node_visits, tri_tests, tri_second_fetch, instance_entries;
    unsigned long long overflow;
};

struct WaveBuffers;  // wavefront queues (kernels_wave.hip)

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
    float4* sky_bake = nullptr;
    float* dfunc = nullptr;
    float* dcdf = nullptr;
    noor_hosek* d_hosek = nullptr;
    DevStats* d_stats = nullptr;
    float4* d_lookat = nullptr;
    WaveBuffers* wave = nullptr;
    uint32_t stack_need = 0;
    std::string last_error;
};

// kernels_mega.hip
hipError_t launch_megakernel(const Context& c, uint32_t frame0, uint32_t spp);
hipError_t launch_trace(const Context& c, const noor_ray* d_rays, uint32_t n, int any_hit, noor_hit* d_hits);
hipError_t launch_debug_skydome(const Context& c, uint32_t frame);
// kernels_env.hip
hipError_t launch_sky_bake(const Context& c, float4* out, int w, int h);
hipError_t launch_dist_build(const Context& c, const float4* tex, int tw, int th, float* func, float* cdf, int W,
                             int H);
// kernels_wave.hip
hipError_t wave_render(Context& c, uint32_t frame0, uint32_t spp);
void wave_free(Context& c);

}  // namespace noor
