// Load-time kernels of the environment light: the physical-sky bake
// (update_skydome, noor/cuda/skydome.cuh:27-37) and the 2D piecewise-constant
// distribution used for importance sampling (update_func / prescan /
// process_cdf, noor/cuda/distribution2d.cuh:63-228).  The scan keeps the
// reference's Blelloch tree order so that every CDF entry rounds the same way
// as the CPU restatement.
#include "ctx.h"

namespace noor {

__global__ void k_sky_bake(const noor_hosek* __restrict__ hk, float4* __restrict__ out, int W, int H) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const float u0 = (float)x / (float)W, u1 = (float)y / (float)H;
    const float phi = u0 * K_2PI;
    const float theta = u1 * K_PI;
    const float cosT = nm_cosf(theta);
    const float sinT = nm_sinf(theta);
    f3 d = f3{sinT * nm_cosf(phi), cosT, sinT * nm_sinf(phi)};
    const float l2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const float inv = 1.0f / sqrtf(l2);
    d = f3{d.x * inv, d.y * inv, d.z * inv};
    const f3 rgb = hosek_query(*hk, d);
    out[(size_t)y * W + x] = make_float4(sinT * rgb.x, sinT * rgb.y, sinT * rgb.z, sinT * 1.0f);
}

hipError_t launch_sky_bake(const Context& c, float4* out, int w, int h) {
    hipLaunchKernelGGL(k_sky_bake, dim3((w + 15) / 16, (h + 15) / 16), dim3(256), 0, c.stream, c.d_hosek, out, w, h);
    return hipGetLastError();
}

// update_func: 3x3 box of bilinear luminance samples per cell
__global__ void k_update_func(const float4* __restrict__ tex, int tw, int th, float* __restrict__ func, int W, int H) {
    const int row = blockIdx.y;
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= H || col >= W) return;
    float f = 0.0f;
    for (int y = -1; y <= 1; ++y) {
        for (int x = -1; x <= 1; ++x) {
            const float u = (float)(x + col) / (float)W;
            const float v = (float)(y + row) / (float)H;
            const f3 c = tex_bilinear(tex, tw, th, u, v);
            f += c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f;
        }
    }
    func[(size_t)row * (W + 1) + col] = f / 9.0f;
}

// prescan: exclusive Blelloch scan of func/n per row (one block per row)
__global__ void k_prescan(float* __restrict__ func, float* __restrict__ cdf, int n, int row_stride) {
    extern __shared__ float temp[];
    float* f = func + (size_t)blockIdx.x * row_stride;
    float* c = cdf + (size_t)blockIdx.x * row_stride;
    for (int i = threadIdx.x; i < n; i += blockDim.x) temp[i] = f[i] / (float)n;
    int offset = 1;
    for (int d = n >> 1; d > 0; d >>= 1) {
        __syncthreads();
        for (int tid = threadIdx.x; tid < d; tid += blockDim.x) {
            const int ai = offset * (2 * tid + 1) - 1;
            const int bi = offset * (2 * tid + 2) - 1;
            temp[bi] += temp[ai];
        }
        offset *= 2;
    }
    __syncthreads();
    if (threadIdx.x == 0) temp[n - 1] = 0.0f;
    for (int d = 1; d < n; d *= 2) {
        offset >>= 1;
        __syncthreads();
        for (int tid = threadIdx.x; tid < d; tid += blockDim.x) {
            const int ai = offset * (2 * tid + 1) - 1;
            const int bi = offset * (2 * tid + 2) - 1;
            const float t = temp[ai];
            temp[ai] = temp[bi];
            temp[bi] += t;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) c[i] = temp[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        const float funcInt = c[n - 1] + f[n - 1] / (float)n;
        c[n] = funcInt;
        f[n] = funcInt;
    }
}

// process_cdf: normalise each row, copy row integrals into the marginal
__global__ void k_process_cdf(float* __restrict__ func, float* __restrict__ cdf, int width, int height) {
    const int row = blockIdx.x;
    __shared__ float funcInt;
    if (threadIdx.x == 0) funcInt = cdf[(size_t)row * (width + 1) + width];
    __syncthreads();
    const float fi = funcInt;
    for (int col = threadIdx.x; col < width; col += blockDim.x) {
        float& e = cdf[(size_t)row * (width + 1) + col + 1];
        if (fi == 0.0f) e = (float)(col + 1) / (float)width;
        else e = e / fi;
    }
    if (threadIdx.x == 0 && height > 1) func[(size_t)height * (width + 1) + row] = fi;
}

hipError_t launch_dist_build(const Context& c, const float4* tex, int tw, int th, float* func, float* cdf, int W,
                             int H) {
    hipLaunchKernelGGL(k_update_func, dim3((W + 255) / 256, H), dim3(256), 0, c.stream, tex, tw, th, func, W, H);
    hipLaunchKernelGGL(k_prescan, dim3(H), dim3(256), (size_t)W * sizeof(float), c.stream, func, cdf, W, W + 1);
    hipLaunchKernelGGL(k_process_cdf, dim3(H), dim3(256), 0, c.stream, func, cdf, W, H);
    float* mf = func + (size_t)H * (W + 1);
    float* mc = cdf + (size_t)H * (W + 1);
    hipLaunchKernelGGL(k_prescan, dim3(1), dim3(256), (size_t)H * sizeof(float), c.stream, mf, mc, H, H + 1);
    hipLaunchKernelGGL(k_process_cdf, dim3(1), dim3(256), 0, c.stream, mf, mc, H, 1);
    return hipGetLastError();
}

}  // namespace noor
