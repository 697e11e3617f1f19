// Adjoint of one render sample per pixel with respect to per-material
// diffuse colour (albedo) and roughness and to the environment texels, by
// path replay (Vicini, Speierer & Jakob, "Path Replay Backpropagation",
// ACM TOG 40(4), 2021).  Nothing is stored between the passes: each path is
// traced again from its RNG stream with the megakernel's per-bounce function
// (shade_bounce), which makes the same decisions as the forward pass, and a
// tape of one bounce gives the quantities the derivatives need.
//
// The estimator of a pixel is L = sum_k beta_k (E_k + D_k): E_k the sky seen
// by an escaping ray, D_k the NEE term, beta_{k+1} = beta_k f_k |cos| / pdf_k.
// Sampling is detached: directions, lobe choice, RR and light selection are
// held fixed, so d/dtheta acts on f and the emitted radiance only.  With
// L_rest(k) = L - (radiance gathered up to and including D_k):
//   d L / d theta  +=  beta_k (d D_k / d theta)  +  L_rest(k) (d f_k / d theta) / f_k
// per colour channel, and the sky terms scatter beta into the four texels of
// their bilinear lookup.  The result is dLoss/dtheta = sum over pixels of
// dLoss/dL . dL/dtheta, accumulated with float atomics (material terms are
// first summed per wave in LDS).
#include <algorithm>

#include "ctx.h"

namespace noor {

struct AdjArgs {
    const float4* radiance;  // per pixel: radiance of the forward sample (xyz)
    const float4* radiance_lo;  // per pixel: rounding error of that sum (xyz), or null
    const float4* grad;      // per pixel: dLoss / dradiance (xyz)
    float* g_albedo;         // num_mats * 3, added to
    float* g_rough;          // num_mats, added to
    float* g_env;            // env_w * env_h * 3, added to (may be null)
    uint32_t num_mats;
    uint32_t stream;         // RNG stream (frame number) of the forward sample
};

constexpr uint32_t kLdsMats = 64;  // materials whose terms are summed in LDS first

__device__ __forceinline__ void add_env(const DevScene& S, float* g, f3 dir, f3 wgt) {
    if (!g || is_black(wgt)) return;
    const EnvTaps t = env_taps(S, dir);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (t.w[k] == 0.0f) continue;
        float* p = g + (size_t)t.idx[k] * 3;
        unsafeAtomicAdd(p + 0, wgt.x * t.w[k]);
        unsafeAtomicAdd(p + 1, wgt.y * t.w[k]);
        unsafeAtomicAdd(p + 2, wgt.z * t.w[k]);
    }
}

__device__ __forceinline__ void add_mat(float* lmat, const AdjArgs& A, uint32_t mat, f3 dalb, float drough) {
    if (mat >= A.num_mats) return;
    if (mat < kLdsMats) {
        float* p = lmat + mat * 4;
        atomicAdd(p + 0, dalb.x);
        atomicAdd(p + 1, dalb.y);
        atomicAdd(p + 2, dalb.z);
        atomicAdd(p + 3, drough);
        return;
    }
    unsafeAtomicAdd(A.g_albedo + (size_t)mat * 3 + 0, dalb.x);
    unsafeAtomicAdd(A.g_albedo + (size_t)mat * 3 + 1, dalb.y);
    unsafeAtomicAdd(A.g_albedo + (size_t)mat * 3 + 2, dalb.z);
    unsafeAtomicAdd(A.g_rough + mat, drough);
}

__device__ __forceinline__ bool is_env_light(const DevScene& S, int idx) {
    int l;
    return idx >= 0 && light_kind(S, idx, l) == NOOR_LIGHT_INFINITE && S.sky_enabled &&
           S.skydome_type == NOOR_SKY_HDR;
}

__global__ __launch_bounds__(64) void k_adjoint(const DevScene* __restrict__ Sp, AdjArgs A) {
    const DevScene& S = *Sp;
    extern __shared__ uint32_t smem[];
    __shared__ float lmat[kLdsMats * 4];
    for (uint32_t i = threadIdx.x; i < kLdsMats * 4; i += blockDim.x) lmat[i] = 0.0f;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t W = S.cam.w, H = S.cam.h;
    const uint32_t tiles_x = (W + 7) / 8, tiles = tiles_x * ((H + 7) / 8);
    const TravStack st = make_stack(S, smem);
    Counters cnt{0, 0, 0, 0};
    bool overflow = false;
    float* genv = (S.sky_enabled && S.skydome_type == NOOR_SKY_HDR) ? A.g_env : nullptr;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t x = (tile % tiles_x) * 8 + (lane & 7);
        const uint32_t y = (tile / tiles_x) * 8 + (lane >> 3);
        if (x >= W || y >= H) continue;
        const uint32_t gid = y * W + x;
        const float4 g4 = A.grad[gid];
        const f3 dl = f3{g4.x, g4.y, g4.z};
        if (is_black(dl)) continue;  // no gradient flows through this pixel
        const float4 r4 = A.radiance[gid];
        const f3 Ltot = f3{r4.x, r4.y, r4.z};
        const float4 l4 = A.radiance_lo ? A.radiance_lo[gid] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const f3 Ltot_lo = f3{l4.x, l4.y, l4.z};
        PathState ps;
        ps.L = f3{0.0f, 0.0f, 0.0f};
        ps.Lc = f3{0.0f, 0.0f, 0.0f};
        ps.beta = f3{1.0f, 1.0f, 1.0f};
        ps.bounce = 0;
        ps.specular = 0;
        rng_init(ps.rng, S.rng_kind, gid, A.stream);
        Ray ray = gen_ray(S.cam, x, y, ps.rng, S.rng_kind);
        for (int b = 0; b < S.bounces; ++b) {
            Hit h;
            const bool hit = bvh_closest<false>(S, ray, h, st, cnt, overflow);
            Nee nee;
            Tape tp;
            const bool cont = shade_bounce<false, true>(S, ps, ray, hit, h, nee, nullptr, &tp);
            if (tp.esc) add_env(S, genv, tp.esc_dir, mul3(dl, tp.esc_beta));
            bool oa = false;
            if (nee.na) oa = bvh_any<false>(S, nee.ra, st, cnt, overflow);
            resolve_nee<true>(ps, nee, oa, false);
            if (tp.nee || tp.scat) {
                const Isect& I = tp.I;
                const DevMat& m = S.mats[I.mat];
                Bsdf bs;
                make_bsdf(I.type, bs);
                if (tp.nee && !oa) {
                    // D = beta Li f |cos| / pdf
                    const float k = absdot3(tp.nee_wi, I.sn) / tp.nee_pdf;
                    if (I.mat < A.num_mats) {
                        f3 dR, dr;
                        bsdf_grad(bs, I, m, I.wo, tp.nee_wi, B_ALL & ~B_SPECULAR, -1, dR, dr);
                        const f3 w = scl3(mul3(mul3(dl, nee.beta), tp.nee_Li), k);
                        const f3 wr = mul3(w, dr);
                        add_mat(lmat, A, I.mat, mul3(w, dR), wr.x + wr.y + wr.z);
                    }
                    if (genv && is_env_light(S, tp.nee_light)) {
                        // Li is the sky times the number of lights selection chose from
                        const f3 f = bsdf_eval(bs, I, m, I.wo, tp.nee_wi, B_ALL & ~B_SPECULAR);
                        const float nl = (float)(S.sky_enabled ? S.num_lights : S.num_lights - 1);
                        add_env(S, genv, tp.nee_wi, scl3(mul3(mul3(dl, nee.beta), f), k * nl));
                    }
                }
                if (tp.scat && I.mat < A.num_mats) {
                    f3 dR, dr;
                    bsdf_grad(bs, I, m, I.wo, tp.scat_wi, B_ALL, tp.scat_lobe, dR, dr);
                    // the radiance gathered after this bounce: (total - gathered so
                    // far), each as a compensated pair (sum, rounding error)
                    const f3 rest = add3(sub3(Ltot, ps.L), sub3(Ltot_lo, ps.Lc));
                    const f3 f = tp.scat_f;
                    const f3 q = f3{f.x != 0.0f ? dl.x * rest.x / f.x : 0.0f, f.y != 0.0f ? dl.y * rest.y / f.y : 0.0f,
                                    f.z != 0.0f ? dl.z * rest.z / f.z : 0.0f};
                    const f3 qr = mul3(q, dr);
                    add_mat(lmat, A, I.mat, mul3(q, dR), qr.x + qr.y + qr.z);
                }
            }
            if (!cont) break;
            ray = mkray(ps.o, ps.d, K_INF);
            ps.bounce++;
        }
    }
    __syncthreads();
    const uint32_t n = min(A.num_mats, kLdsMats);
    for (uint32_t i = threadIdx.x; i < n * 4; i += blockDim.x) {
        const float v = lmat[i];
        if (v == 0.0f) continue;
        const uint32_t mat = i >> 2, c = i & 3;
        if (c < 3) unsafeAtomicAdd(A.g_albedo + (size_t)mat * 3 + c, v);
        else unsafeAtomicAdd(A.g_rough + mat, v);
    }
}

// Parameter updates: albedo (diffuse colour) and isotropic roughness of the
// first n materials, and the environment texels (RGB, alpha kept).
__global__ void k_set_materials(DevMat* mats, const float* albedo, const float* rough, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (albedo) {
        mats[i].diffuse[0] = albedo[3 * (size_t)i];
        mats[i].diffuse[1] = albedo[3 * (size_t)i + 1];
        mats[i].diffuse[2] = albedo[3 * (size_t)i + 2];
    }
    if (rough) mats[i].rough[0] = mats[i].rough[1] = rough[i];
}
__global__ void k_set_env(float4* env, const float* rgb, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    env[i] = make_float4(rgb[3 * (size_t)i], rgb[3 * (size_t)i + 1], rgb[3 * (size_t)i + 2], env[i].w);
}

hipError_t launch_set_params(const Context& c, const float* albedo, const float* rough, uint32_t n, const float* env_rgb) {
    if (n && (albedo || rough))
        hipLaunchKernelGGL(k_set_materials, dim3((n + 255) / 256), dim3(256), 0, c.stream,
                           const_cast<DevMat*>(c.S.mats), albedo, rough, n);
    if (env_rgb) {
        const uint32_t t = c.S.env_w * c.S.env_h;
        hipLaunchKernelGGL(k_set_env, dim3((t + 255) / 256), dim3(256), 0, c.stream, const_cast<float4*>(c.S.env),
                           env_rgb, t);
    }
    return hipGetLastError();
}

hipError_t launch_adjoint(const Context& c, uint32_t stream, const float* radiance, const float4* radiance_lo,
                          const float* grad, float* g_albedo, float* g_rough, uint32_t num_mats, float* g_env) {
    hipError_t se = sync_scene(c);
    if (se != hipSuccess) return se;
    AdjArgs A;
    A.radiance = (const float4*)radiance;
    A.radiance_lo = radiance_lo;
    A.grad = (const float4*)grad;
    A.g_albedo = g_albedo;
    A.g_rough = g_rough;
    A.g_env = g_env;
    A.num_mats = num_mats;
    A.stream = stream;
    const uint32_t tiles = ((c.S.cam.w + 7) / 8) * ((c.S.cam.h + 7) / 8);
    const uint32_t blocks = std::min<uint32_t>(tiles, c.S.gstack_lanes / 64);
    const size_t lds = (size_t)64 * c.S.stack_lds * sizeof(uint32_t);
    KernelTimer& T = const_cast<KernelTimer&>(c.timer);
    T.begin(c.stream, "adjoint");
    hipLaunchKernelGGL(k_adjoint, dim3(blocks), dim3(64), lds, c.stream, c.d_scene, A);
    T.end(c.stream);
    return hipGetLastError();
}

}  // namespace noor
