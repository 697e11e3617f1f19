// Wavefront renderer with material-sorted queues, persistent traversal
// kernels and a hipGraph-captured bounce loop.
//
// One graph replay renders one chunk of one sample layer:
//   k_generate  camera rays for the chunk's items (view, pixel)      -> ray queue
//   per bounce:
//     k_extend  closest hit (persistent waves pull 64-ray batches);
//               each hit is binned by material class with one atomic
//               per class per wave                                    -> class queues
//     k_shade   walks the class queues in class order, so each wave
//               shades one material; runs shade_bounce(), the same
//               per-path arithmetic as the megakernel                 -> shadow queue, next ray queue
//     k_shadow  any-hit for the next-event rays (persistent)         -> L
//   k_finish    running mean into the film
//   k_advance   bumps the device-side chunk index
// Queue counters live in device memory and are cleared by the kernels
// themselves (no memset nodes in the graph).  Because every path runs the
// same functions on the same RNG stream as in the megakernel, the films of
// the two renderers are identical.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.h"

namespace noor {

constexpr int kClasses = 6;  // miss, emitter, lambert, conductor, dielectric, other

struct ViewDesc {
    noor_camera cam;
    uint32_t film_offset;  // float4 offset of this view's film
    uint32_t frame;        // running-mean frame number of this layer
    uint32_t seed;         // per-view RNG stream offset
    uint32_t item_begin;   // first item of this view in the layer
};

struct WaveCtl {
    uint32_t ray_count[2];
    uint32_t class_count[kClasses];
    uint32_t shadow_count;
    uint32_t work_extend, work_shadow;
    uint32_t chunk;          // index of the chunk being rendered
    uint32_t layer_items;    // items in the layer
    uint32_t num_views;
    uint32_t capacity;
    uint32_t pad[3];
};

struct WaveBuffers {
    uint32_t capacity = 0;
    int block = 64;
    int stack_lds = 0;
    // path state (slot = item index within the chunk)
    uint32_t* pix = nullptr;
    uint32_t* view = nullptr;
    float4* L = nullptr;       // L.xyz
    float4* beta = nullptr;    // beta.xyz, w = bounce | specular << 16
    uint64_t* rng = nullptr;
    uint64_t* rng_inc = nullptr;
    float4* ro = nullptr;
    float4* rd = nullptr;
    float4* hit = nullptr;     // t, u, v, prim
    uint2* hit2 = nullptr;     // inst, kind
    float4* nee_a = nullptr;   // candidate.xyz
    float4* nee_beta = nullptr;
    float4* sh_o = nullptr;    // shadow ray origin, tmax
    float4* sh_d = nullptr;
    uint32_t* q_ray[2] = {nullptr, nullptr};
    uint32_t* q_class = nullptr;  // kClasses * capacity
    uint32_t* q_shadow = nullptr;
    WaveCtl* ctl = nullptr;
    ViewDesc* views = nullptr;
    uint32_t max_views = 0;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_bounces = -1;
    DevScene graph_scene{};   // scene arguments baked into the captured graph
    float4* graph_films = nullptr;
    bool graph_stats = false;
    uint32_t graph_grid_persist = 0;
    std::vector<void*> allocs;
};

// ----------------------------------------------------------------------------------------------
// helpers

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63; }

// wave-aggregated append: returns this lane's slot in `queue` (only for lanes with pred)
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    const unsigned long long mask = __ballot(pred);
    const uint32_t lane = lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (mask != 0 && (int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader < 0 ? 0 : leader, 64);
    const unsigned long long below = mask & ((1ull << lane) - 1ull);
    return base + (uint32_t)__popcll(below);
}

__device__ __forceinline__ void flush4(unsigned long long* dst, const Counters& c) {
    unsigned long long v[4] = {c.nodes, c.tris, c.tri2, c.inst};
    for (int i = 0; i < 4; ++i)
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off, 64);
    if (lane_id() == 0)
        for (int i = 0; i < 4; ++i)
            if (v[i]) atomicAdd(&dst[i], v[i]);
}

__device__ __forceinline__ int material_class(const DevScene& S, bool hit, const Hit& h) {
    if (!hit) return 0;
    if (h.kind == 2) return 1;
    const uint32_t t = S.wald2[h.prim].w;
    if (t & M_EMITTER_ANY) return 1;
    if (t == NOOR_MAT_DIFFUSE) return 2;
    if (t == NOOR_MAT_METAL) return 3;
    if (t == NOOR_MAT_GLASS) return 4;
    return 5;
}

__device__ __forceinline__ void load_state(const WaveBuffers& B, uint32_t p, PathState& ps) {
    const float4 L = B.L[p], bt = B.beta[p];
    ps.L = f3{L.x, L.y, L.z};
    ps.beta = f3{bt.x, bt.y, bt.z};
    const uint32_t packed = __float_as_uint(bt.w);
    ps.bounce = packed & 0xFFFFu;
    ps.specular = packed >> 16;
    ps.rng.state = B.rng[p];
    ps.rng.inc = B.rng_inc[p];
}
__device__ __forceinline__ void store_state(const WaveBuffers& B, uint32_t p, const PathState& ps) {
    B.L[p] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
    B.beta[p] = make_float4(ps.beta.x, ps.beta.y, ps.beta.z, __uint_as_float(ps.bounce | (ps.specular << 16)));
    B.rng[p] = ps.rng.state;
}

// ----------------------------------------------------------------------------------------------
// kernels

__global__ void k_generate(const DevScene* __restrict__ Sp, WaveBuffers B) {
    const DevScene& S = *Sp;
    WaveCtl* ctl = B.ctl;
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t chunk = ctl->chunk;
    const uint32_t begin = chunk * B.capacity;
    const uint32_t items = ctl->layer_items;
    const uint32_t n = (items > begin) ? min(B.capacity, items - begin) : 0u;
    if (slot == 0) {
        ctl->ray_count[0] = n;
        ctl->ray_count[1] = 0;
        for (int c = 0; c < kClasses; ++c) ctl->class_count[c] = 0;
        ctl->shadow_count = 0;
        ctl->work_extend = 0;
        ctl->work_shadow = 0;
    }
    if (slot >= n) return;
    const uint32_t item = begin + slot;
    // locate the view of this item (views sorted by item_begin)
    uint32_t lo = 0, hi = ctl->num_views - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (B.views[mid].item_begin <= item) lo = mid;
        else hi = mid - 1;
    }
    const ViewDesc& V = B.views[lo];
    const uint32_t gid = item - V.item_begin;
    const uint32_t x = gid % V.cam.w, y = gid / V.cam.w;
    PathState ps;
    ps.L = f3{0.0f, 0.0f, 0.0f};
    ps.beta = f3{1.0f, 1.0f, 1.0f};
    ps.bounce = 0;
    ps.specular = 0;
    rng_init(ps.rng, S.rng_kind, gid, V.frame + V.seed);
    const Ray r = gen_ray(V.cam, x, y, ps.rng, S.rng_kind);
    B.pix[slot] = gid;
    B.view[slot] = lo;
    B.rng_inc[slot] = ps.rng.inc;
    store_state(B, slot, ps);
    B.ro[slot] = make_float4(r.o.x, r.o.y, r.o.z, 0.0f);
    B.rd[slot] = make_float4(r.d.x, r.d.y, r.d.z, 0.0f);
    B.q_ray[0][slot] = slot;
}

template <bool STATS>
__global__ void k_extend(const DevScene* __restrict__ Sp, WaveBuffers B, int cur, DevStats* stats) {
    const DevScene& S = *Sp;
    extern __shared__ uint32_t smem[];
    WaveCtl* ctl = B.ctl;
    const TravStack st = make_stack(S, smem);
    const uint32_t count = ctl->ray_count[cur];
    const uint32_t* q = B.q_ray[cur];
    Counters cnt{0, 0, 0, 0};
    bool overflow = false;
    unsigned long long ext = 0;
    for (;;) {
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(&ctl->work_extend, 64u);
        base = __shfl(base, 0, 64);
        if (base >= count) break;
        const uint32_t i = base + lane_id();
        const bool active = i < count;
        uint32_t p = 0;
        int cls = -1;
        if (active) {
            p = q[i];
            const float4 o = B.ro[p], d = B.rd[p];
            Ray ray = mkray(f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, K_INF);
            Hit h;
            ++ext;
            const bool hit = bvh_closest<STATS>(S, ray, h, st, cnt, overflow);
            B.hit[p] = make_float4(h.t, h.u, h.v, __uint_as_float(h.prim));
            B.hit2[p] = make_uint2(h.inst, hit ? h.kind : 0u);
            cls = material_class(S, hit, h);
        }
        for (int c = 0; c < kClasses; ++c) {
            const bool mine = (cls == c);
            const uint32_t slot = wave_append(&ctl->class_count[c], mine);
            if (mine) B.q_class[(size_t)c * B.capacity + slot] = p;
        }
    }
    // the shadow work counter of the previous bounce is no longer in use
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->shadow_count = 0;
    const unsigned long long e = [&] {
        unsigned long long v = ext;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        return v;
    }();
    if (lane_id() == 0 && e) atomicAdd(&stats->extension_rays, e);
    if (STATS) flush4(stats->ext, cnt);
    const unsigned long long ov = __ballot(overflow);
    if (lane_id() == 0 && ov) atomicAdd(&stats->overflow, 1ull);
}

__global__ void k_shade(const DevScene* __restrict__ Sp, WaveBuffers B, int cur) {
    const DevScene& S = *Sp;
    WaveCtl* ctl = B.ctl;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    // class-ordered item index -> path slot
    uint32_t prefix = 0;
    int cls = -1;
    uint32_t local = 0;
    for (int c = 0; c < kClasses; ++c) {
        const uint32_t n = ctl->class_count[c];
        if (cls < 0 && t < prefix + n) {
            cls = c;
            local = t - prefix;
        }
        prefix += n;
    }
    if (t == 0) {
        ctl->work_extend = 0;
        ctl->ray_count[cur] = 0;
    }
    const int nxt = cur ^ 1;
    bool cont = false, shadow = false;
    uint32_t p = 0;
    if (cls >= 0) {
        p = B.q_class[(size_t)cls * B.capacity + local];
        PathState ps;
        load_state(B, p, ps);
        const float4 o = B.ro[p], d = B.rd[p];
        Ray ray = mkray(f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, K_INF);
        const float4 hh = B.hit[p];
        const uint2 h2 = B.hit2[p];
        Hit h;
        h.t = hh.x;
        h.u = hh.y;
        h.v = hh.z;
        h.prim = __float_as_uint(hh.w);
        h.inst = h2.x;
        h.kind = h2.y;
        h.key = 0;
        const bool hit = h2.y != 0;
        if (hit) ray.tmax = h.t;
        Nee nee;
        cont = shade_bounce(S, ps, ray, hit, h, nee, nullptr);
        if (nee.has && nee.na) {
            shadow = true;
            B.nee_a[p] = make_float4(nee.a.x, nee.a.y, nee.a.z, 0.0f);
            B.nee_beta[p] = make_float4(nee.beta.x, nee.beta.y, nee.beta.z, 0.0f);
            B.sh_o[p] = make_float4(nee.ra.o.x, nee.ra.o.y, nee.ra.o.z, nee.ra.tmax);
            B.sh_d[p] = make_float4(nee.ra.d.x, nee.ra.d.y, nee.ra.d.z, 0.0f);
        } else {
            resolve_nee(ps, nee, false, false);
        }
        if (cont) {
            ps.bounce++;
            B.ro[p] = make_float4(ps.o.x, ps.o.y, ps.o.z, 0.0f);
            B.rd[p] = make_float4(ps.d.x, ps.d.y, ps.d.z, 0.0f);
        }
        store_state(B, p, ps);
    }
    const uint32_t s1 = wave_append(&ctl->shadow_count, shadow);
    if (shadow) B.q_shadow[s1] = p;
    const uint32_t s2 = wave_append(&ctl->ray_count[nxt], cont);
    if (cont) B.q_ray[nxt][s2] = p;
}

template <bool STATS>
__global__ void k_shadow(const DevScene* __restrict__ Sp, WaveBuffers B, DevStats* stats) {
    const DevScene& S = *Sp;
    extern __shared__ uint32_t smem[];
    WaveCtl* ctl = B.ctl;
    const TravStack st = make_stack(S, smem);
    const uint32_t count = ctl->shadow_count;
    Counters cnt{0, 0, 0, 0};
    bool overflow = false;
    unsigned long long shd = 0;
    for (;;) {
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(&ctl->work_shadow, 64u);
        base = __shfl(base, 0, 64);
        if (base >= count) break;
        const uint32_t i = base + lane_id();
        if (i < count) {
            const uint32_t p = B.q_shadow[i];
            const float4 o = B.sh_o[p], d = B.sh_d[p];
            const Ray ray = mkray(f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, o.w);
            ++shd;
            const bool occ = bvh_any<STATS>(S, ray, st, cnt, overflow);
            Nee nee;
            const float4 a = B.nee_a[p], bt = B.nee_beta[p];
            nee.a = f3{a.x, a.y, a.z};
            nee.beta = f3{bt.x, bt.y, bt.z};
            nee.has = 1;
            nee.na = 1;
            nee.nb = 0;
            PathState ps;
            const float4 L = B.L[p];
            ps.L = f3{L.x, L.y, L.z};
            resolve_nee(ps, nee, occ, false);
            B.L[p] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int c = 0; c < kClasses; ++c) ctl->class_count[c] = 0;
    }
    unsigned long long v = shd;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane_id() == 0 && v) atomicAdd(&stats->shadow_rays, v);
    if (STATS) flush4(stats->shd, cnt);
    const unsigned long long ov = __ballot(overflow);
    if (lane_id() == 0 && ov) atomicAdd(&stats->overflow, 1ull);
}

// after the shadow kernel: reset the shadow work counter for the next bounce
__global__ void k_between(WaveBuffers B) {
    if (threadIdx.x == 0) B.ctl->work_shadow = 0;
}

__global__ void k_finish(const DevScene* __restrict__ Sp, WaveBuffers B, float4* films) {
    const DevScene& S = *Sp;
    WaveCtl* ctl = B.ctl;
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t begin = ctl->chunk * B.capacity;
    const uint32_t items = ctl->layer_items;
    const uint32_t n = (items > begin) ? min(B.capacity, items - begin) : 0u;
    if (slot >= n) return;
    const ViewDesc& V = B.views[B.view[slot]];
    const float4 L = B.L[slot];
    float4* f = films + V.film_offset + B.pix[slot];
    *f = film_lerp(*f, final_radiance(S, f3{L.x, L.y, L.z}), V.frame);
}

__global__ void k_advance(WaveBuffers B) {
    if (threadIdx.x == 0) B.ctl->chunk += 1;
}

// ----------------------------------------------------------------------------------------------
// host side

static hipError_t wave_alloc(Context& c, WaveBuffers& B, uint32_t cap, uint32_t max_views) {
    auto alloc = [&](void** p, size_t bytes) -> hipError_t {
        hipError_t e = hipMalloc(p, bytes);
        if (e == hipSuccess) B.allocs.push_back(*p);
        return e;
    };
    hipError_t e = hipSuccess;
#define WA(ptr, n) \
    if (e == hipSuccess) e = alloc((void**)&(ptr), sizeof(*(ptr)) * (size_t)(n));
    WA(B.pix, cap);
    WA(B.view, cap);
    WA(B.L, cap);
    WA(B.beta, cap);
    WA(B.rng, cap);
    WA(B.rng_inc, cap);
    WA(B.ro, cap);
    WA(B.rd, cap);
    WA(B.hit, cap);
    WA(B.hit2, cap);
    WA(B.nee_a, cap);
    WA(B.nee_beta, cap);
    WA(B.sh_o, cap);
    WA(B.sh_d, cap);
    WA(B.q_ray[0], cap);
    WA(B.q_ray[1], cap);
    WA(B.q_class, (size_t)kClasses * cap);
    WA(B.q_shadow, cap);
    WA(B.ctl, 1);
    WA(B.views, max_views);
#undef WA
    B.capacity = cap;
    B.max_views = max_views;
    (void)c;
    return e;
}

void wave_free(Context& c) {
    if (!c.wave) return;
    WaveBuffers& B = *c.wave;
    if (B.exec) hipGraphExecDestroy(B.exec);
    if (B.graph) hipGraphDestroy(B.graph);
    for (void* p : B.allocs) hipFree(p);
    delete c.wave;
    c.wave = nullptr;
}

// Persistent traversal grid: as many lanes as the global stack part covers
// (32 waves per CU).
static uint32_t persistent_lanes(const Context& c) { return c.S.gstack_lanes; }

// Records one chunk of one layer (generate, bounces, finish, advance) on c.stream.
static hipError_t record_chunk(Context& c, WaveBuffers& B, float4* films) {
    const DevScene& S = c.S;
    const uint32_t cap = B.capacity;
    const dim3 g1((cap + 255) / 256), b1(256);
    const int blk = B.block;
    const size_t lds = (size_t)blk * S.stack_lds * sizeof(uint32_t);
    const dim3 gp(B.graph_grid_persist / (blk / 64)), bp(blk);
    KernelTimer& T = c.timer;
    T.begin(c.stream, "wave_generate");
    hipLaunchKernelGGL(k_generate, g1, b1, 0, c.stream, c.d_scene, B);
    T.end(c.stream);
    for (int b = 0; b < S.bounces; ++b) {
        const int cur = b & 1;
        T.begin(c.stream, "wave_extend");
        if (c.opt.stats)
            hipLaunchKernelGGL(k_extend<true>, gp, bp, lds, c.stream, c.d_scene, B, cur, c.d_stats);
        else
            hipLaunchKernelGGL(k_extend<false>, gp, bp, lds, c.stream, c.d_scene, B, cur, c.d_stats);
        T.end(c.stream);
        T.begin(c.stream, "wave_shade");
        hipLaunchKernelGGL(k_shade, g1, b1, 0, c.stream, c.d_scene, B, cur);
        T.end(c.stream);
        T.begin(c.stream, "wave_shadow");
        if (c.opt.stats)
            hipLaunchKernelGGL(k_shadow<true>, gp, bp, lds, c.stream, c.d_scene, B, c.d_stats);
        else
            hipLaunchKernelGGL(k_shadow<false>, gp, bp, lds, c.stream, c.d_scene, B, c.d_stats);
        T.end(c.stream);
        hipLaunchKernelGGL(k_between, dim3(1), dim3(64), 0, c.stream, B);
    }
    T.begin(c.stream, "wave_finish");
    hipLaunchKernelGGL(k_finish, g1, b1, 0, c.stream, c.d_scene, B, films);
    T.end(c.stream);
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, c.stream, B);
    return hipGetLastError();
}

// Renders the given layers.  `layers[l]` lists the views active in layer l
// (with their frame numbers); items of a layer are processed in chunks of
// the queue capacity.  Uses a captured graph per chunk when use_graph.
hipError_t wave_render_layers(Context& c, const std::vector<std::vector<ViewDesc>>& layers, float4* films,
                              uint32_t capacity, int block, bool use_graph) {
    hipError_t se = sync_scene(c);
    if (se != hipSuccess) return se;
    uint32_t max_items = 0, max_views = 1;
    for (const auto& L : layers) {
        uint32_t n = 0;
        for (const auto& v : L) n += v.cam.w * v.cam.h;
        max_items = std::max(max_items, n);
        max_views = std::max<uint32_t>(max_views, (uint32_t)L.size());
    }
    if (capacity == 0) capacity = std::min<uint32_t>(max_items, 1u << 22);
    capacity = std::max<uint32_t>(64u, (capacity + 63) & ~63u);
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    const size_t lds = (size_t)block * c.S.stack_lds * sizeof(uint32_t);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (!c.wave || c.wave->capacity != capacity || c.wave->max_views < max_views || c.wave->block != block ||
        c.wave->stack_lds != c.S.stack_lds) {
        wave_free(c);
        c.wave = new WaveBuffers();
        c.wave->block = block;
        c.wave->stack_lds = c.S.stack_lds;
        hipError_t e = wave_alloc(c, *c.wave, capacity, std::max<uint32_t>(max_views, 64));
        if (e != hipSuccess) {
            wave_free(c);
            return e;
        }
        c.wave->graph_grid_persist = persistent_lanes(c) / 64u;
    }
    WaveBuffers& B = *c.wave;
    if (c.timer.enabled) use_graph = false;  // per-kernel events need individual launches
    if (use_graph && B.exec &&
        (B.graph_bounces != c.S.bounces || B.graph_films != films || B.graph_stats != (c.opt.stats != 0) ||
         std::memcmp(&B.graph_scene, &c.S, sizeof(DevScene)) != 0)) {
        if (B.exec) hipGraphExecDestroy(B.exec);
        if (B.graph) hipGraphDestroy(B.graph);
        B.exec = nullptr;
        B.graph = nullptr;
    }
    for (const auto& L : layers) {
        if (L.empty()) continue;
        std::vector<ViewDesc> vd = L;
        uint32_t items = 0;
        for (auto& v : vd) {
            v.item_begin = items;
            items += v.cam.w * v.cam.h;
        }
        WaveCtl ctl{};
        ctl.chunk = 0;
        ctl.layer_items = items;
        ctl.num_views = (uint32_t)vd.size();
        ctl.capacity = B.capacity;
        hipError_t e = hipMemcpyAsync(B.views, vd.data(), sizeof(ViewDesc) * vd.size(), hipMemcpyHostToDevice, c.stream);
        if (e == hipSuccess) e = hipMemcpyAsync(B.ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice, c.stream);
        if (e != hipSuccess) return e;
        const uint32_t chunks = (items + B.capacity - 1) / B.capacity;
        for (uint32_t ch = 0; ch < chunks; ++ch) {
            if (use_graph) {
                if (!B.exec) {
                    e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
                    if (e != hipSuccess) return e;
                    hipError_t er = record_chunk(c, B, films);
                    e = hipStreamEndCapture(c.stream, &B.graph);
                    if (er != hipSuccess) return er;
                    if (e != hipSuccess) return e;
                    e = hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0);
                    if (e != hipSuccess) return e;
                    B.graph_bounces = c.S.bounces;
                    B.graph_scene = c.S;
                    B.graph_films = films;
                    B.graph_stats = c.opt.stats != 0;
                }
                e = hipGraphLaunch(B.exec, c.stream);
            } else {
                e = record_chunk(c, B, films);
            }
            if (e != hipSuccess) return e;
        }
    }
    return hipGetLastError();
}

hipError_t wave_render(Context& c, uint32_t frame0, uint32_t spp) {
    std::vector<std::vector<ViewDesc>> layers(spp);
    for (uint32_t s = 0; s < spp; ++s) {
        ViewDesc v{};
        v.cam = c.S.cam;
        v.film_offset = 0;
        v.frame = frame0 + s;
        v.seed = 0;
        layers[s].push_back(v);
    }
    const bool graph = getenv("NOOR_NO_GRAPH") == nullptr;
    return wave_render_layers(c, layers, c.film, c.wave_capacity, c.wave_block, graph);
}

}  // namespace noor
