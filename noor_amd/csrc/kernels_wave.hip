// Wavefront renderer: streaming path regeneration, material-sorted shading
// queues, persistent traversal kernels and hipGraph-captured iterations.
//
// A render is a list of sample items (sample, view, pixel), processed in
// groups of whole sample layers; within a group the items are view-major,
// then pixel-major, so the consecutive items of a wave's refill are the
// samples of one pixel.  A fixed pool of path slots is kept full; each
// iteration runs
//   k_regen   retires the paths that ended in the previous iteration (their
//             radiance goes to the item's entry of the sample buffer) and
//             starts the next items in the freed slots           -> ray queue
//   k_extend  closest hit; persistent waves, and a lane that finishes its
//             ray takes the next one from the queue (half a wave at a time),
//             so lanes do not idle behind the longest ray of a batch
//   k_bin     bins the hits by material class (block-aggregated appends)
//   k_shade   walks the class queues in class order, so each wave shades one
//             material; runs shade_bounce(), the per-bounce function of the
//             megakernel: next rays -> ray queue, NEE rays -> shadow queue,
//             ended paths flagged
//   k_collect_free lists the flagged slots          -> free queue (slot order)
//   k_shadow  any hit for the NEE rays (persistent, lane refill), resolved into L
// Keeping the pool full means every launch has work for the whole GPU (a
// layer-by-layer wavefront spends most of its time in nearly empty launches
// of the late bounces).  When every item has retired, k_finish folds each
// pixel's samples into its film in sample order with the running-mean update
// (noor/cuda/framebuffer.cuh:63-75); every path runs the megakernel's
// arithmetic on the same RNG stream, so the two renderers' films are
// identical.  Queue counters are cleared by the kernels themselves (no memset
// nodes in the graph); the host replays a graph of kIters iterations until
// every item has started and at most kDrainMax paths are in flight, then
// k_drain runs each remaining path to its end in one launch (the late
// iterations carry few rays and cost their launches; the drain measured 4%
// faster on the 1080p terrain than replaying until no path is left).
#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.h"

namespace noor {

constexpr int kClasses = 6;  // miss, emitter, lambert, conductor, dielectric, other
constexpr uint32_t kNoItem = 0xFFFFFFFFu;
// iterations per graph replay (even: each replay starts at parity 0); 6
// measured 1.6% faster than 8 on the 1080p terrain: the drain starts sooner
constexpr int kIters = 6;
constexpr uint64_t kMaxGroupItems = 1ull << 28;  // sample-buffer entries per group of layers (4.3 GB)
constexpr int kRefill = 32;  // idle lanes of a wave that trigger a refill from the ray queue
constexpr uint32_t kDrainMax = 1u << 20;  // paths in flight at which k_drain finishes a group
constexpr uint32_t kMaxAutoPool = 1u << 25;  // path slots of an automatic pool, at most

struct ViewDesc {
    noor_camera cam;
    uint32_t film_offset;  // float4 offset of this view's film
    uint32_t frame0;       // running-mean frame number of this view's first sample
    uint32_t spp;          // samples of this view in this render
    uint32_t seed;         // added to the frame number to form the RNG stream
    uint32_t pix_begin;    // first pixel of this view in the finish enumeration (a multiple of its block)
    uint32_t pad[3];
};

// The items of view `view` in a group: its samples `sample` ..
// `sample + layers - 1`, pixel-major: item item_begin + px * layers + k is
// sample `sample + k` of pixel px.
struct Segment {
    uint32_t item_begin;
    uint32_t view;
    uint32_t sample;
    uint32_t layers;
};

// A queue counter on a 256-byte line of its own: every block of a launch adds
// to the counters, and returning atomics on one line complete one at a time
// (about 11 ns each on MI355X, measured).
struct alignas(256) Ctr {
    uint32_t v;
    __host__ __device__ operator uint32_t() const { return v; }
    __host__ __device__ Ctr& operator=(uint32_t x) {
        v = x;
        return *this;
    }
};

struct WaveCtl {
    Ctr ray_count[2];
    Ctr free_count[2];
    Ctr class_count[kClasses];
    Ctr shadow_count;
    Ctr work_extend, work_shadow;
    uint32_t item_next[2];  // next item to start, by iteration parity
    uint32_t new_rays[2];   // rays k_regen appended after the continuing ones
    uint32_t item_end;      // items of the current group
    uint32_t num_segments;
    uint32_t num_views;
    uint32_t pad[1];
};

// Device view of a context's path pool: device addresses and sizes only
// (trivially copyable).  Kernels take it by value: pointers that arrive as
// kernel arguments are known to be global, so their loads and stores compile
// to global instructions; read from a copy in device memory they became flat
// instructions and the wavefront step ran about 30% slower (measured).
struct WavePool {
    uint32_t capacity = 0;
    int block = 64;
    int stack_lds = 0;
    int steps = 32;            // traversal steps between refills of the persistent kernels
    uint32_t sb_items = 0;     // sample buffer entries
    uint32_t max_views = 0, max_segments = 0;
    // path state (one slot per path in flight)
    uint32_t* item = nullptr;
    // two records per slot, each pair in one 32-byte sector: L.xyz, then
    // beta.xyz with w = bounce | specular << 16; the RNG state, then its
    // increment
    float4* Lb = nullptr;
    uint64_t* rng = nullptr;
    // ray differentials of textured scenes: auxiliary origins (w of dox = has)
    // and directions
    float4* dox = nullptr;
    float4* doy = nullptr;
    float4* ddx = nullptr;
    float4* ddy = nullptr;
    bool tex = false;
    // hit record, two uint4 per slot (32 bytes, one memory line): t, u, v,
    // prim; then instance, kind | fresh << 2 | extension-queue position << 3
    // (fresh: the path's first ray, whose L and beta k_regen leaves
    // unwritten)
    uint4* hit = nullptr;
    uint32_t* cls = nullptr;   // material class of the hit (k_extend -> k_bin)
    float4* nee_a = nullptr;   // candidate.xyz
    float4* nee_beta = nullptr;
    float4* sh_o = nullptr;    // by shadow-queue position: shadow ray origin, tmax
    float4* sh_d = nullptr;    // by shadow-queue position: direction
    // MIS: the BSDF-sampled candidate and its shadow ray; w of nee_a / nee_b
    // flags which of the two candidates need their occlusion test
    float4* nee_b = nullptr;
    float4* sh2_o = nullptr;
    float4* sh2_d = nullptr;
    bool mis = false;
    uint32_t* q_ray[2] = {nullptr, nullptr};
    // the rays of the extension queues, by queue position (origin, direction):
    // a persistent kernel's refill reads them with its queue entries instead
    // of through the path slot, and k_shade's appends write them in order
    float4* qo[2] = {nullptr, nullptr};
    float4* qd[2] = {nullptr, nullptr};
    uint32_t* q_free[2] = {nullptr, nullptr};
    // per slot: its path ended in this iteration's k_shade (k_collect_free
    // lists the flagged slots on the next free queue in slot order)
    uint8_t* freed = nullptr;
    uint32_t* q_class = nullptr;  // kClasses * capacity
    uint32_t* q_shadow = nullptr;
    float4* sb = nullptr;         // final radiance per item of the current group
    WaveCtl* ctl = nullptr;
    ViewDesc* views = nullptr;
    Segment* segs = nullptr;
    int32_t* seg_of = nullptr;    // [view] -> segment of the group, -1 if the view has no sample in it
};

// Host bookkeeping of the pool: the device view, the captured graph and the
// allocations to free.
struct WaveBuffers {
    WavePool P;
    uint32_t requested = 0;  // the capacity asked for (P.capacity may be smaller after an out-of-memory fallback)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_stats = false;
    uint32_t persist_waves = 0;
    std::vector<void*> allocs;
};

// ----------------------------------------------------------------------------------------------
// helpers

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63; }

// Block-aggregated appends to two queues at once (one atomic per queue per
// block): lane results are this lane's slot in each queue where its flag is
// set.  Every thread of the block must call it.
__device__ __forceinline__ void block_append2(uint32_t* c0, bool f0, uint32_t* c1, bool f1, uint32_t& s0,
                                              uint32_t& s1) {
    __shared__ uint32_t cnt[2][16];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const unsigned long long m0 = __ballot(f0), m1 = __ballot(f1);
    if (lane == 0) {
        cnt[0][wave] = (uint32_t)__popcll(m0);
        cnt[1][wave] = (uint32_t)__popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t* ctr = threadIdx.x == 0 ? c0 : c1;
        uint32_t tot = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            const uint32_t n = cnt[threadIdx.x][w];
            cnt[threadIdx.x][w] = tot;
            tot += n;
        }
        const uint32_t base = tot ? atomicAdd(ctr, tot) : 0u;
        for (uint32_t w = 0; w < nw; ++w) cnt[threadIdx.x][w] += base;
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    s0 = cnt[0][wave] + (uint32_t)__popcll(m0 & below);
    s1 = cnt[1][wave] + (uint32_t)__popcll(m1 & below);
}

// Counter slot of this wave for the ray counts (spread over kStatSlots).
__device__ __forceinline__ uint32_t stat_slot() {
    return (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % kStatSlots;
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

__device__ __forceinline__ void load_state(const WavePool& B, uint32_t p, PathState& ps) {
    const float4 L = B.Lb[2 * (size_t)p], bt = B.Lb[2 * (size_t)p + 1];
    ps.L = f3{L.x, L.y, L.z};
    ps.beta = f3{bt.x, bt.y, bt.z};
    const uint32_t packed = __float_as_uint(bt.w);
    ps.bounce = packed & 0xFFFFu;
    ps.specular = packed >> 16;
    ps.rng.state = B.rng[2 * (size_t)p];
    ps.rng.inc = B.rng[2 * (size_t)p + 1];
}
__device__ __forceinline__ void store_state(const WavePool& B, uint32_t p, const PathState& ps) {
    B.Lb[2 * (size_t)p] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
    B.Lb[2 * (size_t)p + 1] = make_float4(ps.beta.x, ps.beta.y, ps.beta.z, __uint_as_float(ps.bounce | (ps.specular << 16)));
    B.rng[2 * (size_t)p] = ps.rng.state;
}

// ----------------------------------------------------------------------------------------------
// kernels

// Slot pool set-up for a group of `items` items.
__global__ void k_stream_init(WavePool B, uint32_t items) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    WaveCtl* ctl = B.ctl;
    if (t < B.capacity) {
        B.q_free[0][t] = t;
        B.item[t] = kNoItem;
        B.freed[t] = 0;
    }
    if (t == 0) {
        ctl->ray_count[0] = ctl->ray_count[1] = 0;
        ctl->free_count[0] = B.capacity;
        ctl->free_count[1] = 0;
        for (int c = 0; c < kClasses; ++c) ctl->class_count[c] = 0;
        ctl->shadow_count = 0;
        ctl->work_extend = ctl->work_shadow = 0;
        ctl->item_next[0] = ctl->item_next[1] = 0;
        ctl->new_rays[0] = ctl->new_rays[1] = 0;
        ctl->item_end = items;
    }
}

// Retires the paths of the free queue and starts new items in their slots.
// Items are handed out in free-queue order, so no atomics are needed: thread
// t starts item item_next + t and appends its ray after the continuing rays.
__device__ __forceinline__ void store_diff(const WavePool& B, uint32_t p, const RayDiff& r) {
    B.dox[p] = make_float4(r.ox.x, r.ox.y, r.ox.z, __int_as_float(r.has));
    B.doy[p] = make_float4(r.oy.x, r.oy.y, r.oy.z, 0.0f);
    B.ddx[p] = make_float4(r.dx.x, r.dx.y, r.dx.z, 0.0f);
    B.ddy[p] = make_float4(r.dy.x, r.dy.y, r.dy.z, 0.0f);
}
__device__ __forceinline__ void load_diff(const WavePool& B, uint32_t p, RayDiff& r) {
    const float4 a = B.dox[p], b = B.doy[p], c = B.ddx[p], d = B.ddy[p];
    r.ox = f3{a.x, a.y, a.z};
    r.has = __float_as_int(a.w);
    r.oy = f3{b.x, b.y, b.z};
    r.dx = f3{c.x, c.y, c.z};
    r.dy = f3{d.x, d.y, d.z};
}

template <bool TEX>
__global__ void k_regen(const DevScene* __restrict__ Sp, WavePool B, int cur) {
    const DevScene& S = *Sp;
    WaveCtl* ctl = B.ctl;
    const int nxt = cur ^ 1;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nfree = ctl->free_count[cur];
    const uint32_t first = ctl->item_next[cur], end = ctl->item_end;
    const uint32_t n_new = first < end ? min(nfree, end - first) : 0u;
    const uint32_t c0 = ctl->ray_count[cur];
    if (t == 0) {
        ctl->item_next[nxt] = first + n_new;
        ctl->new_rays[cur] = S.bounces > 0 ? n_new : 0u;
        // counters the previous iteration's kernels consumed; without bounces
        // no path is traced, so the slots of the items started here go straight
        // back to the next free queue (below) instead of through k_shade
        ctl->free_count[nxt] = S.bounces > 0 ? 0u : n_new;
        ctl->ray_count[nxt] = 0;
        for (int c = 0; c < kClasses; ++c) ctl->class_count[c] = 0;
        ctl->shadow_count = 0;
        ctl->work_extend = 0;
        ctl->work_shadow = 0;
    }
    if (t >= nfree) return;
    const uint32_t p = B.q_free[cur][t];
    const uint32_t it = B.item[p];
    if (it != kNoItem) {
        const float4 L = B.Lb[2 * (size_t)p];
        const f3 c = final_radiance(S, f3{L.x, L.y, L.z});
        B.sb[it] = make_float4(c.x, c.y, c.z, 0.0f);
        B.item[p] = kNoItem;
    }
    if (t >= n_new) return;
    const uint32_t item = first + t;
    if (S.bounces <= 0) {
        // no bounce to trace: the sample is the (clamped) zero radiance, and
        // the slot stays free for the next item
        const f3 c = final_radiance(S, f3{0.0f, 0.0f, 0.0f});
        B.sb[item] = make_float4(c.x, c.y, c.z, 0.0f);
        B.q_free[nxt][t] = p;
        return;
    }
    // view segment of this item (segments are sorted by item_begin; one view
    // has one), then its pixel and sample: the consecutive items of a wave's
    // refill are the samples of one pixel, whose first rays walk the same
    // nodes (11% less k_extend time than pixel-after-pixel items on the 1080p
    // terrain)
    uint32_t lo = 0, hi = ctl->num_segments - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (B.segs[mid].item_begin <= item) lo = mid;
        else hi = mid - 1;
    }
    const Segment sg = B.segs[lo];
    const ViewDesc& V = B.views[sg.view];
    const uint32_t local = item - sg.item_begin;
    const uint32_t gid = local / sg.layers, sample = sg.sample + (local - gid * sg.layers);
    const uint32_t x = gid % V.cam.w, y = gid / V.cam.w;
    PathState ps;
    ps.L = f3{0.0f, 0.0f, 0.0f};
    ps.beta = f3{1.0f, 1.0f, 1.0f};
    ps.bounce = 0;
    ps.specular = 0;
    rng_init(ps.rng, S.rng_kind, gid, V.frame0 + sample + V.seed);
    RayDiff rdf;
    const Ray r = gen_ray<TEX>(V.cam, x, y, ps.rng, S.rng_kind, &rdf);
    if (TEX) store_diff(B, p, rdf);
    B.item[p] = item;
    B.rng[2 * (size_t)p + 1] = ps.rng.inc;
    // L = 0, beta = 1, bounce 0: k_shade knows a path's first ray by the w of
    // its queued origin (carried in the hit record), so only the RNG is stored
    B.rng[2 * (size_t)p] = ps.rng.state;
    B.q_ray[cur][c0 + t] = p;
    B.qo[cur][c0 + t] = make_float4(r.o.x, r.o.y, r.o.z, 1.0f);
    B.qd[cur][c0 + t] = make_float4(r.d.x, r.d.y, r.d.z, 0.0f);
}

// Refill of a persistent wave: when at least kRefill lanes are idle (or all
// are), one atomic hands the idle lanes consecutive queue entries.  Returns
// this lane's entry, or UINT32_MAX.  `exhausted` becomes true (wave-uniform)
// once the queue has run out.
__device__ __forceinline__ uint32_t refill(uint32_t* work, uint32_t count, bool have, bool& exhausted) {
    if (exhausted) return 0xFFFFFFFFu;
    const unsigned long long idle = __ballot(!have);
    const int n = __popcll(idle);
    if (n < kRefill && n < 64) return 0xFFFFFFFFu;
    const uint32_t lane = lane_id();
    const int leader = __ffsll((long long)idle) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(work, (uint32_t)n);
    base = __shfl(base, leader, 64);
    if (base + (uint32_t)n >= count) exhausted = true;
    if (have) return 0xFFFFFFFFu;
    const uint32_t idx = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
    return idx < count ? idx : 0xFFFFFFFFu;
}

// Held to 80 VGPRs (six waves per SIMD, a few spilled registers): measured 6%
// faster than 84 VGPRs at five waves.
template <bool STATS, bool ALPHA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_extend(
    const DevScene* __restrict__ Sp, WavePool B, int cur, DevStats* stats) {
    const DevScene& S = *Sp;
    extern __shared__ uint32_t smem[];
    WaveCtl* ctl = B.ctl;
    const TravStack st = make_stack(S, smem);
    const uint32_t count = ctl->ray_count[cur] + ctl->new_rays[cur];
    const uint32_t* q = B.q_ray[cur];
    Counters cnt{0, 0, 0, 0};
    bool overflow = false, have = false, exhausted = false;
    uint32_t ext = 0, nhit = 0;  // rays this lane started, and (STATS) how many hit
    uint32_t p = 0, qi = 0;
    ClosestTrav t;
    for (;;) {
        const uint32_t idx = refill(&ctl->work_extend.v, count, have, exhausted);
        if (idx != 0xFFFFFFFFu) {
            p = q[idx];
            const float4 o = B.qo[cur][idx], d = B.qd[cur][idx];
            qi = (idx << 3) | (o.w != 0.0f ? 4u : 0u);
            closest_begin(S, mkray(f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, K_INF), t);
            have = true;
            ++ext;
        }
        if (__ballot(have) == 0ull) break;
        if (have && closest_steps<STATS, ALPHA>(S, t, st, cnt, overflow, B.steps)) {
            B.hit[2 * (size_t)p] = make_uint4(__float_as_uint(t.h.t), __float_as_uint(t.h.u), __float_as_uint(t.h.v),
                                              t.h.prim);
            B.hit[2 * (size_t)p + 1] = make_uint4(t.h.inst, (t.hit ? t.h.kind : 0u) | qi, 0u, 0u);
            B.cls[p] = (uint32_t)material_class(S, t.hit, t.h);
            if (STATS && t.hit) ++nhit;
            have = false;
        }
    }
    unsigned long long e = ext;
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);  // 64-bit wave sum
    if (lane_id() == 0 && e) atomicAdd(&stats->slot[stat_slot()].extension_rays, e);
    if (STATS) {
        flush4(stats->ext, cnt);
        unsigned long long h = nhit;
        for (int off = 32; off > 0; off >>= 1) h += __shfl_xor(h, off, 64);
        if (lane_id() == 0 && h) atomicAdd(&stats->ext_hits, h);
    }
    const unsigned long long ov = __ballot(overflow);
    if (lane_id() == 0 && ov) atomicAdd(&stats->overflow, 1ull);
}

// Material-class binning of the extension queue, in queue order: a block of
// up to 16 waves takes kBinPer consecutive stretches of blockDim.x entries,
// counts them per class in LDS and claims its slots with one atomic per class
// (returning atomics on one counter complete about one per 11 ns, and with one
// stretch per block k_bin ran at that rate).
constexpr int kBinPer = 4;
__global__ void k_bin(WavePool B, int cur) {
    WaveCtl* ctl = B.ctl;
    const uint32_t count = ctl->ray_count[cur] + ctl->new_rays[cur];
    const uint32_t first = blockIdx.x * blockDim.x * kBinPer;
    if (first >= count) return;  // whole block idle (uniform per block)
    __shared__ uint32_t cnt[kClasses][16 * kBinPer];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t p[kBinPer];
    int cls[kBinPer];
#pragma unroll
    for (int j = 0; j < kBinPer; ++j) {
        const uint32_t t = first + (uint32_t)j * blockDim.x + threadIdx.x;
        const bool valid = t < count;
        p[j] = valid ? B.q_ray[cur][t] : 0u;
        cls[j] = valid ? (int)B.cls[p[j]] : -1;
    }
#pragma unroll
    for (int j = 0; j < kBinPer; ++j)
#pragma unroll
        for (int c = 0; c < kClasses; ++c) {
            const unsigned long long m = __ballot(cls[j] == c);
            if (lane == 0) cnt[c][j * nw + wave] = (uint32_t)__popcll(m);
        }
    __syncthreads();
    if (threadIdx.x < kClasses) {
        // stretches in queue order, waves in order within a stretch
        const int c = threadIdx.x;
        const uint32_t n_parts = nw * kBinPer;
        uint32_t tot = 0;
        for (uint32_t w = 0; w < n_parts; ++w) {
            const uint32_t n = cnt[c][w];
            cnt[c][w] = tot;
            tot += n;
        }
        const uint32_t base = tot ? atomicAdd(&ctl->class_count[c].v, tot) : 0u;
        for (uint32_t w = 0; w < n_parts; ++w) cnt[c][w] += base;
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int j = 0; j < kBinPer; ++j)
#pragma unroll
        for (int c = 0; c < kClasses; ++c) {
            const unsigned long long m = __ballot(cls[j] == c);
            if (cls[j] == c)
                B.q_class[(size_t)c * B.capacity + cnt[c][j * nw + wave] + (uint32_t)__popcll(m & below)] = p[j];
        }
}

// Untextured: held to 128 VGPRs (four waves per SIMD; twelve spilled
// registers) measured 1.5% faster than 136 VGPRs at three waves.  Textured: 188 VGPRs at
// two waves; holding it to three waves spills 84 registers and gains nothing.
template <bool TEX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TEX ? 2 : 4))) void k_shade(
    const DevScene* __restrict__ Sp, WavePool B, int cur) {
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
    if (blockIdx.x * blockDim.x >= prefix) return;  // whole block idle (uniform per block)
    const int nxt = cur ^ 1;
    bool cont = false, shadow = false, ended = false;
    uint32_t p = 0;
    float4 sho, shd, nro, nrd;  // the shadow ray and the next ray, written at their queue positions
    if (cls >= 0) {
        p = B.q_class[(size_t)cls * B.capacity + local];
        const uint4 h0 = B.hit[2 * (size_t)p], h2 = B.hit[2 * (size_t)p + 1];
        const float4 hh = make_float4(__uint_as_float(h0.x), __uint_as_float(h0.y), __uint_as_float(h0.z),
                                      __uint_as_float(h0.w));
        PathState ps;
        if (h2.y & 4u) {  // the path's first ray: k_regen stored only its RNG
            ps.L = f3{0.0f, 0.0f, 0.0f};
            ps.beta = f3{1.0f, 1.0f, 1.0f};
            ps.bounce = 0;
            ps.specular = 0;
            ps.rng.state = B.rng[2 * (size_t)p];
            ps.rng.inc = B.rng[2 * (size_t)p + 1];
        } else {
            load_state(B, p, ps);
        }
        const float4 o = B.qo[cur][h2.y >> 3], d = B.qd[cur][h2.y >> 3];
        Ray ray = mkray(f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, K_INF);
        Hit h;
        h.t = hh.x;
        h.u = hh.y;
        h.v = hh.z;
        h.prim = __float_as_uint(hh.w);
        h.inst = h2.x;
        h.kind = h2.y & 3u;
        h.key = 0;
        const bool hit = h.kind != 0;
        if (hit) ray.tmax = h.t;
        Nee nee;
        RayDiff rdf;
        if (TEX) load_diff(B, p, rdf);
        cont = shade_bounce<TEX>(S, ps, ray, hit, h, nee, nullptr, nullptr, &rdf);
        // the megakernel traces at most `bounces` extension rays per path
        if (cont && ps.bounce + 1 >= (uint32_t)S.bounces) cont = false;
        ended = !cont;
        if (nee.has && (nee.na || nee.nb)) {
            shadow = true;
            B.nee_a[p] = make_float4(nee.a.x, nee.a.y, nee.a.z, __int_as_float(nee.na));
            B.nee_beta[p] = make_float4(nee.beta.x, nee.beta.y, nee.beta.z, 0.0f);
            sho = make_float4(nee.ra.o.x, nee.ra.o.y, nee.ra.o.z, nee.ra.tmax);
            shd = make_float4(nee.ra.d.x, nee.ra.d.y, nee.ra.d.z, 0.0f);
            if (nee.nb) {  // only with MIS
                B.nee_b[p] = make_float4(nee.b.x, nee.b.y, nee.b.z, 1.0f);
                B.sh2_o[p] = make_float4(nee.rb.o.x, nee.rb.o.y, nee.rb.o.z, nee.rb.tmax);
                B.sh2_d[p] = make_float4(nee.rb.d.x, nee.rb.d.y, nee.rb.d.z, 0.0f);
            } else if (S.mis) {
                B.nee_b[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        } else {
            resolve_nee(ps, nee, false, false);
        }
        if (cont) {
            ps.bounce++;
            nro = make_float4(ps.o.x, ps.o.y, ps.o.z, 0.0f);
            nrd = make_float4(ps.d.x, ps.d.y, ps.d.z, 0.0f);
            if (TEX) store_diff(B, p, rdf);
        }
        if (ended && !shadow) {
            // done, with no occlusion test pending: the sample goes to the
            // sample buffer now and the slot is retired (k_regen then finds
            // kNoItem), instead of storing the state for k_regen to read back
            const uint32_t it = B.item[p];
            const f3 c = final_radiance(S, ps.L);
            B.sb[it] = make_float4(c.x, c.y, c.z, 0.0f);
            B.item[p] = kNoItem;
        } else {
            store_state(B, p, ps);
        }
    }
    uint32_t s1, s2;
    block_append2(&ctl->shadow_count.v, shadow, &ctl->ray_count[nxt].v, cont, s1, s2);
    if (shadow) {
        B.q_shadow[s1] = p;
        B.sh_o[s1] = sho;
        B.sh_d[s1] = shd;
    }
    if (cont) {
        B.q_ray[nxt][s2] = p;
        B.qo[nxt][s2] = nro;
        B.qd[nxt][s2] = nrd;
    }
    if (ended) B.freed[p] = 1;
}

// Lists the slots k_shade flagged on the next free queue, in slot order
// within each block (one atomic per block): the items k_regen starts next
// get ascending slots, so their path state is written and read in fewer
// memory lines than in the order the paths ended (`k_shade` 9% and `k_regen`
// 5% faster on the 1080p terrain).  A thread takes four consecutive slots'
// flags in one 32-bit load; the capacity is a multiple of 256.
__global__ __launch_bounds__(1024) void k_collect_free(WavePool B, int cur) {
    WaveCtl* ctl = B.ctl;
    const int nxt = cur ^ 1;
    __shared__ uint32_t cnt[16];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t p0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4u;
    uint32_t* flags = reinterpret_cast<uint32_t*>(B.freed);
    const uint32_t fl = p0 < B.capacity ? flags[p0 >> 2] : 0u;
    const bool f0 = (fl & 0xFFu) != 0, f1 = (fl & 0xFF00u) != 0, f2 = (fl & 0xFF0000u) != 0,
               f3 = (fl & 0xFF000000u) != 0;
    const unsigned long long m0 = __ballot(f0), m1 = __ballot(f1), m2 = __ballot(f2), m3 = __ballot(f3);
    if (lane == 0)
        cnt[wave] = (uint32_t)(__popcll(m0) + __popcll(m1) + __popcll(m2) + __popcll(m3));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            const uint32_t n = cnt[w];
            cnt[w] = tot;
            tot += n;
        }
        const uint32_t base = tot ? atomicAdd(&ctl->free_count[nxt].v, tot) : 0u;
        for (uint32_t w = 0; w < nw; ++w) cnt[w] += base;
    }
    __syncthreads();
    if (fl == 0u) return;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t pos = cnt[wave] + (uint32_t)(__popcll(m0 & below) + __popcll(m1 & below) + __popcll(m2 & below) +
                                          __popcll(m3 & below));
    uint32_t* q = B.q_free[nxt];
    if (f0) q[pos++] = p0;
    if (f1) q[pos++] = p0 + 1;
    if (f2) q[pos++] = p0 + 2;
    if (f3) q[pos] = p0 + 3;
    flags[p0 >> 2] = 0u;
}

// MIS: a task may hold a second, BSDF-sampled shadow ray (separate
// instantiation, so the kernel without MIS keeps its registers)
template <bool STATS, bool ALPHA, bool MIS>
__global__ void k_shadow(const DevScene* __restrict__ Sp, WavePool B, DevStats* stats) {
    const DevScene& S = *Sp;
    extern __shared__ uint32_t smem[];
    WaveCtl* ctl = B.ctl;
    const TravStack st = make_stack(S, smem);
    const uint32_t count = ctl->shadow_count;
    Counters cnt{0, 0, 0, 0};
    bool overflow = false, have = false, exhausted = false;
    uint32_t shd = 0;  // rays this lane started
    uint32_t p = 0;
    // a task is one path's NEE: its light-sampled ray (a), then, with MIS,
    // its BSDF-sampled ray (b), traced one after the other by the same lane
    bool second = false, occ_a = false;
    AnyTrav t;
    for (;;) {
        const uint32_t idx = refill(&ctl->work_shadow.v, count, have, exhausted);
        if (idx != 0xFFFFFFFFu) {
            p = B.q_shadow[idx];
            occ_a = false;
            second = MIS && __float_as_int(B.nee_a[p].w) == 0;  // no light-sampled ray: the BSDF-sampled one only
            const float4 o = second ? B.sh2_o[p] : B.sh_o[idx], d = second ? B.sh2_d[p] : B.sh_d[idx];
            any_begin(S, mkray(f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, o.w), t);
            have = true;
            ++shd;
        }
        if (__ballot(have) == 0ull) break;
        if (have && any_steps<STATS, ALPHA>(S, t, st, cnt, overflow, B.steps)) {
            const bool nb = MIS && B.nee_b[p].w != 0.0f;
            if (!second && nb) {
                occ_a = t.occluded;
                second = true;
                const float4 o = B.sh2_o[p], d = B.sh2_d[p];
                any_begin(S, mkray(f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, o.w), t);
                ++shd;
                continue;
            }
            Nee nee;
            const float4 a = B.nee_a[p], bt = B.nee_beta[p];
            nee.a = f3{a.x, a.y, a.z};
            nee.beta = f3{bt.x, bt.y, bt.z};
            nee.has = 1;
            nee.na = MIS ? __float_as_int(a.w) : 1;
            nee.nb = nb ? 1 : 0;
            if (nb) {
                const float4 b = B.nee_b[p];
                nee.b = f3{b.x, b.y, b.z};
            }
            const bool oa = nee.na ? (second && nb ? occ_a : t.occluded) : false;
            const bool ob = nb ? t.occluded : false;
            PathState ps;
            const float4 L = B.Lb[2 * (size_t)p];
            ps.L = f3{L.x, L.y, L.z};
            resolve_nee(ps, nee, oa, ob);
            B.Lb[2 * (size_t)p] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
            have = false;
        }
    }
    unsigned long long v = shd;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane_id() == 0 && v) atomicAdd(&stats->slot[stat_slot()].shadow_rays, v);
    if (STATS) flush4(stats->shd, cnt);
    const unsigned long long ov = __ballot(overflow);
    if (lane_id() == 0 && ov) atomicAdd(&stats->overflow, 1ull);
}

// Drains the paths left once every item of the group has started and few
// are still in flight: thread t takes the t-th continuing path of the parity-0
// ray queue and runs its remaining bounces itself, with the megakernel's
// per-bounce sequence (closest hit, shade_bounce, occlusion of the NEE
// candidates, resolve_nee), then stores its sample; the slots the last
// iteration ended are retired as k_regen would.  It replaces the tail of
// nearly empty wavefront iterations, each of which costs five launches.
template <bool STATS, bool TEX>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_drain(
    const DevScene* __restrict__ Sp, WavePool B, DevStats* stats) {
    const DevScene& S = *Sp;
    extern __shared__ uint32_t smem[];
    const WaveCtl* ctl = B.ctl;
    const TravStack st = make_stack(S, smem);
    const uint32_t nray = ctl->ray_count[0], nfree = ctl->free_count[0], n = nray > nfree ? nray : nfree;
    Counters cnt{0, 0, 0, 0}, cns{0, 0, 0, 0};
    bool overflow = false;
    uint32_t ext = 0, shd = 0, nhit = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        if (t < nfree) {
            const uint32_t p = B.q_free[0][t];
            const uint32_t it = B.item[p];
            if (it != kNoItem) {
                const float4 L = B.Lb[2 * (size_t)p];
                const f3 c = final_radiance(S, f3{L.x, L.y, L.z});
                B.sb[it] = make_float4(c.x, c.y, c.z, 0.0f);
                B.item[p] = kNoItem;
            }
        }
        if (t >= nray) continue;
        const uint32_t p = B.q_ray[0][t];
        PathState ps;
        load_state(B, p, ps);
        const float4 o = B.qo[0][t], d = B.qd[0][t];
        Ray ray = mkray(f3{o.x, o.y, o.z}, f3{d.x, d.y, d.z}, K_INF);
        RayDiff rdf;
        if (TEX) load_diff(B, p, rdf);
        for (;;) {
            Hit h;
            ++ext;
            const bool hit = bvh_closest<STATS, TEX>(S, ray, h, st, cnt, overflow);
            if (STATS && hit) ++nhit;
            Nee nee;
            bool cont = shade_bounce<TEX>(S, ps, ray, hit, h, nee, nullptr, nullptr, &rdf);
            if (cont && ps.bounce + 1 >= (uint32_t)S.bounces) cont = false;
            bool oa = false, ob = false;
            if (nee.na) {
                ++shd;
                oa = bvh_any<STATS, TEX>(S, nee.ra, st, cns, overflow);
            }
            if (nee.nb) {
                ++shd;
                ob = bvh_any<STATS, TEX>(S, nee.rb, st, cns, overflow);
            }
            resolve_nee(ps, nee, oa, ob);
            if (!cont) break;
            ps.bounce++;
            ray = mkray(ps.o, ps.d, K_INF);
        }
        const uint32_t it = B.item[p];
        const f3 c = final_radiance(S, ps.L);
        B.sb[it] = make_float4(c.x, c.y, c.z, 0.0f);
        B.item[p] = kNoItem;
    }
    unsigned long long e = ext, sh = shd, hh = nhit;
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off, 64);
        sh += __shfl_xor(sh, off, 64);
    }
    if (lane_id() == 0) {
        DevStats::Slot& sl = stats->slot[stat_slot()];
        if (e) {
            atomicAdd(&sl.extension_rays, e);
            atomicAdd(&sl.drain_extension_rays, e);
        }
        if (sh) {
            atomicAdd(&sl.shadow_rays, sh);
            atomicAdd(&sl.drain_shadow_rays, sh);
        }
    }
    if (STATS) {
        flush4(stats->ext, cnt);
        flush4(stats->shd, cns);
        for (int off = 32; off > 0; off >>= 1) hh += __shfl_xor(hh, off, 64);
        if (lane_id() == 0 && hh) atomicAdd(&stats->ext_hits, hh);
    }
    const unsigned long long ov = __ballot(overflow);
    if (lane_id() == 0 && ov) atomicAdd(&stats->overflow, 1ull);
}

// Folds samples [s0, s1) of every pixel into its film, in sample order.  A
// block takes kFinishBlock pixels of one view (each view's pixels start at a
// multiple of the block in the finish enumeration), reads kFinishLayers
// samples of each of them at a time into LDS (consecutive lanes read
// consecutive items, so a wave reads whole lines) and each lane then folds
// its own pixel's from there; a lane reading its pixel's samples straight
// from memory strides by the view's layer count and ran 5x slower.
constexpr uint32_t kFinishBlock = 256;
constexpr uint32_t kFinishLayers = 8;
__global__ __launch_bounds__(kFinishBlock) void k_finish(WavePool B, float4* films, uint32_t s0, uint32_t s1) {
    __shared__ float4 tile[kFinishBlock * (kFinishLayers + 1)];  // a pixel's row padded by one entry
    const WaveCtl* ctl = B.ctl;
    const uint32_t t0 = blockIdx.x * kFinishBlock;
    uint32_t lo = 0, hi = ctl->num_views - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (B.views[mid].pix_begin <= t0) lo = mid;
        else hi = mid - 1;
    }
    const ViewDesc& V = B.views[lo];
    const uint32_t px0 = t0 - V.pix_begin, P = V.cam.w * V.cam.h;
    const uint32_t end = min(s1, V.spp);
    if (px0 >= P || s0 >= end) return;  // uniform per block
    const uint32_t G = end - s0;          // the view's layers in this group
    const uint32_t npx = min(kFinishBlock, P - px0);
    const uint32_t t = threadIdx.x;
    float4* f = films + V.film_offset + px0 + t;
    float4 acc = t < npx ? *f : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4* src = B.sb + B.segs[B.seg_of[lo]].item_begin + (size_t)px0 * G;
    for (uint32_t c = 0; c < G; c += kFinishLayers) {
        const uint32_t kn = min(kFinishLayers, G - c);
        for (uint32_t e = t; e < npx * kn; e += kFinishBlock) {
            const uint32_t q = kn == kFinishLayers ? e / kFinishLayers : e / kn;
            const uint32_t k = e - q * kn;
            tile[q * (kFinishLayers + 1) + k] = src[(size_t)q * G + c + k];
        }
        __syncthreads();
        if (t < npx)
            for (uint32_t k = 0; k < kn; ++k) {
                const float4 v = tile[t * (kFinishLayers + 1) + k];
                acc = film_lerp(acc, f3{v.x, v.y, v.z}, V.frame0 + s0 + c + k);
            }
        __syncthreads();
    }
    if (t < npx) *f = acc;
}

// ----------------------------------------------------------------------------------------------
// host side

static hipError_t wave_alloc(WaveBuffers& B, uint32_t cap, uint32_t sb_items, uint32_t max_views,
                             uint32_t max_segments, bool tex, bool mis) {
    auto alloc = [&](void** p, size_t bytes) -> hipError_t {
        hipError_t e = hipMalloc(p, bytes);
        if (e == hipSuccess) B.allocs.push_back(*p);
        return e;
    };
    WavePool& P = B.P;
    hipError_t e = hipSuccess;
#define WA(ptr, n) \
    if (e == hipSuccess) e = alloc((void**)&(ptr), sizeof(*(ptr)) * (size_t)(n));
    WA(P.item, cap);
    WA(P.Lb, 2 * (size_t)cap);
    WA(P.rng, 2 * (size_t)cap);
    if (tex) {
        WA(P.dox, cap);
        WA(P.doy, cap);
        WA(P.ddx, cap);
        WA(P.ddy, cap);
    }
    P.tex = tex;
    WA(P.hit, 2 * (size_t)cap);
    WA(P.cls, cap);
    WA(P.nee_a, cap);
    WA(P.nee_beta, cap);
    WA(P.sh_o, cap);
    WA(P.sh_d, cap);
    if (mis) {
        WA(P.nee_b, cap);
        WA(P.sh2_o, cap);
        WA(P.sh2_d, cap);
    }
    P.mis = mis;
    WA(P.q_ray[0], cap);
    WA(P.q_ray[1], cap);
    WA(P.qo[0], cap);
    WA(P.qo[1], cap);
    WA(P.qd[0], cap);
    WA(P.qd[1], cap);
    WA(P.q_free[0], cap);
    WA(P.q_free[1], cap);
    WA(P.freed, cap);
    WA(P.q_class, (size_t)kClasses * cap);
    WA(P.q_shadow, cap);
    WA(P.sb, sb_items);
    WA(P.ctl, 1);
    WA(P.views, max_views);
    WA(P.segs, max_segments);
    WA(P.seg_of, max_segments);
#undef WA
    P.capacity = cap;
    P.sb_items = sb_items;
    P.max_views = max_views;
    P.max_segments = max_segments;
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

// Records `iters` iterations (regen, extend, shade, shadow) on c.stream.
static hipError_t record_iterations(Context& c, WaveBuffers& B, int iters) {
    const uint32_t cap = B.P.capacity;
    const dim3 g1((cap + 255) / 256), b1(256);
    // k_bin takes one atomic per class per block: with returning atomics on
    // one address completing about one per 11 ns (measured), blocks of 256
    // entries left it bound by them; 4096 entries per block (1024 lanes)
    const dim3 gb((cap + 1024 * kBinPer - 1) / (1024 * kBinPer)), bb(1024);
    const int blk = B.P.block;
    const size_t lds = (size_t)blk * B.P.stack_lds * sizeof(uint32_t);
    const dim3 bp(blk);
    KernelTimer& T = c.timer;
    const bool tex = c.S.textured != 0;
    auto regen = tex ? k_regen<true> : k_regen<false>;
    auto shade = tex ? k_shade<true> : k_shade<false>;
    auto extend = c.opt.stats ? (tex ? k_extend<true, true> : k_extend<true, false>)
                              : (tex ? k_extend<false, true> : k_extend<false, false>);
    auto shadow = c.S.mis ? (c.opt.stats ? (tex ? k_shadow<true, true, true> : k_shadow<true, false, true>)
                                         : (tex ? k_shadow<false, true, true> : k_shadow<false, false, true>))
                          : (c.opt.stats ? (tex ? k_shadow<true, true, false> : k_shadow<true, false, false>)
                                         : (tex ? k_shadow<false, true, false> : k_shadow<false, false, false>));
    // Persistent grids of as many blocks as the device keeps resident at once
    // (at most the traversal stack's lanes): a block beyond that starts only
    // when a resident one ends, to find the queue empty (1% faster than a
    // fixed 32 waves per CU).
    auto resident = [&](const void* f) -> dim3 {
        const uint32_t cap_blocks = B.persist_waves / (blk / 64);
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, blk, lds) != hipSuccess || per_cu <= 0 ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess || cus <= 0) {
            (void)hipGetLastError();
            return dim3(cap_blocks);
        }
        return dim3(std::min<uint32_t>(cap_blocks, (uint32_t)per_cu * (uint32_t)cus));
    };
    const dim3 gx = resident(reinterpret_cast<const void*>(extend));
    const dim3 gs = resident(reinterpret_cast<const void*>(shadow));
    for (int i = 0; i < iters; ++i) {
        const int cur = i & 1;
        T.begin(c.stream, "wave_regen");
        hipLaunchKernelGGL(regen, g1, b1, 0, c.stream, c.d_scene, B.P, cur);
        T.end(c.stream);
        T.begin(c.stream, "wave_extend");
        hipLaunchKernelGGL(extend, gx, bp, lds, c.stream, c.d_scene, B.P, cur, c.d_stats);
        T.end(c.stream);
        T.begin(c.stream, "wave_bin");
        hipLaunchKernelGGL(k_bin, gb, bb, 0, c.stream, B.P, cur);
        T.end(c.stream);
        T.begin(c.stream, "wave_shade");
        hipLaunchKernelGGL(shade, g1, b1, 0, c.stream, c.d_scene, B.P, cur);
        T.end(c.stream);
        T.begin(c.stream, "wave_collect");
        hipLaunchKernelGGL(k_collect_free, dim3((cap / 4 + 1023) / 1024), dim3(1024), 0, c.stream, B.P, cur);
        T.end(c.stream);
        T.begin(c.stream, "wave_shadow");
        hipLaunchKernelGGL(shadow, gs, bp, lds, c.stream, c.d_scene, B.P, c.d_stats);
        T.end(c.stream);
    }
    return hipGetLastError();
}

// Renders `views` (each with its own camera, film, first frame and sample
// count) into `films`.  Items are processed in groups of whole sample layers
// so that a group's sample buffer holds at most kMaxGroupItems entries.
hipError_t wave_render_views(Context& c, const std::vector<ViewDesc>& views_in, float4* films, uint32_t capacity,
                             int block, bool use_graph) {
    hipError_t e = sync_scene(c);
    if (e != hipSuccess) return e;
    if (views_in.empty()) return hipSuccess;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    const size_t lds = (size_t)block * c.S.stack_lds * sizeof(uint32_t);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    std::vector<ViewDesc> views = views_in;
    const uint32_t nv = (uint32_t)views.size();
    uint32_t max_spp = 0;
    uint64_t pixels = 0;  // the finish enumeration: each view's pixels from a multiple of the finish block
    for (auto& v : views) {
        v.pix_begin = (uint32_t)pixels;
        pixels += ((uint64_t)v.cam.w * v.cam.h + kFinishBlock - 1) / kFinishBlock * kFinishBlock;
        max_spp = std::max(max_spp, v.spp);
    }
    if (pixels >= (1ull << 31) || (uint64_t)max_spp * nv >= (1ull << 31)) return hipErrorInvalidValue;
    // items per sample layer, and groups of whole layers within the sample buffer
    std::vector<uint64_t> layer_items(max_spp, 0);
    for (uint32_t s = 0; s < max_spp; ++s)
        for (const auto& v : views)
            if (v.spp > s) layer_items[s] += (uint64_t)v.cam.w * v.cam.h;
    std::vector<std::pair<uint32_t, uint32_t>> groups;  // sample ranges [s0, s1)
    uint64_t sb_need = 1;
    for (uint32_t s = 0; s < max_spp;) {
        uint64_t n = layer_items[s];
        if (n > kMaxGroupItems) return hipErrorInvalidValue;
        uint32_t e1 = s + 1;
        while (e1 < max_spp && n + layer_items[e1] <= kMaxGroupItems) n += layer_items[e1++];
        groups.push_back({s, e1});
        sb_need = std::max(sb_need, n);
        s = e1;
    }
    // Automatic pool: as many slots as the largest group has items, up to
    // 2^25 (about 8 GB of path state).  Larger pools measured faster up to
    // that size on the 1080p workloads (4M -> 32M slots: 1.35x on both the
    // 5M-triangle terrain and the 1M-triangle scene): each launch carries more
    // rays, so fewer launches, fuller persistent waves and shorter tails.  If
    // the device cannot hold it, the pool halves down to 2^22 slots.
    const bool automatic = capacity == 0;
    // the hit record holds a queue position in 29 bits
    if (capacity > (1u << 28)) return hipErrorInvalidValue;
    if (automatic) capacity = (uint32_t)std::min<uint64_t>(sb_need, kMaxAutoPool);
    capacity = std::max<uint32_t>(256u, (capacity + 255) & ~255u);
    const uint32_t requested = capacity;
    const uint32_t seg_need = nv;
    const WavePool* pool = c.wave ? &c.wave->P : nullptr;
    if (!pool || c.wave->requested != capacity || pool->sb_items < sb_need || pool->max_views < nv ||
        pool->max_segments < seg_need || pool->block != block || pool->stack_lds != c.S.stack_lds ||
        pool->tex != (c.S.textured != 0) || pool->mis != (c.S.mis != 0)) {
        wave_free(c);
        c.wave = new WaveBuffers();
        c.wave->P.block = block;
        c.wave->P.stack_lds = c.S.stack_lds;
        e = wave_alloc(*c.wave, capacity, (uint32_t)sb_need, std::max<uint32_t>(nv, 64), std::max<uint32_t>(seg_need, 256),
                       c.S.textured != 0, c.S.mis != 0);
        while (e == hipErrorOutOfMemory && automatic && capacity > (1u << 22)) {
            (void)hipGetLastError();
            wave_free(c);
            capacity = std::max<uint32_t>(1u << 22, capacity / 2);
            c.wave = new WaveBuffers();
            c.wave->P.block = block;
            c.wave->P.stack_lds = c.S.stack_lds;
            e = wave_alloc(*c.wave, capacity, (uint32_t)sb_need, std::max<uint32_t>(nv, 64),
                           std::max<uint32_t>(seg_need, 256), c.S.textured != 0, c.S.mis != 0);
        }
        if (e != hipSuccess) {
            wave_free(c);
            return e;
        }
        c.wave->persist_waves = c.S.gstack_lanes / 64u;
        c.wave->requested = requested;
    }
    WaveBuffers& B = *c.wave;
    if (c.timer.enabled) use_graph = false;  // per-kernel events need individual launches
    if (B.exec && B.graph_stats != (c.opt.stats != 0)) {
        hipGraphExecDestroy(B.exec);
        hipGraphDestroy(B.graph);
        B.exec = nullptr;
        B.graph = nullptr;
    }
    e = hipMemcpyAsync(B.P.views, views.data(), sizeof(ViewDesc) * nv, hipMemcpyHostToDevice, c.stream);
    if (e != hipSuccess) return e;
    for (const auto& g : groups) {
        // segments of this group in item order: one per view with samples in
        // it, holding its samples g.first .. min(g.second, spp) - 1
        std::vector<Segment> segs;
        std::vector<int32_t> seg_of(nv, -1);
        uint32_t items = 0;
        for (uint32_t v = 0; v < nv; ++v) {
            if (views[v].spp <= g.first) continue;
            const uint32_t layers = std::min(g.second, views[v].spp) - g.first;
            seg_of[v] = (int32_t)segs.size();
            segs.push_back(Segment{items, v, g.first, layers});
            items += views[v].cam.w * views[v].cam.h * layers;
        }
        WaveCtl ctl{};
        ctl.num_segments = (uint32_t)segs.size();
        ctl.num_views = nv;
        e = hipMemcpyAsync(B.P.segs, segs.data(), sizeof(Segment) * segs.size(), hipMemcpyHostToDevice, c.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(B.P.seg_of, seg_of.data(), sizeof(int32_t) * seg_of.size(), hipMemcpyHostToDevice,
                               c.stream);
        if (e == hipSuccess) e = hipMemcpyAsync(B.P.ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice, c.stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_stream_init, dim3((B.P.capacity + 255) / 256), dim3(256), 0, c.stream, B.P, items);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // Every live path retires within `bounces` iterations, so within that
        // many iterations the whole pool is started afresh: the bound below is
        // never reached unless the device state is corrupt.
        const uint64_t max_iters =
            2ull * ((uint64_t)items / B.P.capacity + 2) * (uint64_t)(std::max(c.S.bounces, 1) + 1) + 4 * kIters;
        for (uint64_t it = 0;; it += kIters) {
            if (use_graph) {
                if (!B.exec) {
                    e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
                    if (e != hipSuccess) return e;
                    hipError_t er = record_iterations(c, B, kIters);
                    e = hipStreamEndCapture(c.stream, &B.graph);
                    if (er != hipSuccess) return er;
                    if (e != hipSuccess) return e;
                    e = hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0);
                    if (e != hipSuccess) return e;
                    B.graph_stats = c.opt.stats != 0;
                }
                e = hipGraphLaunch(B.exec, c.stream);
            } else {
                e = record_iterations(c, B, kIters);
            }
            if (e != hipSuccess) return e;
            WaveCtl h{};
            e = hipMemcpyAsync(&h, B.P.ctl, sizeof(h), hipMemcpyDeviceToHost, c.stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
            if (e != hipSuccess) return e;
            // the last iteration of a replay has parity 1: its shade wrote the parity-0 queues
            if (h.item_next[0] >= h.item_end && h.ray_count[0] == 0 && h.free_count[0] == 0) break;
            if (h.item_next[0] >= h.item_end && h.ray_count[0] <= kDrainMax && c.S.bounces > 0) {
                // every item has started and few paths are left: finish them in one launch
                const uint32_t n = std::max<uint32_t>(h.ray_count[0], h.free_count[0]);
                const uint32_t blocks = std::min<uint32_t>((n + 63) / 64, c.S.gstack_lanes / 64u);
                const size_t dlds = (size_t)64 * c.S.stack_lds * sizeof(uint32_t);
                auto drain = c.opt.stats ? (c.S.textured ? k_drain<true, true> : k_drain<true, false>)
                                         : (c.S.textured ? k_drain<false, true> : k_drain<false, false>);
                c.timer.begin(c.stream, "wave_drain");
                hipLaunchKernelGGL(drain, dim3(std::max(blocks, 1u)), dim3(64), dlds, c.stream, c.d_scene, B.P,
                                   c.d_stats);
                c.timer.end(c.stream);
                if ((e = hipGetLastError()) != hipSuccess) return e;
                break;
            }
            if (it > max_iters) return hipErrorUnknown;
        }
        c.timer.begin(c.stream, "wave_finish");
        hipLaunchKernelGGL(k_finish, dim3((uint32_t)(pixels / kFinishBlock)), dim3(kFinishBlock), 0, c.stream, B.P,
                           films, g.first, g.second);
        c.timer.end(c.stream);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t wave_render(Context& c, uint32_t frame0, uint32_t spp) {
    ViewDesc v{};
    v.cam = c.S.cam;
    v.film_offset = 0;
    v.frame0 = frame0;
    v.spp = spp;
    v.seed = 0;
    const bool graph = getenv("NOOR_NO_GRAPH") == nullptr;
    return wave_render_views(c, {v}, c.film, c.wave_capacity, c.wave_block, graph);
}

hipError_t wave_render_batch(Context& c, const noor_view* v, uint32_t n, float4* films, const size_t* offsets) {
    std::vector<ViewDesc> views(n);
    for (uint32_t i = 0; i < n; ++i) {
        ViewDesc d{};
        d.cam = v[i].camera;
        d.film_offset = (uint32_t)offsets[i];
        d.frame0 = v[i].frame;
        d.spp = v[i].spp;
        d.seed = v[i].seed;
        views[i] = d;
    }
    const bool graph = getenv("NOOR_NO_GRAPH") == nullptr;
    return wave_render_views(c, views, films, c.wave_capacity, c.wave_block, graph);
}

}  // namespace noor
