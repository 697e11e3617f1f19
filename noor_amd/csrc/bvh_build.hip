// GPU binned-SAH builder of bottom-level trees (see bvh_gpu.h).
//
// Per level, over every node of the level at once:
//   k_level_small   one wave per node: leaf rules, and for nodes of up to
//                   kSmall triangles the binning (LDS), the split choice (lane
//                   0, the host's arithmetic) and a stable partition (ballots)
//   k_bin_chunks    larger nodes: one block per kChunk triangles bins into LDS
//                   and writes its partial bins
//   k_split_big     one block per larger node merges the partial bins and
//                   chooses the split
//   k_flag_chunks, k_scan_chunks, k_scatter_chunks
//                   stable partition of the larger nodes: per-block left
//                   counts, their per-node prefix, then the scatter
//   k_children      children of the split nodes, numbered by a scan
// then k_subtree_size (bottom up), k_position (top down) and k_emit lay the
// breadth-first tree out depth first.
//
// Float minima and maxima are exact and order independent; merged through
// LDS atomics they use an order-preserving integer image of the float.
// Negative zero never occurs (the host canonicalises the primitive bounds),
// so the bins equal the host's sequential ones bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/noor_scene.h"
#include "bvh_gpu.h"

namespace noor {
namespace {

constexpr uint32_t kMaxBins = 16;   // bins per axis the device path supports
constexpr uint32_t kSmall = 256;    // nodes up to this size: one wave each
constexpr uint32_t kChunk = 256;    // triangles per block for larger nodes
constexpr uint32_t kWavesPerBlock = 4;
constexpr uint32_t kBinWords = 13;  // count, box min/max, centroid min/max
constexpr float kInfF = __builtin_huge_valf();

enum : uint32_t { NF_PENDING = 0, NF_LEAF = 1, NF_INNER = 2, NF_BIG = 3 };

struct DBox {
    float mn[3], mx[3];
};

struct BNode {
    DBox box, cent;
    uint32_t begin, end, level, flags;
    uint32_t axis, nl, left;  // left: node index of the left child (right = left + 1)
    float bound;
    DBox lbox, lcent, rbox, rcent;
};

struct DParams {
    uint32_t nb, max_height, min_leaf, max_leaf;
    float Ci, Ct;
};

__device__ __forceinline__ float mnf(float a, float b) { return (b < a) ? b : a; }  // std::min(a, b)
__device__ __forceinline__ float mxf(float a, float b) { return (a < b) ? b : a; }  // std::max(a, b)
__device__ __forceinline__ void box_init(DBox& b) {
    for (int i = 0; i < 3; ++i) {
        b.mn[i] = kInfF;
        b.mx[i] = -kInfF;
    }
}
__device__ __forceinline__ void box_merge(DBox& b, const DBox& o) {
    for (int i = 0; i < 3; ++i) {
        b.mn[i] = mnf(b.mn[i], o.mn[i]);
        b.mx[i] = mxf(b.mx[i], o.mx[i]);
    }
}
// Box::area of the host builder: infinite for an empty box
__device__ __forceinline__ float box_area(const DBox& b) {
    for (int i = 0; i < 3; ++i)
        if (!isfinite(b.mn[i]) || !isfinite(b.mx[i])) return kInfF;
    const float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
    return 2.0f * (dx * dy + dy * dz + dx * dz);
}

// order-preserving image of a float in uint32 (no negative zero occurs)
__device__ __forceinline__ uint32_t f2o(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// Bin upper bounds of one axis (Split::reset, noor/header/splitter.h:127-145),
// as the host computes them.
__device__ __forceinline__ void bin_bounds(const DBox& cent, int a, uint32_t nb, float* bmax) {
    const float dt = (cent.mx[a] - cent.mn[a]) * (1.0f / (float)nb);
    bmax[0] = cent.mn[a] + dt;
    for (uint32_t i = 1; i < nb; ++i) bmax[i] = bmax[i - 1] + dt;
    bmax[nb - 1] = cent.mx[a];
}
// smallest bin whose upper bound is >= c
__device__ __forceinline__ uint32_t bin_of(const float* bmax, uint32_t nb, float c) {
    uint32_t l = 0, r = nb - 1;
    while (l < r) {
        const uint32_t m = (l + r) >> 1;
        if (c <= bmax[m]) r = m;
        else l = m + 1;
    }
    return l;
}

// LDS bins: [axis][bin][word], word 0 count, 1-3 box min, 4-6 box max,
// 7-9 centroid min, 10-12 centroid max (ordered images)
__device__ __forceinline__ void bins_clear(uint32_t* b, uint32_t nb, uint32_t lane, uint32_t lanes) {
    for (uint32_t i = lane; i < 3 * nb; i += lanes) {
        uint32_t* w = b + i * kBinWords;
        w[0] = 0;
        for (int k = 0; k < 3; ++k) {
            w[1 + k] = f2o(kInfF);
            w[4 + k] = f2o(-kInfF);
            w[7 + k] = f2o(kInfF);
            w[10 + k] = f2o(-kInfF);
        }
    }
}
__device__ __forceinline__ void bins_add(uint32_t* b, uint32_t nb, const float* bmax3, const GpuPrim& p) {
    for (int a = 0; a < 3; ++a) {
        const uint32_t l = bin_of(bmax3 + a * nb, nb, p.c[a]);
        uint32_t* w = b + (a * nb + l) * kBinWords;
        atomicAdd(w, 1u);
        for (int k = 0; k < 3; ++k) {
            atomicMin(w + 1 + k, f2o(p.mn[k]));
            atomicMax(w + 4 + k, f2o(p.mx[k]));
            atomicMin(w + 7 + k, f2o(p.c[k]));
            atomicMax(w + 10 + k, f2o(p.c[k]));
        }
    }
}
__device__ __forceinline__ void bin_read(const uint32_t* w, uint32_t& cnt, DBox& bb, DBox& bc) {
    cnt = w[0];
    for (int k = 0; k < 3; ++k) {
        bb.mn[k] = o2f(w[1 + k]);
        bb.mx[k] = o2f(w[4 + k]);
        bc.mn[k] = o2f(w[7 + k]);
        bc.mx[k] = o2f(w[10 + k]);
    }
}

// find_split of the host builder (noor/source/splitter.cpp:62-132) from the
// bins; `suffix` is scratch for 13 * nb words.  Fills the node's split.
__device__ void choose_split(const uint32_t* bins, uint32_t* suffix, const DParams& P, BNode& nd,
                             const float* bmax3) {
    const uint32_t nb = P.nb;
    const float parent_area = box_area(nd.box);
    float best = kInfF;
    uint32_t best_nl = 0, best_nr = 0;
    for (int a = 0; a < 3; ++a) {
        const uint32_t n = nb - 1;
        // suffix merges rb, rc, rcnt
        {
            uint32_t c;
            DBox bb, bc;
            bin_read(bins + (a * nb + n) * kBinWords, c, bb, bc);
            float* s = reinterpret_cast<float*>(suffix + n * kBinWords);
            suffix[n * kBinWords] = c;
            for (int k = 0; k < 3; ++k) {
                s[1 + k] = bb.mn[k];
                s[4 + k] = bb.mx[k];
                s[7 + k] = bc.mn[k];
                s[10 + k] = bc.mx[k];
            }
            for (uint32_t i = n - 1; i > 0; --i) {
                bin_read(bins + (a * nb + i) * kBinWords, c, bb, bc);
                const float* q = reinterpret_cast<const float*>(suffix + (i + 1) * kBinWords);
                float* o = reinterpret_cast<float*>(suffix + i * kBinWords);
                suffix[i * kBinWords] = c + suffix[(i + 1) * kBinWords];
                for (int k = 0; k < 3; ++k) {
                    o[1 + k] = mnf(bb.mn[k], q[1 + k]);
                    o[4 + k] = mxf(bb.mx[k], q[4 + k]);
                    o[7 + k] = mnf(bc.mn[k], q[7 + k]);
                    o[10 + k] = mxf(bc.mx[k], q[10 + k]);
                }
            }
        }
        DBox lb, lc;
        box_init(lb);
        box_init(lc);
        uint32_t lcount = 0;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t c;
            DBox bb, bc;
            bin_read(bins + (a * nb + i) * kBinWords, c, bb, bc);
            box_merge(lb, bb);
            box_merge(lc, bc);
            lcount += c;
            const uint32_t rcount = suffix[(i + 1) * kBinWords];
            if (lcount == 0 || rcount == 0) continue;
            const float* q = reinterpret_cast<const float*>(suffix + (i + 1) * kBinWords);
            DBox rb;
            for (int k = 0; k < 3; ++k) {
                rb.mn[k] = q[1 + k];
                rb.mx[k] = q[4 + k];
            }
            const float cost = P.Ct + P.Ci * ((float)lcount * box_area(lb) + (float)rcount * box_area(rb)) / parent_area;
            if (cost < best) {
                best = cost;
                nd.axis = (uint32_t)a;
                best_nl = lcount;
                best_nr = rcount;
                nd.bound = bmax3[a * nb + i];
                nd.lbox = lb;
                nd.lcent = lc;
                nd.rbox = rb;
                for (int k = 0; k < 3; ++k) {
                    nd.rcent.mn[k] = q[7 + k];
                    nd.rcent.mx[k] = q[10 + k];
                }
            }
        }
    }
    const uint32_t count = nd.end - nd.begin;
    const bool valid = best != kInfF && best_nl != 0 && best_nr != 0;
    const float leaf_cost = P.Ci * (float)count;
    if (!valid || (leaf_cost < best && count <= P.max_leaf)) {
        nd.flags = NF_LEAF;
    } else {
        nd.flags = NF_INNER;
        nd.nl = best_nl;
    }
}

// ---- small nodes: one wave each ------------------------------------------------------------
__global__ __launch_bounds__(256) void k_level_small(BNode* __restrict__ nodes, uint32_t lo, uint32_t hi,
                                                     const GpuPrim* __restrict__ prims,
                                                     const uint32_t* __restrict__ perm_in,
                                                     uint32_t* __restrict__ perm_out,
                                                     uint32_t* __restrict__ perm_final, DParams P,
                                                     uint32_t* __restrict__ big_list, uint32_t* __restrict__ big_count) {
    __shared__ uint32_t lds[kWavesPerBlock][(3 * kMaxBins + kMaxBins) * kBinWords];
    __shared__ float lbmax[kWavesPerBlock][3 * kMaxBins];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t ni = lo + blockIdx.x * kWavesPerBlock + wave;
    if (ni >= hi) return;  // uniform per wave; no block-wide barriers below
    BNode nd = nodes[ni];
    const uint32_t count = nd.end - nd.begin;
    const bool rule_leaf = nd.level >= P.max_height || count <= P.min_leaf;
    if (!rule_leaf && count > kSmall) {
        if (lane == 0) {
            nodes[ni].flags = NF_BIG;
            const uint32_t slot = atomicAdd(big_count, 1u);
            big_list[3 * slot] = ni;
            big_list[3 * slot + 1] = nd.begin;
            big_list[3 * slot + 2] = nd.end;
        }
        return;
    }
    uint32_t* bins = lds[wave];
    float* bmax3 = lbmax[wave];
    if (!rule_leaf) {
        if (lane < 3) bin_bounds(nd.cent, (int)lane, P.nb, bmax3 + lane * P.nb);
        bins_clear(bins, P.nb, lane, 64);
        // the wave's own LDS traffic: complete each phase before the next
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        for (uint32_t k = nd.begin + lane; k < nd.end; k += 64) bins_add(bins, P.nb, bmax3, prims[perm_in[k]]);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) choose_split(bins, bins + 3 * P.nb * kBinWords, P, nd, bmax3);
        nd.flags = __shfl(nd.flags, 0, 64);
    } else {
        nd.flags = NF_LEAF;
    }
    if (nd.flags == NF_LEAF) {
        for (uint32_t k = nd.begin + lane; k < nd.end; k += 64) perm_final[k] = perm_in[k];
        if (lane == 0) nodes[ni].flags = NF_LEAF;
        return;
    }
    // stable partition on centroid[axis] <= bound, 64 triangles at a time
    const uint32_t axis = __shfl(nd.axis, 0, 64);
    const float bound = __shfl(nd.bound, 0, 64);
    const uint32_t nl = __shfl(nd.nl, 0, 64);
    uint32_t left_base = nd.begin, right_base = nd.begin + nl;
    for (uint32_t k0 = nd.begin; k0 < nd.end; k0 += 64) {
        const uint32_t k = k0 + lane;
        const bool have = k < nd.end;
        uint32_t id = 0;
        bool left = false;
        if (have) {
            id = perm_in[k];
            left = prims[id].c[axis] <= bound;
        }
        const unsigned long long lm = __ballot(have && left), rm = __ballot(have && !left);
        const unsigned long long below = (1ull << lane) - 1ull;
        if (have) {
            const uint32_t dst = left ? left_base + (uint32_t)__popcll(lm & below)
                                      : right_base + (uint32_t)__popcll(rm & below);
            perm_out[dst] = id;
        }
        left_base += (uint32_t)__popcll(lm);
        right_base += (uint32_t)__popcll(rm);
    }
    if (lane == 0) nodes[ni] = nd;
}

// ---- larger nodes: blocks of kChunk triangles ----------------------------------------------
struct Chunk {
    uint32_t node, first;  // node index, first triangle (absolute position in perm)
    uint32_t local;        // chunk number within its node
    uint32_t big;          // the node's slot among the larger nodes of the level
    uint32_t node_chunk0;  // index of the node's first chunk
};

__global__ __launch_bounds__(256) void k_bin_chunks(const BNode* __restrict__ nodes, const Chunk* __restrict__ chunks,
                                                    const GpuPrim* __restrict__ prims,
                                                    const uint32_t* __restrict__ perm_in, DParams P,
                                                    uint32_t* __restrict__ node_bins) {
    __shared__ uint32_t bins[3 * kMaxBins * kBinWords];
    __shared__ float bmax3[3 * kMaxBins];
    const Chunk ch = chunks[blockIdx.x];
    const BNode& nd = nodes[ch.node];
    if (threadIdx.x < 3) bin_bounds(nd.cent, (int)threadIdx.x, P.nb, bmax3 + threadIdx.x * P.nb);
    bins_clear(bins, P.nb, threadIdx.x, blockDim.x);
    __syncthreads();
    const uint32_t k = ch.first + threadIdx.x;
    if (k < nd.end) bins_add(bins, P.nb, bmax3, prims[perm_in[k]]);
    __syncthreads();
    // merge into the node's bins (min / max of ordered images, sums of counts)
    uint32_t* out = node_bins + (size_t)ch.big * 3 * kMaxBins * kBinWords;
    for (uint32_t i = threadIdx.x; i < 3 * P.nb * kBinWords; i += blockDim.x) {
        const uint32_t w = i % kBinWords, v = bins[i];
        if (w == 0) {
            if (v) atomicAdd(out + i, v);
        } else if (w <= 3 || (w >= 7 && w <= 9)) {
            if (v != f2o(kInfF)) atomicMin(out + i, v);
        } else {
            if (v != f2o(-kInfF)) atomicMax(out + i, v);
        }
    }
}

// identity values of the per-node bins of the larger nodes
__global__ void k_clear_node_bins(uint32_t* __restrict__ node_bins, uint32_t nbig, uint32_t nb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t per = 3 * kMaxBins * kBinWords;
    if (i >= nbig * per) return;
    const uint32_t w = (i % per) % kBinWords;
    node_bins[i] = w == 0 ? 0u : ((w <= 3 || (w >= 7 && w <= 9)) ? f2o(kInfF) : f2o(-kInfF));
    (void)nb;
}

// one block per larger node: choose the split from its merged bins
__global__ __launch_bounds__(64) void k_split_big(BNode* __restrict__ nodes, const uint32_t* __restrict__ big_list,
                                                  const uint32_t* __restrict__ node_bins, DParams P) {
    __shared__ uint32_t bins[(3 * kMaxBins + kMaxBins) * kBinWords];
    __shared__ float bmax3[3 * kMaxBins];
    const uint32_t ni = big_list[blockIdx.x];
    const uint32_t words = 3 * P.nb * kBinWords;
    const uint32_t* src = node_bins + (size_t)blockIdx.x * 3 * kMaxBins * kBinWords;
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) bins[i] = src[i];
    if (threadIdx.x < 3) bin_bounds(nodes[ni].cent, (int)threadIdx.x, P.nb, bmax3 + threadIdx.x * P.nb);
    __syncthreads();
    if (threadIdx.x == 0) {
        BNode nd = nodes[ni];
        choose_split(bins, bins + 3 * P.nb * kBinWords, P, nd, bmax3);
        nodes[ni] = nd;
    }
}

// per chunk: how many of its triangles go left
__global__ __launch_bounds__(256) void k_flag_chunks(const BNode* __restrict__ nodes, const Chunk* __restrict__ chunks,
                                                     const GpuPrim* __restrict__ prims,
                                                     const uint32_t* __restrict__ perm_in,
                                                     uint32_t* __restrict__ chunk_left) {
    __shared__ uint32_t cnt;
    const Chunk ch = chunks[blockIdx.x];
    const BNode& nd = nodes[ch.node];
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const uint32_t k = ch.first + threadIdx.x;
    if (nd.flags == NF_INNER && k < nd.end && prims[perm_in[k]].c[nd.axis] <= nd.bound) atomicAdd(&cnt, 1u);
    __syncthreads();
    if (threadIdx.x == 0) chunk_left[blockIdx.x] = cnt;
}

__global__ __launch_bounds__(256) void k_scatter_chunks(const BNode* __restrict__ nodes,
                                                        const Chunk* __restrict__ chunks,
                                                        const GpuPrim* __restrict__ prims,
                                                        const uint32_t* __restrict__ perm_in,
                                                        const uint32_t* __restrict__ chunk_left,
                                                        uint32_t* __restrict__ perm_out,
                                                        uint32_t* __restrict__ perm_final) {
    __shared__ uint32_t wl[4], wr[4];
    const Chunk ch = chunks[blockIdx.x];
    const BNode& nd = nodes[ch.node];
    const uint32_t k = ch.first + threadIdx.x;
    const bool have = k < nd.end;
    if (nd.flags == NF_LEAF) {  // no valid split: the whole node is one leaf
        if (have) perm_final[k] = perm_in[k];
        return;
    }
    uint32_t id = 0;
    bool left = false;
    if (have) {
        id = perm_in[k];
        left = prims[id].c[nd.axis] <= nd.bound;
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long lm = __ballot(have && left), rm = __ballot(have && !left);
    if (lane == 0) {
        wl[wave] = (uint32_t)__popcll(lm);
        wr[wave] = (uint32_t)__popcll(rm);
    }
    __syncthreads();
    uint32_t lb = 0, rb = 0;
    for (uint32_t w = 0; w < wave; ++w) {
        lb += wl[w];
        rb += wr[w];
    }
    if (!have) return;
    const unsigned long long below = (1ull << lane) - 1ull;
    // chunk_left holds an exclusive scan over all chunks of the level
    const uint32_t left_before = chunk_left[blockIdx.x] - chunk_left[ch.node_chunk0];  // lefts in earlier chunks of the node
    const uint32_t right_before = ch.local * kChunk - left_before;       // rights in earlier chunks
    const uint32_t dst = left ? nd.begin + left_before + lb + (uint32_t)__popcll(lm & below)
                              : nd.begin + nd.nl + right_before + rb + (uint32_t)__popcll(rm & below);
    perm_out[dst] = id;
}

__global__ void k_iota(uint32_t* __restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// ---- children ------------------------------------------------------------------------------
__global__ void k_split_flags(const BNode* __restrict__ nodes, uint32_t lo, uint32_t hi, uint32_t* __restrict__ f) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (lo + i < hi) f[i] = nodes[lo + i].flags == NF_INNER ? 1u : 0u;
}

__global__ void k_children(BNode* __restrict__ nodes, uint32_t lo, uint32_t hi, const uint32_t* __restrict__ rank) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (lo + i >= hi) return;
    BNode& nd = nodes[lo + i];
    if (nd.flags != NF_INNER) return;
    const uint32_t c = hi + 2 * rank[i];
    nd.left = c;
    BNode l, r;
    l.box = nd.lbox;
    l.cent = nd.lcent;
    l.begin = nd.begin;
    l.end = nd.begin + nd.nl;
    r.box = nd.rbox;
    r.cent = nd.rcent;
    r.begin = nd.begin + nd.nl;
    r.end = nd.end;
    l.level = r.level = nd.level + 1;
    l.flags = r.flags = NF_PENDING;
    l.axis = r.axis = l.nl = r.nl = l.left = r.left = 0;
    l.bound = r.bound = 0.0f;
    nodes[c] = l;
    nodes[c + 1] = r;
}

// ---- exclusive scan of uint32 (three passes) -----------------------------------------------
constexpr uint32_t kScanBlock = 1024;

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                            uint32_t* __restrict__ sums) {
    __shared__ uint32_t s[kScanBlock / 64];
    const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    uint32_t v = i < n ? in[i] : 0u;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kScanBlock / 64; ++w) t += s[w];
        sums[blockIdx.x] = t;
    }
}
// single block: exclusive scan of the block sums, total in sums[nblocks]
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(uint32_t* __restrict__ sums, uint32_t nblocks) {
    __shared__ uint32_t s[kScanBlock];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nblocks; base += kScanBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nblocks ? sums[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < kScanBlock; off <<= 1) {
            const uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nblocks) sums[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == kScanBlock - 1) carry += s[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[nblocks] = carry;
}
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const uint32_t* __restrict__ in, uint32_t n,
                                                           const uint32_t* __restrict__ sums,
                                                           uint32_t* __restrict__ out) {
    __shared__ uint32_t s[kScanBlock / 64];
    const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t v = i < n ? in[i] : 0u;
    uint32_t x = v;  // inclusive scan within the wave
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += t;
    }
    if (lane == 63) s[wave] = x;
    __syncthreads();
    uint32_t before = sums[blockIdx.x];
    for (uint32_t w = 0; w < wave; ++w) before += s[w];
    if (i < n) out[i] = before + x - v;
}

// ---- depth-first layout ------------------------------------------------------------------
__global__ void k_subtree_size(const BNode* __restrict__ nodes, uint32_t lo, uint32_t hi, uint32_t* __restrict__ size) {
    const uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    const BNode& nd = nodes[i];
    size[i] = nd.flags == NF_INNER ? 1u + size[nd.left] + size[nd.left + 1] : 1u;
}
__global__ void k_position(const BNode* __restrict__ nodes, uint32_t lo, uint32_t hi, const uint32_t* __restrict__ size,
                           uint32_t* __restrict__ pos) {
    const uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    const BNode& nd = nodes[i];
    if (nd.flags != NF_INNER) return;
    pos[nd.left] = pos[i] + 1u;
    pos[nd.left + 1] = pos[i] + 1u + size[nd.left];
}
__global__ void k_emit(const BNode* __restrict__ nodes, uint32_t n, const uint32_t* __restrict__ pos,
                       uint32_t node_base, uint32_t tri_base, uint4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BNode& nd = nodes[i];
    uint4 a, b;
    a.x = __float_as_uint(nd.box.mn[0]);
    a.y = __float_as_uint(nd.box.mn[1]);
    a.z = __float_as_uint(nd.box.mn[2]);
    b.x = __float_as_uint(nd.box.mx[0]);
    b.y = __float_as_uint(nd.box.mx[1]);
    b.z = __float_as_uint(nd.box.mx[2]);
    if (nd.flags == NF_INNER) {
        a.w = node_base + pos[nd.left + 1];
        b.w = nd.axis;
    } else {
        a.w = tri_base + nd.begin;
        b.w = NOOR_TRI_LEAF_MASK | (nd.end - nd.begin);
    }
    out[2 * (size_t)pos[i]] = a;
    out[2 * (size_t)pos[i] + 1] = b;
}

#define NOOR_HIP_TRY(x)                                                                   \
    do {                                                                                  \
        const hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) return std::string(#x ": ") + hipGetErrorString(e_);        \
    } while (0)

// Device memory of the builder, kept per device for the life of the process
// and grown when a larger mesh comes: allocating the ~2.7 GB a 5M-triangle
// build needs took about 100 ms, as long as the build itself, which repeated
// builds (animation, batches of scenes) should not pay.  Calls on one device
// are serialised by its mutex.
struct Workspace {
    std::mutex m;
    hipStream_t s = nullptr;
    size_t cap = 0;  // triangles the buffers are sized for
    void *prims = nullptr, *nodes = nullptr, *perm0 = nullptr, *perm1 = nullptr, *fin = nullptr, *big = nullptr,
         *bigc = nullptr, *chunks = nullptr, *partial = nullptr, *cleft = nullptr, *flags = nullptr,
         *rank = nullptr, *sums = nullptr, *size = nullptr, *pos = nullptr, *out = nullptr;
};

struct Buf {
    void* p;
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

size_t max_nodes_for(size_t n) { return 2 * n + 1; }
// per level: sum over nodes above kSmall of ceil(count / kChunk) <= n / kChunk + n / kSmall
size_t max_chunks_for(size_t n) { return n / kChunk + n / kSmall + 64; }

std::string ensure(Workspace& w, size_t n) {
    if (!w.s) NOOR_HIP_TRY(hipStreamCreateWithFlags(&w.s, hipStreamNonBlocking));
    if (n <= w.cap) return std::string();
    void** all[] = {&w.prims, &w.nodes, &w.perm0, &w.perm1, &w.fin, &w.big, &w.bigc, &w.chunks,
                    &w.partial, &w.cleft, &w.flags, &w.rank, &w.sums, &w.size, &w.pos, &w.out};
    for (void** q : all)
        if (*q) {
            NOOR_HIP_TRY(hipFree(*q));
            *q = nullptr;
        }
    w.cap = 0;
    const size_t mn = max_nodes_for(n), mc = max_chunks_for(n);
    const size_t bytes[] = {sizeof(GpuPrim) * n, 0, 4 * n, 4 * n, 4 * n, 12 * mc, 16, 0,
                            4 * (n / kSmall + 2) * 3 * kMaxBins * kBinWords, 8 * mc, 4 * mn, 4 * mn,
                            4 * (std::max(mn, mc) / kScanBlock + 2), 4 * mn, 4 * mn, 32 * mn};
    for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k) {
        size_t b = bytes[k];
        if (all[k] == &w.nodes) b = sizeof(BNode) * mn;
        if (all[k] == &w.chunks) b = sizeof(Chunk) * mc;
        NOOR_HIP_TRY(hipMalloc(all[k], std::max<size_t>(b, 16)));
    }
    w.cap = n;
    return std::string();
}

Workspace& workspace(int device) {
    static std::mutex gm;
    static std::vector<Workspace*> ws;  // never freed: the HIP runtime may be gone at exit
    std::lock_guard<std::mutex> g(gm);
    if ((size_t)device >= ws.size()) ws.resize((size_t)device + 1, nullptr);
    if (!ws[device]) ws[device] = new Workspace();
    return *ws[device];
}

std::string scan(hipStream_t s, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* sums, uint32_t& total) {
    const uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (n == 0) {
        total = 0;
        return std::string();
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, s, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, s, sums, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanBlock), 0, s, in, n, sums, out);
    NOOR_HIP_TRY(hipGetLastError());
    NOOR_HIP_TRY(hipMemcpyAsync(&total, sums + nb, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NOOR_HIP_TRY(hipStreamSynchronize(s));
    return std::string();
}

}  // namespace

std::string gpu_build_blas(int device, const GpuPrim* prims, uint32_t n, const float root_box[6],
                           const float root_cent[6], const GpuBvhParams& GP, uint32_t node_base, uint32_t tri_base,
                           GpuBvhResult& out) {
    if (GP.num_bins < 2 || GP.num_bins > kMaxBins) return "the device builder supports 2 to 16 bins";
    if (n == 0) return "empty mesh";
    int cur = 0;
    NOOR_HIP_TRY(hipGetDevice(&cur));
    NOOR_HIP_TRY(hipSetDevice(device));
    struct Restore {
        int d;
        ~Restore() { (void)hipSetDevice(d); }
    } restore{cur};
    const DParams P{GP.num_bins, GP.max_height, GP.min_leaf, GP.max_leaf, GP.Ci, GP.Ct};
    const size_t max_nodes = max_nodes_for(n);
    const size_t max_chunks = max_chunks_for(n);
    Workspace& W = workspace(device);
    std::lock_guard<std::mutex> lock(W.m);
    std::string e;
    if (!(e = ensure(W, n)).empty()) return e;
    hipStream_t s = W.s;
    const Buf d_prims{W.prims}, d_nodes{W.nodes}, d_perm0{W.perm0}, d_perm1{W.perm1}, d_final{W.fin}, d_big{W.big},
        d_bigc{W.bigc}, d_chunks{W.chunks}, d_partial{W.partial}, d_cleft{W.cleft}, d_flags{W.flags},
        d_rank{W.rank}, d_sums{W.sums}, d_size{W.size}, d_pos{W.pos}, d_out{W.out};
    const auto t0 = std::chrono::steady_clock::now();
    NOOR_HIP_TRY(hipMemcpyAsync(d_prims.p, prims, sizeof(GpuPrim) * (size_t)n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, d_perm0.as<uint32_t>(), n);
    {
        BNode root{};
        for (int k = 0; k < 3; ++k) {
            root.box.mn[k] = root_box[k];
            root.box.mx[k] = root_box[3 + k];
            root.cent.mn[k] = root_cent[k];
            root.cent.mx[k] = root_cent[3 + k];
        }
        root.begin = 0;
        root.end = n;
        root.level = 0;
        root.flags = NF_PENDING;
        NOOR_HIP_TRY(hipMemcpyAsync(d_nodes.p, &root, sizeof(BNode), hipMemcpyHostToDevice, s));
        NOOR_HIP_TRY(hipStreamSynchronize(s));
    }
    uint32_t* perm_in = d_perm0.as<uint32_t>();
    uint32_t* perm_out = d_perm1.as<uint32_t>();
    std::vector<uint32_t> level_off{0u, 1u};
    std::vector<uint32_t> big;
    std::vector<Chunk> chunks;
    while (level_off.back() > level_off[level_off.size() - 2]) {
        const uint32_t lo = level_off[level_off.size() - 2], hi = level_off.back();
        const uint32_t K = hi - lo;
        if (level_off.size() > 4096) return "tree too deep";
        NOOR_HIP_TRY(hipMemsetAsync(d_bigc.p, 0, 4, s));
        const uint32_t blocks = (K + kWavesPerBlock - 1) / kWavesPerBlock;
        hipLaunchKernelGGL(k_level_small, dim3(blocks), dim3(64 * kWavesPerBlock), 0, s, d_nodes.as<BNode>(), lo, hi,
                           d_prims.as<GpuPrim>(), perm_in, perm_out, d_final.as<uint32_t>(), P, d_big.as<uint32_t>(),
                           d_bigc.as<uint32_t>());
        NOOR_HIP_TRY(hipGetLastError());
        uint32_t nbig = 0;
        NOOR_HIP_TRY(hipMemcpyAsync(&nbig, d_bigc.p, 4, hipMemcpyDeviceToHost, s));
        NOOR_HIP_TRY(hipStreamSynchronize(s));
        if (nbig) {
            std::vector<uint32_t> trip(3 * (size_t)nbig);
            NOOR_HIP_TRY(hipMemcpy(trip.data(), d_big.p, 12 * (size_t)nbig, hipMemcpyDeviceToHost));
            std::vector<uint32_t> order(nbig);
            for (uint32_t b = 0; b < nbig; ++b) order[b] = b;
            std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return trip[3 * x] < trip[3 * y]; });
            big.resize(nbig);
            chunks.clear();
            for (uint32_t b = 0; b < nbig; ++b) {
                const uint32_t* t = &trip[3 * (size_t)order[b]];
                big[b] = t[0];
                uint32_t local = 0;
                const uint32_t c0 = (uint32_t)chunks.size();
                for (uint32_t f = t[1]; f < t[2]; f += kChunk) chunks.push_back({t[0], f, local++, b, c0});
            }
            const uint32_t nc = (uint32_t)chunks.size();
            if (nc > max_chunks) return "chunk table overflow";
            NOOR_HIP_TRY(hipMemcpyAsync(d_big.p, big.data(), 4 * (size_t)nbig, hipMemcpyHostToDevice, s));
            NOOR_HIP_TRY(hipMemcpyAsync(d_chunks.p, chunks.data(), sizeof(Chunk) * nc, hipMemcpyHostToDevice, s));
            const uint32_t per = 3 * kMaxBins * kBinWords;
            hipLaunchKernelGGL(k_clear_node_bins, dim3((nbig * per + 255) / 256), dim3(256), 0, s,
                               d_partial.as<uint32_t>(), nbig, P.nb);
            hipLaunchKernelGGL(k_bin_chunks, dim3(nc), dim3(kChunk), 0, s, d_nodes.as<BNode>(), d_chunks.as<Chunk>(),
                               d_prims.as<GpuPrim>(), perm_in, P, d_partial.as<uint32_t>());
            hipLaunchKernelGGL(k_split_big, dim3(nbig), dim3(64), 0, s, d_nodes.as<BNode>(), d_big.as<uint32_t>(),
                               d_partial.as<uint32_t>(), P);
            hipLaunchKernelGGL(k_flag_chunks, dim3(nc), dim3(kChunk), 0, s, d_nodes.as<BNode>(), d_chunks.as<Chunk>(),
                               d_prims.as<GpuPrim>(), perm_in, d_cleft.as<uint32_t>());
            uint32_t lefts = 0;
            if (!(e = scan(s, d_cleft.as<uint32_t>(), d_cleft.as<uint32_t>() + max_chunks, nc, d_sums.as<uint32_t>(),
                           lefts)).empty())
                return e;
            hipLaunchKernelGGL(k_scatter_chunks, dim3(nc), dim3(kChunk), 0, s, d_nodes.as<BNode>(),
                               d_chunks.as<Chunk>(), d_prims.as<GpuPrim>(), perm_in,
                               d_cleft.as<uint32_t>() + max_chunks, perm_out, d_final.as<uint32_t>());
            NOOR_HIP_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL(k_split_flags, dim3((K + 255) / 256), dim3(256), 0, s, d_nodes.as<BNode>(), lo, hi,
                           d_flags.as<uint32_t>());
        uint32_t nsplit = 0;
        if (!(e = scan(s, d_flags.as<uint32_t>(), d_rank.as<uint32_t>(), K, d_sums.as<uint32_t>(), nsplit)).empty())
            return e;
        if ((size_t)hi + 2 * (size_t)nsplit > max_nodes) return "node table overflow";
        if (nsplit) {
            hipLaunchKernelGGL(k_children, dim3((K + 255) / 256), dim3(256), 0, s, d_nodes.as<BNode>(), lo, hi,
                               d_rank.as<uint32_t>());
            NOOR_HIP_TRY(hipGetLastError());
        }
        level_off.push_back(hi + 2 * nsplit);
        std::swap(perm_in, perm_out);
    }
    const uint32_t total = level_off.back();
    const uint32_t levels = (uint32_t)level_off.size() - 2;  // the last range is empty
    for (uint32_t L = levels; L-- > 0;) {
        const uint32_t lo = level_off[L], hi = level_off[L + 1];
        hipLaunchKernelGGL(k_subtree_size, dim3((hi - lo + 255) / 256), dim3(256), 0, s, d_nodes.as<BNode>(), lo, hi,
                           d_size.as<uint32_t>());
    }
    NOOR_HIP_TRY(hipMemsetAsync(d_pos.p, 0, 4, s));
    for (uint32_t L = 0; L < levels; ++L) {
        const uint32_t lo = level_off[L], hi = level_off[L + 1];
        hipLaunchKernelGGL(k_position, dim3((hi - lo + 255) / 256), dim3(256), 0, s, d_nodes.as<BNode>(), lo, hi,
                           d_size.as<uint32_t>(), d_pos.as<uint32_t>());
    }
    hipLaunchKernelGGL(k_emit, dim3((total + 255) / 256), dim3(256), 0, s, d_nodes.as<BNode>(), total,
                       d_pos.as<uint32_t>(), node_base, tri_base, d_out.as<uint4>());
    NOOR_HIP_TRY(hipGetLastError());
    out.nodes.resize(8 * (size_t)total);
    out.perm.resize(n);
    NOOR_HIP_TRY(hipMemcpyAsync(out.nodes.data(), d_out.p, 32 * (size_t)total, hipMemcpyDeviceToHost, s));
    NOOR_HIP_TRY(hipMemcpyAsync(out.perm.data(), d_final.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    NOOR_HIP_TRY(hipStreamSynchronize(s));
    out.device_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    out.height = levels ? levels - 1 : 0;
    out.leaves = out.inners = out.max_leaf = 0;
    for (uint32_t i = 0; i < total; ++i) {
        const uint32_t w = out.nodes[8 * (size_t)i + 7];
        if (w & NOOR_TRI_LEAF_MASK) {
            out.leaves++;
            out.max_leaf = std::max(out.max_leaf, w & NOOR_TRI_COUNT_MASK);
        } else {
            out.inners++;
        }
    }
    return std::string();
}

}  // namespace noor
