// Host-side scene assembly for the MI355X render path: binned-SAH two-level
// BVH in the reference node layout, Wald triangle records, payload
// flattening, camera and spec set-up.  See include/noor_scene.h for the
// reference functions each part replaces.
//
// The builder restates the reference's binned SAH (16 equal-width centroid
// bins per axis, cost Ct + Ci (nL SA_L + nR SA_R) / SA, leaf rules of
// noor/source/bvh.cpp:180-214) in its own code and builds the bottom-level
// trees of large meshes in parallel: the top levels of the recursion fork
// into tasks that emit their subtrees into private node arrays, which are
// then relocated and appended depth-first, so the layout stays "left child =
// index + 1, right child index in .w".
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <functional>
#include <future>
#include <limits>
#include <memory>
#include <new>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/noor_scene.h"
#include "bvh_gpu.h"

namespace {

constexpr float kInf = std::numeric_limits<float>::infinity();

struct Box {
    float mn[3] = {kInf, kInf, kInf};
    float mx[3] = {-kInf, -kInf, -kInf};
    void merge(const Box& b) {
        for (int i = 0; i < 3; ++i) {
            mn[i] = std::min(mn[i], b.mn[i]);
            mx[i] = std::max(mx[i], b.mx[i]);
        }
    }
    void include(const float* p) {
        for (int i = 0; i < 3; ++i) {
            mn[i] = std::min(mn[i], p[i]);
            mx[i] = std::max(mx[i], p[i]);
        }
    }
    bool valid() const {
        for (int i = 0; i < 3; ++i)
            if (!std::isfinite(mn[i]) || !std::isfinite(mx[i])) return false;
        return true;
    }
    // BBox::surfaceArea (noor/header/bbox.h): infinite for an empty box
    float area() const {
        if (!valid()) return kInf;
        const float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        return 2.0f * (dx * dy + dy * dz + dx * dz);
    }
    float centroid(int a) const { return (mn[a] + mx[a]) * 0.5f; }
};

struct Node {
    uint32_t a[4];  // min.xyz, right child / first triangle / BLAS root
    uint32_t b[4];  // max.xyz, flags | count | axis | instance | light
};

inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

Node make_node(const Box& bx, uint32_t w0, uint32_t w1) {
    Node n;
    for (int i = 0; i < 3; ++i) {
        n.a[i] = fbits(bx.mn[i]);
        n.b[i] = fbits(bx.mx[i]);
    }
    n.a[3] = w0;
    n.b[3] = w1;
    return n;
}

// A primitive seen by the builder: its bounds and centroid.
struct Prim {
    Box box;
    float c[3];
};

struct Params {
    uint32_t num_bins, max_height, min_leaf, max_leaf;
    float Ci, Ct;
};

struct SplitResult {
    float cost = kInf;
    int axis = 0;
    uint32_t index = 0;
    uint32_t nl = 0, nr = 0;
    float bound = 0.0f;     // centroid[axis] <= bound goes left
    Box lbox, rbox, lcent, rcent;
    bool valid() const { return cost != kInf && nl != 0 && nr != 0; }
};

// Binned SAH over prims[idx[begin..end)] (noor/source/splitter.cpp:62-132).
SplitResult find_split(const Prim* prims, const uint32_t* idx, size_t begin, size_t end,
                       const Box& box, const Box& cent, const Params& P) {
    const uint32_t nb = P.num_bins;
    SplitResult best;
    const float parent_area = box.area();
    // bin boundaries (Split::reset, noor/header/splitter.h:127-145)
    float lo[3], dt[3];
    std::vector<float> bmin(3 * nb), bmax(3 * nb);
    for (int a = 0; a < 3; ++a) {
        dt[a] = (cent.mx[a] - cent.mn[a]) * (1.0f / (float)nb);
        lo[a] = cent.mn[a];
        bmin[a * nb + 0] = lo[a];
        bmax[a * nb + 0] = lo[a] + dt[a];
        for (uint32_t i = 1; i < nb; ++i) {
            bmin[a * nb + i] = bmin[a * nb + i - 1] + dt[a];
            bmax[a * nb + i] = bmax[a * nb + i - 1] + dt[a];
        }
        bmax[a * nb + nb - 1] = cent.mx[a];
    }
    std::vector<Box> bb(nb), bc(nb), rb(nb), rc(nb);
    std::vector<uint32_t> cnt(nb), rcnt(nb);
    for (int a = 0; a < 3; ++a) {
        std::fill(bb.begin(), bb.end(), Box());
        std::fill(bc.begin(), bc.end(), Box());
        std::fill(cnt.begin(), cnt.end(), 0u);
        const float* mxs = &bmax[a * nb];
        for (size_t k = begin; k < end; ++k) {
            const Prim& p = prims[idx[k]];
            const float c = p.c[a];
            // smallest bin whose upper boundary is >= c
            uint32_t l = 0, r = nb - 1;
            while (l < r) {
                uint32_t m = (l + r) >> 1;
                if (c <= mxs[m]) r = m; else l = m + 1;
            }
            bb[l].merge(p.box);
            bc[l].include(p.c);
            ++cnt[l];
        }
        const uint32_t n = nb - 1;
        rb[n] = bb[n];
        rc[n] = bc[n];
        rcnt[n] = cnt[n];
        for (uint32_t i = n - 1; i > 0; --i) {
            rb[i] = bb[i]; rb[i].merge(rb[i + 1]);
            rc[i] = bc[i]; rc[i].merge(rc[i + 1]);
            rcnt[i] = cnt[i] + rcnt[i + 1];
        }
        Box lb, lc;
        uint32_t lcount = 0;
        for (uint32_t i = 0; i < n; ++i) {
            lb.merge(bb[i]);
            lc.merge(bc[i]);
            lcount += cnt[i];
            const uint32_t rcount = rcnt[i + 1];
            if (lcount == 0 || rcount == 0) continue;
            const float cost = P.Ct + P.Ci * ((float)lcount * lb.area() + (float)rcount * rb[i + 1].area()) /
                                           parent_area;
            if (cost < best.cost) {
                best.cost = cost;
                best.axis = a;
                best.index = i;
                best.nl = lcount;
                best.nr = rcount;
                best.bound = mxs[i];
                best.lbox = lb;
                best.lcent = lc;
                best.rbox = rb[i + 1];
                best.rcent = rc[i + 1];
            }
        }
    }
    return best;
}

struct Builder {
    const Prim* prims;
    uint32_t* idx;
    Params P;
    std::atomic<uint32_t> height{0};
    std::atomic<uint32_t> leaves{0}, inners{0}, max_leaf{0};
    int par_depth = 0;

    void note_leaf(uint32_t level, uint32_t count) {
        uint32_t h = height.load();
        while (h < level && !height.compare_exchange_weak(h, level)) {}
        uint32_t m = max_leaf.load();
        while (m < count && !max_leaf.compare_exchange_weak(m, count)) {}
        leaves++;
    }

    // Bottom level over triangles; appends the subtree to `out`.
    void build_tri(size_t begin, size_t end, const Box& box, const Box& cent, uint32_t level,
                   std::vector<Node>& out, int depth) {
        const uint32_t count = (uint32_t)(end - begin);
        if (level >= P.max_height || count <= P.min_leaf) {
            out.push_back(make_node(box, (uint32_t)begin, NOOR_TRI_LEAF_MASK | count));
            note_leaf(level, count);
            return;
        }
        SplitResult s = find_split(prims, idx, begin, end, box, cent, P);
        const float leaf_cost = P.Ci * (float)count;
        if (!s.valid() || (leaf_cost < s.cost && count <= P.max_leaf)) {
            out.push_back(make_node(box, (uint32_t)begin, NOOR_TRI_LEAF_MASK | count));
            note_leaf(level, count);
            return;
        }
        // stable, so that the device builder (which partitions stably) produces the same order
        uint32_t* mid = std::stable_partition(idx + begin, idx + end, [&](uint32_t i) {
            return prims[i].c[s.axis] <= s.bound;
        });
        const size_t split = (size_t)(mid - idx);
        const uint32_t self = (uint32_t)out.size();
        out.push_back(make_node(box, 0, (uint32_t)s.axis));
        inners++;
        if (depth < par_depth && count > 65536) {
            std::vector<Node> right;
            auto fut = std::async(std::launch::async, [&] {
                build_tri(split, end, s.rbox, s.rcent, level + 1, right, depth + 1);
            });
            build_tri(begin, split, s.lbox, s.lcent, level + 1, out, depth + 1);
            fut.get();
            const uint32_t off = (uint32_t)out.size();
            for (Node& n : right)
                if ((n.b[3] & NOOR_TRI_LEAF_MASK) == 0) n.a[3] += off;
            out.insert(out.end(), right.begin(), right.end());
            out[self].a[3] = off;
        } else {
            build_tri(begin, split, s.lbox, s.lcent, level + 1, out, depth + 1);
            out[self].a[3] = (uint32_t)out.size();
            build_tri(split, end, s.rbox, s.rcent, level + 1, out, depth + 1);
        }
    }
};

struct TlasEntry {
    Box box;
    bool light;
    uint32_t light_idx;
    uint32_t instance;   // index into the caller's instance list
    uint32_t mesh;
};

}  // namespace

struct noor_payload {
    std::vector<noor_uint4> attr, wald0, wald1, wald2;
    std::vector<uint32_t> prim_ids;
    std::vector<noor_float4> vertices, normals;
    std::vector<noor_float2> uvs;
    std::vector<noor_uint4> bvh_a, bvh_b;
    uint32_t root = 0;
    std::vector<noor_float4> w2o[3], o2w[3], nrm[3];
    std::vector<uint32_t> instance_source;  // transform index -> caller instance index
    std::vector<std::vector<float>> tex_data;
    std::vector<noor_texture_view> textures;
    std::vector<noor_material> materials;
    std::vector<noor_area_light> area;
    std::vector<noor_point_light> point;
    std::vector<noor_spot_light> spot;
    std::vector<noor_distant_light> distant;
    noor_build_stats stats{};
};

namespace {

// WaldTriangle (noor/header/triangle.h:56-120), object space.
void wald_record(const float* A, const float* B, const float* C, uint32_t mtype, noor_uint4* r0,
                 noor_uint4* r1, noor_uint4* r2) {
    float b[3], c[3];
    for (int i = 0; i < 3; ++i) {
        b[i] = C[i] - A[i];
        c[i] = B[i] - A[i];
    }
    float N[3] = {c[1] * b[2] - c[2] * b[1], c[2] * b[0] - c[0] * b[2], c[0] * b[1] - c[1] * b[0]};
    const float l2 = N[0] * N[0] + N[1] * N[1] + N[2] * N[2];
    const float inv = 1.0f / std::sqrt(l2);
    for (int i = 0; i < 3; ++i) N[i] *= inv;
    static const uint32_t modulo[4] = {1u, 2u, 0u, 1u};
    uint32_t axis = 0;
    for (uint32_t i = 0; i < 3; ++i) axis = std::fabs(N[i]) > std::fabs(N[axis]) ? i : axis;
    const uint32_t u = modulo[axis], v = modulo[axis + 1];
    const float denom = b[u] * c[v] - b[v] * c[u];
    const float Nk = N[axis];
    if (denom == 0.0f || !std::isfinite(denom) || !(l2 > 0.0f)) {
        *r0 = {axis, 0, 0, 0};
        *r1 = {0, 0, 0, 0};
        *r2 = {0, 0, 1u, mtype};
        return;
    }
    const float nu = N[u] / Nk, nv = N[v] / Nk;
    const float nd = (N[0] * A[0] + N[1] * A[1] + N[2] * A[2]) / Nk;
    *r0 = {axis, fbits(nu), fbits(nv), fbits(nd)};
    *r1 = {fbits(b[u] / denom), fbits(-b[v] / denom), fbits(A[u]), fbits(A[v])};
    *r2 = {fbits(c[v] / denom), fbits(-c[u] / denom), 0u, mtype};
}

struct M34 {
    double m[3][4];
};

bool invert34(const M34& a, M34* out) {
    const double(*m)[4] = a.m;
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    if (det == 0.0 || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    double r[3][3];
    r[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
    r[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    r[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    r[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
    r[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    r[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    r[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
    r[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    r[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) out->m[i][j] = r[i][j];
        out->m[i][3] = -(r[i][0] * m[0][3] + r[i][1] * m[1][3] + r[i][2] * m[2][3]);
    }
    return true;
}

// Stack entries needed by the push-both-children traversal in the worst case.
uint32_t stack_need(const std::vector<Node>& nodes, uint32_t ni, std::vector<int32_t>& memo) {
    if (memo[ni] >= 0) return (uint32_t)memo[ni];
    const Node& n = nodes[ni];
    const uint32_t w = n.b[3];
    uint32_t need = 0;
    if (w & NOOR_LIGHT_NODE_MASK) {
        need = 0;
    } else if (w & NOOR_MESH_INSTANCE_MASK) {
        need = std::max(1u, stack_need(nodes, n.a[3], memo));
    } else if (w & NOOR_MESH_LEAF_MASK) {
        const uint32_t k = w & NOOR_COUNT_MASK;
        need = k;
        for (uint32_t j = 0; j < k; ++j)
            need = std::max(need, (k - 1 - j) + stack_need(nodes, ni + 1 + j, memo));
    } else if (w & NOOR_TRI_LEAF_MASK) {
        need = 0;
    } else {
        const uint32_t l = stack_need(nodes, ni + 1, memo), r = stack_need(nodes, n.a[3], memo);
        need = std::max(2u, 1u + std::max(l, r));
    }
    memo[ni] = (int32_t)need;
    return need;
}

}  // namespace

namespace {
noor_status build_impl(const noor_scene_desc* d, const noor_bvh_params* bp, int device, noor_payload** out);
}  // namespace

extern "C" noor_status noor_scene_build(const noor_scene_desc* d, const noor_bvh_params* bp,
                                        noor_payload** out) {
    return build_impl(d, bp, -1, out);
}

extern "C" noor_status noor_scene_build_device(const noor_scene_desc* d, const noor_bvh_params* bp, int device,
                                               noor_payload** out) {
    if (device < 0) return NOOR_ERR_INVALID_ARG;
    return build_impl(d, bp, device, out);
}

namespace {
// Meshes of at least this many triangles are built on the device.
constexpr uint32_t kDeviceMinTris = 16384;

noor_status build_impl(const noor_scene_desc* d, const noor_bvh_params* bp, int device, noor_payload** out) {
    if (!d || !out) return NOOR_ERR_INVALID_ARG;
    *out = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    Params P{16, 32, 2, 2, 1.0f, 0.125f};
    int threads = 0;
    if (bp) {
        P = {bp->num_bins, bp->max_height, bp->min_leaf_tris, bp->max_leaf_tris, bp->Ci, bp->Ct};
        threads = bp->threads;
    }
    if (P.num_bins < 2 || P.num_bins > 256) return NOOR_ERR_INVALID_ARG;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());

    std::unique_ptr<noor_payload> p(new (std::nothrow) noor_payload());
    if (!p) return NOOR_ERR_OUT_OF_MEMORY;

    // ---- validate -----------------------------------------------------------
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        const noor_mesh_desc& md = d->meshes[m];
        if (!md.positions || !md.indices) return NOOR_ERR_INVALID_ARG;
        for (size_t i = 0; i < (size_t)md.num_triangles * 3; ++i)
            if (md.indices[i] >= md.num_vertices) return NOOR_ERR_INVALID_PAYLOAD;
        for (size_t i = 0; i < (size_t)md.num_vertices * 3; ++i)
            if (!std::isfinite(md.positions[i])) return NOOR_ERR_INVALID_PAYLOAD;
        for (uint32_t t = 0; t < md.num_triangles; ++t) {
            const uint32_t mi = md.material ? md.material[t] : md.material_all;
            if (mi >= d->num_materials) return NOOR_ERR_INVALID_PAYLOAD;
        }
    }
    for (uint32_t i = 0; i < d->num_instances; ++i)
        if (d->instances[i].mesh >= d->num_meshes) return NOOR_ERR_INVALID_PAYLOAD;

    // ---- vertices and triangles (object space) --------------------------------
    std::vector<uint32_t> voff(d->num_meshes + 1, 0), toff(d->num_meshes + 1, 0);
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        voff[m + 1] = voff[m] + d->meshes[m].num_vertices;
        toff[m + 1] = toff[m] + d->meshes[m].num_triangles;
    }
    const uint32_t nv = voff[d->num_meshes], nt = toff[d->num_meshes];
    p->vertices.resize(nv);
    p->normals.resize(nv);
    p->uvs.resize(nv);
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        const noor_mesh_desc& md = d->meshes[m];
        std::vector<double> acc;
        if (!md.normals) acc.assign((size_t)md.num_vertices * 3, 0.0);
        if (!md.normals) {
            for (uint32_t t = 0; t < md.num_triangles; ++t) {
                const uint32_t* I = &md.indices[3 * t];
                const float* a = &md.positions[3 * I[0]];
                const float* b = &md.positions[3 * I[1]];
                const float* c = &md.positions[3 * I[2]];
                const double e1[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
                const double e2[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
                const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                     e1[0] * e2[1] - e1[1] * e2[0]};
                for (int k = 0; k < 3; ++k)
                    for (int j = 0; j < 3; ++j) acc[3 * (size_t)I[k] + j] += n[j];
            }
        }
        for (uint32_t v = 0; v < md.num_vertices; ++v) {
            const float* q = &md.positions[3 * (size_t)v];
            p->vertices[voff[m] + v] = {q[0], q[1], q[2], 1.0f};
            if (md.normals) {
                const float* n = &md.normals[3 * (size_t)v];
                p->normals[voff[m] + v] = {n[0], n[1], n[2], 0.0f};
            } else {
                const double* n = &acc[3 * (size_t)v];
                double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                if (l == 0.0) l = 1.0;
                p->normals[voff[m] + v] = {(float)(n[0] / l), (float)(n[1] / l), (float)(n[2] / l), 0.0f};
            }
            if (md.uvs) p->uvs[voff[m] + v] = {md.uvs[2 * (size_t)v], md.uvs[2 * (size_t)v + 1]};
            else p->uvs[voff[m] + v] = {0.0f, 0.0f};
        }
    }

    // ---- per-triangle bounds ------------------------------------------------------
    std::vector<Prim> prims(nt);
    std::vector<uint32_t> tri_mesh(nt);
    {
        std::vector<std::thread> ths;
        const int T = std::min<int>(threads, std::max<int>(1, (int)(nt / 65536) + 1));
        for (int th = 0; th < T; ++th) {
            ths.emplace_back([&, th] {
                for (uint32_t m = 0; m < d->num_meshes; ++m) {
                    const noor_mesh_desc& md = d->meshes[m];
                    // a contiguous share per thread (no two threads write one cache line)
                    const uint32_t t0 = (uint32_t)((uint64_t)md.num_triangles * th / T);
                    const uint32_t t1 = (uint32_t)((uint64_t)md.num_triangles * (th + 1) / T);
                    for (uint32_t t = t0; t < t1; ++t) {
                        Prim& pr = prims[toff[m] + t];
                        pr.box = Box();
                        for (int k = 0; k < 3; ++k) pr.box.include(&md.positions[3 * (size_t)md.indices[3 * (size_t)t + k]]);
                        // -0 becomes +0 (x + 0 = +0 for x = -0): minima and maxima then do
                        // not depend on the order they are taken in
                        for (int a = 0; a < 3; ++a) {
                            pr.box.mn[a] += 0.0f;
                            pr.box.mx[a] += 0.0f;
                            pr.c[a] = pr.box.centroid(a) + 0.0f;
                        }
                        tri_mesh[toff[m] + t] = m;
                    }
                }
            });
        }
        for (auto& t : ths) t.join();
    }

    // ---- bottom levels --------------------------------------------------------------
    std::vector<uint32_t> perm(nt);
    std::iota(perm.begin(), perm.end(), 0u);
    std::vector<Node> nodes;
    nodes.reserve((size_t)nt * 2 + 1024);
    std::vector<uint32_t> blas_root(d->num_meshes, 0);
    std::vector<Box> mesh_box(d->num_meshes);
    double device_ms = 0.0;
    Builder B;
    B.prims = prims.data();
    B.idx = perm.data();
    B.P = P;
    B.par_depth = 0;
    while ((1 << B.par_depth) < threads * 2 && B.par_depth < 8) ++B.par_depth;
    for (uint32_t m = 0; m < d->num_meshes; ++m) {
        Box box, cent;
        for (uint32_t t = toff[m]; t < toff[m + 1]; ++t) {
            box.merge(prims[t].box);
            cent.include(prims[t].c);
        }
        mesh_box[m] = box;
        blas_root[m] = (uint32_t)nodes.size();
        if (toff[m + 1] == toff[m]) {
            // empty mesh: a leaf with no triangles
            nodes.push_back(make_node(box, toff[m], NOOR_TRI_LEAF_MASK | 0u));
            continue;
        }
        const uint32_t mt = toff[m + 1] - toff[m];
        if (device >= 0 && mt >= kDeviceMinTris) {
            // the mesh's primitives in input order (perm is still the identity here), as the
            // device builder's record: bounds then centroid, 9 floats
            static_assert(sizeof(Prim) == sizeof(noor::GpuPrim) && offsetof(Prim, c) == offsetof(noor::GpuPrim, c),
                          "Prim and GpuPrim share a layout");
            const noor::GpuPrim* gp = reinterpret_cast<const noor::GpuPrim*>(&prims[toff[m]]);
            const float rb[6] = {box.mn[0], box.mn[1], box.mn[2], box.mx[0], box.mx[1], box.mx[2]};
            const float rc[6] = {cent.mn[0], cent.mn[1], cent.mn[2], cent.mx[0], cent.mx[1], cent.mx[2]};
            const noor::GpuBvhParams gpar{P.num_bins, P.max_height, P.min_leaf, P.max_leaf, P.Ci, P.Ct};
            noor::GpuBvhResult res;
            const std::string err =
                noor::gpu_build_blas(device, gp, mt, rb, rc, gpar, (uint32_t)nodes.size(), toff[m], res);
            if (!err.empty()) {
                std::fprintf(stderr, "noor_scene_build_device: %s\n", err.c_str());
                return err.find("bins") != std::string::npos ? NOOR_ERR_UNSUPPORTED : NOOR_ERR_HIP;
            }
            for (uint32_t t = 0; t < mt; ++t) perm[toff[m] + t] = toff[m] + res.perm[t];
            const size_t nn = res.nodes.size() / 8, base = nodes.size();
            nodes.resize(base + nn);
            std::memcpy(&nodes[base], res.nodes.data(), nn * sizeof(Node));
            uint32_t h = B.height.load();
            while (h < res.height && !B.height.compare_exchange_weak(h, res.height)) {}
            uint32_t ml = B.max_leaf.load();
            while (ml < res.max_leaf && !B.max_leaf.compare_exchange_weak(ml, res.max_leaf)) {}
            B.leaves += res.leaves;
            B.inners += res.inners;
            device_ms += res.device_ms;
            continue;
        }
        B.build_tri(toff[m], toff[m + 1], box, cent, 0, nodes, 0);
    }

    // ---- top level over instances and area lights ---------------------------------------
    std::vector<noor_instance_desc> inst;
    if (d->num_instances == 0) {
        for (uint32_t m = 0; m < d->num_meshes; ++m) {
            noor_instance_desc id{};
            id.mesh = m;
            id.object_to_world[0] = id.object_to_world[5] = id.object_to_world[10] = 1.0f;
            inst.push_back(id);
        }
    } else {
        inst.assign(d->instances, d->instances + d->num_instances);
    }
    std::vector<TlasEntry> ent;
    for (uint32_t i = 0; i < inst.size(); ++i) {
        const float* M = inst[i].object_to_world;
        const Box& ob = mesh_box[inst[i].mesh];
        Box wb;
        if (ob.valid()) {
            for (int corner = 0; corner < 8; ++corner) {
                const double q[3] = {(corner & 1) ? ob.mx[0] : ob.mn[0], (corner & 2) ? ob.mx[1] : ob.mn[1],
                                     (corner & 4) ? ob.mx[2] : ob.mn[2]};
                for (int r = 0; r < 3; ++r) {
                    const double v = M[4 * r + 0] * q[0] + M[4 * r + 1] * q[1] + M[4 * r + 2] * q[2] + M[4 * r + 3];
                    // round outward by a few ulps so transformed triangles stay inside
                    const double pad = std::fabs(v) * 4e-7 + 1e-30;
                    wb.mn[r] = std::min(wb.mn[r], (float)(v - pad));
                    wb.mx[r] = std::max(wb.mx[r], (float)(v + pad));
                }
            }
        }
        ent.push_back({wb, false, 0u, i, inst[i].mesh});
    }
    for (uint32_t l = 0; l < d->num_area_lights; ++l) {
        const noor_shape& s = d->area_lights[l].shape;
        Box b;
        const float c[3] = {s.center.x, s.center.y, s.center.z};
        if (s.type == NOOR_SHAPE_QUAD) {
            const float u[3] = {s.u.x, s.u.y, s.u.z}, v[3] = {s.v.x, s.v.y, s.v.z};
            for (int k = 0; k < 4; ++k) {
                float q[3];
                for (int a = 0; a < 3; ++a) q[a] = c[a] + ((k & 1) ? u[a] : 0.0f) + ((k & 2) ? v[a] : 0.0f);
                b.include(q);
            }
        } else {
            for (int a = 0; a < 3; ++a) {
                b.mn[a] = c[a] - s.radius;
                b.mx[a] = c[a] + s.radius;
            }
        }
        ent.push_back({b, true, l, 0u, 0u});
    }
    std::vector<Prim> tprims(ent.size());
    for (size_t i = 0; i < ent.size(); ++i) {
        tprims[i].box = ent[i].box;
        for (int a = 0; a < 3; ++a) tprims[i].c[a] = ent[i].box.centroid(a);
    }
    std::vector<uint32_t> tperm(ent.size());
    std::iota(tperm.begin(), tperm.end(), 0u);
    p->root = (uint32_t)nodes.size();
    std::vector<uint32_t> transform_order;  // TLAS order of instance entries
    const uint32_t tl_base_height = B.height.load();
    uint32_t tl_height = tl_base_height;
    // recursive top-level build (noor/source/bvh.cpp:180-214, MESHBIN rules)
    std::function<void(size_t, size_t, const Box&, const Box&, uint32_t)> build_top =
        [&](size_t begin, size_t end, const Box& box, const Box& cent, uint32_t level) {
            const uint32_t count = (uint32_t)(end - begin);
            auto leaf = [&]() {
                nodes.push_back(make_node(box, (uint32_t)nodes.size() + 1, NOOR_MESH_LEAF_MASK | count));
                for (size_t k = begin; k < end; ++k) {
                    const TlasEntry& e = ent[tperm[k]];
                    if (e.light) {
                        nodes.push_back(make_node(e.box, 0u, NOOR_LIGHT_NODE_MASK | e.light_idx));
                    } else {
                        const uint32_t ti = (uint32_t)transform_order.size();
                        transform_order.push_back(e.instance);
                        nodes.push_back(make_node(e.box, blas_root[e.mesh], NOOR_MESH_INSTANCE_MASK | ti));
                    }
                }
                tl_height = std::max(tl_height, level);
            };
            if (count < 2) { leaf(); return; }
            SplitResult s = find_split(tprims.data(), tperm.data(), begin, end, box, cent, P);
            if (!s.valid()) { leaf(); return; }
            uint32_t* mid = std::partition(tperm.data() + begin, tperm.data() + end,
                                           [&](uint32_t i) { return tprims[i].c[s.axis] <= s.bound; });
            const size_t split = (size_t)(mid - tperm.data());
            const uint32_t self = (uint32_t)nodes.size();
            nodes.push_back(make_node(box, 0u, NOOR_MESH_INNER_MASK | (uint32_t)s.axis));
            build_top(begin, split, s.lbox, s.lcent, level + 1);
            nodes[self].a[3] = (uint32_t)nodes.size();
            build_top(split, end, s.rbox, s.rcent, level + 1);
        };
    Box scene_box, scene_cent;
    for (auto& t : tprims) {
        scene_box.merge(t.box);
        scene_cent.include(t.c);
    }
    if (ent.empty()) return NOOR_ERR_INVALID_PAYLOAD;
    build_top(0, ent.size(), scene_box, scene_cent, tl_base_height);

    // ---- payload arrays ------------------------------------------------------------------------
    p->bvh_a.resize(nodes.size());
    p->bvh_b.resize(nodes.size());
    static_assert(sizeof(noor_uint4) == 16 && sizeof(Node) == 32, "node words");
    for (size_t i = 0; i < nodes.size(); ++i) {
        std::memcpy(&p->bvh_a[i], nodes[i].a, 16);
        std::memcpy(&p->bvh_b[i], nodes[i].b, 16);
    }
    p->attr.resize(nt);
    p->wald0.resize(nt);
    p->wald1.resize(nt);
    p->wald2.resize(nt);
    p->prim_ids.resize(nt);
    {
        std::vector<std::thread> ths;
        const int T = std::min<int>(threads, std::max<int>(1, (int)(nt / 65536) + 1));
        for (int th = 0; th < T; ++th) {
            ths.emplace_back([&, th] {
                const uint32_t p0 = (uint32_t)((uint64_t)nt * th / T), p1 = (uint32_t)((uint64_t)nt * (th + 1) / T);
                for (uint32_t pos = p0; pos < p1; ++pos) {
                    const uint32_t t = perm[pos];
                    const uint32_t m = tri_mesh[t];
                    const noor_mesh_desc& md = d->meshes[m];
                    const uint32_t lt = t - toff[m];
                    const uint32_t* I = &md.indices[3 * (size_t)lt];
                    const uint32_t mi = md.material ? md.material[lt] : md.material_all;
                    const uint32_t mtype = d->materials[mi].type;
                    p->attr[pos] = {I[0] + voff[m], I[1] + voff[m], I[2] + voff[m], mi};
                    wald_record(&md.positions[3 * (size_t)I[0]], &md.positions[3 * (size_t)I[1]],
                                &md.positions[3 * (size_t)I[2]], mtype, &p->wald0[pos], &p->wald1[pos],
                                &p->wald2[pos]);
                    p->prim_ids[pos] = t;
                }
            });
        }
        for (auto& t : ths) t.join();
    }
    // transforms in top-level order (noor/source/bvh.cpp:53-86)
    for (int r = 0; r < 3; ++r) {
        p->w2o[r].resize(transform_order.size());
        p->o2w[r].resize(transform_order.size());
        p->nrm[r].resize(transform_order.size());
    }
    p->instance_source = transform_order;
    for (size_t k = 0; k < transform_order.size(); ++k) {
        const float* M = inst[transform_order[k]].object_to_world;
        M34 a, ia;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) a.m[r][c] = M[4 * r + c];
        if (!invert34(a, &ia)) return NOOR_ERR_INVALID_PAYLOAD;
        for (int r = 0; r < 3; ++r) {
            p->o2w[r][k] = {M[4 * r + 0], M[4 * r + 1], M[4 * r + 2], M[4 * r + 3]};
            p->w2o[r][k] = {(float)ia.m[r][0], (float)ia.m[r][1], (float)ia.m[r][2], (float)ia.m[r][3]};
            // normal transform = transpose(inverse(M3x3))
            p->nrm[r][k] = {(float)ia.m[0][r], (float)ia.m[1][r], (float)ia.m[2][r], 0.0f};
        }
    }
    // textures (copied), materials, lights
    p->tex_data.resize(d->num_textures);
    p->textures.resize(d->num_textures);
    for (uint32_t i = 0; i < d->num_textures; ++i) {
        const noor_texture_view& tv = d->textures[i];
        if (!tv.rgba || tv.width == 0 || tv.height == 0) return NOOR_ERR_INVALID_PAYLOAD;
        p->tex_data[i].assign(tv.rgba, tv.rgba + (size_t)tv.width * tv.height * 4);
        p->textures[i] = {tv.width, tv.height, p->tex_data[i].data()};
    }
    p->materials.assign(d->materials, d->materials + d->num_materials);
    if (d->num_area_lights) p->area.assign(d->area_lights, d->area_lights + d->num_area_lights);
    if (d->num_point_lights) p->point.assign(d->point_lights, d->point_lights + d->num_point_lights);
    if (d->num_spot_lights) p->spot.assign(d->spot_lights, d->spot_lights + d->num_spot_lights);
    if (d->num_distant_lights) p->distant.assign(d->distant_lights, d->distant_lights + d->num_distant_lights);
    for (uint32_t l = 0; l < p->area.size(); ++l) p->area[l].shape.light_idx = (int32_t)l;

    // ---- stats ------------------------------------------------------------------------------------
    noor_build_stats& st = p->stats;
    st.height = tl_height + 1;
    std::vector<int32_t> memo(nodes.size(), -1);
    st.stack_depth = std::max(1u, stack_need(nodes, p->root, memo));
    st.num_nodes = (uint32_t)nodes.size();
    st.num_inner = B.inners.load();
    st.num_leaves = B.leaves.load();
    st.max_leaf_tris = B.max_leaf.load();
    for (int a = 0; a < 3; ++a) {
        st.scene_min[a] = scene_box.mn[a];
        st.scene_max[a] = scene_box.mx[a];
    }
    const float dx = scene_box.mx[0] - scene_box.mn[0], dy = scene_box.mx[1] - scene_box.mn[1],
                dz = scene_box.mx[2] - scene_box.mn[2];
    st.scene_radius = 0.5f * std::sqrt(dx * dx + dy * dy + dz * dz);
    st.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    (void)device_ms;
    *out = p.release();
    return NOOR_OK;
}
}  // namespace

extern "C" noor_status noor_payload_view_get(const noor_payload* p, noor_payload_view* v) {
    if (!p || !v) return NOOR_ERR_INVALID_ARG;
    std::memset(v, 0, sizeof(*v));
    v->num_triangles = (uint32_t)p->attr.size();
    v->attribute_indices = p->attr.data();
    v->wald_ax_nu_nv_nd = p->wald0.data();
    v->wald_bnu_bnv_au_av = p->wald1.data();
    v->wald_cnu_cnv_deg_mat = p->wald2.data();
    v->prim_ids = p->prim_ids.data();
    v->num_vertices = (uint32_t)p->vertices.size();
    v->vertices = p->vertices.data();
    v->normals = p->normals.data();
    v->uvs = p->uvs.data();
    v->bvh_root_node = p->root;
    v->num_bvh_nodes = (uint32_t)p->bvh_a.size();
    v->bvh_min_right_start = p->bvh_a.data();
    v->bvh_max_axis_count = p->bvh_b.data();
    v->num_instances = (uint32_t)p->w2o[0].size();
    for (int r = 0; r < 3; ++r) {
        v->world_to_object_row[r] = p->w2o[r].data();
        v->object_to_world_row[r] = p->o2w[r].data();
        v->normal_row[r] = p->nrm[r].data();
    }
    v->num_textures = (uint32_t)p->textures.size();
    v->textures = p->textures.data();
    v->num_materials = (uint32_t)p->materials.size();
    v->materials = p->materials.data();
    v->num_area_lights = (uint32_t)p->area.size();
    v->area_lights = p->area.data();
    v->num_point_lights = (uint32_t)p->point.size();
    v->point_lights = p->point.data();
    v->num_spot_lights = (uint32_t)p->spot.size();
    v->spot_lights = p->spot.data();
    v->num_distant_lights = (uint32_t)p->distant.size();
    v->distant_lights = p->distant.data();
    v->num_infinite_lights = 1;  // Model::loadCudaPayload always adds one (noor/source/model.cpp:450-453)
    return NOOR_OK;
}

extern "C" noor_status noor_payload_stats(const noor_payload* p, noor_build_stats* s) {
    if (!p || !s) return NOOR_ERR_INVALID_ARG;
    *s = p->stats;
    return NOOR_OK;
}

extern "C" void noor_payload_free(noor_payload* p) { delete p; }

namespace {
struct M4 {
    double m[4][4];  // row-major
};
M4 mul(const M4& a, const M4& b) {
    M4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a.m[i][k] * b.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}
M4 ident() {
    M4 r{};
    for (int i = 0; i < 4; ++i) r.m[i][i] = 1.0;
    return r;
}
bool inv4(const M4& a, M4* out) {
    // Gauss-Jordan with partial pivoting
    double m[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) m[i][j] = j < 4 ? a.m[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
        if (m[piv][c] == 0.0) return false;
        for (int j = 0; j < 8; ++j) std::swap(m[c][j], m[piv][j]);
        const double d = m[c][c];
        for (int j = 0; j < 8; ++j) m[c][j] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                const double f = m[r][c];
                for (int j = 0; j < 8; ++j) m[r][j] -= f * m[c][j];
            }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out->m[i][j] = m[i][j + 4];
    return true;
}
noor_transform to_xf(const M4& m) {
    noor_transform t;
    for (int r = 0; r < 3; ++r) {
        t.row[r] = {(float)m.m[r][0], (float)m.m[r][1], (float)m.m[r][2]};
    }
    t.t = {(float)m.m[0][3], (float)m.m[1][3], (float)m.m[2][3]};
    return t;
}
}  // namespace

// Camera (noor/header/camera.h:99-334): look-at frame, projection, raster grid.
extern "C" noor_status noor_camera_look_at(const float eye[3], const float look_at[3], const float up[3],
                                           float fov, float ortho_zoom, float lens_radius, uint32_t w,
                                           uint32_t h, int32_t type, noor_camera* out) {
    if (!eye || !look_at || !up || !out || w == 0 || h == 0) return NOOR_ERR_INVALID_ARG;
    if (type != NOOR_CAMERA_PERSP && type != NOOR_CAMERA_ORTHO && type != NOOR_CAMERA_ENV)
        return NOOR_ERR_INVALID_ARG;
    double f[3] = {(double)eye[0] - look_at[0], (double)eye[1] - look_at[1], (double)eye[2] - look_at[2]};
    const double scale = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (scale == 0.0) return NOOR_ERR_INVALID_ARG;
    for (double& x : f) x /= scale;
    double r[3] = {up[1] * f[2] - up[2] * f[1], up[2] * f[0] - up[0] * f[2], up[0] * f[1] - up[1] * f[0]};
    const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (rl == 0.0) return NOOR_ERR_INVALID_ARG;
    for (double& x : r) x /= rl;
    // cameraToWorld = translate(lookAt) * R * translate(0, 0, scale), R columns (right, up, forward)
    M4 R = ident();
    for (int i = 0; i < 3; ++i) {
        R.m[i][0] = r[i];
        R.m[i][1] = up[i];
        R.m[i][2] = f[i];
    }
    M4 T1 = ident(), T2 = ident();
    for (int i = 0; i < 3; ++i) T1.m[i][3] = look_at[i];
    T2.m[2][3] = scale;
    const M4 c2w = mul(mul(T1, R), T2);
    // screen window (updateProjection)
    const double aspect = (double)w / (double)h;
    double smin[2], smax[2];
    if (aspect > 1.0) {
        smin[0] = -aspect; smax[0] = aspect; smin[1] = -1.0; smax[1] = 1.0;
    } else {
        smin[0] = -1.0; smax[0] = 1.0; smin[1] = -1.0 / aspect; smax[1] = 1.0 / aspect;
    }
    M4 c2s = ident();
    if (type == NOOR_CAMERA_ORTHO) {
        for (int i = 0; i < 2; ++i) { smin[i] *= ortho_zoom; smax[i] *= ortho_zoom; }
        // orthographic(0, 1) = scale(1, 1, 1/(f-n)) * translate(0, 0, -n)
    } else {
        const double n = 0.01, fa = 1000.0;
        const double cot = 1.0 / std::tan((double)fov / 2.0);
        c2s = M4{};
        c2s.m[0][0] = cot;
        c2s.m[1][1] = cot;
        c2s.m[2][2] = (fa + n) / (n - fa);
        c2s.m[2][3] = -2.0 * fa * n / (fa - n);
        c2s.m[3][2] = -1.0;
    }
    // screenToRaster = scale(w, h, 1) * scale(1/(max-min)) * translate(-min)
    M4 s2r = ident();
    s2r.m[0][0] = (double)w / (smax[0] - smin[0]);
    s2r.m[1][1] = (double)h / (smax[1] - smin[1]);
    s2r.m[0][3] = -smin[0] * (double)w / (smax[0] - smin[0]);
    s2r.m[1][3] = -smin[1] * (double)h / (smax[1] - smin[1]);
    M4 ic2s, is2r, iw;
    if (!inv4(c2s, &ic2s) || !inv4(s2r, &is2r)) return NOOR_ERR_INVALID_ARG;
    const M4 r2c = mul(ic2s, is2r);
    M4 ic2w, ir2c;
    if (!inv4(c2w, &ic2w) || !inv4(r2c, &ir2c)) return NOOR_ERR_INVALID_ARG;
    const M4 w2r = mul(ir2c, ic2w);
    (void)iw;
    std::memset(out, 0, sizeof(*out));
    out->camera_to_world = to_xf(c2w);
    out->raster_to_camera = to_xf(r2c);
    out->world_to_raster = to_xf(w2r);
    out->w = w;
    out->h = h;
    out->lens_radius = lens_radius;
    out->focal_length = (float)scale;
    out->type = type;
    out->center = w * (h >> 1) + (w >> 1);
    // CudaCamera::update (noor/cuda/camera.cuh:66-94)
    out->dx_camera = {(float)r2c.m[0][0], (float)r2c.m[1][0], (float)r2c.m[2][0]};
    out->dy_camera = {(float)r2c.m[0][1], (float)r2c.m[1][1], (float)r2c.m[2][1]};
    const double ip0x = r2c.m[0][3], ip0y = r2c.m[1][3];
    const double ip1x = r2c.m[0][0] * w + r2c.m[0][1] * h + r2c.m[0][3];
    const double ip1y = r2c.m[1][0] * w + r2c.m[1][1] * h + r2c.m[1][3];
    out->image_area = (float)std::fabs((ip1x - ip0x) * (ip1y - ip0y));
    return NOOR_OK;
}

// Scene::load spec fill (noor/source/scene.cpp:51-91).
extern "C" noor_status noor_spec_fill(const noor_payload* p, uint16_t bounces, uint16_t rr,
                                      int32_t skydome_type, uint8_t mis, noor_spec* out) {
    if (!p || !out) return NOOR_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    const float radius = p->stats.scene_radius;
    const float bias = 0.0001f * radius;
    out->white = {1.0f, 1.0f, 1.0f};
    out->black = {0.0f, 0.0f, 0.0f};
    out->reflection_bias = bias;
    out->shadow_bias = 0.0001f * radius;
    out->world_radius = radius;
    out->wr2 = radius * radius;
    out->bvh_root_node = p->root;
    out->bounces = bounces;
    out->rr = rr;
    out->bvh_height = (uint16_t)p->stats.height;
    out->num_gpus = 1;
    out->skydome_type = skydome_type;
    out->mis = mis;
    out->outofsync = 1;
    const uint32_t num_lights = (uint32_t)(p->area.size() + p->point.size() + p->spot.size() + p->distant.size()) + 1u;
    if (num_lights < 2) out->lighting_type |= NOOR_LIGHT_SKY_MASK;
    return NOOR_OK;
}
