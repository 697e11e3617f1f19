// Host construction of the four-wide BVH from the payload's binary BVH.
//
// Binary node kinds follow the reference traversal (noor/cuda/bvh.cuh:142-191):
// light node, mesh instance (BLAS root in .w of the first word), mesh leaf
// (its instances follow at +1..+count), triangle leaf (start in .w, count in
// the flags), else an inner node (left child at +1, right child in .w; MESH_INNER
// marks the TLAS).  A wide node is formed by opening, largest surface area
// first, inner children of the same level until four children are collected.
// Boxes are copied bit for bit from the binary nodes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <unordered_map>

#include "wide_bvh.h"

namespace noor {
namespace {

// The decode the GPU evaluates (box4_hit in dev_core.h): with contraction off
// both compute the product exactly (q has 8 bits, s is a power of two) and
// round the sum once.
float q_decode(float org, uint32_t q, float s) {
    const float qs = (float)q * s;
    return org + qs;
}

// The grid value org + q s as a real number, which the GPU's fast slab test
// (box4_hit) measures distances from.  Exact in binary64: a non-empty box is at
// least one binary32 ulp of its planes wide, so s >= ulp(org) 2^-9 and the sum
// spans fewer than 53 bits; an empty extent has q s = 0.
double q_exact(float org, uint32_t q, float s) { return (double)org + (double)q * (double)s; }

// Encodes one axis of one node: the smallest power-of-two step for which every
// slot's rounded-out planes fit in 8 bits.  Returns false if none does.
bool quantize_axis(const WNode& n, int a, uint32_t valid, float& org, uint32_t& ebyte, uint32_t& qlo, uint32_t& qhi) {
    float lo = std::numeric_limits<float>::infinity(), hi = -std::numeric_limits<float>::infinity();
    for (int c = 0; c < 4; ++c)
        if (valid & (1u << c)) {
            lo = std::min(lo, n.lo[a][c]);
            hi = std::max(hi, n.hi[a][c]);
        }
    org = valid ? lo : 0.0f;
    if (!std::isfinite(org) || (valid && !std::isfinite(hi))) return false;
    const double ext = valid ? (double)hi - (double)lo : 0.0;
    int k = -126;  // smallest normal step
    if (ext > 0.0) k = std::max(k, (int)std::ceil(std::log2(ext / 255.0)));
    for (; k <= 127; ++k) {
        const float s = std::ldexp(1.0f, k);
        uint32_t L = 0, H = 0;
        bool ok = true;
        for (int c = 0; c < 4 && ok; ++c) {
            if (!(valid & (1u << c))) continue;
            const float cl = n.lo[a][c], ch = n.hi[a][c];
            double ql = std::floor(((double)cl - (double)org) / (double)s);
            int64_t q = (int64_t)std::max(0.0, std::min(255.0, ql));
            // both decodes (the binary32 one and the exact grid value) below the plane
            auto above = [&](int64_t v) {
                return q_decode(org, (uint32_t)v, s) > cl || q_exact(org, (uint32_t)v, s) > (double)cl;
            };
            auto below = [&](int64_t v) {
                return q_decode(org, (uint32_t)v, s) < ch || q_exact(org, (uint32_t)v, s) < (double)ch;
            };
            while (q > 0 && above(q)) --q;
            if (above(q)) ok = false;
            int64_t r = (int64_t)std::max(0.0, std::ceil(((double)ch - (double)org) / (double)s));
            if (r > 255) {
                ok = false;
                break;
            }
            while (r <= 255 && below(r)) ++r;
            if (r > 255 || r < q) {
                ok = false;
                break;
            }
            L |= (uint32_t)q << (8 * c);
            H |= (uint32_t)r << (8 * c);
        }
        if (ok) {
            ebyte = (uint32_t)(k + 127);
            qlo = L;
            qhi = H;
            return true;
        }
    }
    return false;
}

std::string quantize(const WNode& n, QNode& q) {
    std::memset(&q, 0, sizeof(q));
    uint32_t valid = 0;
    for (int c = 0; c < 4; ++c) {
        q.ref[c] = n.ref[c];
        if (n.ref[c] != REF_EMPTY) valid |= 1u << c;
    }
    q.meta = valid << 24;
    for (int a = 0; a < 3; ++a) {
        uint32_t e = 0;
        if (!quantize_axis(n, a, valid, q.org[a], e, q.qlo[a], q.qhi[a]))
            return "a node box cannot be quantised (non-finite or too large a range)";
        q.meta |= e << (8 * a);
    }
    return std::string();
}

enum BinKind { B_LIGHT, B_INST, B_MLEAF, B_TLEAF, B_INNER };

// Bottom-level subtrees of up to this many triangles are stored as one leaf
// (0: every binary leaf stays a leaf).  Measured on the 5M-triangle terrain:
// 3, 4 and 8 cut the wide nodes by up to 70% but make the traversal 2%, 7%
// and 17% slower (a watertight triangle test costs more than a box test), so
// it is off.
constexpr uint32_t kLeafMerge = 0;

struct Child {
    float lo[3], hi[3];
    uint32_t ref;
};

class Builder {
  public:
    Builder(const noor_payload_view* P, WideBvh& W) : P_(P), W_(W) {}

    std::string run() {
        W_.nodes.clear();
        W_.inst_root.assign(P_->num_instances, REF_EMPTY);
        const Child root = make_child(P_->bvh_root_node);
        if (!err_.empty()) return err_;
        W_.root = root.ref;
        if (W_.nodes.size() >= REF_INDEX_MASK) return "too many wide BVH nodes";
        need_.assign(W_.nodes.size(), -1);
        W_.stack_need = std::max<uint32_t>(1u, need(W_.root));
        W_.qnodes.resize(W_.nodes.size());
        for (size_t i = 0; i < W_.nodes.size() && err_.empty(); ++i) err_ = quantize(W_.nodes[i], W_.qnodes[i]);
        return err_;
    }

  private:
    const noor_payload_view* P_;
    WideBvh& W_;
    std::string err_;
    std::unordered_map<uint32_t, uint32_t> blas_;
    std::vector<int32_t> need_;
    std::vector<uint64_t> range_;

    uint32_t flags(uint32_t i) const { return P_->bvh_max_axis_count[i].w; }
    uint32_t word(uint32_t i) const { return P_->bvh_min_right_start[i].w; }
    BinKind kind(uint32_t i) const {
        const uint32_t w = flags(i);
        if (w & NOOR_LIGHT_NODE_MASK) return B_LIGHT;
        if (w & NOOR_MESH_INSTANCE_MASK) return B_INST;
        if (w & NOOR_MESH_LEAF_MASK) return B_MLEAF;
        if (w & NOOR_TRI_LEAF_MASK) return B_TLEAF;
        return B_INNER;
    }
    bool tlas(uint32_t i) const { return (flags(i) & NOOR_MESH_NODE_MASK) != 0; }
    Child box_of(uint32_t i) const {
        Child c;
        const noor_uint4 a = P_->bvh_min_right_start[i], b = P_->bvh_max_axis_count[i];
        const uint32_t lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
        for (int k = 0; k < 3; ++k) {
            std::memcpy(&c.lo[k], &lo[k], 4);
            std::memcpy(&c.hi[k], &hi[k], 4);
        }
        c.ref = REF_EMPTY;
        return c;
    }
    static float area(const Child& c) {
        const float dx = c.hi[0] - c.lo[0], dy = c.hi[1] - c.lo[1], dz = c.hi[2] - c.lo[2];
        if (!(dx >= 0.0f && dy >= 0.0f && dz >= 0.0f)) return 0.0f;
        return dx * dy + dy * dz + dz * dx;
    }
    std::vector<uint32_t> expand(uint32_t i) const {
        std::vector<uint32_t> ch;
        if (kind(i) == B_MLEAF) {
            const uint32_t n = flags(i) & NOOR_COUNT_MASK;
            for (uint32_t j = 1; j <= n; ++j) ch.push_back(i + j);
        } else {
            ch.push_back(i + 1);
            ch.push_back(word(i));
        }
        return ch;
    }
    // Triangle range of a bottom-level subtree (leaves in depth-first order
    // hold consecutive triangles); false if the range is not contiguous.
    bool tri_range(uint32_t i, uint32_t& start, uint32_t& count) {
        if (range_.empty()) range_.assign(P_->num_bvh_nodes, ~0ull);
        const uint64_t m = range_[i];
        if (m != ~0ull) {
            if (m == ~1ull) return false;
            start = (uint32_t)(m >> 32);
            count = (uint32_t)m;
            return true;
        }
        bool ok = false;
        const BinKind k = kind(i);
        if (k == B_TLEAF) {
            start = word(i);
            count = flags(i) & NOOR_TRI_COUNT_MASK;
            ok = true;
        } else if (k == B_INNER && !tlas(i)) {
            uint32_t s1, c1, s2, c2;
            if (tri_range(i + 1, s1, c1) && tri_range(word(i), s2, c2)) {
                if (c1 == 0) start = s2, count = c2, ok = true;
                else if (c2 == 0) start = s1, count = c1, ok = true;
                else if (s1 + c1 == s2) start = s1, count = c1 + c2, ok = true;
            }
        }
        range_[i] = ok ? ((uint64_t)start << 32 | count) : ~1ull;
        return ok;
    }
    // A bottom-level inner node over at most kLeafMerge triangles becomes one leaf.
    bool small(uint32_t i) {
        uint32_t s, n;
        return kLeafMerge > 0 && kind(i) == B_INNER && !tlas(i) && tri_range(i, s, n) && n > 0 && n <= kLeafMerge;
    }
    bool openable(uint32_t i) {
        const BinKind k = kind(i);
        return (k == B_INNER && !small(i)) || k == B_MLEAF;
    }
    size_t fanout(uint32_t i) const { return kind(i) == B_MLEAF ? (flags(i) & NOOR_COUNT_MASK) : 2u; }

    uint32_t new_node() {
        WNode n;
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 4; ++c) {
                n.lo[k][c] = std::numeric_limits<float>::infinity();
                n.hi[k][c] = -std::numeric_limits<float>::infinity();
            }
        for (int c = 0; c < 4; ++c) {
            n.ref[c] = REF_EMPTY;
            n.pad[c] = 0;
        }
        W_.nodes.push_back(n);
        return (uint32_t)(W_.nodes.size() - 1);
    }
    void fill(uint32_t idx, const std::vector<Child>& cs) {
        WNode& n = W_.nodes[idx];
        for (size_t c = 0; c < cs.size(); ++c) {
            for (int k = 0; k < 3; ++k) {
                n.lo[k][c] = cs[c].lo[k];
                n.hi[k][c] = cs[c].hi[k];
            }
            n.ref[c] = cs[c].ref;
        }
    }
    // A wide node (or a small tree of them) over more than four children.
    Child group(std::vector<Child> cs, bool is_tlas) {
        while (cs.size() > 4) {
            std::vector<Child> up;
            const size_t per = (cs.size() + 3) / 4;
            for (size_t s = 0; s < cs.size(); s += per) {
                const size_t e = std::min(cs.size(), s + per);
                std::vector<Child> part(cs.begin() + s, cs.begin() + e);
                Child u = part[0];
                for (const Child& p : part)
                    for (int k = 0; k < 3; ++k) {
                        u.lo[k] = std::min(u.lo[k], p.lo[k]);
                        u.hi[k] = std::max(u.hi[k], p.hi[k]);
                    }
                const uint32_t idx = new_node();
                fill(idx, part);
                u.ref = (is_tlas ? REF_TLAS : 0u) | (REF_INNER << REF_KIND_SHIFT) | idx;
                up.push_back(u);
            }
            cs.swap(up);
        }
        const uint32_t idx = new_node();
        fill(idx, cs);
        Child u = cs.empty() ? Child{} : cs[0];
        for (const Child& p : cs)
            for (int k = 0; k < 3; ++k) {
                u.lo[k] = std::min(u.lo[k], p.lo[k]);
                u.hi[k] = std::max(u.hi[k], p.hi[k]);
            }
        u.ref = (is_tlas ? REF_TLAS : 0u) | (REF_INNER << REF_KIND_SHIFT) | idx;
        return u;
    }

    uint32_t blas_ref(uint32_t b) {
        auto it = blas_.find(b);
        if (it != blas_.end()) return it->second;
        const BinKind k = kind(b);
        if (k != B_INNER && k != B_TLEAF) {
            err_ = "instance points at a node that is not a BLAS root";
            return REF_EMPTY;
        }
        const uint32_t r = make_child(b).ref;
        blas_[b] = r;
        return r;
    }

    Child make_child(uint32_t i) {
        Child c = box_of(i);
        if (!err_.empty()) return c;
        const uint32_t w = flags(i);
        switch (kind(i)) {
        case B_LIGHT:
            c.ref = REF_TLAS | (REF_LIGHT << REF_KIND_SHIFT) | (w & NOOR_LIGHT_MASK);
            return c;
        case B_INST: {
            const uint32_t ins = w & NOOR_TRANS_MASK;
            const uint32_t r = blas_ref(word(i));
            if (ins >= W_.inst_root.size()) {
                err_ = "instance index out of range";
                return c;
            }
            if (W_.inst_root[ins] != REF_EMPTY && W_.inst_root[ins] != r) {
                err_ = "one transform used with two different BLAS roots";
                return c;
            }
            W_.inst_root[ins] = r;
            c.ref = REF_TLAS | (REF_INST << REF_KIND_SHIFT) | ins;
            return c;
        }
        case B_TLEAF: {
            const uint32_t start = word(i), count = w & NOOR_TRI_COUNT_MASK;
            if (count == 0) return c;  // empty leaf: no reference
            if ((uint64_t)start + count > (uint64_t)REF_LEAF_START_MASK + 1) {
                err_ = "more than 67108864 triangles";
                return c;
            }
            if (count <= REF_LEAF_MAX) {
                c.ref = (REF_LEAF << REF_KIND_SHIFT) | ((count - 1) << REF_LEAF_COUNT_SHIFT) | start;
                return c;
            }
            std::vector<Child> parts;
            for (uint32_t s = 0; s < count; s += REF_LEAF_MAX) {
                Child p = c;
                const uint32_t n = std::min(REF_LEAF_MAX, count - s);
                p.ref = (REF_LEAF << REF_KIND_SHIFT) | ((n - 1) << REF_LEAF_COUNT_SHIFT) | (start + s);
                parts.push_back(p);
            }
            return group(parts, false);
        }
        case B_MLEAF:
        case B_INNER:
            break;
        }
        if (small(i)) {
            uint32_t start = 0, count = 0;
            tri_range(i, start, count);
            c.ref = (REF_LEAF << REF_KIND_SHIFT) | ((count - 1) << REF_LEAF_COUNT_SHIFT) | start;
            return c;
        }
        const bool is_tlas = tlas(i);
        std::vector<uint32_t> ch = expand(i);
        if (ch.size() > 4) {
            std::vector<Child> cs;
            for (uint32_t j : ch) {
                Child cc = make_child(j);
                if (cc.ref != REF_EMPTY) cs.push_back(cc);
            }
            return group(cs, is_tlas);
        }
        // open the largest inner children while the node has room
        for (;;) {
            int best = -1;
            float best_area = -1.0f;
            for (size_t j = 0; j < ch.size(); ++j) {
                if (!openable(ch[j]) || ch.size() - 1 + fanout(ch[j]) > 4) continue;
                const float a = area(box_of(ch[j]));
                if (a > best_area) {
                    best_area = a;
                    best = (int)j;
                }
            }
            if (best < 0) break;
            const std::vector<uint32_t> sub = expand(ch[best]);
            ch.erase(ch.begin() + best);
            ch.insert(ch.begin() + best, sub.begin(), sub.end());
        }
        const uint32_t idx = new_node();
        std::vector<Child> cs;
        for (uint32_t j : ch) {
            Child cc = make_child(j);
            if (cc.ref != REF_EMPTY) cs.push_back(cc);
        }
        fill(idx, cs);
        c.ref = (is_tlas ? REF_TLAS : 0u) | (REF_INNER << REF_KIND_SHIFT) | idx;
        return c;
    }

    // Most entries a traversal holds on its stack below `ref`: at a node with
    // k children up to k - 1 are pushed while one is followed.
    uint32_t need(uint32_t ref) {
        if (ref == REF_EMPTY) return 0;
        const uint32_t kind = (ref >> REF_KIND_SHIFT) & 3u;
        if (kind == REF_INST) {
            const uint32_t ins = ref & REF_INDEX_MASK;
            return ins < W_.inst_root.size() ? need(W_.inst_root[ins]) : 0u;
        }
        if (kind != REF_INNER) return 0;
        const uint32_t idx = ref & REF_INDEX_MASK;
        if (need_[idx] >= 0) return (uint32_t)need_[idx];
        const WNode n = W_.nodes[idx];
        uint32_t k = 0, m = 0;
        for (int c = 0; c < 4; ++c) {
            if (n.ref[c] == REF_EMPTY) continue;
            ++k;
            m = std::max(m, need(n.ref[c]));
        }
        const uint32_t r = k ? (k - 1 + m) : 0u;
        need_[idx] = (int32_t)r;
        return r;
    }
};

}  // namespace

namespace {

bool inside(const float lo[3], const float hi[3], const WNode& n, int c) {
    for (int k = 0; k < 3; ++k)
        if (lo[k] < n.lo[k][c] || hi[k] > n.hi[k][c]) return false;
    return true;
}

// Triangles under binary node b of one BLAS.
void binary_tris(const noor_payload_view* P, uint32_t b, std::vector<uint32_t>& out) {
    std::vector<uint32_t> st{b};
    while (!st.empty()) {
        const uint32_t i = st.back();
        st.pop_back();
        const uint32_t w = P->bvh_max_axis_count[i].w;
        if (w & (NOOR_LIGHT_NODE_MASK | NOOR_MESH_NODE_MASK)) continue;
        if (w & NOOR_TRI_LEAF_MASK) {
            const uint32_t s = P->bvh_min_right_start[i].w, n = w & NOOR_TRI_COUNT_MASK;
            for (uint32_t t = s; t < s + n; ++t) out.push_back(t);
            continue;
        }
        st.push_back(i + 1);
        st.push_back(P->bvh_min_right_start[i].w);
    }
}

}  // namespace

std::string check_wide_bvh(const noor_payload_view* P, const WideBvh& W) {
    const size_t nn = W.nodes.size();
    // BLAS roots: one per transform, plus the root when it is not a TLAS node
    std::vector<std::pair<uint32_t, uint32_t>> roots;  // (wide ref, binary root)
    for (uint32_t i = 0; i < P->num_bvh_nodes; ++i) {
        const uint32_t w = P->bvh_max_axis_count[i].w;
        if ((w & NOOR_LIGHT_NODE_MASK) || !(w & NOOR_MESH_INSTANCE_MASK)) continue;
        const uint32_t ins = w & NOOR_TRANS_MASK;
        if (ins >= W.inst_root.size()) return "instance out of range";
        roots.push_back({W.inst_root[ins], P->bvh_min_right_start[i].w});
    }
    if (!(P->bvh_max_axis_count[P->bvh_root_node].w & (NOOR_MESH_NODE_MASK | NOOR_LIGHT_NODE_MASK)))
        roots.push_back({W.root, P->bvh_root_node});
    std::sort(roots.begin(), roots.end());
    roots.erase(std::unique(roots.begin(), roots.end()), roots.end());
    for (const auto& r : roots) {
        std::vector<uint32_t> want;
        binary_tris(P, r.second, want);
        std::vector<uint32_t> got;
        // walk the wide BLAS: every child inside its parent's entry, leaves around their triangles
        std::vector<uint32_t> st{r.first};
        while (!st.empty()) {
            const uint32_t ref = st.back();
            st.pop_back();
            const uint32_t kind = (ref >> REF_KIND_SHIFT) & 3u;
            if (ref == REF_EMPTY) continue;
            if (kind == REF_LEAF) {
                const uint32_t s = ref & REF_LEAF_START_MASK, n = ((ref >> REF_LEAF_COUNT_SHIFT) & 7u) + 1u;
                for (uint32_t t = s; t < s + n; ++t) got.push_back(t);
                continue;
            }
            if (kind != REF_INNER) return "a BLAS references a TLAS node";
            const uint32_t idx = ref & REF_INDEX_MASK;
            if (idx >= nn) return "wide node index out of range";
            const WNode& n = W.nodes[idx];
            for (int c = 0; c < 4; ++c) {
                const uint32_t cr = n.ref[c];
                if (cr == REF_EMPTY) continue;
                const uint32_t ck = (cr >> REF_KIND_SHIFT) & 3u;
                if (ck == REF_INNER) {
                    const WNode& m = W.nodes[cr & REF_INDEX_MASK];
                    for (int d = 0; d < 4; ++d) {
                        if (m.ref[d] == REF_EMPTY) continue;
                        const float lo[3] = {m.lo[0][d], m.lo[1][d], m.lo[2][d]};
                        const float hi[3] = {m.hi[0][d], m.hi[1][d], m.hi[2][d]};
                        if (!inside(lo, hi, n, c)) return "child box outside its parent's";
                    }
                } else if (ck == REF_LEAF) {
                    const uint32_t s = cr & REF_LEAF_START_MASK, cnt = ((cr >> REF_LEAF_COUNT_SHIFT) & 7u) + 1u;
                    for (uint32_t t = s; t < s + cnt; ++t) {
                        const noor_uint4 a = P->attribute_indices[t];
                        for (uint32_t v : {a.x, a.y, a.z}) {
                            const noor_float4 p = P->vertices[v];
                            const float q[3] = {p.x, p.y, p.z};
                            if (!inside(q, q, n, c)) return "triangle outside its leaf box";
                        }
                    }
                }
                st.push_back(cr);
            }
        }
        std::sort(want.begin(), want.end());
        std::sort(got.begin(), got.end());
        if (want != got) return "a BLAS does not cover its triangles exactly once";
    }
    if (W.stack_need < 1) return "stack bound";
    // compressed nodes: same references, decoded boxes around the exact ones
    if (W.qnodes.size() != nn) return "compressed node count";
    for (size_t i = 0; i < nn; ++i) {
        const WNode& n = W.nodes[i];
        const QNode& q = W.qnodes[i];
        for (int c = 0; c < 4; ++c) {
            if (q.ref[c] != n.ref[c]) return "compressed node reference";
            const bool valid = ((q.meta >> 24) >> c) & 1u;
            if (valid != (n.ref[c] != REF_EMPTY)) return "compressed node slot mask";
            if (!valid) continue;
            for (int a = 0; a < 3; ++a) {
                const float s = std::ldexp(1.0f, (int)((q.meta >> (8 * a)) & 0xFFu) - 127);
                const uint32_t ql = (q.qlo[a] >> (8 * c)) & 0xFFu, qh = (q.qhi[a] >> (8 * c)) & 0xFFu;
                const float lo = q_decode(q.org[a], ql, s);
                const float hi = q_decode(q.org[a], qh, s);
                if (!(lo <= n.lo[a][c] && hi >= n.hi[a][c])) return "compressed box does not contain its child";
                if (!(q_exact(q.org[a], ql, s) <= (double)n.lo[a][c] && q_exact(q.org[a], qh, s) >= (double)n.hi[a][c]))
                    return "compressed box (exact grid values) does not contain its child";
            }
        }
    }
    return std::string();
}

std::string build_wide_bvh(const noor_payload_view* P, WideBvh& out) {
    Builder b(P, out);
    return b.run();
}

}  // namespace noor
