// Four-wide BVH used by the GPU traversal.
//
// It is derived on the host, at context creation, from the binary two-level
// BVH of the payload (the reference node format, noor/cuda/bvh.cuh:28-140):
// each wide node holds the boxes of up to four children, taken unchanged from
// the binary nodes they replace, so one 128-byte fetch tests four boxes and a
// ray needs about half the dependent node fetches of the binary traversal.
// Together with the conservative slab test this returns the same hits as the
// binary traversal (see box4_hit in dev_core.h).
//
// Child references (also the traversal stack entries), 32 bits:
//   bit 31      TLAS level: the world-space ray applies below this entry
//   bits 30-29  kind: 0 inner (wide node index), 1 triangle leaf,
//               2 instance (transform index), 3 area-light node (light index)
//   leaf:       bits 0-25 first triangle (leaf order), bits 26-28 count - 1
//   others:     bits 0-28 index
#ifndef NOOR_WIDE_BVH_H
#define NOOR_WIDE_BVH_H

#include <stdint.h>

namespace noor {

constexpr uint32_t REF_TLAS = 0x80000000u;
constexpr uint32_t REF_KIND_SHIFT = 29;
constexpr uint32_t REF_INNER = 0u, REF_LEAF = 1u, REF_INST = 2u, REF_LIGHT = 3u;
constexpr uint32_t REF_INDEX_MASK = 0x1FFFFFFFu;
constexpr uint32_t REF_LEAF_START_MASK = 0x03FFFFFFu;
constexpr uint32_t REF_LEAF_COUNT_SHIFT = 26;
constexpr uint32_t REF_LEAF_MAX = 8u;
constexpr uint32_t REF_EMPTY = 0xFFFFFFFFu;

// Host form of a wide node: SoA boxes of the four children (an empty slot has
// lo = +inf, hi = -inf), then their references.  The builder and the checks
// work on this form; the GPU reads the compressed QNode below.
struct alignas(128) WNode {
    float lo[3][4];
    float hi[3][4];
    uint32_t ref[4];
    uint32_t pad[4];
};
static_assert(sizeof(WNode) == 128, "WNode is one 128-byte line");

// Device form, 64 bytes (56 read per visit): the four child boxes quantised
// to 8 bits per plane on a grid local to the node (as in Ylitie, Karras &
// Laine, "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide
// BVHs", HPG 2017).  On axis a a plane decodes to
//     org[a] + (float)q * 2^(e[a] - 127)        (binary32, no contraction)
// with e[a] = byte a of `meta`.  The encoder rounds lower planes down and upper
// planes up and checks the decoded values, so every decoded box contains its
// child's exact box; with the conservative slab test the traversal therefore
// returns the same hits as over the exact boxes.  Byte 3 of `meta` has bit c
// set when slot c holds a child.  Words: [org.xyz, meta] [qlo.xyz, qhi.x]
// [qhi.yz, ref0, ref1] [ref2, ref3, pad, pad]; slot c is byte c of each q word.
struct alignas(64) QNode {
    float org[3];
    uint32_t meta;
    uint32_t qlo[3];
    uint32_t qhi[3];
    uint32_t ref[4];
    uint32_t pad[2];
};
static_assert(sizeof(QNode) == 64, "QNode is 64 bytes");

}  // namespace noor

#include <string>
#include <vector>

#include "noor_hip.h"

namespace noor {

struct WideBvh {
    std::vector<WNode> nodes;         // exact boxes (host)
    std::vector<QNode> qnodes;        // the same nodes, compressed (uploaded)
    std::vector<uint32_t> inst_root;  // per transform: reference of its BLAS root
    uint32_t root = REF_EMPTY;
    uint32_t stack_need = 0;          // most stack entries any traversal can hold
};

// Builds the wide BVH of a validated payload.  Returns an empty string on
// success, else the reason the payload cannot be represented.
std::string build_wide_bvh(const noor_payload_view* P, WideBvh& out);

// Structural check of a built wide BVH against the payload (see
// noor_wide_bvh_check).  Returns an empty string when it holds.
std::string check_wide_bvh(const noor_payload_view* P, const WideBvh& W);

}  // namespace noor

#endif
