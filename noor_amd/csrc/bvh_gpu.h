// GPU binned-SAH build of one bottom-level tree, in the reference node layout.
//
// The builder makes the same decisions as the host builder of scene_build.cpp
// (16 centroid bins per axis, cost Ct + Ci (nL SA_L + nR SA_R) / SA, the leaf
// rules of noor/source/bvh.cpp:180-214, a stable partition on
// centroid <= bin bound), with the same binary32 arithmetic, so the two
// produce identical node arrays and triangle orders.  It works level by level
// over all nodes of a level at once: nodes of up to 256 triangles are split
// by one thread each, larger ones by blocks of 256 triangles that bin into
// LDS and merge their bins per node; the breadth-first tree is then laid out
// depth first (left child = index + 1, right child index in .w).
#ifndef NOOR_BVH_GPU_H
#define NOOR_BVH_GPU_H

#include <stdint.h>

#include <string>
#include <vector>

namespace noor {

struct GpuPrim {
    float mn[3], mx[3];  // triangle bounds
    float c[3];          // centroid (mn + mx) * 0.5
};

struct GpuBvhParams {
    uint32_t num_bins, max_height, min_leaf, max_leaf;
    float Ci, Ct;
};

struct GpuBvhResult {
    std::vector<uint32_t> nodes;  // 8 words per node: min.xyz, right/start, max.xyz, flags
    std::vector<uint32_t> perm;   // leaf order: perm[k] = prim index of the k-th triangle
    uint32_t height = 0, leaves = 0, inners = 0, max_leaf = 0;
    double device_ms = 0.0;       // time of the device work (build and layout)
};

// Builds the tree over prims[0, n) on `device`.  Node references are offset
// by node_base and leaf starts by tri_base.  Returns an empty string on
// success, else the reason (the caller then builds on the host).
std::string gpu_build_blas(int device, const GpuPrim* prims, uint32_t n, const float root_box[6],
                           const float root_cent[6], const GpuBvhParams& P, uint32_t node_base, uint32_t tri_base,
                           GpuBvhResult& out);

}  // namespace noor

#endif
