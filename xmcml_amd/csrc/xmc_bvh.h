// Host-side BVH for the GPU engine: the reference's binned-SAH build and its
// flattening into the layout the transport kernel walks.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/xmcml.h"

namespace xmc {

// A node in the reference's layout (BVHNode, mcml_intersection.h:37-51):
// internal nodes have count 0 and offset -1, leaves left = right = -1.
struct RefNode {
    double lo[3], hi[3];
    int left, right, offset, count;
};

struct RefBvh {
    std::vector<RefNode> nodes;  // post-order, as GenerateBVHTree emits them
    std::vector<int> tri_surface, tri_index;  // leaf order
    int root = 0;
    int depth = 0;  // number of nodes on the longest root-to-leaf path
};

// GenerateBVHTree (mcml_intersection.cpp:316-344) on the caller's surfaces.
// Returns false (and fills err) on a degenerate input.
bool build_reference_bvh(const xmc_surface* surfaces, int n_surfaces, RefBvh& out, const char** err);

// GPU node: the boxes of both children (float), then the child references.
// A reference r >= 0 is an internal node; r < 0 is a leaf whose triangles
// start at ~r in the triangle array and run until the one flagged last.
// A child reference in an entry's restricted tree (xmc_entry.h) for a child
// none of whose leaves is in reach: the walk neither visits nor pushes it (a
// walk from the root would visit only nodes inside it, push and pop only its
// own entries, and change neither the hit nor the stack below).
constexpr int kDeadRef = 0x7FFFFFFF;

struct alignas(16) GpuNode {
    float c0lo[3];
    float c0hi[3];
    float c1lo[3];
    float c1hi[3];
    int child[2];
    int pad[2];
};
static_assert(sizeof(GpuNode) == 64, "GpuNode is one 64-byte record");

// GPU triangle: vertex a, edges b-a and c-a in float, surface id, index of
// the triangle in its surface, and a flag on the last triangle of each leaf.
struct alignas(16) GpuTri {
    float a[3];
    int surface;
    float e1[3];
    int index;
    float e2[3];
    int last;
};
static_assert(sizeof(GpuTri) == 48, "GpuTri is three 16-byte words");

struct GpuBvh {
    std::vector<GpuNode> nodes;  // pre-order, node 0 is the root if it is internal
    std::vector<GpuTri> tris;
    int root_ref = 0;
    float root_lo[3], root_hi[3];
    int depth = 0;
};

void flatten_bvh(const RefBvh& ref, const xmc_surface* surfaces, GpuBvh& out);

}  // namespace xmc
