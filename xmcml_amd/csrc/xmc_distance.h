// Conservative distance grid for skipping BVH walks that cannot hit.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/xmcml.h"

namespace xmc {

// Cell c of a dims[0] x dims[1] x dims[2] grid with origin `lo` and cubic
// cells of edge h holds q[c] such that the centre of the cell,
// lo + (i + 1/2) h, is at least q[c] * step_q away from every triangle.  A
// point p is then at least q[c] * step_q - |p - centre| away, and a
// step-limited query from p shorter than that cannot hit anything.
struct DistanceGrid {
    float lo[3] = {0, 0, 0};
    float h = 1.0f;
    float step_q = 0.125f;  // distance quantum (h / 8)
    float reach = 0.0f;     // R: distances up to R are exact (before quantisation)
    int dims[3] = {0, 0, 0};
    std::vector<uint8_t> q;
};

// Builds the grid over [lo, hi] with about `target_cells` cells.  The bound
// is the exact distance from the cell centre to the nearest triangle when
// that is below R (4 cells, or fewer on very large meshes), otherwise the
// larger of R and the distance transform of the cells that meet triangle
// bounding boxes minus half a cell diagonal; floored to the quantum.
void build_distance_grid(const xmc_surface* surfaces, int n_surfaces, const double lo[3], const double hi[3],
                         size_t target_cells, DistanceGrid& out);

// Separating slabs for the cells near the mesh, where the distance bound
// skips few steps.  Cell c's word holds a radius R, a direction n and up to
// two bands [a1, b1] <= [a2, b2] of n.(x - c) that hold every triangle
// point within R of the cell centre c; the rest of the ball, the free
// intervals n.(x - c) < a1, b1 < n.(x - c) < a2 and n.(x - c) > b2, holds
// none.  A straight step whose two ends lie in one free interval (the
// interval is convex) and stay within R of c (the ball is convex) crosses no
// triangle.  Layout, low bits first: n as two int8 octahedral coordinates,
// a1, b1, a2, b2 as int8 multiples of R / 127, R as a uint8 multiple of
// h / 8; 0 = no slab (R = 0 skips nothing).  Built for cells whose distance
// bound is below 3.5 cells; per cell, R and n (the nearest triangle's normal
// or the mean normal within R) are those that skip most of 64 sample steps
// (uniform start in the cell, isotropic direction, exponential length of
// mean `mean_step`).
struct SlabGrid {
    float unit = 0.0f;  // h / 8
    std::vector<uint64_t> words;
};

void build_slab_grid(const xmc_surface* surfaces, int n_surfaces, const DistanceGrid& dg, float mean_step,
                     SlabGrid& out);

// Octahedral direction of a slab word: (u, v, 1 - |u| - |v|) with u, v = a, b / 127,
// folded over the lower half-space; not unit length (|n| <= 1).
inline void slab_direction(int a, int b, float n[3]) {
    float u = (float)a * (1.0f / 127.0f), v = (float)b * (1.0f / 127.0f);
    float z = 1.0f - (u < 0.0f ? -u : u) - (v < 0.0f ? -v : v);
    if (z < 0.0f) {
        float fu = (1.0f - (v < 0.0f ? -v : v)) * (u >= 0.0f ? 1.0f : -1.0f);
        float fv = (1.0f - (u < 0.0f ? -u : u)) * (v >= 0.0f ? 1.0f : -1.0f);
        u = fu;
        v = fv;
    }
    n[0] = u;
    n[1] = v;
    n[2] = z;
}

}  // namespace xmc
