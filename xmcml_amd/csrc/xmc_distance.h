// Conservative distance grid for skipping BVH walks that cannot hit.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/xmcml.h"

namespace xmc {

// Cell c of a dims[0] x dims[1] x dims[2] grid with origin `lo` and cubic
// cells of edge h holds q[c] such that every point of the cell is at least
// q[c] * step_q away from every triangle.  A step-limited query from inside
// cell c cannot hit anything when its length is below that bound.
struct DistanceGrid {
    float lo[3] = {0, 0, 0};
    float h = 1.0f;
    float step_q = 0.25f;  // distance quantum (h / 4)
    int dims[3] = {0, 0, 0};
    std::vector<uint8_t> q;
};

// Builds the grid over [lo, hi] with about `target_cells` cells.  The bound
// is |c_C - c_M| - sqrt(3) h minimised over cells M that intersect some
// triangle's bounding box (exact Euclidean distance transform of cell
// centres), floored to the quantum.
void build_distance_grid(const xmc_surface* surfaces, int n_surfaces, const double lo[3], const double hi[3],
                         size_t target_cells, DistanceGrid& out);

}  // namespace xmc
