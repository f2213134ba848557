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

}  // namespace xmc
