// Entry-node grid: where a short step-limited walk may start instead of the
// BVH root.
#pragma once

#include <cstdint>
#include <vector>

#include "xmc_bvh.h"

namespace xmc {

// Cell c of a dims[0] x dims[1] x dims[2] grid with origin `lo` and cubic
// cells of edge h holds an index into the entry table, or -1 when no leaf
// box comes within `margin` (+ a rounding guard) of the cell.  The entry is
// the deepest node whose subtree holds every leaf whose box comes that
// close.  A walk from a point of the cell whose step moves it at most
// `margin` along every axis cannot hit a triangle outside that subtree, so
// starting there (after testing the entry's own box, which implies every
// ancestor's test) visits the same leaves in the same order as a walk from
// the root and returns the same hit.
struct EntryGrid {
    float margin = 0.0f;
    std::vector<int32_t> cell;       // entry index or -1
    std::vector<int32_t> ref;        // entry node reference (>= 0 node, < 0 leaf)
    std::vector<float> box;          // 6 floats per entry: lo[3], hi[3]
};

// Entries for the cells of the grid (lo, h, dims) -- the distance grid's, so
// that one cell word serves both -- with margin = h.
void build_entry_grid(const GpuBvh& bvh, const float lo[3], float h, const int dims[3], EntryGrid& out);

}  // namespace xmc
