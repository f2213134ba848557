// Entry grid: where, and on which nodes, a short step-limited walk runs
// instead of the whole BVH from its root.
#pragma once

#include <cstdint>
#include <vector>

#include "xmc_bvh.h"

namespace xmc {

// Cell c of a dims[0] x dims[1] x dims[2] grid with origin `lo` and cubic
// cells of edge h holds an index into the entry table, or -1 when no leaf
// box comes within `margin` (+ a rounding guard) of the cell.  A walk from a
// point of the cell whose step moves it at most `margin` along every axis
// cannot hit a triangle in any other leaf.  Its entry is the BVH restricted
// to those leaves: the walk starts at the deepest node holding all of them
// (after testing that node's box, which implies every ancestor's test) and
// runs on copies of the nodes below it in which a chain of nodes with only
// one child in reach is skipped -- the chain's end is entered directly (its
// children's boxes lie inside every box of the chain, so their tests decide
// as the chain's would), a leaf at a chain's end behind a gate node that
// tests the leaf's box, and a child none of whose leaves is in reach is
// kDeadRef, which the walk neither visits nor pushes.  The copies keep their children's boxes, so the walk
// orders, pushes and prunes as the walk from the root does and returns the
// same hit, with fewer visits.  Copies and gates are appended to the BVH's
// node array (`extra`, references from bvh.nodes.size() on), equal nodes
// shared between cells.
struct EntryGrid {
    float margin = 0.0f;
    std::vector<int32_t> cell;       // entry index or -1
    // the same for steps that move at most margin2 = 3 h along every axis (their
    // own trees, in the same entry table)
    float margin2 = 0.0f;
    std::vector<int32_t> cell2;
    std::vector<int32_t> ref;        // entry's start reference (>= 0 node, < 0 leaf)
    std::vector<float> box;          // 6 floats per entry: the start node's box lo[3], hi[3]
    std::vector<GpuNode> extra;      // restricted-tree nodes, after the BVH's own
};

// Entries for the cells of the grid (lo, h, dims) -- the distance grid's, so
// that one cell word serves both -- with margin = h, and a second level with
// margin2 = 3 h.  With the distance grid's
// bounds (q[c] * step_q, or q = NULL for none), cells whose every short step
// the bound skips get no entry.
void build_entry_grid(const GpuBvh& bvh, const float lo[3], float h, const int dims[3], const uint8_t* q, float step_q,
                      EntryGrid& out);

}  // namespace xmc
