// The surface forms of the transport's surface rule (host side, no HIP):
// axis-aligned rectangles, whose nearest hit is found analytically, and
// heightfields on a shared uniform lattice, whose nearest hit is found by
// testing the triangles of the lattice cells a step's bounding box covers.
#pragma once

#include <vector>

#include "../../include/xmcml.h"

namespace xmc {

// A surface of two triangles over the four corners of an axis-aligned
// rectangle at one z, sharing a diagonal (as the reference's plane generator
// writes them).
struct RectForm {
    float z, x0, x1, y0, y1;
};
bool planar_rectangle(const xmc_surface& s, RectForm& r);

// Heightfield surfaces on one uniform lattice: the distinct x (and y)
// coordinates of their vertices are x0 + i dx, i = 0..nx (y0 + j dy,
// j = 0..ny); every triangle spans exactly one lattice cell along x and y
// (so it lies inside that cell's rectangle), at most two triangles per cell
// and surface.  cells holds, per (cell, surface), four floats: the z-range of
// the cell's triangles of that surface (+inf, -inf when it has none) and
// their two leaf indices as int bits (lower first, -1 = none).  The margins
// cover the float rounding of the cell arithmetic and of the triangle test.
struct LatticeForm {
    int ns = 0, nx = 0, ny = 0;
    float x0 = 0.0f, y0 = 0.0f, inv_dx = 0.0f, inv_dy = 0.0f;
    float mxy = 0.0f, mz = 0.0f, zlo = 0.0f, zhi = 0.0f;
    std::vector<float> cells;  // [((iy * nx + ix) * ns + s) * 4 + k]
};

// Builds the lattice form of surfaces[which[0..]] (in that order); leaf_of
// maps base[s] + t (base = the running triangle count over surfaces) to the
// leaf index of triangle t of surface s.  Returns false when they do not
// share such a lattice (or there are more than 4 of them, or it would exceed
// 2^26 records).
bool lattice_form(const xmc_surface* surfaces, int n_surfaces, const std::vector<int>& which,
                  const std::vector<int>& leaf_of, LatticeForm& out);

}  // namespace xmc
