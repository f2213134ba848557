// Launch parameters of the transport and query kernels (host and device).
#pragma once

#include <cstdint>

#include "../../include/xmcml.h"
#include "xmc_bvh.h"

namespace xmc {

constexpr int kMaxLayers = XMC_MAX_LAYERS;
constexpr int kStreamLog2 = 24;        // MCG59 draws reserved per photon
// photons < 2^33 per seed: MCG59 with an odd state has period 2^57
// (a = 13^13 = 5 mod 8), so 2^57 / 2^24 = 2^33 disjoint blocks; photon
// p + 2^33 would repeat photon p's numbers
constexpr int kJumpBits = 57 - kStreamLog2;
constexpr uint64_t kMcgA = 302875106592253ULL;     // 13^13, mcml_mcg59.cpp:3
constexpr uint64_t kMcgMask = 576460752303423487ULL;  // 2^59 - 1
constexpr int kStackLds = 8;           // LDS entries of the per-lane short stack
// Allocation-free ray queries: the per-lane stack spill lives in a static
// device array of this size; at most kQueryLanes lanes share it.
constexpr size_t kQueryScratchBytes = 32u << 20;
constexpr int kQueryLanes = 131072;
constexpr int kMaxDevices = 64;
// A photon whose stream block is nearly used up is retired (its weight goes
// to the "other" ledger entry): drawing past 2^24 would reuse the next
// photon's numbers.  One interaction draws at most 4 numbers plus redraws.
constexpr uint32_t kDrawLimit = (1u << kStreamLog2) - 64u;

struct Layers {
    float n[kMaxLayers];
    float mua[kMaxLayers];
    float mus[kMaxLayers];
    float g[kMaxLayers];
    float inv_mut[kMaxLayers];            // 1 / (mua + mus), in float
    float abs_frac[kMaxLayers];           // mua / (mua + mus), in float
    float crit[kMaxLayers][kMaxLayers];  // critical cosine for (from, to)
    uint32_t glass_mask;                  // bit l: layer l has mua == mus == 0
    uint32_t target_mask;                 // bit l: targetRangeLayers[l]
    int count;
};

struct Detectors {
    int n_cube, n_ring;
    const float* cube;  // 7 per cube: lo.xyz, hi.xyz (center -/+ length/2), cos target
    const float* ring;  // 5 per ring: center.xy, rs, rb, cos target
    const uint32_t* filter;  // per detector (cubes first): bit l = filterLayers[l]
    int T;                   // time bins per detector
    double t0, dt, inv_dt;   // inv_dt = 1 / dt (0 when dt = 0)
    const double* t_start;   // T
    const double* t_finish;  // T
};

struct Tallies {
    double* absorption;      // gridSize, or null
    double* det_weight;      // per detector
    double* det_target;
    unsigned long long* det_count;
    double* bin_weight;      // ndet * T
    double* bin_target;
    unsigned long long* bin_count;
    unsigned long long* stats;  // 6 event counters, trajectory overflows, entry walks, then 8 ledger doubles from word 8
};

// Voxel-centre plane crossings of the photon path (UpdatePhotonTrajectory,
// mcml_kernel.cpp:750-792): the scattering grid and, for detected photons,
// the per-detector voxel counts (UpdateDetectorTrajectory, :794-810).
struct Paths {
    double* scattering;              // gridSize, or null
    unsigned long long* det_traj;    // ndet * gridSize, or null
    uint32_t* lane_voxels;           // lanes * capacity voxel indices (the photon's trajectory)
    int capacity;                    // voxels kept per photon
    int grid_size;
};

struct Area {
    float corner[3];
    float length[3];
    float partf[3];
    float scale[3];  // partf / length, in float
    int part[3];
};

struct Trace {
    xmc_trace_event* events;
    int* counts;
    int max_events;
    uint64_t first;  // photons [first, first + n) are traced
    uint64_t n;
};

// Conservative distance grid (see xmc_distance.h); q == nullptr disables it.
// Per-cell word of the walk-skipping grid: bits 0-7 hold q, a lower bound
// q * step_q on the distance from the cell's centre to any triangle;
// bits 8-31 hold the cell's entry for walks that move at most `margin`
// along every axis: 0 = no triangle within reach, kEntryRoot = start at the
// root, else 1 + an index into entry_ref / entry_box (node reference and its
// box, 6 floats).  cells2 holds, per cell, the entry (same encoding, all 32
// bits) for walks that move more than `margin` but at most `margin2`.
constexpr uint32_t kEntryRoot = 0xFFFFFFu;

struct DistGrid {
    const uint32_t* cells;  // null when the scene has no grid
    float lo[3];
    float h;
    float inv_h;
    float step_q;
    float margin;
    int dims[3];
    const int* entry_ref;
    const float* entry_box;
    const uint32_t* cells2;  // null without entries
    float margin2;
    // per cell a separating slab near the mesh (64-bit words, layout of
    // SlabGrid in xmc_distance.h): radius unit h / 8; null = none
    const uint64_t* slabs;
    float slab_unit;
};

// Surfaces that are axis-aligned rectangles in a plane z = const (two
// triangles sharing a diagonal, as the reference's plane generator makes
// them): their nearest hit is found analytically.
constexpr int kPlanarMax = 16;
struct Planar {
    int n;
    float4 rect[kPlanarMax];  // x0, x1, y0, y1 (one load per rectangle where a batch lane reads them)
    float z[kPlanarMax];
    int tri[kPlanarMax];  // the surface's first triangle in BVH leaf order (its normal and surface id)
};

// Heightfield surfaces on a shared uniform lattice (every triangle inside one
// lattice cell, at most two per cell and surface): the step's nearest surface
// among them is found by testing the triangles of the cells that the step's xy
// bounding box (widened by mxy) covers, in cells whose z-range (per surface)
// meets the step's z-range (widened by mz).  Triangles are tested exactly as
// the BVH leaves test them, so a hit is the walk's hit (on equal distances the
// lower leaf index).  ns = 0: none.
constexpr int kLatticeMaxBox = 4;       // cells along x and y; larger boxes (long steps) take the BVH walk
constexpr int kLatticeMaxSurfaces = 4;  // heightfields per lattice
struct Lattice {
    int ns;              // lattice surfaces
    int nx, ny;          // cells along x and y
    float x0, y0;        // lattice origin
    float inv_dx, inv_dy;
    float mxy, mz;       // margins of the bounding box
    float zlo, zhi;      // z-range of every lattice triangle
    const float4* cells; // [(iy * nx + ix) * ns + s] = (zmin, zmax, triangle a, triangle b) (bits; -1 = none)
};

// A box of absorption voxels (around the beam entry, where most weight is
// deposited) that each block accumulates in LDS and adds to the grid once at
// the end; deposits elsewhere go straight to the grid.  count = 0: no tile.
struct AbsTile {
    int lo[3];
    int size[3];
    int count;
};

// Attractive-vector biasing (mcml_kernel.cpp:325-440) in float: the
// attractive vector of a photon at p in layer l is
// normalize(t2 / |t2|^3 - t1 / |t1|^3), t_k = (p - start) - (rd.x, rd.y, rz_k[l]),
// rd = target - start, r1z = rd.z + ltr, r2z = rd.z - (ltr + 2 delta) with
// ltr = 3 / (3 ((1 - g) mus + mua)), delta = 0.74 ltr; the weight factor of a
// deflection from direction d is table[off[l] + round((d.a + 1) half_n[l])] / p.
struct Bias {
    int on;
    float factor;               // attractiveFactor (the acceptance probability's exponent)
    float rd[3];
    float r1z[kMaxLayers], r2z[kMaxLayers];
    const float* table;         // every layer's weight-integral table, float
    int off[kMaxLayers];        // first entry of layer l's table
    int last[kMaxLayers];       // its last index (entries - 1)
    float half_n[kMaxLayers];   // 0.5 * last
};

struct Params {
    const GpuNode* nodes;
    const GpuTri* tris;
    const float* normals;  // 4 floats per triangle: the unit normal normalize(cross(e1, e2)), in float
    DistGrid dist;
    int root_ref;
    float root_lo[3], root_hi[3];
    const int8_t* next_layer;  // [layer * n_surfaces + surface] -> layer or -1
    int n_surfaces;
    Layers layers;
    Area area;
    Detectors det;
    Tallies tally;
    // photons and streams
    uint64_t first_photon;
    uint64_t n_photons;
    unsigned long long* queue;  // next photon to hand out (relative to first)
    uint64_t seed_value;        // (2 seed + 1) mod 2^59
    uint64_t jump[kJumpBits];   // (a^(2^24))^(2^i) mod 2^59
    float start_pos[3], start_dir[3];
    float w0;
    float min_weight, chance;
    int roulette;
    // short-stack spill (global) for entries beyond kStackLds
    int2* spill;
    int spill_depth;  // entries per lane in the spill
    // the stacks' rings of kStackLds entries: null = in LDS, else in global memory
    // ([entry][lane], every lane of the grid); plain single scenes only
    int2* ring;
    Trace trace;
    Paths paths;
    AbsTile abs_tile;
    Bias bias;
    // detectors whose weight, target weight and count each block sums in LDS
    // and adds to the tallies once at its end (0: every detection goes
    // straight to the tallies); the time bins always go straight
    int det_lds;
    // the step's nearest surface from the planar and lattice forms instead of the
    // BVH walk: 0 no, 1 rectangles only (planar.n), 2 rectangles and heightfields
    // (lattice.ns > 0)
    int surface_rule;
    Planar planar;
    Lattice lattice;
};

struct QueryParams {
    const GpuNode* nodes;
    const GpuTri* tris;
    int root_ref;
    float root_lo[3], root_hi[3];
    const float* origins;     // 3 per ray
    const float* directions;  // 3 per ray
    const float* steps;       // <= 0: unlimited
    int n;
    int brute;
    int n_tris;
    int* found;
    float* distance;
    int* surface;
    int* triangle;
    float* normal;
    int2* spill;
    int spill_depth;
};

}  // namespace xmc
