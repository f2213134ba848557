// Host side of the C-ABI (include/xmcml.h): scene preparation, launches,
// tally collection.  Replaces GenerateBVHTree + LaunchOMP
// (base/xmcml/xmcml/mcml_intersection.cpp:316-344,
//  base/xmcml/xmcmlLauncher/launcher_omp.cpp:10-163).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/xmcml.h"
#include "xmc_bvh.h"
#include "xmc_distance.h"
#include "xmc_lattice.h"
#include "xmc_entry.h"
#include "xmc_params.h"

extern "C" {
hipError_t xmc_launch_transport(const xmc::Params* p, int grid, int block, size_t lds, hipStream_t s, int trace);
hipError_t xmc_launch_query(const xmc::QueryParams* q, const int* perm, int block, size_t lds, hipStream_t s);
hipError_t xmc_launch_query_stride(const xmc::QueryParams* q, const int* perm, int lanes, size_t lds, hipStream_t s);
hipError_t xmc_query_scratch(int device, int2** ptr, size_t* entries);
hipError_t xmc_launch_zero(unsigned long long* p, size_t n, hipStream_t s);
hipError_t xmc_transport_occupancy(int block, size_t lds, int* blocks_per_cu);
hipError_t xmc_transport_occupancy_plain(int block, size_t lds, int* blocks_per_cu);
hipError_t xmc_transport_occupancy_merge(int block, size_t lds, int* blocks_per_cu);
hipError_t xmc_launch_batch(const xmc::Params* p, const xmc::Params* cfgs, const uint64_t* prefix, int ncfg, int grid,
                            int block, size_t lds, hipStream_t s);
hipError_t xmc_batch_occupancy(int block, size_t lds, int* blocks_per_cu);
}

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

}  // namespace

namespace xmc {
// Shared with the other translation units of the library so that every
// entry point reports through xmc_last_error().
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace xmc

namespace {

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return fail(XMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));       \
    } while (0)

constexpr double kPi = 3.14159265358979323846;

// dynamic LDS of the transport kernels: short stacks (unless their rings are in global
// memory), the block's ledger, its detector sums (three words per detector), then its
// absorption tile
size_t transport_lds(int block, int tile_count = 0, int det_lds = 0, bool stack_in_lds = true) {
    return (stack_in_lds ? (size_t)xmc::kStackLds * block * sizeof(int2) : 0) + 8 * sizeof(double) +
           (size_t)det_lds * 3 * sizeof(double) + (size_t)tile_count * sizeof(double);
}

// detectors up to which a block keeps its detector sums in LDS
constexpr int kDetLds = 64;

// LDS absorption tile: at most this many voxels per block (24 kB; measured on the slab
// faster than 13.5 kB at five blocks, 4.2e7 vs 3.8e7 photons/s, against 2.3e7 without a
// tile -- when the stacks' rings were in LDS; with them in global memory and merged
// deposits, six blocks per CU, 2,752 and 2,048 voxels measured the same)
#ifndef XMC_ABS_TILE_VOXELS
#define XMC_ABS_TILE_VOXELS 3072
#endif
constexpr int kAbsTileVoxels = XMC_ABS_TILE_VOXELS;

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void**)&p, count * sizeof(T));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

uint64_t mcg_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r *= a;
        a *= a;
        e >>= 1;
    }
    return r;
}

}  // namespace

struct xmc_scene {
    int device = 0;
    int n_layers = 0, n_surfaces = 0, n_det = 0, T = 0, grid = 0;
    int block = 256, blocks = 0;
    bool absorption = true;
    double rsp = 0.0;
    hipStream_t stream = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;  // transport launches not yet timed
    double kernel_ms = 0.0;
    xmc::RefBvh ref;
    std::vector<int> perm;  // (surface, index) order -> leaf-order position
    DevBuf<xmc::GpuNode> nodes;
    DevBuf<xmc::GpuTri> tris;
    DevBuf<float> bias_table;  // weight-integral tables of the biasing, float
    DevBuf<float> normals;  // 4 per triangle
    DevBuf<int8_t> next_layer;
    DevBuf<float> cube, ring;
    DevBuf<uint32_t> filter;
    DevBuf<double> t_start, t_finish;
    // tallies: one 64-bit slab, views below
    DevBuf<unsigned long long> slab;
    size_t off_abs = 0, off_dw = 0, off_dt = 0, off_dc = 0, off_bw = 0, off_bt = 0, off_bc = 0, off_st = 0,
           off_q = 0, slab_words = 0;
    DevBuf<int2> spill;
    DevBuf<int> perm_dev;
    DevBuf<uint32_t> dist;  // walk-skipping grid cells
    DevBuf<uint32_t> dist2;  // second-level entries (steps of up to 3 cells)
    DevBuf<uint64_t> slabs;  // separating slabs of the cells near the mesh
    DevBuf<float> lattice_cells;  // the surface rule's heightfield cell records
    DevBuf<int> entry_ref;
    DevBuf<float> entry_box;
    // voxel-plane crossings: scattering grid, detector trajectory counts and
    // the per-lane trajectory lists (allocated only when asked for)
    DevBuf<double> scat;
    DevBuf<unsigned long long> dtraj;
    DevBuf<uint32_t> lane_vox;
    int traj_cap = 0;
    xmc::Params base;
    std::vector<double> h_t_start, h_t_finish;
};

namespace {

int validate(const xmc_input* in) {
    if (!in) return fail(XMC_ERR_ARG, "input is NULL");
    if (in->numberOfLayers < 2 || in->numberOfLayers > XMC_MAX_LAYERS)
        return fail(XMC_ERR_LIMIT, "numberOfLayers must be in [2, 8] (MAX_LAYERS)");
    if (!in->layerInfo || !in->area) return fail(XMC_ERR_ARG, "layerInfo / area is NULL");
    const xmc_layer& l1 = in->layerInfo[1];
    if (l1.absorptionCoefficient == 0.0 && l1.scatteringCoefficient == 0.0 && in->numberOfLayers < 3)
        return fail(XMC_ERR_ARG, "a glass layer 1 needs a layer 2 (ComputeSpecularReflectance)");
    const xmc_int3& p = in->area->partitionNumber;
    if (p.x < 1 || p.y < 1 || p.z < 1 || p.x > 255 || p.y > 255 || p.z > 255)
        return fail(XMC_ERR_LIMIT, "area partitions must be in [1, 255] per axis (byte3 voxel indices)");
    if (!(in->area->length.x > 0 && in->area->length.y > 0 && in->area->length.z > 0))
        return fail(XMC_ERR_ARG, "area lengths must be positive");
    if (in->numberOfSurfaces < 1 || !in->surface) return fail(XMC_ERR_ARG, "no surfaces");
    if (in->numberOfSurfaces > 127) return fail(XMC_ERR_LIMIT, "at most 127 surfaces");
    for (int i = 0; i < in->numberOfLayers; i++) {
        const xmc_layer& l = in->layerInfo[i];
        if (l.numberOfSurfaces > 0 && !l.surfaceId) return fail(XMC_ERR_ARG, "layer surfaceId is NULL");
        for (int j = 0; j < l.numberOfSurfaces; j++)
            if (l.surfaceId[j] < 0 || l.surfaceId[j] >= in->numberOfSurfaces)
                return fail(XMC_ERR_ARG, "layer surfaceId out of range");
        if (!(l.refractiveIndex > 0)) return fail(XMC_ERR_ARG, "refractive index must be positive");
    }
    if (in->timeScaleSize < 1) return fail(XMC_ERR_ARG, "timeScaleSize must be >= 1");
    if (in->numberOfCubeDetectors < 0 || in->numberOfRingDetectors < 0)
        return fail(XMC_ERR_ARG, "negative detector count");
    if (in->numberOfCubeDetectors > 0 && !in->cubeDetector) return fail(XMC_ERR_ARG, "cubeDetector is NULL");
    if (in->numberOfRingDetectors > 0 && !in->ringDetector) return fail(XMC_ERR_ARG, "ringDetector is NULL");
    if (in->terminationMode != XMC_TERMINATE_KILL && in->terminationMode != XMC_TERMINATE_ROULETTE)
        return fail(XMC_ERR_ARG, "unknown terminationMode");
    if (in->terminationMode == XMC_TERMINATE_ROULETTE && !(in->rouletteChance > 0 && in->rouletteChance <= 1))
        return fail(XMC_ERR_ARG, "rouletteChance must be in (0, 1]");
    return XMC_OK;
}

// Attractive-vector biasing: the weight-integral tables (the caller's, or computed on
// the GPU for every distinct anisotropy of layers 1.., GetDifferentAnisotropies,
// weight_integral.cpp:67-90), each layer's table (the last one whose anisotropy equals
// the layer's, GetWeightTableIndex, mcml_kernel.cpp:436-446) and the per-layer float
// constants of GetLocalAttractiveVector (:325-362).
int setup_bias(const xmc_input* in, xmc_scene* sc, xmc::Params& P) {
    if (!(in->attractiveFactor == in->attractiveFactor)) return fail(XMC_ERR_ARG, "attractiveFactor is NaN");
    std::vector<double> gs;
    for (int l = 1; l < in->numberOfLayers; l++) {
        double g = in->layerInfo[l].anisotropy;
        if (std::find(gs.begin(), gs.end(), g) == gs.end()) gs.push_back(g);
    }
    std::vector<xmc_weight_table> tables;
    std::vector<std::vector<double>> own;
    if (in->weightTable && in->numberOfWeightTables > 0) {
        for (int t = 0; t < in->numberOfWeightTables; t++) {
            const xmc_weight_table& w = in->weightTable[t];
            if (!w.elements || w.numberOfElements < 2) return fail(XMC_ERR_ARG, "weight table with < 2 elements");
            tables.push_back(w);
        }
    } else {
        const int bp = in->weightIntegralPrecision > 0 ? in->weightIntegralPrecision : 128;
        const int tp = in->weightTablePrecision > 0 ? in->weightTablePrecision : 262144;
        for (double g : gs) {
            own.emplace_back((size_t)tp + 1);
            int rc = xmc_weight_table_compute(g, in->attractiveFactor, bp, tp, own.back().data(), sc->device);
            if (rc) return rc;
            tables.push_back({g, tp + 1, own.back().data()});
        }
    }
    std::vector<float> flat;
    std::vector<int> first(tables.size());
    for (size_t t = 0; t < tables.size(); t++) {
        first[t] = (int)flat.size();
        for (int e = 0; e < tables[t].numberOfElements; e++) flat.push_back((float)tables[t].elements[e]);
    }
    xmc::Bias& B = P.bias;
    std::memset(&B, 0, sizeof(B));
    for (int l = 1; l < in->numberOfLayers; l++) {
        const xmc_layer& L = in->layerInfo[l];
        if (L.absorptionCoefficient == 0.0 && L.scatteringCoefficient == 0.0) continue;  // glass: no scattering
        int k = -1;
        for (size_t t = 0; t < tables.size(); t++)
            if (tables[t].anisotropy == L.anisotropy) k = (int)t;
        if (k < 0) return fail(XMC_ERR_ARG, "no weight table for the anisotropy of layer " + std::to_string(l));
        B.off[l] = first[k];
        B.last[l] = tables[k].numberOfElements - 1;
        B.half_n[l] = 0.5f * (float)B.last[l];
    }
    HIP_TRY(sc->bias_table.alloc(flat.size()));
    HIP_TRY(hipMemcpy(sc->bias_table.p, flat.data(), flat.size() * sizeof(float), hipMemcpyHostToDevice));
    B.table = sc->bias_table.p;
    B.on = 1;
    B.factor = (float)in->attractiveFactor;
    for (int a = 0; a < 3; a++) {
        const double t = a == 0 ? in->targetPoint.x : (a == 1 ? in->targetPoint.y : in->targetPoint.z);
        B.rd[a] = (float)t - P.start_pos[a];
    }
    for (int l = 0; l < in->numberOfLayers; l++) {  // the float steps of :329-347
        float g = P.layers.g[l], mus = P.layers.mus[l], mua = P.layers.mua[l];
        float D = 1.0f / (3.0f * ((1.0f - g) * mus + mua));
        float ltr = 3.0f * D;
        float delta = 0.74f * ltr;
        B.r1z[l] = B.rd[2] + ltr;
        B.r2z[l] = B.rd[2] - (ltr + 2.0f * delta);
    }
    return XMC_OK;
}

// critical cosine in float arithmetic (GetCriticalCos, mcml_kernel.cpp:644-661)
float crit_cos(float n1, float n2) {
    if (n1 > n2) return std::sqrt(1.0f - n2 * n2 / (n1 * n1));
    return 0.0f;
}

}  // namespace

extern "C" {

const char* xmc_last_error(void) { return g_error.c_str(); }
const char* xmc_version(void) { return "xmcml_amd 0.1 gfx950"; }

double xmc_specular_reflectance(const xmc_layer* layer) {  // mcml_kernel.cpp:16-37
    double t = (layer[0].refractiveIndex - layer[1].refractiveIndex) /
               (layer[0].refractiveIndex + layer[1].refractiveIndex);
    double r1 = t * t;
    if (layer[1].absorptionCoefficient == 0.0 && layer[1].scatteringCoefficient == 0.0) {
        t = (layer[1].refractiveIndex - layer[2].refractiveIndex) /
            (layer[1].refractiveIndex + layer[2].refractiveIndex);
        double r2 = t * t;
        t = 1 - r1;
        r1 = r1 + t * t * r2 / (1 - r1 * r2);
    }
    return r1;
}

static int scene_init(const xmc_input* in, const xmc_options* opt, bool launchable, xmc_scene** out);

int xmc_scene_create(const xmc_input* in, const xmc_options* opt, xmc_scene** out) {
    return scene_init(in, opt, true, out);
}

// The surface rule of a scene (Planar and Lattice in xmc_params.h): every
// surface an axis-aligned rectangle or a heightfield on one shared lattice.
// planar gets the rectangles in surface order, each with its first triangle in
// leaf order; lattice gets the heightfields' cell records (uploaded to `cells`).
// Returns false, with both forms empty, for any other scene (or for a scene with
// heightfields when with_lattice is false).
static int surface_rule(const xmc_input* in, const xmc::GpuBvh& gb, const std::vector<int>& leaf_of,
                        bool with_lattice, DevBuf<float>& cells, xmc::Params& P, bool* on) {
    *on = false;
    std::memset(&P.planar, 0, sizeof(P.planar));
    std::memset(&P.lattice, 0, sizeof(P.lattice));
    P.surface_rule = 0;
    const int ns = in->numberOfSurfaces;
    std::vector<int> rect_of(ns, -1), heights;
    xmc::Planar pl;
    std::memset(&pl, 0, sizeof(pl));
    for (int s = 0; s < ns; s++) {
        xmc::RectForm r;
        if (xmc::planar_rectangle(in->surface[s], r)) {
            if (pl.n == xmc::kPlanarMax) return XMC_OK;
            rect_of[s] = pl.n;
            pl.z[pl.n] = r.z;
            pl.rect[pl.n] = make_float4(r.x0, r.x1, r.y0, r.y1);
            pl.tri[pl.n] = -1;
            pl.n++;
        } else {
            heights.push_back(s);
        }
    }
    xmc::LatticeForm lf;
    if (!heights.empty() && (!with_lattice || !xmc::lattice_form(in->surface, ns, heights, leaf_of, lf)))
        return XMC_OK;
    for (size_t k = 0; k < gb.tris.size(); k++) {
        const int r = rect_of[gb.tris[k].surface];
        if (r >= 0 && pl.tri[r] < 0) pl.tri[r] = (int)k;
    }
    P.planar = pl;
    if (lf.ns > 0) {
        HIP_TRY(cells.alloc(lf.cells.size()));
        HIP_TRY(hipMemcpy(cells.p, lf.cells.data(), lf.cells.size() * sizeof(float), hipMemcpyHostToDevice));
        xmc::Lattice& L = P.lattice;
        L.ns = lf.ns;
        L.nx = lf.nx;
        L.ny = lf.ny;
        L.x0 = lf.x0;
        L.y0 = lf.y0;
        L.inv_dx = lf.inv_dx;
        L.inv_dy = lf.inv_dy;
        L.mxy = lf.mxy;
        L.mz = lf.mz;
        L.zlo = lf.zlo;
        L.zhi = lf.zhi;
        L.cells = reinterpret_cast<const float4*>(cells.p);
    }
    P.surface_rule = lf.ns > 0 ? 2 : 1;
    *on = true;
    return XMC_OK;
}

// Persistent launch shape and the stack spill it needs.  kernel_occupancy is
// the occupancy query of the kernel that will run.
static int configure_launch(int device, int depth, int block, int blocks_per_cu,
                            hipError_t (*kernel_occupancy)(int, size_t, int*), int* blocks_out,
                            DevBuf<int2>& spill, xmc::Params& P) {
    if (block <= 0) block = 256;
    if (block % 64 != 0 || block > 256) return fail(XMC_ERR_ARG, "blockSize must be 64, 128, 192 or 256");
    // a single scene under the surface rule walks only the rare steps too long for its
    // rule: its stacks' rings go to global memory, leaving the LDS to the tallies (more
    // blocks per CU where an absorption tile fills it); batches keep them in LDS
    const bool ring_global = kernel_occupancy != xmc_batch_occupancy && P.surface_rule != 0;
    size_t lds = transport_lds(block, P.abs_tile.count, P.det_lds, !ring_global);
    int per_cu = 0;
    HIP_TRY(kernel_occupancy(block, lds, &per_cu));
    if (per_cu < 1) per_cu = 1;
    if (blocks_per_cu > 0 && blocks_per_cu < per_cu) per_cu = blocks_per_cu;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    *blocks_out = prop.multiProcessorCount * per_cu;
    // the walk holds at most depth - 1 entries; the top kStackLds live in LDS
    int need = depth - xmc::kStackLds;
    P.spill_depth = need > 0 ? need : 0;
    size_t lanes = (size_t)(*blocks_out) * block;
    const size_t ring = ring_global ? (size_t)xmc::kStackLds * lanes : 0;
    HIP_TRY(spill.alloc(std::max<size_t>((size_t)P.spill_depth * lanes + ring, 1)));
    P.spill = spill.p;
    P.ring = ring_global ? spill.p + (size_t)P.spill_depth * lanes : nullptr;
    return XMC_OK;
}

// The occupancy query of the transport kernel xmc_launch_transport picks for
// this scene (interleaved when it has a walk-skipping grid; the merging plain variant
// when it tallies an absorption grid).
static auto transport_occupancy_of(const xmc::Params& P) -> hipError_t (*)(int, size_t, int*) {
    if (P.dist.cells != nullptr) return xmc_transport_occupancy;
    return P.tally.absorption != nullptr ? xmc_transport_occupancy_merge : xmc_transport_occupancy_plain;
}

// One trajectory list of traj_cap voxel indices per persistent lane.
static int alloc_lane_voxels(xmc_scene* sc) {
    xmc::Params& P = sc->base;
    if (!P.paths.det_traj) return XMC_OK;
    size_t lanes = (size_t)sc->blocks * (size_t)(sc->block > 0 ? sc->block : 256);
    HIP_TRY(sc->lane_vox.alloc(lanes * (size_t)sc->traj_cap));
    P.paths.lane_voxels = sc->lane_vox.p;
    return XMC_OK;
}

static int scene_init(const xmc_input* in, const xmc_options* opt, bool launchable, xmc_scene** out) {
    if (!out) return fail(XMC_ERR_ARG, "scene pointer is NULL");
    *out = nullptr;
    int rc = validate(in);
    if (rc) return rc;
    xmc_options o;
    std::memset(&o, 0, sizeof(o));
    if (opt) o = *opt;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(XMC_ERR_NOGPU, "no HIP device");
    if (o.device < 0 || o.device >= ndev) return fail(XMC_ERR_ARG, "device ordinal out of range");
    HIP_TRY(hipSetDevice(o.device));

    auto* sc = new xmc_scene();
    std::unique_ptr<xmc_scene> guard(sc);
    sc->device = o.device;
    sc->n_layers = in->numberOfLayers;
    sc->n_surfaces = in->numberOfSurfaces;
    sc->n_det = in->numberOfCubeDetectors + in->numberOfRingDetectors;
    sc->T = in->timeScaleSize;
    const xmc_int3& part = in->area->partitionNumber;
    sc->grid = part.x * part.y * part.z;
    sc->absorption = o.skipAbsorption == 0;
    sc->rsp = xmc_specular_reflectance(in->layerInfo);

    // BVH (GenerateBVHTree) and its GPU layout
    const char* err = nullptr;
    if (!xmc::build_reference_bvh(in->surface, in->numberOfSurfaces, sc->ref, &err))
        return fail(XMC_ERR_ARG, std::string("BVH: ") + (err ? err : "failed"));
    xmc::GpuBvh gb;
    xmc::flatten_bvh(sc->ref, in->surface, gb);
    size_t nt = gb.tris.size();
    {
        std::vector<int> base(in->numberOfSurfaces + 1, 0);
        for (int i = 0; i < in->numberOfSurfaces; i++) base[i + 1] = base[i] + in->surface[i].numberOfTriangles;
        sc->perm.assign(nt, 0);
        for (size_t k = 0; k < nt; k++) sc->perm[base[sc->ref.tri_surface[k]] + sc->ref.tri_index[k]] = (int)k;
    }
    HIP_TRY(sc->nodes.alloc(gb.nodes.size() > 0 ? gb.nodes.size() : 1));
    if (!gb.nodes.empty())
        HIP_TRY(hipMemcpy(sc->nodes.p, gb.nodes.data(), gb.nodes.size() * sizeof(xmc::GpuNode), hipMemcpyHostToDevice));
    HIP_TRY(sc->tris.alloc(nt));
    HIP_TRY(hipMemcpy(sc->tris.p, gb.tris.data(), nt * sizeof(xmc::GpuTri), hipMemcpyHostToDevice));
    HIP_TRY(sc->perm_dev.alloc(nt));
    HIP_TRY(hipMemcpy(sc->perm_dev.p, sc->perm.data(), nt * sizeof(int), hipMemcpyHostToDevice));
    {
        // unit normals with the kernel's float operations (cross and dot with fused
        // multiply-adds, sqrt, three divisions): the bits the kernel would compute at each hit
        std::vector<float> nrm(4 * nt, 0.0f);
        for (size_t k = 0; k < nt; k++) {
            const xmc::GpuTri& t = gb.tris[k];
            float cx = std::fmaf(t.e1[1], t.e2[2], -(t.e1[2] * t.e2[1]));
            float cy = std::fmaf(t.e1[2], t.e2[0], -(t.e1[0] * t.e2[2]));
            float cz = std::fmaf(t.e1[0], t.e2[1], -(t.e1[1] * t.e2[0]));
            float l = std::sqrt(std::fmaf(cz, cz, std::fmaf(cy, cy, cx * cx)));
            nrm[4 * k] = cx / l;
            nrm[4 * k + 1] = cy / l;
            nrm[4 * k + 2] = cz / l;
        }
        HIP_TRY(sc->normals.alloc(4 * nt));
        HIP_TRY(hipMemcpy(sc->normals.p, nrm.data(), nrm.size() * sizeof(float), hipMemcpyHostToDevice));
    }

    xmc::Params& P = sc->base;
    std::memset(&P, 0, sizeof(P));
    P.nodes = sc->nodes.p;
    P.tris = sc->tris.p;
    P.normals = sc->normals.p;
    bool rule = false;
    if ((rc = surface_rule(in, gb, sc->perm, o.noLatticeRule == 0, sc->lattice_cells, P, &rule))) return rc;
    P.root_ref = gb.root_ref;
    for (int a = 0; a < 3; a++) {
        P.root_lo[a] = gb.root_lo[a];
        P.root_hi[a] = gb.root_hi[a];
    }

    // walk-skipping grid: per cell a distance bound (skip walks that cannot hit) and an
    // entry (short walks run on the cell's restricted tree)
    // a scene under the surface rule has no grid: its plain kernel ran slower with the
    // distance bound ahead of the rule (one more dependent load per step); its few walks
    // (steps too long for the lattice) start at the root
    const bool want_dist = o.noDistanceGrid == 0 && !rule;
    const bool want_entry = o.noEntryGrid == 0 && gb.root_ref >= 0 && !rule;
    if (nt > 256 && (want_dist || want_entry)) {
        xmc::DistanceGrid dg;
        double lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            double pad = 1e-3 * (gb.root_hi[a] - gb.root_lo[a]) + 1e-3;
            lo[a] = gb.root_lo[a] - pad;
            hi[a] = gb.root_hi[a] + pad;
        }
        // about 256 cells per triangle, at most 32M (measured on the heightfield: 12.6M cells
        // 3.37e7 photons/s, 25M 3.42e7, 50M 3.43e7; 6.3M 3.27e7)
        size_t target = std::min<size_t>(std::max<size_t>((size_t)nt * 256, (size_t)1 << 18), (size_t)32 << 20);
        xmc::build_distance_grid(in->surface, in->numberOfSurfaces, lo, hi, target, dg);
        xmc::EntryGrid eg;
        if (want_entry)
            xmc::build_entry_grid(gb, dg.lo, dg.h, dg.dims, want_dist ? dg.q.data() : nullptr, dg.step_q, eg);
        std::vector<uint32_t> cells(dg.q.size());
        for (size_t c = 0; c < cells.size(); c++) {
            uint32_t q = want_dist ? dg.q[c] : 0u;
            uint32_t e = !want_entry ? xmc::kEntryRoot : (uint32_t)(eg.cell[c] + 1);
            cells[c] = q | (e << 8);
        }
        if (eg.ref.size() >= xmc::kEntryRoot - 1) return fail(XMC_ERR_LIMIT, "too many BVH entry nodes");
        if (want_entry) {
            std::vector<uint32_t> cells2(eg.cell2.size());
            for (size_t c = 0; c < cells2.size(); c++) cells2[c] = (uint32_t)(eg.cell2[c] + 1);
            HIP_TRY(sc->dist2.alloc(cells2.size()));
            HIP_TRY(hipMemcpy(sc->dist2.p, cells2.data(), cells2.size() * 4, hipMemcpyHostToDevice));
        }
        if (gb.nodes.size() + eg.extra.size() > (size_t)INT32_MAX) return fail(XMC_ERR_LIMIT, "too many BVH nodes");
        if (!eg.extra.empty()) {
            // the entries' restricted trees follow the BVH's own nodes in one array
            std::vector<xmc::GpuNode> all(gb.nodes);
            all.insert(all.end(), eg.extra.begin(), eg.extra.end());
            HIP_TRY(sc->nodes.alloc(all.size()));
            HIP_TRY(hipMemcpy(sc->nodes.p, all.data(), all.size() * sizeof(xmc::GpuNode), hipMemcpyHostToDevice));
            P.nodes = sc->nodes.p;
        }
        HIP_TRY(sc->dist.alloc(cells.size()));
        HIP_TRY(hipMemcpy(sc->dist.p, cells.data(), cells.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(sc->entry_ref.alloc(std::max<size_t>(eg.ref.size(), 1)));
        HIP_TRY(sc->entry_box.alloc(std::max<size_t>(eg.box.size(), 1)));
        if (!eg.ref.empty()) {
            HIP_TRY(hipMemcpy(sc->entry_ref.p, eg.ref.data(), eg.ref.size() * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(sc->entry_box.p, eg.box.data(), eg.box.size() * 4, hipMemcpyHostToDevice));
        }
        P.dist.cells = sc->dist.p;
        for (int a = 0; a < 3; a++) {
            P.dist.lo[a] = dg.lo[a];
            P.dist.dims[a] = dg.dims[a];
        }
        P.dist.h = dg.h;
        P.dist.inv_h = 1.0f / dg.h;
        P.dist.step_q = dg.step_q;
        P.dist.margin = want_entry ? eg.margin : 0.0f;
        P.dist.entry_ref = sc->entry_ref.p;
        P.dist.entry_box = sc->entry_box.p;
        P.dist.cells2 = want_entry ? sc->dist2.p : nullptr;
        P.dist.margin2 = want_entry ? eg.margin2 : 0.0f;
        if (want_dist && o.noSlabGrid == 0) {
            // slabs chosen for steps of the scene's typical length: the median mean free path
            // of its scattering layers
            std::vector<double> mfp;
            for (int l = 1; l < in->numberOfLayers; l++) {
                double mut = in->layerInfo[l].absorptionCoefficient + in->layerInfo[l].scatteringCoefficient;
                if (mut > 0.0) mfp.push_back(1.0 / mut);
            }
            std::sort(mfp.begin(), mfp.end());
            double ms = mfp.empty() ? (double)dg.h : mfp[mfp.size() / 2];
            ms = std::clamp(ms, 0.25 * (double)dg.h, 8.0 * (double)dg.h);
            xmc::SlabGrid sg;
            xmc::build_slab_grid(in->surface, in->numberOfSurfaces, dg, (float)ms, sg);
            HIP_TRY(sc->slabs.alloc(sg.words.size()));
            HIP_TRY(hipMemcpy(sc->slabs.p, sg.words.data(), sg.words.size() * 8, hipMemcpyHostToDevice));
            P.dist.slabs = sc->slabs.p;
            P.dist.slab_unit = sg.unit;
        }
    }

    // layer boundary table: FindIntersectionLayer (mcml_kernel.cpp:566-580)
    std::vector<int8_t> nl((size_t)sc->n_layers * sc->n_surfaces, -1);
    for (int l = 0; l < sc->n_layers; l++)
        for (int s = 0; s < sc->n_surfaces; s++) {
            int found = -1;
            for (int i = 0; i < in->numberOfLayers && found < 0; ++i)
                for (int j = 0; j < in->layerInfo[i].numberOfSurfaces; ++j)
                    if (in->layerInfo[i].surfaceId[j] == s && i != l) {
                        found = i;
                        break;
                    }
            nl[(size_t)l * sc->n_surfaces + s] = (int8_t)found;
        }
    HIP_TRY(sc->next_layer.alloc(nl.size()));
    HIP_TRY(hipMemcpy(sc->next_layer.p, nl.data(), nl.size(), hipMemcpyHostToDevice));
    P.next_layer = sc->next_layer.p;
    P.n_surfaces = sc->n_surfaces;

    for (int l = 0; l < XMC_MAX_LAYERS; l++) {
        if (l < sc->n_layers) {
            const xmc_layer& L = in->layerInfo[l];
            P.layers.n[l] = (float)L.refractiveIndex;
            P.layers.mua[l] = (float)L.absorptionCoefficient;
            P.layers.mus[l] = (float)L.scatteringCoefficient;
            P.layers.g[l] = (float)L.anisotropy;
            // per-layer constants of the step and the deposit, in the kernel's float operations
            float mut = P.layers.mua[l] + P.layers.mus[l];
            P.layers.inv_mut[l] = 1.0f / mut;
            P.layers.abs_frac[l] = P.layers.mua[l] / mut;
            if (L.absorptionCoefficient == 0.0 && L.scatteringCoefficient == 0.0) P.layers.glass_mask |= 1u << l;
        } else {
            P.layers.n[l] = 1.0f;
        }
        if (in->targetRangeLayers[l]) P.layers.target_mask |= 1u << l;
    }
    for (int a = 0; a < XMC_MAX_LAYERS; a++)
        for (int b = 0; b < XMC_MAX_LAYERS; b++) P.layers.crit[a][b] = crit_cos(P.layers.n[a], P.layers.n[b]);
    P.layers.count = sc->n_layers;

    P.area.corner[0] = (float)in->area->corner.x;
    P.area.corner[1] = (float)in->area->corner.y;
    P.area.corner[2] = (float)in->area->corner.z;
    P.area.length[0] = (float)in->area->length.x;
    P.area.length[1] = (float)in->area->length.y;
    P.area.length[2] = (float)in->area->length.z;
    P.area.part[0] = part.x;
    P.area.part[1] = part.y;
    P.area.part[2] = part.z;
    for (int a = 0; a < 3; a++) {
        P.area.partf[a] = (float)P.area.part[a];
        P.area.scale[a] = P.area.partf[a] / P.area.length[a];
    }

    // detectors (cube bounds in float arithmetic as the float walk evaluates them)
    std::vector<float> cube, ring;
    std::vector<uint32_t> filt;
    for (int i = 0; i < in->numberOfCubeDetectors; i++) {
        const xmc_cube_detector& c = in->cubeDetector[i];
        float cx = (float)c.center.x, cy = (float)c.center.y, cz = (float)c.center.z;
        float hx = (float)c.length.x / 2.0f, hy = (float)c.length.y / 2.0f, hz = (float)c.length.z / 2.0f;
        cube.insert(cube.end(), {cx - hx, cy - hy, cz - hz, cx + hx, cy + hy, cz + hz,
                                 (float)std::cos(c.targetAngle * kPi / 180.0)});
        uint32_t m = 0;
        for (int j = 0; j < XMC_MAX_LAYERS; j++)
            if (c.filterLayers[j]) m |= 1u << j;
        filt.push_back(m);
    }
    for (int i = 0; i < in->numberOfRingDetectors; i++) {
        const xmc_ring_detector& r = in->ringDetector[i];
        ring.insert(ring.end(), {(float)r.center.x, (float)r.center.y, (float)r.smallRadius, (float)r.bigRadius,
                                 (float)std::cos(r.targetAngle * kPi / 180.0)});
        uint32_t m = 0;
        for (int j = 0; j < XMC_MAX_LAYERS; j++)
            if (r.filterLayers[j]) m |= 1u << j;
        filt.push_back(m);
    }
    HIP_TRY(sc->cube.alloc(cube.size() ? cube.size() : 1));
    HIP_TRY(sc->ring.alloc(ring.size() ? ring.size() : 1));
    HIP_TRY(sc->filter.alloc(filt.size() ? filt.size() : 1));
    if (!cube.empty()) HIP_TRY(hipMemcpy(sc->cube.p, cube.data(), cube.size() * 4, hipMemcpyHostToDevice));
    if (!ring.empty()) HIP_TRY(hipMemcpy(sc->ring.p, ring.data(), ring.size() * 4, hipMemcpyHostToDevice));
    if (!filt.empty()) HIP_TRY(hipMemcpy(sc->filter.p, filt.data(), filt.size() * 4, hipMemcpyHostToDevice));
    P.det.n_cube = in->numberOfCubeDetectors;
    P.det.n_ring = in->numberOfRingDetectors;
    P.det_lds = (sc->n_det > 0 && sc->n_det <= kDetLds) ? sc->n_det : 0;
    P.det.cube = sc->cube.p;
    P.det.ring = sc->ring.p;
    P.det.filter = sc->filter.p;

    // time bins exactly as InitOutput lays them out (launcher_omp.cpp:38-55)
    int T = sc->T;
    double dt = (in->timeFinish - in->timeStart) / T;
    if (dt < 0.0) dt = 0.0;
    sc->h_t_start.assign(T, 0.0);
    sc->h_t_finish.assign(T, 0.0);
    for (int j = 0; j < T - 1; j++) {
        sc->h_t_start[j] = in->timeStart + j * dt;
        sc->h_t_finish[j] = sc->h_t_start[j] + dt;
    }
    sc->h_t_start[T - 1] = in->timeStart + (T - 1) * dt;
    sc->h_t_finish[T - 1] = in->timeFinish;
    HIP_TRY(sc->t_start.alloc(T));
    HIP_TRY(sc->t_finish.alloc(T));
    HIP_TRY(hipMemcpy(sc->t_start.p, sc->h_t_start.data(), T * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(sc->t_finish.p, sc->h_t_finish.data(), T * 8, hipMemcpyHostToDevice));
    P.det.T = T;
    P.det.t0 = in->timeStart;
    P.det.dt = dt;
    P.det.inv_dt = dt > 0.0 ? 1.0 / dt : 0.0;
    P.det.t_start = sc->t_start.p;
    P.det.t_finish = sc->t_finish.p;

    // tallies in one zeroed slab of 64-bit words
    size_t w = 0;
    sc->off_abs = w;
    w += sc->absorption ? (size_t)sc->grid : 0;
    sc->off_dw = w;
    w += sc->n_det;
    sc->off_dt = w;
    w += sc->n_det;
    sc->off_dc = w;
    w += sc->n_det;
    sc->off_bw = w;
    w += (size_t)sc->n_det * T;
    sc->off_bt = w;
    w += (size_t)sc->n_det * T;
    sc->off_bc = w;
    w += (size_t)sc->n_det * T;
    sc->off_st = w;
    w += 16;
    sc->off_q = w;
    w += 1;
    sc->slab_words = w;
    HIP_TRY(sc->slab.alloc(w));
    HIP_TRY(hipMemset(sc->slab.p, 0, w * 8));
    unsigned long long* s = sc->slab.p;
    P.tally.absorption = sc->absorption ? reinterpret_cast<double*>(s + sc->off_abs) : nullptr;
    P.tally.det_weight = reinterpret_cast<double*>(s + sc->off_dw);
    P.tally.det_target = reinterpret_cast<double*>(s + sc->off_dt);
    P.tally.det_count = s + sc->off_dc;
    P.tally.bin_weight = reinterpret_cast<double*>(s + sc->off_bw);
    P.tally.bin_target = reinterpret_cast<double*>(s + sc->off_bt);
    P.tally.bin_count = s + sc->off_bc;
    P.tally.stats = s + sc->off_st;
    P.queue = s + sc->off_q;

    // LDS absorption tile: up to kAbsTileVoxels voxels from the beam's entry voxel down,
    // centred on it across (most of the weight is deposited within a few voxels of it)
    std::memset(&P.abs_tile, 0, sizeof(P.abs_tile));
    if (launchable && sc->absorption && P.dist.cells == nullptr) {  // the kernel variant without the grid
        const xmc_area* ar = in->area;
        const int pn[3] = {part.x, part.y, part.z};
        const double sp[3] = {in->startPosition.x, in->startPosition.y, in->startPosition.z};
        const double co[3] = {ar->corner.x, ar->corner.y, ar->corner.z};
        const double ln[3] = {ar->length.x, ar->length.y, ar->length.z};
        int size[3];
        size[2] = std::min(pn[2], 12);
        int across = (int)std::sqrt((double)kAbsTileVoxels / size[2]);
        size[0] = std::min(pn[0], across);
        size[1] = std::min(pn[1], across);
        for (int a = 0; a < 3; a++) {
            int c = (int)std::floor(pn[a] * (sp[a] - co[a]) / ln[a]);
            c = std::min(std::max(c, 0), pn[a] - 1);
            int lo = a < 2 ? c - size[a] / 2 : c - 1;
            P.abs_tile.lo[a] = std::min(std::max(lo, 0), pn[a] - size[a]);
            P.abs_tile.size[a] = size[a];
        }
        P.abs_tile.count = size[0] * size[1] * size[2];
    }

    // scattering grid and detector trajectory maps (mcml_kernel.cpp:750-810)
    if (o.trajectoryCapacity < 0) return fail(XMC_ERR_ARG, "trajectoryCapacity must be >= 0");
    std::memset(&P.paths, 0, sizeof(P.paths));
    sc->traj_cap = o.trajectoryCapacity > 0 ? o.trajectoryCapacity : 20480;
    P.paths.grid_size = sc->grid;
    P.paths.capacity = sc->traj_cap;
    if (o.scattering) {
        HIP_TRY(sc->scat.alloc(sc->grid));
        HIP_TRY(hipMemset(sc->scat.p, 0, (size_t)sc->grid * 8));
        P.paths.scattering = sc->scat.p;
    }
    if (o.trajectories && sc->n_det > 0) {
        HIP_TRY(sc->dtraj.alloc((size_t)sc->n_det * sc->grid));
        HIP_TRY(hipMemset(sc->dtraj.p, 0, (size_t)sc->n_det * sc->grid * 8));
        P.paths.det_traj = sc->dtraj.p;
    }

    // photons and streams
    const uint64_t jump0 = mcg_pow(xmc::kMcgA, 1ull << xmc::kStreamLog2) & xmc::kMcgMask;
    uint64_t j = jump0;
    for (int i = 0; i < xmc::kJumpBits; i++) {
        P.jump[i] = j;
        j = (j * j) & xmc::kMcgMask;
    }
    P.start_pos[0] = (float)in->startPosition.x;
    P.start_pos[1] = (float)in->startPosition.y;
    P.start_pos[2] = (float)in->startPosition.z;
    P.start_dir[0] = (float)in->startDirection.x;
    P.start_dir[1] = (float)in->startDirection.y;
    P.start_dir[2] = (float)in->startDirection.z;
    P.w0 = (float)(1.0 - sc->rsp);
    P.min_weight = (float)in->minWeight;
    P.chance = (float)in->rouletteChance;
    P.roulette = in->terminationMode == XMC_TERMINATE_ROULETTE;
    if (in->useBiasing && (rc = setup_bias(in, sc, P))) return rc;

    sc->block = o.blockSize > 0 ? o.blockSize : 256;
    if (launchable) {
        rc = configure_launch(sc->device, sc->ref.depth, sc->block, o.blocksPerCU, transport_occupancy_of(P),
                              &sc->blocks, sc->spill, P);
        if (rc) return rc;
        if ((rc = alloc_lane_voxels(sc))) return rc;
        HIP_TRY(hipStreamCreate(&sc->stream));
    }
    *out = guard.release();
    return XMC_OK;
}

int xmc_scene_set_launch(xmc_scene* sc, int block, int blocks_per_cu) {
    if (!sc || !sc->stream) return fail(XMC_ERR_ARG, "scene is NULL or not launchable");
    HIP_TRY(hipSetDevice(sc->device));
    HIP_TRY(hipStreamSynchronize(sc->stream));
    int blocks = 0;
    int rc = configure_launch(sc->device, sc->ref.depth, block, blocks_per_cu, transport_occupancy_of(sc->base),
                              &blocks, sc->spill, sc->base);
    if (rc) return rc;
    sc->block = block;
    sc->blocks = blocks;
    return alloc_lane_voxels(sc);
}

int xmc_scene_get_launch(xmc_scene* sc, int* block, int* blocks) {
    if (!sc || !block || !blocks) return fail(XMC_ERR_ARG, "NULL argument");
    *block = sc->block;
    *blocks = sc->blocks;
    return XMC_OK;
}

int xmc_scene_destroy(xmc_scene* sc) {
    if (!sc) return XMC_OK;
    (void)hipSetDevice(sc->device);
    if (sc->stream) (void)hipStreamSynchronize(sc->stream);
    for (auto& e : sc->pending) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (sc->stream) (void)hipStreamDestroy(sc->stream);
    delete sc;
    return XMC_OK;
}

static int run_impl(xmc_scene* sc, uint64_t seed, uint64_t first, uint64_t n, hipStream_t s, const xmc::Trace* tr) {
    if (n == 0) return XMC_OK;
    if (first + n < first || first + n > (1ull << xmc::kJumpBits))
        return fail(XMC_ERR_LIMIT, "photon index beyond 2^33 (a seed has 2^33 disjoint streams; use another seed)");
    HIP_TRY(hipSetDevice(sc->device));
    xmc::Params P = sc->base;
    P.first_photon = first;
    P.n_photons = n;
    P.seed_value = (2 * seed + 1) & xmc::kMcgMask;
    if (tr) P.trace = *tr;
    HIP_TRY(xmc_launch_zero(P.queue, 1, s));
    size_t lds = transport_lds(sc->block, P.abs_tile.count, P.det_lds, P.ring == nullptr);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    sc->pending.emplace_back(e0, e1);
    HIP_TRY(hipEventRecord(e0, s));
    HIP_TRY(xmc_launch_transport(&P, sc->blocks, sc->block, lds, s, tr != nullptr));
    HIP_TRY(hipEventRecord(e1, s));
    return XMC_OK;
}

int xmc_scene_run(xmc_scene* sc, uint64_t seed, uint64_t first, uint64_t n, void* stream) {
    if (!sc) return fail(XMC_ERR_ARG, "scene is NULL");
    hipStream_t s = stream ? (hipStream_t)stream : sc->stream;
    return run_impl(sc, seed, first, n, s, nullptr);
}

int xmc_scene_sync(xmc_scene* sc) {
    if (!sc) return fail(XMC_ERR_ARG, "scene is NULL");
    HIP_TRY(hipSetDevice(sc->device));
    HIP_TRY(hipDeviceSynchronize());
    return XMC_OK;
}

static int collect_impl(xmc_scene* sc, xmc_output* out, hipStream_t stream);

int xmc_scene_collect(xmc_scene* sc, xmc_output* out) {
    if (!sc || !out) return fail(XMC_ERR_ARG, "scene/output is NULL");
    HIP_TRY(hipSetDevice(sc->device));
    HIP_TRY(hipStreamSynchronize(sc->stream));
    return collect_impl(sc, out, sc->stream);
}

static int collect_impl(xmc_scene* sc, xmc_output* out, hipStream_t stream) {
    HIP_TRY(hipDeviceSynchronize());
    if (out->numberOfDetectors != sc->n_det) return fail(XMC_ERR_ARG, "output detector count does not match input");
    std::vector<unsigned long long> h(sc->slab_words);
    HIP_TRY(hipMemcpy(h.data(), sc->slab.p, sc->slab_words * 8, hipMemcpyDeviceToHost));
    auto dv = [&](size_t off) { return reinterpret_cast<const double*>(h.data() + off); };
    if (sc->absorption && out->absorption) {
        if (out->gridSize != sc->grid) return fail(XMC_ERR_ARG, "output gridSize does not match area");
        const double* a = dv(sc->off_abs);
        for (int i = 0; i < sc->grid; i++) out->absorption[i] += a[i];
    }
    for (int d = 0; d < sc->n_det; d++) {
        xmc_detector_info& di = out->detectorInfo[d];
        di.weight += dv(sc->off_dw)[d];
        di.targetWeight += dv(sc->off_dt)[d];
        di.numberOfPhotons += h[sc->off_dc + d];
        for (int b = 0; b < sc->T && b < di.timeScaleSize; b++) {
            size_t k = (size_t)d * sc->T + b;
            di.timeScale[b].weight += dv(sc->off_bw)[k];
            di.timeScale[b].targetWeight += dv(sc->off_bt)[k];
            di.timeScale[b].numberOfPhotons += h[sc->off_bc + k];
        }
    }
    const size_t G = (size_t)sc->grid;
    if (sc->scat.p) {
        if (out->scattering) {
            if (out->gridSize != sc->grid) return fail(XMC_ERR_ARG, "output gridSize does not match area");
            std::vector<double> g(G);
            HIP_TRY(hipMemcpy(g.data(), sc->scat.p, G * 8, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < G; i++) out->scattering[i] += g[i];
        }
        HIP_TRY(xmc_launch_zero(reinterpret_cast<unsigned long long*>(sc->scat.p), G, stream));
    }
    if (sc->dtraj.p) {
        std::vector<unsigned long long> g(G);
        for (int d = 0; d < sc->n_det; d++) {
            uint64_t* t = out->detectorInfo[d].trajectory;
            if (!t) continue;
            if (out->detectorInfo[d].trajectorySize != sc->grid)
                return fail(XMC_ERR_ARG, "detector trajectorySize does not match area");
            HIP_TRY(hipMemcpy(g.data(), sc->dtraj.p + d * G, G * 8, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < G; i++) t[i] += g[i];
        }
        HIP_TRY(xmc_launch_zero(sc->dtraj.p, (size_t)sc->n_det * G, stream));
    }
    out->specularReflectance = sc->rsp;
    // zero everything except the counters of xmc_scene_stats
    HIP_TRY(xmc_launch_zero(sc->slab.p, sc->off_st, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return XMC_OK;
}

static int read_stats(unsigned long long* dev, hipStream_t stream, xmc_stats* st) {
    unsigned long long c[16];
    HIP_TRY(hipMemcpy(c, dev, sizeof(c), hipMemcpyDeviceToHost));
    st->photons = c[0];
    st->scatterEvents = c[1];
    st->boundaryEvents = c[2];
    st->rouletteEvents = c[3];
    st->stackSpills = c[4];
    st->culledWalks = c[5];
    double led[8];
    std::memcpy(led, c + 8, sizeof(led));
    st->wAbsorbed = led[0];
    st->wEscaped = led[1];
    st->wDetected = led[2];
    st->wLeftArea = led[3];
    st->wKilled = led[4];
    st->wRouletteLost = led[5];
    st->wRouletteGain = led[6];
    st->wOther = led[7];
    st->trajectoryOverflows = c[6];
    st->entryWalks = c[7];
    HIP_TRY(xmc_launch_zero(dev, 16, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return XMC_OK;
}

static int drain_timers(std::vector<std::pair<hipEvent_t, hipEvent_t>>& pending, double* ms_total) {
    for (auto& e : pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
        *ms_total += ms;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    pending.clear();
    return XMC_OK;
}

int xmc_scene_stats(xmc_scene* sc, xmc_stats* st) {
    if (!sc || !st) return fail(XMC_ERR_ARG, "scene/stats is NULL");
    HIP_TRY(hipSetDevice(sc->device));
    HIP_TRY(hipDeviceSynchronize());
    double ms = sc->kernel_ms;
    int rc = drain_timers(sc->pending, &ms);
    if (rc) return rc;
    rc = read_stats(sc->slab.p + sc->off_st, sc->stream, st);
    st->kernelMs = ms;
    sc->kernel_ms = 0.0;
    return rc;
}

int xmc_scene_trace(xmc_scene* sc, uint64_t seed, uint64_t first, uint64_t n, xmc_trace_event* events,
                    int max_events, int* counts) {
    if (!sc || !events || !counts || max_events <= 0) return fail(XMC_ERR_ARG, "bad trace arguments");
    HIP_TRY(hipSetDevice(sc->device));
    DevBuf<xmc_trace_event> ev;
    DevBuf<int> cnt;
    HIP_TRY(ev.alloc(n * (size_t)max_events));
    HIP_TRY(cnt.alloc(n));
    xmc::Trace tr;
    tr.events = ev.p;
    tr.counts = cnt.p;
    tr.max_events = max_events;
    tr.first = first;
    tr.n = n;
    int rc = run_impl(sc, seed, first, n, sc->stream, &tr);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(sc->stream));
    HIP_TRY(hipMemcpy(events, ev.p, n * (size_t)max_events * sizeof(xmc_trace_event), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(counts, cnt.p, n * sizeof(int), hipMemcpyDeviceToHost));
    return XMC_OK;
}

int xmc_scene_query(xmc_scene* sc, const double* o, const double* d, const double* steps, int nr, int brute,
                    int* found, double* dist, int* surf, int* tri, double* normal) {
    if (!sc || nr < 0) return fail(XMC_ERR_ARG, "bad query arguments");
    if (nr == 0) return XMC_OK;
    HIP_TRY(hipSetDevice(sc->device));
    std::vector<float> ho(3 * (size_t)nr), hd(3 * (size_t)nr), hs(nr);
    for (size_t i = 0; i < 3 * (size_t)nr; i++) {
        ho[i] = (float)o[i];
        hd[i] = (float)d[i];
    }
    for (int i = 0; i < nr; i++) hs[i] = (float)steps[i];
    DevBuf<float> fo, fd, fs, fdist, fn;
    DevBuf<int> ff, fsurf, ftri;
    HIP_TRY(fo.alloc(3 * (size_t)nr));
    HIP_TRY(fd.alloc(3 * (size_t)nr));
    HIP_TRY(fs.alloc(nr));
    HIP_TRY(fdist.alloc(nr));
    HIP_TRY(fn.alloc(3 * (size_t)nr));
    HIP_TRY(ff.alloc(nr));
    HIP_TRY(fsurf.alloc(nr));
    HIP_TRY(ftri.alloc(nr));
    HIP_TRY(hipMemcpy(fo.p, ho.data(), ho.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(fd.p, hd.data(), hd.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(fs.p, hs.data(), hs.size() * 4, hipMemcpyHostToDevice));
    int block = 256;
    int grid = (nr + block - 1) / block;
    int need = sc->ref.depth - xmc::kStackLds;
    int sdepth = need > 0 ? need : 0;
    DevBuf<int2> spill;
    HIP_TRY(spill.alloc(sdepth > 0 ? (size_t)sdepth * grid * block : 1));
    xmc::QueryParams Q;
    std::memset(&Q, 0, sizeof(Q));
    Q.nodes = sc->base.nodes;
    Q.tris = sc->base.tris;
    Q.root_ref = sc->base.root_ref;
    for (int a = 0; a < 3; a++) {
        Q.root_lo[a] = sc->base.root_lo[a];
        Q.root_hi[a] = sc->base.root_hi[a];
    }
    Q.origins = fo.p;
    Q.directions = fd.p;
    Q.steps = fs.p;
    Q.n = nr;
    Q.brute = brute;
    Q.n_tris = (int)sc->perm.size();
    Q.found = ff.p;
    Q.distance = fdist.p;
    Q.surface = fsurf.p;
    Q.triangle = ftri.p;
    Q.normal = fn.p;
    Q.spill = spill.p;
    Q.spill_depth = sdepth;
    HIP_TRY(xmc_launch_query(&Q, sc->perm_dev.p, block, (size_t)xmc::kStackLds * block * sizeof(int2), sc->stream));
    HIP_TRY(hipStreamSynchronize(sc->stream));
    std::vector<float> hdist(nr), hn(3 * (size_t)nr);
    HIP_TRY(hipMemcpy(found, ff.p, nr * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(surf, fsurf.p, nr * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(tri, ftri.p, nr * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hdist.data(), fdist.p, nr * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hn.data(), fn.p, hn.size() * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < nr; i++) dist[i] = hdist[i];
    for (size_t i = 0; i < hn.size(); i++) normal[i] = hn[i];
    return XMC_OK;
}

int xmc_scene_query_device(xmc_scene* sc, const float* o, const float* d, const float* steps, int nr, int brute,
                           int* found, float* dist, int* surf, int* tri, float* normal, void* stream) {
    if (!sc || nr < 0) return fail(XMC_ERR_ARG, "bad query arguments");
    if (nr == 0) return XMC_OK;
    if (!o || !d || !steps || !found || !dist || !surf || !tri || !normal)
        return fail(XMC_ERR_ARG, "NULL device buffer");
    HIP_TRY(hipSetDevice(sc->device));
    int2* scratch = nullptr;
    size_t entries = 0;
    HIP_TRY(xmc_query_scratch(sc->device, &scratch, &entries));
    // the stack entries beyond the LDS ones spill to the scratch, [entry][lane]
    const int need = sc->ref.depth - xmc::kStackLds;
    const int sdepth = need > 0 ? need : 0;
    int lanes = xmc::kQueryLanes;
    if (sdepth > 0 && entries / (size_t)sdepth < (size_t)lanes) lanes = (int)(entries / (size_t)sdepth / 256 * 256);
    if (lanes < 256)
        return fail(XMC_ERR_LIMIT, "BVH of depth " + std::to_string(sc->ref.depth) +
                                       ": 256 lanes' stack spill exceeds the query scratch of " +
                                       std::to_string(xmc::kQueryScratchBytes >> 20) + " MiB");
    xmc::QueryParams Q;
    std::memset(&Q, 0, sizeof(Q));
    Q.nodes = sc->base.nodes;
    Q.tris = sc->base.tris;
    Q.root_ref = sc->base.root_ref;
    for (int a = 0; a < 3; a++) {
        Q.root_lo[a] = sc->base.root_lo[a];
        Q.root_hi[a] = sc->base.root_hi[a];
    }
    Q.origins = o;
    Q.directions = d;
    Q.steps = steps;
    Q.n = nr;
    Q.brute = brute;
    Q.n_tris = (int)sc->perm.size();
    Q.found = found;
    Q.distance = dist;
    Q.surface = surf;
    Q.triangle = tri;
    Q.normal = normal;
    Q.spill = scratch;
    Q.spill_depth = sdepth;
    hipStream_t s = stream ? (hipStream_t)stream : sc->stream;
    HIP_TRY(xmc_launch_query_stride(&Q, sc->perm_dev.p, lanes, (size_t)xmc::kStackLds * 256 * sizeof(int2), s));
    return XMC_OK;
}

int xmc_scene_bvh(xmc_scene* sc, int* nn, int* root, int* nt, double* boxes, int* links, int* tsurf, int* tidx) {
    if (!sc || !nn || !root || !nt) return fail(XMC_ERR_ARG, "bad bvh arguments");
    *nn = (int)sc->ref.nodes.size();
    *root = sc->ref.root;
    *nt = (int)sc->ref.tri_surface.size();
    if (boxes && links) {
        for (size_t i = 0; i < sc->ref.nodes.size(); i++) {
            const xmc::RefNode& n = sc->ref.nodes[i];
            for (int a = 0; a < 3; a++) {
                boxes[6 * i + a] = n.lo[a];
                boxes[6 * i + 3 + a] = n.hi[a];
            }
            links[4 * i] = n.left;
            links[4 * i + 1] = n.right;
            links[4 * i + 2] = n.offset;
            links[4 * i + 3] = n.count;
        }
    }
    if (tsurf && tidx)
        for (size_t i = 0; i < sc->ref.tri_surface.size(); i++) {
            tsurf[i] = sc->ref.tri_surface[i];
            tidx[i] = sc->ref.tri_index[i];
        }
    return XMC_OK;
}

int xmc_device_name(int device, char* name, int capacity, int* cus) {
    if (!name || capacity < 1) return fail(XMC_ERR_ARG, "name buffer is NULL or empty");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(XMC_ERR_NOGPU, "no HIP device");
    if (device < 0 || device >= ndev) return fail(XMC_ERR_ARG, "device ordinal out of range");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    std::string s = std::string(prop.name) + " " + prop.gcnArchName;
    std::snprintf(name, (size_t)capacity, "%s", s.c_str());
    if (cus) *cus = prop.multiProcessorCount;
    return XMC_OK;
}

int xmc_launch(const xmc_input* in, xmc_output* out, xmc_rng* rng, const xmc_options* opt) {
    if (!in || !out || !rng) return fail(XMC_ERR_ARG, "input/output/rng is NULL");
    xmc_options o;
    std::memset(&o, 0, sizeof(o));
    if (opt) o = *opt;
    if (!out->absorption) o.skipAbsorption = 1;
    if (out->scattering) o.scattering = 1;
    for (int d = 0; d < out->numberOfDetectors; d++)
        if (out->detectorInfo && out->detectorInfo[d].trajectory) o.trajectories = 1;
    xmc_scene* sc = nullptr;
    int rc = xmc_scene_create(in, &o, &sc);
    if (rc) return rc;
    rc = xmc_scene_run(sc, rng->seed, rng->nextPhoton, in->numberOfPhotons, nullptr);
    if (rc == XMC_OK) rc = xmc_scene_collect(sc, out);
    if (rc == XMC_OK) rng->nextPhoton += in->numberOfPhotons;
    xmc_scene_destroy(sc);
    return rc;
}

// the distance grid a scene of `in` builds (over the BVH root's box, padded)
static int scene_distance_grid(const xmc_input* in, uint64_t target_cells, xmc::DistanceGrid& dg) {
    xmc::RefBvh ref;
    const char* err = nullptr;
    if (!xmc::build_reference_bvh(in->surface, in->numberOfSurfaces, ref, &err))
        return fail(XMC_ERR_ARG, std::string("BVH: ") + (err ? err : "failed"));
    xmc::GpuBvh gb;
    xmc::flatten_bvh(ref, in->surface, gb);
    double glo[3], ghi[3];
    for (int a = 0; a < 3; a++) {
        double pad = 1e-3 * (gb.root_hi[a] - gb.root_lo[a]) + 1e-3;
        glo[a] = gb.root_lo[a] - pad;
        ghi[a] = gb.root_hi[a] + pad;
    }
    xmc::build_distance_grid(in->surface, in->numberOfSurfaces, glo, ghi, target_cells ? target_cells : (16u << 20), dg);
    return XMC_OK;
}

int xmc_distance_grid(const xmc_input* in, uint64_t target_cells, float* lo, float* h, float* step_q, int* dims,
                      uint8_t* q) {
    if (!in || !lo || !h || !step_q || !dims) return fail(XMC_ERR_ARG, "NULL argument");
    xmc::DistanceGrid dg;
    int rc = scene_distance_grid(in, target_cells, dg);
    if (rc != XMC_OK) return rc;
    for (int a = 0; a < 3; a++) {
        lo[a] = dg.lo[a];
        dims[a] = dg.dims[a];
    }
    *h = dg.h;
    *step_q = dg.step_q;
    if (q) std::memcpy(q, dg.q.data(), dg.q.size());
    return XMC_OK;
}

int xmc_slab_grid(const xmc_input* in, uint64_t target_cells, float mean_step, float* lo, float* h, float* unit,
                  int* dims, uint64_t* words) {
    if (!in || !lo || !h || !unit || !dims) return fail(XMC_ERR_ARG, "NULL argument");
    if (!(mean_step > 0.0f)) return fail(XMC_ERR_ARG, "meanStep must be > 0");
    xmc::DistanceGrid dg;
    int rc = scene_distance_grid(in, target_cells, dg);
    if (rc != XMC_OK) return rc;
    for (int a = 0; a < 3; a++) {
        lo[a] = dg.lo[a];
        dims[a] = dg.dims[a];
    }
    *h = dg.h;
    if (!words) {
        *unit = dg.h * 0.125f;
        return XMC_OK;
    }
    xmc::SlabGrid sg;
    xmc::build_slab_grid(in->surface, in->numberOfSurfaces, dg, mean_step, sg);
    *unit = sg.unit;
    std::memcpy(words, sg.words.data(), sg.words.size() * 8);
    return XMC_OK;
}

int xmc_init_output(const xmc_input* in, xmc_output* out) {  // InitOutput, launcher_omp.cpp:10-58
    if (!in || !out || !in->area) return fail(XMC_ERR_ARG, "input/output is NULL");
    if (in->timeScaleSize < 1) return fail(XMC_ERR_ARG, "timeScaleSize must be >= 1");
    std::memset(out, 0, sizeof(*out));
    const xmc_int3& p = in->area->partitionNumber;
    int G = p.x * p.y * p.z;
    out->gridSize = G;
    out->absorption = (double*)calloc(G, sizeof(double));
    out->scattering = (double*)calloc(G, sizeof(double));
    int nd = in->numberOfCubeDetectors + in->numberOfRingDetectors;
    out->numberOfDetectors = nd;
    out->detectorInfo = (xmc_detector_info*)calloc(nd > 0 ? nd : 1, sizeof(xmc_detector_info));
    if (!out->absorption || !out->scattering || !out->detectorInfo) return fail(XMC_ERR_ARG, "out of host memory");
    int T = in->timeScaleSize;
    double dt = (in->timeFinish - in->timeStart) / T;
    if (dt < 0.0) dt = 0.0;
    for (int d = 0; d < nd; d++) {
        xmc_detector_info& di = out->detectorInfo[d];
        di.trajectorySize = G;
        di.trajectory = (uint64_t*)calloc(G, sizeof(uint64_t));  // launcher_omp.cpp:33-34
        di.timeScaleSize = T;
        di.timeScale = (xmc_time_info*)calloc(T, sizeof(xmc_time_info));
        if (!di.timeScale || !di.trajectory) return fail(XMC_ERR_ARG, "out of host memory");
        for (int j = 0; j < T - 1; j++) {
            di.timeScale[j].timeStart = in->timeStart + j * dt;
            di.timeScale[j].timeFinish = di.timeScale[j].timeStart + dt;
        }
        di.timeScale[T - 1].timeStart = in->timeStart + (T - 1) * dt;
        di.timeScale[T - 1].timeFinish = in->timeFinish;
    }
    return XMC_OK;
}

void xmc_free_output(xmc_output* out) {
    if (!out) return;
    free(out->absorption);
    free(out->scattering);
    if (out->detectorInfo) {
        for (int d = 0; d < out->numberOfDetectors; d++) {
            free(out->detectorInfo[d].timeScale);
            free(out->detectorInfo[d].trajectory);
        }
        free(out->detectorInfo);
    }
    std::memset(out, 0, sizeof(*out));
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: many configurations, one persistent launch, replayed as a graph.
// ---------------------------------------------------------------------------
struct xmc_batch {
    int device = 0;
    int n = 0;
    std::vector<xmc_scene*> members;
    std::vector<uint64_t> prefix;  // photons of configuration c: [prefix[c], prefix[c+1])
    DevBuf<xmc::Params> cfgs;
    DevBuf<uint64_t> dprefix;
    DevBuf<unsigned long long> counters;  // 16 stats words + queue
    DevBuf<int2> spill;
    int depth = 0, block = 256, blocks = 0;
    bool use_graph = true;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t graph_seed = 0;
    bool graph_valid = false;
    xmc::Params run;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double kernel_ms = 0.0;
};

static void drop_graph(xmc_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

extern "C" {

int xmc_batch_destroy(xmc_batch* b) {
    if (!b) return XMC_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (auto& e : b->pending) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (xmc_scene* m : b->members) xmc_scene_destroy(m);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return XMC_OK;
}

int xmc_batch_set_launch(xmc_batch* b, int block, int blocks_per_cu, int use_graph) {
    if (!b) return fail(XMC_ERR_ARG, "batch is NULL");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    int blocks = 0;
    int rc = configure_launch(b->device, b->depth, block, blocks_per_cu, xmc_batch_occupancy, &blocks, b->spill,
                              b->run);
    if (rc) return rc;
    b->block = block > 0 ? block : 256;
    b->blocks = blocks;
    b->use_graph = use_graph != 0;
    drop_graph(b);
    return XMC_OK;
}

int xmc_batch_create(const xmc_input* configs, int n, const xmc_options* opt, xmc_batch** out) {
    if (!configs || n <= 0 || !out) return fail(XMC_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    xmc_options o;
    std::memset(&o, 0, sizeof(o));
    if (opt) o = *opt;
    if (o.scattering || o.trajectories)
        return fail(XMC_ERR_ARG, "a batch tallies no scattering grid or trajectories (use one scene per input)");
    auto* b = new xmc_batch();
    std::unique_ptr<xmc_batch, int (*)(xmc_batch*)> guard(b, xmc_batch_destroy);
    // the batch kernel carries the rectangles' rule only: a configuration with
    // heightfields walks them (keeping the lattice code out of its registers)
    o.noLatticeRule = 1;
    b->device = o.device;
    b->n = n;
    b->prefix.assign(n + 1, 0);
    std::vector<xmc::Params> host(n);
    for (int c = 0; c < n; c++) {
        xmc_scene* m = nullptr;
        int rc = scene_init(&configs[c], &o, false, &m);
        if (rc) return fail(rc, "configuration " + std::to_string(c) + ": " + g_error);
        b->members.push_back(m);
        host[c] = m->base;
        b->prefix[c + 1] = b->prefix[c] + configs[c].numberOfPhotons;
        if (m->ref.depth > b->depth) b->depth = m->ref.depth;
    }
    if (b->prefix[n] > (1ull << xmc::kJumpBits)) return fail(XMC_ERR_LIMIT, "batch photon total beyond 2^33 (a seed has 2^33 disjoint streams)");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(b->cfgs.alloc(n));
    HIP_TRY(hipMemcpy(b->cfgs.p, host.data(), n * sizeof(xmc::Params), hipMemcpyHostToDevice));
    HIP_TRY(b->dprefix.alloc(n + 1));
    HIP_TRY(hipMemcpy(b->dprefix.p, b->prefix.data(), (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
    HIP_TRY(b->counters.alloc(17));
    HIP_TRY(hipMemset(b->counters.p, 0, 17 * 8));
    b->run = host[0];
    b->run.det_lds = 0;  // a batch lane's detectors are its configuration's: sums go straight to its tallies
    b->run.tally.stats = b->counters.p;
    b->run.queue = b->counters.p + 16;
    b->run.first_photon = 0;
    b->run.n_photons = b->prefix[n];
    std::memset(&b->run.trace, 0, sizeof(b->run.trace));
    HIP_TRY(hipStreamCreate(&b->stream));
    int rc = configure_launch(b->device, b->depth, o.blockSize, o.blocksPerCU, xmc_batch_occupancy, &b->blocks,
                              b->spill, b->run);
    if (rc) return rc;
    b->block = o.blockSize > 0 ? o.blockSize : 256;
    *out = guard.release();
    return XMC_OK;
}

// Runs every configuration's whole photon budget (seed `seed`, photon p of
// configuration c drawing stream block prefix[c] + p) as one launch.
int xmc_batch_run(xmc_batch* b, uint64_t seed, void* stream) {
    if (!b) return fail(XMC_ERR_ARG, "batch is NULL");
    HIP_TRY(hipSetDevice(b->device));
    hipStream_t s = stream ? (hipStream_t)stream : b->stream;
    b->run.seed_value = (2 * seed + 1) & xmc::kMcgMask;
    size_t lds = transport_lds(b->block);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    b->pending.emplace_back(e0, e1);
    if (b->use_graph) {
        if (!b->graph_valid || b->graph_seed != seed) {
            drop_graph(b);
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e1c = xmc_launch_zero(b->run.queue, 1, b->stream);
            hipError_t e2c = xmc_launch_batch(&b->run, b->cfgs.p, b->dprefix.p, b->n, b->blocks, b->block, lds,
                                              b->stream);
            hipError_t ec = hipStreamEndCapture(b->stream, &b->graph);
            if (e1c != hipSuccess || e2c != hipSuccess || ec != hipSuccess)
                return fail(XMC_ERR_HIP, "graph capture failed");
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graph_valid = true;
            b->graph_seed = seed;
        }
        HIP_TRY(hipEventRecord(e0, s));
        HIP_TRY(hipGraphLaunch(b->exec, s));
        HIP_TRY(hipEventRecord(e1, s));
    } else {
        HIP_TRY(xmc_launch_zero(b->run.queue, 1, s));
        HIP_TRY(hipEventRecord(e0, s));
        HIP_TRY(xmc_launch_batch(&b->run, b->cfgs.p, b->dprefix.p, b->n, b->blocks, b->block, lds, s));
        HIP_TRY(hipEventRecord(e1, s));
    }
    return XMC_OK;
}

int xmc_batch_collect(xmc_batch* b, xmc_output* outputs) {
    if (!b || !outputs) return fail(XMC_ERR_ARG, "batch/outputs is NULL");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (int c = 0; c < b->n; c++) {
        int rc = collect_impl(b->members[c], &outputs[c], b->stream);
        if (rc) return rc;
    }
    return XMC_OK;
}

int xmc_batch_stats(xmc_batch* b, xmc_stats* st) {
    if (!b || !st) return fail(XMC_ERR_ARG, "batch/stats is NULL");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipDeviceSynchronize());
    double ms = 0.0;
    int rc = drain_timers(b->pending, &ms);
    if (rc) return rc;
    rc = read_stats(b->counters.p, b->stream, st);
    st->kernelMs = ms;
    return rc;
}

int xmc_batch_get_launch(xmc_batch* b, int* block, int* blocks) {
    if (!b || !block || !blocks) return fail(XMC_ERR_ARG, "NULL argument");
    *block = b->block;
    *blocks = b->blocks;
    return XMC_OK;
}

}  // extern "C"
