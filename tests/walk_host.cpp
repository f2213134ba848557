// Test driver (CPU): the transport kernel's own BVH walk code compiled for the
// host.  tests/test_walk_on_host.py copies the device functions of
// xmcml_amd/csrc/xmc_transport.hip (from its constants to the photon physics:
// special functions, box and triangle tests, the short stack, walk_begin /
// walk_visit and walk) verbatim into walk_region.inc; the few HIP names they
// use get host definitions here.  walk_compare() runs every ray through
// walk(), through walk_begin + walk_visit until the walk ends, and through all
// triangles, and counts disagreements.  Built with sanitizers or with
// pattern-initialised locals, it checks the walk for undefined behaviour and
// reads of uninitialised values.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#define __device__
#define __forceinline__ inline
struct int2 { int x, y; };
struct int4 { int x, y, z, w; };
struct float4 { float x, y, z, w; };
struct float2 { float x, y; };
inline int2 make_int2(int a, int b) { return {a, b}; }
inline int __float_as_int(float f) { int i; std::memcpy(&i, &f, 4); return i; }
inline float __int_as_float(int i) { float f; std::memcpy(&f, &i, 4); return f; }
inline uint32_t __float_as_uint(float f) { uint32_t i; std::memcpy(&i, &f, 4); return i; }
inline float __uint_as_float(uint32_t i) { float f; std::memcpy(&f, &i, 4); return f; }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
#include "xmc_params.h"
#include "xmc_bvh.h"
#include "xmc_lattice.h"
namespace xmc {
#include "walk_region.inc"
// the whole walk as a loop over walk_visit (the form the interleaved transport
// runs one visit per round)
Hit walk_via_visit(const Geo& g, const Ray& r, float step, bool limited, Stack& st) {
    Hit h;
    int cur = 0, best = -1;
    if (!walk_begin(g, r, step, limited, st, h, cur, best)) return h;
    while (walk_visit(g.nodes, g.tris, r, step, limited, st, h, cur, best)) {
    }
    return h;
}
}  // namespace xmc
using namespace xmc;

extern "C" int walk_compare(const xmc_surface* surf, int ns, const double* o, const double* d, const double* steps,
                            int n, int* stats) {
    RefBvh ref;
    const char* err = nullptr;
    if (!build_reference_bvh(surf, ns, ref, &err)) return -1;
    GpuBvh gb;
    flatten_bvh(ref, surf, gb);
    int depth = gb.depth + 4;
    std::vector<int2> lds(kStackLds), spill(depth > 0 ? depth : 1);
    Geo geo{gb.nodes.data(), gb.tris.data(), gb.root_ref, {gb.root_lo[0], gb.root_lo[1], gb.root_lo[2]},
            {gb.root_hi[0], gb.root_hi[1], gb.root_hi[2]}};
    int ntri = (int)gb.tris.size();
    int bad_walk = 0, bad_visit = 0, found = 0;
    for (int i = 0; i < n; i++) {
        V3 oo = {(float)o[3 * i], (float)o[3 * i + 1], (float)o[3 * i + 2]};
        V3 dd = {(float)d[3 * i], (float)d[3 * i + 1], (float)d[3 * i + 2]};
        float step = (float)steps[i];
        bool limited = step > 0.0f;
        Ray r = make_ray(oo, dd);
        Stack s1{lds.data(), 1, 0, spill.data(), 1, 0, depth, 0, 0};
        Hit a = walk(geo, r, limited ? step : kInf, limited, s1);
        Stack s2{lds.data(), 1, 0, spill.data(), 1, 0, depth, 0, 0};
        Hit b = walk_via_visit(geo, r, limited ? step : kInf, limited, s2);
        // brute force
        float bd = kInf;
        int bt = -1, last;
        for (int k = 0; k < ntri; k++) {
            float t = tri_distance(gb.tris.data(), k, r, last);
            if (t >= 0.0f && t < bd) { bd = t; bt = k; }
        }
        bool bf = bt >= 0 && (!limited || bd <= step);
        found += bf;
        if (a.found != bf || (bf && a.dist != bd)) bad_walk++;
        if (b.found != bf || (bf && b.dist != bd)) bad_visit++;
    }
    stats[0] = found;
    stats[1] = bad_walk;
    stats[2] = bad_visit;
    return 0;
}

// The transport's surface rule (rule_hit) for every ray: the scene's forms
// built as the engine builds them (xmc_lattice.cpp), found / distance /
// surface / triangle index within the surface per ray.  Returns 1 when the
// scene is not under the rule.
extern "C" int rule_query(const xmc_surface* surf, int ns, const double* o, const double* d, const double* steps,
                          int n, int* found, double* dist, int* surface, int* index) {
    RefBvh ref;
    const char* err = nullptr;
    if (!build_reference_bvh(surf, ns, ref, &err)) return -1;
    GpuBvh gb;
    flatten_bvh(ref, surf, gb);
    std::vector<int> base(ns + 1, 0);
    for (int s = 0; s < ns; s++) base[s + 1] = base[s] + surf[s].numberOfTriangles;
    std::vector<int> leaf_of(base[ns]);
    for (size_t k = 0; k < gb.tris.size(); k++) leaf_of[base[ref.tri_surface[k]] + ref.tri_index[k]] = (int)k;
    static Params P;
    std::memset(&P, 0, sizeof(P));
    std::vector<int> rect_of(ns, -1), heights;
    for (int s = 0; s < ns; s++) {
        RectForm r;
        if (planar_rectangle(surf[s], r)) {
            const int k = P.planar.n++;
            rect_of[s] = k;
            P.planar.z[k] = r.z;
            P.planar.rect[k] = {r.x0, r.x1, r.y0, r.y1};
            P.planar.tri[k] = -1;
        } else {
            heights.push_back(s);
        }
    }
    LatticeForm lf;
    if (!heights.empty() && !lattice_form(surf, ns, heights, leaf_of, lf)) return 1;
    for (size_t k = 0; k < gb.tris.size(); k++) {
        const int r = rect_of[gb.tris[k].surface];
        if (r >= 0 && P.planar.tri[r] < 0) P.planar.tri[r] = (int)k;
    }
    P.surface_rule = lf.ns > 0 ? 2 : 1;
    Lattice& L = P.lattice;
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
    L.cells = reinterpret_cast<const float4*>(lf.cells.data());
    int depth = gb.depth + 4;
    std::vector<int2> lds(kStackLds), spill(depth > 0 ? depth : 1);
    Geo geo{gb.nodes.data(), gb.tris.data(), gb.root_ref, {gb.root_lo[0], gb.root_lo[1], gb.root_lo[2]},
            {gb.root_hi[0], gb.root_hi[1], gb.root_hi[2]}};
    for (int i = 0; i < n; i++) {
        V3 oo = {(float)o[3 * i], (float)o[3 * i + 1], (float)o[3 * i + 2]};
        V3 dd = {(float)d[3 * i], (float)d[3 * i + 1], (float)d[3 * i + 2]};
        const float step = steps[i] > 0.0 ? (float)steps[i] : kInf;
        Ray r = make_ray(oo, dd);
        Stack st{lds.data(), 1, 0, spill.data(), 1, 0, depth, 0, 0};
        Hit h = rule_hit(P, geo, oo, dd, step, st);
        found[i] = h.found ? 1 : 0;
        dist[i] = h.found ? (double)h.dist : 0.0;
        surface[i] = h.found ? ref.tri_surface[h.tri] : -1;
        index[i] = h.found ? ref.tri_index[h.tri] : -1;
    }
    return 0;
}
