// Test driver (CPU): walks from the entry grid's restricted trees equal walks
// from the BVH root.
//
// Builds the scene's BVH, distance grid and entry grid with the engine's own
// host code (xmcml_amd/csrc/xmc_bvh.cpp, xmc_distance.cpp, xmc_entry.cpp) and
// runs step-limited walks of random rays near the surfaces twice: from the
// root of the real BVH, and from the cell's entry (level 1 for steps moving at
// most `margin` along every axis, level 2 up to `margin2`; no entry = nothing
// in reach).  The walk is the transport kernel's walk_visit loop restated on
// the host with the kernel's float operations (fused dot/cross products and
// slab distances, kDeadRef children neither visited nor pushed).  Prints one
// JSON line: query counts, mismatches (hit triangle or distance differing),
// mean visits.
//
// usage: entry_walks <file.surface> <queries> <seed>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <utility>
#include <vector>

#include "xmc_bvh.h"
#include "xmc_distance.h"
#include "xmc_entry.h"

using namespace xmc;

namespace {

const float kInf = INFINITY, kEpsF = 1.0e-6f;

struct V3 {
    float x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
float dot(V3 a, V3 b) { return std::fmaf(a.z, b.z, std::fmaf(a.y, b.y, a.x * b.x)); }
V3 cross(V3 a, V3 b) {
    return {std::fmaf(a.y, b.z, -(a.z * b.y)), std::fmaf(a.z, b.x, -(a.x * b.z)), std::fmaf(a.x, b.y, -(a.y * b.x))};
}

struct Ray {
    float o[3], d[3], inv[3], oi[3];
    bool ax[3];
};

Ray make_ray(V3 o, V3 d) {
    Ray r;
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    for (int a = 0; a < 3; a++) {
        r.o[a] = oo[a];
        r.d[a] = dd[a];
        r.ax[a] = std::fabs(dd[a]) > kEpsF;
        r.inv[a] = r.ax[a] ? 1.0f / dd[a] : 0.0f;
        r.oi[a] = r.o[a] * r.inv[a];
    }
    return r;
}

bool box_hit(const Ray& r, const float* lo, const float* hi, float& tn) {
    float tmin = -kInf, tmax = kInf;
    bool ok = true;
    for (int a = 0; a < 3; a++) {
        if (r.ax[a]) {
            float t1 = std::fmaf(lo[a], r.inv[a], -r.oi[a]), t2 = std::fmaf(hi[a], r.inv[a], -r.oi[a]);
            tmin = std::fmax(tmin, std::fmin(t1, t2));
            tmax = std::fmin(tmax, std::fmax(t1, t2));
        } else {
            ok = ok && !((lo[a] - r.o[a] > 0.0f) || (hi[a] - r.o[a] < 0.0f));
        }
    }
    ok = ok && !(tmin > tmax) && !(tmax < 0.0f);
    tn = ok ? tmin : kInf;
    return ok;
}

float tri_distance(const GpuTri& T, const Ray& r) {
    V3 a = {T.a[0], T.a[1], T.a[2]}, e1 = {T.e1[0], T.e1[1], T.e1[2]}, e2 = {T.e2[0], T.e2[1], T.e2[2]};
    V3 o = {r.o[0], r.o[1], r.o[2]}, d = {r.d[0], r.d[1], r.d[2]};
    V3 p = cross(d, e2);
    float det = dot(e1, p);
    if (det < kEpsF && det > -kEpsF) return kInf;
    float inv = 1.0f / det;
    V3 t = sub(o, a);
    float u = dot(t, p) * inv;
    if (u < 0.0f || u > 1.0f) return kInf;
    V3 q = cross(t, e1);
    float v = dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return kInf;
    return dot(e2, q) * inv;
}

struct Result {
    int tri = -1;
    float dist = kInf;
    long visits = 0;
};

// walk_begin + walk_visit until the walk ends (limited walks)
Result walk(const std::vector<GpuNode>& nodes, const std::vector<GpuTri>& tris, int start, const float* slo,
            const float* shi, const Ray& r, float step) {
    Result R;
    float tn;
    if (!box_hit(r, slo, shi, tn) || step < tn) return R;
    std::vector<std::pair<int, float>> st;
    int cur = start, best = -1;
    float hd = kInf;
    while (true) {
        bool use_stack;
        R.visits++;
        if (cur < 0) {
            int i = ~cur;
            float t = tri_distance(tris[i], r);
            if (t >= 0.0f && t < hd) {
                hd = t;
                best = i;
            }
            if (!tris[i].last) {
                cur = ~(i + 1);
                continue;
            }
            if (best >= 0 && hd <= step) R.tri = best;
            best = -1;
            use_stack = true;
        } else {
            const GpuNode& N = nodes[cur];
            float t1, t2;
            bool f1 = box_hit(r, N.c0lo, N.c0hi, t1), f2 = box_hit(r, N.c1lo, N.c1hi, t2);
            int L = N.child[0], Rr = N.child[1];
            use_stack = false;
            if (f1 && f2) {
                if (t1 <= t2) {
                    if (t1 <= step) {
                        cur = L;
                        if (t2 <= step && Rr != kDeadRef) st.push_back({Rr, t2});
                    } else {
                        use_stack = true;
                    }
                } else {
                    if (t2 <= step) {
                        cur = Rr;
                        if (t1 <= step && L != kDeadRef) st.push_back({L, t1});
                    } else {
                        use_stack = true;
                    }
                }
            } else if (f1) {
                if (t1 <= step) cur = L;
                else use_stack = true;
            } else {
                if (f2 && t2 <= step) cur = Rr;
                else use_stack = true;
            }
            if (cur == kDeadRef) use_stack = true;
        }
        if (use_stack) {
            if (st.empty()) break;
            std::pair<int, float> e = st.back();
            st.pop_back();
            float t = e.second;
            cur = e.first;
            if (hd < t) {
                while (!st.empty() && hd < t) {
                    e = st.back();
                    st.pop_back();
                    t = e.second;
                }
                if (hd < t || step < t) break;
                cur = e.first;
            }
        }
    }
    R.dist = hd;
    return R;
}

bool read_surfaces(const char* path, std::vector<std::vector<xmc_double3>>& V, std::vector<std::vector<xmc_int3>>& T,
                   std::vector<xmc_surface>& S) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    uint32_t id = 0;
    int ns = 0;
    bool ok = std::fread(&id, 4, 1, f) == 1 && std::fread(&ns, 4, 1, f) == 1 && id == 0x0100 && ns > 0;
    V.resize(ok ? ns : 0);
    T.resize(ok ? ns : 0);
    for (int i = 0; ok && i < ns; i++) {
        int nv = 0, nt = 0;
        ok = std::fread(&nv, 4, 1, f) == 1 && nv > 0;
        if (ok) V[i].resize(nv), ok = std::fread(V[i].data(), sizeof(xmc_double3), nv, f) == (size_t)nv;
        ok = ok && std::fread(&nt, 4, 1, f) == 1 && nt > 0;
        if (ok) T[i].resize(nt), ok = std::fread(T[i].data(), sizeof(xmc_int3), nt, f) == (size_t)nt;
    }
    std::fclose(f);
    if (!ok) return false;
    for (int i = 0; i < ns; i++)
        S.push_back({V[i].data(), (int)V[i].size(), T[i].data(), (int)T[i].size()});
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: entry_walks <file.surface> <queries> <seed>\n");
        return 2;
    }
    std::vector<std::vector<xmc_double3>> V;
    std::vector<std::vector<xmc_int3>> T;
    std::vector<xmc_surface> S;
    if (!read_surfaces(argv[1], V, T, S)) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    const long nq = std::atol(argv[2]);
    std::mt19937_64 rng(std::strtoull(argv[3], nullptr, 10));

    // the scene set-up of xmc_scene_create (xmc_host.cpp)
    RefBvh rb;
    const char* err = nullptr;
    if (!build_reference_bvh(S.data(), (int)S.size(), rb, &err)) return 2;
    GpuBvh gb;
    flatten_bvh(rb, S.data(), gb);
    const size_t nt = gb.tris.size();
    DistanceGrid dg;
    double lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        double pad = 1e-3 * (gb.root_hi[a] - gb.root_lo[a]) + 1e-3;
        lo[a] = gb.root_lo[a] - pad;
        hi[a] = gb.root_hi[a] + pad;
    }
    size_t target = std::min<size_t>(std::max<size_t>(nt * 256, (size_t)1 << 18), (size_t)32 << 20);
    build_distance_grid(S.data(), (int)S.size(), lo, hi, target, dg);
    EntryGrid eg;
    build_entry_grid(gb, dg.lo, dg.h, dg.dims, dg.q.data(), dg.step_q, eg);
    std::vector<GpuNode> all(gb.nodes);
    all.insert(all.end(), eg.extra.begin(), eg.extra.end());

    // rays from points within three cells of a random point of a random triangle, random
    // directions, steps of 0.3, 1 and 3 cells on average
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long n1 = 0, n2 = 0, nnone = 0, nlong = 0, mism = 0, hits = 0, vroot = 0, ventry = 0;
    const double h = dg.h;
    for (long k = 0; k < nq; k++) {
        const GpuTri& t = gb.tris[(size_t)(U(rng) * nt) % nt];
        double a = U(rng), b = U(rng);
        if (a + b > 1.0) a = 1.0 - a, b = 1.0 - b;
        V3 p;
        float* pp = &p.x;
        for (int i = 0; i < 3; i++)
            pp[i] = (float)(t.a[i] + a * t.e1[i] + b * t.e2[i] + (2.0 * U(rng) - 1.0) * 3.0 * h);
        double z = 2.0 * U(rng) - 1.0, phi = 2.0 * M_PI * U(rng), s = std::sqrt(1.0 - z * z);
        V3 d = {(float)(s * std::cos(phi)), (float)(s * std::sin(phi)), (float)z};
        const double mean = (k % 3 == 0 ? 0.3 : (k % 3 == 1 ? 1.0 : 3.0)) * h;
        float step = (float)(-std::log(1.0 - U(rng)) * mean);
        Ray r = make_ray(p, d);
        Result root = walk(all, gb.tris, gb.root_ref, gb.root_lo, gb.root_hi, r, step);
        vroot += root.visits;
        if (root.tri >= 0) hits++;
        // the cell as the kernel files p (walk_start)
        float fx = (p.x - dg.lo[0]) * (1.0f / dg.h), fy = (p.y - dg.lo[1]) * (1.0f / dg.h),
              fz = (p.z - dg.lo[2]) * (1.0f / dg.h);
        if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f)) continue;
        int ix = (int)fx, iy = (int)fy, iz = (int)fz;
        if (ix >= dg.dims[0] || iy >= dg.dims[1] || iz >= dg.dims[2]) continue;
        size_t c = ((size_t)iz * dg.dims[1] + iy) * dg.dims[0] + ix;
        float reach = step * std::fmax(std::fabs(d.x), std::fmax(std::fabs(d.y), std::fabs(d.z)));
        int e;
        if (reach <= eg.margin) {
            e = eg.cell[c];
            n1++;
        } else if (reach <= eg.margin2) {
            e = eg.cell2[c];
            n2++;
        } else {
            nlong++;
            continue;
        }
        Result got;
        if (e < 0) {
            nnone++;  // nothing in reach: the walk is skipped
        } else {
            got = walk(all, gb.tris, eg.ref[e], &eg.box[6 * e], &eg.box[6 * e + 3], r, step);
            ventry += got.visits;
        }
        if (got.tri != root.tri || (got.tri >= 0 && got.dist != root.dist)) mism++;
    }
    std::printf(
        "{\"triangles\": %zu, \"queries\": %ld, \"level1\": %ld, \"level2\": %ld, \"no_entry\": %ld, \"long\": %ld, "
        "\"hits\": %ld, \"mismatches\": %ld, \"visits_root\": %.3f, \"visits_entry\": %.3f, \"trees\": %zu, "
        "\"extra_nodes\": %zu}\n",
        nt, nq, n1, n2, nnone, nlong, hits, mism, (double)vroot / nq,
        (double)ventry / std::max(1L, n1 + n2 - nnone), eg.ref.size(), eg.extra.size());
    return 0;
}
