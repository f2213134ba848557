// Conservative distance grid (see xmc_distance.h).
//
// Each cell stores a lower bound D_c on the distance from its centre c to
// the nearest triangle.  For any point p, |p - t| >= D_c - |p - c| for every
// triangle point t (triangle inequality), so the kernel, which knows p,
// subtracts |p - c| itself and never relies on p lying inside the cell.
//
// Two bounds are combined, D_c = max(near, far):
//   near: min(R, exact distance from c to every triangle whose bounding box
//         grown by R contains c).  When the nearest triangle is within R of
//         c it is one of those, so the minimum is exact; otherwise every
//         triangle is farther than R and R itself is a bound.  Closest point
//         on a triangle by Voronoi regions (Ericson, Real-Time Collision
//         Detection, 2005, 5.1.5).
//   far:  every cell that intersects the bounding box of a triangle is a
//         source; a triangle point in source cell M is within sqrt(3) h / 2
//         of its centre c_M, so |c - t| >= |c - c_M| - sqrt(3) h / 2.  The
//         exact Euclidean distance transform of the source set (Felzenszwalb
//         & Huttenlocher, "Distance Transforms of Sampled Functions", 2012:
//         separable lower envelopes of parabolas, one pass per axis) gives
//         min_M |c - c_M|.
#include "xmc_distance.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <atomic>
#include <thread>

namespace xmc {
namespace {

constexpr double kInf = 1e30;

// 1-D squared distance transform of f (length n) into d: lower envelope of
// the parabolas (x - q)^2 + f[q] (F&H 2012, algorithm DT).
void edt_1d(const double* f, double* d, int n, std::vector<int>& v, std::vector<double>& z) {
    v.resize(n);
    z.resize(n + 1);
    int k = 0;
    v[0] = 0;
    z[0] = -kInf;
    z[1] = kInf;
    for (int q = 1; q < n; q++) {
        double s = ((f[q] + (double)q * q) - (f[v[k]] + (double)v[k] * v[k])) / (2.0 * q - 2.0 * v[k]);
        while (s <= z[k]) {
            k--;
            s = ((f[q] + (double)q * q) - (f[v[k]] + (double)v[k] * v[k])) / (2.0 * q - 2.0 * v[k]);
        }
        k++;
        v[k] = q;
        z[k] = s;
        z[k + 1] = kInf;
    }
    k = 0;
    for (int q = 0; q < n; q++) {
        while (z[k + 1] < q) k++;
        double dq = q - v[k];
        d[q] = dq * dq + f[v[k]];
    }
}

struct D3 {
    double x, y, z;
};
inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 axpy(D3 a, double t, D3 b) { return {a.x + t * b.x, a.y + t * b.y, a.z + t * b.z}; }

// Squared distance from p to triangle (a, b, c): the closest point is a
// vertex, a point of an edge or of the face, told apart by the signs of the
// projections of p onto the edges (the triangle's Voronoi regions).
double point_triangle_sq(D3 p, D3 a, D3 b, D3 c) {
    D3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return dot(ap, ap);
    D3 bp = sub(p, b);
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return dot(bp, bp);
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        D3 q = axpy(a, d1 / (d1 - d3), ab);
        D3 e = sub(p, q);
        return dot(e, e);
    }
    D3 cp = sub(p, c);
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return dot(cp, cp);
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        D3 q = axpy(a, d2 / (d2 - d6), ac);
        D3 e = sub(p, q);
        return dot(e, e);
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        D3 q = axpy(b, (d4 - d3) / ((d4 - d3) + (d5 - d6)), sub(c, b));
        D3 e = sub(p, q);
        return dot(e, e);
    }
    double den = va + vb + vc;
    if (!(den > 0.0)) {  // degenerate (zero-area) triangle: nearest of its three edges
        double best = std::min({dot(ap, ap), dot(bp, bp), dot(cp, cp)});
        const D3 ends[3][2] = {{a, b}, {a, c}, {b, c}};
        for (const auto& s : ends) {
            D3 e = sub(s[1], s[0]);
            double ee = dot(e, e);
            if (ee <= 0.0) continue;
            double t = std::clamp(dot(sub(p, s[0]), e) / ee, 0.0, 1.0);
            D3 r = sub(p, axpy(s[0], t, e));
            best = std::min(best, dot(r, r));
        }
        return best;
    }
    double v = vb / den, w = vc / den;
    D3 q = axpy(axpy(a, v, ab), w, ac);
    D3 e = sub(p, q);
    return dot(e, e);
}

// Sutherland-Hodgman clip of a convex polygon (n <= 9 vertices) to the box
// [lo, hi]; returns the vertex count left.
int clip_to_box(D3* poly, int n, const double lo[3], const double hi[3]) {
    D3 tmp[12];
    for (int ax = 0; ax < 3 && n > 0; ax++)
        for (int side = 0; side < 2 && n > 0; side++) {
            auto val = [&](const D3& p) {
                double c = ax == 0 ? p.x : (ax == 1 ? p.y : p.z);
                return side == 0 ? c - lo[ax] : hi[ax] - c;
            };
            int m = 0;
            for (int i = 0; i < n; i++) {
                const D3 a = poly[i], b = poly[(i + 1) % n];
                double fa = val(a), fb = val(b);
                if (fa >= 0.0) tmp[m++] = a;
                if ((fa >= 0.0) != (fb >= 0.0)) tmp[m++] = axpy(a, fa / (fa - fb), sub(b, a));
            }
            for (int i = 0; i < m; i++) poly[i] = tmp[i];
            n = m;
        }
    return n;
}

void octahedral(D3 n, int& a, int& b) {
    double s = std::fabs(n.x) + std::fabs(n.y) + std::fabs(n.z);
    double u = n.x / s, v = n.y / s;
    if (n.z < 0.0) {
        double fu = (1.0 - std::fabs(v)) * (u >= 0.0 ? 1.0 : -1.0), fv = (1.0 - std::fabs(u)) * (v >= 0.0 ? 1.0 : -1.0);
        u = fu;
        v = fv;
    }
    a = (int)std::lrint(std::clamp(u, -1.0, 1.0) * 127.0);
    b = (int)std::lrint(std::clamp(v, -1.0, 1.0) * 127.0);
}

}  // namespace

void build_distance_grid(const xmc_surface* s, int ns, const double lo[3], const double hi[3], size_t target,
                         DistanceGrid& out) {
    double ext[3], vol = 1.0;
    for (int a = 0; a < 3; a++) {
        ext[a] = std::max(hi[a] - lo[a], 1e-6);
        vol *= ext[a];
    }
    double h = std::cbrt(vol / (double)std::max<size_t>(target, 1));
    int dims[3];
    size_t total = 1;
    for (int a = 0; a < 3; a++) {
        dims[a] = std::max(1, (int)std::ceil(ext[a] / h));
        total *= (size_t)dims[a];
    }
    // float-exact grid parameters: the kernel recomputes cell indices in float
    float hf = (float)h;
    h = hf;
    std::vector<double> f(total, kInf);
    auto idx = [&](int x, int y, int z) { return ((size_t)z * dims[1] + y) * dims[0] + x; };
    for (int i = 0; i < ns; i++) {
        for (int j = 0; j < s[i].numberOfTriangles; j++) {
            const xmc_int3& t = s[i].triangles[j];
            const xmc_double3* v = s[i].vertices;
            double tl[3] = {std::min({v[t.x].x, v[t.y].x, v[t.z].x}), std::min({v[t.x].y, v[t.y].y, v[t.z].y}),
                            std::min({v[t.x].z, v[t.y].z, v[t.z].z})};
            double th[3] = {std::max({v[t.x].x, v[t.y].x, v[t.z].x}), std::max({v[t.x].y, v[t.y].y, v[t.z].y}),
                            std::max({v[t.x].z, v[t.y].z, v[t.z].z})};
            int c0[3], c1[3];
            for (int a = 0; a < 3; a++) {
                c0[a] = std::max(0, (int)std::floor((tl[a] - lo[a]) / h));
                c1[a] = std::min(dims[a] - 1, (int)std::floor((th[a] - lo[a]) / h));
            }
            for (int z = c0[2]; z <= c1[2]; z++)
                for (int y = c0[1]; y <= c1[1]; y++)
                    for (int x = c0[0]; x <= c1[0]; x++) f[idx(x, y, z)] = 0.0;
        }
    }
    // separable squared EDT along x, y, z; the rows of one pass are independent
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (int a = 0; a < 3; a++) {
        const int n = dims[a], o1 = (a + 1) % 3, o2 = (a + 2) % 3;
        const int rows = dims[o1] * dims[o2];
        std::vector<std::thread> pool;
        for (unsigned th = 0; th < nthreads; th++)
            pool.emplace_back([&, th]() {
                std::vector<int> v;
                std::vector<double> zb, in(n), res(n);
                for (int row = (int)th; row < rows; row += (int)nthreads) {
                    int c[3];
                    c[o1] = row / dims[o2];
                    c[o2] = row % dims[o2];
                    for (int i = 0; i < n; i++) {
                        c[a] = i;
                        in[i] = f[idx(c[0], c[1], c[2])];
                    }
                    edt_1d(in.data(), res.data(), n, v, zb);
                    for (int i = 0; i < n; i++) {
                        c[a] = i;
                        f[idx(c[0], c[1], c[2])] = res[i];
                    }
                }
            });
        for (auto& t : pool) t.join();
    }
    // near bound: exact centre-to-triangle distances within R of each triangle.  R is
    // 4 cells, less when that would cost more than ~4e8 distance evaluations.
    struct Tri {
        D3 v[3];
        double lo[3], hi[3];
        int c0[3], c1[3];
    };
    std::vector<Tri> tris;
    for (int i = 0; i < ns; i++)
        for (int j = 0; j < s[i].numberOfTriangles; j++) {
            const xmc_int3& t = s[i].triangles[j];
            const xmc_double3* v = s[i].vertices;
            Tri T;
            T.v[0] = {v[t.x].x, v[t.x].y, v[t.x].z};
            T.v[1] = {v[t.y].x, v[t.y].y, v[t.y].z};
            T.v[2] = {v[t.z].x, v[t.z].y, v[t.z].z};
            tris.push_back(T);
        }
    auto cell_range = [&](double R) {
        double work = 0.0;
        for (Tri& T : tris) {
            const double tl[3] = {std::min({T.v[0].x, T.v[1].x, T.v[2].x}), std::min({T.v[0].y, T.v[1].y, T.v[2].y}),
                                  std::min({T.v[0].z, T.v[1].z, T.v[2].z})};
            const double th[3] = {std::max({T.v[0].x, T.v[1].x, T.v[2].x}), std::max({T.v[0].y, T.v[1].y, T.v[2].y}),
                                  std::max({T.v[0].z, T.v[1].z, T.v[2].z})};
            double cells = 1.0;
            for (int a = 0; a < 3; a++) {
                T.lo[a] = tl[a];
                T.hi[a] = th[a];
                // centres lo + (i + 1/2) h within [tl - R, th + R]
                T.c0[a] = std::max(0, (int)std::ceil((tl[a] - R - lo[a]) / h - 0.5));
                T.c1[a] = std::min(dims[a] - 1, (int)std::floor((th[a] + R - lo[a]) / h - 0.5));
                cells *= std::max(0, T.c1[a] - T.c0[a] + 1);
            }
            work += cells;
        }
        return work;
    };
    // cell centres as the kernel forms them: from the float origin
    const double flo[3] = {(double)(float)lo[0], (double)(float)lo[1], (double)(float)lo[2]};
    double R = 4.0 * h;
    while (R > h && cell_range(R) > 4e8) R -= 0.5 * h;
    cell_range(R);
    std::vector<float> near(total, (float)R);
    {
        const unsigned nth = nthreads;
        std::vector<std::thread> pool;
        for (unsigned th = 0; th < nth; th++)
            pool.emplace_back([&, th]() {
                // this thread owns the z-slices z = th (mod nth): the surfaces may fill only a
                // few slices, and interleaving shares those among all threads
                for (const Tri& T : tris) {
                    int za = T.c0[2] + (int)((th + nth - (unsigned)T.c0[2] % nth) % nth);
                    for (int z = za; z <= T.c1[2]; z += (int)nth)
                        for (int y = T.c0[1]; y <= T.c1[1]; y++)
                            for (int x = T.c0[0]; x <= T.c1[0]; x++) {
                                D3 c = {flo[0] + (x + 0.5) * h, flo[1] + (y + 0.5) * h, flo[2] + (z + 0.5) * h};
                                float& m = near[idx(x, y, z)];
                                // the triangle's bounding box is no farther than the triangle:
                                // skip it when the box is already beyond the cell's minimum
                                double bx = std::max({T.lo[0] - c.x, 0.0, c.x - T.hi[0]});
                                double by = std::max({T.lo[1] - c.y, 0.0, c.y - T.hi[1]});
                                double bz = std::max({T.lo[2] - c.z, 0.0, c.z - T.hi[2]});
                                if (bx * bx + by * by + bz * bz >= (double)m * (double)m) continue;
                                double d = std::sqrt(point_triangle_sq(c, T.v[0], T.v[1], T.v[2]));
                                // stored rounded down, so that the float never overstates d
                                float df = (float)d;
                                if ((double)df > d) df = std::nextafter(df, 0.0f);
                                if (df < m) m = df;
                            }
                }
            });
        for (auto& t : pool) t.join();
    }
    out.dims[0] = dims[0];
    out.dims[1] = dims[1];
    out.dims[2] = dims[2];
    for (int a = 0; a < 3; a++) out.lo[a] = (float)lo[a];
    out.h = hf;
    out.step_q = hf * 0.125f;
    out.reach = (float)R;
    out.q.assign(total, 0);
    const double half_diag = 0.5 * std::sqrt(3.0);
    for (size_t c = 0; c < total; c++) {
        double far = f[c] >= kInf * 0.5 ? 1e9 : (std::sqrt(f[c]) - half_diag) * h;
        double d = std::max((double)near[c], far);
        if (d <= 0.0) continue;
        double qv = std::floor(d / (double)out.step_q);
        out.q[c] = (uint8_t)std::min(qv, 255.0);
    }
}

}  // namespace xmc

namespace xmc {

void build_slab_grid(const xmc_surface* s, int ns, const DistanceGrid& dg, float mean_step, SlabGrid& out) {
    const double h = dg.h;
    const int* dims = dg.dims;
    const size_t total = (size_t)dims[0] * dims[1] * dims[2];
    out.unit = dg.h * 0.125f;
    out.words.assign(total, 0);
    struct Tri {
        D3 v[3];
        D3 n;  // unit normal
        double lo[3], hi[3];
    };
    std::vector<Tri> tris;
    for (int i = 0; i < ns; i++)
        for (int j = 0; j < s[i].numberOfTriangles; j++) {
            const xmc_int3& t = s[i].triangles[j];
            const xmc_double3* v = s[i].vertices;
            Tri T;
            T.v[0] = {v[t.x].x, v[t.x].y, v[t.x].z};
            T.v[1] = {v[t.y].x, v[t.y].y, v[t.y].z};
            T.v[2] = {v[t.z].x, v[t.z].y, v[t.z].z};
            D3 e1 = sub(T.v[1], T.v[0]), e2 = sub(T.v[2], T.v[0]);
            D3 c = {e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
            double l = std::sqrt(dot(c, c));
            T.n = l > 0.0 ? D3{c.x / l, c.y / l, c.z / l} : D3{0.0, 0.0, 1.0};
            for (int a = 0; a < 3; a++) {
                const double x0 = a == 0 ? T.v[0].x : (a == 1 ? T.v[0].y : T.v[0].z);
                const double x1 = a == 0 ? T.v[1].x : (a == 1 ? T.v[1].y : T.v[1].z);
                const double x2 = a == 0 ? T.v[2].x : (a == 1 ? T.v[2].y : T.v[2].z);
                T.lo[a] = std::min({x0, x1, x2});
                T.hi[a] = std::max({x0, x1, x2});
            }
            tris.push_back(T);
        }
    if (tris.empty()) return;
    // radius candidates in units of h / 8 (1, 1.5, 2, 3 and 4 cells)
    const int rs[5] = {8, 12, 16, 24, 32};
    const double Rmax = (double)((float)rs[4] * out.unit);
    // triangles bucketed by bounding box, buckets of 4 cells
    const int bc = 4;
    int bd[3];
    for (int a = 0; a < 3; a++) bd[a] = (dims[a] + bc - 1) / bc;
    std::vector<std::vector<int>> bucket((size_t)bd[0] * bd[1] * bd[2]);
    const double flo[3] = {(double)dg.lo[0], (double)dg.lo[1], (double)dg.lo[2]};
    const double bh = bc * h;
    for (size_t i = 0; i < tris.size(); i++) {
        int c0[3], c1[3];
        for (int a = 0; a < 3; a++) {
            c0[a] = std::clamp((int)std::floor((tris[i].lo[a] - flo[a]) / bh), 0, bd[a] - 1);
            c1[a] = std::clamp((int)std::floor((tris[i].hi[a] - flo[a]) / bh), 0, bd[a] - 1);
        }
        for (int z = c0[2]; z <= c1[2]; z++)
            for (int y = c0[1]; y <= c1[1]; y++)
                for (int x = c0[0]; x <= c1[0]; x++) bucket[((size_t)z * bd[1] + y) * bd[0] + x].push_back((int)i);
    }
    // cells near the mesh (bound below 3.5 cells)
    std::vector<size_t> cells;
    const double limit = 3.5 * h;
    for (size_t c = 0; c < total; c++)
        if ((double)dg.q[c] * (double)dg.step_q < limit) cells.push_back(c);
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned th = 0; th < nthreads; th++)
        pool.emplace_back([&]() {
            std::vector<uint32_t> stamp(tris.size(), 0xFFFFFFFFu);
            std::vector<std::pair<int, double>> near;  // triangle, squared distance (<= Rmax^2)
            std::vector<std::pair<double, double>> iv, bands;
            const int kSamples = 64;
            float sp[kSamples][3], sd[kSamples][3], ss[kSamples], soff[kSamples];
            for (;;) {
                const size_t chunk = next.fetch_add(256);
                if (chunk >= cells.size()) break;
                const size_t cend = std::min(cells.size(), chunk + 256);
                for (size_t ci = chunk; ci < cend; ci++) {
                    const size_t cell = cells[ci];
                    const int ix = (int)(cell % dims[0]), iy = (int)((cell / dims[0]) % dims[1]),
                              iz = (int)(cell / ((size_t)dims[0] * dims[1]));
                    const D3 c = {flo[0] + (ix + 0.5) * h, flo[1] + (iy + 0.5) * h, flo[2] + (iz + 0.5) * h};
                    const double cc[3] = {c.x, c.y, c.z};
                    near.clear();
                    int b0[3], b1[3];
                    for (int a = 0; a < 3; a++) {
                        b0[a] = std::clamp((int)std::floor((cc[a] - Rmax - flo[a]) / bh), 0, bd[a] - 1);
                        b1[a] = std::clamp((int)std::floor((cc[a] + Rmax - flo[a]) / bh), 0, bd[a] - 1);
                    }
                    for (int z = b0[2]; z <= b1[2]; z++)
                        for (int y = b0[1]; y <= b1[1]; y++)
                            for (int x = b0[0]; x <= b1[0]; x++)
                                for (int t : bucket[((size_t)z * bd[1] + y) * bd[0] + x]) {
                                    if (stamp[t] == (uint32_t)ci) continue;
                                    stamp[t] = (uint32_t)ci;
                                    const Tri& T = tris[t];
                                    double bx = std::max({T.lo[0] - c.x, 0.0, c.x - T.hi[0]});
                                    double by = std::max({T.lo[1] - c.y, 0.0, c.y - T.hi[1]});
                                    double bz = std::max({T.lo[2] - c.z, 0.0, c.z - T.hi[2]});
                                    if (bx * bx + by * by + bz * bz > Rmax * Rmax) continue;
                                    double d2 = point_triangle_sq(c, T.v[0], T.v[1], T.v[2]);
                                    if (d2 <= Rmax * Rmax) near.push_back({t, d2});
                                }
                    // sample steps of this cell (deterministic per cell)
                    uint64_t st = (uint64_t)cell * 0x9E3779B97F4A7C15ull + 1;
                    auto rnd = [&]() {
                        st = st * 6364136223846793005ull + 1442695040888963407ull;
                        return (double)(st >> 11) * 0x1p-53;
                    };
                    for (int k = 0; k < kSamples; k++) {
                        for (int a = 0; a < 3; a++) sp[k][a] = (float)((rnd() - 0.5) * h);
                        double zz = 2.0 * rnd() - 1.0, ph = 6.283185307179586 * rnd(), rr = std::sqrt(1.0 - zz * zz);
                        sd[k][0] = (float)(rr * std::cos(ph));
                        sd[k][1] = (float)(rr * std::sin(ph));
                        sd[k][2] = (float)zz;
                        ss[k] = (float)(-std::log(1.0 - rnd()) * mean_step);
                        soff[k] = std::sqrt(sp[k][0] * sp[k][0] + sp[k][1] * sp[k][1] + sp[k][2] * sp[k][2]);
                    }
                    const double cm = std::max({std::fabs(c.x), std::fabs(c.y), std::fabs(c.z)});
                    const float m = 1.0e-4f + 16.0f * (float)(cm * 0x1p-20 + 0x1p-22);
                    uint64_t best_word = 0;
                    int best_score = 0;
                    for (int ri = 0; ri < 5; ri++) {
                        const float Rf = (float)rs[ri] * out.unit;
                        const double R = Rf, q = R / 127.0;
                        int nearest = -1;
                        double nd2 = 1e300;
                        D3 mean = {0.0, 0.0, 0.0};
                        for (const auto& e : near)
                            if (e.second <= R * R && e.second < nd2) {
                                nd2 = e.second;
                                nearest = e.first;
                            }
                        if (nearest >= 0) {
                            for (const auto& e : near)
                                if (e.second <= R * R) {
                                    D3 n = tris[e.first].n;
                                    mean = axpy(mean, dot(n, tris[nearest].n) < 0.0 ? -1.0 : 1.0, n);
                                }
                        }
                        const double dbox_lo[3] = {c.x - R, c.y - R, c.z - R}, dbox_hi[3] = {c.x + R, c.y + R, c.z + R};
                        for (int ni = 0; ni < (nearest >= 0 ? 2 : 1); ni++) {
                            int na = 0, nb = 0;
                            int a1 = 127, b1v = 127, a2 = 127, b2v = 127;
                            if (nearest >= 0) {
                                D3 dir = tris[nearest].n;
                                if (ni == 1) {
                                    double l = std::sqrt(dot(mean, mean));
                                    if (!(l > 0.0)) continue;
                                    dir = {mean.x / l, mean.y / l, mean.z / l};
                                }
                                octahedral(dir, na, nb);
                                float nf[3];
                                slab_direction(na, nb, nf);
                                const D3 n = {nf[0], nf[1], nf[2]};
                                iv.clear();
                                for (const auto& e : near) {
                                    if (e.second > R * R) continue;
                                    D3 poly[12] = {tris[e.first].v[0], tris[e.first].v[1], tris[e.first].v[2]};
                                    int pn = clip_to_box(poly, 3, dbox_lo, dbox_hi);
                                    if (pn == 0) continue;
                                    double mn = 1e300, mx = -1e300;
                                    for (int k = 0; k < pn; k++) {
                                        double t = dot(n, sub(poly[k], c));
                                        mn = std::min(mn, t);
                                        mx = std::max(mx, t);
                                    }
                                    iv.push_back({mn, mx});
                                }
                                std::sort(iv.begin(), iv.end());
                                bands.clear();
                                for (const auto& b : iv) {
                                    if (!bands.empty() && b.first <= bands.back().second)
                                        bands.back().second = std::max(bands.back().second, b.second);
                                    else
                                        bands.push_back(b);
                                }
                                // at most two bands: join the two closest neighbours
                                while (bands.size() > 2) {
                                    size_t k = 0;
                                    double gap = 1e300;
                                    for (size_t i = 0; i + 1 < bands.size(); i++)
                                        if (bands[i + 1].first - bands[i].second < gap) {
                                            gap = bands[i + 1].first - bands[i].second;
                                            k = i;
                                        }
                                    bands[k].second = bands[k + 1].second;
                                    bands.erase(bands.begin() + (long)k + 1);
                                }
                                // bands widened outward to multiples of R / 127
                                auto lo_q = [&](double v) { return (int)std::clamp(std::floor(v / q), -127.0, 127.0); };
                                auto hi_q = [&](double v) { return (int)std::clamp(std::ceil(v / q), -127.0, 127.0); };
                                if (!bands.empty()) {
                                    a1 = lo_q(bands[0].first);
                                    b1v = hi_q(bands[0].second);
                                }
                                if (bands.size() == 2) {
                                    a2 = lo_q(bands[1].first);
                                    b2v = hi_q(bands[1].second);
                                }
                            }
                            // the kernel's test on the samples
                            float nf[3];
                            slab_direction(na, nb, nf);
                            const float qf = Rf * (1.0f / 127.0f);
                            int score = 0;
                            for (int k = 0; k < kSamples; k++) {
                                const float room = Rf - m - soff[k];
                                if (!(room > 0.0f && ss[k] < room)) continue;
                                float t0 = nf[0] * sp[k][0] + nf[1] * sp[k][1] + nf[2] * sp[k][2];
                                float t1 = t0 + ss[k] * (nf[0] * sd[k][0] + nf[1] * sd[k][1] + nf[2] * sd[k][2]);
                                float tmin = std::min(t0, t1), tmax = std::max(t0, t1);
                                if (tmax < (float)a1 * qf - m || (tmin > (float)b1v * qf + m && tmax < (float)a2 * qf - m) ||
                                    tmin > (float)b2v * qf + m)
                                    score++;
                            }
                            if (score > best_score) {
                                best_score = score;
                                best_word = (uint64_t)(uint8_t)(int8_t)na | (uint64_t)(uint8_t)(int8_t)nb << 8 |
                                            (uint64_t)(uint8_t)(int8_t)a1 << 16 | (uint64_t)(uint8_t)(int8_t)b1v << 24 |
                                            (uint64_t)(uint8_t)(int8_t)a2 << 32 | (uint64_t)(uint8_t)(int8_t)b2v << 40 |
                                            (uint64_t)(uint8_t)rs[ri] << 48;
                            }
                        }
                    }
                    out.words[cell] = best_word;
                }
            }
        });
    for (auto& t : pool) t.join();
}

}  // namespace xmc
