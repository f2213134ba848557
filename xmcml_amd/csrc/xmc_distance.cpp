// Conservative distance grid (see xmc_distance.h).
//
// Marking: every cell that intersects the bounding box of a triangle is a
// source, so every triangle point lies in a source cell.  For a point p in
// cell C and a triangle point t in source cell M,
//   |p - t| >= |c_C - c_M| - sqrt(3) h
// (each point is within sqrt(3) h / 2 of its cell centre).  The exact
// Euclidean distance transform of the source set (Felzenszwalb &
// Huttenlocher, "Distance Transforms of Sampled Functions", 2012: separable
// lower envelopes of parabolas, one pass per axis) gives min_M |c_C - c_M|.
#include "xmc_distance.h"

#include <algorithm>
#include <cmath>
#include <limits>

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
    // separable squared EDT along x, y, z
    std::vector<int> v;
    std::vector<double> zb, in, res;
    for (int a = 0; a < 3; a++) {
        int n = dims[a];
        in.resize(n);
        res.resize(n);
        int o1 = (a + 1) % 3, o2 = (a + 2) % 3;
        for (int p = 0; p < dims[o1]; p++)
            for (int r = 0; r < dims[o2]; r++) {
                int c[3];
                c[o1] = p;
                c[o2] = r;
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
    }
    out.dims[0] = dims[0];
    out.dims[1] = dims[1];
    out.dims[2] = dims[2];
    for (int a = 0; a < 3; a++) out.lo[a] = (float)lo[a];
    out.h = hf;
    out.step_q = hf * 0.25f;
    out.q.assign(total, 0);
    const double sq3 = std::sqrt(3.0);
    for (size_t c = 0; c < total; c++) {
        double d = f[c] >= kInf * 0.5 ? 1e9 : (std::sqrt(f[c]) - sq3) * h;
        if (d <= 0.0) continue;
        double qv = std::floor(d / (double)out.step_q);
        out.q[c] = (uint8_t)std::min(qv, 255.0);
    }
}

}  // namespace xmc
