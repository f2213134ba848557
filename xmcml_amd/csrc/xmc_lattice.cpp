// Surface forms of the transport's surface rule (see xmc_lattice.h).
#include "xmc_lattice.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace xmc {

bool planar_rectangle(const xmc_surface& S, RectForm& r) {
    if (S.numberOfTriangles != 2 || !S.triangles || !S.vertices) return false;
    const xmc_int3 t[2] = {S.triangles[0], S.triangles[1]};
    const int used[6] = {t[0].x, t[0].y, t[0].z, t[1].x, t[1].y, t[1].z};
    for (int k = 0; k < 6; k++)
        if (used[k] < 0 || used[k] >= S.numberOfVertices) return false;
    const xmc_double3& v0 = S.vertices[used[0]];
    double xs[2] = {v0.x, v0.x}, ys[2] = {v0.y, v0.y};
    for (int k = 0; k < 6; k++) {
        const xmc_double3& v = S.vertices[used[k]];
        if (v.z != v0.z) return false;
        xs[0] = std::min(xs[0], v.x);
        xs[1] = std::max(xs[1], v.x);
        ys[0] = std::min(ys[0], v.y);
        ys[1] = std::max(ys[1], v.y);
    }
    if (!(xs[0] < xs[1] && ys[0] < ys[1])) return false;
    // each vertex a corner; each triangle three distinct corners; the shared two opposite
    auto corner = [&](int i) {
        const xmc_double3& v = S.vertices[i];
        const int cx = v.x == xs[0] ? 0 : (v.x == xs[1] ? 1 : -1);
        const int cy = v.y == ys[0] ? 0 : (v.y == ys[1] ? 1 : -1);
        return (cx < 0 || cy < 0) ? -1 : cx + 2 * cy;
    };
    int mask[2] = {0, 0};
    for (int k = 0; k < 6; k++) {
        const int c = corner(used[k]);
        if (c < 0) return false;
        mask[k / 3] |= 1 << c;
    }
    auto bits = [](int m) { return __builtin_popcount((unsigned)m); };
    if (bits(mask[0]) != 3 || bits(mask[1]) != 3 || (mask[0] | mask[1]) != 15) return false;
    const int shared = mask[0] & mask[1];
    if (shared != (1 | 8) && shared != (2 | 4)) return false;  // corners 0-3 or 1-2: a diagonal
    r.z = (float)v0.z;
    r.x0 = (float)xs[0];
    r.x1 = (float)xs[1];
    r.y0 = (float)ys[0];
    r.y1 = (float)ys[1];
    return true;
}

namespace {

// The distinct values of a lattice axis: uniform spacing within 1e-6 of a step.
bool uniform_axis(std::vector<double>& v, double& origin, double& step) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    if (v.size() < 2 || v.size() > 65537) return false;
    const int n = (int)v.size() - 1;
    origin = v.front();
    step = (v.back() - v.front()) / n;
    if (!(step > 0.0) || !std::isfinite(step)) return false;
    for (int i = 0; i <= n; i++)
        if (std::fabs(v[i] - (origin + i * step)) > 1e-6 * step) return false;
    return true;
}

}  // namespace

bool lattice_form(const xmc_surface* surfaces, int n_surfaces, const std::vector<int>& which,
                  const std::vector<int>& leaf_of, LatticeForm& out) {
    out = LatticeForm();
    if (which.empty() || which.size() > 4) return false;  // kLatticeMaxSurfaces
    std::vector<int> base(n_surfaces + 1, 0);
    for (int s = 0; s < n_surfaces; s++) base[s + 1] = base[s] + surfaces[s].numberOfTriangles;
    std::vector<double> xs, ys;
    double maxabs_xy = 0.0, maxabs_z = 0.0;
    for (int s : which) {
        const xmc_surface& S = surfaces[s];
        if (S.numberOfTriangles < 1 || !S.triangles || !S.vertices) return false;
        for (int t = 0; t < S.numberOfTriangles; t++) {
            const int id[3] = {S.triangles[t].x, S.triangles[t].y, S.triangles[t].z};
            for (int k = 0; k < 3; k++) {
                if (id[k] < 0 || id[k] >= S.numberOfVertices) return false;
                const xmc_double3& v = S.vertices[id[k]];
                if (!std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z)) return false;
                xs.push_back(v.x);
                ys.push_back(v.y);
                maxabs_xy = std::max(maxabs_xy, std::max(std::fabs(v.x), std::fabs(v.y)));
                maxabs_z = std::max(maxabs_z, std::fabs(v.z));
            }
        }
    }
    double x0, y0, dx, dy;
    if (!uniform_axis(xs, x0, dx) || !uniform_axis(ys, y0, dy)) return false;
    const int nx = (int)xs.size() - 1, ny = (int)ys.size() - 1, ns = (int)which.size();
    if ((double)nx * ny * ns > (double)(1 << 26)) return false;
    auto index = [](double v, double origin, double step) { return (int)std::lround((v - origin) / step); };

    const float inf = std::numeric_limits<float>::infinity();
    std::vector<float> cells((size_t)nx * ny * ns * 4);
    for (size_t c = 0; c < (size_t)nx * ny * ns; c++) {
        cells[4 * c] = inf;
        cells[4 * c + 1] = -inf;
        const int none = -1;
        std::memcpy(&cells[4 * c + 2], &none, 4);
        std::memcpy(&cells[4 * c + 3], &none, 4);
    }
    float zlo = inf, zhi = -inf;
    for (int ls = 0; ls < ns; ls++) {
        const int s = which[ls];
        const xmc_surface& S = surfaces[s];
        for (int t = 0; t < S.numberOfTriangles; t++) {
            const int id[3] = {S.triangles[t].x, S.triangles[t].y, S.triangles[t].z};
            int ix[3], iy[3];
            float tz[2] = {inf, -inf};
            for (int k = 0; k < 3; k++) {
                const xmc_double3& v = S.vertices[id[k]];
                ix[k] = index(v.x, x0, dx);
                iy[k] = index(v.y, y0, dy);
                const float z = (float)v.z;
                tz[0] = std::min(tz[0], z);
                tz[1] = std::max(tz[1], z);
            }
            const int cx = std::min({ix[0], ix[1], ix[2]}), cy = std::min({iy[0], iy[1], iy[2]});
            if (std::max({ix[0], ix[1], ix[2]}) != cx + 1 || std::max({iy[0], iy[1], iy[2]}) != cy + 1) return false;
            float* rec = &cells[(((size_t)cy * nx + cx) * ns + ls) * 4];
            int a, b;
            std::memcpy(&a, rec + 2, 4);
            std::memcpy(&b, rec + 3, 4);
            if (b >= 0) return false;  // a third triangle in the cell
            const int leaf = leaf_of[base[s] + t];
            if (a < 0) {
                a = leaf;
            } else {
                b = std::max(a, leaf);
                a = std::min(a, leaf);
            }
            std::memcpy(rec + 2, &a, 4);
            std::memcpy(rec + 3, &b, 4);
            rec[0] = std::min(rec[0], tz[0]);
            rec[1] = std::max(rec[1], tz[1]);
            zlo = std::min(zlo, tz[0]);
            zhi = std::max(zhi, tz[1]);
        }
    }
    // margins: 1e-3 of a cell, the float error of (x - x0) / dx over nx cells,
    // and a few ulps of the coordinates (vertex rounding, the triangle test)
    const double ulp = std::ldexp(1.0, -22);
    const double mxy = 1e-3 * std::min(dx, dy) + 4.0 * std::max(nx, ny) * ulp * std::max(dx, dy) + 8.0 * ulp * maxabs_xy;
    const double mz = 1e-4 * (1.0 + ((double)zhi - (double)zlo)) + 8.0 * ulp * maxabs_z;
    out.ns = ns;
    out.nx = nx;
    out.ny = ny;
    out.x0 = (float)x0;
    out.y0 = (float)y0;
    out.inv_dx = (float)(1.0 / dx);
    out.inv_dy = (float)(1.0 / dy);
    out.mxy = (float)mxy;
    out.mz = (float)mz;
    out.zlo = zlo;
    out.zhi = zhi;
    out.cells = std::move(cells);
    return true;
}

}  // namespace xmc
