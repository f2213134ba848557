// Entry-node grid (see xmc_entry.h).
#include "xmc_entry.h"

#include <algorithm>
#include <cmath>
#include <thread>
#include <unordered_map>

namespace xmc {

namespace {

inline bool overlaps(const float lo[3], const float hi[3], const double blo[3], const double bhi[3]) {
    for (int a = 0; a < 3; a++)
        if ((double)hi[a] < blo[a] || (double)lo[a] > bhi[a]) return false;
    return true;
}

// Deepest node holding every leaf whose box meets [blo, bhi]: its reference
// and box, or false when no leaf does.
bool entry_node(const GpuBvh& bvh, const double blo[3], const double bhi[3], int& ref, const float*& elo,
                const float*& ehi) {
    if (!overlaps(bvh.root_lo, bvh.root_hi, blo, bhi)) return false;
    ref = bvh.root_ref;
    elo = bvh.root_lo;
    ehi = bvh.root_hi;
    while (ref >= 0) {
        const GpuNode& nd = bvh.nodes[ref];
        bool o0 = overlaps(nd.c0lo, nd.c0hi, blo, bhi);
        bool o1 = overlaps(nd.c1lo, nd.c1hi, blo, bhi);
        if (o0 && o1) break;
        if (!o0 && !o1) return false;
        if (o0) {
            ref = nd.child[0];
            elo = nd.c0lo;
            ehi = nd.c0hi;
        } else {
            ref = nd.child[1];
            elo = nd.c1lo;
            ehi = nd.c1hi;
        }
    }
    return true;
}

}  // namespace

void build_entry_grid(const GpuBvh& bvh, const float lo[3], float h, const int dims[3], EntryGrid& out) {
    double scale = 0.0;
    for (int a = 0; a < 3; a++)
        scale = std::max(scale, std::max(std::fabs((double)bvh.root_lo[a]), std::fabs((double)bvh.root_hi[a])));
    out.margin = h;
    // the kernel locates a point's cell in float, so a point may sit just outside the cell
    // it is filed under; its float hit tests also reach slightly beyond the exact segment
    const double guard = 1e-3 + 1e-5 * scale;
    const double grow = (double)out.margin + guard;
    const double hh = (double)h;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    out.cell.assign(n, -1);
    // per cell: the entry's node reference (or INT32_MIN = none); boxes looked up after
    std::vector<int32_t> node_of(n, INT32_MIN);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++)
        pool.emplace_back([&, t]() {
            for (int iz = (int)t; iz < dims[2]; iz += (int)nt)
                for (int iy = 0; iy < dims[1]; iy++)
                    for (int ix = 0; ix < dims[0]; ix++) {
                        double blo[3], bhi[3];
                        const int ii[3] = {ix, iy, iz};
                        for (int a = 0; a < 3; a++) {
                            blo[a] = (double)lo[a] + ii[a] * hh - grow;
                            bhi[a] = (double)lo[a] + (ii[a] + 1) * hh + grow;
                        }
                        int ref;
                        const float *elo, *ehi;
                        if (entry_node(bvh, blo, bhi, ref, elo, ehi))
                            node_of[((size_t)iz * dims[1] + iy) * dims[0] + ix] = ref;
                    }
        });
    for (auto& th : pool) th.join();
    // number the distinct entry nodes and record their boxes
    std::unordered_map<int, int> index;
    out.ref.clear();
    out.box.clear();
    for (size_t c = 0; c < n; c++) {
        int ref = node_of[c];
        if (ref == INT32_MIN) continue;
        auto it = index.find(ref);
        if (it == index.end()) {
            int e = (int)out.ref.size();
            it = index.emplace(ref, e).first;
            out.ref.push_back(ref);
        }
        out.cell[c] = it->second;
    }
    // boxes: the root's, or the one stored with the node in its parent
    std::vector<const float*> blo(out.ref.size(), nullptr), bhi(out.ref.size(), nullptr);
    for (size_t e = 0; e < out.ref.size(); e++)
        if (out.ref[e] == bvh.root_ref) {
            blo[e] = bvh.root_lo;
            bhi[e] = bvh.root_hi;
        }
    for (const GpuNode& nd : bvh.nodes)
        for (int k = 0; k < 2; k++) {
            auto it = index.find(nd.child[k]);
            if (it == index.end()) continue;
            blo[it->second] = k == 0 ? nd.c0lo : nd.c1lo;
            bhi[it->second] = k == 0 ? nd.c0hi : nd.c1hi;
        }
    out.box.resize(6 * out.ref.size());
    for (size_t e = 0; e < out.ref.size(); e++)
        for (int a = 0; a < 3; a++) {
            out.box[6 * e + a] = blo[e][a];
            out.box[6 * e + 3 + a] = bhi[e][a];
        }
}

}  // namespace xmc
