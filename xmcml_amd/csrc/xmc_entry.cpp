// Entry grid (see xmc_entry.h): per cell, the BVH restricted to the leaves a
// short step from the cell can reach.
#include "xmc_entry.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>
#include <unordered_map>

namespace xmc {

namespace {

constexpr int32_t kDead = INT32_MIN;      // token: a child none of whose leaves is in reach
constexpr int32_t kGate = INT32_MIN + 1;  // token: a leaf at the end of a chain, next token

inline bool overlaps(const float lo[3], const float hi[3], const double blo[3], const double bhi[3]) {
    for (int a = 0; a < 3; a++)
        if ((double)hi[a] < blo[a] || (double)lo[a] > bhi[a]) return false;
    return true;
}

struct VecHash {
    size_t operator()(const std::vector<int32_t>& v) const {
        uint64_t h = 1469598103934665603ull;
        for (int32_t x : v) h = (h ^ (uint32_t)x) * 1099511628211ull;
        return (size_t)(h ^ (h >> 29));
    }
};

// A restricted-tree node is fixed by (source, child 0, child 1): source = the
// real node it copies, or kGate with the leaf in child 0, or kDead.
struct NodeKey {
    int32_t src, c0, c1;
    bool operator==(const NodeKey& o) const { return src == o.src && c0 == o.c0 && c1 == o.c1; }
};
struct NodeHash {
    size_t operator()(const NodeKey& k) const {
        uint64_t h = ((uint64_t)(uint32_t)k.src * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)(uint32_t)k.c0 << 32 | (uint32_t)k.c1);
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ull;
        return (size_t)(h ^ (h >> 29));
    }
};

// The restricted tree of one region [blo, bhi] as tokens in preorder: a
// branching node (both children's boxes meet the region) is its reference
// followed by its two children; a leaf met as a child of a branching node is
// its (negative) reference; a leaf at the end of a chain of nodes with one
// child in the region is kGate, reference; a child with no leaf in the region
// is kDead.  Chains of internal nodes are skipped: a walk entering the node
// at the chain's end tests that node's children, whose boxes lie inside every
// box of the chain.
class Restrictor {
public:
    Restrictor(const GpuBvh& b) : bvh(b) {}

    // The chain's end below child `ref`: the first node with both children in
    // the region, or a leaf; kDead when the chain leaves the region.
    int bottom(int ref) const {
        while (ref >= 0) {
            const GpuNode& nd = bvh.nodes[ref];
            bool o0 = overlaps(nd.c0lo, nd.c0hi, lo, hi), o1 = overlaps(nd.c1lo, nd.c1hi, lo, hi);
            if (o0 && o1) return ref;
            if (o0) ref = nd.child[0];
            else if (o1) ref = nd.child[1];
            else return kDead;
        }
        return ref;
    }

    // Tokens of the tree below the branching node or leaf `b`.
    void emit(int b, std::vector<int32_t>& out) const {
        if (b < 0) {
            out.push_back(b);
            return;
        }
        out.push_back(b);
        const GpuNode& nd = bvh.nodes[b];
        for (int k = 0; k < 2; k++) {
            int ch = nd.child[k];
            int bt = bottom(ch);
            if (bt == kDead) {
                out.push_back(kDead);
            } else if (bt != ch && bt < 0) {
                out.push_back(kGate);
                out.push_back(bt);
            } else {
                emit(bt, out);
            }
        }
    }

    const GpuBvh& bvh;
    double lo[3], hi[3];
};

}  // namespace

void build_entry_grid(const GpuBvh& bvh, const float lo[3], float h, const int dims[3], const uint8_t* q, float step_q,
                      EntryGrid& out) {
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
    // A step that moves at most `margin` along every axis is at most sqrt(3) margin long;
    // from a point at most sqrt(3)/2 h from its cell's centre (plus the kernel's rounding
    // allowance, at most 1e-4 + 16 (|p| 2^-20 + 2^-22)) it is skipped by the distance
    // bound when the bound exceeds the sum.  Such cells need no entry.
    const double never = std::sqrt(3.0) * (double)out.margin + 0.5 * std::sqrt(3.0) * (double)h + 1e-4 +
                         16.0 * (scale * 0x1p-20 + 0x1p-22) + 0.01 * (double)h;
    out.cell.assign(n, -1);
    out.ref.clear();
    out.box.clear();
    out.extra.clear();
    if (bvh.root_ref < 0) return;
    // the box of every node and leaf as its parent stores it (the root's: the root box)
    const size_t nn = bvh.nodes.size(), nt = bvh.tris.size();
    std::vector<const float*> nlo(nn, nullptr), nhi(nn, nullptr), llo(nt, nullptr), lhi(nt, nullptr);
    nlo[bvh.root_ref] = bvh.root_lo;
    nhi[bvh.root_ref] = bvh.root_hi;
    for (const GpuNode& nd : bvh.nodes)
        for (int k = 0; k < 2; k++) {
            int c = nd.child[k];
            const float* bl = k == 0 ? nd.c0lo : nd.c1lo;
            const float* bh = k == 0 ? nd.c0hi : nd.c1hi;
            if (c >= 0) {
                nlo[c] = bl;
                nhi[c] = bh;
            } else {
                llo[~c] = bl;
                lhi[~c] = bh;
            }
        }

    // pass 1, in parallel over z-slices: each cell's restricted tree, deduplicated per
    // thread; the cell keeps (thread, local tree)
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    using Map = std::unordered_map<std::vector<int32_t>, int32_t, VecHash>;
    std::vector<Map> maps(nth);
    std::vector<std::vector<const std::vector<int32_t>*>> lists(nth);
    std::vector<int32_t> local(n, -1);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++)
        pool.emplace_back([&, t]() {
            Restrictor R(bvh);
            std::vector<int32_t> tok;
            for (int iz = (int)t; iz < dims[2]; iz += (int)nth)
                for (int iy = 0; iy < dims[1]; iy++)
                    for (int ix = 0; ix < dims[0]; ix++) {
                        const int ii[3] = {ix, iy, iz};
                        for (int a = 0; a < 3; a++) {
                            R.lo[a] = (double)lo[a] + ii[a] * hh - grow;
                            R.hi[a] = (double)lo[a] + (ii[a] + 1) * hh + grow;
                        }
                        const size_t c = ((size_t)iz * dims[1] + iy) * dims[0] + ix;
                        if (q && (double)((float)q[c] * step_q) > never) continue;
                        if (!overlaps(bvh.root_lo, bvh.root_hi, R.lo, R.hi)) continue;
                        int top = R.bottom(bvh.root_ref);
                        if (top == kDead) continue;
                        tok.clear();
                        R.emit(top, tok);
                        auto it = maps[t].find(tok);
                        if (it == maps[t].end()) {
                            it = maps[t].emplace(tok, (int32_t)lists[t].size()).first;
                            lists[t].push_back(&it->first);
                        }
                        local[c] = it->second;
                    }
        });
    for (auto& th : pool) th.join();

    // pass 2: number the distinct trees and lay them out as nodes appended to the BVH's,
    // sharing equal nodes.  A copy of a branching node keeps its children's boxes (so
    // the walk orders and prunes as the walk from the root does) and points to the
    // children's restricted trees.  A gate tests a chain-ending leaf's box in its first
    // slot; its second box and a dead child's boxes lie at 1e30, beyond any step.
    std::unordered_map<NodeKey, int32_t, NodeHash> interned;
    interned.reserve(1024);
    auto far_node = [](GpuNode& g) {
        std::memset(&g, 0, sizeof(g));
        for (int a = 0; a < 3; a++) {
            g.c0lo[a] = g.c0hi[a] = 1e30f;
            g.c1lo[a] = g.c1hi[a] = 1e30f;
        }
    };
    auto intern = [&](NodeKey k) -> int32_t {
        auto it = interned.find(k);
        if (it != interned.end()) return it->second;
        GpuNode g;
        if (k.src >= 0) {
            g = bvh.nodes[k.src];
            g.pad[0] = g.pad[1] = 0;
        } else {
            far_node(g);
            if (k.src == kGate)
                for (int a = 0; a < 3; a++) {
                    g.c0lo[a] = llo[~k.c0][a];
                    g.c0hi[a] = lhi[~k.c0][a];
                }
        }
        g.child[0] = k.c0;
        g.child[1] = k.c1;
        int32_t ref = (int32_t)(nn + out.extra.size());
        out.extra.push_back(g);
        interned.emplace(k, ref);
        return ref;
    };
    // builds the nodes of the tokens from position i on; returns the reference
    std::function<int32_t(const std::vector<int32_t>&, size_t&)> make = [&](const std::vector<int32_t>& tok,
                                                                           size_t& i) -> int32_t {
        int32_t t = tok[i++];
        if (t == kDead) return intern({kDead, 0, 0});
        if (t == kGate) {
            int32_t leaf = tok[i++];
            return intern({kGate, leaf, leaf});
        }
        if (t < 0) return t;
        int32_t c0 = make(tok, i);
        int32_t c1 = make(tok, i);
        return intern({t, c0, c1});
    };
    std::unordered_map<std::vector<int32_t>, int32_t, VecHash> global;
    std::vector<std::vector<int32_t>> to_global(nth);
    for (unsigned t = 0; t < nth; t++) {
        to_global[t].resize(lists[t].size());
        for (size_t j = 0; j < lists[t].size(); j++) {
            const std::vector<int32_t>& tok = *lists[t][j];
            auto it = global.find(tok);
            if (it == global.end()) {
                int32_t e = (int32_t)out.ref.size();
                it = global.emplace(tok, e).first;
                int32_t top = tok[0] == kGate ? tok[1] : tok[0];
                size_t i = 0;
                int32_t ref = tok[0] >= 0 ? make(tok, i) : top;
                out.ref.push_back(ref);
                const float* bl = top >= 0 ? nlo[top] : llo[~top];
                const float* bh = top >= 0 ? nhi[top] : lhi[~top];
                for (int a = 0; a < 3; a++) out.box.push_back(bl[a]);
                for (int a = 0; a < 3; a++) out.box.push_back(bh[a]);
            }
            to_global[t][j] = it->second;
        }
    }
    for (int iz = 0; iz < dims[2]; iz++) {
        unsigned t = (unsigned)iz % nth;
        for (size_t c = (size_t)iz * dims[1] * dims[0], e = c + (size_t)dims[1] * dims[0]; c < e; c++)
            if (local[c] >= 0) out.cell[c] = to_global[t][local[c]];
    }
}

}  // namespace xmc
