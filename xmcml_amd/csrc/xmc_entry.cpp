// Entry grid (see xmc_entry.h): per cell, the BVH restricted to the leaves a
// short step from the cell can reach.
#include "xmc_entry.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <thread>
#include <unordered_map>

namespace xmc {

namespace {

constexpr int32_t kDead = INT32_MIN;      // token: a child none of whose leaves is in reach
constexpr int32_t kGate = INT32_MIN + 1;  // token: a leaf at the end of a chain, next token
constexpr int32_t kWhole = INT32_MIN + 2; // token: the walk runs on the real subtree of the next token
#ifndef XMC_ENTRY_COPY_MIN_LEAVES
#define XMC_ENTRY_COPY_MIN_LEAVES 64
#endif
// Deepest common nodes with fewer leaves than this are used as they are:
// their chains are short, and copying every cell's tree would cost memory.
constexpr int kCopyMinLeaves = XMC_ENTRY_COPY_MIN_LEAVES;

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
// real node it copies, or kGate with the leaf in child 0.  Interned in an
// open-addressing table (millions of lookups per scene).
struct NodeKey {
    int32_t src, c0, c1;
};

class NodeTable {
public:
    // the value stored for k, or -1 with the slot to fill
    int32_t* find(const NodeKey& k) {
        if (2 * (count + 1) > keys.size()) grow();
        size_t m = keys.size() - 1;
        for (size_t i = hash(k) & m;; i = (i + 1) & m) {
            if (vals[i] < 0) return &vals[i];
            const NodeKey& o = keys[i];
            if (o.src == k.src && o.c0 == k.c0 && o.c1 == k.c1) return &vals[i];
        }
    }
    void set(int32_t* slot, const NodeKey& k, int32_t v) {
        keys[(size_t)(slot - vals.data())] = k;
        *slot = v;
        count++;
    }

private:
    static size_t hash(const NodeKey& k) {
        uint64_t h = (uint64_t)(uint32_t)k.src * 0x9E3779B97F4A7C15ull;
        h ^= ((uint64_t)(uint32_t)k.c0 << 32 | (uint32_t)k.c1) * 0xC2B2AE3D27D4EB4Full;
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ull;
        return (size_t)(h ^ (h >> 29));
    }
    void grow() {
        std::vector<NodeKey> ok(std::move(keys));
        std::vector<int32_t> ov(std::move(vals));
        size_t cap = ok.empty() ? (size_t)1 << 16 : 2 * ok.size();
        keys.assign(cap, NodeKey{0, 0, 0});
        vals.assign(cap, -1);
        count = 0;
        for (size_t i = 0; i < ok.size(); i++)
            if (ov[i] >= 0) set(find(ok[i]), ok[i], ov[i]);
    }
    std::vector<NodeKey> keys;
    std::vector<int32_t> vals;
    size_t count = 0;
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
    out.margin2 = 3.0f * h;
    const float margins[2] = {out.margin, out.margin2};
    // the kernel locates a point's cell in float, so a point may sit just outside the cell
    // it is filed under; its float hit tests also reach slightly beyond the exact segment
    const double guard = 1e-3 + 1e-5 * scale;
    const double hh = (double)h;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    // A step that moves at most `margin` along every axis is at most sqrt(3) margin long;
    // from a point at most sqrt(3)/2 h from its cell's centre (plus the kernel's rounding
    // allowance, at most 1e-4 + 16 (|p| 2^-20 + 2^-22)) it is skipped by the distance
    // bound when the bound exceeds the sum.  Such cells need no entry at that level.
    double grow[2], never[2];
    for (int L = 0; L < 2; L++) {
        grow[L] = (double)margins[L] + guard;
        never[L] = std::sqrt(3.0) * (double)margins[L] + 0.5 * std::sqrt(3.0) * (double)h + 1e-4 +
                   16.0 * (scale * 0x1p-20 + 0x1p-22) + 0.01 * (double)h;
    }
    out.cell.assign(n, -1);
    out.cell2.assign(n, -1);
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

    // leaves below every node (children come after their parent in the array)
    std::vector<int32_t> leaves(nn, 0);
    for (size_t k = nn; k-- > 0;)
        for (int c : bvh.nodes[k].child) leaves[k] += c >= 0 ? leaves[c] : 1;
    // pass 1, in parallel over z-slices: each cell's restricted trees (both levels),
    // deduplicated per thread; the cell keeps (thread, local tree) per level
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    using Map = std::unordered_map<std::vector<int32_t>, int32_t, VecHash>;
    std::vector<Map> maps(nth);
    std::vector<std::vector<const std::vector<int32_t>*>> lists(nth);
    std::vector<int32_t> local[2] = {std::vector<int32_t>(n, -1), std::vector<int32_t>(n, -1)};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++)
        pool.emplace_back([&, t]() {
            Restrictor R(bvh);
            std::vector<int32_t> tok;
            for (int iz = (int)t; iz < dims[2]; iz += (int)nth)
                for (int iy = 0; iy < dims[1]; iy++)
                    for (int ix = 0; ix < dims[0]; ix++)
                        for (int L = 0; L < 2; L++) {
                            const int ii[3] = {ix, iy, iz};
                            for (int a = 0; a < 3; a++) {
                                R.lo[a] = (double)lo[a] + ii[a] * hh - grow[L];
                                R.hi[a] = (double)lo[a] + (ii[a] + 1) * hh + grow[L];
                            }
                            const size_t c = ((size_t)iz * dims[1] + iy) * dims[0] + ix;
                            if (q && (double)((float)q[c] * step_q) > never[L]) continue;
                            if (!overlaps(bvh.root_lo, bvh.root_hi, R.lo, R.hi)) continue;
                            int top = R.bottom(bvh.root_ref);
                            if (top == kDead) continue;
                            tok.clear();
                            if (top >= 0 && leaves[top] < kCopyMinLeaves) {
                                tok.push_back(kWhole);
                                tok.push_back(top);
                            } else {
                                R.emit(top, tok);
                            }
                            auto it = maps[t].find(tok);
                            if (it == maps[t].end()) {
                                it = maps[t].emplace(tok, (int32_t)lists[t].size()).first;
                                lists[t].push_back(&it->first);
                            }
                            local[L][c] = it->second;
                        }
        });
    for (auto& th : pool) th.join();

    // pass 2: number the distinct trees and lay them out as nodes appended to the BVH's,
    // sharing equal nodes.  A copy of a branching node keeps its children's boxes (so
    // the walk orders and prunes as the walk from the root does) and points to the
    // children's restricted trees, or kDeadRef for a child with no leaf in reach.  A gate
    // tests a chain-ending leaf's box in its first slot; its second box lies at 1e30,
    // beyond any step.
    // In parallel over shards: shard s owns the distinct trees whose start node is s mod
    // nth (equal trees meet in one shard, and trees below the same node, which share most
    // of their nodes, too), numbers them and builds their nodes with its own table; a node
    // shared between shards is stored once per shard.  Shard-local node references (from
    // nn on) are moved into place afterwards.
    auto far_node = [](GpuNode& g) {
        std::memset(&g, 0, sizeof(g));
        for (int a = 0; a < 3; a++) {
            g.c0lo[a] = g.c0hi[a] = 1e30f;
            g.c1lo[a] = g.c1hi[a] = 1e30f;
        }
    };
    struct PtrHash {
        size_t operator()(const std::vector<int32_t>* v) const { return VecHash()(*v); }
    };
    struct PtrEq {
        bool operator()(const std::vector<int32_t>* a, const std::vector<int32_t>* b) const { return *a == *b; }
    };
    struct Shard {
        std::vector<GpuNode> nodes;
        std::vector<int32_t> ref;
        std::vector<float> box;
    };
    std::vector<Shard> shards(nth);
    std::vector<std::vector<int32_t>> tree_of(nth);  // per thread's list: shard-local tree id
    std::vector<std::vector<size_t>> hashes(nth);
    for (unsigned t = 0; t < nth; t++) {
        tree_of[t].assign(lists[t].size(), -1);
        hashes[t].resize(lists[t].size());
    }
    pool.clear();
    for (unsigned t = 0; t < nth; t++)
        pool.emplace_back([&, t]() {
            for (size_t j = 0; j < lists[t].size(); j++) {
                const std::vector<int32_t>& tok = *lists[t][j];
                int32_t top = tok[0] == kWhole ? tok[1] : tok[0];
                hashes[t][j] = (size_t)(((uint64_t)(uint32_t)top * 0x9E3779B97F4A7C15ull) >> 32);
            }
        });
    for (auto& th : pool) th.join();
    pool.clear();
    for (unsigned sh = 0; sh < nth; sh++)
        pool.emplace_back([&, sh]() {
            Shard& S = shards[sh];
            NodeTable interned;
            auto intern = [&](NodeKey k) -> int32_t {
                int32_t* slot = interned.find(k);
                if (*slot >= 0) return *slot;
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
                int32_t ref = (int32_t)(nn + S.nodes.size());
                S.nodes.push_back(g);
                interned.set(slot, k, ref);
                return ref;
            };
            // builds the nodes of the tokens from position i on; returns the reference
            std::function<int32_t(const std::vector<int32_t>&, size_t&)> make = [&](const std::vector<int32_t>& tok,
                                                                                   size_t& i) -> int32_t {
                int32_t t = tok[i++];
                if (t == kDead) return kDeadRef;
                if (t == kGate) {
                    int32_t leaf = tok[i++];
                    return intern({kGate, leaf, leaf});
                }
                if (t < 0) return t;
                int32_t c0 = make(tok, i);
                int32_t c1 = make(tok, i);
                return intern({t, c0, c1});
            };
            std::unordered_map<const std::vector<int32_t>*, int32_t, PtrHash, PtrEq> seen;
            for (unsigned t = 0; t < nth; t++)
                for (size_t j = 0; j < lists[t].size(); j++) {
                    if (hashes[t][j] % nth != sh) continue;
                    const std::vector<int32_t>& tok = *lists[t][j];
                    auto it = seen.find(&tok);
                    if (it == seen.end()) {
                        it = seen.emplace(&tok, (int32_t)S.ref.size()).first;
                        const bool whole = tok[0] == kWhole;
                        int32_t top = whole ? tok[1] : tok[0];
                        size_t i = 0;
                        S.ref.push_back(top >= 0 && !whole ? make(tok, i) : top);
                        const float* bl = top >= 0 ? nlo[top] : llo[~top];
                        const float* bh = top >= 0 ? nhi[top] : lhi[~top];
                        for (int a = 0; a < 3; a++) S.box.push_back(bl[a]);
                        for (int a = 0; a < 3; a++) S.box.push_back(bh[a]);
                    }
                    tree_of[t][j] = it->second;
                }
        });
    for (auto& th : pool) th.join();
    // lay the shards out one after another
    std::vector<size_t> node_base(nth + 1, 0), tree_base(nth + 1, 0);
    for (unsigned sh = 0; sh < nth; sh++) {
        node_base[sh + 1] = node_base[sh] + shards[sh].nodes.size();
        tree_base[sh + 1] = tree_base[sh] + shards[sh].ref.size();
    }
    if (nn + node_base[nth] >= (size_t)kDeadRef) throw std::length_error("entry trees: too many nodes");
    out.extra.resize(node_base[nth]);
    out.ref.resize(tree_base[nth]);
    out.box.resize(6 * tree_base[nth]);
    auto move = [&](int32_t r, unsigned sh) -> int32_t {
        return (r >= (int32_t)nn && r != kDeadRef) ? (int32_t)(r + node_base[sh]) : r;
    };
    pool.clear();
    for (unsigned sh = 0; sh < nth; sh++)
        pool.emplace_back([&, sh]() {
            Shard& S = shards[sh];
            for (size_t k = 0; k < S.nodes.size(); k++) {
                GpuNode g = S.nodes[k];
                g.child[0] = move(g.child[0], sh);
                g.child[1] = move(g.child[1], sh);
                out.extra[node_base[sh] + k] = g;
            }
            for (size_t k = 0; k < S.ref.size(); k++) out.ref[tree_base[sh] + k] = move(S.ref[k], sh);
            std::copy(S.box.begin(), S.box.end(), out.box.begin() + 6 * tree_base[sh]);
            Shard().nodes.swap(S.nodes);
        });
    for (auto& th : pool) th.join();
    std::vector<std::vector<int32_t>> to_global(nth);
    for (unsigned t = 0; t < nth; t++) {
        to_global[t].resize(lists[t].size());
        for (size_t j = 0; j < lists[t].size(); j++)
            to_global[t][j] = (int32_t)(tree_base[hashes[t][j] % nth] + tree_of[t][j]);
    }
    for (int iz = 0; iz < dims[2]; iz++) {
        unsigned t = (unsigned)iz % nth;
        for (size_t c = (size_t)iz * dims[1] * dims[0], e = c + (size_t)dims[1] * dims[0]; c < e; c++) {
            if (local[0][c] >= 0) out.cell[c] = to_global[t][local[0][c]];
            if (local[1][c] >= 0) out.cell2[c] = to_global[t][local[1][c]];
        }
    }
    pool.clear();
    for (unsigned t = 0; t < nth; t++) pool.emplace_back([&, t]() { Map().swap(maps[t]); });
    for (auto& th : pool) th.join();
}

}  // namespace xmc
