// Host-side BVH build for the GPU engine.
//
// The tree is the one xmcml's GenerateBVHTree builds
// (base/xmcml/xmcml/mcml_intersection.cpp:88-344): 16-bin centroid SAH over the
// three axes, leaves of at most 5 triangles or where SAH >= triangle count,
// the same in-place two-pointer partition, children emitted before their
// parent.  Reproducing the same tree (node boxes, partition order) keeps the
// GPU walk's visiting order -- and with it the choice between equally distant
// triangles -- identical to the reference's.  This builder precomputes
// per-triangle centroids and bounds once instead of re-reading the vertices
// at every level; the values are the same expressions, so the bins are too.
#include "xmc_bvh.h"

#include <cstring>

namespace xmc {
namespace {

constexpr int kBins = 16;        // BIN_NUMBER, mcml_intersection.h:8
constexpr int kLeafMax = 5;      // MAX_TRIANGLES_IN_LEAVES, :9
constexpr double kEps = 1.0e-6;  // EPSILON, :6
constexpr double kBig = 1.0e256; // MAX_DISTANCE, :7

struct Builder {
    std::vector<int> order;        // global triangle ids, permuted in place
    std::vector<double> centroid;  // 3 per global id
    std::vector<double> tlo, thi;  // per-triangle bounds, 3 per global id
    std::vector<RefNode>* nodes = nullptr;
    int max_depth = 0;
    bool degenerate = false;

    void bounds(int off, int n, double lo[3], double hi[3]) const {
        for (int a = 0; a < 3; a++) {
            lo[a] = kBig;
            hi[a] = -kBig;
        }
        for (int i = 0; i < n; i++) {
            int g = order[off + i];
            for (int a = 0; a < 3; a++) {
                if (tlo[3 * g + a] < lo[a]) lo[a] = tlo[3 * g + a];
                if (thi[3 * g + a] > hi[a]) hi[a] = thi[3 * g + a];
            }
        }
    }

    int bin(int g, int axis, double lo, double d) const {
        int b = (int)((centroid[3 * g + axis] - lo) / d);
        if (b == kBins) b--;
        return b;
    }

    // FindBVHSplit, mcml_intersection.cpp:125-200
    void split(const double lo[3], const double hi[3], int off, int n, double& min_sah, int& split_bin,
               int& split_axis) const {
        int bins[3][kBins];
        std::memset(bins, 0, sizeof(bins));
        double size[3], d[3];
        for (int a = 0; a < 3; a++) {
            size[a] = hi[a] - lo[a];
            d[a] = size[a] / kBins;
        }
        for (int i = 0; i < n; i++) {
            int g = order[off + i];
            for (int a = 0; a < 3; a++) bins[a][d[a] > kEps ? bin(g, a, lo[a], d[a]) : 0]++;
        }
        double inv_area = 0.5 / (size[0] * size[1] + size[1] * size[2] + size[0] * size[2]);
        min_sah = kBig;
        for (int a = 0; a < 3; a++) {
            if (!(d[a] > kEps)) continue;
            int left = 0, right = n;
            double full = size[a];
            for (int i = 1; i < kBins; i++) {
                left += bins[a][i - 1];
                right -= bins[a][i - 1];
                if (left > 0 && right > 0) {
                    size[a] = i * d[a];
                    double la = 2.0 * (size[0] * size[1] + size[1] * size[2] + size[2] * size[0]);
                    size[a] = full - size[a];
                    double ra = 2.0 * (size[0] * size[1] + size[1] * size[2] + size[2] * size[0]);
                    double sah = inv_area * (left * la + ra * right);
                    if (sah < min_sah) {
                        min_sah = sah;
                        split_axis = a;
                        split_bin = i;
                    }
                }
            }
            size[a] = full;
        }
    }

    int emit_leaf(const double lo[3], const double hi[3], int off, int n) {
        RefNode nd;
        std::memcpy(nd.lo, lo, sizeof(nd.lo));
        std::memcpy(nd.hi, hi, sizeof(nd.hi));
        nd.left = nd.right = -1;
        nd.offset = off;
        nd.count = n;
        nodes->push_back(nd);
        return (int)nodes->size() - 1;
    }

    // GenerateBVHNode, mcml_intersection.cpp:202-314
    int node(int off, int n, int depth) {
        if (depth > max_depth) max_depth = depth;
        double lo[3], hi[3];
        bounds(off, n, lo, hi);
        if (n <= kLeafMax) return emit_leaf(lo, hi, off, n);
        int sb = 0, sa = 0;
        double min_sah;
        split(lo, hi, off, n, min_sah, sb, sa);
        if ((double)n <= min_sah) return emit_leaf(lo, hi, off, n);
        // two-pointer partition by bin index, exactly as the reference walks it
        double d = (hi[sa] - lo[sa]) / kBins;
        int* t = order.data() + off;
        int k = 0, j = n - 1, lc = 0;
        int ik = bin(t[k], sa, lo[sa], d), ij = bin(t[j], sa, lo[sa], d);
        while (k < j) {
            if (ik < sb) {
                k++;
                lc++;
                if (k < n) ik = bin(t[k], sa, lo[sa], d);
            }
            if (ij >= sb) {
                j--;
                if (j >= 0) ij = bin(t[j], sa, lo[sa], d);
            }
            if (ik >= sb && ij < sb) {
                int tmp = t[k];
                t[k] = t[j];
                t[j] = tmp;
                k++;
                j--;
                lc++;
                if (k < n) ik = bin(t[k], sa, lo[sa], d);
                if (j >= 0) ij = bin(t[j], sa, lo[sa], d);
            }
        }
        if (lc <= 0 || lc >= n) {
            // the reference would recurse forever here; keep a leaf instead
            degenerate = true;
            return emit_leaf(lo, hi, off, n);
        }
        RefNode nd;
        std::memcpy(nd.lo, lo, sizeof(nd.lo));
        std::memcpy(nd.hi, hi, sizeof(nd.hi));
        nd.count = 0;
        nd.offset = -1;
        nd.left = node(off, lc, depth + 1);
        nd.right = node(off + lc, n - lc, depth + 1);
        nodes->push_back(nd);
        return (int)nodes->size() - 1;
    }
};

}  // namespace

bool build_reference_bvh(const xmc_surface* s, int ns, RefBvh& out, const char** err) {
    Builder b;
    b.nodes = &out.nodes;
    out.nodes.clear();
    std::vector<int> base(ns + 1, 0);
    for (int i = 0; i < ns; i++) base[i + 1] = base[i] + s[i].numberOfTriangles;
    int nt = base[ns];
    if (nt <= 0) {
        *err = "no triangles";
        return false;
    }
    b.order.resize(nt);
    b.centroid.resize(3 * (size_t)nt);
    b.tlo.resize(3 * (size_t)nt);
    b.thi.resize(3 * (size_t)nt);
    out.tri_surface.resize(nt);
    out.tri_index.resize(nt);
    for (int i = 0; i < ns; i++) {
        for (int j = 0; j < s[i].numberOfTriangles; j++) {
            int g = base[i] + j;
            b.order[g] = g;
            const xmc_int3& tr = s[i].triangles[j];
            if (tr.x < 0 || tr.y < 0 || tr.z < 0 || tr.x >= s[i].numberOfVertices ||
                tr.y >= s[i].numberOfVertices || tr.z >= s[i].numberOfVertices) {
                *err = "triangle vertex index out of range";
                return false;
            }
            const xmc_double3 v[3] = {s[i].vertices[tr.x], s[i].vertices[tr.y], s[i].vertices[tr.z]};
            for (int a = 0; a < 3; a++) {
                double c0 = a == 0 ? v[0].x : (a == 1 ? v[0].y : v[0].z);
                double c1 = a == 0 ? v[1].x : (a == 1 ? v[1].y : v[1].z);
                double c2 = a == 0 ? v[2].x : (a == 1 ? v[2].y : v[2].z);
                b.centroid[3 * (size_t)g + a] = (c0 + c1 + c2) / 3.0;
                double lo = c0, hi = c0;
                if (c1 < lo) lo = c1;
                if (c2 < lo) lo = c2;
                if (c1 > hi) hi = c1;
                if (c2 > hi) hi = c2;
                b.tlo[3 * (size_t)g + a] = lo;
                b.thi[3 * (size_t)g + a] = hi;
            }
            out.tri_surface[g] = i;
            out.tri_index[g] = j;
        }
    }
    out.root = b.node(0, nt, 1);
    out.depth = b.max_depth;
    std::vector<int> ts(nt), ti(nt);
    for (int k = 0; k < nt; k++) {
        ts[k] = out.tri_surface[b.order[k]];
        ti[k] = out.tri_index[b.order[k]];
    }
    out.tri_surface.swap(ts);
    out.tri_index.swap(ti);
    if (b.degenerate) *err = nullptr;  // tree still valid; documented divergence
    return true;
}

namespace {

void set_box(float lo[3], float hi[3], const RefNode& n) {
    for (int a = 0; a < 3; a++) {
        lo[a] = (float)n.lo[a];
        hi[a] = (float)n.hi[a];
    }
}

struct Flattener {
    const RefBvh& ref;
    GpuBvh& out;
    int child_ref(int r) {
        const RefNode& n = ref.nodes[r];
        if (n.offset >= 0) return ~n.offset;
        return emit(r);
    }
    int emit(int r) {
        int idx = (int)out.nodes.size();
        out.nodes.emplace_back();
        const RefNode& n = ref.nodes[r];
        GpuNode g;
        std::memset(&g, 0, sizeof(g));
        set_box(g.c0lo, g.c0hi, ref.nodes[n.left]);
        set_box(g.c1lo, g.c1hi, ref.nodes[n.right]);
        g.child[0] = child_ref(n.left);
        g.child[1] = child_ref(n.right);
        out.nodes[idx] = g;
        return idx;
    }
};

}  // namespace

void flatten_bvh(const RefBvh& ref, const xmc_surface* s, GpuBvh& out) {
    out.nodes.clear();
    out.tris.clear();
    out.depth = ref.depth;
    const RefNode& root = ref.nodes[ref.root];
    set_box(out.root_lo, out.root_hi, root);
    Flattener f{ref, out};
    out.root_ref = f.child_ref(ref.root);
    size_t nt = ref.tri_surface.size();
    out.tris.resize(nt);
    for (size_t i = 0; i < nt; i++) {
        const xmc_surface& sf = s[ref.tri_surface[i]];
        const xmc_int3& tr = sf.triangles[ref.tri_index[i]];
        float a[3] = {(float)sf.vertices[tr.x].x, (float)sf.vertices[tr.x].y, (float)sf.vertices[tr.x].z};
        float b[3] = {(float)sf.vertices[tr.y].x, (float)sf.vertices[tr.y].y, (float)sf.vertices[tr.y].z};
        float c[3] = {(float)sf.vertices[tr.z].x, (float)sf.vertices[tr.z].y, (float)sf.vertices[tr.z].z};
        GpuTri& g = out.tris[i];
        for (int k = 0; k < 3; k++) {
            g.a[k] = a[k];
            g.e1[k] = b[k] - a[k];
            g.e2[k] = c[k] - a[k];
        }
        g.surface = ref.tri_surface[i];
        g.index = ref.tri_index[i];
        g.last = 0;
    }
    for (const RefNode& n : ref.nodes)
        if (n.offset >= 0 && n.count > 0) out.tris[n.offset + n.count - 1].last = 1;
}

}  // namespace xmc
