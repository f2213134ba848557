// gfx950 kernels of the photon random-walk engine.
//
// transport_kernel is a persistent kernel: every lane carries one photon and,
// when its photon dies, takes the next photon index from a device counter
// (one atomic per wavefront per refill, ranks from mbcnt), so wavefronts stay
// full while photon lifetimes differ by orders of magnitude.  Each lane runs
// xmcml's ComputePhoton loop (base/xmcml/xmcml/mcml_kernel.cpp:39-126) in
// float: step, nearest surface hit through the BVH, then either a Fresnel
// boundary interaction or hop/drop/spin, then termination.  The arithmetic of
// each formula follows the reference's expression order with no FMA
// contraction, so that a float CPU restatement reproduces it bit for bit; the
// float-only rules (24-bit uniforms, own log and sincos(2 pi u), accurate
// 1-x^2, coordinate-scaled surface offsets) are stated where they apply.
//
// Random numbers: photon p of seed s owns the MCG59 block that starts at
// (2s+1) * a^(p * 2^24) mod 2^59 (a = 13^13, mcml_mcg59.cpp:3-41); the lane
// jumps there with a table of squarings and then draws x <- a x mod 2^59.
//
// BVH walk: ordered front-to-back traversal with the reference's visiting
// and pruning rules (mcml_intersection.cpp:354-584) over nodes that hold both
// child boxes; the per-lane stack keeps its top 8 entries in LDS and spills
// deeper entries to a per-lane global area.
#include <hip/hip_runtime.h>

#include <mutex>

#include "xmc_params.h"

namespace xmc {

constexpr float kInf = __builtin_inff();
#ifndef XMC_PARK_HITS
#define XMC_PARK_HITS 1
#endif
#ifndef XMC_PARK_MIN
#define XMC_PARK_MIN 16
#endif
#ifndef XMC_PARK_WAIT
#define XMC_PARK_WAIT 16
#endif
#ifndef XMC_START_MIN
#define XMC_START_MIN 8
#endif
#ifndef XMC_START_WAIT
#define XMC_START_WAIT 8
#endif
#ifndef XMC_VISITS_PER_ROUND
#define XMC_VISITS_PER_ROUND 1
#endif
constexpr float kEpsF = 1.0e-6f;                 // EPSILON
constexpr float kCosZeroF = 1.0f - 1.0e-6f;      // float "cosine of zero"
constexpr float kCos90F = 1.0e-12f;              // COSINE_OF_90D
constexpr float kMinDistF = 1.0e-8f;             // MIN_DISTANCE

// --------------------------------------------------------------------------
// float special functions
// log: x = 2^k (1+f), 1+f in [sqrt(2)/2, sqrt(2)); log(1+f) = 2 atanh(s),
// s = f/(2+f) = f - (f^2/2 - s (f^2/2 + R)), R the atanh series to s^8.
// --------------------------------------------------------------------------
__device__ __forceinline__ float xlogf(float x) {
    uint32_t ix = __float_as_uint(x);
    int32_t t = (int32_t)(ix - 0x3f3504f3u);
    int32_t k = t >> 23;
    uint32_t im = ix - ((uint32_t)k << 23);
    float m = __uint_as_float(im);
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = __builtin_fmaf(z, 2.0f / 9.0f, 2.0f / 7.0f);
    R = __builtin_fmaf(z, R, 2.0f / 5.0f);
    R = __builtin_fmaf(z, R, 2.0f / 3.0f);
    R = z * R;
    float hfsq = 0.5f * (f * f);
    float l1p = f - (hfsq - s * (hfsq + R));
    float kf = (float)k;
    return __builtin_fmaf(kf, 0.693145751953125f, __builtin_fmaf(kf, 1.42860677e-06f, l1p));
}

// exp(x): x = k ln2 + r with k = rint(x log2 e) and the same two-part ln2 as
// xlogf (Cody-Waite), exp(r) for |r| <= ln2/2 by its Taylor series to r^7
// (truncation below 6e-9 relative), scaled by 2^k through the exponent bits.
// Below -87 the result is taken as 0 (the normal range ends near -87.3).
__device__ __forceinline__ float xexpf(float x) {
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) return kInf;
    float kf = __builtin_rintf(x * 1.44269504f);
    float r = __builtin_fmaf(kf, -0.693145751953125f, x);
    r = __builtin_fmaf(kf, -1.42860677e-06f, r);
    float p = __builtin_fmaf(r, 1.0f / 5040.0f, 1.0f / 720.0f);
    p = __builtin_fmaf(r, p, 1.0f / 120.0f);
    p = __builtin_fmaf(r, p, 1.0f / 24.0f);
    p = __builtin_fmaf(r, p, 1.0f / 6.0f);
    p = __builtin_fmaf(r, p, 0.5f);
    p = __builtin_fmaf(r, p, 1.0f);
    p = __builtin_fmaf(r, p, 1.0f);
    return p * __uint_as_float((uint32_t)((int)kf + 127) << 23);
}

// sin and cos of 2 pi u, u in [0, 1): exact quadrant split q = rint(4u),
// theta = (4u - q) pi/2 in [-pi/4, pi/4], Taylor series to theta^9 / theta^10.
__device__ __forceinline__ void xsincos2pi(float u, float& sp, float& cp) {
    float u4 = 4.0f * u;
    float q = __builtin_rintf(u4);
    float r = u4 - q;
    int iq = ((int)q) & 3;
    float th = r * 1.57079637f;
    float t2 = th * th;
    float ps = __builtin_fmaf(t2, 1.0f / 362880.0f, -1.0f / 5040.0f);
    ps = __builtin_fmaf(t2, ps, 1.0f / 120.0f);
    ps = __builtin_fmaf(t2, ps, -1.0f / 6.0f);
    float s = __builtin_fmaf(th * t2, ps, th);
    float pc = __builtin_fmaf(t2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
    pc = __builtin_fmaf(t2, pc, -1.0f / 720.0f);
    pc = __builtin_fmaf(t2, pc, 1.0f / 24.0f);
    pc = __builtin_fmaf(t2, pc, -0.5f);
    float c = __builtin_fmaf(t2, pc, 1.0f);
    float ss = (iq & 1) ? c : s;
    float cc = (iq & 1) ? s : c;
    sp = (iq == 2 || iq == 3) ? -ss : ss;
    cp = (iq == 1 || iq == 2) ? -cc : cc;
}

__device__ __forceinline__ float one_minus_sq(float x) { return (1.0f - x) * (1.0f + x); }

// MCG59 draw: top 24 bits of the 59-bit state.
__device__ __forceinline__ float next_u(uint64_t& x) {
    x = (x * kMcgA) & kMcgMask;
    return (float)(uint32_t)(x >> 35) * 0x1p-24f;
}

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// float rule: dot and cross products are evaluated with fused multiply-adds,
// dot = fma(a.z, b.z, fma(a.y, b.y, a.x b.x)), cross_i = fma(a_j, b_k, -(a_k b_j))
__device__ __forceinline__ float dot(V3 a, V3 b) {
    return __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x));
}
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {__builtin_fmaf(a.y, b.z, -(a.z * b.y)), __builtin_fmaf(a.z, b.x, -(a.x * b.z)),
            __builtin_fmaf(a.x, b.y, -(a.y * b.x))};
}
__device__ __forceinline__ V3 normalize(V3 a) {
    float l = __builtin_sqrtf(dot(a, a));
    return {a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ float offset_scale(V3 p) {
    float m = fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z)));
    return __builtin_fmaf(m, 0x1p-20f, 0x1p-22f);
}

// --------------------------------------------------------------------------
// geometry
// --------------------------------------------------------------------------
struct Ray {
    V3 o, d, inv;
    bool ax, ay, az;  // |d| > EPSILON per axis
    V3 oi;            // o * inv per axis, for the slab distances fma(bound, inv, -o inv)
};

__device__ __forceinline__ void set_oi(Ray& r) {
    r.oi = {r.o.x * r.inv.x, r.o.y * r.inv.y, r.o.z * r.inv.z};
}

__device__ __forceinline__ Ray make_ray(V3 o, V3 d) {
    Ray r;
    r.o = o;
    r.d = d;
    r.ax = fabsf(d.x) > kEpsF;
    r.ay = fabsf(d.y) > kEpsF;
    r.az = fabsf(d.z) > kEpsF;
    r.inv.x = r.ax ? 1.0f / d.x : 0.0f;
    r.inv.y = r.ay ? 1.0f / d.y : 0.0f;
    r.inv.z = r.az ? 1.0f / d.z : 0.0f;
    set_oi(r);
    return r;
}

// IntersectAABB (mcml_intersection.cpp:11-42).  Slab tests of the three axes
// without the early exit: tmin only grows and tmax only shrinks, so the
// result and t_near are the same.  Float rule: the slab distance (b - o) / d
// is fma(b, 1/d, -(o * (1/d))).
__device__ __forceinline__ bool box_hit(const Ray& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                        float& tnear) {
    float tmin = -kInf, tmax = kInf;
    bool ok = true;
    if (r.ax) {
        float t1 = __builtin_fmaf(lx, r.inv.x, -r.oi.x), t2 = __builtin_fmaf(hx, r.inv.x, -r.oi.x);
        tmin = fmaxf(tmin, fminf(t1, t2));
        tmax = fminf(tmax, fmaxf(t1, t2));
    } else {
        ok = ok && !((lx - r.o.x > 0.0f) || (hx - r.o.x < 0.0f));
    }
    if (r.ay) {
        float t1 = __builtin_fmaf(ly, r.inv.y, -r.oi.y), t2 = __builtin_fmaf(hy, r.inv.y, -r.oi.y);
        tmin = fmaxf(tmin, fminf(t1, t2));
        tmax = fminf(tmax, fmaxf(t1, t2));
    } else {
        ok = ok && !((ly - r.o.y > 0.0f) || (hy - r.o.y < 0.0f));
    }
    if (r.az) {
        float t1 = __builtin_fmaf(lz, r.inv.z, -r.oi.z), t2 = __builtin_fmaf(hz, r.inv.z, -r.oi.z);
        tmin = fmaxf(tmin, fminf(t1, t2));
        tmax = fminf(tmax, fmaxf(t1, t2));
    } else {
        ok = ok && !((lz - r.o.z > 0.0f) || (hz - r.o.z < 0.0f));
    }
    ok = ok && !(tmin > tmax) && !(tmax < 0.0f);
    tnear = ok ? tmin : kInf;
    return ok;
}

// GetTriangleIntersectionDistance (mcml_intersection.cpp:44-86) on the
// precomputed edges; returns +inf for a miss.
__device__ __forceinline__ float tri_distance(const GpuTri* tris, int i, const Ray& r, int& last) {
    const float4* T = reinterpret_cast<const float4*>(tris) + 3 * i;
    float4 A = T[0], B = T[1], C = T[2];
    last = __float_as_int(C.w);
    V3 a = {A.x, A.y, A.z}, e1 = {B.x, B.y, B.z}, e2 = {C.x, C.y, C.z};
    V3 p = cross(r.d, e2);
    float det = dot(e1, p);
    if (det < kEpsF && det > -kEpsF) return kInf;
    float inv = 1.0f / det;
    V3 t = sub(r.o, a);
    float u = dot(t, p) * inv;
    if (u < 0.0f || u > 1.0f) return kInf;
    V3 q = cross(t, e1);
    float v = dot(r.d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return kInf;
    return dot(e2, q) * inv;
}

__device__ __forceinline__ V3 tri_normal(const GpuTri* tris, int i) {
    const float4* T = reinterpret_cast<const float4*>(tris) + 3 * i;
    float4 B = T[1], C = T[2];
    V3 e1 = {B.x, B.y, B.z}, e2 = {C.x, C.y, C.z};
    return normalize(cross(e1, e2));
}

// Per-lane stack: entries [max(0, sp-K), sp) live in LDS slots (i mod K),
// older entries in the lane's spill area (layout [entry][lane]).
struct Stack {
    int2* lds;
    int stride;
    int tid;
    int2* spill;
    int sstride;
    int lane;
    int cap;
    int sp;
    uint32_t spills;

    __device__ __forceinline__ void push(int ref, float t) {
        if (sp >= kStackLds) {
            int j = sp - kStackLds;
            spills++;
            if (j < cap) spill[(size_t)j * sstride + lane] = lds[(j % kStackLds) * stride + tid];
        }
        lds[(sp % kStackLds) * stride + tid] = make_int2(ref, __float_as_int(t));
        sp++;
    }
    __device__ __forceinline__ int2 pop() {
        sp--;
        int2 e = lds[(sp % kStackLds) * stride + tid];
        if (sp >= kStackLds) {
            int j = sp - kStackLds;
            if (j < cap) lds[(j % kStackLds) * stride + tid] = spill[(size_t)j * sstride + lane];
        }
        return e;
    }
};

struct Geo {
    const GpuNode* nodes;
    const GpuTri* tris;
    int root_ref;
    float rlo[3], rhi[3];
};

struct Hit {
    float dist;
    int tri;
    bool found;
};

// ComputeBVHIntersection (limited = true, :354-466) or
// ComputeBVHIntersectionWithoutStep (:468-584), as a resumable walk:
// walk_begin tests the start node's box, walk_visit processes one node or
// leaf and finds the next one.  Both return false when the walk is over; h
// then holds the result.  Run to completion, they are the reference's loop.
__device__ __forceinline__ bool walk_begin(const Geo& g, const Ray& r, float step, bool limited, Stack& st, Hit& h,
                                           int& cur, int& best) {
    h.dist = kInf;
    h.tri = -1;
    h.found = false;
    best = -1;
    float tn;
    if (!box_hit(r, g.rlo[0], g.rlo[1], g.rlo[2], g.rhi[0], g.rhi[1], g.rhi[2], tn)) return false;
    if (limited && step < tn) return false;
    st.sp = 0;
    cur = g.root_ref;
    return true;
}

// A visit is one node (both child boxes) or one triangle of a leaf: inside a
// leaf, cur = ~(next triangle) and best is the leaf's closest so far, so a
// lane testing a triangle and a lane testing a node cost about the same.
__device__ __forceinline__ bool walk_visit(const GpuNode* nodes, const GpuTri* tris, const Ray& r, float step,
                                           bool limited, Stack& st, Hit& h, int& cur, int& best) {
    bool use_stack;
    float tn1, tn2;
    if (cur < 0) {
        int i = ~cur, last;
        float t = tri_distance(tris, i, r, last);
        if (t >= 0.0f && t < h.dist) {
            h.dist = t;
            best = i;
        }
        if (!last) {
            cur = ~(i + 1);
            return true;
        }
        if (best >= 0 && (!limited || h.dist <= step)) {
            h.found = true;
            h.tri = best;
        }
        best = -1;
        use_stack = true;
    } else {
        const float4* N = reinterpret_cast<const float4*>(nodes) + 4 * cur;
        float4 a = N[0], b = N[1], c = N[2];
        int4 d = reinterpret_cast<const int4*>(N)[3];
        bool f1 = box_hit(r, a.x, a.y, a.z, a.w, b.x, b.y, tn1);
        bool f2 = box_hit(r, b.z, b.w, c.x, c.y, c.z, c.w, tn2);
        int left = d.x, right = d.y;
        use_stack = false;
        if (limited) {
            if (f1 && f2) {
                if (tn1 <= tn2) {
                    if (tn1 <= step) {
                        cur = left;
                        if (tn2 <= step && right != kDeadRef) st.push(right, tn2);
                    } else {
                        use_stack = true;
                    }
                } else {
                    if (tn2 <= step) {
                        cur = right;
                        if (tn1 <= step && left != kDeadRef) st.push(left, tn1);
                    } else {
                        use_stack = true;
                    }
                }
            } else if (f1) {
                if (tn1 <= step) cur = left;
                else use_stack = true;
            } else {
                if (f2 && tn2 <= step) cur = right;
                else use_stack = true;
            }
            // a child with nothing in reach (entry trees only): go on as its walk would end
            if (cur == kDeadRef) use_stack = true;
        } else {
            if (f1 && f2) {
                if (tn1 <= tn2) {
                    cur = left;
                    st.push(right, tn2);
                } else {
                    cur = right;
                    st.push(left, tn1);
                }
            } else if (f1) {
                cur = left;
            } else if (f2) {
                cur = right;
            } else {
                use_stack = true;
            }
        }
    }
    if (use_stack) {
        if (st.sp == 0) return false;
        int2 e = st.pop();
        float tn = __int_as_float(e.y);
        cur = e.x;
        if (h.dist < tn) {
            while (st.sp > 0 && h.dist < tn) {
                e = st.pop();
                tn = __int_as_float(e.y);
            }
            if (h.dist < tn || (limited && step < tn)) return false;
            cur = e.x;
        }
    }
    return true;
}

// The whole walk in one loop: the same visits as walk_begin/walk_visit, with
// no state kept between them (ray queries, and transport scenes whose walks
// are a handful of visits).
__device__ __forceinline__ Hit walk(const Geo& g, const Ray& r, float step, bool limited, Stack& st) {
    Hit h;
    h.dist = kInf;
    h.tri = -1;
    h.found = false;
    float tn1, tn2;
    if (!box_hit(r, g.rlo[0], g.rlo[1], g.rlo[2], g.rhi[0], g.rhi[1], g.rhi[2], tn1)) return h;
    if (limited && step < tn1) return h;
    st.sp = 0;
    int cur = g.root_ref;
    while (true) {
        bool use_stack;
        if (cur < 0) {
            int i = ~cur, best = -1, last;
            do {
                float t = tri_distance(g.tris, i, r, last);
                if (t >= 0.0f && t < h.dist) {
                    h.dist = t;
                    best = i;
                }
                i++;
            } while (!last);
            if (best >= 0 && (!limited || h.dist <= step)) {
                h.found = true;
                h.tri = best;
            }
            use_stack = true;
        } else {
            const float4* N = reinterpret_cast<const float4*>(g.nodes) + 4 * cur;
            float4 a = N[0], b = N[1], c = N[2];
            int4 d = reinterpret_cast<const int4*>(N)[3];
            bool f1 = box_hit(r, a.x, a.y, a.z, a.w, b.x, b.y, tn1);
            bool f2 = box_hit(r, b.z, b.w, c.x, c.y, c.z, c.w, tn2);
            int left = d.x, right = d.y;
            use_stack = false;
            if (limited) {
                if (f1 && f2) {
                    if (tn1 <= tn2) {
                        if (tn1 <= step) {
                            cur = left;
                            if (tn2 <= step) st.push(right, tn2);
                        } else {
                            use_stack = true;
                        }
                    } else {
                        if (tn2 <= step) {
                            cur = right;
                            if (tn1 <= step) st.push(left, tn1);
                        } else {
                            use_stack = true;
                        }
                    }
                } else if (f1) {
                    if (tn1 <= step) cur = left;
                    else use_stack = true;
                } else {
                    if (f2 && tn2 <= step) cur = right;
                    else use_stack = true;
                }
            } else {
                if (f1 && f2) {
                    if (tn1 <= tn2) {
                        cur = left;
                        st.push(right, tn2);
                    } else {
                        cur = right;
                        st.push(left, tn1);
                    }
                } else if (f1) {
                    cur = left;
                } else if (f2) {
                    cur = right;
                } else {
                    use_stack = true;
                }
            }
        }
        if (use_stack) {
            if (st.sp == 0) break;
            int2 e = st.pop();
            float tn = __int_as_float(e.y);
            cur = e.x;
            if (h.dist < tn) {
                while (st.sp > 0 && h.dist < tn) {
                    e = st.pop();
                    tn = __int_as_float(e.y);
                }
                if (h.dist < tn || (limited && step < tn)) break;
                cur = e.x;
            }
        }
    }
    return h;
}

// Float rule of the planar scenes (Planar in xmc_params.h): the nearest
// surface is found analytically.  For |d.z| > EPSILON, surface s at distance
// t = (z_s - p.z) * (1 / d.z) is hit when t >= 0 and (p.x + t d.x,
// p.y + t d.y) lies in its rectangle (edges included); the smallest such t
// wins (the first surface on ties), and a step-limited search reports it
// when t <= step.  Parallel to the planes nothing is hit.  The hit reports
// the surface's first triangle (its normal is (0, 0, +-1), whose sign the
// boundary interaction does not use).
__device__ __forceinline__ Hit planar_hit(const Planar& S, V3 p, V3 d, float step) {
    Hit h;
    h.found = false;
    h.dist = kInf;
    h.tri = -1;
    if (fabsf(d.z) > kEpsF) {
        const float iz = 1.0f / d.z;
        for (int s = 0; s < S.n; ++s) {
            const float t = (S.z[s] - p.z) * iz;
            if (t >= 0.0f && t < h.dist) {
                const float x = p.x + t * d.x, y = p.y + t * d.y;
                const float4 r = S.rect[s];
                if (x >= r.x && x <= r.y && y >= r.z && y <= r.w) {
                    h.dist = t;
                    h.tri = S.tri[s];
                }
            }
        }
    }
    h.found = h.tri >= 0 && h.dist <= step;
    return h;
}

// Float rule of the scenes whose every surface is a rectangle or a lattice
// heightfield (surface_rule = 2): the rectangles' nearest hit as planar_hit
// finds it (for the steps whose z-range reaches one of their planes; for the
// others there is none within the step, and the found hit is the same either
// way); then every triangle of the heightfields whose cell lies in the
// step's xy bounding box widened by L.mxy and whose cell z-range (per
// heightfield) meets the step's z-range widened by L.mz, tested as a BVH leaf
// tests it: a triangle nearer than the hit, or as near with a lower leaf
// index, replaces it (so the order of the tests does not matter).  A step
// whose box spans more than kLatticeMaxBox cells along x or y, or an
// unlimited one, takes the BVH walk over every surface instead.
__device__ __forceinline__ void lattice_take(const GpuTri* tris, int k, const Ray& ray, Hit& h) {
    int last;
    const float t = tri_distance(tris, k, ray, last);
    if (t >= 0.0f && (t < h.dist || (t == h.dist && k < h.tri))) {
        h.dist = t;
        h.tri = k;
    }
}

__device__ __forceinline__ Hit rule_hit(const Params& C, const Geo& geo, V3 p, V3 d, float step, Stack& st) {
    const Lattice& L = C.lattice;
    if (L.ns == 0) return planar_hit(C.planar, p, d, step);
    if (!(step < kInf)) return walk(geo, make_ray(p, d), step, true, st);
    const float ez = __builtin_fmaf(step, d.z, p.z);
    const float az = fminf(p.z, ez) - L.mz, bz = fmaxf(p.z, ez) + L.mz;
    // a rectangle whose plane lies outside the step's z-range (widened further by 1e-5 of
    // its |z|) cannot be hit within the step: planar_hit runs only for steps that reach one
    bool reach = false;
    for (int s = 0; s < C.planar.n; ++s) {
        const float z = C.planar.z[s], m = 1.0e-5f * fabsf(z);
        reach = reach || (z >= az - m && z <= bz + m);
    }
    Hit h;
    if (reach) {
        h = planar_hit(C.planar, p, d, step);
    } else {
        h.found = false;
        h.dist = kInf;
        h.tri = -1;
    }
    if (bz < L.zlo || az > L.zhi) return h;
    const float ex = __builtin_fmaf(step, d.x, p.x), ey = __builtin_fmaf(step, d.y, p.y);
    const float fx = (float)L.nx, fy = (float)L.ny;
    const int i0 = max(0, (int)floorf(fmaxf((fminf(p.x, ex) - L.mxy - L.x0) * L.inv_dx, -1.0f)));
    const int i1 = min(L.nx - 1, (int)floorf(fminf((fmaxf(p.x, ex) + L.mxy - L.x0) * L.inv_dx, fx)));
    const int j0 = max(0, (int)floorf(fmaxf((fminf(p.y, ey) - L.mxy - L.y0) * L.inv_dy, -1.0f)));
    const int j1 = min(L.ny - 1, (int)floorf(fminf((fmaxf(p.y, ey) + L.mxy - L.y0) * L.inv_dy, fy)));
    if (i0 > i1 || j0 > j1) return h;
    if (i1 - i0 >= kLatticeMaxBox || j1 - j0 >= kLatticeMaxBox) return walk(geo, make_ray(p, d), step, true, st);
    Ray ray;  // the triangle test reads the origin and direction only
    ray.o = p;
    ray.d = d;
    // the scan marks the (cell, heightfield) slots whose z-ranges meet the step's in a
    // 64-bit mask, bit ((j - j0) * 4 + (i - i0)) * 4 + s; the tests then run over the set
    // bits, so that a wave runs as many test rounds as its lane with the most candidates
    uint64_t mask = 0;
    for (int j = j0; j <= j1; ++j) {
        for (int i = i0; i <= i1; ++i) {
            const float4* R = L.cells + (size_t)(j * L.nx + i) * L.ns;
            const int b0 = ((j - j0) * 4 + (i - i0)) * 4;
#pragma unroll
            for (int s = 0; s < kLatticeMaxSurfaces; ++s) {
                if (s < L.ns) {
                    const float2 z = *reinterpret_cast<const float2*>(R + s);
                    if (z.x <= bz && z.y >= az) mask |= 1ull << (b0 + s);
                }
            }
        }
    }
    while (mask != 0) {
        const int b = __builtin_ctzll(mask);
        mask &= mask - 1;
        const int s = b & 3, i = i0 + ((b >> 2) & 3), j = j0 + (b >> 4);
        const float4 c = L.cells[(size_t)(j * L.nx + i) * L.ns + s];
        const int ta = __float_as_int(c.z), tb = __float_as_int(c.w);
        lattice_take(geo.tris, ta, ray, h);
        if (tb >= 0) lattice_take(geo.tris, tb, ray, h);
    }
    h.found = h.tri >= 0 && h.dist <= step;
    return h;
}

// --------------------------------------------------------------------------
// photon physics (mcml_kernel.cpp), float
// --------------------------------------------------------------------------
__device__ __forceinline__ float transmit_cos(float n1, float n2, float ci) {  // :497-529
    if (n1 == n2) return ci;
    if (ci > kCosZeroF) return ci;
    if (ci < kCos90F) return 0.0f;
    float si = __builtin_sqrtf(one_minus_sq(ci));
    float st = n1 * si / n2;
    if (st >= 1.0f) return 0.0f;
    return __builtin_sqrtf(one_minus_sq(st));
}

__device__ __forceinline__ float fresnel(float n1, float n2, float ci, float ct) {  // :531-564
    if (n1 == n2) return 0.0f;
    if (ci > kCosZeroF) {
        float r = (n2 - n1) / (n2 + n1);
        return r * r;
    }
    if (ci < kCos90F) return 1.0f;
    float si = __builtin_sqrtf(one_minus_sq(ci));
    float st = __builtin_sqrtf(one_minus_sq(ct));
    float cs = ci * ct - si * st;
    float cd = ci * ct + si * st;
    float ss = si * ct + ci * st;
    float sd = si * ct - ci * st;
    return 0.5f * sd * sd * (cd * cd + cs * cs) / (ss * ss * cd * cd);
}

__device__ __forceinline__ V3 reflect(V3 i, V3 n) {  // :582-593
    V3 sn = normalize(n);
    V3 si = normalize(i);
    float tmp = 2.0f * dot(si, sn);
    sn.x *= tmp;
    sn.y *= tmp;
    sn.z *= tmp;
    return normalize(sub(si, sn));
}

__device__ __forceinline__ V3 refract(float n1, float n2, V3 i, V3 n) {  // :595-616
    V3 si = normalize(i);
    si.x *= n1;
    si.y *= n1;
    si.z *= n1;
    V3 sn = normalize(n);
    float t1 = dot(si, sn);
    float t2 = __builtin_sqrtf((n2 * n2 - n1 * n1) / (t1 * t1) + 1.0f);
    t2 = (t2 - 1.0f) * t1;
    V3 t = {si.x + t2 * sn.x, si.y + t2 * sn.y, si.z + t2 * sn.z};
    return normalize(t);
}

// GetNewPhotonDirection + ComputeCosineTheta (:442-495).  The
// Henyey-Greenstein inversion is rearranged so that 1 - cos(theta) and
// 1 + cos(theta) come out without cancellation (D = 1 - g + 2gu,
// t = (1 - g^2)/D: 1 - cos = (1-g)(1-u)(t+1-g)/D, 1 + cos = (1+g)u(1+g+t)/D;
// 2(1-u) and 2u at g = 0): with cos computed first, float keeps only a few
// digits of sin(theta) for the strongly forward scatters of g ~ 0.9.
__device__ __forceinline__ V3 spin(V3 d, float g, uint64_t& x, uint32_t& draws) {
    float u = next_u(x);
    float a = 1.0f - g, b = 1.0f + g;
    float inv = 1.0f / (a + 2.0f * g * u);
    float t = (1.0f - g * g) * inv;
    float vers = a * (1.0f - u) * (t + a) * inv;
    float cove = b * u * (b + t) * inv;
    float c = vers < 1.0f ? 1.0f - vers : cove - 1.0f;
    float st = __builtin_sqrtf(vers * cove);
    float sp, cp;
    xsincos2pi(next_u(x), sp, cp);
    draws += 2;
    float ux = d.x, uy = d.y, uz = d.z;
    V3 r;
    // sin(theta_u) from the components (frame stays orthonormal near the
    // pole); rotate when ux^2 + uy^2 > 2e-12, the reference's |uz| <= 1 - 1e-12
    float s2 = ux * ux + uy * uy;
    if (s2 > 2e-12f) {
        float tmp = __builtin_sqrtf(s2);
        float itmp = 1.0f / tmp;
        r.x = st * (ux * uz * cp - uy * sp) * itmp + ux * c;
        r.y = st * (uy * uz * cp + ux * sp) * itmp + uy * c;
        r.z = -st * cp * tmp + uz * c;
    } else {
        float tmp = (uz >= 0.0f) ? 1.0f : -1.0f;
        r.x = st * cp;
        r.y = st * sp;
        r.z = tmp * c;
    }
    // one first-order renormalisation step, d *= (3 - |d|^2)/2: the rotation
    // does not preserve |d| = 1 and float rounding would drift it
    float f = __builtin_fmaf(-0.5f, dot(r, r), 1.5f);
    r.x = r.x * f;
    r.y = r.y * f;
    r.z = r.z * f;
    return r;
}

// GetLocalAttractiveVector (:325-362) in float: the unit vector along
// t2 / |t2|^3 - t1 / |t1|^3 of the photon at p in layer l (constants in Bias).
__device__ __forceinline__ V3 attractive(const Bias& B, V3 start, V3 p, int l) {
    float rx = p.x - start.x, ry = p.y - start.y, rz = p.z - start.z;
    V3 t1 = {rx - B.rd[0], ry - B.rd[1], rz - B.r1z[l]};
    float l1 = __builtin_sqrtf(dot(t1, t1));
    float c1 = l1 * l1 * l1;
    t1 = {t1.x / c1, t1.y / c1, t1.z / c1};
    V3 t2 = {rx - B.rd[0], ry - B.rd[1], rz - B.r2z[l]};
    float l2 = __builtin_sqrtf(dot(t2, t2));
    float c2 = l2 * l2 * l2;
    t2 = {t2.x / c2, t2.y / c2, t2.z / c2};
    return normalize(sub(t2, t1));
}

// GetAreaIndex (:190-227) in float, with partitions per unit length (scale)
// in place of the division; -1 outside.  Partitions are <= 255.
__device__ __forceinline__ int area_index(const Area& A, V3 p, int& bx, int& by, int& bz) {
    float ix = (p.x - A.corner[0]) * A.scale[0];
    float iy = (p.y - A.corner[1]) * A.scale[1];
    float iz = (p.z - A.corner[2]) * A.scale[2];
    bool inside = !((ix < 0.0f || ix >= A.partf[0]) || (iy < 0.0f || iy >= A.partf[1]) ||
                    (iz < 0.0f || iz >= A.partf[2]));
    if (!inside) return -1;
    bx = (int)ix;
    by = (int)iy;
    bz = (int)iz;
    return bx * A.part[1] * A.part[2] + by * A.part[2] + bz;
}

__device__ __forceinline__ int area_index(const Area& A, V3 p) {
    int bx, by, bz;
    return area_index(A, p, bx, by, bz);
}

// GetPlaneSegmentIntersectionPoint (mcml_math.cpp:59-94) for the plane
// z = const: the two dot products with e_z are exactly z - a.z and b.z - a.z.
__device__ __forceinline__ V3 plane_point(V3 a, V3 b, float z) {
    const float kFar = 1.0e30f;
    float d = z - a.z;
    V3 w = sub(b, a);
    float e = w.z;
    V3 r = {kFar, kFar, kFar};
    if (e != 0.0f) {
        float t = d / e;
        r = {a.x + w.x * t, a.y + w.y * t, a.z + w.z * t};
        if (dot(sub(a, r), sub(b, r)) > 0.0f) r = {kFar, kFar, kFar};
    }
    return r;
}

// UpdatePhotonTrajectory (:750-792): every voxel-centre plane z_k = corner +
// (k + 1/2) step crossed by the segment p0-p1 adds the weight to the
// scattering grid and the voxel to the photon's trajectory (kept up to
// `capacity` voxels; n counts them all).  Stops at the first crossing outside
// the Area, as the reference does.
__device__ void path_crossings(const Params& C, V3 p0, V3 p1, float w, uint32_t* vox, int& n) {
    V3 a, b;
    if (p1.z > p0.z) {
        a = p0;
        b = p1;
    } else if (p1.z < p0.z) {
        a = p1;
        b = p0;
    } else {
        return;
    }
    const Area& A = C.area;
    const float step = A.length[2] / A.partf[2];
    int plane_index = (int)((a.z - A.corner[2] + step / 2.0f) / step);
    float plane = A.corner[2] + plane_index * step + step / 2.0f;
    const float top = A.corner[2] + A.length[2];
    while ((plane <= b.z) && (plane < top)) {
        int idx = area_index(A, plane_point(a, b, plane));
        if (idx < 0) break;
        if (C.paths.scattering) unsafeAtomicAdd(C.paths.scattering + idx, (double)w);
        if (vox) {
            if (n < C.paths.capacity) vox[n] = (uint32_t)idx;
            n++;
        }
        plane += step;
    }
}

// GetDetectorId (:663-742): first matching cube, then ring.  The direction
// test's cosine -d.z / |d|^2 does not depend on the detector and a ring's
// distance only on its centre, so each is computed once (concentric rings
// share one distance); the values and the first match are the same.
__device__ int detector_id(const Detectors& D, V3 p, V3 d, uint32_t visited) {
    float len = d.x * d.x + d.y * d.y + d.z * d.z;
    float ac = (-d.z) / len;
    for (int i = 0; i < D.n_cube; ++i) {
        const float* c = D.cube + 7 * i;
        bool in = (p.x >= c[0]) && (p.x < c[3]) && (p.y >= c[1]) && (p.y < c[4]) && (p.z >= c[2]) && (p.z < c[5]);
        if (in && (visited & D.filter[i]) && ac > c[6]) return i;
    }
    float cx = __builtin_nanf(""), cy = 0.0f, dist = 0.0f;
    for (int i = 0; i < D.n_ring; ++i) {
        const float* r = D.ring + 5 * i;
        if (!(r[0] == cx && r[1] == cy)) {
            cx = r[0];
            cy = r[1];
            dist = __builtin_sqrtf(((p.x - cx) * (p.x - cx)) + ((p.y - cy) * (p.y - cy)));
        }
        bool in = (dist >= r[2]) && (dist < r[3]);
        if (in && (visited & D.filter[D.n_cube + i]) && ac > r[4]) return D.n_cube + i;
    }
    return -1;
}

// UpdateDetectorTimeScale (:812-849): first bin i < T-1 with
// start_i <= t < finish_i (candidates around (t - t0)/dt: a bin overlaps at
// most its neighbours, so the first match is among them), else the open
// last bin when t >= start_{T-1}.
__device__ __forceinline__ int time_bin(const Detectors& D, double t) {
    int T = D.T;
    if (T <= 0) return -1;
    if (T > 1) {
        if (D.dt > 0.0) {
            // the guess only picks the three candidates; the bin is decided by the comparisons,
            // so a reciprocal (within an ulp of the quotient) picks the same bin
            double xb = (t - D.t0) * D.inv_dt;
            int k = xb < 0.0 ? -1 : (xb >= (double)(T - 1) ? T - 1 : (int)xb);
            int lo = k - 1 < 0 ? 0 : k - 1;
            int hi = k + 1 > T - 2 ? T - 2 : k + 1;
            for (int i = lo; i <= hi; ++i)
                if (t >= D.t_start[i] && t < D.t_finish[i]) return i;
        } else {
            for (int i = 0; i < T - 1; ++i)
                if (t >= D.t_start[i] && t < D.t_finish[i]) return i;
        }
    }
    if (t >= D.t_start[T - 1]) return T - 1;
    return -1;
}

// Where a step-limited walk from p along d over `step` starts: -1 nothing
// can be hit (skip the walk), -2 at the root, e >= 0 at entry e.  The cell
// bounds the distance from its centre to every triangle, so p is at least
// that bound less |p - centre| away (whichever cell float rounding files p
// under).  Skipping needs that to exceed the step by a margin that covers
// float rounding of the lookup and of the triangle test, or a step that moves
// at most the grid's margin along every axis from a cell with no triangle
// within reach; a short step from a cell with an entry starts there.  Steps
// that move at most margin2 use the cell's second-level entry likewise.
__device__ __forceinline__ int walk_start(const DistGrid& G, V3 p, V3 d, float step) {
    float fx = (p.x - G.lo[0]) * G.inv_h;
    float fy = (p.y - G.lo[1]) * G.inv_h;
    float fz = (p.z - G.lo[2]) * G.inv_h;
    if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f)) return -2;
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    if (ix >= G.dims[0] || iy >= G.dims[1] || iz >= G.dims[2]) return -2;
    uint32_t word = G.cells[((size_t)iz * G.dims[1] + iy) * G.dims[0] + ix];
    float ox = (fx - ((float)ix + 0.5f)) * G.h;
    float oy = (fy - ((float)iy + 0.5f)) * G.h;
    float oz = (fz - ((float)iz + 0.5f)) * G.h;
    // |p - centre| only enters the bound, whose margin is 1e-4 mm and more: the hardware
    // square root (within an ulp) serves, in place of the correctly rounded sequence
    float off = __builtin_amdgcn_sqrtf(ox * ox + oy * oy + oz * oz);
    float L = (float)(word & 0xFFu) * G.step_q;
    float room = L - off - (1.0e-4f + 16.0f * offset_scale(p));
    if (room > 0.0f && step * step * dot(d, d) < room * room) return -1;
    float reach = step * fmaxf(fabsf(d.x), fmaxf(fabsf(d.y), fabsf(d.z)));
    uint32_t e = word >> 8;
    if (!(reach <= G.margin)) {
        if (G.cells2 == nullptr || !(reach <= G.margin2)) return -2;
        e = G.cells2[((size_t)iz * G.dims[1] + iy) * G.dims[0] + ix];
    }
    if (e == kEntryRoot) return -2;
    return (int)e - 1;
}

// The separating slab of p's cell (see SlabGrid in xmc_distance.h): true
// when the step from p along d over `step` stays within the cell's ball of
// radius R about its centre c and both of its ends lie in one free interval
// of n.(x - c), each by the margin of walk_start's distance bound -- the step
// then crosses no triangle and its walk would find nothing.
__device__ __forceinline__ int sbyte(uint32_t w, int k) { return (int)(w << (24 - 8 * k)) >> 24; }

__device__ __forceinline__ bool slab_skips(const DistGrid& G, V3 p, V3 d, float step) {
    float fx = (p.x - G.lo[0]) * G.inv_h;
    float fy = (p.y - G.lo[1]) * G.inv_h;
    float fz = (p.z - G.lo[2]) * G.inv_h;
    if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f)) return false;
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    if (ix >= G.dims[0] || iy >= G.dims[1] || iz >= G.dims[2]) return false;
    const uint64_t w64 = G.slabs[((size_t)iz * G.dims[1] + iy) * G.dims[0] + ix];
    const uint2 w = make_uint2((uint32_t)w64, (uint32_t)(w64 >> 32));
    float ox = (fx - ((float)ix + 0.5f)) * G.h;
    float oy = (fy - ((float)iy + 0.5f)) * G.h;
    float oz = (fz - ((float)iz + 0.5f)) * G.h;
    float off = __builtin_amdgcn_sqrtf(ox * ox + oy * oy + oz * oz);  // within an ulp, as in walk_start
    float m = 1.0e-4f + 16.0f * offset_scale(p);
    float R = (float)((w.y >> 16) & 0xFFu) * G.slab_unit;
    float room = R - m - off;
    bool in_ball = room > 0.0f && step * step * dot(d, d) < room * room;
    // octahedral direction (slab_direction in xmc_distance.h)
    float u = (float)sbyte(w.x, 0) * (1.0f / 127.0f), v = (float)sbyte(w.x, 1) * (1.0f / 127.0f);
    float nz = 1.0f - fabsf(u) - fabsf(v);
    if (nz < 0.0f) {
        float fu = (1.0f - fabsf(v)) * (u >= 0.0f ? 1.0f : -1.0f);
        float fv = (1.0f - fabsf(u)) * (v >= 0.0f ? 1.0f : -1.0f);
        u = fu;
        v = fv;
    }
    float t0 = u * ox + v * oy + nz * oz;
    float t1 = t0 + step * (u * d.x + v * d.y + nz * d.z);
    float tmin = fminf(t0, t1), tmax = fmaxf(t0, t1);
    float qf = R * (1.0f / 127.0f);
    bool free = tmax < (float)sbyte(w.x, 2) * qf - m ||
                (tmin > (float)sbyte(w.x, 3) * qf + m && tmax < (float)sbyte(w.y, 0) * qf - m) ||
                tmin > (float)sbyte(w.y, 1) * qf + m;
    return in_ball && free;
}

template <bool TRACE>
struct Recorder {
    xmc_trace_event* ev;
    int n;
    int cap;
    __device__ __forceinline__ void put(int type, int layer, int aux, uint32_t draws, V3 p, V3 d, float w,
                                        float t) {
        if (TRACE) {
            if (ev != nullptr && n < cap) {
                xmc_trace_event& e = ev[n];
                e.type = type;
                e.layer = layer;
                e.aux = aux;
                e.draws = draws;
                e.pos[0] = p.x; e.pos[1] = p.y; e.pos[2] = p.z;
                e.dir[0] = d.x; e.dir[1] = d.y; e.dir[2] = d.z;
                e.weight = w;
                e.time = t;
            }
            n++;
        }
    }
};

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int lane) {
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, lane);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

// Weight ledger slots (per block in LDS, flushed to stats words 8..15).
enum { kAbs = 0, kEsc, kDet, kArea, kKill, kRLost, kRGain, kOther, kLedger };

// The photon loop.  Run-wide fields (photon range, queue, streams, stack
// spill, trace, counters) come from P; per-configuration fields (geometry,
// optics, detectors, tallies, source, termination) from CFG(), which is P
// itself for a single scene and the lane's entry of `cfgs` in a batch.
// Lane phases of the persistent loop.  Each round a lane either starts an
// interaction (step and walk start), advances its BVH walk by a few visits,
// or resolves the interaction (boundary or scatter, termination) -- so a
// long walk in one lane no longer holds the wave's other lanes.  Scenes whose
// walks are a handful of visits (no walk-skipping grid) run the walks to the
// end within the round instead (INTERLEAVE = false): fewer registers live
// across rounds, more waves per SIMD.
constexpr int kBegin = 0, kWalk = 1, kResolve = 2, kStart = 3;
constexpr int kVisitsPerRound = XMC_VISITS_PER_ROUND;
// Boundary interactions (Fresnel, refraction, detectors: the longest path of
// a round) wait, up to kParkWait rounds, until kParkMin lanes of the wave
// have one: a round pays for every path any lane takes, and about half of
// all rounds would otherwise carry one.  Each photon's own operations are
// unchanged.  Measured at four waves per SIMD (heightfield / slab photons/s):
// none 2.73e7 / 4.46e7; 2 lanes or 1 round 2.78e7 / 4.45e7; 6 or 5 2.90e7 /
// 4.58e7; 12 or 8 2.92e7 / 4.68e7; 16 or 16 2.91e7 / 4.79e7.  At five waves
// with one visit per round, 16 or 16 is 0.7 % ahead of 12 or 8.
constexpr bool kParkHits = XMC_PARK_HITS != 0;
constexpr int kParkMin = XMC_PARK_MIN;    // lanes that make a boundary round worth running
constexpr uint32_t kParkWait = XMC_PARK_WAIT;  // rounds a boundary interaction may wait
// Starting a walk (reciprocal direction, entry lookup, separating-slab and first box
// tests) is parked the same way.  Heightfield photons/s with the slabs (two launches of 1e8
// each): none 3.76e7; 2 lanes or 2 rounds 3.86e7; 4 or 4 4.01e7; 6 or 6 4.08e7; 8 or 8
// 4.09e7; 10 or 10 4.08e7; 12 or 12 4.03e7; 16 or 16 3.92e7; 24 or 24 3.67e7.
constexpr int kStartMin = XMC_START_MIN;
constexpr uint32_t kStartWait = XMC_START_WAIT;

template <bool TRACE, bool BATCHED, bool PATHS, bool INTERLEAVE, bool MERGE = false>
__device__ __forceinline__ void transport_body(const Params& P, const Params* __restrict__ cfgs,
                                               const uint64_t* __restrict__ prefix, int ncfg) {
    extern __shared__ int2 lds_stack[];
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const int nlanes = gridDim.x * blockDim.x;
    const int glane = (int)(blockIdx.x * blockDim.x + tid);
    // the stacks' rings: in LDS, or (a plain single scene under the surface rule, which
    // walks only its rare long steps) in global memory, leaving the LDS to the tallies
    bool ring_global = false;
    if constexpr (!INTERLEAVE && !BATCHED) ring_global = P.ring != nullptr;
    Stack st{ring_global ? P.ring : lds_stack, ring_global ? nlanes : (int)blockDim.x, ring_global ? glane : tid,
             P.spill, nlanes, glane, P.spill_depth, 0, 0};
    double* ledger = reinterpret_cast<double*>(lds_stack + (ring_global ? 0 : kStackLds * (int)blockDim.x));
    if (tid < kLedger) ledger[tid] = 0.0;
    // the block's detector sums (weight, target weight, count per detector) follow the
    // ledger in LDS, then its absorption tile
    const int ndl = BATCHED ? 0 : P.det_lds;
    double* dsum = ledger + kLedger;
    unsigned long long* dcount = reinterpret_cast<unsigned long long*>(dsum + 2 * ndl);
    for (int i = tid; i < 3 * ndl; i += blockDim.x) dsum[i] = 0.0;
    double* tile = dsum + 3 * ndl;
    const AbsTile& T = P.abs_tile;
    constexpr bool kTile = !BATCHED && !INTERLEAVE;  // the slab-like scenes where deposits crowd
    if constexpr (kTile)
        for (int i = tid; i < T.count; i += blockDim.x) tile[i] = 0.0;
    __syncthreads();
    double w_abs = 0.0;
    // MERGE (a single scene with an absorption grid): a lane holds its last voxel's deposits
    // and adds them to the tile or the grid when it deposits into another voxel (and at the
    // end), which saves the atomics -- contended ones around the beam -- of consecutive
    // deposits into one voxel
    [[maybe_unused]] int pend_idx = -1, pend_tile = -1;
    [[maybe_unused]] double pend_w = 0.0;
    auto flush_pending = [&]() {
        if (pend_idx >= 0) {
            if (pend_tile >= 0) atomicAdd(tile + pend_tile, pend_w);
            else unsafeAtomicAdd(P.tally.absorption + pend_idx, pend_w);
        }
    };
    // never take P's address: that would move the kernel-argument struct
    // into private memory
    const Params* cfg = cfgs;
    auto CFG = [&]() -> const Params& {
        if constexpr (BATCHED) return *cfg;
        else return P;
    };

    V3 pos = {0, 0, 0}, dir = {0, 0, 1};
    float w = 0.0f, tim = 0.0f, trange = 0.0f, orange = 0.0f;
    int layer = 1;
    uint32_t visited = 0;
    uint64_t x = 0;
    bool alive = false, done = false;
    float Ln = 1.0f, Linv = 0.0f, Labs = 0.0f, Lg = 0.0f;  // n, 1/(mua+mus), mua/(mua+mus), g
    bool Lglass = false, Ltarget = false;
    uint32_t draws = 0;
    Recorder<TRACE> rec{nullptr, 0, 0};
    // per-lane event counters, 32-bit (a lane sees far fewer than 2^32 events per launch,
    // and flush() empties them long before): two registers fewer each than 64-bit ones
    uint32_t n_scatter = 0, n_boundary = 0, n_roulette = 0, n_photons = 0, n_culled = 0;
    uint32_t n_overflow = 0, n_entry = 0;
    uint32_t* vox = nullptr;  // this lane's trajectory (voxel indices)
    int nvox = 0;
    // the interaction in progress: its step, walk and hit
    int phase = kBegin;
    float step = kInf;
    Hit h;
    h.found = false;
    h.dist = kInf;
    h.tri = -1;
    V3 rinv = {0.0f, 0.0f, 0.0f};  // the walking ray is (pos, dir) with these inverses and axis flags
    uint32_t raxes = 0;
    int cur = 0, best = -1;
    uint32_t hit_wait = 0;    // (wave-uniform) rounds the wave's boundary interactions have waited
    uint32_t start_wait = 0;  // (wave-uniform) rounds its walk starts have waited
    if constexpr (PATHS) {
        if (P.paths.det_traj)
            vox = P.paths.lane_voxels + (size_t)(blockIdx.x * blockDim.x + tid) * (size_t)P.paths.capacity;
    }

    auto load_layer = [&](int l) {
        const Layers& L = CFG().layers;
        Ln = L.n[l];
        Linv = L.inv_mut[l];   // 1 / (mua + mus) and mua / (mua + mus), computed once per scene
        Labs = L.abs_frac[l];
        Lg = L.g[l];
        Lglass = (L.glass_mask >> l) & 1u;
        Ltarget = (L.target_mask >> l) & 1u;
    };

    while (true) {
        // ---- refill retired lanes from the photon queue
        bool need = !alive && !done;
        uint64_t need_mask = __ballot(need);
        if (need_mask) {
            int leader = __ffsll((unsigned long long)need_mask) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(P.queue, (unsigned long long)__popcll(need_mask));
            base = readlane64(base, leader);
            if (need) {
                uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need_mask >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)need_mask, 0u));
                uint64_t idx = base + rank;
                if (idx < P.n_photons) {
                    uint64_t pid = P.first_photon + idx;
                    if constexpr (BATCHED) {
                        // configuration c owns photons [prefix[c], prefix[c+1])
                        int lo = 0, hi = ncfg;
                        while (hi - lo > 1) {
                            int mid = (lo + hi) >> 1;
                            if (prefix[mid] <= idx) lo = mid;
                            else hi = mid;
                        }
                        cfg = cfgs + lo;
                    }
                    x = P.seed_value;
                    for (int i = 0; i < kJumpBits; ++i)
                        if ((pid >> i) & 1ull) x = (x * P.jump[i]) & kMcgMask;
                    // LaunchPhoton (:128-149); float start offset along +z
                    const Params& C = CFG();
                    V3 s = {C.start_pos[0], C.start_pos[1], C.start_pos[2]};
                    pos = s;
                    pos.z = s.z + offset_scale(s);
                    dir = {C.start_dir[0], C.start_dir[1], C.start_dir[2]};
                    w = C.w0;
                    tim = 0.0f;
                    trange = 0.0f;
                    orange = 0.0f;
                    layer = 1;
                    visited = 1u << 1;
                    load_layer(1);
                    alive = true;
                    n_photons++;
                    draws = 0;
                    nvox = 0;
                    phase = kBegin;
                    if (TRACE) {
                        uint64_t k = pid - P.trace.first;
                        rec.ev = (k < P.trace.n) ? P.trace.events + k * (uint64_t)P.trace.max_events : nullptr;
                        rec.n = 0;
                        rec.cap = P.trace.max_events;
                        rec.put(XMC_EV_LAUNCH, layer, 0, draws, pos, dir, w, tim);
                        if (k < P.trace.n) P.trace.counts[k] = 0;
                    }
                } else {
                    done = true;
                }
            }
        }
        if (__ballot(alive) == 0) break;
        if (!alive) continue;

        if constexpr (!INTERLEAVE) {
            // every round starts an interaction: nothing of the last one is carried over,
            // which tells the compiler that none of its registers live across rounds
            // (but a parked boundary interaction's hit)
            if (!(kParkHits && phase == kResolve && h.found)) {
                phase = kBegin;
                step = kInf;
                h.found = false;
                h.dist = kInf;
                h.tri = -1;
            }
            rinv = {0.0f, 0.0f, 0.0f};
            raxes = 0;
            cur = 0;
            best = -1;
        }
        if (phase == kBegin) {
            if (draws > kDrawLimit) {  // stream block exhausted
                alive = false;
                atomicAdd(ledger + kOther, (double)w);
                continue;
            }
            // ---- one interaction (mcml_kernel.cpp:47-124): the step, then the walk.  A glass
            // layer's walk is unlimited (ComputeBVHIntersectionWithoutStep, :468-584): the
            // step-limited walk with step = +inf takes every branch the unlimited one takes
            // (each tn <= step holds, each step < tn fails), so one walk serves both.
            step = kInf;
            if (!Lglass) {
                float u;
                do {
                    u = next_u(x);
                    draws++;
                } while (u == 0.0f);
                step = -xlogf(u) * Linv;
            }
            const Params& C = CFG();
            h.found = false;  // until a walk finds something within the step
            h.dist = kInf;
            h.tri = -1;
            // (a single scene's plain variant runs only scenes without a grid)
            int e = -2;
            if constexpr (INTERLEAVE || BATCHED)
                e = (!Lglass && C.dist.cells != nullptr) ? walk_start(C.dist, pos, dir, step) : -2;
            if (e == -1) {
                n_culled++;
                phase = kResolve;
            } else {
                phase = kStart;
                cur = e;  // the entry, until the walk starts
            }
        }
        // ---- start the walks: lanes whose step needs one wait (interleaved variant) until
        // kStartMin lanes of the wave do, or kStartWait rounds pass
        bool go_start = true;
        if constexpr (INTERLEAVE && kStartMin > 1) {
            const uint64_t sm = __ballot(phase == kStart);
            go_start = __popcll(sm) >= kStartMin || start_wait >= kStartWait;
            start_wait = (sm != 0 && !go_start) ? start_wait + 1u : 0u;
        }
        if (phase == kStart && go_start) {
            const Params& C = CFG();
            const int e = cur;
            // the root's box as values.  A single scene's are kernel arguments, uniform over
            // the wave: readfirstlane stops the compiler from turning the choice between them
            // and an entry's box into a load through a chosen pointer (which would copy the
            // kernel arguments into private memory).  A batch lane reads its own configuration.
            auto u = [](float v) {
                if constexpr (BATCHED) return v;
                else return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
            };
            int root = C.root_ref;
            if constexpr (!BATCHED) root = __builtin_amdgcn_readfirstlane(root);
            Geo geo{C.nodes, C.tris, root, {u(C.root_lo[0]), u(C.root_lo[1]), u(C.root_lo[2])},
                    {u(C.root_hi[0]), u(C.root_hi[1]), u(C.root_hi[2])}};
            if constexpr (INTERLEAVE) {
                // start below the root (same leaves, same order, same hit); the entry table
                // always has an entry 0 when this variant runs
                const int ei = e >= 0 ? e : 0;
                const int ref = C.dist.entry_ref[ei];
                const float* b = C.dist.entry_box + 6 * ei;
                const float b0 = b[0], b1 = b[1], b2 = b[2], b3 = b[3], b4 = b[4], b5 = b[5];
                if (e >= 0) {
                    geo.root_ref = ref;
                    geo.rlo[0] = b0;
                    geo.rlo[1] = b1;
                    geo.rlo[2] = b2;
                    geo.rhi[0] = b3;
                    geo.rhi[1] = b4;
                    geo.rhi[2] = b5;
                    n_entry++;
                }
            }
            // a step that cannot cross the surfaces near its cell skips its walk
            bool skip = false;
            if constexpr (INTERLEAVE) {
                if (C.dist.slabs != nullptr) skip = slab_skips(C.dist, pos, dir, step);
            }
            if constexpr (INTERLEAVE) {
                Ray ray = make_ray(pos, dir);
                rinv = ray.inv;
                raxes = (ray.ax ? 1u : 0u) | (ray.ay ? 2u : 0u) | (ray.az ? 4u : 0u);
                if (skip) {
                    n_culled++;
                    phase = kResolve;
                } else {
                    phase = walk_begin(geo, ray, step, true, st, h, cur, best) ? kWalk : kResolve;
                }
            } else {
                // the ray's reciprocal direction serves the box tests of a walk only
                if constexpr (BATCHED) {
                    // batched configurations keep no lattice form (their heightfields walk)
                    h = C.surface_rule != 0 ? planar_hit(C.planar, pos, dir, step)
                                            : walk(geo, make_ray(pos, dir), step, true, st);
                } else {
                    h = C.surface_rule != 0 ? rule_hit(C, geo, pos, dir, step, st)
                                            : walk(geo, make_ray(pos, dir), step, true, st);
                }
                phase = kResolve;
            }
        }
        // ---- up to kVisitsPerRound BVH visits for the lanes still walking; a lane
        // whose walk is long keeps walking next round while the others carry on
        if constexpr (INTERLEAVE) {
            for (int k = 0; k < kVisitsPerRound; ++k) {
                if (phase == kWalk) {
                    const Params& C = CFG();
                    Ray ray{pos, dir, rinv, (raxes & 1u) != 0, (raxes & 2u) != 0, (raxes & 4u) != 0, {}};
                    set_oi(ray);
                    if (!walk_visit(C.nodes, C.tris, ray, step, true, st, h, cur, best)) phase = kResolve;
                }
                if (__ballot(phase == kWalk) == 0) break;
            }
        }
        if constexpr (kParkHits) {
            const uint64_t hm = __ballot(phase == kResolve && h.found);
            const bool do_hits = __popcll(hm) >= kParkMin || hit_wait >= kParkWait;
            hit_wait = (hm != 0 && !do_hits) ? hit_wait + 1u : 0u;
            if (h.found && !do_hits) continue;
        }
        if (phase != kResolve) continue;
        phase = kBegin;
        const Params& C = CFG();
        if (h.found) {
            step = h.dist - 0.01f * kMinDistF;
            tim += step * Ln;
            if (Ltarget) trange += step;
            else orange += step;
            V3 start = pos;
            pos.x += step * dir.x;
            pos.y += step * dir.y;
            pos.z += step * dir.z;
            V3 hit = pos;
            const float4 nv = reinterpret_cast<const float4*>(C.normals)[h.tri];  // tri_normal(), once per scene
            V3 nrm = {nv.x, nv.y, nv.z};
            float ofs = offset_scale(hit);
            if (dot(dir, nrm) < 0.0f) ofs = -ofs;
            pos.x = hit.x - ofs * nrm.x;
            pos.y = hit.y - ofs * nrm.y;
            pos.z = hit.z - ofs * nrm.z;
            // the path ends at the incident-side point short of the surface
            if constexpr (PATHS) path_crossings(C, start, pos, w, vox, nvox);
            if (area_index(C.area, pos) >= 0) {
                // CrossBoundary (:229-284)
                float ci = dot(dir, nrm);
                if (ci < 0.0f) ci = -ci;
                int surface = __float_as_int(reinterpret_cast<const float4*>(C.tris)[3 * h.tri].w);
                int nxt = C.next_layer[layer * C.n_surfaces + surface];
                if (nxt < 0) {
                    alive = false;
                    atomicAdd(ledger + kOther, (double)w);
                } else {
                    float n1 = Ln, n2 = C.layers.n[nxt];
                    float cc = C.layers.crit[layer][nxt];
                    float refl;
                    if (ci <= cc) {
                        refl = 1.0f;
                    } else {
                        float ct = transmit_cos(n1, n2, ci);
                        refl = fresnel(n1, n2, ci, ct);
                    }
                    n_boundary++;
                    float u = next_u(x);
                    draws++;
                    if (u > refl) {
                        if (nxt == 0) {
                            int id = detector_id(C.det, pos, dir, visited);
                            if (id >= 0) {
                                double wd = (double)w;
                                float total = trange + orange;
                                float tw = trange * (w / total);
                                if (ndl) {
                                    atomicAdd(dsum + id, wd);
                                    atomicAdd(dsum + ndl + id, (double)tw);
                                    atomicAdd(dcount + id, 1ull);
                                } else {
                                    unsafeAtomicAdd(C.tally.det_weight + id, wd);
                                    unsafeAtomicAdd(C.tally.det_target + id, (double)tw);
                                    atomicAdd(C.tally.det_count + id, 1ull);
                                }
                                int b = time_bin(C.det, (double)tim);
                                if (b >= 0) {
                                    size_t k = (size_t)id * C.det.T + b;
                                    unsafeAtomicAdd(C.tally.bin_target + k, (double)tw);
                                    atomicAdd(C.tally.bin_count + k, 1ull);
                                    unsafeAtomicAdd(C.tally.bin_weight + k, wd);
                                }
                                atomicAdd(ledger + kDet, wd);
                                if constexpr (PATHS) {
                                    if (vox) {  // UpdateDetectorTrajectory (:794-810)
                                        unsigned long long* dt =
                                            C.paths.det_traj + (size_t)id * (size_t)C.paths.grid_size;
                                        int m = nvox < C.paths.capacity ? nvox : C.paths.capacity;
                                        for (int i = 0; i < m; ++i) atomicAdd(dt + vox[i], 1ull);
                                        if (nvox > C.paths.capacity) n_overflow++;
                                    }
                                }
                            }
                            atomicAdd(ledger + kEsc, (double)w);
                            alive = false;
                            rec.put(XMC_EV_ESCAPE, layer, id, draws, pos, dir, w, tim);
                        } else {
                            dir = refract(n1, n2, dir, nrm);
                            layer = nxt;
                            visited |= 1u << nxt;
                            load_layer(nxt);
                            float o2 = offset_scale(hit);
                            if (dot(dir, nrm) < 0.0f) o2 = -o2;
                            V3 before = pos;
                            pos.x = hit.x + o2 * nrm.x;
                            pos.y = hit.y + o2 * nrm.y;
                            pos.z = hit.z + o2 * nrm.z;
                            if constexpr (PATHS) path_crossings(C, before, pos, w, vox, nvox);
                            rec.put(XMC_EV_TRANSMIT, layer, surface, draws, pos, dir, w, tim);
                        }
                    } else {
                        dir = reflect(dir, nrm);
                        float o2 = offset_scale(hit);
                        if (dot(dir, nrm) < 0.0f) o2 = -o2;
                        V3 before = pos;
                        pos.x = hit.x + o2 * nrm.x;
                        pos.y = hit.y + o2 * nrm.y;
                        pos.z = hit.z + o2 * nrm.z;
                        if constexpr (PATHS) path_crossings(C, before, pos, w, vox, nvox);
                        rec.put(XMC_EV_REFLECT, layer, surface, draws, pos, dir, w, tim);
                    }
                }
            } else {
                alive = false;
                atomicAdd(ledger + kArea, (double)w);
                rec.put(XMC_EV_DIE_AREA, layer, 0, draws, pos, dir, w, tim);
            }
        } else if (Lglass) {
            alive = false;
            atomicAdd(ledger + kOther, (double)w);
            rec.put(XMC_EV_GLASS_MISS, layer, 0, draws, pos, dir, w, tim);
        } else {
            tim += step * Ln;
            if (Ltarget) trange += step;
            else orange += step;
            V3 prev = pos;
            pos.x += step * dir.x;
            pos.y += step * dir.y;
            pos.z += step * dir.z;
            if constexpr (PATHS) path_crossings(C, prev, pos, w, vox, nvox);
            int bx, by, bz;
            int idx = area_index(C.area, pos, bx, by, bz);
            if (idx >= 0) {
                float dw = w * Labs;  // ComputeDroppedWeightOfPhoton (:633-642)
                w -= dw;
                if constexpr (MERGE) {
                    if (idx != pend_idx) {
                        flush_pending();
                        pend_idx = idx;
                        pend_tile = -1;
                        unsigned tx = (unsigned)(bx - T.lo[0]), ty = (unsigned)(by - T.lo[1]),
                                 tz = (unsigned)(bz - T.lo[2]);
                        if (tx < (unsigned)T.size[0] && ty < (unsigned)T.size[1] && tz < (unsigned)T.size[2])
                            pend_tile = ((int)tz * T.size[1] + (int)ty) * T.size[0] + (int)tx;
                        pend_w = 0.0;
                    }
                    pend_w += (double)dw;
                } else if (C.tally.absorption) {
                    bool in_tile = false;
                    if constexpr (kTile) {
                        unsigned tx = (unsigned)(bx - T.lo[0]), ty = (unsigned)(by - T.lo[1]),
                                 tz = (unsigned)(bz - T.lo[2]);
                        in_tile = tx < (unsigned)T.size[0] && ty < (unsigned)T.size[1] && tz < (unsigned)T.size[2];
                        if (in_tile) atomicAdd(tile + ((int)tz * T.size[1] + (int)ty) * T.size[0] + (int)tx, (double)dw);
                    }
                    if (!in_tile) unsafeAtomicAdd(C.tally.absorption + idx, (double)dw);
                }
                w_abs += (double)dw;
                if (C.bias.on) {
                    // ComputePhotonDirectionWithBiasing (:398-418): draw deflections until one is
                    // accepted with probability (0.1 + 0.45 (1 + d.a))^factor, then weight it by
                    // the weight integral of the old direction (GetNewPhotonWeight, :420-434)
                    V3 A = attractive(C.bias, V3{C.start_pos[0], C.start_pos[1], C.start_pos[2]}, pos, layer);
                    float d0 = dot(dir, A);
                    V3 nd;
                    float p, u;
                    do {
                        nd = spin(dir, Lg, x, draws);
                        p = 0.1f + 0.9f * ((1.0f + dot(nd, A)) / 2.0f);
                        if (C.bias.factor != 1.0f) p = xexpf(C.bias.factor * xlogf(p));
                        u = next_u(x);
                        draws++;
                    } while (p <= u && draws <= kDrawLimit);
                    int k = (int)__builtin_floorf((d0 + 1.0f) * C.bias.half_n[layer] + 0.5f);
                    k = k < 0 ? 0 : (k > C.bias.last[layer] ? C.bias.last[layer] : k);
                    w = w * C.bias.table[C.bias.off[layer] + k] / p;
                    dir = nd;
                } else {
                    dir = spin(dir, Lg, x, draws);
                }
                n_scatter++;
                rec.put(XMC_EV_SCATTER, layer, idx, draws, pos, dir, w, tim);
            } else {
                alive = false;
                atomicAdd(ledger + kArea, (double)w);
                rec.put(XMC_EV_DIE_AREA, layer, 0, draws, pos, dir, w, tim);
            }
        }
        // ---- termination: xmcml kill (:121-124) or MCML roulette
        if (C.roulette) {
            if (w < C.min_weight && alive) {
                n_roulette++;
                if (w == 0.0f) {
                    alive = false;
                    rec.put(XMC_EV_ROULETTE_DIE, layer, 0, draws, pos, dir, w, tim);
                } else {
                    float u = next_u(x);
                    draws++;
                    if (u < C.chance) {
                        float before = w;
                        w /= C.chance;
                        atomicAdd(ledger + kRGain, (double)w - (double)before);
                        rec.put(XMC_EV_ROULETTE_LIVE, layer, 0, draws, pos, dir, w, tim);
                    } else {
                        alive = false;
                        atomicAdd(ledger + kRLost, (double)w);
                        rec.put(XMC_EV_ROULETTE_DIE, layer, 0, draws, pos, dir, w, tim);
                    }
                }
            }
        } else if (w < C.min_weight) {
            if (alive) {
                atomicAdd(ledger + kKill, (double)w);
                rec.put(XMC_EV_DIE_WEIGHT, layer, 0, draws, pos, dir, w, tim);
            }
            alive = false;
        }
        if (TRACE && !alive && rec.ev != nullptr) {
            uint64_t k = (uint64_t)(rec.ev - P.trace.events) / (uint64_t)P.trace.max_events;
            P.trace.counts[k] = rec.n;
        }
    }
    if constexpr (MERGE) flush_pending();
    if (n_photons) atomicAdd(P.tally.stats + 0, (unsigned long long)n_photons);
    if (n_scatter) atomicAdd(P.tally.stats + 1, (unsigned long long)n_scatter);
    if (n_boundary) atomicAdd(P.tally.stats + 2, (unsigned long long)n_boundary);
    if (n_roulette) atomicAdd(P.tally.stats + 3, (unsigned long long)n_roulette);
    if (st.spills) atomicAdd(P.tally.stats + 4, (unsigned long long)st.spills);
    if (n_culled) atomicAdd(P.tally.stats + 5, (unsigned long long)n_culled);
    if (n_overflow) atomicAdd(P.tally.stats + 6, (unsigned long long)n_overflow);
    if (n_entry) atomicAdd(P.tally.stats + 7, (unsigned long long)n_entry);
    atomicAdd(ledger + kAbs, w_abs);
    __syncthreads();
    if (tid < kLedger) unsafeAtomicAdd(reinterpret_cast<double*>(P.tally.stats + 8) + tid, ledger[tid]);
    for (int i = tid; i < ndl; i += blockDim.x) {
        if (dcount[i] == 0) continue;
        unsafeAtomicAdd(P.tally.det_weight + i, dsum[i]);
        unsafeAtomicAdd(P.tally.det_target + i, dsum[ndl + i]);
        atomicAdd(P.tally.det_count + i, dcount[i]);
    }
    if constexpr (kTile) {
        const Area& A = P.area;
        for (int i = tid; i < T.count; i += blockDim.x) {
            double v = tile[i];
            if (v == 0.0) continue;
            int tx = i % T.size[0], ty = (i / T.size[0]) % T.size[1], tz = i / (T.size[0] * T.size[1]);
            int g = ((T.lo[0] + tx) * A.part[1] + (T.lo[1] + ty)) * A.part[2] + (T.lo[2] + tz);
            unsafeAtomicAdd(P.tally.absorption + g, v);
        }
    }
}

// The production variants (no trace, no path tallies) trade a few spills for more
// waves per SIMD.  Interleaved: 96 VGPRs, five waves instead of four (heightfield on
// the walk 2.93e7 -> 3.29e7 photons/s; six waves, 80 VGPRs, spill enough to fall to
// 2.54e7).  Plain: 72 VGPRs, seven waves (the heightfield under its lattice rule:
// 5.25e7 at five, 5.49e7 at six, 5.58e7 at seven, 4.64e7 at eight; the slab, whose
// absorption tile's LDS caps its blocks per CU, is the same at five to seven).  The
// trace and path-tally variants keep the registers they need.
#ifndef XMC_PLAIN_WAVES
#define XMC_PLAIN_WAVES 7
#endif
#ifndef XMC_INTERLEAVED_WAVES
#define XMC_INTERLEAVED_WAVES 5
#endif
template <bool TRACE, bool PATHS, bool INTERLEAVE>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu((!TRACE && !PATHS) ? (INTERLEAVE ? XMC_INTERLEAVED_WAVES : XMC_PLAIN_WAVES) : 1)))
    transport_kernel(Params P) {
    transport_body<TRACE, false, PATHS, INTERLEAVE>(P, nullptr, nullptr, 0);
}

// The plain variant of a single scene that tallies an absorption grid, with its
// deposits merged per voxel (MERGE in transport_body).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XMC_PLAIN_WAVES)))
    transport_merge_kernel(Params P) {
    transport_body<false, false, false, false, true>(P, nullptr, nullptr, 0);
}

// Many configurations in one launch: P carries the run-wide fields and the
// photon range [0, prefix[ncfg]); cfgs[c] the geometry, optics, detectors and
// tallies of configuration c.
// six waves per SIMD for the batch kernel (four kernel-time measurements on the 1024
// layered configurations at a quarter of their budgets: 4345 ms at its natural five,
// 4246 ms at six, 4803 ms at seven)
#ifndef XMC_BATCH_WAVES
#define XMC_BATCH_WAVES 6
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XMC_BATCH_WAVES)))
    batch_kernel(Params P, const Params* __restrict__ cfgs, const uint64_t* __restrict__ prefix, int ncfg) {
    transport_body<false, true, false, false>(P, cfgs, prefix, ncfg);
}

template __global__ void transport_kernel<false, false, true>(Params);
template __global__ void transport_kernel<true, false, true>(Params);
template __global__ void transport_kernel<false, true, true>(Params);
template __global__ void transport_kernel<true, true, true>(Params);
template __global__ void transport_kernel<false, false, false>(Params);
template __global__ void transport_kernel<true, false, false>(Params);
template __global__ void transport_kernel<false, true, false>(Params);
template __global__ void transport_kernel<true, true, false>(Params);

// The nearest hit of ray i, through the BVH or over every triangle, written to
// the query's outputs.
__device__ __forceinline__ void query_one(const QueryParams& Q, const int* perm, int i, Stack& st) {
    V3 o = {Q.origins[3 * i], Q.origins[3 * i + 1], Q.origins[3 * i + 2]};
    V3 d = {Q.directions[3 * i], Q.directions[3 * i + 1], Q.directions[3 * i + 2]};
    float step = Q.steps[i];
    bool limited = step > 0.0f;
    Ray r = make_ray(o, d);
    Hit h;
    if (Q.brute) {
        h.dist = kInf;
        h.tri = -1;
        int last;
        for (int g = 0; g < Q.n_tris; ++g) {
            int k = perm[g];
            float t = tri_distance(Q.tris, k, r, last);
            if (t >= 0.0f && t < h.dist) {
                h.dist = t;
                h.tri = k;
            }
        }
        h.found = h.tri >= 0 && (!limited || h.dist <= step);
    } else {
        Geo geo{Q.nodes, Q.tris, Q.root_ref, {Q.root_lo[0], Q.root_lo[1], Q.root_lo[2]},
                {Q.root_hi[0], Q.root_hi[1], Q.root_hi[2]}};
        h = walk(geo, r, limited ? step : kInf, limited, st);
    }
    Q.found[i] = h.found ? 1 : 0;
    Q.distance[i] = h.found ? h.dist : 0.0f;
    if (h.found) {
        const float4* T = reinterpret_cast<const float4*>(Q.tris) + 3 * h.tri;
        Q.surface[i] = __float_as_int(T[0].w);
        Q.triangle[i] = __float_as_int(T[1].w);
        V3 n = tri_normal(Q.tris, h.tri);
        Q.normal[3 * i] = n.x;
        Q.normal[3 * i + 1] = n.y;
        Q.normal[3 * i + 2] = n.z;
    } else {
        Q.surface[i] = -1;
        Q.triangle[i] = -1;
        Q.normal[3 * i] = Q.normal[3 * i + 1] = Q.normal[3 * i + 2] = 0.0f;
    }
}

// One ray per thread.
__global__ void __launch_bounds__(256) query_kernel(QueryParams Q, const int* perm) {
    extern __shared__ int2 lds_stack[];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st{lds_stack, (int)blockDim.x, (int)threadIdx.x, Q.spill, (int)(gridDim.x * blockDim.x), i,
             Q.spill_depth, 0, 0};
    if (i >= Q.n) return;
    query_one(Q, perm, i, st);
}

// A fixed grid of lanes, each taking rays i, i + lanes, ...: the stack spill
// area is sized by the lanes, not by the rays.
__global__ void __launch_bounds__(256) query_stride_kernel(QueryParams Q, const int* perm) {
    extern __shared__ int2 lds_stack[];
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    const int lanes = gridDim.x * blockDim.x;
    Stack st{lds_stack, (int)blockDim.x, (int)threadIdx.x, Q.spill, lanes, lane, Q.spill_depth, 0, 0};
    for (int i = lane; i < Q.n; i += lanes) query_one(Q, perm, i, st);
}

// Fills n 64-bit words with zero (capture-safe replacement for memset).
__global__ void zero_kernel(unsigned long long* p, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t s = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = 0ull;
}

// The BVH-stack spill of the allocation-free ray query (xmc_scene_query_device):
// 32 MiB per GPU, reserved when the library's code object loads and held until
// the process ends.  Host code never names it: it takes the per-device address
// from query_scratch() below.
static __device__ __align__(256) int2 scratch[kQueryScratchBytes / sizeof(int2)];

}  // namespace xmc

namespace {
std::mutex g_scratch_mutex;
int2* g_scratch_ptr[xmc::kMaxDevices] = {};
}  // namespace

template <bool TRACE, bool PATHS>
static void launch_transport(const xmc::Params* p, bool interleave, int grid, int block, size_t lds, hipStream_t s) {
    if (interleave)
        hipLaunchKernelGGL((xmc::transport_kernel<TRACE, PATHS, true>), dim3(grid), dim3(block), lds, s, *p);
    else
        hipLaunchKernelGGL((xmc::transport_kernel<TRACE, PATHS, false>), dim3(grid), dim3(block), lds, s, *p);
}


// Host-callable launchers (the host translation unit has no kernel syntax).
extern "C" {

// The interleaved loop for scenes with a walk-skipping grid (long walks),
// walks run to the end within the round otherwise.
hipError_t xmc_launch_transport(const xmc::Params* p, int grid, int block, size_t lds, hipStream_t s, int trace) {
    const bool paths = p->paths.scattering != nullptr || p->paths.det_traj != nullptr;
    const bool interleave = p->dist.cells != nullptr;
    if (trace && paths) launch_transport<true, true>(p, interleave, grid, block, lds, s);
    else if (trace) launch_transport<true, false>(p, interleave, grid, block, lds, s);
    else if (paths) launch_transport<false, true>(p, interleave, grid, block, lds, s);
    else if (!interleave && p->tally.absorption != nullptr)
        hipLaunchKernelGGL(xmc::transport_merge_kernel, dim3(grid), dim3(block), lds, s, *p);
    else launch_transport<false, false>(p, interleave, grid, block, lds, s);
    return hipGetLastError();
}

// The address of `scratch` on `device`: looked up once per device, with that
// device current, and kept (each GPU holds its own copy of the array).
hipError_t xmc_query_scratch(int device, int2** ptr, size_t* entries) {
    if (device < 0 || device >= xmc::kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device] == nullptr) {
        hipError_t e = hipSetDevice(device);
        if (e != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(xmc::scratch));
        if (e != hipSuccess) return e;
        if (p == nullptr) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = static_cast<int2*>(p);
    }
    *ptr = g_scratch_ptr[device];
    *entries = xmc::kQueryScratchBytes / sizeof(int2);
    return hipSuccess;
}

// Lanes of query_stride_kernel: at most kQueryLanes, fewer when the tree is so
// deep that their spill would not fit in the scratch array.
hipError_t xmc_launch_query_stride(const xmc::QueryParams* q, const int* perm, int lanes, size_t lds, hipStream_t s) {
    const int block = 256;
    int grid = lanes / block;
    const int need = (q->n + block - 1) / block;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(xmc::query_stride_kernel, dim3(grid), dim3(block), lds, s, *q, perm);
    return hipGetLastError();
}

hipError_t xmc_launch_query(const xmc::QueryParams* q, const int* perm, int block, size_t lds, hipStream_t s) {
    int grid = (q->n + block - 1) / block;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(xmc::query_kernel, dim3(grid), dim3(block), lds, s, *q, perm);
    return hipGetLastError();
}

hipError_t xmc_launch_zero(unsigned long long* p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(xmc::zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t xmc_transport_occupancy(int block, size_t lds, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, xmc::transport_kernel<false, false, true>,
                                                        block, lds);
}

hipError_t xmc_transport_occupancy_plain(int block, size_t lds, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, xmc::transport_kernel<false, false, false>,
                                                        block, lds);
}

hipError_t xmc_transport_occupancy_merge(int block, size_t lds, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, xmc::transport_merge_kernel, block, lds);
}

hipError_t xmc_launch_batch(const xmc::Params* p, const xmc::Params* cfgs, const uint64_t* prefix, int ncfg, int grid,
                            int block, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(xmc::batch_kernel, dim3(grid), dim3(block), lds, s, *p, cfgs, prefix, ncfg);
    return hipGetLastError();
}

hipError_t xmc_batch_occupancy(int block, size_t lds, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, xmc::batch_kernel, block, lds);
}

}  // extern "C"
