/*
 * xmc_oracle.cpp -- CPU oracle for the photon random-walk path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference's CPU algorithm
 * so that tests can check the GPU engine against it and bench.py can time it
 * as the CPU baseline.  The product library never links it.
 *
 * What each part follows (paths relative to the reference repository):
 *   MCG59 generator ............ base/xmcml/xmcml/mcml_mcg59.cpp:3-41
 *   vector math ................ base/xmcml/xmcml/mcml_math.cpp:6-94
 *   slab box test .............. base/xmcml/xmcml/mcml_intersection.cpp:11-42
 *   Moller-Trumbore ............ mcml_intersection.cpp:44-86
 *   SAH BVH build .............. mcml_intersection.cpp:88-344
 *   BVH walk (step-limited) .... mcml_intersection.cpp:354-466
 *   BVH walk (unlimited) ....... mcml_intersection.cpp:468-584
 *   photon loop ................ base/xmcml/xmcml/mcml_kernel.cpp:39-126
 *   launch, move, area index ... mcml_kernel.cpp:128-227
 *   boundary / Fresnel ......... mcml_kernel.cpp:229-284, 497-616, 644-661
 *   spin, step, drop ........... mcml_kernel.cpp:442-495, 618-642
 *   detectors, TOF, ranges ..... mcml_kernel.cpp:663-858
 *   scattering grid/trajectory . mcml_kernel.cpp:750-810
 *   output init, thread grids .. base/xmcml/xmcmlLauncher/launcher_omp.cpp:10-163
 *   roulette ................... stable/plane_parallel/1.0/McmlGpu/kernel/kernel.cu:214-222,
 *                                called as in kernel.cu:449-452
 *
 * Parity pins: the 64-bit mode is compared bit for bit with the reference
 * sources compiled under oracle/_ref (same streams, same libm), and with the
 * golden values of stable/1.2/xmcml/xmcml_tests (tests/golden).  The 32-bit
 * mode is the same template instantiated in float; its float-only rules are
 * listed at struct Prec<float> and its special functions are checked against
 * libm over every input the engine can produce.
 */
#include "xmc_oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

namespace {

// ----------------------------------------------------------------------------
// MCG59 (mcml_mcg59.cpp:3-41)
// ----------------------------------------------------------------------------
constexpr uint64_t kMcgC = 302875106592253ULL;      // 13^13
constexpr uint64_t kMcgM = 576460752303423488ULL;   // 2^59
constexpr uint64_t kMcgMask = 576460752303423487ULL;
constexpr int kStreamLog2 = 24;                     // draws reserved per photon
// photons about to run past their block of draws are retired
constexpr uint32_t kDrawLimit = (1u << kStreamLog2) - 64u;

uint64_t mcg_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r *= a;
        a *= a;
        e >>= 1;
    }
    return r;
}

struct Mcg59 {
    uint64_t value;
    uint64_t offset;
    uint32_t draws;
    uint64_t next() {
        value = (value * offset) & kMcgMask;
        ++draws;
        return value;
    }
};

uint64_t photon_stream(uint64_t seed, uint64_t photon) {
    const uint64_t jump = mcg_pow(kMcgC, 1ULL << kStreamLog2);
    return ((2 * seed + 1) * mcg_pow(jump, photon)) & kMcgMask;
}

// ----------------------------------------------------------------------------
// float special functions of the 32-bit mode.
// log: x = 2^k (1+f), 1+f in [sqrt(2)/2, sqrt(2)); log(1+f) = 2 atanh(s),
// s = f/(2+f), written as f - (f^2/2 - s (f^2/2 + R)), R = series of
// 2 atanh(s)/s - 2 truncated after s^8 (Abramowitz & Stegun 4.1.27).
// sincos(2 pi u): q = rint(4u), r = 4u - q (exact), theta = r pi/2 in
// [-pi/4, pi/4], Taylor series to theta^9 / theta^10, quadrant rotation by q.
// ----------------------------------------------------------------------------
inline float f_logf(float x) {
    uint32_t ix;
    std::memcpy(&ix, &x, 4);
    int32_t t = (int32_t)(ix - 0x3f3504f3u);
    int32_t k = t >> 23;
    uint32_t im = ix - ((uint32_t)k << 23);
    float m;
    std::memcpy(&m, &im, 4);
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = std::fmaf(z, 2.0f / 9.0f, 2.0f / 7.0f);
    R = std::fmaf(z, R, 2.0f / 5.0f);
    R = std::fmaf(z, R, 2.0f / 3.0f);
    R = z * R;
    float hfsq = 0.5f * (f * f);
    float l1p = f - (hfsq - s * (hfsq + R));
    float kf = (float)k;
    return std::fmaf(kf, 0.693145751953125f, std::fmaf(kf, 1.42860677e-06f, l1p));
}

// exp(x) of the 32-bit mode (the kernel's xexpf): k = rint(x log2 e), r = x - k ln2
// with ln2 split as in f_logf (Cody-Waite), exp(r) by its Taylor series to r^7,
// scaled by 2^k through the exponent bits; 0 below -87.
inline float f_expf(float x) {
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) return std::numeric_limits<float>::infinity();
    float kf = std::rint(x * 1.44269504f);
    float r = std::fmaf(kf, -0.693145751953125f, x);
    r = std::fmaf(kf, -1.42860677e-06f, r);
    float p = std::fmaf(r, 1.0f / 5040.0f, 1.0f / 720.0f);
    p = std::fmaf(r, p, 1.0f / 120.0f);
    p = std::fmaf(r, p, 1.0f / 24.0f);
    p = std::fmaf(r, p, 1.0f / 6.0f);
    p = std::fmaf(r, p, 0.5f);
    p = std::fmaf(r, p, 1.0f);
    p = std::fmaf(r, p, 1.0f);
    uint32_t bits = (uint32_t)((int)kf + 127) << 23;
    float s;
    std::memcpy(&s, &bits, 4);
    return p * s;
}

inline void f_sincos2pi(float u, float* sp, float* cp) {
    float u4 = 4.0f * u;
    float q = std::rint(u4);
    float r = u4 - q;
    int iq = ((int)q) & 3;
    float th = r * 1.57079637f;
    float t2 = th * th;
    float ps = std::fmaf(t2, 1.0f / 362880.0f, -1.0f / 5040.0f);
    ps = std::fmaf(t2, ps, 1.0f / 120.0f);
    ps = std::fmaf(t2, ps, -1.0f / 6.0f);
    float s = std::fmaf(th * t2, ps, th);
    float pc = std::fmaf(t2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
    pc = std::fmaf(t2, pc, -1.0f / 720.0f);
    pc = std::fmaf(t2, pc, 1.0f / 24.0f);
    pc = std::fmaf(t2, pc, -0.5f);
    float c = std::fmaf(t2, pc, 1.0f);
    switch (iq) {
        case 0: *sp = s; *cp = c; break;
        case 1: *sp = c; *cp = -s; break;
        case 2: *sp = -s; *cp = -c; break;
        default: *sp = -c; *cp = s; break;
    }
}

// ComputeWeightIntegral (mcml_integration.cpp:7-82): 2-point Gauss-Legendre on
// n x n cells of (theta, phi) in [0, pi] x [0, 2 pi], statement for statement.
double weight_integral(double dot, double anisotropy, double eps, unsigned n) {
    const double coeff = 2.0 * sqrt(3.0);
    const double pi = 3.14159265358979323846;
    double anisotropy2 = anisotropy * anisotropy;
    double sqrt_dot = sqrt(1 - dot * dot);
    double xh = pi / n, yh = 2 * pi / n;
    double sum = 0.0;
    for (unsigned i = 0; i < n; ++i) {
        double a = xh * i, b = a + xh;
        double xp = 0.5 * (a + b) + (b - a) / coeff;
        double xm = 0.5 * (a + b) - (b - a) / coeff;
        for (unsigned j = 0; j < n; ++j) {
            double c = yh * j, d = c + yh;
            double yp = 0.5 * (c + d) + (d - c) / coeff;
            double ym = 0.5 * (c + d) - (d - c) / coeff;
            double stxp = sin(xp), stxm = sin(xm), spyp = sin(yp), spym = sin(ym), cpyp = cos(yp), cpym = cos(ym);
            auto term = [&](double st, double cp, double sp) {
                double px = st * cp;
                double fx = dot * px + sqrt_dot * st * sp;
                double tmp = 1 + anisotropy2 - 2 * anisotropy * fx;
                double f = (1 - anisotropy2) / (2 * tmp * sqrt(tmp));
                double p = pow(0.1 + 0.9 * 0.5 * (1 + px), eps);
                return f * p * st;
            };
            double f_xp_yp = term(stxp, cpyp, spyp);
            double f_xm_yp = term(stxm, cpyp, spyp);
            double f_xp_ym = term(stxp, cpym, spym);
            double f_xm_ym = term(stxm, cpym, spym);
            sum += 0.25 * (b - a) * (d - c) * (f_xp_yp + f_xp_ym + f_xm_yp + f_xm_ym);
        }
    }
    return sum / (2.0 * pi);
}

// ----------------------------------------------------------------------------
// precision policies
// ----------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846;

template <class R> struct Prec;

// Precision 65: double arithmetic driven by the 24-bit uniforms the float
// walk draws (the same recorded sequence), for float-vs-double comparisons
// that isolate arithmetic from uniform quantisation.
static bool g_uniform24 = false;
// the 32-bit mode's lattice rule (xmco_set_lattice_rule); rectangles keep theirs
static bool g_lattice_rule = true;

template <> struct Prec<double> {
    static constexpr double kCosZero = 1.0 - 1.0e-12;  // mcml_kernel.cpp:11
    static constexpr double kCos90 = 1.0e-12;           // :12
    static constexpr double kMinDistance = 1.0e-8;      // :13
    static constexpr double kMaxDistance = 1.0e256;     // mcml_intersection.h:7
    static constexpr double kEps = 1.0e-6;              // mcml_intersection.h:6
    static double uniform(uint64_t x) {  // mcml_mcg59.cpp:40
        return g_uniform24 ? (double)(x >> 35) * 0x1p-24 : (double)x / (double)kMcgM;
    }
    static double log_(double u) { return log(u); }
    static double pow_(double p, double e) { return pow(p, e); }
    static void sincos_2pi(double u, double* s, double* c) {
        double psi = 2.0 * kPi * u;  // mcml_kernel.cpp:453-455
        // glibc's sincos: g++ turns the reference's adjacent cos(psi), sin(psi)
        // into this call, and it differs from separate sin and cos by an ulp on
        // some angles; calling it by name keeps every compiler on the reference's
        // arithmetic (the sanitized build uses clang)
        sincos(psi, s, c);
    }
    static double one_minus_sq(double x) { return 1.0 - x * x; }
};

// Float-only rules (mirrored by the GPU kernel):
//  - uniforms carry the top 24 bits of the MCG59 state: (x >> 35) * 2^-24;
//  - log and sincos(2 pi u) are f_logf / f_sincos2pi above;
//  - after each spin the direction gets one first-order renormalisation
//    step, d *= 1.5 - 0.5 |d|^2 (the rotation formula does not preserve
//    |d| = 1 and float rounding would otherwise drift it by ~1e-4);
//  - the scattering angle comes from hg_float (1 - cos and 1 + cos without
//    cancellation) instead of cos first, which loses sin(theta) for the
//    strongly forward scatters of g ~ 0.9;
//  - dot and cross products use fused multiply-adds (dot<float>, cross<float>)
//    and the slab distance (b - o) / d is fma(b, 1/d, -(o * (1/d)));
//  - 1 - x^2 is evaluated as (1 - x)(1 + x); in the spin, sin(theta_u) is
//    sqrt(ux^2 + uy^2) and the rotation is used when that square is above
//    2e-12 (the reference's |uz| <= 1 - 1e-12, which rounds to 1.0f in
//    float); in the Fresnel formulas the "cosine of zero" is 1 - 1e-6;
//  - the 1e-8 mm nudges of the reference (launch, the 1e-10 mm stop short of
//    a surface, and the step after every reflection or refraction) are below
//    the float spacing of coordinates, so they are replaced by an offset
//    scaled to the coordinates, delta = max(|x|,|y|,|z|) * 2^-20 + 2^-22 mm,
//    applied along the unit surface normal: the photon moves to the hit
//    point h, is placed at h - delta*n on the side it came from for the Area
//    test and the detectors, and after a reflection or refraction at
//    h + delta*n on the side its new direction points to (at launch: +delta
//    along z).  It therefore starts every ray strictly on the correct side of
//    the surface it stands on (and of coplanar neighbours sharing an edge).
template <> struct Prec<float> {
    static constexpr float kCosZero = 1.0f - 1.0e-6f;
    static constexpr float kCos90 = 1.0e-12f;
    static constexpr float kMinDistance = 1.0e-8f;
    static constexpr float kMaxDistance = std::numeric_limits<float>::infinity();
    static constexpr float kEps = 1.0e-6f;
    static float uniform(uint64_t x) { return (float)(x >> 35) * 0x1p-24f; }
    static float log_(float u) { return f_logf(u); }
    // the acceptance probability's power (0.1 <= p <= 1): exact for exponent 1
    static float pow_(float p, float e) { return e == 1.0f ? p : f_expf(e * f_logf(p)); }
    static void sincos_2pi(float u, float* s, float* c) { f_sincos2pi(u, s, c); }
    static float one_minus_sq(float x) { return (1.0f - x) * (1.0f + x); }
};

// ----------------------------------------------------------------------------
// vectors (mcml_math.cpp)
// ----------------------------------------------------------------------------
template <class R> struct V3 {
    R x, y, z;
    R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    const R& operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> sub(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> divv(V3<R> a, R b) { return {a.x / b, a.y / b, a.z / b}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> normalize(V3<R> a) { return divv(a, length(a)); }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// float rule: dot and cross products with fused multiply-adds (the GPU kernel's
// dot and cross), dot = fma(a.z, b.z, fma(a.y, b.y, a.x b.x)),
// cross_i = fma(a_j, b_k, -(a_k b_j)); double mode keeps the reference's sums
template <> inline float dot<float>(V3<float> a, V3<float> b) {
    return std::fmaf(a.z, b.z, std::fmaf(a.y, b.y, a.x * b.x));
}
template <> inline V3<float> cross<float>(V3<float> a, V3<float> b) {
    return {std::fmaf(a.y, b.z, -(a.z * b.y)), std::fmaf(a.z, b.x, -(a.x * b.z)), std::fmaf(a.x, b.y, -(a.y * b.x))};
}
template <class R> inline V3<R> plane_normal(V3<R> a, V3<R> b, V3<R> c) {
    return cross(sub(a, b), sub(a, c));
}
// GetPlaneSegmentIntersectionPoint, mcml_math.cpp:59-94
inline V3<double> plane_segment_point(V3<double> a, V3<double> b, double z) {
    const double kMax = 1.0e256;
    V3<double> n = {0.0, 0.0, 1.0};
    V3<double> v = {-a.x, -a.y, z - a.z};
    double d = dot(n, v);
    V3<double> w = sub(b, a);
    double e = dot(n, w);
    V3<double> r;
    if (e != 0.0) {
        r.x = a.x + w.x * (d / e);
        r.y = a.y + w.y * (d / e);
        r.z = a.z + w.z * (d / e);
        if (dot(sub(a, r), sub(b, r)) > 0.0) r = {kMax, kMax, kMax};
    } else {
        r = {kMax, kMax, kMax};
    }
    return r;
}

// The same steps in float, written as the GPU kernel evaluates them: with
// n = e_z the two dot products are exactly z - a.z and b.z - a.z.
inline V3<float> plane_segment_point_f(V3<float> a, V3<float> b, float z) {
    const float kFar = 1.0e30f;
    float d = z - a.z;
    V3<float> w = sub(b, a);
    float e = w.z;
    V3<float> r = {kFar, kFar, kFar};
    if (e != 0.0f) {
        float t = d / e;
        r.x = a.x + w.x * t;
        r.y = a.y + w.y * t;
        r.z = a.z + w.z * t;
        if (dot(sub(a, r), sub(b, r)) > 0.0f) r = {kFar, kFar, kFar};
    }
    return r;
}

// Precision 65 also feeds the double walk the float-rounded inputs the float
// walk sees (optics, geometry, detectors, source), so that a float-vs-double
// comparison isolates the arithmetic.
inline double in_val(double v) { return g_uniform24 ? (double)(float)v : v; }

template <class R> V3<R> cvt(const xmc_double3& v) { return {(R)in_val(v.x), (R)in_val(v.y), (R)in_val(v.z)}; }

// ----------------------------------------------------------------------------
// geometry: triangle test, box test, BVH
// ----------------------------------------------------------------------------
template <class R>
R triangle_distance(V3<R> o, V3<R> d, V3<R> a, V3<R> b, V3<R> c) {  // :44-86
    const R eps = Prec<R>::kEps;
    V3<R> e1 = sub(b, a);
    V3<R> e2 = sub(c, a);
    V3<R> p = cross(d, e2);
    R det = dot(e1, p);
    if (det < eps && det > -eps) return Prec<R>::kMaxDistance;
    R inv = (R)1.0 / det;
    V3<R> t = sub(o, a);
    R u = dot(t, p) * inv;
    if (u < (R)0.0 || u > (R)1.0) return Prec<R>::kMaxDistance;
    V3<R> q = cross(t, e1);
    R v = dot(d, q) * inv;
    if (v < (R)0.0 || u + v > (R)1.0) return Prec<R>::kMaxDistance;
    return dot(e2, q) * inv;
}

template <class R> struct Box { V3<R> lo, hi; };

template <class R>
bool box_hit(const Box<R>& b, const V3<R>& o, const V3<R>& d, R& tnear, R& tfar) {  // :11-42
    const R kMax = Prec<R>::kMaxDistance;
    R tmin = -kMax, tmax = kMax;
    for (int i = 0; i < 3; i++) {
        if (std::fabs(d[i]) > Prec<R>::kEps) {
            R t = (R)1.0 / d[i];
            R t1, t2;
            if constexpr (sizeof(R) == 4) {
                // float rule: (b - o) / d as fma(b, 1/d, -(o * (1/d)))
                const R oi = o[i] * t;
                t1 = std::fmaf(b.lo[i], t, -oi);
                t2 = std::fmaf(b.hi[i], t, -oi);
            } else {
                t1 = (b.lo[i] - o[i]) * t;
                t2 = (b.hi[i] - o[i]) * t;
            }
            if (t2 < t1) { R tt = t1; t1 = t2; t2 = tt; }
            if (t1 > tmin) tmin = t1;
            if (t2 < tmax) tmax = t2;
            if (tmin > tmax || tmax < 0) { tnear = kMax; tfar = kMax; return false; }
        } else {
            if ((b.lo[i] - o[i] > 0) || (b.hi[i] - o[i] < 0)) { tnear = kMax; tfar = kMax; return false; }
        }
    }
    tnear = tmin;
    tfar = tmax;
    return true;
}

struct Node {
    Box<double> box;
    int right, left, offset, count;
};
struct Tri { int surf, index; };

struct Bvh {
    std::vector<Node> nodes;
    std::vector<Tri> tris;
    int root = 0;
};

struct Mesh {  // double geometry as given by the caller
    const xmc_surface* s;
    int n;
    const xmc_double3& vert(const Tri& t, int k) const {
        const xmc_int3& tr = s[t.surf].triangles[t.index];
        int vi = k == 0 ? tr.x : (k == 1 ? tr.y : tr.z);
        return s[t.surf].vertices[vi];
    }
};

inline double comp(const xmc_double3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

Box<double> create_box(const Mesh& m, const Tri* t, int n) {  // :88-123
    Box<double> b;
    const double kMax = 1.0e256;
    b.lo = {kMax, kMax, kMax};
    b.hi = {-kMax, -kMax, -kMax};
    for (int i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++)
            for (int k = 0; k < 3; k++)
                if (comp(m.vert(t[i], k), a) < b.lo[a]) b.lo[a] = comp(m.vert(t[i], k), a);
        for (int a = 0; a < 3; a++)
            for (int k = 0; k < 3; k++)
                if (comp(m.vert(t[i], k), a) > b.hi[a]) b.hi[a] = comp(m.vert(t[i], k), a);
    }
    return b;
}

constexpr int kBins = 16;   // BIN_NUMBER
constexpr int kLeafMax = 5; // MAX_TRIANGLES_IN_LEAVES

inline double centroid(const Mesh& m, const Tri& t, int a) {
    return (comp(m.vert(t, 0), a) + comp(m.vert(t, 1), a) + comp(m.vert(t, 2), a)) / 3.0;
}

void find_split(const Box<double>& box, const Mesh& m, const Tri* t, int n, double& min_sah,
                int& split, int& axis_out) {  // :125-200
    int bins[3][kBins] = {};
    double size[3], d[3];
    for (int a = 0; a < 3; a++) {
        size[a] = box.hi[a] - box.lo[a];
        d[a] = size[a] / kBins;
    }
    for (int i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            int bi;
            if (d[a] > 1e-6) {
                double c = centroid(m, t[i], a);
                bi = (int)((c - box.lo[a]) / d[a]);
                if (bi == kBins) bi--;
            } else {
                bi = 0;
            }
            bins[a][bi]++;
        }
    }
    double inv_area = 0.5 / (size[0] * size[1] + size[1] * size[2] + size[0] * size[2]);
    min_sah = 1.0e256;
    for (int a = 0; a < 3; a++) {
        if (d[a] > 1e-6) {
            int lc = 0, rc = n;
            double tmp = size[a];
            for (int i = 1; i < kBins; i++) {
                lc += bins[a][i - 1];
                rc -= bins[a][i - 1];
                if (lc > 0 && rc > 0) {
                    size[a] = i * d[a];
                    double la = 2.0 * (size[0] * size[1] + size[1] * size[2] + size[2] * size[0]);
                    size[a] = tmp - size[a];
                    double ra = 2.0 * (size[0] * size[1] + size[1] * size[2] + size[2] * size[0]);
                    double sah = inv_area * (lc * la + ra * rc);
                    if (sah < min_sah) {
                        min_sah = sah;
                        axis_out = a;
                        split = i;
                    }
                }
            }
            size[a] = tmp;
        }
    }
}

int build_node(Bvh& bvh, const Mesh& m, int offset, int n) {  // :202-314
    Tri* t = bvh.tris.data() + offset;
    Box<double> box = create_box(m, t, n);
    auto leaf = [&]() {
        Node nd;
        nd.box = box;
        nd.right = -1;
        nd.left = -1;
        nd.count = n;
        nd.offset = offset;
        bvh.nodes.push_back(nd);
        return (int)bvh.nodes.size() - 1;
    };
    if (n <= kLeafMax) return leaf();
    int split = 0, axis = 0;
    double min_sah;
    find_split(box, m, t, n, min_sah, split, axis);
    double sah = n;
    if (sah <= min_sah) return leaf();
    double d = (box.hi[axis] - box.lo[axis]) / kBins;
    auto bin_of = [&](int i) {
        int b = (int)((centroid(m, t[i], axis) - box.lo[axis]) / d);
        if (b == kBins) b--;
        return b;
    };
    int lc = 0, k = 0, j = n - 1;
    int i1 = bin_of(k), i2 = bin_of(j);
    while (k < j) {
        if (i1 < split) {
            k++;
            lc++;
            if (k < n) i1 = bin_of(k);
        }
        if (i2 >= split) {
            j--;
            if (j >= 0) i2 = bin_of(j);
        }
        if (i1 >= split && i2 < split) {
            Tri tmp = t[k];
            t[k] = t[j];
            t[j] = tmp;
            k++;
            j--;
            lc++;
            if (k < n) i1 = bin_of(k);
            if (j >= 0) i2 = bin_of(j);
        }
    }
    Node nd;
    nd.box = box;
    nd.count = 0;
    nd.offset = -1;
    nd.left = build_node(bvh, m, offset, lc);
    nd.right = build_node(bvh, m, offset + lc, n - lc);
    bvh.nodes.push_back(nd);
    return (int)bvh.nodes.size() - 1;
}

void build_bvh(Bvh& bvh, const xmc_surface* s, int ns) {  // :316-344
    bvh.nodes.clear();
    bvh.tris.clear();
    for (int i = 0; i < ns; i++)
        for (int j = 0; j < s[i].numberOfTriangles; j++) bvh.tris.push_back({i, j});
    Mesh m{s, ns};
    bvh.root = build_node(bvh, m, 0, (int)bvh.tris.size());
}

// ----------------------------------------------------------------------------
// the scene in precision R
// ----------------------------------------------------------------------------
template <class R> struct Hit {
    V3<R> normal;
    R distance;
    int surface;
    int found;
    int tri;  // position in bvh.tris
};

template <class R> struct Scene {
    const xmc_input* in = nullptr;
    Bvh bvh;
    std::vector<Box<R>> boxes;
    std::vector<V3<R>> tv;  // 3 vertices per bvh triangle, in bvh.tris order
    struct Layer { R n, mua, mus, g; bool glass; };
    std::vector<Layer> layers;
    V3<R> area_corner, area_length;
    R area_part[3];
    int part[3];
    struct Cube { V3<R> center, length; uint8_t filter[XMC_MAX_LAYERS]; R cos_target; };
    struct Ring { V3<R> center; R rs, rb; uint8_t filter[XMC_MAX_LAYERS]; R cos_target; };
    std::vector<Cube> cubes;
    std::vector<Ring> rings;
    double rsp = 0.0;
    R w0;
    R min_weight, chance;
    V3<R> start_pos, start_dir;
    // time bins as InitOutput builds them (launcher_omp.cpp:38-55)
    std::vector<double> t_start, t_finish;
    // attractive-vector biasing (mcml_kernel.cpp:325-440): the tables and each
    // layer's (GetWeightTableIndex, :436-446: the last one of equal anisotropy)
    bool biasing = false;
    R factor = (R)1;
    V3<R> target;
    std::vector<std::vector<double>> tables;
    std::vector<std::vector<float>> tables_f;  // the 32-bit mode's float copies
    std::vector<int> layer_table;
    // the 32-bit mode's surface rule (the engine's float rule for scenes whose every
    // surface is an axis-aligned rectangle at one z -- two triangles sharing a
    // diagonal -- or a heightfield on one uniform lattice shared by all of them)
    bool rule = false;
    struct PlanarForm {
        int n = 0;
        float z[16], x0[16], x1[16], y0[16], y1[16];
        int tri[16];
    } planar;
    struct LatticeRule {
        int ns = 0, nx = 0, ny = 0;
        float x0 = 0, y0 = 0, inv_dx = 0, inv_dy = 0, mxy = 0, mz = 0, zlo = 0, zhi = 0;
        std::vector<float> zmin, zmax;  // per (cell, heightfield)
        std::vector<int> ta, tb;        // its triangles (bvh positions, lower first; -1 = none)
    } lattice;

    // GetLocalAttractiveVector (:325-362), statement for statement
    V3<R> attractive(const V3<R>& pos, int l) const {
        const Layer& L = layers[l];
        R D = (R)1 / ((R)3 * (((R)1 - L.g) * L.mus + L.mua));
        R ltr = (R)3 * D;
        R delta = (R)0.74 * ltr;
        V3<R> r = sub(pos, start_pos);
        V3<R> rd = sub(target, start_pos);
        V3<R> r1 = rd;
        r1.z += ltr;
        V3<R> r2 = rd;
        r2.z -= ltr + (R)2 * delta;
        V3<R> t1 = sub(r, r1);
        R l1 = length(t1);
        t1 = divv(t1, l1 * l1 * l1);
        V3<R> t2 = sub(r, r2);
        R l2 = length(t2);
        t2 = divv(t2, l2 * l2 * l2);
        return normalize(sub(t2, t1));
    }

    // GetNewPhotonWeight (:420-434): the weight times the table entry of the old
    // direction's cosine to the attractive vector, over the acceptance probability
    R biased_weight(R w, const V3<R>& old_dir, const V3<R>& a, R p, int l) const {
        int t = layer_table[l];
        int n = (int)tables[t].size() - 1;
        R d = dot(old_dir, a);
        int k;
        if constexpr (sizeof(R) == 4) {
            k = (int)std::floor((d + 1.0f) * (0.5f * (float)n) + 0.5f);
        } else {
            double lo = -1.0, hi = 1.0, step = (hi - lo) / n;
            k = (int)floor((d - lo) / step + 0.5);
        }
        k = k < 0 ? 0 : (k > n ? n : k);
        if constexpr (sizeof(R) == 4) return w * tables_f[t][k] / p;
        return w * (R)tables[t][k] / p;
    }

    void init(const xmc_input* input) {
        in = input;
        build_bvh(bvh, input->surface, input->numberOfSurfaces);
        boxes.resize(bvh.nodes.size());
        for (size_t i = 0; i < bvh.nodes.size(); i++) {
            const Node& n = bvh.nodes[i];
            boxes[i].lo = {(R)in_val(n.box.lo.x), (R)in_val(n.box.lo.y), (R)in_val(n.box.lo.z)};
            boxes[i].hi = {(R)in_val(n.box.hi.x), (R)in_val(n.box.hi.y), (R)in_val(n.box.hi.z)};
        }
        Mesh m{input->surface, input->numberOfSurfaces};
        tv.resize(bvh.tris.size() * 3);
        for (size_t i = 0; i < bvh.tris.size(); i++)
            for (int k = 0; k < 3; k++) tv[3 * i + k] = cvt<R>(m.vert(bvh.tris[i], k));
        layers.resize(input->numberOfLayers);
        for (int i = 0; i < input->numberOfLayers; i++) {
            const xmc_layer& l = input->layerInfo[i];
            layers[i] = {(R)in_val(l.refractiveIndex), (R)in_val(l.absorptionCoefficient),
                         (R)in_val(l.scatteringCoefficient), (R)in_val(l.anisotropy),
                         l.absorptionCoefficient == 0.0 && l.scatteringCoefficient == 0.0};
        }
        area_corner = cvt<R>(input->area->corner);
        area_length = cvt<R>(input->area->length);
        part[0] = input->area->partitionNumber.x;
        part[1] = input->area->partitionNumber.y;
        part[2] = input->area->partitionNumber.z;
        for (int a = 0; a < 3; a++) area_part[a] = (R)part[a];
        (void)0;
        for (int i = 0; i < input->numberOfCubeDetectors; i++) {
            const xmc_cube_detector& c = input->cubeDetector[i];
            Cube q;
            q.center = cvt<R>(c.center);
            q.length = cvt<R>(c.length);
            std::memcpy(q.filter, c.filterLayers, XMC_MAX_LAYERS);
            q.cos_target = (R)cos(c.targetAngle * kPi / 180.0);
            cubes.push_back(q);
        }
        for (int i = 0; i < input->numberOfRingDetectors; i++) {
            const xmc_ring_detector& c = input->ringDetector[i];
            Ring q;
            q.center = cvt<R>(c.center);
            q.rs = (R)in_val(c.smallRadius);
            q.rb = (R)in_val(c.bigRadius);
            std::memcpy(q.filter, c.filterLayers, XMC_MAX_LAYERS);
            q.cos_target = (R)cos(c.targetAngle * kPi / 180.0);
            rings.push_back(q);
        }
        if (input->useBiasing) init_bias(input);
        if constexpr (sizeof(R) == 4) init_rule(input);
        rsp = xmco_specular(input->layerInfo);
        w0 = (R)in_val(1.0 - rsp);
        min_weight = (R)in_val(input->minWeight);
        chance = (R)in_val(input->rouletteChance);
        start_pos = cvt<R>(input->startPosition);
        start_dir = cvt<R>(input->startDirection);
        int T = input->timeScaleSize;
        t_start.assign(T > 0 ? T : 0, 0.0);
        t_finish.assign(T > 0 ? T : 0, 0.0);
        if (T > 0) {
            double dt = (input->timeFinish - input->timeStart) / T;
            if (dt < 0.0) dt = 0.0;
            for (int j = 0; j < T - 1; j++) {
                t_start[j] = input->timeStart + j * dt;
                t_finish[j] = t_start[j] + dt;
            }
            t_start[T - 1] = input->timeStart + (T - 1) * dt;
            t_finish[T - 1] = input->timeFinish;
        }
    }

    void init_bias(const xmc_input* input) {
        biasing = true;
        factor = (R)input->attractiveFactor;
        target = cvt<R>(input->targetPoint);
        std::vector<double> gs;  // GetDifferentAnisotropies (weight_integral.cpp:67-90)
        for (int l = 1; l < input->numberOfLayers; l++) {
            double g = input->layerInfo[l].anisotropy;
            if (std::find(gs.begin(), gs.end(), g) == gs.end()) gs.push_back(g);
        }
        std::vector<double> tg;
        if (input->weightTable && input->numberOfWeightTables > 0) {
            for (int t = 0; t < input->numberOfWeightTables; t++) {
                const xmc_weight_table& w = input->weightTable[t];
                tables.emplace_back(w.elements, w.elements + w.numberOfElements);
                tg.push_back(w.anisotropy);
            }
        } else {  // ComputeWeightIntegral (weight_integral.cpp:33-64)
            const int bp = input->weightIntegralPrecision > 0 ? input->weightIntegralPrecision : 128;
            const int tp = input->weightTablePrecision > 0 ? input->weightTablePrecision : 262144;
            for (double g : gs) {
                unsigned n = bp;
                if (g > 0.98) n *= 16;
                else if (g > 0.97) n *= 8;
                else if (g > 0.95) n *= 4;
                else if (g > 0.90) n *= 2;
                std::vector<double> e(tp + 1);
                const double lo = -1.0, hi = 1.0, step = (hi - lo) / tp;
#pragma omp parallel for schedule(dynamic, 64)
                for (int j = 0; j <= tp; j++) e[j] = weight_integral(lo + j * step, g, input->attractiveFactor, n);
                tables.push_back(e);
                tg.push_back(g);
            }
        }
        for (const auto& e : tables) tables_f.emplace_back(e.begin(), e.end());
        layer_table.assign(input->numberOfLayers, -1);
        for (int l = 0; l < input->numberOfLayers; l++)
            for (size_t t = 0; t < tg.size(); t++)
                if (tg[t] == input->layerInfo[l].anisotropy) layer_table[l] = (int)t;
    }

    // a rectangle: two triangles over the four corners of an axis-aligned
    // rectangle at one z, sharing a diagonal (z, x0, x1, y0, y1 in float)
    static bool rect_form(const xmc_surface& S, float out[5]) {
        if (S.numberOfTriangles != 2) return false;
        const int idx[6] = {S.triangles[0].x, S.triangles[0].y, S.triangles[0].z,
                            S.triangles[1].x, S.triangles[1].y, S.triangles[1].z};
        for (int k = 0; k < 6; k++)
            if (idx[k] < 0 || idx[k] >= S.numberOfVertices) return false;
        const double z0 = S.vertices[idx[0]].z;
        double xlo = S.vertices[idx[0]].x, xhi = xlo, ylo = S.vertices[idx[0]].y, yhi = ylo;
        for (int k = 0; k < 6; k++) {
            const xmc_double3& v = S.vertices[idx[k]];
            if (v.z != z0) return false;
            xlo = std::min(xlo, v.x);
            xhi = std::max(xhi, v.x);
            ylo = std::min(ylo, v.y);
            yhi = std::max(yhi, v.y);
        }
        if (!(xlo < xhi && ylo < yhi)) return false;
        int cover[2] = {0, 0};
        for (int k = 0; k < 6; k++) {
            const xmc_double3& v = S.vertices[idx[k]];
            const int cx = v.x == xlo ? 0 : (v.x == xhi ? 1 : -1);
            const int cy = v.y == ylo ? 0 : (v.y == yhi ? 1 : -1);
            if (cx < 0 || cy < 0) return false;
            cover[k / 3] |= 1 << (cx + 2 * cy);
        }
        auto count = [](int m) { int c = 0; for (int b = 0; b < 4; b++) c += (m >> b) & 1; return c; };
        if (count(cover[0]) != 3 || count(cover[1]) != 3 || (cover[0] | cover[1]) != 15) return false;
        const int shared = cover[0] & cover[1];
        if (shared != 9 && shared != 6) return false;
        out[0] = (float)z0;
        out[1] = (float)xlo;
        out[2] = (float)xhi;
        out[3] = (float)ylo;
        out[4] = (float)yhi;
        return true;
    }

    // the distinct values of a lattice axis, uniformly spaced within 1e-6 of a step
    static bool lattice_axis(std::vector<double>& v, double& origin, double& step) {
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

    // the heightfields' lattice: every triangle spans one cell along x and y, at
    // most two per (cell, heightfield); margins as the engine states them
    bool init_lattice(const xmc_input* input, const std::vector<int>& hs) {
        const xmc_surface* S = input->surface;
        std::vector<double> xs, ys;
        double axy = 0.0, az = 0.0;
        for (int s : hs)
            for (int t = 0; t < S[s].numberOfTriangles; t++) {
                const int id[3] = {S[s].triangles[t].x, S[s].triangles[t].y, S[s].triangles[t].z};
                for (int k = 0; k < 3; k++) {
                    if (id[k] < 0 || id[k] >= S[s].numberOfVertices) return false;
                    const xmc_double3& v = S[s].vertices[id[k]];
                    if (!std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z)) return false;
                    xs.push_back(v.x);
                    ys.push_back(v.y);
                    axy = std::max(axy, std::max(std::fabs(v.x), std::fabs(v.y)));
                    az = std::max(az, std::fabs(v.z));
                }
            }
        double x0, y0, dx, dy;
        if (!lattice_axis(xs, x0, dx) || !lattice_axis(ys, y0, dy)) return false;
        LatticeRule& L = lattice;
        L.nx = (int)xs.size() - 1;
        L.ny = (int)ys.size() - 1;
        L.ns = (int)hs.size();
        if ((double)L.nx * L.ny * L.ns > (double)(1 << 26)) return false;
        const size_t nrec = (size_t)L.nx * L.ny * L.ns;
        const float inf = std::numeric_limits<float>::infinity();
        L.zmin.assign(nrec, inf);
        L.zmax.assign(nrec, -inf);
        L.ta.assign(nrec, -1);
        L.tb.assign(nrec, -1);
        std::vector<int> first(input->numberOfSurfaces + 1, 0);
        for (int s = 0; s < input->numberOfSurfaces; s++) first[s + 1] = first[s] + S[s].numberOfTriangles;
        std::vector<int> pos(first.back(), -1);
        for (size_t i = 0; i < bvh.tris.size(); i++) pos[first[bvh.tris[i].surf] + bvh.tris[i].index] = (int)i;
        float zlo = inf, zhi = -inf;
        for (int l = 0; l < L.ns; l++) {
            const int s = hs[l];
            for (int t = 0; t < S[s].numberOfTriangles; t++) {
                const int id[3] = {S[s].triangles[t].x, S[s].triangles[t].y, S[s].triangles[t].z};
                int ix[3], iy[3];
                float lo = inf, hi = -inf;
                for (int k = 0; k < 3; k++) {
                    const xmc_double3& v = S[s].vertices[id[k]];
                    ix[k] = (int)std::lround((v.x - x0) / dx);
                    iy[k] = (int)std::lround((v.y - y0) / dy);
                    lo = std::min(lo, (float)v.z);
                    hi = std::max(hi, (float)v.z);
                }
                const int cx = std::min(ix[0], std::min(ix[1], ix[2])), cy = std::min(iy[0], std::min(iy[1], iy[2]));
                if (std::max(ix[0], std::max(ix[1], ix[2])) != cx + 1 || std::max(iy[0], std::max(iy[1], iy[2])) != cy + 1)
                    return false;
                const size_t r = ((size_t)cy * L.nx + cx) * L.ns + l;
                const int p = pos[first[s] + t];
                if (L.tb[r] >= 0) return false;
                if (L.ta[r] < 0) {
                    L.ta[r] = p;
                } else {
                    L.tb[r] = std::max(L.ta[r], p);
                    L.ta[r] = std::min(L.ta[r], p);
                }
                L.zmin[r] = std::min(L.zmin[r], lo);
                L.zmax[r] = std::max(L.zmax[r], hi);
                zlo = std::min(zlo, lo);
                zhi = std::max(zhi, hi);
            }
        }
        const double ulp = std::ldexp(1.0, -22);
        L.x0 = (float)x0;
        L.y0 = (float)y0;
        L.inv_dx = (float)(1.0 / dx);
        L.inv_dy = (float)(1.0 / dy);
        L.mxy = (float)(1e-3 * std::min(dx, dy) + 4.0 * std::max(L.nx, L.ny) * ulp * std::max(dx, dy) + 8.0 * ulp * axy);
        L.mz = (float)(1e-4 * (1.0 + ((double)zhi - (double)zlo)) + 8.0 * ulp * az);
        L.zlo = zlo;
        L.zhi = zhi;
        return true;
    }

    // the surface rule: rectangles in surface order (each with its first triangle
    // in bvh order) and the other surfaces as heightfields on one lattice; off
    // (rule false, both forms empty) for any other scene, and for scenes with
    // heightfields while the lattice rule is switched off
    void init_rule(const xmc_input* input) {
        rule = false;
        planar = PlanarForm();
        lattice = LatticeRule();
        const int ns = input->numberOfSurfaces;
        if (ns < 1) return;
        std::vector<int> rect_of(ns, -1), hs;
        for (int si = 0; si < ns; si++) {
            float f[5];
            if (rect_form(input->surface[si], f)) {
                if (planar.n == 16) { planar = PlanarForm(); return; }
                rect_of[si] = planar.n;
                planar.z[planar.n] = f[0];
                planar.x0[planar.n] = f[1];
                planar.x1[planar.n] = f[2];
                planar.y0[planar.n] = f[3];
                planar.y1[planar.n] = f[4];
                planar.tri[planar.n] = -1;
                planar.n++;
            } else {
                hs.push_back(si);
            }
        }
        if (!hs.empty() && (!g_lattice_rule || hs.size() > 4 || !init_lattice(input, hs))) {
            planar = PlanarForm();
            lattice = LatticeRule();
            return;
        }
        for (size_t k = 0; k < bvh.tris.size(); k++) {
            const int r = rect_of[bvh.tris[k].surf];
            if (r >= 0 && planar.tri[r] < 0) planar.tri[r] = (int)k;
        }
        rule = true;
    }

    // the 32-bit mode's nearest surface of a planar scene: t = (z_s - o.z) * (1 / d.z)
    // for |d.z| > EPSILON, hit when t >= 0 and the point lies in the rectangle (edges
    // included); the smallest t, the first surface on ties; reported when t <= step
    Hit<float> planar_hit(const V3<float>& o, const V3<float>& d, float step, bool limited) const {
        Hit<float> res;
        res.found = 0;
        res.distance = Prec<float>::kMaxDistance;
        res.surface = -1;
        res.tri = -1;
        res.normal = {0, 0, 0};
        float best = std::numeric_limits<float>::infinity();
        int tri = -1;
        if (std::fabs(d.z) > 1.0e-6f) {
            const float iz = 1.0f / d.z;
            for (int si = 0; si < planar.n; si++) {
                const float t = (planar.z[si] - o.z) * iz;
                if (t >= 0.0f && t < best) {
                    const float x = o.x + t * d.x, y = o.y + t * d.y;
                    if (x >= planar.x0[si] && x <= planar.x1[si] && y >= planar.y0[si] && y <= planar.y1[si]) {
                        best = t;
                        tri = planar.tri[si];
                    }
                }
            }
        }
        if (tri >= 0 && (!limited || best <= step)) {
            res.found = 1;
            res.distance = best;
            res.tri = tri;
            res.surface = bvh.tris[tri].surf;
            res.normal = tri_normal(tri);
        }
        return res;
    }

    // the 32-bit mode's surface rule: the rectangles' nearest hit as planar_hit
    // finds it; then, for a limited step, over the lattice cells of the step's xy
    // bounding box widened by mxy, per heightfield whose cell z-range meets the
    // step's z-range widened by mz, the cell's triangles with the BVH leaves' test:
    // a nearer one, or one as near with a lower bvh position, replaces the hit.
    // Unlimited steps, and boxes of more than 4 cells along x or y, take the BVH
    // walk over every surface.
    Hit<float> rule_hit(const V3<float>& o, const V3<float>& d, float step, bool limited) const {
        const LatticeRule& L = lattice;
        if (L.ns == 0) return planar_hit(o, d, step, limited);
        if (!limited) return intersect(o, d, step, false);
        float best = std::numeric_limits<float>::infinity();
        int tri = -1;
        if (std::fabs(d.z) > 1.0e-6f) {
            const float iz = 1.0f / d.z;
            for (int si = 0; si < planar.n; si++) {
                const float t = (planar.z[si] - o.z) * iz;
                if (t >= 0.0f && t < best) {
                    const float x = o.x + t * d.x, y = o.y + t * d.y;
                    if (x >= planar.x0[si] && x <= planar.x1[si] && y >= planar.y0[si] && y <= planar.y1[si]) {
                        best = t;
                        tri = planar.tri[si];
                    }
                }
            }
        }
        auto search = [&]() -> bool {  // false: the box is too large for the rule
            const float ez = std::fmaf(step, d.z, o.z);
            const float az = std::fmin(o.z, ez) - L.mz, bz = std::fmax(o.z, ez) + L.mz;
            if (bz < L.zlo || az > L.zhi) return true;
            const float ex = std::fmaf(step, d.x, o.x), ey = std::fmaf(step, d.y, o.y);
            const int i0 = std::max(0, (int)std::floor(std::fmax((std::fmin(o.x, ex) - L.mxy - L.x0) * L.inv_dx, -1.0f)));
            const int i1 = std::min(L.nx - 1, (int)std::floor(std::fmin((std::fmax(o.x, ex) + L.mxy - L.x0) * L.inv_dx,
                                                                        (float)L.nx)));
            const int j0 = std::max(0, (int)std::floor(std::fmax((std::fmin(o.y, ey) - L.mxy - L.y0) * L.inv_dy, -1.0f)));
            const int j1 = std::min(L.ny - 1, (int)std::floor(std::fmin((std::fmax(o.y, ey) + L.mxy - L.y0) * L.inv_dy,
                                                                        (float)L.ny)));
            if (i0 > i1 || j0 > j1) return true;
            if (i1 - i0 >= 4 || j1 - j0 >= 4) return false;
            for (int j = j0; j <= j1; j++)
                for (int i = i0; i <= i1; i++)
                    for (int l = 0; l < L.ns; l++) {
                        const size_t r = ((size_t)j * L.nx + i) * L.ns + l;
                        if (!(L.zmin[r] <= bz && L.zmax[r] >= az)) continue;
                        for (int k : {L.ta[r], L.tb[r]}) {
                            if (k < 0) continue;
                            const float t = tri_t(k, o, d);
                            if (t >= 0.0f && (t < best || (t == best && k < tri))) {
                                best = t;
                                tri = k;
                            }
                        }
                    }
            return true;
        };
        if (!search()) return intersect(o, d, step, true);
        Hit<float> res;
        res.found = 0;
        res.distance = Prec<float>::kMaxDistance;
        res.surface = -1;
        res.tri = -1;
        res.normal = {0, 0, 0};
        if (tri >= 0 && best <= step) {
            res.found = 1;
            res.distance = best;
            res.tri = tri;
            res.surface = bvh.tris[tri].surf;
            res.normal = tri_normal(tri);
        }
        return res;
    }

    V3<R> tri_normal(int ti) const {
        return normalize(plane_normal(tv[3 * ti], tv[3 * ti + 1], tv[3 * ti + 2]));
    }

    R tri_t(int ti, const V3<R>& o, const V3<R>& d) const {
        return triangle_distance(o, d, tv[3 * ti], tv[3 * ti + 1], tv[3 * ti + 2]);
    }

    // ComputeBVHIntersection (:354-466) when limited, else
    // ComputeBVHIntersectionWithoutStep (:468-584).
    Hit<R> intersect(const V3<R>& o, const V3<R>& d, R step, bool limited) const {
        Hit<R> res;
        res.found = 0;
        res.distance = Prec<R>::kMaxDistance;
        res.surface = -1;
        res.tri = -1;
        res.normal = {0, 0, 0};
        R tn1, tn2, tf;
        int cur = bvh.root;
        if (!box_hit(boxes[cur], o, d, tn1, tf)) return res;
        if (limited && step < tn1) return res;
        struct Entry { int node; R tnear; };
        std::vector<Entry> stack;
        stack.reserve(64);
        auto pop = [&](bool& done) {
            if (stack.empty()) { done = true; return; }
            Entry info = stack.back();
            cur = info.node;
            stack.pop_back();
            if (res.distance < info.tnear) {
                while (!stack.empty() && res.distance < info.tnear) {
                    info = stack.back();
                    stack.pop_back();
                }
                if (res.distance < info.tnear || (limited && step < info.tnear)) {
                    done = true;
                    return;
                }
                cur = info.node;
            }
        };
        while (true) {
            const Node& nd = bvh.nodes[cur];
            bool use_stack;
            if (nd.offset >= 0) {
                int best = -1;
                for (int i = 0; i < nd.count; i++) {
                    int ti = nd.offset + i;
                    R t = tri_t(ti, o, d);
                    if (t >= (R)0.0 && t < res.distance) {
                        res.distance = t;
                        best = ti;
                    }
                }
                if (best >= 0 && (!limited || res.distance <= step)) {
                    res.surface = bvh.tris[best].surf;
                    res.tri = best;
                    res.found = 1;
                    res.normal = tri_normal(best);
                }
                use_stack = true;
            } else {
                int left = nd.left, right = nd.right;
                bool f1 = box_hit(boxes[left], o, d, tn1, tf);
                bool f2 = box_hit(boxes[right], o, d, tn2, tf);
                if (limited) {
                    use_stack = false;
                    if (f1 && f2) {
                        if (tn1 <= tn2) {
                            if (tn1 <= step) {
                                cur = left;
                                if (tn2 <= step) stack.push_back({right, tn2});
                            } else use_stack = true;
                        } else {
                            if (tn2 <= step) {
                                cur = right;
                                if (tn1 <= step) stack.push_back({left, tn1});
                            } else use_stack = true;
                        }
                    } else if (f1) {
                        if (tn1 <= step) cur = left;
                        else use_stack = true;
                    } else {
                        if (f2 && tn2 <= step) cur = right;
                        else use_stack = true;
                    }
                } else {
                    use_stack = false;
                    if (f1 && f2) {
                        if (tn1 <= tn2) {
                            cur = left;
                            stack.push_back({right, tn2});
                        } else {
                            cur = right;
                            stack.push_back({left, tn1});
                        }
                    } else if (f1) cur = left;
                    else if (f2) cur = right;
                    else use_stack = true;
                }
            }
            if (use_stack) {
                bool done = false;
                pop(done);
                if (done) return res;
            }
        }
    }

    // brute force over every triangle in (surface, triangle) order
    Hit<R> brute(const V3<R>& o, const V3<R>& d, R step, bool limited) const {
        Hit<R> res;
        res.found = 0;
        res.distance = Prec<R>::kMaxDistance;
        res.surface = -1;
        res.tri = -1;
        res.normal = {0, 0, 0};
        // visit in the caller's order, not bvh order
        std::vector<int> pos_of;  // bvh position of (surface, index)
        const xmc_surface* s = in->surface;
        std::vector<int> base(in->numberOfSurfaces + 1, 0);
        for (int i = 0; i < in->numberOfSurfaces; i++) base[i + 1] = base[i] + s[i].numberOfTriangles;
        pos_of.assign(base.back(), 0);
        for (size_t i = 0; i < bvh.tris.size(); i++) pos_of[base[bvh.tris[i].surf] + bvh.tris[i].index] = (int)i;
        for (int g = 0; g < base.back(); g++) {
            int ti = pos_of[g];
            R t = tri_t(ti, o, d);
            if (t >= (R)0.0 && t < res.distance) {
                res.distance = t;
                res.tri = ti;
            }
        }
        if (res.tri >= 0 && (!limited || res.distance <= step)) {
            res.found = 1;
            res.surface = bvh.tris[res.tri].surf;
            res.normal = tri_normal(res.tri);
        }
        return res;
    }

    // GetAreaIndexVector / GetAreaIndex (:190-227); returns -1 outside
    int area_index(const V3<R>& p, int v[3]) const {
        R ix, iy, iz;
        if constexpr (sizeof(R) == 4) {  // float: partitions per unit length once per scene
            ix = (p.x - area_corner.x) * (area_part[0] / area_length.x);
            iy = (p.y - area_corner.y) * (area_part[1] / area_length.y);
            iz = (p.z - area_corner.z) * (area_part[2] / area_length.z);
        } else {
            ix = area_part[0] * (p.x - area_corner.x) / area_length.x;
            iy = area_part[1] * (p.y - area_corner.y) / area_length.y;
            iz = area_part[2] * (p.z - area_corner.z) / area_length.z;
        }
        bool inside = !((ix < (R)0.0 || ix >= area_part[0]) || (iy < (R)0.0 || iy >= area_part[1]) ||
                        (iz < (R)0.0 || iz >= area_part[2]));
        uint8_t bx = 255, by = 255, bz = 255;
        if (inside) {
            bx = (uint8_t)ix;
            by = (uint8_t)iy;
            bz = (uint8_t)iz;
        }
        v[0] = bx;
        v[1] = by;
        v[2] = bz;
        if (bx == 255 && by == 255 && bz == 255) return -1;
        return bx * part[1] * part[2] + by * part[2] + bz;
    }

    int find_layer(int surface, int current) const {  // :566-580
        for (int i = 0; i < in->numberOfLayers; ++i)
            for (int j = 0; j < in->layerInfo[i].numberOfSurfaces; ++j)
                if (in->layerInfo[i].surfaceId[j] == surface && i != current) return i;
        return -1;
    }
};

// ----------------------------------------------------------------------------
// physics helpers (mcml_kernel.cpp)
// ----------------------------------------------------------------------------
template <class R> R transmit_cos(R n1, R n2, R ci) {  // :497-529
    if (n1 == n2) return ci;
    if (ci > Prec<R>::kCosZero) return ci;
    if (ci < Prec<R>::kCos90) return (R)0.0;
    R si = std::sqrt(Prec<R>::one_minus_sq(ci));
    R st = n1 * si / n2;
    if (st >= (R)1.0) return (R)0.0;
    return std::sqrt(Prec<R>::one_minus_sq(st));
}

template <class R> R fresnel(R n1, R n2, R ci, R ct) {  // :531-564
    if (n1 == n2) return (R)0.0;
    if (ci > Prec<R>::kCosZero) {
        R r = (n2 - n1) / (n2 + n1);
        return r * r;
    }
    if (ci < Prec<R>::kCos90) return (R)1.0;
    R si = std::sqrt(Prec<R>::one_minus_sq(ci));
    R st = std::sqrt(Prec<R>::one_minus_sq(ct));
    R cs = ci * ct - si * st;
    R cd = ci * ct + si * st;
    R ss = si * ct + ci * st;
    R sd = si * ct - ci * st;
    return (R)0.5 * sd * sd * (cd * cd + cs * cs) / (ss * ss * cd * cd);
}

template <class R> R critical_cos(R n1, R n2) {  // :644-661
    if (n1 > n2) return std::sqrt((R)1.0 - n2 * n2 / (n1 * n1));
    return (R)0.0;
}

template <class R> V3<R> reflect(V3<R> i, V3<R> n) {  // :582-593
    V3<R> sn = normalize(n);
    V3<R> si = normalize(i);
    R tmp = (R)2 * dot(si, sn);
    sn.x *= tmp;
    sn.y *= tmp;
    sn.z *= tmp;
    return normalize(sub(si, sn));
}

template <class R> V3<R> refract(R n1, R n2, V3<R> i, V3<R> n) {  // :595-616
    V3<R> si = normalize(i);
    si.x *= n1;
    si.y *= n1;
    si.z *= n1;
    V3<R> sn = normalize(n);
    R t1 = dot(si, sn);
    R t2 = std::sqrt((n2 * n2 - n1 * n1) / (t1 * t1) + (R)1);
    t2 = (t2 - (R)1) * t1;
    V3<R> t = {si.x + t2 * sn.x, si.y + t2 * sn.y, si.z + t2 * sn.z};
    return normalize(t);
}

template <class R> R cosine_theta(R g, Mcg59& rng) {  // :475-495
    R u = Prec<R>::uniform(rng.next());
    R c;
    if (g != (R)0.0) {
        R tmp = ((R)1 - g * g) / ((R)1 - g + (R)2 * g * u);
        c = ((R)1 + g * g - tmp * tmp) / ((R)2 * g);
    } else {
        c = (R)2.0 * u - (R)1.0;
    }
    if (c > (R)1.0) c = (R)1.0;
    else if (c < (R)-1.0) c = (R)-1.0;
    return c;
}

// Float only: the Henyey-Greenstein inversion of ComputeCosineTheta
// (:475-495) rewritten so that 1 - cos(theta) and 1 + cos(theta) come out
// without cancellation.  With D = 1 - g + 2 g u and t = (1 - g^2)/D:
//   1 - cos = (1 - g)(1 - u)(t + 1 - g)/D,   1 + cos = (1 + g) u (1 + g + t)/D,
// exact rearrangements of cos = (1 + g^2 - t^2)/(2g) that also give the
// isotropic 2(1 - u), 2u at g = 0.  sin(theta) = sqrt((1 - cos)(1 + cos)).
inline void hg_float(float g, float u, float* c, float* s) {
    float a = 1.0f - g, b = 1.0f + g;
    float inv = 1.0f / (a + 2.0f * g * u);
    float t = (1.0f - g * g) * inv;
    float vers = a * (1.0f - u) * (t + a) * inv;
    float cove = b * u * (b + t) * inv;
    *c = vers < 1.0f ? 1.0f - vers : cove - 1.0f;
    *s = std::sqrt(vers * cove);
}

template <class R> V3<R> new_direction(V3<R> d, R g, Mcg59& rng) {  // :442-473
    V3<R> r;
    R ux = d.x, uy = d.y, uz = d.z;
    R ct, st;
    if constexpr (sizeof(R) == 4) {
        hg_float(g, Prec<R>::uniform(rng.next()), &ct, &st);
    } else {
        ct = cosine_theta(g, rng);
        st = std::sqrt(Prec<R>::one_minus_sq(ct));
    }
    R sp, cp;
    Prec<R>::sincos_2pi(Prec<R>::uniform(rng.next()), &sp, &cp);
    // float: sin(theta_u) from the components themselves, sqrt(ux^2 + uy^2),
    // so the local frame stays orthonormal near the pole; the branch is the
    // reference's |uz| <= 1 - 1e-12, i.e. ux^2 + uy^2 >= 2e-12
    bool rotate;
    R tmp;
    if constexpr (sizeof(R) == 4) {
        R s2 = ux * ux + uy * uy;
        rotate = s2 > (R)2e-12;
        tmp = std::sqrt(s2);
    } else {
        rotate = std::fabs(uz) <= Prec<R>::kCosZero;
        tmp = rotate ? std::sqrt(Prec<R>::one_minus_sq(uz)) : (R)0;
    }
    if (rotate) {
        if constexpr (sizeof(R) == 4) {  // float: one reciprocal for both components
            R itmp = (R)1.0 / tmp;
            r.x = st * (ux * uz * cp - uy * sp) * itmp + ux * ct;
            r.y = st * (uy * uz * cp + ux * sp) * itmp + uy * ct;
        } else {
            r.x = st * (ux * uz * cp - uy * sp) / tmp + ux * ct;
            r.y = st * (uy * uz * cp + ux * sp) / tmp + uy * ct;
        }
        r.z = -st * cp * tmp + uz * ct;
    } else {
        R tmp = (uz >= (R)0.0) ? (R)1.0 : (R)-1.0;
        r.x = st * cp;
        r.y = st * sp;
        r.z = tmp * ct;
    }
    if constexpr (sizeof(R) == 4) {
        // float only: one first-order renormalisation step, d *= (3 - |d|^2)/2,
        // keeps |d| within float rounding of 1 over long walks
        R f = std::fmaf((R)-0.5, dot(r, r), (R)1.5);
        r.x = r.x * f;
        r.y = r.y * f;
        r.z = r.z * f;
    }
    return r;
}

// MCML roulette (stable/plane_parallel/1.0/McmlGpu/kernel/kernel.cu:214-222),
// applied only to a live photon below the threshold (kernel.cu:449-452): a
// zero weight dies without a draw; otherwise one uniform u, survive when
// u < chance with the weight divided by chance, else die (the weight is kept
// so that the caller can book it as lost).  Expected weight afterwards equals
// the weight before.
enum { kRouletteSkip = 0, kRouletteZero = 1, kRouletteLive = 2, kRouletteDie = 3 };
template <class R> int roulette_step(R& weight, bool& dead, R min_weight, R chance, Mcg59& rng) {
    if (!(weight < min_weight) || dead) return kRouletteSkip;
    if (weight == (R)0.0) {
        dead = true;
        return kRouletteZero;
    }
    if (Prec<R>::uniform(rng.next()) < chance) {
        weight /= chance;
        return kRouletteLive;
    }
    dead = true;
    return kRouletteDie;
}

template <class R> R step_size(R mua, R mus, Mcg59& rng) {  // :618-631
    R u;
    do {
        u = Prec<R>::uniform(rng.next());
    } while (u == (R)0.0);
    if constexpr (sizeof(R) == 4) return -Prec<R>::log_(u) * ((R)1.0 / (mua + mus));  // float: reciprocal per layer
    return -Prec<R>::log_(u) / (mua + mus);
}

// ----------------------------------------------------------------------------
// per-thread tallies (launcher_omp.cpp:60-74 keeps absorption/scattering per
// thread; detector fields are shared with omp atomic in the reference -- the
// oracle keeps them per thread and sums at the end)
// ----------------------------------------------------------------------------
struct Tally {
    std::vector<double> absorption, scattering;
    std::vector<double> det_weight, det_target;
    std::vector<uint64_t> det_count;
    std::vector<std::vector<uint64_t>> det_traj;
    std::vector<double> bin_weight, bin_target;  // ndet * T
    std::vector<uint64_t> bin_count;
    uint64_t scatter = 0, boundary = 0, roulette = 0, draws = 0, photons = 0;
    double w_abs = 0, w_esc = 0, w_det = 0, w_area = 0, w_kill = 0, w_rlost = 0, w_rgain = 0, w_other = 0;
    void init(int grid, int ndet, int T, bool abs, bool scat) {
        if (abs) absorption.assign(grid, 0.0);
        if (scat) {
            scattering.assign(grid, 0.0);
            det_traj.assign(ndet, std::vector<uint64_t>(grid, 0));
        }
        det_weight.assign(ndet, 0.0);
        det_target.assign(ndet, 0.0);
        det_count.assign(ndet, 0);
        bin_weight.assign((size_t)ndet * T, 0.0);
        bin_target.assign((size_t)ndet * T, 0.0);
        bin_count.assign((size_t)ndet * T, 0);
    }
};

template <class R> struct Photon {
    V3<R> pos, dir;
    R step, weight, time, target_range, other_range;
    int layer;
    bool visited[XMC_MAX_LAYERS];
    bool dead;
};

struct Traj {  // PhotonTrajectory, grown on demand
    std::vector<uint8_t> x, y, z;
    void clear() { x.clear(); y.clear(); z.clear(); }
};

template <class R> struct Walker {
    const Scene<R>& sc;
    Tally& tl;
    bool scat;
    Traj traj;
    xmc_trace_event* ev = nullptr;
    int max_ev = 0, n_ev = 0;

    Walker(const Scene<R>& s, Tally& t, bool sc_) : sc(s), tl(t), scat(sc_) {}

    void record(int type, const Photon<R>& p, int aux, const Mcg59& rng) {
        if (ev) {
            if (n_ev < max_ev) {
                xmc_trace_event& e = ev[n_ev];
                e.type = type;
                e.layer = p.layer;
                e.aux = aux;
                e.draws = rng.draws;
                e.pos[0] = p.pos.x; e.pos[1] = p.pos.y; e.pos[2] = p.pos.z;
                e.dir[0] = p.dir.x; e.dir[1] = p.dir.y; e.dir[2] = p.dir.z;
                e.weight = p.weight;
                e.time = p.time;
            }
            n_ev++;
        }
    }

    // UpdatePhotonTrajectory (:750-792).  Double: the reference's arithmetic
    // on the raw Area.  Float: the same steps in float on the float-rounded
    // Area, as the GPU kernel performs them.
    void update_trajectory(const Photon<R>& p, V3<R> prev) {
        if (!scat) return;
        if constexpr (sizeof(R) == 4) {
            V3<float> a, b;
            if (p.pos.z > prev.z) { a = prev; b = p.pos; }
            else if (p.pos.z < prev.z) { a = p.pos; b = prev; }
            else return;
            const float step = sc.area_length.z / sc.area_part[2];
            int plane_index = (int)((a.z - sc.area_corner.z + step / 2.0f) / step);
            float plane = sc.area_corner.z + plane_index * step + step / 2.0f;
            const float top = sc.area_corner.z + sc.area_length.z;
            while ((plane <= b.z) && (plane < top)) {
                V3<float> q = plane_segment_point_f(a, b, plane);
                int v[3];
                int idx = sc.area_index(q, v);
                if (idx < 0) break;
                traj.x.push_back((uint8_t)v[0]);
                traj.y.push_back((uint8_t)v[1]);
                traj.z.push_back((uint8_t)v[2]);
                tl.scattering[idx] += (double)p.weight;
                plane += step;
            }
            return;
        }
        V3<double> cur = {(double)p.pos.x, (double)p.pos.y, (double)p.pos.z};
        V3<double> pv = {(double)prev.x, (double)prev.y, (double)prev.z};
        V3<double> a, b;
        if (cur.z > pv.z) { a = pv; b = cur; }
        else if (cur.z < pv.z) { a = cur; b = pv; }
        else return;
        const xmc_area* ar = sc.in->area;
        double step = ar->length.z / ar->partitionNumber.z;
        int plane_index = (int)((a.z - ar->corner.z + step / 2.0) / step);
        double plane = ar->corner.z + plane_index * step + step / 2.0;
        while ((plane <= b.z) && (plane < ar->corner.z + ar->length.z)) {
            V3<double> q = plane_segment_point(a, b, plane);
            int v[3];
            int idx = area_index_d(q, v);
            if ((v[0] == 255 && v[1] == 255 && v[2] == 255) || idx == -1) break;
            traj.x.push_back((uint8_t)v[0]);
            traj.y.push_back((uint8_t)v[1]);
            traj.z.push_back((uint8_t)v[2]);
            tl.scattering[idx] += (double)p.weight;
            plane += step;
        }
    }

    int area_index_d(const V3<double>& p, int v[3]) const {
        const xmc_area* a = sc.in->area;
        double ix = a->partitionNumber.x * (p.x - a->corner.x) / a->length.x;
        double iy = a->partitionNumber.y * (p.y - a->corner.y) / a->length.y;
        double iz = a->partitionNumber.z * (p.z - a->corner.z) / a->length.z;
        bool inside = !((ix < 0.0 || ix >= a->partitionNumber.x) || (iy < 0.0 || iy >= a->partitionNumber.y) ||
                        (iz < 0.0 || iz >= a->partitionNumber.z));
        uint8_t bx = 255, by = 255, bz = 255;
        if (inside) { bx = (uint8_t)ix; by = (uint8_t)iy; bz = (uint8_t)iz; }
        v[0] = bx; v[1] = by; v[2] = bz;
        if (bx == 255 && by == 255 && bz == 255) return -1;
        return bx * a->partitionNumber.y * a->partitionNumber.z + by * a->partitionNumber.z + bz;
    }

    // MovePhotonAndUpdateTrajectory (:158-174); `track` = false leaves the
    // trajectory update to the caller
    void move(Photon<R>& p, bool track = true) {
        V3<R> prev = p.pos;
        p.time += p.step * sc.layers[p.layer].n;
        if (sc.in->targetRangeLayers[p.layer]) p.target_range += p.step;
        else p.other_range += p.step;
        p.pos.x += p.step * p.dir.x;
        p.pos.y += p.step * p.dir.y;
        p.pos.z += p.step * p.dir.z;
        if (track) update_trajectory(p, prev);
    }

    // MovePhotonOnMinDistanceAndUpdateTrajectory (:176-188) in double; in
    // float the photon is placed delta off the hit point `hit` on the side
    // its (new) direction points to (see Prec<float>).
    void nudge(Photon<R>& p, const V3<R>& n, const V3<R>& hit) {
        if constexpr (sizeof(R) == 8) {
            V3<R> prev = p.pos;
            const R md = Prec<R>::kMinDistance;
            p.pos.x += md * p.dir.x;
            p.pos.y += md * p.dir.y;
            p.pos.z += md * p.dir.z;
            update_trajectory(p, prev);
        } else {
            R d = offset_scale(hit);
            if (dot(p.dir, n) < (R)0) d = -d;
            V3<R> prev = p.pos;  // the incident-side point the hit move ended at
            p.pos.x = hit.x + d * n.x;
            p.pos.y = hit.y + d * n.y;
            p.pos.z = hit.z + d * n.z;
            update_trajectory(p, prev);
        }
    }

    static R offset_scale(const V3<R>& p) {
        R m = std::fmax(std::fabs(p.x), std::fmax(std::fabs(p.y), std::fabs(p.z)));
        return std::fmaf(m, 0x1p-20f, 0x1p-22f);
    }

    int detector_id(const Photon<R>& p) const {  // :663-742
        int nc = (int)sc.cubes.size();
        for (int i = 0; i < nc; ++i) {
            const auto& c = sc.cubes[i];
            V3<R> h = {c.length.x / (R)2.0, c.length.y / (R)2.0, c.length.z / (R)2.0};
            bool in = (p.pos.x >= (c.center.x - h.x)) && (p.pos.x < (c.center.x + h.x)) &&
                      (p.pos.y >= (c.center.y - h.y)) && (p.pos.y < (c.center.y + h.y)) &&
                      (p.pos.z >= (c.center.z - h.z)) && (p.pos.z < (c.center.z + h.z));
            if (in && visited_filter(p, c.filter)) {
                R len = p.dir.x * p.dir.x + p.dir.y * p.dir.y + p.dir.z * p.dir.z;
                R ac = (-p.dir.z) / len;
                if (ac > c.cos_target) return i;
            }
        }
        for (int i = 0; i < (int)sc.rings.size(); ++i) {
            const auto& r = sc.rings[i];
            R dist = std::sqrt(((p.pos.x - r.center.x) * (p.pos.x - r.center.x)) +
                               ((p.pos.y - r.center.y) * (p.pos.y - r.center.y)));
            bool in = (dist >= r.rs) && (dist < r.rb);
            if (in && visited_filter(p, r.filter)) {
                R len = p.dir.x * p.dir.x + p.dir.y * p.dir.y + p.dir.z * p.dir.z;
                R ac = (-p.dir.z) / len;
                if (ac > r.cos_target) return nc + i;
            }
        }
        return -1;
    }

    static bool visited_filter(const Photon<R>& p, const uint8_t* f) {
        for (int j = 0; j < XMC_MAX_LAYERS; ++j)
            if (p.visited[j] && f[j]) return true;
        return false;
    }

    void detect(const Photon<R>& p, int id) {  // :744-858
        double w = (double)p.weight;
        tl.det_weight[id] += w;
        if (scat) {
            const xmc_area* a = sc.in->area;
            for (size_t i = 0; i < traj.x.size(); ++i) {
                int idx = traj.x[i] * a->partitionNumber.y * a->partitionNumber.z +
                          traj.y[i] * a->partitionNumber.z + traj.z[i];
                tl.det_traj[id][idx]++;
            }
        }
        tl.det_count[id]++;
        int T = sc.in->timeScaleSize;
        // ranges and weights in precision R, summed in double
        R total = p.target_range + p.other_range;
        R tw = p.target_range * (p.weight / total);
        double t = (double)p.time;
        size_t base = (size_t)id * T;
        bool binned = false;
        for (int i = 0; i < T - 1; ++i) {
            if ((t >= sc.t_start[i]) && (t < sc.t_finish[i])) {
                tl.bin_target[base + i] += (double)tw;
                tl.bin_count[base + i]++;
                tl.bin_weight[base + i] += w;
                binned = true;
                break;
            }
        }
        if (!binned && T > 0 && t >= sc.t_start[T - 1]) {
            tl.bin_target[base + T - 1] += (double)tw;
            tl.bin_count[base + T - 1]++;
            tl.bin_weight[base + T - 1] += w;
        }
        tl.det_target[id] += (double)tw;
    }

    // CrossBoundary (:229-284)
    void cross(Photon<R>& p, const Hit<R>& h, const V3<R>& hit, Mcg59& rng) {
        R ci = dot(p.dir, h.normal);
        if (ci < (R)0.0) ci = -ci;
        int cur = p.layer;
        int nxt = sc.find_layer(h.surface, cur);
        if (nxt < 0) {  // surface not shared with another layer: undefined in the reference
            p.dead = true;
            tl.w_other += (double)p.weight;
            return;
        }
        R n1 = sc.layers[cur].n, n2 = sc.layers[nxt].n;
        R cc = critical_cos(n1, n2);
        R refl;
        if (ci <= cc) refl = (R)1.0;
        else {
            R ct = transmit_cos(n1, n2, ci);
            refl = fresnel(n1, n2, ci, ct);
        }
        tl.boundary++;
        if (Prec<R>::uniform(rng.next()) > refl) {
            if (nxt == 0) {
                int id = detector_id(p);
                if (id >= 0) {
                    detect(p, id);
                    tl.w_det += (double)p.weight;
                }
                tl.w_esc += (double)p.weight;
                p.dead = true;
                record(XMC_EV_ESCAPE, p, id, rng);
            } else {
                p.dir = refract(n1, n2, p.dir, h.normal);
                p.layer = nxt;
                p.visited[nxt] = true;
                nudge(p, h.normal, hit);
                record(XMC_EV_TRANSMIT, p, h.surface, rng);
            }
        } else {
            p.dir = reflect(p.dir, h.normal);
            nudge(p, h.normal, hit);
            record(XMC_EV_REFLECT, p, h.surface, rng);
        }
    }

    // ComputePhoton (:39-126) with the termination policy
    void photon(Mcg59& rng) {
        Photon<R> p;
        const xmc_input* in = sc.in;
        traj.clear();
        p.weight = sc.w0;  // LaunchPhoton (:128-149)
        p.dead = false;
        p.layer = 1;
        p.step = 0;
        p.time = 0;
        p.target_range = 0;
        p.other_range = 0;
        p.pos = sc.start_pos;
        if constexpr (sizeof(R) == 8) p.pos.z = in->startPosition.z + 1.0e-8;
        else p.pos.z = sc.start_pos.z + offset_scale(sc.start_pos);
        p.dir = sc.start_dir;
        for (int i = 0; i < XMC_MAX_LAYERS; ++i) p.visited[i] = false;
        p.visited[p.layer] = true;
        record(XMC_EV_LAUNCH, p, 0, rng);
        while (!p.dead) interact(p, rng);
        tl.draws += rng.draws;
        tl.photons++;
    }

    // One iteration of ComputePhoton's loop (:47-124) plus termination.
    void interact(Photon<R>& p, Mcg59& rng) {
        const xmc_input* in = sc.in;
        const R step_unlimited = Prec<R>::kMaxDistance;

        if (rng.draws > kDrawLimit) {  // stream block exhausted: retire the photon
            p.dead = true;
            tl.w_other += (double)p.weight;
            return;
        }
        const auto& L = sc.layers[p.layer];
        Hit<R> h;
        // the nearest surface: the BVH walk, or in the 32-bit mode of a scene of
        // rectangles and lattice heightfields the engine's surface rule
        auto nearest = [&](R step, bool limited) {
            if constexpr (sizeof(R) == 4) {
                if (sc.rule) return sc.rule_hit(p.pos, p.dir, step, limited);
            }
            return sc.intersect(p.pos, p.dir, step, limited);
        };
        if (L.glass) {
            h = nearest(step_unlimited, false);
        } else {
            p.step = step_size(L.mua, L.mus, rng);
            h = nearest(p.step, true);
        }
        int v[3];
        if (h.found) {
            p.step = h.distance - (R)0.01 * Prec<R>::kMinDistance;
            V3<R> start = p.pos;
            move(p, sizeof(R) == 8);
            V3<R> hit = p.pos;
            if constexpr (sizeof(R) == 4) {
                // the float walk stops delta short of the surface on the incident
                // side (the double walk stops 1e-10 short along the direction);
                // its path, and so its plane crossings, ends there
                R d = offset_scale(hit);
                if (dot(p.dir, h.normal) < (R)0) d = -d;
                p.pos.x = hit.x - d * h.normal.x;
                p.pos.y = hit.y - d * h.normal.y;
                p.pos.z = hit.z - d * h.normal.z;
                update_trajectory(p, start);
            }
            if (sc.area_index(p.pos, v) >= 0) {
                cross(p, h, hit, rng);
            } else {
                p.dead = true;
                tl.w_area += (double)p.weight;
                record(XMC_EV_DIE_AREA, p, 0, rng);
            }
        } else if (L.glass) {
            p.dead = true;
            tl.w_other += (double)p.weight;
            record(XMC_EV_GLASS_MISS, p, 0, rng);
        } else {
            move(p);
            int idx = sc.area_index(p.pos, v);
            if (idx >= 0) {
                // :633-642; float: the layer's absorbed fraction mua / (mua + mus) once per layer
                R dw = sizeof(R) == 4 ? p.weight * (L.mua / (L.mua + L.mus)) : p.weight * L.mua / (L.mua + L.mus);
                p.weight -= dw;
                if (!tl.absorption.empty()) tl.absorption[idx] += (double)dw;
                tl.w_abs += (double)dw;
                if (sc.biasing) {
                    // ComputePhotonDirectionWithBiasing (:398-418)
                    V3<R> a = sc.attractive(p.pos, p.layer);
                    V3<R> nd;
                    R pr;
                    do {
                        nd = new_direction(p.dir, L.g, rng);
                        pr = (R)0.1 + (R)0.9 * (((R)1 + dot(nd, a)) / (R)2);
                        pr = Prec<R>::pow_(pr, sc.factor);
                    } while (pr <= Prec<R>::uniform(rng.next()) && rng.draws <= kDrawLimit);
                    p.weight = sc.biased_weight(p.weight, p.dir, a, pr, p.layer);
                    p.dir = nd;
                } else {
                    p.dir = new_direction(p.dir, L.g, rng);
                }
                tl.scatter++;
                record(XMC_EV_SCATTER, p, idx, rng);
            } else {
                p.dead = true;
                tl.w_area += (double)p.weight;
                record(XMC_EV_DIE_AREA, p, 0, rng);
            }
        }
        if (in->terminationMode == XMC_TERMINATE_ROULETTE) {
            const R before = p.weight;
            const int r = roulette_step(p.weight, p.dead, sc.min_weight, sc.chance, rng);
            if (r != kRouletteSkip) tl.roulette++;
            if (r == kRouletteLive) {
                tl.w_rgain += (double)p.weight - (double)before;
                record(XMC_EV_ROULETTE_LIVE, p, 0, rng);
            } else if (r == kRouletteDie) {
                tl.w_rlost += (double)p.weight;
                record(XMC_EV_ROULETTE_DIE, p, 0, rng);
            } else if (r == kRouletteZero) {
                record(XMC_EV_ROULETTE_DIE, p, 0, rng);
            }
        } else if (p.weight < sc.min_weight) {
            if (!p.dead) {
                tl.w_kill += (double)p.weight;
                record(XMC_EV_DIE_WEIGHT, p, 0, rng);
            }
            p.dead = true;
        }
        
    }
};

Mcg59 stream_for(uint64_t seed, uint64_t photon) {
    Mcg59 g;
    g.value = photon_stream(seed, photon);
    g.offset = kMcgC;
    g.draws = 0;
    return g;
}

void add_into(xmc_output* out, const Tally& t, int T, bool scat) {
    if (out->absorption && !t.absorption.empty())
        for (int j = 0; j < out->gridSize; ++j) out->absorption[j] += t.absorption[j];
    if (scat && out->scattering && !t.scattering.empty())
        for (int j = 0; j < out->gridSize; ++j) out->scattering[j] += t.scattering[j];
    for (int d = 0; d < out->numberOfDetectors; ++d) {
        xmc_detector_info& di = out->detectorInfo[d];
        di.weight += t.det_weight[d];
        di.targetWeight += t.det_target[d];
        di.numberOfPhotons += t.det_count[d];
        if (scat && di.trajectory && !t.det_traj.empty())
            for (int j = 0; j < out->gridSize; ++j) di.trajectory[j] += t.det_traj[d][j];
        for (int b = 0; b < T && b < di.timeScaleSize; ++b) {
            size_t k = (size_t)d * T + b;
            di.timeScale[b].weight += t.bin_weight[k];
            di.timeScale[b].targetWeight += t.bin_target[k];
            di.timeScale[b].numberOfPhotons += t.bin_count[k];
        }
    }
}

template <class R>
int run_impl(const xmc_input* in, xmc_output* out, uint64_t seed, uint64_t first, uint64_t n,
             int nthreads, bool scat, xmco_counters* c) {
    Scene<R> sc;
    sc.init(in);
    int T = in->timeScaleSize;
    int ndet = in->numberOfCubeDetectors + in->numberOfRingDetectors;
    if (out->numberOfDetectors != ndet) return XMC_ERR_ARG;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    std::vector<Tally> tallies(nthreads);
    for (auto& t : tallies) t.init(out->gridSize, ndet, T, out->absorption != nullptr, scat);
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(nthreads)
    {
        int tid = omp_get_thread_num();
        Walker<R> w(sc, tallies[tid], scat);
#pragma omp for schedule(dynamic, 16)
        for (long long i = 0; i < (long long)n; ++i) {
            Mcg59 g = stream_for(seed, first + (uint64_t)i);
            w.photon(g);
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    out->specularReflectance = sc.rsp;
    for (auto& t : tallies) add_into(out, t, T, scat);
    if (c) {
        std::memset(c, 0, sizeof(*c));
        for (auto& t : tallies) {
            c->photons += t.photons;
            c->scatterEvents += t.scatter;
            c->boundaryEvents += t.boundary;
            c->rouletteEvents += t.roulette;
            c->draws += t.draws;
            c->wAbsorbed += t.w_abs;
            c->wEscaped += t.w_esc;
            c->wDetected += t.w_det;
            c->wLeftArea += t.w_area;
            c->wKilled += t.w_kill;
            c->wRouletteLost += t.w_rlost;
            c->wRouletteGain += t.w_rgain;
            c->wOther += t.w_other;
        }
        c->seconds = std::chrono::duration<double>(t1 - t0).count();
        c->threads = nthreads;
    }
    return XMC_OK;
}

template <class R>
int trace_impl(const xmc_input* in, uint64_t seed, uint64_t photon, xmc_trace_event* ev, int max_ev) {
    Scene<R> sc;
    sc.init(in);
    int ndet = in->numberOfCubeDetectors + in->numberOfRingDetectors;
    Tally t;
    t.init(in->area->partitionNumber.x * in->area->partitionNumber.y * in->area->partitionNumber.z, ndet,
           in->timeScaleSize, false, false);
    Walker<R> w(sc, t, false);
    w.ev = ev;
    w.max_ev = max_ev;
    Mcg59 g = stream_for(seed, photon);
    w.photon(g);
    return w.n_ev <= max_ev ? w.n_ev : -w.n_ev;
}

// One interaction from each recorded state of a trace (a float walk of
// `photon`): position, direction, weight, time and layer from the event,
// visited layers accumulated along the trace, random stream advanced by the
// recorded number of draws.  out[2i..2i+1] receive the events of the
// iteration started from trace[i], counts[i] their number.
template <class R>
int shadow_impl(const xmc_input* in, uint64_t seed, uint64_t photon, const xmc_trace_event* trace, int n,
                xmc_trace_event* out, int* counts) {
    Scene<R> sc;
    sc.init(in);
    int ndet = in->numberOfCubeDetectors + in->numberOfRingDetectors;
    Tally t;
    t.init(1, ndet, in->timeScaleSize, false, false);
    Walker<R> w(sc, t, false);
    uint32_t visited = 1u << 1;
    Mcg59 g = stream_for(seed, photon);
    uint32_t drawn = 0;
    for (int i = 0; i < n; ++i) {
        const xmc_trace_event* st = trace + i;
        visited |= 1u << st->layer;
        while (drawn < st->draws) {
            g.next();
            drawn++;
        }
        Mcg59 gi = g;
        gi.draws = st->draws;
        Photon<R> p;
        p.pos = {(R)st->pos[0], (R)st->pos[1], (R)st->pos[2]};
        p.dir = {(R)st->dir[0], (R)st->dir[1], (R)st->dir[2]};
        // the recorded float direction is unit only to float rounding; the
        // double formulas (sqrt(1 - uz^2) in the spin) need it unit to double
        // rounding, so replay from the normalised direction
        if constexpr (sizeof(R) == 8) p.dir = normalize(p.dir);
        p.weight = (R)st->weight;
        p.time = (R)st->time;
        p.step = 0;
        p.target_range = 0;
        p.other_range = 0;
        p.layer = st->layer;
        p.dead = st->type == XMC_EV_ESCAPE || st->type == XMC_EV_DIE_AREA || st->type == XMC_EV_DIE_WEIGHT ||
                 st->type == XMC_EV_ROULETTE_DIE || st->type == XMC_EV_GLASS_MISS;
        for (int k = 0; k < XMC_MAX_LAYERS; ++k) p.visited[k] = (visited >> k) & 1u;
        w.ev = out + 2 * i;
        w.max_ev = 2;
        w.n_ev = 0;
        if (!p.dead) w.interact(p, gi);
        counts[i] = w.n_ev;
    }
    return XMC_OK;
}

template <class R>
int query_impl(const xmc_input* in, const double* o, const double* d, const double* steps, int nr,
               int mode, int* found, double* dist, int* surf, int* tri, double* normal) {
    Scene<R> sc;
    sc.init(in);
    if (mode < 0 || mode > 2 || (mode == 2 && (sizeof(R) != 4 || !sc.rule))) return XMC_ERR_ARG;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nr; ++i) {
        V3<R> oo = {(R)o[3 * i], (R)o[3 * i + 1], (R)o[3 * i + 2]};
        V3<R> dd = {(R)d[3 * i], (R)d[3 * i + 1], (R)d[3 * i + 2]};
        bool limited = steps[i] > 0.0;
        R st = limited ? (R)steps[i] : Prec<R>::kMaxDistance;
        Hit<R> h = mode == 1 ? sc.brute(oo, dd, st, limited) : sc.intersect(oo, dd, st, limited);
        if constexpr (sizeof(R) == 4) {
            if (mode == 2) h = sc.rule_hit(oo, dd, st, limited);
        }
        found[i] = h.found;
        dist[i] = h.found ? (double)h.distance : 0.0;
        surf[i] = h.found ? h.surface : -1;
        tri[i] = h.found ? sc.bvh.tris[h.tri].index : -1;
        normal[3 * i] = h.found ? (double)h.normal.x : 0.0;
        normal[3 * i + 1] = h.found ? (double)h.normal.y : 0.0;
        normal[3 * i + 2] = h.found ? (double)h.normal.z : 0.0;
    }
    return XMC_OK;
}

V3<double> d3(const double* p) { return {p[0], p[1], p[2]}; }

}  // namespace

extern "C" {

uint64_t xmco_photon_stream(uint64_t seed, uint64_t photon) { return photon_stream(seed, photon); }

int xmco_run(const xmc_input* input, xmc_output* output, uint64_t seed, uint64_t first, uint64_t n,
             int precision, int nthreads, int scattering, xmco_counters* counters) {
    if (!input || !output) return XMC_ERR_ARG;
    g_uniform24 = precision == 65;
    if (precision == 32)
        return run_impl<float>(input, output, seed, first, n, nthreads, scattering != 0, counters);
    return run_impl<double>(input, output, seed, first, n, nthreads, scattering != 0, counters);
}

int xmco_trace(const xmc_input* input, uint64_t seed, uint64_t photon, int precision,
               xmc_trace_event* events, int maxEvents) {
    g_uniform24 = precision == 65;
    if (precision == 32) return trace_impl<float>(input, seed, photon, events, maxEvents);
    return trace_impl<double>(input, seed, photon, events, maxEvents);
}

int xmco_shadow_trace(const xmc_input* input, uint64_t seed, uint64_t photon, int precision,
                      const xmc_trace_event* trace, int n, xmc_trace_event* out, int* counts) {
    g_uniform24 = precision == 65;
    if (precision == 32) return shadow_impl<float>(input, seed, photon, trace, n, out, counts);
    return shadow_impl<double>(input, seed, photon, trace, n, out, counts);
}

// The walk's roulette step on n photons: photon i has weight weights[i] and is
// dead when dead[i]; it draws from stream i of `seed` (precision 32: float
// weights and 24-bit uniforms; 64: double and 59-bit).  Returns the weight
// after the step in out_weight (0 for a photon that died), the outcome
// (kRoulette*) and the uniforms drawn.
int xmco_roulette_many(int precision, const double* weights, const uint8_t* dead, double min_weight, double chance,
                       uint64_t seed, int64_t n, double* out_weight, int* outcome, int* draws) {
    g_uniform24 = false;
    for (int64_t i = 0; i < n; ++i) {
        Mcg59 g = stream_for(seed, (uint64_t)i);
        bool d = dead[i] != 0;
        int r;
        if (precision == 32) {
            float w = (float)weights[i];
            r = roulette_step<float>(w, d, (float)min_weight, (float)chance, g);
            out_weight[i] = d ? 0.0 : (double)w;
        } else {
            double w = weights[i];
            r = roulette_step<double>(w, d, min_weight, chance, g);
            out_weight[i] = d ? 0.0 : w;
        }
        outcome[i] = r;
        draws[i] = (int)g.draws;
    }
    return XMC_OK;
}

int xmco_bvh(const xmc_input* input, int* nn, int* root, int* nt, double* boxes, int* links,
             int* tsurf, int* tidx) {
    Bvh b;
    build_bvh(b, input->surface, input->numberOfSurfaces);
    *nn = (int)b.nodes.size();
    *root = b.root;
    *nt = (int)b.tris.size();
    if (boxes) {
        for (size_t i = 0; i < b.nodes.size(); i++) {
            const Node& n = b.nodes[i];
            double* bx = boxes + 6 * i;
            bx[0] = n.box.lo.x; bx[1] = n.box.lo.y; bx[2] = n.box.lo.z;
            bx[3] = n.box.hi.x; bx[4] = n.box.hi.y; bx[5] = n.box.hi.z;
            links[4 * i] = n.left;
            links[4 * i + 1] = n.right;
            links[4 * i + 2] = n.offset;
            links[4 * i + 3] = n.count;
        }
        for (size_t i = 0; i < b.tris.size(); i++) {
            tsurf[i] = b.tris[i].surf;
            tidx[i] = b.tris[i].index;
        }
    }
    return XMC_OK;
}

void xmco_set_lattice_rule(int on) { g_lattice_rule = on != 0; }

int xmco_query(const xmc_input* input, const double* o, const double* d, const double* steps, int nr,
               int precision, int mode, int* found, double* dist, int* surf, int* tri, double* normal) {
    if (precision == 32) return query_impl<float>(input, o, d, steps, nr, mode, found, dist, surf, tri, normal);
    return query_impl<double>(input, o, d, steps, nr, mode, found, dist, surf, tri, normal);
}

uint64_t xmco_mcg59_init(uint64_t seed, uint32_t id, uint32_t step, uint64_t* offset) {
    uint64_t value = 2 * seed + 1;  // InitMCG59, mcml_mcg59.cpp:28-35
    value = (value * mcg_pow(kMcgC, id)) & kMcgMask;
    *offset = mcg_pow(kMcgC, step);
    return value;
}

double xmco_mcg59_next(uint64_t* value, uint64_t offset) {
    g_uniform24 = false;
    Mcg59 g{*value, offset, 0};
    double u = Prec<double>::uniform(g.next());
    *value = g.value;
    return u;
}

double xmco_specular(const xmc_layer* layer) {  // :16-37
    double t = (layer[0].refractiveIndex - layer[1].refractiveIndex) /
               (layer[0].refractiveIndex + layer[1].refractiveIndex);
    double r1 = t * t;
    if (layer[1].absorptionCoefficient == 0.0 && layer[1].scatteringCoefficient == 0.0) {
        t = (layer[1].refractiveIndex - layer[2].refractiveIndex) /
            (layer[1].refractiveIndex + layer[2].refractiveIndex);
        double r2 = t * t;
        t = 1 - r1;
        r1 = r1 + t * t * r2 / (1 - r1 * r2);
    }
    return r1;
}

double xmco_cosine_theta(double g, uint64_t* value, uint64_t offset) {
    Mcg59 r{*value, offset, 0};
    double c = cosine_theta<double>(g, r);
    *value = r.value;
    return c;
}

void xmco_new_direction(const double* dir, double g, uint64_t* value, uint64_t offset, double* out) {
    Mcg59 r{*value, offset, 0};
    V3<double> v = new_direction<double>(d3(dir), g, r);
    *value = r.value;
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

double xmco_transmit_cos(double n1, double n2, double ci) { return transmit_cos<double>(n1, n2, ci); }
double xmco_fresnel(double n1, double n2, double ci, double ct) { return fresnel<double>(n1, n2, ci, ct); }

int xmco_find_layer(const xmc_input* input, int surfaceId, int layer) {
    for (int i = 0; i < input->numberOfLayers; ++i)
        for (int j = 0; j < input->layerInfo[i].numberOfSurfaces; ++j)
            if (input->layerInfo[i].surfaceId[j] == surfaceId && i != layer) return i;
    return -1;
}

void xmco_reflect(const double* i, const double* n, double* out) {
    V3<double> r = reflect<double>(d3(i), d3(n));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void xmco_refract(double n1, double n2, const double* i, const double* n, double* out) {
    V3<double> r = refract<double>(n1, n2, d3(i), d3(n));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

double xmco_step(double mua, double mus, uint64_t* value, uint64_t offset) {
    Mcg59 r{*value, offset, 0};
    double s = step_size<double>(mua, mus, r);
    *value = r.value;
    return s;
}

double xmco_critical_cos(double n1, double n2) { return critical_cos<double>(n1, n2); }

int xmco_area_index(const xmc_area* a, const double* p) {
    double ix = a->partitionNumber.x * (p[0] - a->corner.x) / a->length.x;
    double iy = a->partitionNumber.y * (p[1] - a->corner.y) / a->length.y;
    double iz = a->partitionNumber.z * (p[2] - a->corner.z) / a->length.z;
    bool inside = !((ix < 0.0 || ix >= a->partitionNumber.x) || (iy < 0.0 || iy >= a->partitionNumber.y) ||
                    (iz < 0.0 || iz >= a->partitionNumber.z));
    uint8_t bx = 255, by = 255, bz = 255;
    if (inside) { bx = (uint8_t)ix; by = (uint8_t)iy; bz = (uint8_t)iz; }
    if (bx == 255 && by == 255 && bz == 255) return -1;
    return bx * a->partitionNumber.y * a->partitionNumber.z + by * a->partitionNumber.z + bz;
}

int xmco_detector_id(const xmc_input* input, const double* pos, const double* dir, const uint8_t* visited) {
    Scene<double> sc;
    sc.in = input;
    for (int i = 0; i < input->numberOfCubeDetectors; i++) {
        const xmc_cube_detector& c = input->cubeDetector[i];
        Scene<double>::Cube q;
        q.center = cvt<double>(c.center);
        q.length = cvt<double>(c.length);
        std::memcpy(q.filter, c.filterLayers, XMC_MAX_LAYERS);
        q.cos_target = cos(c.targetAngle * kPi / 180.0);
        sc.cubes.push_back(q);
    }
    for (int i = 0; i < input->numberOfRingDetectors; i++) {
        const xmc_ring_detector& c = input->ringDetector[i];
        Scene<double>::Ring q;
        q.center = cvt<double>(c.center);
        q.rs = c.smallRadius;
        q.rb = c.bigRadius;
        std::memcpy(q.filter, c.filterLayers, XMC_MAX_LAYERS);
        q.cos_target = cos(c.targetAngle * kPi / 180.0);
        sc.rings.push_back(q);
    }
    Tally t;
    Walker<double> w(sc, t, false);
    Photon<double> p;
    p.pos = d3(pos);
    p.dir = d3(dir);
    for (int i = 0; i < XMC_MAX_LAYERS; i++) p.visited[i] = visited[i] != 0;
    return w.detector_id(p);
}

double xmco_triangle_distance(const double* o, const double* d, const double* a, const double* b,
                              const double* c) {
    return triangle_distance<double>(d3(o), d3(d), d3(a), d3(b), d3(c));
}

float xmco_logf(float x) { return f_logf(x); }
void xmco_expf_many(const float* x, float* y, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) y[i] = f_expf(x[i]);
}
double xmco_weight_integral(double dot, double anisotropy, double eps, unsigned n) {
    return weight_integral(dot, anisotropy, eps, n);
}
void xmco_sincos2pi(float u, float* s, float* c) { f_sincos2pi(u, s, c); }
void xmco_hg_float_many(float g, const float* u, float* c, float* s, int64_t n) {
    for (int64_t i = 0; i < n; i++) hg_float(g, u[i], c + i, s + i);
}
void xmco_logf_many(const float* x, float* y, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) y[i] = f_logf(x[i]);
}
void xmco_sincos2pi_many(const float* u, float* s, float* c, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) f_sincos2pi(u[i], s + i, c + i);
}

}  // extern "C"
