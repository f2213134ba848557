/*
 * ref_tape.cpp -- makes the reference's own sources, compiled unmodified,
 * draw the GPU engine's random numbers and record every nearest-surface
 * query their photon loop makes.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links this.
 *
 * oracle/Makefile links this file with ref_driver.cpp and the reference
 * objects into oracle/_ref/libxmcml_ref_tape.so with
 *   -Wl,--wrap=<NextMCG59>, --wrap=<ComputeBVHIntersection>,
 *   --wrap=<ComputeBVHIntersectionWithoutStep>
 * (their mangled names).  Every call the reference's ComputePhoton
 * (base/xmcml/xmcml/mcml_kernel.cpp:39-126) makes to these functions, which
 * live in other translation units (mcml_mcg59.cpp:37-41,
 * mcml_intersection.cpp:354-584), then lands here:
 *
 *   NextMCG59      advances the generator exactly as the reference does and
 *                  returns the top 24 bits of the 59-bit state,
 *                  (x >> 35) * 2^-24 -- the uniform the GPU engine draws from
 *                  the same state -- instead of x / 2^59;
 *   ComputeBVH...  calls the reference's walk and, while a log is open,
 *                  appends (origin, direction, step, draws so far, found,
 *                  distance, surface id) of the query.
 *
 * The sequence of queries is the photon's walk: one query per interaction,
 * from the position and direction the previous interaction left.
 */
#include "mcml_kernel.h"
#include "mcml_mcg59.h"
#include "mcml_intersection.h"

#include <cstdint>
#include <cstring>

extern "C" {

// one nearest-surface query of the reference's photon loop
typedef struct xref_query_record {
    double origin[3];
    double direction[3];
    double step;       // -1 for an unlimited query (glass layer)
    double distance;   // valid when found
    int32_t found;
    int32_t surface;
    uint32_t draws;    // uniforms drawn by the photon before the query
    int32_t pad;
} xref_query_record;

double __real__Z9NextMCG59P7__MCG59(MCG59* g);
IntersectionInfo __real__Z22ComputeBVHIntersection9__double3S_dP9__BVHTreeP9__Surface(double3 o, double3 d,
                                                                                     double step, BVHTree* t,
                                                                                     Surface* s);
IntersectionInfo __real__Z33ComputeBVHIntersectionWithoutStep9__double3S_P9__BVHTreeP9__Surface(double3 o, double3 d,
                                                                                               BVHTree* t,
                                                                                               Surface* s);
}

namespace {

thread_local xref_query_record* g_log = nullptr;
thread_local int g_log_cap = 0;
thread_local int g_log_n = 0;
thread_local uint32_t g_draws = 0;

void log_query(const double3& o, const double3& d, double step, const IntersectionInfo& h) {
    if (!g_log) return;
    if (g_log_n < g_log_cap) {
        xref_query_record& r = g_log[g_log_n];
        r.origin[0] = o.x; r.origin[1] = o.y; r.origin[2] = o.z;
        r.direction[0] = d.x; r.direction[1] = d.y; r.direction[2] = d.z;
        r.step = step;
        r.found = h.isFindIntersection;
        r.distance = h.isFindIntersection ? h.distance : 0.0;
        r.surface = h.isFindIntersection ? h.surfaceId : -1;
        r.draws = g_draws;
        r.pad = 0;
    }
    g_log_n++;
}

}  // namespace

extern "C" {

double __wrap__Z9NextMCG59P7__MCG59(MCG59* g) {
    (void)__real__Z9NextMCG59P7__MCG59(g);  // x <- x * offset mod 2^59
    ++g_draws;
    return (double)(uint32_t)(g->value >> 35) * (1.0 / 16777216.0);
}

IntersectionInfo __wrap__Z22ComputeBVHIntersection9__double3S_dP9__BVHTreeP9__Surface(double3 o, double3 d,
                                                                                     double step, BVHTree* t,
                                                                                     Surface* s) {
    IntersectionInfo h = __real__Z22ComputeBVHIntersection9__double3S_dP9__BVHTreeP9__Surface(o, d, step, t, s);
    log_query(o, d, step, h);
    return h;
}

IntersectionInfo __wrap__Z33ComputeBVHIntersectionWithoutStep9__double3S_P9__BVHTreeP9__Surface(double3 o, double3 d,
                                                                                               BVHTree* t,
                                                                                               Surface* s) {
    IntersectionInfo h = __real__Z33ComputeBVHIntersectionWithoutStep9__double3S_P9__BVHTreeP9__Surface(o, d, t, s);
    log_query(o, d, -1.0, h);
    return h;
}

// Opens a query log on the calling thread (and resets its draw count).
void xref_tape_log_begin(xref_query_record* buffer, int capacity) {
    g_log = buffer;
    g_log_cap = capacity;
    g_log_n = 0;
    g_draws = 0;
}

// Closes the log; returns the number of queries made (may exceed capacity).
int xref_tape_log_end(uint32_t* draws) {
    int n = g_log_n;
    if (draws) *draws = g_draws;
    g_log = nullptr;
    g_log_cap = 0;
    g_log_n = 0;
    return n;
}

// Uniforms drawn on this thread since the last log_begin.
uint32_t xref_tape_draws(void) { return g_draws; }

}  // extern "C"
