/*
 * xmc_oracle.h -- C API of the CPU oracle (test infrastructure only).
 *
 * The oracle is a CPU restatement of xmcml's photon random-walk path
 * (base/xmcml/xmcml/mcml_kernel.cpp, mcml_intersection.cpp, mcml_math.cpp,
 * mcml_mcg59.cpp, base/xmcml/xmcmlLauncher/launcher_omp.cpp).  It is used by
 * the tests as the checker and by bench.py as the timed CPU baseline.  The
 * product (xmcml_amd) never links or calls it.
 *
 * precision = 64: the reference's own double arithmetic, operation for
 *                 operation (libm log/sin/cos/sqrt), no FMA contraction.
 * precision = 32: the same algorithm in float, with the float-specific rules
 *                 the GPU kernel uses (own log / sincos(2*pi*u), 24-bit
 *                 uniforms, coordinate-scaled surface offsets, accurate
 *                 1-x^2).  This mode is bit-comparable with the GPU kernel.
 * precision = 65: the 64-bit arithmetic driven by the 24-bit uniforms the
 *                 32-bit mode draws ("the same recorded sequence").
 */
#ifndef XMC_ORACLE_H
#define XMC_ORACLE_H

#include <stdint.h>
#include "../include/xmcml.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xmco_counters {
    uint64_t photons;
    uint64_t scatterEvents;
    uint64_t boundaryEvents;
    uint64_t rouletteEvents;
    uint64_t draws;
    double seconds; /* wall time of the photon loop */
    int threads;
    /* weight ledger: where the launched weight went */
    double wAbsorbed;      /* dropped weight (absorption grid total) */
    double wEscaped;       /* weight of photons transmitted into layer 0 */
    double wDetected;      /* part of wEscaped that hit a detector */
    double wLeftArea;      /* weight of photons that left the Area */
    double wKilled;        /* weight of photons killed below minWeight */
    double wRouletteLost;  /* weight of photons lost in roulette */
    double wRouletteGain;  /* weight added to roulette survivors */
    double wOther;         /* glass misses, surfaces without a next layer */
} xmco_counters;

/* Per-photon stream: state before the first draw of photon p. */
uint64_t xmco_photon_stream(uint64_t seed, uint64_t photon);

/* LaunchOMP-equivalent over photons [first, first+n) with per-photon streams.
 * Adds into output (+=).  scattering != 0 also produces the scattering grid
 * and detector trajectory maps (the reference always does; the GPU never). */
int xmco_run(const xmc_input* input, xmc_output* output, uint64_t seed, uint64_t first,
             uint64_t n, int precision, int nthreads, int scattering, xmco_counters* counters);

/* Trace one photon; returns the number of events written (<= maxEvents) or
 * the negated total when it did not fit. */
int xmco_trace(const xmc_input* input, uint64_t seed, uint64_t photon, int precision,
               xmc_trace_event* events, int maxEvents);

/* One loop iteration of ComputePhoton from every recorded state of a trace
 * (e.g. the GPU trace of photon `photon`): position, direction, weight, time
 * and layer from trace[i], visited layers accumulated along the trace, random
 * stream advanced by trace[i].draws.  out[2i], out[2i+1] receive the events of
 * that iteration (interaction, then termination if any), counts[i] their
 * number (0 after a terminal event). */
int xmco_shadow_trace(const xmc_input* input, uint64_t seed, uint64_t photon, int precision,
                      const xmc_trace_event* trace, int n, xmc_trace_event* out, int* counts);

/* BVH (GenerateBVHTree restated): sizes first (NULL arrays), then contents. */
int xmco_bvh(const xmc_input* input, int* numberOfNodes, int* root, int* numberOfTriangles,
             double* boxes, int* links, int* triangleSurface, int* triangleIndex);

/* Nearest-hit queries. mode 0 = BVH walk (ComputeBVHIntersection when
 * step > 0, ComputeBVHIntersectionWithoutStep when step <= 0), mode 1 = brute
 * force over every triangle with the same triangle test and tie rule, mode 2
 * (precision 32, scenes under the surface rule only) = the engine's surface
 * rule for scenes whose every surface is an axis-aligned rectangle at one z or
 * a heightfield on a uniform lattice shared by all of them.
 * triangleId is the triangle's index inside its surface. */
/* Switches the 32-bit mode's lattice rule on (default) or off for the calls
 * that follow (process-wide): off, scenes with heightfields walk the BVH, as
 * the engine does with xmc_options.noLatticeRule; scenes of rectangles only
 * keep their rule either way. */
void xmco_set_lattice_rule(int on);

int xmco_query(const xmc_input* input, const double* origins, const double* directions,
               const double* steps, int numberOfRays, int precision, int mode, int* found,
               double* distance, int* surfaceId, int* triangleId, double* normal);

/* Single-function entry points used to check the golden values of the
 * reference's unit tests (double precision). */
uint64_t xmco_mcg59_init(uint64_t seed, uint32_t id, uint32_t step, uint64_t* offset);
double xmco_mcg59_next(uint64_t* value, uint64_t offset);
double xmco_specular(const xmc_layer* layers);
double xmco_cosine_theta(double g, uint64_t* value, uint64_t offset);
void xmco_new_direction(const double* dir, double g, uint64_t* value, uint64_t offset, double* out);
double xmco_transmit_cos(double n1, double n2, double ci);
double xmco_fresnel(double n1, double n2, double ci, double ct);
int xmco_find_layer(const xmc_input* input, int surfaceId, int layer);
void xmco_reflect(const double* i, const double* n, double* out);
void xmco_refract(double n1, double n2, const double* i, const double* n, double* out);
double xmco_step(double mua, double mus, uint64_t* value, uint64_t offset);
double xmco_critical_cos(double n1, double n2);
int xmco_area_index(const xmc_area* area, const double* pos);
int xmco_detector_id(const xmc_input* input, const double* pos, const double* dir,
                     const uint8_t* visited);
double xmco_triangle_distance(const double* o, const double* d, const double* a, const double* b,
                              const double* c);
/* float special functions of the 32-bit mode (and bulk forms for tests) */
float xmco_logf(float x);
void xmco_sincos2pi(float u, float* s, float* c);
void xmco_logf_many(const float* x, float* y, int64_t n);
/* Henyey-Greenstein cos/sin of the scattering angle as the 32-bit mode
 * computes them (1 - cos and 1 + cos without cancellation). */
void xmco_hg_float_many(float g, const float* u, float* c, float* s, int64_t n);
void xmco_sincos2pi_many(const float* u, float* s, float* c, int64_t n);
/* exp of the 32-bit mode (the acceptance power of the biasing, p^f = exp(f log p)) */
void xmco_expf_many(const float* x, float* y, int64_t n);
/* ComputeWeightIntegral (mcml_integration.cpp:7-82), double, restated */
double xmco_weight_integral(double dot, double anisotropy, double eps, unsigned n);
/* The walk's MCML roulette step on n photons (stream i of seed each); out_weight
 * is the weight after the step (0 when dead), outcome 0 skipped / 1 zero weight
 * died / 2 survived / 3 died, draws the uniforms used. */
int xmco_roulette_many(int precision, const double* weights, const uint8_t* dead, double min_weight,
                       double chance, uint64_t seed, int64_t n, double* out_weight, int* outcome, int* draws);

#ifdef __cplusplus
}
#endif

#endif
