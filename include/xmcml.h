/*
 * xmcml.h -- C-ABI of the MI355X photon random-walk engine.
 *
 * The structures below are plain-old-data mirrors of xmcml's InputInfo /
 * OutputInfo family (base/xmcml/xmcml/mcml_kernel_types.h, mcml_types.h) so
 * that a caller which fills InputInfo today can fill xmc_input field for field.
 * Every pointer is owned by the caller; the library copies what it needs to the
 * GPU.  No C++ or HIP types appear in any signature.
 *
 * The entry points replace the photon random-walk seam of xmcml:
 *   GenerateBVHTree()  (base/xmcml/xmcml/mcml_intersection.h:78,
 *                       called from base/xmcml/xmcmlLauncher/launcher_mpi.cpp:396)
 *       -> xmc_scene_create()   builds the same SAH BVH on the host, flattens it
 *                               and uploads geometry, layers and detectors.
 *   LaunchOMP()        (base/xmcml/xmcmlLauncher/launcher_omp.h:10,
 *                       called from launcher_mpi.cpp:369,374)
 *       -> xmc_launch()         runs numberOfPhotons photons on the GPU and adds
 *                               the tallies into the caller's xmc_output.
 *   InitOutput()/FreeOutput() (launcher_omp.h:8-9)
 *       -> xmc_init_output()/xmc_free_output().
 *   ComputeSpecularReflectance() (base/xmcml/xmcml/mcml_kernel.h:10)
 *       -> xmc_specular_reflectance().
 *
 * Attractive-vector biasing (InputInfo.useBiasing..targetPoint) runs on the
 * GPU; its weight-integral tables are computed there too (xmc_weight_table_compute).
 *
 * Differences from the reference seam, all deliberate:
 *   - Functions return an int status (XMC_OK or a negative code) instead of
 *     void; xmc_last_error() gives the message.  LaunchOMP reports nothing.
 *   - Random streams are per photon, not per thread: photon p of a run with
 *     seed s draws the MCG59 sequence that starts at (2s+1) * a^(p * 2^24)
 *     mod 2^59 (a = 13^13).  The result therefore does not depend on how
 *     photons are scheduled.  A seed has 2^33 such blocks (the generator's
 *     period is 2^57), so photon indices of one seed stay below 2^33
 *     (8.6e9); runs that would pass it fail with XMC_ERR_LIMIT.  xmc_rng carries (seed, next photon) where
 *     LaunchOMP carries MCG59[numThreads]; both are advanced in place.
 *   - Termination can follow xmcml (die when weight < minWeight,
 *     mcml_kernel.cpp:121-124) or MCML Russian roulette
 *     (stable/plane_parallel/1.0/McmlGpu/kernel/kernel.cu:214-222).
 *   - The scattering grid and the per-detector voxel trajectory maps
 *     (mcml_kernel.cpp:750-810) cost memory and time, so a scene produces them
 *     only when xmc_options asks (xmc_launch: when the output has the arrays);
 *     otherwise their output arrays are left untouched.
 */
#ifndef XMCML_AMD_XMCML_H
#define XMCML_AMD_XMCML_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XMC_MAX_LAYERS 8 /* MAX_LAYERS, mcml_kernel_types.h:7 */

/* Status codes. */
#define XMC_OK 0
#define XMC_ERR_ARG (-1)     /* invalid argument / input */
#define XMC_ERR_HIP (-2)     /* HIP runtime failure */
#define XMC_ERR_LIMIT (-3)   /* input exceeds a documented limit */
#define XMC_ERR_NOGPU (-4)   /* no usable GPU */

/* double3 / int3, mcml_types.h:8-24 */
typedef struct xmc_double3 { double x, y, z; } xmc_double3;
typedef struct xmc_int3 { int x, y, z; } xmc_int3;

/* Area, mcml_types.h:26-31: voxel grid of the absorption tally.  A photon
 * whose position leaves the Area dies (mcml_kernel.cpp:63-71,90-117). */
typedef struct xmc_area {
    xmc_double3 corner;
    xmc_double3 length;
    xmc_int3 partitionNumber;
} xmc_area;

/* Surface, mcml_types.h:33-39: a triangle mesh. */
typedef struct xmc_surface {
    const xmc_double3* vertices;
    int numberOfVertices;
    const xmc_int3* triangles;
    int numberOfTriangles;
} xmc_surface;

/* LayerInfo, mcml_kernel_types.h:54-62.  Layer 0 is the ambient medium; a
 * layer with absorption == scattering == 0 is "glass". */
typedef struct xmc_layer {
    double refractiveIndex;
    double absorptionCoefficient;
    double scatteringCoefficient;
    double anisotropy;
    const int* surfaceId;
    int numberOfSurfaces;
} xmc_layer;

/* CubeDetector / RingDetector, mcml_kernel_types.h:19-34 (bool -> uint8_t). */
typedef struct xmc_cube_detector {
    xmc_double3 center;
    xmc_double3 length;
    uint8_t filterLayers[XMC_MAX_LAYERS];
    double targetAngle; /* degrees */
} xmc_cube_detector;

typedef struct xmc_ring_detector {
    xmc_double3 center;
    double smallRadius;
    double bigRadius;
    uint8_t filterLayers[XMC_MAX_LAYERS];
    double targetAngle; /* degrees */
} xmc_ring_detector;

#define XMC_TERMINATE_KILL 0     /* die when weight < minWeight (xmcml) */
#define XMC_TERMINATE_ROULETTE 1 /* MCML roulette: below minWeight survive with
                                    probability rouletteChance, weight /= chance */

/* WeightIntegralTable, mcml_kernel_types.h:47-52: for one anisotropy g,
 * elements[j] = the weight integral at dot = -1 + 2 j / (numberOfElements - 1)
 * (mcml_integration.cpp:7-82, weight_integral.cpp:33-64). */
typedef struct xmc_weight_table {
    double anisotropy;
    int numberOfElements;
    const double* elements;
} xmc_weight_table;

/* InputInfo, mcml_kernel_types.h:64-91.  The biasing fields follow the
 * others (appended, so that callers which zero the struct get no biasing). */
typedef struct xmc_input {
    uint64_t numberOfPhotons; /* photons for this call (launch) */
    const xmc_area* area;
    const xmc_layer* layerInfo;
    int numberOfLayers;
    const xmc_surface* surface;
    int numberOfSurfaces;
    const xmc_cube_detector* cubeDetector;
    int numberOfCubeDetectors;
    const xmc_ring_detector* ringDetector;
    int numberOfRingDetectors;
    double minWeight;
    double timeStart;  /* time unit: optical path n*s in mm (mcml_kernel.cpp:161) */
    double timeFinish;
    int timeScaleSize;
    xmc_double3 startPosition;
    xmc_double3 startDirection; /* used as given, not normalised (mcml_kernel.cpp:142-144) */
    uint8_t targetRangeLayers[XMC_MAX_LAYERS];
    int terminationMode; /* XMC_TERMINATE_* */
    double rouletteChance;
    /* Attractive-vector biasing (ComputePhotonDirectionWithBiasing,
     * mcml_kernel.cpp:325-440): each scattering draws directions until one is
     * accepted with probability (0.1 + 0.45 (1 + d.a))^attractiveFactor, a the
     * attractive vector towards targetPoint, and multiplies the weight by the
     * weight integral of the old direction over that probability.  The tables
     * (one per distinct anisotropy of layers 1.., GetDifferentAnisotropies,
     * weight_integral.cpp:67-90) are the caller's, or, when weightTable is
     * NULL, computed on the GPU by xmc_scene_create with weightIntegralPrecision
     * and weightTablePrecision (defaults 128 and 262144 when 0, parser.cpp:10-11). */
    int useBiasing;
    double attractiveFactor;
    xmc_double3 targetPoint;
    int weightIntegralPrecision;
    int weightTablePrecision;
    const xmc_weight_table* weightTable;
    int numberOfWeightTables;
} xmc_input;

/* TimeInfo, mcml_kernel_types.h:10-17. */
typedef struct xmc_time_info {
    double timeStart;
    double timeFinish;
    uint64_t numberOfPhotons;
    double weight;
    double targetWeight;
} xmc_time_info;

/* DetectorInfo, mcml_kernel_types.h:36-45. */
typedef struct xmc_detector_info {
    uint64_t numberOfPhotons;
    uint64_t* trajectory; /* trajectorySize counts per voxel, or NULL (map not tallied) */
    int trajectorySize;
    xmc_time_info* timeScale;
    int timeScaleSize;
    double targetWeight;
    double weight;
} xmc_detector_info;

/* OutputInfo, mcml_kernel_types.h:93-102.  Detector index order: cube
 * detectors first, then ring detectors (mcml_kernel.cpp:663-742). */
typedef struct xmc_output {
    uint64_t numberOfPhotons;
    double specularReflectance;
    double* absorption; /* gridSize doubles; may be NULL (tally skipped) */
    double* scattering; /* gridSize doubles, or NULL (grid not tallied) */
    int gridSize;
    xmc_detector_info* detectorInfo;
    int numberOfDetectors;
} xmc_output;

/* Random stream state of a run (replaces MCG59[numThreads] of LaunchOMP). */
typedef struct xmc_rng {
    uint64_t seed;
    uint64_t nextPhoton;
} xmc_rng;

/* Event counters and weight ledger of the work since the last stats call. */
typedef struct xmc_stats {
    uint64_t photons;
    uint64_t scatterEvents;  /* hop + drop + spin */
    uint64_t boundaryEvents; /* Fresnel reflect / refract / escape */
    uint64_t rouletteEvents;
    uint64_t stackSpills;    /* BVH stack pushes that spilled beyond the LDS entries */
    uint64_t culledWalks;    /* step-limited walks skipped by the distance grid */
    double kernelMs;         /* device time of the transport kernels (HIP events) */
    /* where the launched weight went; with the specular part they sum to the
     * number of photons, up to float rounding */
    double wAbsorbed;
    double wEscaped;      /* transmitted into layer 0 */
    double wDetected;     /* part of wEscaped that hit a detector */
    double wLeftArea;
    double wKilled;       /* below minWeight, terminationMode KILL */
    double wRouletteLost;
    double wRouletteGain;
    double wOther;        /* glass misses, surfaces without a next layer */
    uint64_t trajectoryOverflows; /* detected photons whose trajectory exceeded trajectoryCapacity */
    uint64_t entryWalks;          /* walks started at a cell's entry node instead of the root */
} xmc_stats;

/* One recorded step of a traced photon (debug / parity).  Positions and
 * weights are stored as double even when the transport runs in float. */
#define XMC_EV_LAUNCH 0
#define XMC_EV_SCATTER 1
#define XMC_EV_REFLECT 2
#define XMC_EV_TRANSMIT 3
#define XMC_EV_ESCAPE 4       /* transmitted into layer 0; aux = detector id or -1 */
#define XMC_EV_DIE_AREA 5     /* left the Area */
#define XMC_EV_DIE_WEIGHT 6   /* weight < minWeight, terminationMode KILL */
#define XMC_EV_ROULETTE_LIVE 7
#define XMC_EV_ROULETTE_DIE 8
#define XMC_EV_GLASS_MISS 9   /* no surface ahead in a glass layer */
typedef struct xmc_trace_event {
    int32_t type;
    int32_t layer;
    int32_t aux;
    uint32_t draws; /* uniforms drawn by this photon so far */
    double pos[3];
    double dir[3];
    double weight;
    double time;
} xmc_trace_event;

/* Launch options (all zero = defaults). */
typedef struct xmc_options {
    int device;          /* HIP device ordinal */
    int blockSize;       /* threads per workgroup (64..256, multiple of 64); 0 = 256 */
    int blocksPerCU;     /* cap on persistent workgroups per CU; 0 = occupancy limit */
    int skipAbsorption;  /* 1 = do not tally the absorption grid */
    int noDistanceGrid;  /* 1 = always walk the BVH (no conservative distance culling) */
    int scattering;      /* 1 = tally the scattering grid (UpdatePhotonTrajectory, mcml_kernel.cpp:750-792) */
    int trajectories;    /* 1 = tally the per-detector voxel counts (UpdateDetectorTrajectory, :794-810) */
    int trajectoryCapacity; /* voxels kept per photon, 0 = 20480 (MAX_TRAJECTORY_SIZE, mcml_kernel_types.h:8);
                             * a detected photon with more crossings counts only the first ones and is
                             * reported in xmc_stats.trajectoryOverflows */
    int noEntryGrid;     /* 1 = step-limited walks start at the BVH root, not on a cell's entry tree (the entry
                          * trees of a 1.6M-triangle mesh take ~30 s of host time and ~5 GB of device
                          * memory to build; results are the same either way) */
    int noSlabGrid;      /* 1 = no separating slabs near the mesh (steps that cannot cross the nearby surfaces
                          * skip their walk; results are the same either way) */
    int noLatticeRule;   /* 1 = scenes of heightfields on a shared uniform lattice (and axis-aligned rectangles)
                          * walk the BVH instead of testing the triangles of the lattice cells each step covers
                          * (hits are the walk's but for float ties; scenes of rectangles only keep their rule).
                          * Batches always walk heightfields. */
} xmc_options;

typedef struct xmc_scene xmc_scene;

/* --- Scene: geometry + optics uploaded once (GenerateBVHTree + copy). --- */
int xmc_scene_create(const xmc_input* input, const xmc_options* options, xmc_scene** scene);
int xmc_scene_destroy(xmc_scene* scene);

/* Runs photons [firstPhoton, firstPhoton + numberOfPhotons) of stream `seed`
 * asynchronously on `hipStream` (NULL = the scene's stream), accumulating
 * into the scene's device tallies.  Inputs stay resident in HBM. */
int xmc_scene_run(xmc_scene* scene, uint64_t seed, uint64_t firstPhoton,
                  uint64_t numberOfPhotons, void* hipStream);

/* Blocks until every launch queued on the scene has finished. */
int xmc_scene_sync(xmc_scene* scene);

/* Waits for the scene's work, adds the device tallies into `output`
 * (+=, like LaunchOMP), sets output->specularReflectance and zeroes the
 * device tallies.  output->numberOfPhotons is not touched. */
int xmc_scene_collect(xmc_scene* scene, xmc_output* output);

/* Counters of the work since the last collect/reset. */
int xmc_scene_stats(xmc_scene* scene, xmc_stats* stats);

/* Traced run (debug / parity): photons [firstPhoton, firstPhoton+n) run with
 * tracing; photon firstPhoton+i writes at most maxEvents events at
 * events[i*maxEvents ...] and its event count in counts[i].  Host buffers. */
int xmc_scene_trace(xmc_scene* scene, uint64_t seed, uint64_t firstPhoton, uint64_t numberOfPhotons,
                    xmc_trace_event* events, int maxEvents, int* counts);

/* Nearest-hit ray queries against the scene's flattened BVH, on the GPU.
 * rays: origin[3], direction[3] as double (converted to float like the
 * transport does).  step[i] <= 0 means unlimited.  Outputs: found (0/1),
 * distance, surface id, triangle id (in the surface), unit normal[3].
 * bruteForce = 1 tests every triangle instead of walking the BVH. */
int xmc_scene_query(xmc_scene* scene, const double* origins, const double* directions,
                    const double* steps, int numberOfRays, int bruteForce,
                    int* found, double* distance, int* surfaceId, int* triangleId,
                    double* normal);

/* xmc_scene_query on device memory, allocating nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory --
 * code under HIP graph capture, tight serving loops.  The rays (float
 * origin[3], direction[3], step; step <= 0 = unlimited) and the outputs are
 * the caller's device buffers, laid out as xmc_scene_query's; the walk and
 * its results are the same.  The BVH stack spill lives in a static 32 MiB
 * device array per GPU (held until the process ends) shared by at most
 * 131072 lanes; a tree too deep for 256 lanes' spill to fit is rejected with
 * XMC_ERR_LIMIT.  Queued on hipStream (NULL = the scene's stream) and not
 * waited for.  Every call on one GPU shares that GPU's array: calls that
 * overlap in time on one GPU race. */
int xmc_scene_query_device(xmc_scene* scene, const float* origins, const float* directions, const float* steps,
                           int numberOfRays, int bruteForce, int* found, float* distance, int* surfaceId,
                           int* triangleId, float* normal, void* hipStream);

/* BVH built by the scene, in the reference's layout (for parity tests):
 * node arrays in the order GenerateBVHTree emits them (post-order, root
 * last).  Pass NULL arrays to query the sizes. */
int xmc_scene_bvh(xmc_scene* scene, int* numberOfNodes, int* root, int* numberOfTriangles,
                  double* boxes /* 6 per node */, int* links /* left,right,offset,count */,
                  int* triangleSurface, int* triangleIndex);

/* Launch shape of a scene (threads per block, resident blocks per CU cap);
 * used by the autotuner.  get returns the block size and total blocks. */
int xmc_scene_set_launch(xmc_scene* scene, int blockSize, int blocksPerCU);
int xmc_scene_get_launch(xmc_scene* scene, int* blockSize, int* blocks);

/* --- Batches: many configurations in one persistent launch. ---
 * No counterpart in the reference (it runs one input per process).  configs[c]
 * is a complete input; configs[c].numberOfPhotons is its photon budget.  A run
 * processes every budget in one launch fed by a device work queue (photon p of
 * configuration c draws the stream block sum_{c'<c} budget + p of `seed`),
 * replayed from a captured HIP graph when useGraph is set (default). */
typedef struct xmc_batch xmc_batch;
int xmc_batch_create(const xmc_input* configs, int numberOfConfigs, const xmc_options* options,
                     xmc_batch** batch);
int xmc_batch_destroy(xmc_batch* batch);
int xmc_batch_run(xmc_batch* batch, uint64_t seed, void* hipStream);
/* outputs[c] receives configuration c's tallies (+=). */
int xmc_batch_collect(xmc_batch* batch, xmc_output* outputs);
int xmc_batch_stats(xmc_batch* batch, xmc_stats* stats);
int xmc_batch_set_launch(xmc_batch* batch, int blockSize, int blocksPerCU, int useGraph);
int xmc_batch_get_launch(xmc_batch* batch, int* blockSize, int* blocks);

/* The conservative distance grid the scene would build (host only; for
 * tests): the centre lo + ((x, y, z) + 1/2) * h of cell (x, y, z) is at least
 * q[(z*dims[1]+y)*dims[0]+x] * stepQ away from every triangle, so a point p
 * is at least that less |p - centre| away.  q may be NULL to query the
 * sizes. */
int xmc_distance_grid(const xmc_input* input, uint64_t targetCells, float* lo, float* h, float* stepQ,
                      int* dims, uint8_t* q);

/* The separating slabs the scene would build on that grid for steps of mean
 * length meanStep (host only; for tests).  words[(z*dims[1]+y)*dims[0]+x]
 * packs, low bits first: a direction n as two int8 octahedral coordinates
 * (u, v) / 127 -> (u, v, 1 - |u| - |v|), folded where that is negative; the
 * bands a1 <= b1 <= a2 <= b2 as int8 multiples of R / 127; R as a uint8
 * multiple of *unit.  Every triangle point within R of the cell centre c
 * has n.(x - c) in [a1, b1] or [a2, b2]; 0 = no slab.  words may be NULL to
 * query the sizes. */
int xmc_slab_grid(const xmc_input* input, uint64_t targetCells, float meanStep, float* lo, float* h, float* unit,
                  int* dims, uint64_t* words);

/* --- Weight integrals of the biasing (FP64, on the GPU). --- */
/* ComputeWeightIntegral (mcml_integration.cpp:7-82) at numberOfDots values of
 * dot: 2-point Gauss-Legendre on precision x precision cells of
 * (theta, phi) in [0, pi] x [0, 2 pi].  The sines, cosines and powers that do
 * not depend on dot come from the host's libm (as in the reference); the sums
 * run on the GPU, one dot per thread, in the reference's order, so the values
 * equal the reference's. */
int xmc_weight_integrals(const double* dots, int numberOfDots, double anisotropy, double attractiveFactor,
                         int precision, double* out, int device);
/* The table ComputeWeightIntegral (weight_integral.cpp:33-64) builds for one
 * anisotropy: elements[j] at dot = -1 + 2 j / tablePrecision, j = 0..tablePrecision,
 * with the integral precision scaled by GetWeightIntegralPrecision
 * (weight_integral.cpp:92-103: x16 above g = 0.98, x8 above 0.97, x4 above
 * 0.95, x2 above 0.90).  elements holds tablePrecision + 1 doubles. */
int xmc_weight_table_compute(double anisotropy, double attractiveFactor, int integralPrecision, int tablePrecision,
                             double* elements, int device);

/* --- The LaunchOMP replacement. --- */
/* Builds a scene, runs input->numberOfPhotons photons starting at
 * rng->nextPhoton, adds tallies into output, advances rng->nextPhoton. */
int xmc_launch(const xmc_input* input, xmc_output* output, xmc_rng* rng,
               const xmc_options* options);

/* InitOutput / FreeOutput equivalents (malloc-based).  Like InitOutput
 * (launcher_omp.cpp:10-58), xmc_init_output allocates every tally: the
 * absorption and scattering grids and each detector's trajectory map, so
 * that xmc_launch on it produces all of them, as LaunchOMP does.  The
 * scattering grid and trajectory maps need the path-tally kernel variant
 * (about 2.5x slower on the 3-D surface configuration); a caller that does
 * not want them frees and NULLs those arrays first. */
int xmc_init_output(const xmc_input* input, xmc_output* output);
void xmc_free_output(xmc_output* output);

double xmc_specular_reflectance(const xmc_layer* layers);
/* The HIP device's name and architecture ("AMD Instinct MI355X gfx950:sramecc+:xnack-"),
 * NUL-terminated in name[capacity], and its compute-unit count; for caches of
 * tuned launch shapes.  No counterpart in the reference. */
int xmc_device_name(int device, char* name, int capacity, int* computeUnits);
const char* xmc_last_error(void);
const char* xmc_version(void);

#ifdef __cplusplus
}
#endif

#endif /* XMCML_AMD_XMCML_H */
