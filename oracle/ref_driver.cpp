/*
 * ref_driver.cpp -- drives the reference's own sources (compiled unmodified
 * from the reference tree by oracle/Makefile into oracle/_ref/) so that tests
 * can check the oracle, and through it the GPU engine, against the reference.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links this.
 *
 * All physics comes from the reference objects: ComputePhoton, LaunchOMP,
 * GenerateBVHTree, ComputeBVHIntersection and the exported kernel helpers
 * (base/xmcml/xmcml/mcml_kernel.h:14-44).  This file only converts the
 * xmc_* structures into InputInfo/OutputInfo, seeds MCG59 streams, and, for
 * traces, repeats the control flow of ComputePhoton (mcml_kernel.cpp:39-126)
 * around the reference helpers so that each iteration can be recorded.
 */
#include "mcml_kernel.h"
#include "mcml_math.h"
#include "mcml_integration.h"
#include "launcher_omp.h"

#include <omp.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/xmcml.h"

namespace {

const uint64_t kMcgC = 302875106592253ULL;
const uint64_t kMcgMask = 576460752303423487ULL;

uint64_t upow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r *= a;
        a *= a;
        e >>= 1;
    }
    return r;
}

// Per-photon block split of the MCG59 sequence: photon p starts at
// (2s+1) * a^(p * 2^24) and draws with offset a.
MCG59 photon_stream(uint64_t seed, uint64_t p) {
    MCG59 g;
    g.value = ((2 * seed + 1) * upow(upow(kMcgC, 1ULL << 24), p)) & kMcgMask;
    g.offset = kMcgC;
    return g;
}

struct RefInput {
    InputInfo in;
    Area area;
    std::vector<LayerInfo> layers;
    std::vector<Surface> surfaces;
    std::vector<CubeDetector> cubes;
    std::vector<RingDetector> rings;
    WeightIntegralTable table;
    double table_elem;
    std::vector<WeightIntegralTable> tables;

    explicit RefInput(const xmc_input* x, bool with_bvh = true) {
        std::memset(&in, 0, sizeof(in));
        area.corner = {{x->area->corner.x, x->area->corner.y, x->area->corner.z}};
        area.length = {{x->area->length.x, x->area->length.y, x->area->length.z}};
        area.partitionNumber.x = x->area->partitionNumber.x;
        area.partitionNumber.y = x->area->partitionNumber.y;
        area.partitionNumber.z = x->area->partitionNumber.z;
        in.area = &area;
        layers.resize(x->numberOfLayers);
        for (int i = 0; i < x->numberOfLayers; i++) {
            const xmc_layer& l = x->layerInfo[i];
            layers[i].refractiveIndex = l.refractiveIndex;
            layers[i].absorptionCoefficient = l.absorptionCoefficient;
            layers[i].scatteringCoefficient = l.scatteringCoefficient;
            layers[i].anisotropy = l.anisotropy;
            layers[i].surfaceId = const_cast<int*>(l.surfaceId);
            layers[i].numberOfSurfaces = l.numberOfSurfaces;
        }
        in.layerInfo = layers.data();
        in.numberOfLayers = x->numberOfLayers;
        surfaces.resize(x->numberOfSurfaces);
        for (int i = 0; i < x->numberOfSurfaces; i++) {
            surfaces[i].vertices = reinterpret_cast<double3*>(const_cast<xmc_double3*>(x->surface[i].vertices));
            surfaces[i].numberOfVertices = x->surface[i].numberOfVertices;
            surfaces[i].triangles = reinterpret_cast<int3*>(const_cast<xmc_int3*>(x->surface[i].triangles));
            surfaces[i].numberOfTriangles = x->surface[i].numberOfTriangles;
        }
        in.surface = surfaces.data();
        in.numberOfSurfaces = x->numberOfSurfaces;
        cubes.resize(x->numberOfCubeDetectors);
        for (int i = 0; i < x->numberOfCubeDetectors; i++) {
            const xmc_cube_detector& c = x->cubeDetector[i];
            cubes[i].center = {{c.center.x, c.center.y, c.center.z}};
            cubes[i].length = {{c.length.x, c.length.y, c.length.z}};
            for (int j = 0; j < MAX_LAYERS; j++) cubes[i].filterLayers[j] = c.filterLayers[j] != 0;
            cubes[i].targetAngle = c.targetAngle;
        }
        rings.resize(x->numberOfRingDetectors);
        for (int i = 0; i < x->numberOfRingDetectors; i++) {
            const xmc_ring_detector& c = x->ringDetector[i];
            rings[i].center = {{c.center.x, c.center.y, c.center.z}};
            rings[i].smallRadius = c.smallRadius;
            rings[i].bigRadius = c.bigRadius;
            for (int j = 0; j < MAX_LAYERS; j++) rings[i].filterLayers[j] = c.filterLayers[j] != 0;
            rings[i].targetAngle = c.targetAngle;
        }
        in.cubeDetector = cubes.data();
        in.numberOfCubeDetectors = x->numberOfCubeDetectors;
        in.ringDetector = rings.data();
        in.numberOfRingDetectors = x->numberOfRingDetectors;
        in.minWeight = x->minWeight;
        in.timeStart = x->timeStart;
        in.timeFinish = x->timeFinish;
        in.timeScaleSize = x->timeScaleSize;
        in.startPosition = {{x->startPosition.x, x->startPosition.y, x->startPosition.z}};
        in.startDirection = {{x->startDirection.x, x->startDirection.y, x->startDirection.z}};
        in.useBiasing = x->useBiasing ? 1 : 0;
        if (x->useBiasing) {  // the caller's tables (the reference computes them in weight_integral.cpp)
            in.attractiveFactor = x->attractiveFactor;
            in.targetPoint = {{x->targetPoint.x, x->targetPoint.y, x->targetPoint.z}};
            for (int t = 0; t < x->numberOfWeightTables; t++) {
                WeightIntegralTable w;
                w.anisotropy = x->weightTable[t].anisotropy;
                w.numberOfElements = x->weightTable[t].numberOfElements;
                w.elements = const_cast<double*>(x->weightTable[t].elements);
                tables.push_back(w);
            }
            in.weightTable = tables.data();
            in.numberOfWeightTables = (int)tables.size();
        } else {
            // main.cpp:146-152 gives a one-element table when biasing is off
            table.anisotropy = 0.0;
            table.numberOfElements = 1;
            table_elem = 0.0;
            table.elements = &table_elem;
            in.weightTable = &table;
            in.numberOfWeightTables = 1;
        }
        for (int j = 0; j < MAX_LAYERS; j++) in.targetRangeLayers[j] = x->targetRangeLayers[j];
        in.numberOfPhotons = x->numberOfPhotons;
        in.bvhTree = with_bvh ? GenerateBVHTree(in.surface, in.numberOfSurfaces) : nullptr;
    }
    ~RefInput() { delete in.bvhTree; }
};

// OutputInfo built by the reference's own InitOutput, added into xmc_output.
struct RefOutput {
    OutputInfo out;
    explicit RefOutput(RefInput& r) { InitOutput(&r.in, &out); }
    ~RefOutput() { FreeOutput(&out); }
    void add_into(xmc_output* x) const {
        if (x->absorption)
            for (int j = 0; j < out.gridSize; j++) x->absorption[j] += out.absorption[j];
        if (x->scattering)
            for (int j = 0; j < out.gridSize; j++) x->scattering[j] += out.scattering[j];
        x->specularReflectance = out.specularReflectance;
        for (int d = 0; d < out.numberOfDetectors && d < x->numberOfDetectors; d++) {
            const DetectorInfo& s = out.detectorInfo[d];
            xmc_detector_info& t = x->detectorInfo[d];
            t.weight += s.weight;
            t.targetWeight += s.targetWeight;
            t.numberOfPhotons += s.numberOfPhotons;
            if (t.trajectory)
                for (int j = 0; j < out.gridSize; j++) t.trajectory[j] += s.trajectory[j];
            for (int b = 0; b < s.timeScaleSize && b < t.timeScaleSize; b++) {
                t.timeScale[b].weight += s.timeScale[b].weight;
                t.timeScale[b].targetWeight += s.timeScale[b].targetWeight;
                t.timeScale[b].numberOfPhotons += s.timeScale[b].numberOfPhotons;
            }
        }
    }
};

void clear_output(OutputInfo* o, int T) {
    // InitOutput leaves the last bin's targetWeight unset (launcher_omp.cpp:50-55)
    for (int d = 0; d < o->numberOfDetectors; d++) {
        o->detectorInfo[d].targetWeight = 0.0;
        for (int b = 0; b < T; b++) o->detectorInfo[d].timeScale[b].targetWeight = 0.0;
    }
}

int count_draws(uint64_t before, uint64_t after) {
    uint64_t v = before;
    for (int k = 0; k < 64; k++) {
        if (v == after) return k;
        v = (v * kMcgC) & kMcgMask;
    }
    return -1;
}

struct Recorder {
    xmc_trace_event* ev;
    int max_ev;
    int n = 0;
    uint32_t draws = 0;
    void put(int type, const PhotonState& p, int aux) {
        if (ev && n < max_ev) {
            xmc_trace_event& e = ev[n];
            e.type = type;
            e.layer = p.layerId;
            e.aux = aux;
            e.draws = draws;
            e.pos[0] = p.position.x; e.pos[1] = p.position.y; e.pos[2] = p.position.z;
            e.dir[0] = p.direction.x; e.dir[1] = p.direction.y; e.dir[2] = p.direction.z;
            e.weight = p.weight;
            e.time = p.time;
        }
        n++;
    }
};

// ComputePhoton's loop (mcml_kernel.cpp:39-126) around the reference
// helpers, with MCML roulette (kernel.cu:214-222, applied as kernel.cu:449-452)
// when requested.  Without roulette it must agree bit for bit with
// ComputePhoton itself; the tests check that.
void driver_photon(double rsp, InputInfo* input, OutputInfo* output, MCG59* rng, PhotonTrajectory* traj,
                   const xmc_input* x, Recorder* rec) {
    PhotonState photon;
    traj->position = 0;
    LaunchPhoton(input, &photon, rsp);
    if (rec) rec->put(XMC_EV_LAUNCH, photon, 0);
    while (!photon.isDead) {
        uint64_t v0 = rng->value;
        int layerId = photon.layerId;
        bool glass = (input->layerInfo[layerId].absorptionCoefficient == 0.0) &&
                     (input->layerInfo[layerId].scatteringCoefficient == 0.0);
        int ev = -1, aux = 0;
        IntersectionInfo hit;
        if (glass) {
            hit = ComputeBVHIntersectionWithoutStep(photon.position, photon.direction, input->bvhTree,
                                                    input->surface);
        } else {
            ComputeStepSizeInTissue(&photon, input, rng);
            hit = ComputeBVHIntersection(photon.position, photon.direction, photon.step, input->bvhTree,
                                         input->surface);
        }
        if (hit.isFindIntersection) {
            photon.step = hit.distance - 0.01 * 1.0E-8;
            MovePhotonAndUpdateTrajectory(&photon, input, output->scattering, traj);
            if (GetAreaIndex(photon.position, input->area) >= 0) {
                int before = photon.layerId;
                CrossBoundary(&photon, input, output, &hit, rng, traj);
                if (photon.isDead) {
                    ev = XMC_EV_ESCAPE;
                    aux = GetDetectorId(&photon, input);
                } else if (photon.layerId != before) {
                    ev = XMC_EV_TRANSMIT;
                    aux = hit.surfaceId;
                } else {
                    ev = XMC_EV_REFLECT;
                    aux = hit.surfaceId;
                }
            } else {
                photon.isDead = true;
                ev = XMC_EV_DIE_AREA;
            }
        } else if (glass) {
            photon.isDead = true;
            ev = XMC_EV_GLASS_MISS;
        } else {
            MovePhotonAndUpdateTrajectory(&photon, input, output->scattering, traj);
            int idx = GetAreaIndex(photon.position, input->area);
            if (idx >= 0) {
                ComputeDroppedWeightOfPhoton(&photon, input, output, idx);
                if (input->useBiasing)
                    ComputePhotonDirectionWithBiasing(&photon, input, rng);
                else
                    photon.direction = GetNewPhotonDirection(&photon, input->layerInfo[photon.layerId].anisotropy, rng);
                ev = XMC_EV_SCATTER;
                aux = idx;
            } else {
                photon.isDead = true;
                ev = XMC_EV_DIE_AREA;
            }
        }
        if (rec) {
            rec->draws += count_draws(v0, rng->value);
            rec->put(ev, photon, aux);
            v0 = rng->value;
        }
        if (x->terminationMode == XMC_TERMINATE_ROULETTE) {
            if (photon.weight < input->minWeight && !photon.isDead) {
                int t;
                if (photon.weight == 0.0) {
                    photon.isDead = true;
                    t = XMC_EV_ROULETTE_DIE;
                } else if (NextMCG59(rng) < x->rouletteChance) {
                    photon.weight /= x->rouletteChance;
                    t = XMC_EV_ROULETTE_LIVE;
                } else {
                    photon.isDead = true;
                    t = XMC_EV_ROULETTE_DIE;
                }
                if (rec) {
                    rec->draws += count_draws(v0, rng->value);
                    rec->put(t, photon, 0);
                }
            }
        } else if (photon.weight < input->minWeight) {
            if (rec && !photon.isDead) rec->put(XMC_EV_DIE_WEIGHT, photon, 0);
            photon.isDead = true;
        }
    }
}

}  // namespace

extern "C" {

// Photons [first, first+n), one at a time, per-photon streams.  use_driver=0
// calls ComputePhoton itself (termination must be KILL); 1 runs the driver loop.
int xref_run(const xmc_input* x, xmc_output* o, uint64_t seed, uint64_t first, uint64_t n, int use_driver) {
    if (!use_driver && x->terminationMode != XMC_TERMINATE_KILL) return -1;
    RefInput r(x);
    RefOutput ro(r);
    clear_output(&ro.out, x->timeScaleSize);
    double rsp = ComputeSpecularReflectance(r.in.layerInfo);
    ro.out.specularReflectance = rsp;
    PhotonTrajectory* traj = new PhotonTrajectory;
    for (uint64_t i = 0; i < n; i++) {
        MCG59 g = photon_stream(seed, first + i);
        if (use_driver) driver_photon(rsp, &r.in, &ro.out, &g, traj, x, nullptr);
        else ComputePhoton(rsp, &r.in, &ro.out, &g, traj);
    }
    delete traj;
    ro.add_into(o);
    return 0;
}

// The reference's LaunchOMP unchanged, with the thread streams main.cpp:106-112
// builds for one MPI rank.  Returns wall seconds of the LaunchOMP call.
double xref_launch_omp(const xmc_input* x, xmc_output* o, uint64_t seed, int nthreads) {
    RefInput r(x);
    RefOutput ro(r);
    clear_output(&ro.out, x->timeScaleSize);
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    std::vector<MCG59> g(nthreads);
    for (int i = 0; i < nthreads; i++) InitMCG59(&g[i], seed, i, nthreads);
    r.in.numberOfPhotons = x->numberOfPhotons;
    auto t0 = std::chrono::steady_clock::now();
    LaunchOMP(&r.in, &ro.out, g.data(), nthreads);
    auto t1 = std::chrono::steady_clock::now();
    ro.add_into(o);
    return std::chrono::duration<double>(t1 - t0).count();
}

int xref_trace(const xmc_input* x, uint64_t seed, uint64_t photon, xmc_trace_event* ev, int max_ev) {
    RefInput r(x);
    RefOutput ro(r);
    double rsp = ComputeSpecularReflectance(r.in.layerInfo);
    PhotonTrajectory* traj = new PhotonTrajectory;
    MCG59 g = photon_stream(seed, photon);
    Recorder rec;
    rec.ev = ev;
    rec.max_ev = max_ev;
    driver_photon(rsp, &r.in, &ro.out, &g, traj, x, &rec);
    delete traj;
    return rec.n <= max_ev ? rec.n : -rec.n;
}

int xref_bvh(const xmc_input* x, int* nn, int* root, int* nt, double* boxes, int* links, int* tsurf, int* tidx) {
    RefInput r(x);
    BVHTree* t = r.in.bvhTree;
    *nn = t->numOfNodes;
    *root = t->root;
    *nt = t->numOfTriangles;
    if (boxes) {
        for (int i = 0; i < t->numOfNodes; i++) {
            const BVHNode& n = t->nodes[i];
            double* b = boxes + 6 * i;
            b[0] = n.Box.Ver1.x; b[1] = n.Box.Ver1.y; b[2] = n.Box.Ver1.z;
            b[3] = n.Box.Ver2.x; b[4] = n.Box.Ver2.y; b[5] = n.Box.Ver2.z;
            links[4 * i] = n.leftNode;
            links[4 * i + 1] = n.rightNode;
            links[4 * i + 2] = n.offset;
            links[4 * i + 3] = n.numOfTriangles;
        }
        for (int i = 0; i < t->numOfTriangles; i++) {
            tsurf[i] = t->triangles[i].surfId;
            tidx[i] = t->triangles[i].trIndex;
        }
    }
    return 0;
}

int xref_query(const xmc_input* x, const double* o, const double* d, const double* steps, int nr, int* found,
               double* dist, int* surf, double* normal) {
    RefInput r(x);
    for (int i = 0; i < nr; i++) {
        double3 oo = {{o[3 * i], o[3 * i + 1], o[3 * i + 2]}};
        double3 dd = {{d[3 * i], d[3 * i + 1], d[3 * i + 2]}};
        IntersectionInfo h = steps[i] > 0.0
                                 ? ComputeBVHIntersection(oo, dd, steps[i], r.in.bvhTree, r.in.surface)
                                 : ComputeBVHIntersectionWithoutStep(oo, dd, r.in.bvhTree, r.in.surface);
        found[i] = h.isFindIntersection;
        dist[i] = h.isFindIntersection ? h.distance : 0.0;
        surf[i] = h.isFindIntersection ? h.surfaceId : -1;
        normal[3 * i] = h.isFindIntersection ? h.normal.x : 0.0;
        normal[3 * i + 1] = h.isFindIntersection ? h.normal.y : 0.0;
        normal[3 * i + 2] = h.isFindIntersection ? h.normal.z : 0.0;
    }
    return 0;
}

// Single reference functions, for the golden values of the unit tests.
double xref_specular(const xmc_layer* x) {
    LayerInfo l[3];
    for (int i = 0; i < 3; i++) {
        l[i].refractiveIndex = x[i].refractiveIndex;
        l[i].absorptionCoefficient = x[i].absorptionCoefficient;
        l[i].scatteringCoefficient = x[i].scatteringCoefficient;
        l[i].anisotropy = x[i].anisotropy;
    }
    return ComputeSpecularReflectance(l);
}
double xref_cosine_theta(double g, uint64_t* value, uint64_t offset) {
    MCG59 m = {*value, offset};
    double c = ComputeCosineTheta(g, &m);
    *value = m.value;
    return c;
}
void xref_new_direction(const double* dir, double g, uint64_t* value, uint64_t offset, double* out) {
    MCG59 m = {*value, offset};
    PhotonState p;
    p.direction = {{dir[0], dir[1], dir[2]}};
    double3 r = GetNewPhotonDirection(&p, g, &m);
    *value = m.value;
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
double xref_transmit_cos(double n1, double n2, double ci) { return ComputeTransmitCosine(n1, n2, ci); }
double xref_fresnel(double n1, double n2, double ci, double ct) { return ComputeFrenselReflectance(n1, n2, ci, ct); }
void xref_reflect(const double* i, const double* n, double* out) {
    double3 r = ReflectVector({{i[0], i[1], i[2]}}, {{n[0], n[1], n[2]}});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void xref_refract(double n1, double n2, const double* i, const double* n, double* out) {
    double3 r = RefractVector(n1, n2, {{i[0], i[1], i[2]}}, {{n[0], n[1], n[2]}});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
double xref_step(double mua, double mus, uint64_t* value, uint64_t offset) {
    MCG59 m = {*value, offset};
    LayerInfo l[2];
    l[1].absorptionCoefficient = mua;
    l[1].scatteringCoefficient = mus;
    InputInfo in;
    in.layerInfo = l;
    PhotonState p;
    p.layerId = 1;
    ComputeStepSizeInTissue(&p, &in, &m);
    *value = m.value;
    return p.step;
}
double xref_triangle_distance(const double* o, const double* d, const double* a, const double* b, const double* c) {
    return GetTriangleIntersectionDistance({{o[0], o[1], o[2]}}, {{d[0], d[1], d[2]}}, {{a[0], a[1], a[2]}},
                                           {{b[0], b[1], b[2]}}, {{c[0], c[1], c[2]}});
}
int xref_detector_id(const xmc_input* x, const double* pos, const double* dir, const uint8_t* visited) {
    RefInput r(x, false);
    PhotonState p;
    p.position = {{pos[0], pos[1], pos[2]}};
    p.direction = {{dir[0], dir[1], dir[2]}};
    for (int i = 0; i < MAX_LAYERS; i++) p.visitedLayers[i] = visited[i] != 0;
    return GetDetectorId(&p, &r.in);
}

// ComputeWeightIntegral (mcml_integration.cpp:7-82) of the reference build.
double ref_weight_integral(double dot, double anisotropy, double eps, unsigned n) {
    return ComputeWeightIntegral(dot, anisotropy, eps, n);
}

}  // extern "C"
