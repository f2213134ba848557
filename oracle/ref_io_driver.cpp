/*
 * ref_io_driver.cpp -- drives the reference's own file readers and writers
 * (ParseInputFile, ParseSurfaceFile, WriteOutputToFile, WriteBackupToFile,
 * ParseBackupFile, WriteWeightIntegralTableToFile, ParseWeightIntegralFile;
 * base/xmcml/xmcmlLauncher/{parser,writer,reader,weight_integral}.cpp, with
 * the tinyxml sources of the same repository) compiled unmodified by
 * oracle/Makefile into oracle/_ref/libxmcml_ref_io.so.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links this.
 *
 * This file only converts between the xmc_* structures and InputInfo /
 * OutputInfo; every byte read or written comes from the reference code.
 */
#include "mcml_kernel.h"
#include "launcher_omp.h"
#include "parser.h"
#include "reader.h"
#include "writer.h"
#include "weight_integral.h"

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../include/xmcml_io.h"

// Defined in weight_integral.cpp without a header declaration (:9-14).
bool WriteWeightIntegralTableToFile(InputInfo* input);
bool ParseWeightIntegralFile(InputInfo* input, CmdArguments* args);

namespace {

// The parsed input of the last refio_parse call, as xmc_* arrays.
struct Parsed {
    xmc_area area;
    std::vector<xmc_layer> layers;
    std::vector<std::vector<int>> ids;
    std::vector<xmc_surface> surfaces;
    std::vector<std::vector<xmc_double3>> verts;
    std::vector<std::vector<xmc_int3>> tris;
    std::vector<xmc_cube_detector> cubes;
    std::vector<xmc_ring_detector> rings;
} g_parsed;

xmc_double3 d3(const double3& v) { return {v.x, v.y, v.z}; }
double3 r3(const xmc_double3& v) {
    double3 r;
    r.x = v.x;
    r.y = v.y;
    r.z = v.z;
    return r;
}

// InputInfo of an xmc_input, for the writers (area, detectors, photons).
struct RefIn {
    InputInfo in;
    Area area;
    std::vector<CubeDetector> cubes;
    std::vector<RingDetector> rings;
    explicit RefIn(const xmc_input* x) {
        std::memset(&in, 0, sizeof(in));
        area.corner = r3(x->area->corner);
        area.length = r3(x->area->length);
        area.partitionNumber.x = x->area->partitionNumber.x;
        area.partitionNumber.y = x->area->partitionNumber.y;
        area.partitionNumber.z = x->area->partitionNumber.z;
        in.area = &area;
        cubes.resize(x->numberOfCubeDetectors);
        for (int i = 0; i < x->numberOfCubeDetectors; i++) {
            cubes[i].center = r3(x->cubeDetector[i].center);
            cubes[i].length = r3(x->cubeDetector[i].length);
            for (int j = 0; j < MAX_LAYERS; j++) cubes[i].filterLayers[j] = x->cubeDetector[i].filterLayers[j] != 0;
            cubes[i].targetAngle = x->cubeDetector[i].targetAngle;
        }
        rings.resize(x->numberOfRingDetectors);
        for (int i = 0; i < x->numberOfRingDetectors; i++) {
            rings[i].center = r3(x->ringDetector[i].center);
            rings[i].smallRadius = x->ringDetector[i].smallRadius;
            rings[i].bigRadius = x->ringDetector[i].bigRadius;
            for (int j = 0; j < MAX_LAYERS; j++) rings[i].filterLayers[j] = x->ringDetector[i].filterLayers[j] != 0;
            rings[i].targetAngle = x->ringDetector[i].targetAngle;
        }
        in.cubeDetector = cubes.data();
        in.numberOfCubeDetectors = x->numberOfCubeDetectors;
        in.ringDetector = rings.data();
        in.numberOfRingDetectors = x->numberOfRingDetectors;
        in.numberOfPhotons = x->numberOfPhotons;
    }
};

// OutputInfo viewing an xmc_output's arrays (same layouts: TimeInfo and
// xmc_time_info are both 5 x 8 bytes).
struct RefOut {
    OutputInfo out;
    std::vector<DetectorInfo> dets;
    std::vector<double> zeros_abs, zeros_scat;
    std::vector<std::vector<uint64>> zeros_traj;
    explicit RefOut(const xmc_output* x) {
        out.numberOfPhotons = x->numberOfPhotons;
        out.specularReflectance = x->specularReflectance;
        out.gridSize = x->gridSize;
        if (!x->absorption) zeros_abs.assign(x->gridSize, 0.0);
        if (!x->scattering) zeros_scat.assign(x->gridSize, 0.0);
        out.absorption = x->absorption ? x->absorption : zeros_abs.data();
        out.scattering = x->scattering ? x->scattering : zeros_scat.data();
        dets.resize(x->numberOfDetectors);
        zeros_traj.resize(x->numberOfDetectors);
        for (int i = 0; i < x->numberOfDetectors; i++) {
            const xmc_detector_info& s = x->detectorInfo[i];
            DetectorInfo& d = dets[i];
            d.numberOfPhotons = s.numberOfPhotons;
            if (!s.trajectory) zeros_traj[i].assign(s.trajectorySize, 0);
            d.trajectory = s.trajectory ? reinterpret_cast<uint64*>(s.trajectory) : zeros_traj[i].data();
            d.trajectorySize = s.trajectorySize;
            d.timeScale = reinterpret_cast<TimeInfo*>(s.timeScale);
            d.timeScaleSize = s.timeScaleSize;
            d.targetWeight = s.targetWeight;
            d.weight = s.weight;
        }
        out.detectorInfo = dets.data();
        out.numberOfDetectors = x->numberOfDetectors;
    }
    void copy_back(xmc_output* x) const {
        x->numberOfPhotons = out.numberOfPhotons;
        for (int i = 0; i < x->numberOfDetectors; i++) {
            x->detectorInfo[i].numberOfPhotons = dets[i].numberOfPhotons;
            x->detectorInfo[i].targetWeight = dets[i].targetWeight;
            x->detectorInfo[i].weight = dets[i].weight;
        }
    }
};

}  // namespace

extern "C" {

// ParseInputFile (+ ParseSurfaceFile when surface != NULL) into *x; the arrays
// stay valid until the next call.  Returns 1 on success, 0 when the reference
// reports failure.
int refio_parse(const char* xml, const char* surface, xmc_input* x, xmc_biasing* bias) {
    InputInfo in;
    std::memset(&in, 0, sizeof(in));
    std::string xs(xml);
    if (!ParseInputFile(&xs[0], &in)) return 0;
    if (surface) {
        std::string ss(surface);
        if (!ParseSurfaceFile(&ss[0], &in)) return 0;
    }
    Parsed& P = g_parsed;
    P = Parsed();
    std::memset(x, 0, sizeof(*x));
    x->numberOfPhotons = in.numberOfPhotons;
    if (in.area) {
        P.area.corner = d3(in.area->corner);
        P.area.length = d3(in.area->length);
        P.area.partitionNumber = {in.area->partitionNumber.x, in.area->partitionNumber.y, in.area->partitionNumber.z};
        x->area = &P.area;
    }
    P.layers.resize(in.numberOfLayers);
    P.ids.resize(in.numberOfLayers);
    for (int i = 0; i < in.numberOfLayers; i++) {
        const LayerInfo& l = in.layerInfo[i];
        P.ids[i].assign(l.surfaceId, l.surfaceId + l.numberOfSurfaces);
        P.layers[i] = {l.refractiveIndex, l.absorptionCoefficient, l.scatteringCoefficient, l.anisotropy,
                       P.ids[i].data(), l.numberOfSurfaces};
    }
    x->layerInfo = P.layers.data();
    x->numberOfLayers = in.numberOfLayers;
    x->numberOfSurfaces = in.numberOfSurfaces;
    if (surface) {
        P.surfaces.resize(in.numberOfSurfaces);
        P.verts.resize(in.numberOfSurfaces);
        P.tris.resize(in.numberOfSurfaces);
        for (int i = 0; i < in.numberOfSurfaces; i++) {
            const Surface& s = in.surface[i];
            for (int k = 0; k < s.numberOfVertices; k++) P.verts[i].push_back(d3(s.vertices[k]));
            for (int k = 0; k < s.numberOfTriangles; k++)
                P.tris[i].push_back({s.triangles[k].x, s.triangles[k].y, s.triangles[k].z});
            P.surfaces[i] = {P.verts[i].data(), s.numberOfVertices, P.tris[i].data(), s.numberOfTriangles};
        }
        x->surface = P.surfaces.data();
    }
    for (int i = 0; i < in.numberOfCubeDetectors; i++) {
        const CubeDetector& c = in.cubeDetector[i];
        xmc_cube_detector d{};
        d.center = d3(c.center);
        d.length = d3(c.length);
        for (int j = 0; j < MAX_LAYERS; j++) d.filterLayers[j] = c.filterLayers[j] ? 1 : 0;
        d.targetAngle = c.targetAngle;
        P.cubes.push_back(d);
    }
    for (int i = 0; i < in.numberOfRingDetectors; i++) {
        const RingDetector& r = in.ringDetector[i];
        xmc_ring_detector d{};
        d.center = d3(r.center);
        d.smallRadius = r.smallRadius;
        d.bigRadius = r.bigRadius;
        for (int j = 0; j < MAX_LAYERS; j++) d.filterLayers[j] = r.filterLayers[j] ? 1 : 0;
        d.targetAngle = r.targetAngle;
        P.rings.push_back(d);
    }
    x->cubeDetector = P.cubes.data();
    x->numberOfCubeDetectors = in.numberOfCubeDetectors;
    x->ringDetector = P.rings.data();
    x->numberOfRingDetectors = in.numberOfRingDetectors;
    x->minWeight = in.minWeight;
    x->timeStart = in.timeStart;
    x->timeFinish = in.timeFinish;
    x->timeScaleSize = in.timeScaleSize;
    x->startPosition = d3(in.startPosition);
    x->startDirection = d3(in.startDirection);
    for (int j = 0; j < MAX_LAYERS; j++) x->targetRangeLayers[j] = in.targetRangeLayers[j];
    x->terminationMode = XMC_TERMINATE_KILL;
    bias->useBiasing = in.useBiasing;
    bias->weightIntegralPrecision = (int)in.weightIntegralPrecision;
    bias->weightTablePrecision = (int)in.weightTablePrecision;
    bias->attractiveFactor = in.attractiveFactor;
    bias->targetPoint = d3(in.targetPoint);
    return 1;
}

// WriteOutputToFile of the same input / output.
int refio_write_output(const xmc_input* x, const xmc_output* o, const char* path) {
    RefIn ri(x);
    RefOut ro(o);
    std::string p(path);
    return WriteOutputToFile(&ri.in, &ro.out, &p[0]) ? 1 : 0;
}

// WriteBackupToFile with one MCG59 record {value, offset}; the file lands in
// the working directory under the reference's own name.
int refio_write_backup(const xmc_input* x, const xmc_output* o, uint64_t value, uint64_t offset) {
    RefIn ri(x);
    RefOut ro(o);
    MCG59 g;
    g.value = value;
    g.offset = offset;
    return WriteBackupToFile(&ri.in, &ro.out, &g, 1, 1) ? 1 : 0;
}

// ParseBackupFile for one thread and one process into an output laid out by
// xmc_init_output (arrays are filled in place; counts and weights copied back).
int refio_read_backup(const char* path, xmc_output* o, uint64_t* value, uint64_t* offset) {
    RefOut ro(o);
    MCG59 g;
    std::string p(path);
    if (!ParseBackupFile(&p[0], &ro.out, &g, 1, 1)) return 0;
    ro.copy_back(o);
    *value = g.value;
    *offset = g.offset;
    return 1;
}

// ReadOutputFormBackupFile (reader.cpp:366-392): the tallies of a backup or
// result file, without the generator section.
int refio_read_output(const char* path, xmc_output* o) {
    RefOut ro(o);
    std::string p(path);
    if (ReadOutputFormBackupFile(&p[0], &ro.out) != 0) return 0;
    ro.copy_back(o);
    return 1;
}

// WriteWeightIntegralTableToFile for the given tables (elements concatenated);
// the file lands in the working directory under the reference's own name.
int refio_write_weight_table(double factor, int integralPrecision, int numberOfTables, const double* anisotropy,
                             const int* numberOfElements, const double* elements) {
    InputInfo in;
    std::memset(&in, 0, sizeof(in));
    in.attractiveFactor = factor;
    in.weightIntegralPrecision = integralPrecision;
    in.numberOfWeightTables = numberOfTables;
    std::vector<WeightIntegralTable> t(numberOfTables);
    for (int i = 0; i < numberOfTables; i++) {
        t[i].anisotropy = anisotropy[i];
        t[i].numberOfElements = numberOfElements[i];
        t[i].elements = const_cast<double*>(elements);
        elements += numberOfElements[i];
    }
    in.weightTable = t.data();
    return WriteWeightIntegralTableToFile(&in) ? 1 : 0;
}

// ParseWeightIntegralFile against an input of numberOfLayers layers with the
// given anisotropies and biasing settings.  On success the tables are copied
// out (at most capTables tables and capElements elements in all).
int refio_parse_weight_table(const char* path, int numberOfLayers, const double* layerAnisotropy, double factor,
                             int integralPrecision, int tablePrecision, int capTables, int capElements,
                             int* numberOfTables, double* anisotropy, int* numberOfElements, double* elements) {
    InputInfo in;
    std::memset(&in, 0, sizeof(in));
    std::vector<LayerInfo> layers(numberOfLayers);
    for (int i = 0; i < numberOfLayers; i++) layers[i].anisotropy = layerAnisotropy[i];
    in.layerInfo = layers.data();
    in.numberOfLayers = numberOfLayers;
    in.attractiveFactor = factor;
    in.weightIntegralPrecision = integralPrecision;
    in.weightTablePrecision = tablePrecision;
    CmdArguments args;
    std::memset(&args, 0, sizeof(args));
    std::string p(path);
    args.weightTableFileName = &p[0];
    if (!ParseWeightIntegralFile(&in, &args)) return 0;
    *numberOfTables = in.numberOfWeightTables;
    int used = 0;
    for (int i = 0; i < in.numberOfWeightTables && i < capTables; i++) {
        const WeightIntegralTable& w = in.weightTable[i];
        anisotropy[i] = w.anisotropy;
        numberOfElements[i] = w.numberOfElements;
        for (int k = 0; k < w.numberOfElements && used < capElements; k++) elements[used++] = w.elements[k];
        delete[] w.elements;
    }
    delete[] in.weightTable;
    return 1;
}

}  // extern "C"
