"""ctypes mirror of include/xmcml.h and helpers that keep the numpy arrays
behind an ``xmc_input`` / ``xmc_output`` alive.

The structures follow xmcml's InputInfo / OutputInfo family
(base/xmcml/xmcml/mcml_kernel_types.h:10-102, mcml_types.h:8-39) field for
field; see include/xmcml.h for the C declarations.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

MAX_LAYERS = 8
TERMINATE_KILL = 0
TERMINATE_ROULETTE = 1

EV_LAUNCH, EV_SCATTER, EV_REFLECT, EV_TRANSMIT, EV_ESCAPE = 0, 1, 2, 3, 4
EV_DIE_AREA, EV_DIE_WEIGHT, EV_ROULETTE_LIVE, EV_ROULETTE_DIE, EV_GLASS_MISS = 5, 6, 7, 8, 9


class Double3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class Area(C.Structure):
    _fields_ = [("corner", Double3), ("length", Double3), ("partitionNumber", Int3)]


class Surface(C.Structure):
    _fields_ = [
        ("vertices", C.POINTER(Double3)),
        ("numberOfVertices", C.c_int),
        ("triangles", C.POINTER(Int3)),
        ("numberOfTriangles", C.c_int),
    ]


class Layer(C.Structure):
    _fields_ = [
        ("refractiveIndex", C.c_double),
        ("absorptionCoefficient", C.c_double),
        ("scatteringCoefficient", C.c_double),
        ("anisotropy", C.c_double),
        ("surfaceId", C.POINTER(C.c_int)),
        ("numberOfSurfaces", C.c_int),
    ]


class CubeDetector(C.Structure):
    _fields_ = [
        ("center", Double3),
        ("length", Double3),
        ("filterLayers", C.c_uint8 * MAX_LAYERS),
        ("targetAngle", C.c_double),
    ]


class RingDetector(C.Structure):
    _fields_ = [
        ("center", Double3),
        ("smallRadius", C.c_double),
        ("bigRadius", C.c_double),
        ("filterLayers", C.c_uint8 * MAX_LAYERS),
        ("targetAngle", C.c_double),
    ]


class WeightTable(C.Structure):
    _fields_ = [("anisotropy", C.c_double), ("numberOfElements", C.c_int), ("elements", C.POINTER(C.c_double))]


class Input(C.Structure):
    _fields_ = [
        ("numberOfPhotons", C.c_uint64),
        ("area", C.POINTER(Area)),
        ("layerInfo", C.POINTER(Layer)),
        ("numberOfLayers", C.c_int),
        ("surface", C.POINTER(Surface)),
        ("numberOfSurfaces", C.c_int),
        ("cubeDetector", C.POINTER(CubeDetector)),
        ("numberOfCubeDetectors", C.c_int),
        ("ringDetector", C.POINTER(RingDetector)),
        ("numberOfRingDetectors", C.c_int),
        ("minWeight", C.c_double),
        ("timeStart", C.c_double),
        ("timeFinish", C.c_double),
        ("timeScaleSize", C.c_int),
        ("startPosition", Double3),
        ("startDirection", Double3),
        ("targetRangeLayers", C.c_uint8 * MAX_LAYERS),
        ("terminationMode", C.c_int),
        ("rouletteChance", C.c_double),
        ("useBiasing", C.c_int),
        ("attractiveFactor", C.c_double),
        ("targetPoint", Double3),
        ("weightIntegralPrecision", C.c_int),
        ("weightTablePrecision", C.c_int),
        ("weightTable", C.POINTER(WeightTable)),
        ("numberOfWeightTables", C.c_int),
    ]


class TimeInfo(C.Structure):
    _fields_ = [
        ("timeStart", C.c_double),
        ("timeFinish", C.c_double),
        ("numberOfPhotons", C.c_uint64),
        ("weight", C.c_double),
        ("targetWeight", C.c_double),
    ]


TIME_INFO_DTYPE = np.dtype([("timeStart", "<f8"), ("timeFinish", "<f8"), ("numberOfPhotons", "<u8"),
                            ("weight", "<f8"), ("targetWeight", "<f8")])
assert TIME_INFO_DTYPE.itemsize == C.sizeof(TimeInfo)


class DetectorInfo(C.Structure):
    _fields_ = [
        ("numberOfPhotons", C.c_uint64),
        ("trajectory", C.POINTER(C.c_uint64)),
        ("trajectorySize", C.c_int),
        ("timeScale", C.POINTER(TimeInfo)),
        ("timeScaleSize", C.c_int),
        ("targetWeight", C.c_double),
        ("weight", C.c_double),
    ]


class Output(C.Structure):
    _fields_ = [
        ("numberOfPhotons", C.c_uint64),
        ("specularReflectance", C.c_double),
        ("absorption", C.POINTER(C.c_double)),
        ("scattering", C.POINTER(C.c_double)),
        ("gridSize", C.c_int),
        ("detectorInfo", C.POINTER(DetectorInfo)),
        ("numberOfDetectors", C.c_int),
    ]


class Rng(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("nextPhoton", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [
        ("photons", C.c_uint64),
        ("scatterEvents", C.c_uint64),
        ("boundaryEvents", C.c_uint64),
        ("rouletteEvents", C.c_uint64),
        ("stackSpills", C.c_uint64),
        ("culledWalks", C.c_uint64),
        ("kernelMs", C.c_double),
        ("wAbsorbed", C.c_double),
        ("wEscaped", C.c_double),
        ("wDetected", C.c_double),
        ("wLeftArea", C.c_double),
        ("wKilled", C.c_double),
        ("wRouletteLost", C.c_double),
        ("wRouletteGain", C.c_double),
        ("wOther", C.c_double),
        ("trajectoryOverflows", C.c_uint64),
        ("entryWalks", C.c_uint64),
    ]


class TraceEvent(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("layer", C.c_int32),
        ("aux", C.c_int32),
        ("draws", C.c_uint32),
        ("pos", C.c_double * 3),
        ("dir", C.c_double * 3),
        ("weight", C.c_double),
        ("time", C.c_double),
    ]


TRACE_DTYPE = np.dtype(
    [
        ("type", "<i4"),
        ("layer", "<i4"),
        ("aux", "<i4"),
        ("draws", "<u4"),
        ("pos", "<f8", (3,)),
        ("dir", "<f8", (3,)),
        ("weight", "<f8"),
        ("time", "<f8"),
    ]
)
assert TRACE_DTYPE.itemsize == C.sizeof(TraceEvent)


class Options(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("blockSize", C.c_int),
        ("blocksPerCU", C.c_int),
        ("skipAbsorption", C.c_int),
        ("noDistanceGrid", C.c_int),
        ("scattering", C.c_int),
        ("trajectories", C.c_int),
        ("trajectoryCapacity", C.c_int),
        ("noEntryGrid", C.c_int),
        ("noSlabGrid", C.c_int),
        ("noLatticeRule", C.c_int),
    ]


def _d3(v) -> Double3:
    return Double3(float(v[0]), float(v[1]), float(v[2]))


@dataclass
class LayerSpec:
    n: float
    mua: float
    mus: float
    g: float
    surfaces: Sequence[int]


@dataclass
class RingSpec:
    r_small: float
    r_big: float
    center: Sequence[float] = (0.0, 0.0, 0.0)
    filter_layers: Sequence[int] = tuple(range(1, MAX_LAYERS))
    target_angle: float = 90.0


@dataclass
class CubeSpec:
    center: Sequence[float]
    length: Sequence[float]
    filter_layers: Sequence[int] = tuple(range(1, MAX_LAYERS))
    target_angle: float = 90.0


@dataclass
class SceneSpec:
    """A complete xmcml input: optics, geometry, detectors, time gates."""

    layers: List[LayerSpec]
    surfaces: List[tuple]  # (vertices float64 (V,3), triangles int32 (T,3))
    area_corner: Sequence[float]
    area_length: Sequence[float]
    area_partitions: Sequence[int]
    rings: List[RingSpec] = field(default_factory=list)
    cubes: List[CubeSpec] = field(default_factory=list)
    min_weight: float = 1e-4
    termination: int = TERMINATE_ROULETTE
    roulette_chance: float = 0.1
    time_start: float = 0.0
    time_finish: float = 1.0
    time_bins: int = 1
    start_position: Sequence[float] = (0.0, 0.0, 0.0)
    start_direction: Sequence[float] = (0.0, 0.0, 1.0)
    target_range_layers: Sequence[int] = ()
    photons: int = 0
    name: str = "scene"
    # attractive-vector biasing (InputInfo.useBiasing..targetPoint); weight_tables:
    # [(anisotropy, elements)], empty = computed by the library
    use_biasing: bool = False
    attractive_factor: float = 1.0
    target_point: Sequence[float] = (0.0, 0.0, 0.0)
    weight_integral_precision: int = 128
    weight_table_precision: int = 262144
    weight_tables: List[tuple] = field(default_factory=list)

    @property
    def grid_size(self) -> int:
        p = self.area_partitions
        return int(p[0]) * int(p[1]) * int(p[2])

    @property
    def num_detectors(self) -> int:
        return len(self.cubes) + len(self.rings)

    @property
    def num_triangles(self) -> int:
        return sum(int(t.shape[0]) for _, t in self.surfaces)


class InputHandle:
    """Builds an ``xmc_input`` from a SceneSpec and owns every array behind it."""

    def __init__(self, spec: SceneSpec, photons: Optional[int] = None):
        self.spec = spec
        keep = []
        self._area = Area(_d3(spec.area_corner), _d3(spec.area_length), Int3(*[int(v) for v in spec.area_partitions]))
        surfs = (Surface * max(1, len(spec.surfaces)))()
        for i, (v, t) in enumerate(spec.surfaces):
            v = np.ascontiguousarray(v, dtype=np.float64)
            t = np.ascontiguousarray(t, dtype=np.int32)
            keep += [v, t]
            surfs[i].vertices = v.ctypes.data_as(C.POINTER(Double3))
            surfs[i].numberOfVertices = v.shape[0]
            surfs[i].triangles = t.ctypes.data_as(C.POINTER(Int3))
            surfs[i].numberOfTriangles = t.shape[0]
        layers = (Layer * len(spec.layers))()
        for i, l in enumerate(spec.layers):
            ids = np.ascontiguousarray(list(l.surfaces), dtype=np.int32)
            keep.append(ids)
            layers[i].refractiveIndex = l.n
            layers[i].absorptionCoefficient = l.mua
            layers[i].scatteringCoefficient = l.mus
            layers[i].anisotropy = l.g
            layers[i].surfaceId = ids.ctypes.data_as(C.POINTER(C.c_int))
            layers[i].numberOfSurfaces = len(ids)
        rings = (RingDetector * max(1, len(spec.rings)))()
        for i, r in enumerate(spec.rings):
            rings[i].center = _d3(r.center)
            rings[i].smallRadius = r.r_small
            rings[i].bigRadius = r.r_big
            for j in r.filter_layers:
                rings[i].filterLayers[j] = 1
            rings[i].targetAngle = r.target_angle
        cubes = (CubeDetector * max(1, len(spec.cubes)))()
        for i, c in enumerate(spec.cubes):
            cubes[i].center = _d3(c.center)
            cubes[i].length = _d3(c.length)
            for j in c.filter_layers:
                cubes[i].filterLayers[j] = 1
            cubes[i].targetAngle = c.target_angle
        inp = Input()
        inp.numberOfPhotons = int(photons if photons is not None else spec.photons)
        inp.area = C.pointer(self._area)
        inp.layerInfo = layers
        inp.numberOfLayers = len(spec.layers)
        inp.surface = surfs
        inp.numberOfSurfaces = len(spec.surfaces)
        inp.cubeDetector = cubes
        inp.numberOfCubeDetectors = len(spec.cubes)
        inp.ringDetector = rings
        inp.numberOfRingDetectors = len(spec.rings)
        inp.minWeight = spec.min_weight
        inp.timeStart = spec.time_start
        inp.timeFinish = spec.time_finish
        inp.timeScaleSize = spec.time_bins
        inp.startPosition = _d3(spec.start_position)
        inp.startDirection = _d3(spec.start_direction)
        for j in spec.target_range_layers:
            inp.targetRangeLayers[j] = 1
        inp.terminationMode = spec.termination
        inp.rouletteChance = spec.roulette_chance
        inp.useBiasing = int(bool(spec.use_biasing))
        inp.attractiveFactor = float(spec.attractive_factor)
        inp.targetPoint = _d3(spec.target_point)
        inp.weightIntegralPrecision = int(spec.weight_integral_precision)
        inp.weightTablePrecision = int(spec.weight_table_precision)
        if spec.weight_tables:
            tabs = (WeightTable * len(spec.weight_tables))()
            for i, (g, el) in enumerate(spec.weight_tables):
                el = np.ascontiguousarray(el, dtype=np.float64)
                keep.append(el)
                tabs[i].anisotropy = float(g)
                tabs[i].numberOfElements = el.size
                tabs[i].elements = el.ctypes.data_as(C.POINTER(C.c_double))
            inp.weightTable = tabs
            inp.numberOfWeightTables = len(spec.weight_tables)
            self._tabs = tabs
        self._keep = keep
        self._surfs, self._layers, self._rings, self._cubes = surfs, layers, rings, cubes
        self.c = inp

    @property
    def ptr(self):
        return C.byref(self.c)


class OutputHandle:
    """``xmc_output`` with numpy storage, time bins laid out as InitOutput does
    (base/xmcml/xmcmlLauncher/launcher_omp.cpp:10-58)."""

    def __init__(self, spec: SceneSpec, absorption: bool = True, scattering: bool = False,
                 trajectory: bool = False):
        G = spec.grid_size
        nd = spec.num_detectors
        T = spec.time_bins
        self.absorption = np.zeros(G, np.float64) if absorption else None
        self.scattering = np.zeros(G, np.float64) if scattering else None
        self.trajectory = np.zeros((max(nd, 1), G), np.uint64) if trajectory else None
        self.bins = (TimeInfo * max(1, nd * T))()
        dt = (spec.time_finish - spec.time_start) / T if T > 0 else 0.0
        if dt < 0.0:
            dt = 0.0
        self.dets = (DetectorInfo * max(1, nd))()
        for d in range(nd):
            base = d * T
            for j in range(T - 1):
                self.bins[base + j].timeStart = spec.time_start + j * dt
                self.bins[base + j].timeFinish = self.bins[base + j].timeStart + dt
            if T > 0:
                self.bins[base + T - 1].timeStart = spec.time_start + (T - 1) * dt
                self.bins[base + T - 1].timeFinish = spec.time_finish
            self.dets[d].timeScale = C.cast(C.byref(self.bins, base * C.sizeof(TimeInfo)), C.POINTER(TimeInfo))
            self.dets[d].timeScaleSize = T
            self.dets[d].trajectorySize = G
            if self.trajectory is not None:
                self.dets[d].trajectory = self.trajectory[d].ctypes.data_as(C.POINTER(C.c_uint64))
        out = Output()
        out.gridSize = G
        out.numberOfDetectors = nd
        out.detectorInfo = self.dets
        if self.absorption is not None:
            out.absorption = self.absorption.ctypes.data_as(C.POINTER(C.c_double))
        if self.scattering is not None:
            out.scattering = self.scattering.ctypes.data_as(C.POINTER(C.c_double))
        self.c = out
        self.nd, self.T = nd, T

    @property
    def ptr(self):
        return C.byref(self.c)

    # numpy views of the detector tallies
    def det_weight(self) -> np.ndarray:
        return np.array([self.dets[d].weight for d in range(self.nd)])

    def det_target(self) -> np.ndarray:
        return np.array([self.dets[d].targetWeight for d in range(self.nd)])

    def det_count(self) -> np.ndarray:
        return np.array([self.dets[d].numberOfPhotons for d in range(self.nd)], dtype=np.uint64)

    def _bins(self, name, dtype=np.float64) -> np.ndarray:
        rec = np.frombuffer(self.bins, dtype=TIME_INFO_DTYPE, count=self.nd * self.T)
        return rec[name].astype(dtype).reshape(self.nd, self.T)

    def bin_weight(self) -> np.ndarray:
        return self._bins("weight")

    def bin_target(self) -> np.ndarray:
        return self._bins("targetWeight")

    def bin_count(self) -> np.ndarray:
        return self._bins("numberOfPhotons", np.uint64)

    def bin_edges(self) -> np.ndarray:
        return np.array([[self.bins[i].timeStart, self.bins[i].timeFinish] for i in range(self.nd * self.T)])

    @property
    def specular(self) -> float:
        return self.c.specularReflectance
