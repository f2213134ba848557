"""Helpers of the file-format tests: the reference's own readers and writers
(oracle/_ref/libxmcml_ref_io.so, built from the reference sources by
oracle/Makefile) and a canonical dump of a parsed input.  Test infrastructure."""
from __future__ import annotations

import ctypes as C
import hashlib
from pathlib import Path

import numpy as np

from xmcml_amd.abi import MAX_LAYERS, Input
from xmcml_amd.files import Biasing

ROOT = Path(__file__).resolve().parents[1]
REF_IO_SO = ROOT / "oracle" / "_ref" / "libxmcml_ref_io.so"
FIXTURES = ROOT / "tests" / "golden" / "files"

_ref_io = None


def ref_io_available() -> bool:
    return REF_IO_SO.exists()


def ref_io():
    global _ref_io
    if _ref_io is None:
        L = C.CDLL(str(REF_IO_SO))
        P, U64 = C.c_void_p, C.c_uint64
        L.refio_parse.restype = C.c_int
        L.refio_parse.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Input), C.POINTER(Biasing)]
        L.refio_write_output.restype = C.c_int
        L.refio_write_output.argtypes = [P, P, C.c_char_p]
        L.refio_write_backup.restype = C.c_int
        L.refio_write_backup.argtypes = [P, P, U64, U64]
        L.refio_read_backup.restype = C.c_int
        L.refio_read_backup.argtypes = [C.c_char_p, P, C.POINTER(U64), C.POINTER(U64)]
        L.refio_read_output.restype = C.c_int
        L.refio_read_output.argtypes = [C.c_char_p, P]
        _ref_io = L
    return _ref_io


def ref_parse(xml, surface=None):
    """(Input, Biasing) from the reference's ParseInputFile / ParseSurfaceFile,
    or None when the reference reports failure."""
    x, b = Input(), Biasing()
    ok = ref_io().refio_parse(str(xml).encode(), None if surface is None else str(surface).encode(),
                              C.byref(x), C.byref(b))
    return (x, b) if ok else None


def canonical(x: Input, b: Biasing) -> dict:
    """Every value of a parsed input, floats as their exact hex form."""
    h = float.hex

    def v3(v):
        return [h(v.x), h(v.y), h(v.z)]

    d = {
        "numberOfPhotons": int(x.numberOfPhotons), "minWeight": h(x.minWeight), "timeStart": h(x.timeStart),
        "timeFinish": h(x.timeFinish), "timeScaleSize": x.timeScaleSize, "startPosition": v3(x.startPosition),
        "startDirection": v3(x.startDirection), "targetRangeLayers": list(x.targetRangeLayers),
        "numberOfSurfaces": x.numberOfSurfaces,
        "area": None, "layers": [], "cubes": [], "rings": [], "surfaces": [],
        "biasing": {"useBiasing": b.useBiasing, "weightIntegralPrecision": b.weightIntegralPrecision,
                    "weightTablePrecision": b.weightTablePrecision, "attractiveFactor": h(b.attractiveFactor),
                    "targetPoint": v3(b.targetPoint)},
    }
    if x.area:
        a = x.area.contents
        d["area"] = {"corner": v3(a.corner), "length": v3(a.length),
                     "partitions": [a.partitionNumber.x, a.partitionNumber.y, a.partitionNumber.z]}
    for i in range(x.numberOfLayers):
        l = x.layerInfo[i]
        d["layers"].append([h(l.refractiveIndex), h(l.absorptionCoefficient), h(l.scatteringCoefficient),
                            h(l.anisotropy), [l.surfaceId[j] for j in range(l.numberOfSurfaces)]])
    for i in range(x.numberOfCubeDetectors):
        c = x.cubeDetector[i]
        d["cubes"].append([v3(c.center), v3(c.length), list(c.filterLayers[:MAX_LAYERS]), h(c.targetAngle)])
    for i in range(x.numberOfRingDetectors):
        r = x.ringDetector[i]
        d["rings"].append([v3(r.center), h(r.smallRadius), h(r.bigRadius), list(r.filterLayers[:MAX_LAYERS]),
                           h(r.targetAngle)])
    if x.surface:
        for i in range(x.numberOfSurfaces):
            s = x.surface[i]
            v = np.ctypeslib.as_array(C.cast(s.vertices, C.POINTER(C.c_double)), (s.numberOfVertices * 3,))
            t = np.ctypeslib.as_array(C.cast(s.triangles, C.POINTER(C.c_int)), (s.numberOfTriangles * 3,))
            d["surfaces"].append([s.numberOfVertices, s.numberOfTriangles,
                                  hashlib.sha256(v.tobytes() + t.astype("<i4").tobytes()).hexdigest()])
    return d


def sha256(path) -> str:
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def filled_output(spec, seed: int):
    """An output laid out as InitOutput does for `spec`, every tally filled
    with seeded values (targetWeight <= weight, as the ranges check wants)."""
    from xmcml_amd.abi import OutputHandle
    rng = np.random.default_rng(seed)
    o = OutputHandle(spec, absorption=True, scattering=True, trajectory=True)
    o.c.numberOfPhotons = int(rng.integers(1, 1 << 40))
    o.c.specularReflectance = float(rng.random())
    o.absorption[:] = rng.random(o.absorption.size)
    o.scattering[:] = rng.random(o.scattering.size) * 7.0
    o.trajectory[:] = rng.integers(0, 1 << 50, o.trajectory.shape, dtype=np.uint64)
    for d in range(o.nd):
        w = float(rng.random() * 100.0)
        o.dets[d].weight = w
        o.dets[d].targetWeight = w * float(rng.random())
        o.dets[d].numberOfPhotons = int(rng.integers(0, 1 << 33))
        for j in range(o.T):
            t = o.bins[d * o.T + j]
            t.numberOfPhotons = int(rng.integers(0, 1 << 30))
            t.weight = float(rng.random())
            t.targetWeight = float(rng.random())
    return o
