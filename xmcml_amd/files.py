"""xmcml's file formats through the C-ABI (include/xmcml_io.h).

The reference launcher (base/xmcml/xmcmlLauncher/main.cpp:63-177) reads an XML
<Input> description and a binary .surface file, optionally resumes from a
.mcml.bk backup, runs the photons in backup portions and writes a .mcml.out
result.  ``run_files`` is that flow on the GPU engine; the readers and writers
themselves are C++ in the library (xmcml_amd/csrc/xmc_io.cpp), this module only
binds them and converts documents to and from ``SceneSpec``.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional

import numpy as np

from . import engine
from .abi import (MAX_LAYERS, TERMINATE_KILL, CubeSpec, Double3, Input, InputHandle, LayerSpec, Options,
                  OutputHandle, RingSpec, Rng, SceneSpec)


class Biasing(C.Structure):
    _fields_ = [
        ("useBiasing", C.c_int),
        ("weightIntegralPrecision", C.c_int),
        ("weightTablePrecision", C.c_int),
        ("attractiveFactor", C.c_double),
        ("targetPoint", Double3),
    ]


def _b(path) -> bytes:
    return str(path).encode()


class Document:
    """An input read from an XML file (and a .surface file): owns the C
    document; ``input`` is its ``xmc_input``."""

    def __init__(self, xml_path, surface_path=None):
        L = engine.lib()
        h = C.c_void_p()
        engine._check(L.xmc_read_input_file(_b(xml_path), C.byref(h)), f"xmc_read_input_file({xml_path})")
        self._h = h
        if surface_path is not None:
            engine._check(L.xmc_read_surface_file(h, _b(surface_path)), f"xmc_read_surface_file({surface_path})")

    @property
    def input(self) -> Input:
        return C.cast(engine.lib().xmc_document_input(self._h), C.POINTER(Input)).contents

    @property
    def ptr(self):
        return engine.lib().xmc_document_input(self._h)

    def read_weight_table_file(self, path) -> None:
        """ParseWeightIntegralFile: the file's tables become the input's
        weightTable (the file must match the input's attractive factor and
        precisions and cover every distinct anisotropy of layers 1..)."""
        engine._check(engine.lib().xmc_read_weight_table_file(self._h, _b(path)),
                      f"xmc_read_weight_table_file({path})")

    def biasing(self) -> Biasing:
        b = Biasing()
        engine._check(engine.lib().xmc_document_biasing(self._h, C.byref(b)), "xmc_document_biasing")
        return b

    def close(self):
        if self._h:
            engine.lib().xmc_document_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def spec(self, termination: int = TERMINATE_KILL, roulette_chance: float = 0.1) -> SceneSpec:
        """The document as a SceneSpec (xmcml files carry no roulette: the
        reference kills photons below MinWeight)."""
        return input_to_spec(self.input, termination, roulette_chance)


def input_to_spec(x: Input, termination: int = TERMINATE_KILL, roulette_chance: float = 0.1) -> SceneSpec:
    """An ``xmc_input`` (from any source) as a SceneSpec, copying every array."""
    layers = []
    for i in range(x.numberOfLayers):
        l = x.layerInfo[i]
        layers.append(LayerSpec(l.refractiveIndex, l.absorptionCoefficient, l.scatteringCoefficient, l.anisotropy,
                                [l.surfaceId[j] for j in range(l.numberOfSurfaces)]))
    surfaces = []
    if x.surface:
        for i in range(x.numberOfSurfaces):
            s = x.surface[i]
            v = np.ctypeslib.as_array(C.cast(s.vertices, C.POINTER(C.c_double)), (s.numberOfVertices * 3,))
            t = np.ctypeslib.as_array(C.cast(s.triangles, C.POINTER(C.c_int)), (s.numberOfTriangles * 3,))
            surfaces.append((v.reshape(-1, 3).copy(), t.reshape(-1, 3).astype(np.int32)))

    def filt(f):
        return [j for j in range(MAX_LAYERS) if f[j]]

    rings = [RingSpec(r.smallRadius, r.bigRadius, (r.center.x, r.center.y, r.center.z), filt(r.filterLayers),
                      r.targetAngle) for r in (x.ringDetector[i] for i in range(x.numberOfRingDetectors))]
    cubes = [CubeSpec((c.center.x, c.center.y, c.center.z), (c.length.x, c.length.y, c.length.z),
                      filt(c.filterLayers), c.targetAngle)
             for c in (x.cubeDetector[i] for i in range(x.numberOfCubeDetectors))]
    a = x.area.contents
    return SceneSpec(
        layers=layers, surfaces=surfaces,
        area_corner=(a.corner.x, a.corner.y, a.corner.z), area_length=(a.length.x, a.length.y, a.length.z),
        area_partitions=(a.partitionNumber.x, a.partitionNumber.y, a.partitionNumber.z),
        rings=rings, cubes=cubes, min_weight=x.minWeight, termination=termination,
        roulette_chance=roulette_chance, time_start=x.timeStart, time_finish=x.timeFinish,
        time_bins=x.timeScaleSize, start_position=(x.startPosition.x, x.startPosition.y, x.startPosition.z),
        start_direction=(x.startDirection.x, x.startDirection.y, x.startDirection.z),
        target_range_layers=[j for j in range(MAX_LAYERS) if x.targetRangeLayers[j]],
        photons=int(x.numberOfPhotons), name="file",
        use_biasing=bool(x.useBiasing), attractive_factor=x.attractiveFactor,
        target_point=(x.targetPoint.x, x.targetPoint.y, x.targetPoint.z),
        weight_integral_precision=x.weightIntegralPrecision, weight_table_precision=x.weightTablePrecision,
        weight_tables=[(x.weightTable[i].anisotropy,
                        np.ctypeslib.as_array(x.weightTable[i].elements, (x.weightTable[i].numberOfElements,)).copy())
                       for i in range(x.numberOfWeightTables if x.weightTable else 0)])


def _xyz(tag: str, v, ind: str) -> str:
    return (f"{ind}<{tag}>\n{ind}    <X>{v[0]!r}</X>\n{ind}    <Y>{v[1]!r}</Y>\n{ind}    <Z>{v[2]!r}</Z>\n"
            f"{ind}</{tag}>\n")


def write_input_xml(spec: SceneSpec, path, photons: Optional[int] = None) -> None:
    """A SceneSpec as an xmcml <Input> file (the elements parser.cpp:57-486
    reads; floats written with repr so they read back exactly).  Detectors
    are written in the spec's order; the readers sort them by filter count."""
    n = int(photons if photons is not None else spec.photons)
    o = ['<?xml version="1.0"?>\n<Input>\n', f"    <NumberOfPhotons>{n}</NumberOfPhotons>\n",
         f"    <MinWeight>{float(spec.min_weight)!r}</MinWeight>\n",
         f"    <TimeStart>{float(spec.time_start)!r}</TimeStart>\n",
         f"    <TimeFinish>{float(spec.time_finish)!r}</TimeFinish>\n",
         f"    <TimeScaleSize>{int(spec.time_bins)}</TimeScaleSize>\n"]
    if spec.target_range_layers:
        o.append("    <TargetRange>\n" + "".join(f"        <Layer>{int(j)}</Layer>\n"
                                                for j in spec.target_range_layers) + "    </TargetRange>\n")
    if spec.use_biasing:
        t = [float(v) for v in spec.target_point]
        if any(v != int(v) for v in t):
            raise ValueError("the reader takes TargetPoint coordinates as integers (atoi, parser.cpp:233-241)")
        o.append("    <Biasing>\n        <WeightIntegral>\n"
                 f"            <IntegralPrecision>{int(spec.weight_integral_precision)}</IntegralPrecision>\n"
                 f"            <TablePrecision>{int(spec.weight_table_precision)}</TablePrecision>\n"
                 f"            <AttractiveFactor>{float(spec.attractive_factor)!r}</AttractiveFactor>\n"
                 "        </WeightIntegral>\n        <TargetPoint>\n"
                 + "".join(f"            <{a}>{int(v)}</{a}>\n" for a, v in zip("XYZ", t))
                 + "        </TargetPoint>\n    </Biasing>\n")
    o.append(_xyz("StartPosition", [float(v) for v in spec.start_position], "    "))
    o.append(_xyz("StartDirection", [float(v) for v in spec.start_direction], "    "))
    o.append("    <Area>\n" + _xyz("Corner", [float(v) for v in spec.area_corner], "        ")
             + _xyz("Length", [float(v) for v in spec.area_length], "        ")
             + "        <PartitionNumber>\n"
             + "".join(f"            <{a}>{int(v)}</{a}>\n" for a, v in zip("XYZ", spec.area_partitions))
             + "        </PartitionNumber>\n    </Area>\n")
    o.append(f"    <NumberOfLayers>{len(spec.layers)}</NumberOfLayers>\n    <Layers>\n")
    for l in spec.layers:
        o.append("        <Layer>\n"
                 f"            <RefractiveIndex>{float(l.n)!r}</RefractiveIndex>\n"
                 f"            <AbsorptionCoefficient>{float(l.mua)!r}</AbsorptionCoefficient>\n"
                 f"            <ScatteringCoefficient>{float(l.mus)!r}</ScatteringCoefficient>\n"
                 f"            <Anisotropy>{float(l.g)!r}</Anisotropy>\n"
                 f"            <NumberOfSurfaces>{len(l.surfaces)}</NumberOfSurfaces>\n"
                 "            <SurfaceIds>\n"
                 + "".join(f"                <Id>{int(s)}</Id>\n" for s in l.surfaces)
                 + "            </SurfaceIds>\n        </Layer>\n")
    o.append("    </Layers>\n")

    def filt(f, ind):
        if list(f) == list(range(MAX_LAYERS)):
            return ""
        return f"{ind}<Filter>\n" + "".join(f"{ind}    <Layer>{int(j)}</Layer>\n" for j in f) + f"{ind}</Filter>\n"

    if spec.cubes:
        o.append(f"    <NumberOfCubeDetectors>{len(spec.cubes)}</NumberOfCubeDetectors>\n    <CubeDetectors>\n")
        for c in spec.cubes:
            o.append("        <CubeDetector>\n" + _xyz("Center", [float(v) for v in c.center], "            ")
                     + _xyz("Length", [float(v) for v in c.length], "            ")
                     + filt(c.filter_layers, "            ")
                     + f"            <TargetAngle>{float(c.target_angle)!r}</TargetAngle>\n        </CubeDetector>\n")
        o.append("    </CubeDetectors>\n")
    if spec.rings:
        o.append(f"    <NumberOfRingDetectors>{len(spec.rings)}</NumberOfRingDetectors>\n    <RingDetectors>\n")
        for r in spec.rings:
            o.append("        <RingDetector>\n" + _xyz("Center", [float(v) for v in r.center], "            ")
                     + f"            <SmallRadius>{float(r.r_small)!r}</SmallRadius>\n"
                     f"            <BigRadius>{float(r.r_big)!r}</BigRadius>\n"
                     + filt(r.filter_layers, "            ")
                     + f"            <TargetAngle>{float(r.target_angle)!r}</TargetAngle>\n        </RingDetector>\n")
        o.append("    </RingDetectors>\n")
    o.append(f"    <NumberOfSurfaces>{len(spec.surfaces)}</NumberOfSurfaces>\n</Input>\n")
    Path(path).write_text("".join(o))


def write_surface_file(spec: SceneSpec, path) -> None:
    inp = InputHandle(spec)
    engine._check(engine.lib().xmc_write_surface_file(inp.c.surface, inp.c.numberOfSurfaces, _b(path)),
                  "xmc_write_surface_file")


def write_output_file(inp, out: OutputHandle, path) -> None:
    """WriteOutputToFile of an input (InputHandle or Document) and output."""
    engine._check(engine.lib().xmc_write_output_file(_ptr(inp), out.ptr, _b(path)), "xmc_write_output_file")


def write_backup_file(inp, out: OutputHandle, rng: Rng, path=None) -> None:
    engine._check(engine.lib().xmc_write_backup_file(_ptr(inp), out.ptr, C.byref(rng),
                                                     None if path is None else _b(path)), "xmc_write_backup_file")


def write_weight_table_file(inp, path=None) -> None:
    """WriteWeightIntegralTableToFile of the input's weightTable (path None =
    "xmcml_weight_integral_table.bin" in the working directory)."""
    engine._check(engine.lib().xmc_write_weight_table_file(_ptr(inp), None if path is None else _b(path)),
                  "xmc_write_weight_table_file")


def backup_file_name(photons: int) -> str:
    buf = C.create_string_buffer(64)
    engine._check(engine.lib().xmc_backup_file_name(int(photons), buf, 64), "xmc_backup_file_name")
    return buf.value.decode()


def read_backup_file(path, out: OutputHandle, with_rng: bool = True) -> Optional[Rng]:
    """Fills `out` from a backup (or, with_rng=False, a .mcml.out result)."""
    rng = Rng()
    engine._check(engine.lib().xmc_read_backup_file(_b(path), out.ptr, C.byref(rng) if with_rng else None),
                  f"xmc_read_backup_file({path})")
    return rng if with_rng else None


def launch_with_backups(inp, out: OutputHandle, rng: Rng, portion: int, backup_path=None,
                        options: Optional[Options] = None) -> None:
    engine._check(engine.lib().xmc_launch_with_backups(_ptr(inp), out.ptr, C.byref(rng),
                                                       None if options is None else C.byref(options), int(portion),
                                                       None if backup_path is None else _b(backup_path)),
                  "xmc_launch_with_backups")


def _ptr(inp):
    """The xmc_input pointer of an InputHandle or a Document."""
    return inp.ptr


WEIGHT_TABLE_FILE = "xmcml_weight_integral_table.bin"


def get_weight_tables(doc: Document, weight_table_path=None, device: int = 0) -> None:
    """GetWeightIntegralTable (weight_integral.cpp:16-31) for a document that
    asks for biasing: read the tables from a table file, or compute them on
    the GPU for every distinct anisotropy at the document's precisions, write
    them to "xmcml_weight_integral_table.bin" in the working directory and
    take them from there."""
    if weight_table_path is not None:
        doc.read_weight_table_file(weight_table_path)
        return
    from .scenes import distinct_anisotropies
    x = doc.input
    spec = doc.spec()
    spec.weight_tables = [(g, engine.weight_table(g, x.attractiveFactor, x.weightIntegralPrecision,
                                                  x.weightTablePrecision, device))
                          for g in distinct_anisotropies(spec)]
    write_weight_table_file(InputHandle(spec), None)
    doc.read_weight_table_file(WEIGHT_TABLE_FILE)


def run_files(xml_path, surface_path, output_path, seed: int = 777, portion: int = 10_000_000,
              backup_path=None, resume_from=None, options: Optional[Options] = None,
              weight_table_path=None) -> OutputHandle:
    """main.cpp:63-177 on the GPU: parse, take the biasing's weight tables
    (from weight_table_path, the launcher's -t file, or computed and written
    to the working directory), optionally resume from a backup, run in backup
    portions, write the result file.  Returns the output."""
    doc = Document(xml_path, surface_path)
    x = doc.input
    if x.useBiasing:
        get_weight_tables(doc, weight_table_path)
    spec = doc.spec()
    # every tally InitOutput allocates (launcher_omp.cpp:10-58), trajectory maps included
    out = OutputHandle(spec, absorption=True, scattering=True, trajectory=True)
    rng = Rng(seed, 0)
    total = int(x.numberOfPhotons)
    if resume_from is not None:
        rng = read_backup_file(resume_from, out)
        if total <= out.c.numberOfPhotons:
            raise engine.EngineError("number of photons in the backup is not below the input's: nothing to do")
        x.numberOfPhotons = total - out.c.numberOfPhotons  # main.cpp:127
    launch_with_backups(doc, out, rng, portion, backup_path, options)
    x.numberOfPhotons = out.c.numberOfPhotons  # main.cpp:166
    write_output_file(doc, out, output_path)
    x.numberOfPhotons = total
    doc.close()
    return out
