"""xmcml's file formats (include/xmcml_io.h) against the reference's own
readers and writers: ParseInputFile / ParseSurfaceFile
(base/xmcml/xmcmlLauncher/parser.cpp:18-587), WriteOutputToFile /
WriteBackupToFile (writer.cpp:431-522) and ParseBackupFile
(parser.cpp:589-611, reader.cpp:27-392).

Pins, strongest first: the reference build oracle/_ref/libxmcml_ref_io.so
(its sources compiled unmodified) when present; SHA-256 digests and parsed
values that build wrote into tests/golden/file_vectors.json
(tests/golden/make_file_vectors.py), which hold without the reference tree.
The input fixtures are the reference's own example inputs
(branches/ovsyuhno/OCT/xmcml/diffuse_830_ring.xml and the exp.xml/exp.surface
pairs of branches/gerasimov) plus quirks.xml, written for these tests.
No GPU: every function here is host code."""
import ctypes as C
import json
import os
from pathlib import Path

import numpy as np
import pytest

import files_lib as fl
from xmcml_amd import engine, files, scenes
from xmcml_amd.abi import InputHandle, OutputHandle, Rng

GOLD = json.loads((fl.ROOT / "tests" / "golden" / "file_vectors.json").read_text())
INPUTS = {
    "quirks": ("quirks.xml", None),
    "diffuse_830_ring": ("diffuse_830_ring.xml", None),
    "cosine_exp": ("cosine_exp.xml", "cosine_exp.surface"),
    "stripes_exp": ("stripes_exp.xml", "stripes_exp.surface"),
}
needs_ref = pytest.mark.skipif(not fl.ref_io_available(), reason="reference build oracle/_ref is not present")


def ours(name):
    xml, surf = INPUTS[name]
    return files.Document(fl.FIXTURES / xml, None if surf is None else fl.FIXTURES / surf)


@pytest.mark.parametrize("name", list(INPUTS))
def test_parsed_input_matches_reference_golden(name):
    d = ours(name)
    assert fl.canonical(d.input, d.biasing()) == GOLD["parsed"][name]


@needs_ref
@pytest.mark.parametrize("name", list(INPUTS))
def test_parsed_input_matches_reference_build(name):
    xml, surf = INPUTS[name]
    r = fl.ref_parse(fl.FIXTURES / xml, None if surf is None else fl.FIXTURES / surf)
    d = ours(name)
    assert fl.canonical(d.input, d.biasing()) == fl.canonical(*r)


def test_quirks_the_reader_keeps():
    """parser.cpp semantics that look like bugs but define the format."""
    d = ours("quirks")
    x, b = d.input, d.biasing()
    assert x.numberOfPhotons == 123456789012  # atoll
    assert (b.targetPoint.x, b.targetPoint.y, b.targetPoint.z) == (1.0, -2.0, 3.0)  # atoi, :233-241
    assert list(x.targetRangeLayers) == [0, 0, 1, 1, 0, 0, 0, 0]  # layer 9 out of range: ignored
    assert (x.startDirection.x, x.startDirection.y, x.startDirection.z) == (0.3, 0.4, 2.0)  # not normalised
    # detectors ordered by number of filter layers, ties in file order (:489-509)
    assert [x.ringDetector[i].smallRadius for i in range(4)] == [1.0, 3.0, 2.0, 0.0]
    assert [sum(x.cubeDetector[i].filterLayers) for i in range(3)] == [1, 2, 8]
    # a detector record is reused across elements: the second cube keeps the first's Length and
    # Y/Z centre, the third the second's centre (parser.cpp:323-335)
    c = {sum(x.cubeDetector[i].filterLayers): x.cubeDetector[i] for i in range(3)}
    assert (c[2].center.x, c[2].center.y, c[2].length.x) == (-1.0, 2.0, 0.5)
    assert (c[1].center.x, c[1].length.x, c[8].targetAngle, c[2].targetAngle) == (-1.0, 4.0, 45.0, 90.0)


@pytest.mark.parametrize("bad, message", [
    ("<Input><NumberOfLayers>1</NumberOfLayers><Layers><Layer/><Layer/></Layers></Input>", b"number of layers"),
    ("<Input><Layers><Layer/></Layers></Input>", b"NumberOfLayers"),
    ("<Input><NumberOfPhotons>1</NumberOfPhotons>", b"missing </Input>"),
    ("<Input><NumberOfPhotons></NumberOfPhotons></Input>", b"has no text"),
    ("<Input><NumberOfRingDetectors>2</NumberOfRingDetectors><RingDetectors><RingDetector/></RingDetectors>"
     "<Area/><NumberOfLayers>0</NumberOfLayers><Layers/></Input>", b"fewer <RingDetector>"),
    ("<Input><NumberOfCubeDetectors>1</NumberOfCubeDetectors><CubeDetectors><CubeDetector><Filter>"
     "<Layer>8</Layer></Filter></CubeDetector></CubeDetectors></Input>", b"FilterLayers invalid"),
    ("<Input><A></B></Input>", b"closes <A>"),
])
def test_malformed_inputs_are_rejected(tmp_path, bad, message):
    p = tmp_path / "bad.xml"
    p.write_text(bad)
    h = C.c_void_p()
    assert engine.lib().xmc_read_input_file(str(p).encode(), C.byref(h)) == -1
    assert message in engine.lib().xmc_last_error()


def test_surface_file_round_trip(tmp_path):
    spec = scenes.heightfield_scene(0)
    path = tmp_path / "h.surface"
    files.write_surface_file(spec, path)
    xml = tmp_path / "h.xml"
    files.write_input_xml(spec, xml, photons=1000)
    back = files.Document(xml, path).spec(termination=spec.termination, roulette_chance=spec.roulette_chance)
    assert len(back.surfaces) == len(spec.surfaces)
    for (v0, t0), (v1, t1) in zip(spec.surfaces, back.surfaces):
        assert np.array_equal(np.asarray(v0, np.float64), v1) and np.array_equal(np.asarray(t0, np.int32), t1)
    for a, b in zip(spec.layers, back.layers):
        assert (a.n, a.mua, a.mus, a.g, list(a.surfaces)) == (b.n, b.mua, b.mus, b.g, list(b.surfaces))
    assert [r.r_small for r in back.rings] == [r.r_small for r in spec.rings]
    assert (back.time_start, back.time_finish, back.time_bins) == (spec.time_start, spec.time_finish,
                                                                    spec.time_bins)
    # the surface count must match the input's NumberOfSurfaces (parser.cpp:570-572)
    doc = files.Document(fl.FIXTURES / "quirks.xml")
    with pytest.raises(engine.EngineError, match="surface count"):
        engine._check(engine.lib().xmc_read_surface_file(doc._h, str(path).encode()), "read")


@needs_ref
def test_written_surface_file_loads_in_the_reference(tmp_path):
    spec = scenes.slab_scene(0)
    files.write_surface_file(spec, tmp_path / "s.surface")
    files.write_input_xml(spec, tmp_path / "s.xml", photons=10)
    r = fl.ref_parse(tmp_path / "s.xml", tmp_path / "s.surface")
    assert r is not None
    d = files.Document(tmp_path / "s.xml", tmp_path / "s.surface")
    assert fl.canonical(d.input, d.biasing()) == fl.canonical(*r)


def _quirks_output():
    d = ours("quirks")
    spec = d.spec()
    return d, fl.filled_output(spec, GOLD["writers"]["seed"])


def test_output_and_backup_bytes_match_reference_golden(tmp_path, monkeypatch):
    d, o = _quirks_output()
    files.write_output_file(d, o, tmp_path / "result.mcml.out")
    assert fl.sha256(tmp_path / "result.mcml.out") == GOLD["writers"]["output_sha256"]
    monkeypatch.chdir(tmp_path)
    seed, nxt = GOLD["writers"]["rng"]
    files.write_backup_file(d, o, Rng(seed, nxt))  # the reference's own file name
    name = GOLD["writers"]["backup_name"]
    assert files.backup_file_name(d.input.numberOfPhotons) == name
    assert fl.sha256(tmp_path / name) == GOLD["writers"]["backup_sha256"]


@needs_ref
def test_output_and_backup_bytes_match_reference_build(tmp_path, monkeypatch):
    d, o = _quirks_output()
    o.c.numberOfPhotons = 7  # the section holds input->numberOfPhotons, not the output's
    files.write_output_file(d, o, tmp_path / "ours.mcml.out")
    assert fl.ref_io().refio_write_output(d.ptr, o.ptr, str(tmp_path / "ref.mcml.out").encode()) == 1
    assert (tmp_path / "ours.mcml.out").read_bytes() == (tmp_path / "ref.mcml.out").read_bytes()
    for n in (999, 1000, 25_000, 1_500_000, 2_000_000_000):
        d.input.numberOfPhotons = n
        sub = tmp_path / str(n)
        sub.mkdir()
        monkeypatch.chdir(sub)
        assert fl.ref_io().refio_write_backup(d.ptr, o.ptr, 5, 77) == 1
        (ref_name,) = os.listdir(sub)
        assert files.backup_file_name(n) == ref_name
        files.write_backup_file(d, o, Rng(5, 77), sub / "ours.bk")
        assert (sub / "ours.bk").read_bytes() == (sub / ref_name).read_bytes()


def _same_outputs(a: OutputHandle, b: OutputHandle):
    assert a.c.numberOfPhotons == b.c.numberOfPhotons
    assert np.array_equal(a.absorption, b.absorption) and np.array_equal(a.scattering, b.scattering)
    assert np.array_equal(a.trajectory, b.trajectory)
    assert np.array_equal(a.det_weight(), b.det_weight()) and np.array_equal(a.det_target(), b.det_target())
    assert np.array_equal(a.det_count(), b.det_count())
    ra = np.frombuffer(a.bins, dtype=np.uint8)
    rb = np.frombuffer(b.bins, dtype=np.uint8)
    assert np.array_equal(ra, rb)


def test_backup_round_trip(tmp_path):
    d, o = _quirks_output()
    files.write_backup_file(d, o, Rng(12, 34), tmp_path / "b.bk")
    spec = d.spec()
    back = OutputHandle(spec, absorption=True, scattering=True, trajectory=True)
    rng = files.read_backup_file(tmp_path / "b.bk", back)
    assert (rng.seed, rng.nextPhoton) == (12, 34)
    back.c.numberOfPhotons = back.c.numberOfPhotons  # read from the file's first section
    assert back.c.numberOfPhotons == d.input.numberOfPhotons
    o.c.numberOfPhotons = d.input.numberOfPhotons
    _same_outputs(o, back)
    # a .mcml.out result reads without a generator section
    files.write_output_file(d, o, tmp_path / "r.mcml.out")
    again = OutputHandle(spec, absorption=True, scattering=True, trajectory=True)
    files.read_backup_file(tmp_path / "r.mcml.out", again, with_rng=False)
    _same_outputs(o, again)
    with pytest.raises(engine.EngineError, match="random generator"):
        files.read_backup_file(tmp_path / "r.mcml.out", again)


@needs_ref
def test_backup_reads_like_the_reference(tmp_path, monkeypatch):
    d, o = _quirks_output()
    spec = d.spec()
    files.write_backup_file(d, o, Rng(3, 1 << 40), tmp_path / "b.bk")
    ours_out = OutputHandle(spec, absorption=True, scattering=True, trajectory=True)
    rng = files.read_backup_file(tmp_path / "b.bk", ours_out)
    assert (rng.seed, rng.nextPhoton) == (3, 1 << 40)
    # the tallies as the reference's ReadOutputFormBackupFile reads them
    ref_out = OutputHandle(spec, absorption=True, scattering=True, trajectory=True)
    assert fl.ref_io().refio_read_output(str(tmp_path / "b.bk").encode(), ref_out.ptr) == 1
    _same_outputs(ours_out, ref_out)
    # The reference's ParseBackupFile cannot read the generator section of its own backups:
    # ReadThreadsFromBackupFile skips 10 sections before it (reader.cpp:34-43) where
    # WriteBackupToFile writes 11 (the scattering section 0x0C).  Ours reads them.
    monkeypatch.chdir(tmp_path)
    assert fl.ref_io().refio_write_backup(d.ptr, o.ptr, 9, 99) == 1
    name = files.backup_file_name(d.input.numberOfPhotons)
    v, off = C.c_uint64(), C.c_uint64()
    assert fl.ref_io().refio_read_backup(name.encode(), ref_out.ptr, C.byref(v), C.byref(off)) == 0
    again = OutputHandle(spec, absorption=True, scattering=True, trajectory=True)
    rng = files.read_backup_file(tmp_path / name, again)
    assert (rng.seed, rng.nextPhoton) == (9, 99)
    _same_outputs(ours_out, again)


def test_backup_with_other_layout_is_rejected_and_leaves_output(tmp_path):
    d, o = _quirks_output()
    files.write_backup_file(d, o, Rng(1, 2), tmp_path / "b.bk")
    other = scenes.slab_scene(0)  # different grid and detectors
    out = OutputHandle(other, absorption=True, scattering=True)
    out.absorption[:] = 5.0
    with pytest.raises(engine.EngineError, match="grid size"):
        files.read_backup_file(tmp_path / "b.bk", out)
    assert np.all(out.absorption == 5.0)


def test_writer_rejects_null_arguments(tmp_path):
    spec = scenes.slab_scene(0)
    o = OutputHandle(spec)
    assert engine.lib().xmc_write_output_file(None, o.ptr, str(tmp_path / "x").encode()) == -1
    assert engine.lib().xmc_launch_with_backups(InputHandle(spec).ptr, o.ptr, C.byref(Rng(1, 0)), None, 0,
                                                None) == -1
