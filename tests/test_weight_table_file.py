"""The biasing's weight-table file (the launcher's -t option) against the
reference's own writer and reader: WriteWeightIntegralTableToFile and
ParseWeightIntegralFile (base/xmcml/xmcmlLauncher/weight_integral.cpp:105-247),
and the <Biasing> block of the input file (parser.cpp:201-243).

Pins: the byte layout packed here field by field with struct (section id
0x0200 of sections.h, a 32-bit length, then factor, integral precision,
table count and per table anisotropy, element count, elements); the
reference build oracle/_ref/libxmcml_ref_io.so, when present, writes the same
bytes and accepts and rejects the same files.  No GPU: host code only."""
import copy
import struct

import numpy as np
import pytest

import files_lib as fl
from xmcml_amd import engine, files, scenes
from xmcml_amd.abi import InputHandle

needs_ref = pytest.mark.skipif(not fl.ref_io_available(), reason="reference build oracle/_ref is not present")

FACTOR, IPREC, TPREC = 1.5, 16, 8


def _spec(tables=None, factor=FACTOR, iprec=IPREC, tprec=TPREC, gs=(0.9, 0.8, 0.9)):
    """A 4-layer slab whose layers 1..3 have anisotropies gs (2 distinct for
    the default), biasing towards (2, -1, 0)."""
    base = scenes.slab_scene(1000)
    assert len(base.layers) == 4
    layers = [copy.copy(l) for l in base.layers]
    for l, g in zip(layers[1:], gs):
        l.g = g
    base.layers = layers
    rng = np.random.default_rng(5)
    if tables is None:
        tables = [(g, rng.random(tprec + 1)) for g in scenes.distinct_anisotropies(base)]
    return scenes.biased(base, target=(2.0, -1.0, 0.0), factor=factor, tables=tables, integral_precision=iprec,
                         table_precision=tprec)


def _packed(factor, iprec, tables):
    body = struct.pack("<dii", factor, iprec, len(tables))
    for g, el in tables:
        body += struct.pack("<di", g, len(el)) + np.asarray(el, dtype="<f8").tobytes()
    return struct.pack("<II", 0x0200, len(body)) + body


def test_written_file_has_the_reference_layout(tmp_path):
    spec = _spec()
    files.write_weight_table_file(InputHandle(spec), tmp_path / "t.bin")
    assert (tmp_path / "t.bin").read_bytes() == _packed(FACTOR, IPREC, spec.weight_tables)


def test_default_file_name(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    files.write_weight_table_file(InputHandle(_spec()))
    assert (tmp_path / files.WEIGHT_TABLE_FILE).exists()


@needs_ref
def test_written_file_bytes_match_reference_build(tmp_path, monkeypatch):
    import ctypes as C
    spec = _spec()
    files.write_weight_table_file(InputHandle(spec), tmp_path / "ours.bin")
    monkeypatch.chdir(tmp_path)
    g = np.array([t[0] for t in spec.weight_tables])
    n = np.array([len(t[1]) for t in spec.weight_tables], dtype=np.int32)
    el = np.concatenate([t[1] for t in spec.weight_tables])
    L = fl.ref_io()
    L.refio_write_weight_table.restype = C.c_int
    L.refio_write_weight_table.argtypes = [C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    assert L.refio_write_weight_table(FACTOR, IPREC, len(g), g.ctypes.data, n.ctypes.data, el.ctypes.data) == 1
    assert (tmp_path / "ours.bin").read_bytes() == (tmp_path / files.WEIGHT_TABLE_FILE).read_bytes()


def _doc(tmp_path, spec, name="in.xml"):
    files.write_input_xml(spec, tmp_path / name)
    return files.Document(tmp_path / name)


def test_input_file_biasing_block_reads_back(tmp_path):
    spec = _spec()
    d = _doc(tmp_path, spec)
    b = d.biasing()
    assert (b.useBiasing, b.weightIntegralPrecision, b.weightTablePrecision, b.attractiveFactor) == (1, IPREC, TPREC,
                                                                                                     FACTOR)
    assert (b.targetPoint.x, b.targetPoint.y, b.targetPoint.z) == (2.0, -1.0, 0.0)
    x = d.input
    assert (x.useBiasing, x.weightIntegralPrecision, x.weightTablePrecision, x.attractiveFactor) == (1, IPREC, TPREC,
                                                                                                     FACTOR)
    assert x.numberOfWeightTables == 0  # tables come from the table file or are computed
    with pytest.raises(ValueError):
        files.write_input_xml(scenes.biased(spec, target=(0.5, 0.0, 0.0)), tmp_path / "bad.xml")


@needs_ref
def test_input_file_biasing_block_reads_like_the_reference(tmp_path):
    spec = _spec()
    d = _doc(tmp_path, spec)
    ref = fl.ref_parse(tmp_path / "in.xml")
    assert ref is not None
    assert fl.canonical(d.input, d.biasing()) == fl.canonical(*ref)


def test_round_trip_sets_the_input_tables(tmp_path):
    spec = _spec()
    files.write_weight_table_file(InputHandle(spec), tmp_path / "t.bin")
    d = _doc(tmp_path, spec)
    d.read_weight_table_file(tmp_path / "t.bin")
    x = d.input
    assert x.numberOfWeightTables == len(spec.weight_tables)
    back = d.spec().weight_tables
    for (g0, e0), (g1, e1) in zip(spec.weight_tables, back):
        assert g0 == g1 and np.array_equal(e0, e1)


def _ref_parse_table(path, spec):
    import ctypes as C
    L = fl.ref_io()
    L.refio_parse_weight_table.restype = C.c_int
    L.refio_parse_weight_table.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    gl = np.array([l.g for l in spec.layers])
    cap_t, cap_e = 16, 4096
    nt = np.zeros(1, dtype=np.int32)
    g = np.zeros(cap_t)
    ne = np.zeros(cap_t, dtype=np.int32)
    el = np.zeros(cap_e)
    ok = L.refio_parse_weight_table(str(path).encode(), len(gl), gl.ctypes.data, spec.attractive_factor,
                                    spec.weight_integral_precision, spec.weight_table_precision, cap_t, cap_e,
                                    nt.ctypes.data, g.ctypes.data, ne.ctypes.data, el.ctypes.data)
    if not ok:
        return None
    out, k = [], 0
    for i in range(nt[0]):
        out.append((g[i], el[k:k + ne[i]].copy()))
        k += ne[i]
    return out


@needs_ref
def test_reads_like_the_reference(tmp_path):
    spec = _spec()
    files.write_weight_table_file(InputHandle(spec), tmp_path / "t.bin")
    d = _doc(tmp_path, spec)
    d.read_weight_table_file(tmp_path / "t.bin")
    ref = _ref_parse_table(tmp_path / "t.bin", spec)
    ours = d.spec().weight_tables
    assert ref is not None and len(ref) == len(ours)
    for (g0, e0), (g1, e1) in zip(ref, ours):
        assert g0 == g1 and np.array_equal(e0, e1)


def _bad_files(spec):
    """(name, bytes, input spec) triples that ParseWeightIntegralFile rejects."""
    good = spec.weight_tables
    return [
        ("other factor", _packed(FACTOR + 0.5, IPREC, good), spec),
        ("other integral precision", _packed(FACTOR, IPREC * 2, good), spec),
        ("other table precision", _packed(FACTOR, IPREC, [(g, e[:-1]) for g, e in good]), spec),
        ("anisotropy missing", _packed(FACTOR, IPREC, [good[0], (0.5, good[1][1])]), spec),
        ("fewer tables than anisotropies", _packed(FACTOR, IPREC, good[:1]), spec),
        ("section id", struct.pack("<I", 0x0100) + _packed(FACTOR, IPREC, good)[4:], spec),
        ("truncated", _packed(FACTOR, IPREC, good)[:-8], spec),
        ("huge table count", _packed(FACTOR, IPREC, good)[:20] + struct.pack("<i", 2 ** 30)
         + _packed(FACTOR, IPREC, good)[24:], spec),
    ]


@pytest.mark.parametrize("case", range(8))
def test_rejected_files_leave_the_input(tmp_path, case):
    spec = _spec()
    name, data, s = _bad_files(spec)[case]
    (tmp_path / "bad.bin").write_bytes(data)
    d = _doc(tmp_path, s)
    with pytest.raises(engine.EngineError):
        d.read_weight_table_file(tmp_path / "bad.bin")
    assert d.input.numberOfWeightTables == 0, name


@needs_ref
@pytest.mark.parametrize("case", range(7))  # the reference sizes its table list from an unchecked count
def test_rejected_like_the_reference(tmp_path, case):
    spec = _spec()
    name, data, s = _bad_files(spec)[case]
    (tmp_path / "bad.bin").write_bytes(data)
    assert _ref_parse_table(tmp_path / "bad.bin", s) is None, name


def test_tables_within_the_tolerance_are_accepted(tmp_path):
    """A table counts for an anisotropy within 1e-6 (EPSILON,
    mcml_intersection.h:6; weight_integral.cpp:232)."""
    spec = _spec()
    near = [(g + 5e-7, e) for g, e in spec.weight_tables]
    (tmp_path / "near.bin").write_bytes(_packed(FACTOR, IPREC, near))
    d = _doc(tmp_path, spec)
    d.read_weight_table_file(tmp_path / "near.bin")
    assert d.input.numberOfWeightTables == 2
    far = [(g + 2e-6, e) for g, e in spec.weight_tables]
    (tmp_path / "far.bin").write_bytes(_packed(FACTOR, IPREC, far))
    d2 = _doc(tmp_path, spec, "in2.xml")
    with pytest.raises(engine.EngineError):
        d2.read_weight_table_file(tmp_path / "far.bin")
