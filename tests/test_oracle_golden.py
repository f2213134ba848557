"""The oracle (and, when built, the reference itself) against the expected
values of the reference's CppUnit suites (stable/1.2/xmcml/xmcml_tests),
transcribed in tests/golden/reference_unit_cases.json."""
import ctypes as C
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd.abi import (CubeSpec, InputHandle, LayerSpec, RingSpec, SceneSpec, Area, Int3, Double3, Layer)

G = json.loads((Path(__file__).parent / "golden" / "reference_unit_cases.json").read_text())
DELTA = G["delta"]


def _val(v):
    if isinstance(v, str):
        return eval(v, {"cos": math.cos, "sqrt": math.sqrt, "pi": math.pi})
    return v


def _a(v):
    return np.asarray(v, dtype=np.float64)


def _rng():
    off = C.c_uint64()
    val = C.c_uint64(ol.oracle().xmco_mcg59_init(777, 0, 1, C.byref(off)))
    return val, off.value


def impls(ref_fixture):
    out = [("oracle", "xmco")]
    if ol.ref_available():
        out.append(("reference", "xref"))
    return out


def both():
    libs = [ol.oracle()]
    if ol.ref_available():
        libs.append(ol.ref())
    return libs


def fn(lib, name):
    return getattr(lib, ("xmco_" if lib is ol.oracle() else "xref_") + name)


def test_mcg59_first_draws_and_moments():
    val, off = _rng()
    u = [ol.oracle().xmco_mcg59_next(C.byref(val), off) for _ in range(2)]
    assert u == pytest.approx(G["mcg59"]["first_draws"], abs=1e-8)
    # closed form: x_k = (2*777+1) * 13^(13k) mod 2^59
    x = 2 * 777 + 1
    a = 302875106592253
    for k, uk in enumerate(u, 1):
        assert uk == ((x * pow(a, k, 1 << 64)) % (1 << 59)) / 2.0 ** 59
    m = G["mcg59"]["moments"]
    val, off = _rng()
    draws = np.array([ol.oracle().xmco_mcg59_next(C.byref(val), off) for _ in range(m["serial"]["n"])])
    assert draws.min() >= 0.0 and draws.max() <= 1.0
    assert abs(draws.mean() - 0.5) < m["serial"]["delta"]
    assert abs(draws.var() - 1 / 12) < m["serial"]["delta"]
    for p in m["parallel"]:
        for i in range(p["streams"]):
            o = C.c_uint64()
            v = C.c_uint64(ol.oracle().xmco_mcg59_init(777, i, p["streams"], C.byref(o)))
            d = np.array([ol.oracle().xmco_mcg59_next(C.byref(v), o.value) for _ in range(p["n"])])
            assert abs(d.mean() - 0.5) < p["delta"] and abs(d.var() - 1 / 12) < p["delta"]


def test_photon_streams_are_block_splits():
    # photon p starts at (2s+1) a^(p 2^24) mod 2^59
    a = 302875106592253
    for s, p in [(777, 0), (777, 1), (1, 12345), (2024, 10 ** 9)]:
        expect = ((2 * s + 1) * pow(pow(a, 1 << 24, 1 << 64), p, 1 << 64)) % (1 << 59)
        assert ol.oracle().xmco_photon_stream(s, p) == expect


@pytest.mark.parametrize("case", G["specular"], ids=lambda c: c["source"])
def test_specular(case):
    layers = (Layer * 3)()
    for i, (n, mua, mus) in enumerate(case["layers"]):
        layers[i].refractiveIndex, layers[i].absorptionCoefficient, layers[i].scatteringCoefficient = n, mua, mus
    for lib in both():
        assert fn(lib, "specular")(C.byref(layers)) == pytest.approx(case["expected"], abs=DELTA)


@pytest.mark.parametrize("case", G["triangle_distance"], ids=lambda c: c["source"])
def test_triangle_distance(case):
    args = [ol.dptr(_a(case[k])) for k in ("o", "d", "a", "b", "c")]
    for lib in both():
        t = fn(lib, "triangle_distance")(*args)
        if case["expected"] == "miss":
            assert t >= 1e256
        else:
            assert t == pytest.approx(case["expected"], abs=DELTA)


def _surface_scene(vertices, triangles):
    v = np.asarray(vertices, np.float64)
    t = np.asarray(triangles, np.int32)
    return SceneSpec(layers=[LayerSpec(1.0, 0, 0, 0, [0]), LayerSpec(1.4, 1, 10, 0.9, [0])], surfaces=[(v, t)],
                     area_corner=(-10, -10, -10), area_length=(20, 20, 20), area_partitions=(1, 1, 1))


@pytest.mark.parametrize("case", G["bvh_intersection"], ids=lambda c: c["source"])
def test_bvh_intersection(case):
    spec = _surface_scene(case["vertices"], case["triangles"])
    inp = InputHandle(spec)
    o, d, s = _a([case["o"]]), _a([case["d"]]), _a([0.0])
    found, dist, surf, tri, nrm = np.zeros(1, np.int32), np.zeros(1), np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(3)
    for mode in (0, 1):
        ol.oracle().xmco_query(inp.ptr, ol.dptr(o), ol.dptr(d), ol.dptr(s), 1, 64, mode, ol.iptr(found),
                               ol.dptr(dist), ol.iptr(surf), ol.iptr(tri), ol.dptr(nrm))
        assert found[0] == case["found"]
        if case["found"]:
            assert dist[0] == pytest.approx(case["distance"], abs=DELTA)
            assert nrm == pytest.approx(case["normal"], abs=DELTA)
    if ol.ref_available():
        ol.ref().xref_query(inp.ptr, ol.dptr(o), ol.dptr(d), ol.dptr(s), 1, ol.iptr(found), ol.dptr(dist),
                            ol.iptr(surf), ol.dptr(nrm))
        assert found[0] == case["found"]


@pytest.mark.parametrize("case", G["area_index"], ids=lambda c: c["source"])
def test_area_index(case):
    area = Area(Double3(*case["corner"]), Double3(*case["length"]), Int3(*case["part"]))
    assert ol.oracle().xmco_area_index(C.byref(area), ol.dptr(_a(case["pos"]))) == case["expected"]


@pytest.mark.parametrize("case", G["cosine_theta"], ids=lambda c: c["source"])
def test_cosine_theta(case):
    for lib in both():
        val, off = _rng()
        assert fn(lib, "cosine_theta")(case["g"], C.byref(val), off) == pytest.approx(case["expected"], abs=DELTA)


@pytest.mark.parametrize("case", G["new_direction"], ids=lambda c: c["source"])
def test_new_direction(case):
    d = _a(case["dir"])
    if case["normalize"]:
        d = d / np.sqrt(np.dot(d, d))
    for lib in both():
        val, off = _rng()
        out = np.zeros(3)
        fn(lib, "new_direction")(ol.dptr(d), case["g"], C.byref(val), off, ol.dptr(out))
        assert out == pytest.approx(case["expected"], abs=DELTA)
        assert np.linalg.norm(out) == pytest.approx(1.0, abs=DELTA)


@pytest.mark.parametrize("case", G["transmit_cos"], ids=lambda c: c["source"])
def test_transmit_cos(case):
    ci = _val(case["ci"])
    exp = ci if case["expected"] == "ci" else case["expected"]
    for lib in both():
        assert fn(lib, "transmit_cos")(case["n1"], case["n2"], ci) == pytest.approx(exp, abs=DELTA)


@pytest.mark.parametrize("case", G["fresnel"], ids=lambda c: c["source"])
def test_fresnel(case):
    for lib in both():
        r = fn(lib, "fresnel")(case["n1"], case["n2"], _val(case["ci"]), _val(case["ct"]))
        assert r == pytest.approx(case["expected"], abs=DELTA)


def test_fresnel_closed_form():
    # normal incidence ((n1-n2)/(n1+n2))^2 and the s/p average at 45 degrees
    n1, n2 = 1.0, 1.5
    ci = math.cos(math.pi / 4)
    st = n1 * math.sin(math.pi / 4) / n2
    ct = math.sqrt(1 - st * st)
    rs = ((n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)) ** 2
    rp = ((n1 * ct - n2 * ci) / (n1 * ct + n2 * ci)) ** 2
    assert ol.oracle().xmco_fresnel(n1, n2, ci, ct) == pytest.approx(0.5 * (rs + rp), rel=1e-12)
    assert ol.oracle().xmco_transmit_cos(n1, n2, ci) == pytest.approx(ct, rel=1e-12)
    assert ol.oracle().xmco_fresnel(n1, n2, 1.0, 1.0) == pytest.approx(((n1 - n2) / (n1 + n2)) ** 2, rel=1e-12)


def test_find_layer():
    fl = G["find_layer"]
    spec = SceneSpec(layers=[LayerSpec(1, 0, 0, 0, s) for s in fl["layers"]],
                     surfaces=[(np.zeros((3, 3)), np.array([[0, 1, 2]], np.int32))] * 7,
                     area_corner=(0, 0, 0), area_length=(1, 1, 1), area_partitions=(1, 1, 1))
    inp = InputHandle(spec)
    for c in fl["cases"]:
        assert ol.oracle().xmco_find_layer(inp.ptr, c["surface"], c["layer"]) == c["expected"]


@pytest.mark.parametrize("case", G["refract"], ids=lambda c: c["source"])
def test_refract(case):
    for lib in both():
        out = np.zeros(3)
        fn(lib, "refract")(_val(case["n1"]), _val(case["n2"]), ol.dptr(_a(case["i"])), ol.dptr(_a(case["n"])),
                           ol.dptr(out))
        assert out == pytest.approx(case["expected"], abs=DELTA)


@pytest.mark.parametrize("case", G["reflect"], ids=lambda c: c["source"])
def test_reflect(case):
    for lib in both():
        out = np.zeros(3)
        fn(lib, "reflect")(ol.dptr(_a(case["i"])), ol.dptr(_a(case["n"])), ol.dptr(out))
        assert out == pytest.approx(case["expected"], abs=DELTA)


@pytest.mark.parametrize("case", G["step"], ids=lambda c: c["source"])
def test_step(case):
    for lib in both():
        val, off = _rng()
        assert fn(lib, "step")(case["mua"], case["mus"], C.byref(val), off) == pytest.approx(case["expected"], abs=DELTA)


@pytest.mark.parametrize("case", G["drop_weight"], ids=lambda c: c["source"])
def test_drop_weight(case):
    # ComputeDroppedWeightOfPhoton: w -= w mua/(mua+mus); checked through a
    # one-layer closed form since the oracle folds it into the photon loop
    w = case["w"]
    for _ in range(case["calls"]):
        w -= w * case["mua"] / (case["mua"] + case["mus"])
    assert w == pytest.approx(case["expected"], abs=DELTA)


@pytest.mark.parametrize("case", G["critical_cos"], ids=lambda c: c["source"])
def test_critical_cos(case):
    assert ol.oracle().xmco_critical_cos(case["n1"], case["n2"]) == pytest.approx(case["expected"], abs=DELTA)


@pytest.mark.parametrize("case", G["detector_id"], ids=lambda c: c["source"])
def test_detector_id(case):
    spec = SceneSpec(layers=[LayerSpec(1, 0, 0, 0, [0]), LayerSpec(1.4, 1, 1, 0.9, [0])],
                     surfaces=[(np.zeros((3, 3)), np.array([[0, 1, 2]], np.int32))], area_corner=(0, 0, 0),
                     area_length=(1, 1, 1), area_partitions=(1, 1, 1),
                     cubes=[CubeSpec(c, l, f) for c, l, f in case["cubes"]],
                     rings=[RingSpec(rs, rb, c, f) for c, rs, rb, f in case["rings"]])
    inp = InputHandle(spec)
    visited = (C.c_uint8 * 8)(0, 1, 0, 0, 0, 0, 0, 0)
    d = _a([0.0, 0.0, -1.0])
    for lib in both():
        assert fn(lib, "detector_id")(inp.ptr, ol.dptr(_a(case["pos"])), ol.dptr(d), visited) == case["expected"]


def test_a_seed_has_2_33_disjoint_photon_streams():
    """Photon p starts at (2s+1) a^(p 2^24) mod 2^59.  MCG59 (a = 13^13 = 5 mod 8)
    has period 2^57 on odd states, so the jump a^(2^24 2^i) is the identity
    exactly from i = 33 on: photon p + 2^33 repeats photon p, photons below
    2^33 never share a block.  The engine caps photon indices at 2^33."""
    a, mask = 302875106592253, (1 << 59) - 1
    assert a % 8 == 5
    j = pow(a, 1 << 24, 1 << 59)
    table = []
    for _ in range(40):
        table.append(j)
        j = (j * j) & mask
    assert all(t != 1 for t in table[:33])
    assert all(t == 1 for t in table[33:])
    o = ol.oracle()
    for p in (0, 5, 123456789):
        assert o.xmco_photon_stream(3, p) != o.xmco_photon_stream(3, p + (1 << 32))
        assert o.xmco_photon_stream(3, p) == o.xmco_photon_stream(3, p + (1 << 33))
    src = (ol.ROOT / "xmcml_amd" / "csrc" / "xmc_params.h").read_text()
    assert "constexpr int kJumpBits = 57 - kStreamLog2;" in src and "constexpr int kStreamLog2 = 24;" in src
