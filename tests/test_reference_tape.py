"""The reference's own photon loop driven by the GPU engine's random numbers.

oracle/_ref/libxmcml_ref_tape.so is the reference's unmodified sources linked
with oracle/ref_tape.cpp through --wrap: NextMCG59 (mcml_mcg59.cpp:37-41)
returns the top 24 bits of the same per-photon MCG59 states the engine draws,
and every ComputeBVHIntersection[WithoutStep] call of the photon loop is
logged.  On the same uniforms:

* the reference's ComputePhoton (mcml_kernel.cpp:39-126) and the driver loop
  that records traces (oracle/ref_driver.cpp) make the same BVH queries, bit
  for bit, and produce the same tallies -- so the driver's traces are the
  reference's walk;
* the float walk (the 32-bit oracle, which the GPU reproduces bit for bit)
  takes the reference's integer decisions -- event type, layer, uniforms
  drawn -- over whole photon lives, with weights within 1e-5 relative
  throughout and positions and directions within 2e-5 (1.2e-5 measured over
  512 lives) until the photon's second boundary interaction.  After that
  the float walk's surface offset (it stands ~1e-6 mm off a surface where the
  reference stands 1e-8 mm off; an offset of delta shifts the next hit by
  delta / |cos theta| at grazing incidence) and accumulated direction
  rounding leave a bounded drift (below 1e-3 over lives of up to thousands of
  events);
* the committed fixture tests/golden/tape_traces.npz, which the GPU test
  reads, is this library's output.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd import scenes
from xmcml_amd.abi import InputHandle, OutputHandle, TERMINATE_KILL, TRACE_DTYPE

pytestmark = pytest.mark.skipif(not ol.ref_tape_available(), reason="reference tape build oracle/_ref is not present")

SCENES = {
    "slab": lambda: scenes.slab_scene(64),
    "glass": lambda: scenes.small_scene("glass", 64),
    "mismatch": lambda: scenes.small_scene("mismatch", 64),
    "heightfield": lambda: scenes.small_scene("heightfield", 64),
}
FIXTURE = Path(__file__).parent / "golden" / "tape_traces.npz"


@pytest.mark.parametrize("name", list(SCENES))
def test_compute_photon_and_driver_make_the_same_queries(name):
    spec = SCENES[name]()
    spec.termination = TERMINATE_KILL  # ComputePhoton's own termination
    inp = InputHandle(spec)
    lib = ol.ref_tape()
    for p in range(24):
        o = OutputHandle(spec, absorption=False)
        qa, da = ol.tape_queries(lambda: lib.xref_run(inp.ptr, o.ptr, 5, p, 1, 0))
        tr = np.zeros(200000, dtype=TRACE_DTYPE)
        qb, db = ol.tape_queries(lambda: lib.xref_trace(inp.ptr, 5, p, tr.ctypes.data, 200000))
        assert len(qa) > 0 and da == db, (p, da, db)
        assert qa.tobytes() == qb.tobytes(), p
    # and the tallies of 300 photons
    a, b = OutputHandle(spec), OutputHandle(spec)
    assert lib.xref_run(inp.ptr, a.ptr, 6, 0, 300, 0) == 0
    assert lib.xref_run(inp.ptr, b.ptr, 6, 0, 300, 1) == 0
    assert np.array_equal(a.absorption, b.absorption)
    assert np.array_equal(a.det_weight(), b.det_weight())
    assert np.array_equal(a.bin_weight(), b.bin_weight())


def test_tape_draws_the_engine_uniforms():
    """The first step of photon p is -ln(u) / mu_t with u the top 24 bits of
    the photon's first MCG59 state (the engine's draw), not x / 2^59."""
    spec = scenes.slab_scene(8)
    spec.termination = TERMINATE_KILL
    inp = InputHandle(spec)
    lib = ol.ref_tape()
    mut = spec.layers[1].mua + spec.layers[1].mus
    for p in range(8):
        o = OutputHandle(spec, absorption=False)
        q, _ = ol.tape_queries(lambda: lib.xref_run(inp.ptr, o.ptr, 3, p, 1, 0))
        x = (ol.oracle().xmco_photon_stream(3, p) * 302875106592253) & ((1 << 59) - 1)
        u = (x >> 35) / 2.0 ** 24
        assert q[0]["step"] == -np.log(u) / mut
        assert q[0]["draws"] == 1


@pytest.mark.parametrize("name", list(SCENES))
def test_float_walk_takes_the_reference_decisions(name):
    spec = SCENES[name]()
    inp = InputHandle(spec)
    n = 128
    full, w, early, whole = 0, 0.0, 0.0, 0.0
    for p in range(n):
        a = ol.trace(ol.ref_tape().xref_trace, inp, 99, p, 200000)
        b = ol.trace(ol.oracle().xmco_trace, inp, 99, p, 200000, precision=32)
        r = ol.compare_walks(a, b)
        full += r["full"]
        w, early, whole = max(w, r["weight"]), max(early, r["early"]), max(whole, r["whole"])
    print(f"{name}: {full}/{n} whole lives with identical decisions; weight {w:.2e}, "
          f"before the second boundary {early:.2e}, whole prefix {whole:.2e}")
    assert full >= 0.98 * n
    assert w < 1e-5
    assert early < 2e-5
    assert whole < 1e-3


def test_fixture_is_the_tape_reference():
    """tests/golden/tape_traces.npz holds this library's traces."""
    z = np.load(FIXTURE, allow_pickle=False)
    seed, photons = int(z["seed"]), int(z["photons"])
    for name, make in SCENES.items():
        inp = InputHandle(make())
        for p in range(photons):
            f = z[f"{name}_{p}"]
            t = ol.trace(ol.ref_tape().xref_trace, inp, seed, p, 200000)
            assert len(f) == len(t), (name, p)
            for k in f.dtype.names:
                assert np.array_equal(f[k], t[k]), (name, p, k)
