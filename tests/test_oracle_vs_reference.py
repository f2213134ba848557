"""The 64-bit oracle against the reference sources compiled unmodified
(oracle/_ref): identical per-photon MCG59 streams must give bit-identical
tallies and per-step traces.  Where the reference build is absent, the
committed vectors in tests/golden/ref_vectors.npz (written by
tests/golden/make_ref_vectors.py from that build) are the check."""
import hashlib
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd import scenes
from xmcml_amd.abi import InputHandle, OutputHandle, TERMINATE_KILL

GOLD = Path(__file__).parent / "golden" / "ref_vectors.npz"


def digest(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


def kill_slab(n=300):
    s = scenes.slab_scene(n)
    s.termination = TERMINATE_KILL
    return s


SCENES = {
    "slab-kill": (kill_slab, 0),
    "slab-roulette": (lambda n: scenes.slab_scene(n), 1),
    "glass": (lambda n: scenes.small_scene("glass", n), 1),
    "mismatch": (lambda n: scenes.small_scene("mismatch", n), 0),
    "heightfield": (lambda n: scenes.small_scene("heightfield", n), 1),
}


def _outputs(spec):
    return OutputHandle(spec, scattering=True, trajectory=True)


def _same_outputs(a, b):
    assert np.array_equal(a.absorption, b.absorption)
    assert np.array_equal(a.scattering, b.scattering)
    assert np.array_equal(a.trajectory, b.trajectory)
    assert np.array_equal(a.det_weight(), b.det_weight())
    assert np.array_equal(a.det_target(), b.det_target())
    assert np.array_equal(a.det_count(), b.det_count())
    assert np.array_equal(a.bin_weight(), b.bin_weight())
    assert np.array_equal(a.bin_target(), b.bin_target())
    assert np.array_equal(a.bin_count(), b.bin_count())
    assert a.specular == b.specular


@pytest.mark.parametrize("name", list(SCENES))
def test_tallies_bit_identical(ref, name):
    make, driver = SCENES[name]
    n = 200
    spec = make(n)
    inp = InputHandle(spec)
    r = _outputs(spec)
    assert ref.xref_run(inp.ptr, r.ptr, 777, 5, n, driver) == 0
    o = _outputs(spec)
    ol.run_oracle(spec, inp, o, 777, 5, n, 64, 1, True)
    _same_outputs(r, o)


def test_driver_loop_equals_compute_photon(ref):
    # the trace driver repeats ComputePhoton's control flow; pin it to the real thing
    spec = kill_slab(300)
    inp = InputHandle(spec)
    a, b = _outputs(spec), _outputs(spec)
    ref.xref_run(inp.ptr, a.ptr, 11, 0, 300, 0)
    ref.xref_run(inp.ptr, b.ptr, 11, 0, 300, 1)
    _same_outputs(a, b)


@pytest.mark.parametrize("name", list(SCENES))
def test_traces_bit_identical(ref, name):
    make, _ = SCENES[name]
    inp = InputHandle(make(10))
    for p in range(6):
        a = ol.trace(ref.xref_trace, inp, 2024, p, 50000)
        b = ol.trace(ol.oracle().xmco_trace, inp, 2024, p, 50000, precision=64)
        assert len(a) == len(b)
        for k in a.dtype.names:
            assert np.array_equal(a[k], b[k]), (p, k)


@pytest.mark.parametrize("name", ["heightfield", "slab-kill"])
def test_bvh_identical(ref, name):
    inp = InputHandle(SCENES[name][0](10))
    A = ol.bvh_arrays(ref.xref_bvh, inp)
    B = ol.bvh_arrays(ol.oracle().xmco_bvh, inp)
    for k in A:
        assert np.array_equal(A[k], B[k]), k


def test_committed_vectors():
    """Oracle against the vectors the reference build wrote (runs without it)."""
    if not GOLD.exists():
        pytest.skip("tests/golden/ref_vectors.npz not generated")
    g = np.load(GOLD, allow_pickle=False)
    spec = kill_slab(int(g["n"]))
    inp = InputHandle(spec)
    o = _outputs(spec)
    ol.run_oracle(spec, inp, o, int(g["seed"]), 0, int(g["n"]), 64, 1, True)
    assert np.array_equal(digest(o.absorption), g["absorption"])
    assert np.array_equal(digest(o.scattering), g["scattering"])
    assert np.array_equal(digest(o.trajectory), g["trajectory"])
    assert np.array_equal(o.det_weight(), g["det_weight"])
    assert np.array_equal(o.bin_weight(), g["bin_weight"])
    assert np.array_equal(o.det_count(), g["det_count"])
    for p in range(int(g["n_traces"])):
        b = ol.trace(ol.oracle().xmco_trace, inp, int(g["seed"]), p, 50000, precision=64)
        a = g[f"trace{p}"]
        assert len(a) == len(b)
        for k in a.dtype.names:
            assert np.array_equal(a[k], b[k])
    hs = scenes.small_scene("heightfield", 10)
    A = ol.bvh_arrays(ol.oracle().xmco_bvh, InputHandle(hs))
    for k in ("boxes", "links", "tsurf", "tidx"):
        assert np.array_equal(digest(A[k]), g["bvh_" + k])


def test_reference_launch_omp_runs(ref):
    # the unchanged LaunchOMP with leapfrog thread streams (the timed baseline)
    spec = kill_slab(2000)
    inp = InputHandle(spec, photons=2000)
    o = OutputHandle(spec, scattering=True, trajectory=True)
    secs = ref.xref_launch_omp(inp.ptr, o.ptr, 777, 4)
    assert secs > 0
    r = o.det_weight().sum() / 2000
    assert 0.3 < r < 0.5
