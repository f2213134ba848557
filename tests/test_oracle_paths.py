"""The oracle's float restatement of the scattering grid and detector
trajectory maps (UpdatePhotonTrajectory / UpdateDetectorTrajectory,
mcml_kernel.cpp:750-810) against its double path, which is bit-identical to
the reference build.  On the same photon streams float and double walks take
the same decisions for most photons, so the two grids must agree closely; a
wrong plane rule (an off-by-one plane, a missing nudge segment, crossings
outside the Area) moves the totals by far more than the bounds below."""
import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd import scenes
from xmcml_amd.abi import InputHandle, OutputHandle


@pytest.mark.parametrize("name", ["slab", "heightfield"])
def test_float_paths_agree_with_double(name):
    n = 3000
    spec = scenes.slab_scene(n) if name == "slab" else scenes.small_scene("heightfield", n)
    inp = InputHandle(spec)
    grids = {}
    for prec in (32, 64):
        o = OutputHandle(spec, scattering=True, trajectory=True)
        ol.run_oracle(spec, inp, o, 9, 0, n, prec, 8, scattering=True)
        grids[prec] = o
    s32, s64 = grids[32].scattering, grids[64].scattering
    t32, t64 = grids[32].trajectory.astype(np.float64), grids[64].trajectory.astype(np.float64)
    assert s64.sum() > 0 and t64.sum() > 0
    assert s32.sum() == pytest.approx(s64.sum(), rel=5e-3)
    assert np.abs(s32 - s64).sum() <= 0.05 * s64.sum()
    assert t32.sum() == pytest.approx(t64.sum(), rel=0.01)
    assert np.abs(t32 - t64).sum() <= 0.05 * t64.sum()
    # only detectors that detected packets carry trajectory counts
    assert np.all(t32.sum(axis=1)[grids[32].det_count() == 0] == 0)


def test_crossings_of_a_straight_beam():
    # a non-scattering glass layer over a clear slab: a normal-incidence photon
    # crosses the voxel-centre planes of the Area exactly once each on its
    # way down, so the scattering grid holds its weight in one column
    spec = scenes.small_scene("glass", 1)
    spec.area_corner = (-1.0, -1.0, 0.0)
    spec.area_length = (2.0, 2.0, 1.0)
    spec.area_partitions = (2, 2, 10)  # planes at z = 0.05, 0.15, ..., 0.95 inside the glass (0..1 mm)
    inp = InputHandle(spec)
    for prec in (32, 64):
        o = OutputHandle(spec, scattering=True, trajectory=True)
        ol.run_oracle(spec, inp, o, 1, 0, 1, prec, 1, scattering=True)
        g = o.scattering.reshape(2, 2, 10)
        w0 = 1.0 - o.specular
        # the photon starts at x = y = 0 (a voxel corner), so voxel (1, 1, z)
        assert np.count_nonzero(g) >= 10, prec
        col = g[1, 1]
        assert np.allclose(col[col > 0], w0, rtol=1e-6), (prec, col)
