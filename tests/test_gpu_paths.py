"""Scattering grid and detector trajectory maps on the GPU (voxel-plane
crossings, mcml_kernel.cpp:750-810) against the float oracle on the same
photon streams: trajectory counts identical, scattering grid equal up to the
order of the FP64 additions."""
import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd import scenes
from xmcml_amd.abi import InputHandle, OutputHandle

pytestmark = pytest.mark.gpu

SMALL = {
    "slab": lambda n: scenes.slab_scene(n),
    "glass": lambda n: scenes.small_scene("glass", n),
    "mismatch": lambda n: scenes.small_scene("mismatch", n),
    "heightfield": lambda n: scenes.small_scene("heightfield", n),
}


@pytest.mark.parametrize("name", list(SMALL))
def test_paths_match_float_oracle(engine, name):
    n = 3000
    spec = SMALL[name](n)
    inp = InputHandle(spec)
    with engine.Scene(spec, scattering=True, trajectories=True, trajectory_capacity=16384) as sc:
        sc.run(13, 0, n)
        g = sc.collect()
        st = sc.stats()
    o = OutputHandle(spec, scattering=True, trajectory=True)
    ol.run_oracle(spec, inp, o, 13, 0, n, 32, 8, scattering=True)
    assert st["trajectoryOverflows"] == 0
    assert o.scattering.sum() > 0
    np.testing.assert_allclose(g.scattering, o.scattering, rtol=1e-10, atol=1e-12)
    assert np.array_equal(g.trajectory, o.trajectory)
    # the other tallies are unchanged by tracking the paths
    assert np.array_equal(g.det_count(), o.det_count())
    np.testing.assert_allclose(g.absorption, o.absorption, rtol=1e-10, atol=1e-12)


def test_trajectory_capacity_overflow_is_counted(engine):
    n = 2000
    spec = scenes.slab_scene(n)
    with engine.Scene(spec, scattering=True, trajectories=True) as full:
        full.run(3, 0, n)
        a = full.collect()
        full.stats()
    with engine.Scene(spec, scattering=True, trajectories=True, trajectory_capacity=4) as small:
        small.run(3, 0, n)
        b = small.collect()
        st = small.stats()
    assert st["trajectoryOverflows"] > 0
    # the first 4 crossings of every detected photon are still counted
    assert 0 < b.trajectory.sum() < a.trajectory.sum()
    assert np.all(b.trajectory <= a.trajectory)
    np.testing.assert_allclose(b.scattering, a.scattering, rtol=1e-10, atol=1e-12)


def test_paths_off_leave_outputs_untouched(engine):
    spec = scenes.slab_scene(500)
    with engine.Scene(spec) as sc:
        sc.run(1, 0, 500)
        out = OutputHandle(spec, scattering=True, trajectory=True)
        out.scattering[:] = 7.0
        sc.collect(out)
    assert np.all(out.scattering == 7.0)
    assert out.trajectory.sum() == 0
    assert out.det_count().sum() > 0
