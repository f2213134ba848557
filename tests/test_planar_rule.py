"""The engine's float rule for planar scenes (every surface an axis-aligned
rectangle at one z, two triangles sharing a diagonal): the nearest surface is
found analytically instead of by the BVH walk.  On the CPU oracle's 32-bit
mode (which the GPU reproduces bit for bit) the rule must pick the surface
the BVH walk picks -- in float and in the reference's double arithmetic --
at distances equal to float rounding, and it must apply to exactly the
planar scenes."""
import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd import scenes
from xmcml_amd.abi import InputHandle


def query(spec, o, d, s, precision, mode):
    inp = InputHandle(spec)
    n = len(s)
    found = np.zeros(n, np.int32)
    dist = np.zeros(n)
    surf = np.zeros(n, np.int32)
    tri = np.zeros(n, np.int32)
    nrm = np.zeros(3 * n)
    rc = ol.oracle().xmco_query(inp.ptr, ol.dptr(o), ol.dptr(d), ol.dptr(s), n, precision, mode, ol.iptr(found),
                                ol.dptr(dist), ol.iptr(surf), ol.iptr(tri), ol.dptr(nrm))
    return rc, found, dist, surf, nrm.reshape(n, 3)


PLANAR = {
    "slab": lambda: scenes.slab_scene(10),
    "glass": lambda: scenes.small_scene("glass", 10),
    "mismatch": lambda: scenes.small_scene("mismatch", 10),
    "layered": lambda: scenes.batched_configs(3, seed=5)[2],
}


def rays(spec, n, seed):
    """Origins spread over the layers (within 20 mm of the axis), isotropic
    directions, three in four steps exponential and the rest unlimited."""
    rng = np.random.default_rng(seed)
    zs = sorted({float(v[0, 2]) for v, _ in spec.surfaces})
    z = rng.uniform(zs[0] - 0.5, zs[-1] + 0.5, n)
    o = np.stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n), z], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # float inputs, as the engine has them
    o = o.astype(np.float32).astype(np.float64)
    d = d.astype(np.float32).astype(np.float64)
    s = rng.exponential(1.0, n)
    s[::4] = 0.0  # unlimited
    return o, d, s


@pytest.mark.parametrize("kind", sorted(PLANAR))
def test_planar_rule_picks_the_surface_the_walk_picks(kind):
    spec = PLANAR[kind]()
    o, d, s = rays(spec, 20000, 3)
    rc, fp, tp, sp, np_ = query(spec, o, d, s, 32, 2)
    assert rc == 0
    rc, fb, tb, sb, nb = query(spec, o, d, s, 32, 0)
    assert rc == 0
    rc, fd, td, sd, _ = query(spec, o, d, s, 64, 0)
    assert rc == 0
    assert fp.sum() > 2000
    # a disagreement needs the hit within float rounding of the step's end
    near_end = (s > 0) & (np.abs(np.where(fp | fb, np.maximum(tp, tb), 0.0) - s) < 1e-5 * np.maximum(s, 1.0))
    # distances agree to the float resolution of the geometry along the ray: the
    # planes' z and the origin carry half an ulp each, divided by |d.z|
    zmax = max(float(np.abs(v[:, 2]).max()) for v, _ in spec.surfaces)
    ulp = np.spacing(np.float32(max(zmax, np.abs(o[:, 2]).max()))).astype(np.float64)
    tol = 4.0 * ulp / np.maximum(np.abs(d[:, 2]), 1e-6) + 2e-5 * np.abs(tp)
    for found, dist, surf in ((fb, tb, sb), (fd, td, sd)):
        same = (found == fp) | near_end
        assert same.all(), np.nonzero(~same)[0][:5]
        both = (found == 1) & (fp == 1)
        assert np.array_equal(surf[both], sp[both])
        assert np.all(np.abs(tp[both] - dist[both]) <= tol[both]), np.max(np.abs(tp[both] - dist[both]) / tol[both])
    # the normal is the surface's: (0, 0, +-1) within float rounding
    hit = fp == 1
    np.testing.assert_allclose(np.abs(np_[hit, 2]), 1.0, atol=1e-6)
    assert np.all(np_[hit, :2] == 0.0)


def test_planar_rule_applies_only_to_planar_scenes():
    spec = scenes.small_scene("heightfield", 10)
    o, d, s = rays(spec, 10, 4)
    rc, *_ = query(spec, o, d, s, 32, 2)
    assert rc != 0
    # a rectangle split along a side, not a diagonal, is not planar in this sense
    slab = scenes.slab_scene(10)
    v, t = slab.surfaces[1]
    slab.surfaces[1] = (v, np.array([[0, 1, 2], [0, 1, 3]], np.int32))
    rc, *_ = query(slab, o, d, s, 32, 2)
    assert rc != 0
    # nor is a tilted one
    slab = scenes.slab_scene(10)
    v, t = slab.surfaces[2]
    v = v.copy()
    v[3, 2] += 1e-9
    slab.surfaces[2] = (v, t)
    rc, *_ = query(slab, o, d, s, 32, 2)
    assert rc != 0
