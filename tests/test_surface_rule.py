"""The engine's float rule for scenes whose every surface is an axis-aligned
rectangle at one z (two triangles sharing a diagonal) or a heightfield on a
uniform lattice shared by all of them: the nearest rectangle is found
analytically, the nearest heightfield triangle by testing the triangles of the
lattice cells the step covers, instead of by the BVH walk.  On the CPU
oracle's 32-bit mode (which the GPU reproduces bit for bit) the rule must pick
the surface the BVH walk picks -- in float and in the reference's double
arithmetic -- at distances equal to float rounding (bit-identical for the
heightfields, whose triangles the rule tests as the walk does), and it must
apply to exactly those scenes."""
import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd import scenes
from xmcml_amd.abi import InputHandle, OutputHandle


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
LATTICE = {
    "heightfield": lambda: scenes.heightfield_scene(10),
    "heightfield-small": lambda: scenes.small_scene("heightfield", 10),
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


@pytest.mark.parametrize("kind", sorted(LATTICE))
def test_lattice_rule_finds_the_walks_hit(kind):
    """Heightfields: the same hit, distance and triangle as the 32-bit BVH
    walk on every ray (the rule tests the walk's triangles with the walk's
    test), and the surface of the reference's double walk but within float
    rounding of the step's end."""
    spec = LATTICE[kind]()
    rng = np.random.default_rng(9)
    n = 60000
    half = 21.0 if kind == "heightfield" else 5.0
    o = np.stack([rng.uniform(-half, half, n), rng.uniform(-half, half, n), rng.uniform(-1.6, 3.6, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = o.astype(np.float32).astype(np.float64)
    d = d.astype(np.float32).astype(np.float64)
    s = rng.exponential(0.2, n) * np.where(rng.random(n) < 0.1, 10.0, 1.0)
    s[::7] = 0.0
    rc, fr, tr, sr, nr = query(spec, o, d, s, 32, 2)
    assert rc == 0
    rc, fb, tb, sb, nb = query(spec, o, d, s, 32, 0)
    assert rc == 0
    assert fr.sum() > 5000
    assert np.array_equal(fr, fb)
    assert np.array_equal(tr, tb)
    assert np.array_equal(sr, sb)
    assert np.array_equal(nr, nb)
    rc, fd, td, sd, _ = query(spec, o, d, s, 64, 0)
    assert rc == 0
    near_end = (s > 0) & (np.abs(np.where(fr | fd, np.maximum(tr, td), 0.0) - s) < 1e-5 * np.maximum(s, 1.0))
    assert np.all((fd == fr) | near_end)
    both = (fd == 1) & (fr == 1)
    assert np.mean(sd[both] == sr[both]) > 0.9999


def test_rule_applies_only_to_its_scenes():
    spec = scenes.small_scene("heightfield", 10)
    o, d, s = rays(spec, 10, 4)
    rc, *_ = query(spec, o, d, s, 32, 2)
    assert rc == 0
    # switched off, heightfield scenes walk
    with ol.lattice_rule(False):
        rc, *_ = query(spec, o, d, s, 32, 2)
        assert rc != 0
        rc, *_ = query(scenes.slab_scene(10), o, d, s, 32, 2)
        assert rc == 0
    # a lattice line moved off the uniform spacing
    spec = scenes.small_scene("heightfield", 10)
    v, t = spec.surfaces[1]
    v = v.copy()
    v[np.isclose(v[:, 0], v[:, 0].min()), 0] -= 1e-3
    spec.surfaces[1] = (v, t)
    rc, *_ = query(spec, o, d, s, 32, 2)
    assert rc != 0
    # a triangle across two cells
    spec = scenes.small_scene("heightfield", 10)
    v, t = spec.surfaces[0]
    t = t.copy()
    t[0] = [t[0, 0], t[0, 1], t[2, 2]]  # (0, 0), (1, 0), (0, 2)
    spec.surfaces[0] = (v, t)
    rc, *_ = query(spec, o, d, s, 32, 2)
    assert rc != 0
    # a third triangle in a cell
    spec = scenes.small_scene("heightfield", 10)
    v, t = spec.surfaces[0]
    spec.surfaces[0] = (v, np.concatenate([t, t[:1]]))
    rc, *_ = query(spec, o, d, s, 32, 2)
    assert rc != 0
    # a closed mesh (no lattice)
    rc, *_ = query(scenes.head_scene(10, level=2), o, d, s, 32, 2)
    assert rc != 0
    # a rectangle split along a side, not a diagonal, is not a rectangle in this
    # sense, nor is a tilted one: each is a one-cell lattice, tested as the walk tests it
    split = scenes.slab_scene(10)
    v, t = split.surfaces[1]
    split.surfaces[1] = (v, np.array([[0, 1, 2], [0, 1, 3]], np.int32))
    tilted = scenes.slab_scene(10)
    v, t = tilted.surfaces[2]
    v = v.copy()
    v[3, 2] += 1e-9
    tilted.surfaces[2] = (v, t)
    for spec, lattice_surface in ((split, 1), (tilted, 2)):
        o, d, s = rays(spec, 20000, 5)
        rc, fr, tr, sr, _ = query(spec, o, d, s, 32, 2)
        assert rc == 0
        with ol.lattice_rule(False):
            rc, *_ = query(spec, o, d, s, 32, 2)
            assert rc != 0
        rc, fb, tb, sb, _ = query(spec, o, d, s, 32, 0)
        lattice_hit = (fb == 1) & (sb == lattice_surface) & (fr == 1) & (sr == sb)
        assert lattice_hit.sum() > 200
        assert np.array_equal(tr[lattice_hit], tb[lattice_hit])


@pytest.mark.parametrize("kind", sorted(LATTICE))
def test_lattice_rule_changes_no_tally(kind):
    """The 32-bit oracle's transport with the lattice rule and with the BVH
    walk: 3,000 photons give the same event counts, detector counts and
    weights, bin for bin (the rule finds the walk's hits)."""
    n = 3000
    spec = LATTICE[kind]()
    spec.photons = n
    res = []
    for on in (True, False):
        with ol.lattice_rule(on):
            inp = InputHandle(spec)
            o = OutputHandle(spec)
            c = ol.run_oracle(spec, inp, o, 5, 0, n, 32, 1)  # one thread: a fixed summation order
            res.append((c.scatterEvents, c.boundaryEvents, c.rouletteEvents, o.det_count().copy(),
                        o.det_weight().copy(), o.bin_weight().copy(), o.absorption.copy()))
    a, b = res
    assert a[0] == b[0] and a[1] == b[1] and a[2] == b[2]
    assert a[1] > 8 * n
    for x, y in zip(a[3:], b[3:]):
        assert np.array_equal(x, y)
