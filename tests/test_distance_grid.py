"""The distance grid used to skip BVH walks must never overstate the distance
to the nearest triangle, and near the mesh it must be the exact distance from
each cell centre (host computation, no GPU)."""
import numpy as np

from xmcml_amd import engine, scenes


def closest_point_distance(p, a, b, c):
    """Distance from point p to every triangle (a, b, c): closest point on a
    triangle by Voronoi regions (Ericson, Real-Time Collision Detection 5.1.5)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = (ab * ap).sum(1), (ac * ap).sum(1)
    bp = p - b
    d3, d4 = (ab * bp).sum(1), (ac * bp).sum(1)
    cp = p - c
    d5, d6 = (ab * cp).sum(1), (ac * cp).sum(1)
    vc = d1 * d4 - d3 * d2
    vb = d5 * d2 - d1 * d6
    va = d3 * d6 - d5 * d4
    denom = va + vb + vc
    with np.errstate(divide="ignore", invalid="ignore"):
        v = vb / denom
        w = vc / denom
        q = a + ab * v[:, None] + ac * w[:, None]
        # edges and vertices
        t_ab = d1 / (d1 - d3)
        t_ac = d2 / (d2 - d6)
        t_bc = (d4 - d3) / ((d4 - d3) + (d5 - d6))
    q = np.where(((va <= 0) & (d4 - d3 >= 0) & (d5 - d6 >= 0))[:, None], b + (c - b) * t_bc[:, None], q)
    q = np.where(((vb <= 0) & (d2 >= 0) & (d6 <= 0))[:, None], a + ac * t_ac[:, None], q)
    q = np.where(((vc <= 0) & (d1 >= 0) & (d3 <= 0))[:, None], a + ab * t_ab[:, None], q)
    q = np.where(((d6 >= 0) & (d5 <= d6))[:, None], c, q)
    q = np.where(((d3 >= 0) & (d4 <= d3))[:, None], b, q)
    q = np.where(((d1 <= 0) & (d2 <= 0))[:, None], a, q)
    return np.sqrt(((p - q) ** 2).sum(1))


def _grid_and_triangles(target=200_000):
    spec = scenes.small_scene("heightfield", 10)
    g = engine.distance_grid(spec, target_cells=target)
    tri = np.concatenate([v[t] for v, t in spec.surfaces])  # (T, 3, 3)
    return g, tri[:, 0], tri[:, 1], tri[:, 2]


def test_bound_never_exceeds_true_distance():
    """For any point p, q * step_q of its cell less |p - centre| never
    exceeds p's distance to the nearest triangle (brute force)."""
    g, a, b, c = _grid_and_triangles()
    rng = np.random.default_rng(1)
    dims = np.array(g["dims"])
    hi = g["lo"] + dims * g["h"]
    pts = rng.uniform(g["lo"], hi, size=(3000, 3))
    # points concentrated near the surfaces too
    pts[:1500, 2] = rng.uniform(-1.5, 3.5, 1500)
    idx = np.floor((pts - g["lo"]) / g["h"]).astype(int)
    ok = np.all((idx >= 0) & (idx < dims), axis=1)
    pts, idx = pts[ok], idx[ok]
    centre = g["lo"].astype(np.float64) + (idx + 0.5) * g["h"]
    bound = (g["q"][idx[:, 2], idx[:, 1], idx[:, 0]].astype(np.float64) * g["step_q"]
             - np.sqrt(((pts - centre) ** 2).sum(1)))
    culled = 0
    for p, L in zip(pts, bound):
        if L <= 0:
            continue
        d = closest_point_distance(np.broadcast_to(p, a.shape), a, b, c).min()
        assert L <= d + 1e-9, (p, L, d)
        culled += 1
    assert culled > 1000  # the bound is useful: most random points are far from every surface


def test_centre_distance_exact_near_the_surfaces():
    """Within a few cells of the mesh the stored bound is the exact centre
    distance floored to the quantum (a box distance or a misplaced centre
    would be off by far more), and farther out it never undercuts the old
    point-of-cell bound by more than a cell."""
    g, a, b, c = _grid_and_triangles()
    dims = np.array(g["dims"])
    rng = np.random.default_rng(2)
    idx = np.stack([rng.integers(0, dims[0], 4000), rng.integers(0, dims[1], 4000),
                    rng.integers(0, dims[2], 4000)], axis=1)
    centre = g["lo"].astype(np.float64) + (idx + 0.5) * g["h"]
    keep = (centre[:, 2] > -1.5) & (centre[:, 2] < 3.5)
    idx, centre = idx[keep][:800], centre[keep][:800]
    stored = g["q"][idx[:, 2], idx[:, 1], idx[:, 0]].astype(np.float64) * g["step_q"]
    true = np.array([closest_point_distance(np.broadcast_to(p, a.shape), a, b, c).min() for p in centre])
    assert np.all(stored <= true + 1e-9)
    near = true < 3.5 * g["h"]
    assert near.sum() > 50
    # float storage of the distance loses at most a few ulps before flooring
    assert np.all(stored[near] > true[near] - g["step_q"] - 1e-5)
    far = ~near & (true < 200 * g["step_q"])
    assert np.all(stored[far] > true[far] - 1.8 * g["h"])
