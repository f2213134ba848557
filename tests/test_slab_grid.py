"""The separating slabs of the cells near the mesh (host computation, no
GPU).  A cell's word claims that every triangle point within R of the cell
centre c has n.(x - c) inside one of its two bands; a step whose two ends lie
in one free interval and stay in the ball then crosses no triangle.  The
tests sample triangle points densely around many cells and check the claim,
and check on random steps that the transport's skip rule (restated here in
float32) never skips a step that crosses a triangle."""
import numpy as np
import pytest

from xmcml_amd import engine, scenes

MEAN_STEP = 0.0666


def sbyte(w, k):
    return ((w >> np.uint64(8 * k)) & np.uint64(0xFF)).astype(np.uint8).view(np.int8).astype(np.int64)


def direction(a, b):
    """slab_direction (xmc_distance.h) in float32."""
    f = np.float32
    u = a.astype(f) * f(1.0 / 127.0)
    v = b.astype(f) * f(1.0 / 127.0)
    z = f(1.0) - np.abs(u) - np.abs(v)
    fu = (f(1.0) - np.abs(v)) * np.where(u >= 0, f(1.0), f(-1.0))
    fv = (f(1.0) - np.abs(u)) * np.where(v >= 0, f(1.0), f(-1.0))
    neg = z < 0
    return np.stack([np.where(neg, fu, u), np.where(neg, fv, v), z], -1).astype(f)


@pytest.fixture(scope="module")
def grid():
    spec = scenes.small_scene("heightfield", 10)
    g = engine.slab_grid(spec, MEAN_STEP, target_cells=300_000)
    tri = np.concatenate([v[t] for v, t in spec.surfaces])  # (T, 3, 3)
    return g, tri


def decode(g, idx):
    w = g["words"][idx[:, 2], idx[:, 1], idx[:, 0]]
    n = direction(sbyte(w, 0), sbyte(w, 1))
    bands = np.stack([sbyte(w, 2), sbyte(w, 3), sbyte(w, 4), sbyte(w, 5)], 1)
    r = sbyte(w, 6) & 0xFF
    R = (r.astype(np.float32) * np.float32(g["unit"])).astype(np.float64)
    return w, n.astype(np.float64), bands, R


def test_words_are_built_near_the_mesh(grid):
    g, tri = grid
    w = g["words"]
    assert (w != 0).sum() > 1000
    r = (w >> np.uint64(48)) & np.uint64(0xFF)
    assert set(np.unique(r[w != 0]).tolist()) <= {8, 12, 16, 24, 32}
    assert (w >> np.uint64(56)).max() == 0


def test_triangle_points_within_the_ball_lie_in_the_bands(grid):
    """Dense samples of every triangle near 400 cells: each sample within R
    of the centre projects into a band (a sign, offset, unit or byte-order
    mistake moves the bands away from the mesh)."""
    g, tri = grid
    rng = np.random.default_rng(11)
    nz = np.argwhere(g["words"] != 0)[:, ::-1]  # (x, y, z)
    pick = nz[rng.choice(len(nz), 400, replace=False)]
    w, n, bands, R = decode(g, pick)
    centre = g["lo"].astype(np.float64) + (pick + 0.5) * np.float64(g["h"])
    # barycentric samples: vertices, edges and the inside
    uv = rng.random((3000, 2))
    flip = uv.sum(1) > 1
    uv[flip] = 1 - uv[flip]
    uv = np.concatenate([uv, [[0, 0], [1, 0], [0, 1], [0.5, 0], [0, 0.5], [0.5, 0.5]],
                         np.stack([np.linspace(0, 1, 64), np.zeros(64)], 1),
                         np.stack([np.zeros(64), np.linspace(0, 1, 64)], 1),
                         np.stack([np.linspace(0, 1, 64), 1 - np.linspace(0, 1, 64)], 1)])
    lo_t, hi_t = tri.min(1), tri.max(1)
    checked = 0
    for k in range(len(pick)):
        c, Rk = centre[k], R[k]
        near = np.all((lo_t <= c + Rk) & (hi_t >= c - Rk), axis=1)
        q = Rk / 127.0
        a1, b1, a2, b2 = bands[k] * q
        for t in tri[near]:
            pts = t[0] + uv[:, :1] * (t[1] - t[0]) + uv[:, 1:] * (t[2] - t[0])
            inside = ((pts - c) ** 2).sum(1) <= Rk * Rk
            if not inside.any():
                continue
            s = (pts[inside] - c) @ n[k]
            ok = ((s >= a1 - 1e-9) & (s <= b1 + 1e-9)) | ((s >= a2 - 1e-9) & (s <= b2 + 1e-9))
            assert ok.all(), (pick[k], s[~ok][:3], (a1, b1, a2, b2))
            checked += int(inside.sum())
    assert checked > 100_000


def segment_hits(o, d, s, tri):
    """Does the segment o + t d, t in [0, s], cross any triangle (double
    Moller-Trumbore, no tolerance)?"""
    a = tri[:, 0]
    e1, e2 = tri[:, 1] - a, tri[:, 2] - a
    p = np.cross(d, e2)
    det = (e1 * p).sum(1)
    ok = np.abs(det) > 1e-14
    inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
    tv = o - a
    u = (tv * p).sum(1) * inv
    q = np.cross(tv, e1)
    v = (q @ d) * inv
    t = (e2 * q).sum(1) * inv
    return bool(np.any(ok & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t >= 0) & (t <= s)))


def test_skip_rule_never_skips_a_crossing_step(grid):
    """Random steps from points near the surfaces: the transport's slab test
    (float32, margin 1e-4 + 16 * offset scale) skips many, and none of the
    skipped ones crosses a triangle."""
    g, tri = grid
    f = np.float32
    rng = np.random.default_rng(5)
    N = 20000
    lo, h = g["lo"], f(g["h"])
    dims = np.array(g["dims"])
    p = np.stack([rng.uniform(-3, 3, N), rng.uniform(-3, 3, N), rng.uniform(-0.8, 2.9, N)], 1).astype(f)
    d = rng.normal(size=(N, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(f)
    s = rng.exponential(MEAN_STEP, N).astype(f)
    fxyz = (p - lo) * f(1.0 / g["h"])
    idx = np.floor(fxyz).astype(int)
    ok = np.all((idx >= 0) & (idx < dims), axis=1)
    p, d, s, fxyz, idx = p[ok], d[ok], s[ok], fxyz[ok], idx[ok]
    w, n, bands, R = decode(g, idx)
    n = n.astype(f)
    o = (fxyz - (idx.astype(f) + f(0.5))) * h
    off = np.sqrt((o * o).sum(1))
    scale = np.abs(p).max(1) * f(2.0 ** -20) + f(2.0 ** -22)
    m = f(1e-4) + f(16.0) * scale
    Rf = R.astype(f)
    room = Rf - m - off
    in_ball = (room > 0) & (s * s * (d * d).sum(1) < room * room)
    t0 = (n * o).sum(1)
    t1 = t0 + s * (n * d).sum(1)
    tmin, tmax = np.minimum(t0, t1), np.maximum(t0, t1)
    qf = Rf * f(1.0 / 127.0)
    b = bands.astype(f)
    free = (tmax < b[:, 0] * qf - m) | ((tmin > b[:, 1] * qf + m) & (tmax < b[:, 2] * qf - m)) | (tmin > b[:, 3] * qf + m)
    skip = in_ball & free
    assert skip.sum() > 2000
    hits = 0
    for k in np.nonzero(skip)[0]:
        assert not segment_hits(p[k].astype(np.float64), d[k].astype(np.float64), float(s[k]), tri), k
    for k in np.nonzero(~skip)[0][:3000]:
        hits += segment_hits(p[k].astype(np.float64), d[k].astype(np.float64), float(s[k]), tri)
    assert hits > 50  # the sample does reach the surfaces
