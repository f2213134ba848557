"""Synthetic inputs of the documented workloads, as xmcml scenes.

xmcml describes every boundary as a triangulated surface (Surface,
base/xmcml/xmcml/mcml_types.h:33-39) and every layer by the list of surfaces
that bound it (LayerInfo, mcml_kernel_types.h:54-62); layer 0 is the ambient
medium.  The generators here build those meshes:

* ``plane`` -- a 2-triangle square in the vertex/triangle order of the
  reference's surface generator (base/tools/xmcmlSurfaceGenerator/
  xmcmlSurfaceGenerator/plane.cpp:3-27).
* ``heightfield`` -- z = z0 + h(x, y) on a regular grid, two triangles per cell.
* ``ellipsoid`` -- an icosphere scaled to an ellipsoid with seeded radial noise.

The heightfield workload uses one shared perturbation h for the three
perturbed boundaries (conformal layers): layer 1 is 0.1 mm thick while the sum
of the four amplitudes can reach 1.2 mm, so independent surfaces would cut
through each other.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .abi import (CubeSpec, LayerSpec, RingSpec, SceneSpec, TERMINATE_KILL, TERMINATE_ROULETTE)

SPEED_OF_LIGHT_MM_PER_NS = 299.792458

# The three tissue layers of the slab and heightfield workloads: (n, mua, mus, g).
SLAB_LAYERS = [(1.4, 0.3, 20.0, 0.8), (1.4, 0.05, 15.0, 0.9), (1.37, 0.02, 10.0, 0.9)]
SLAB_DEPTHS = [0.0, 0.1, 2.1, 22.1]


def plane(z: float, side: float = 2000.0, cx: float = 0.0, cy: float = 0.0) -> Tuple[np.ndarray, np.ndarray]:
    h = side / 2.0
    v = np.array([[cx - h, cy - h, z], [cx - h, cy + h, z], [cx + h, cy - h, z], [cx + h, cy + h, z]],
                 dtype=np.float64)
    t = np.array([[0, 1, 2], [1, 2, 3]], dtype=np.int32)
    return v, t


def grid_triangles(n: int) -> np.ndarray:
    """Two triangles per cell of an (n+1) x (n+1) vertex grid (row-major in y)."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    v00 = (i * (n + 1) + j).ravel()
    v10 = ((i + 1) * (n + 1) + j).ravel()
    v01 = (i * (n + 1) + j + 1).ravel()
    v11 = ((i + 1) * (n + 1) + j + 1).ravel()
    t = np.empty((2 * n * n, 3), dtype=np.int32)
    t[0::2] = np.stack([v00, v10, v01], 1)
    t[1::2] = np.stack([v10, v11, v01], 1)
    return t


class SineHeight:
    """h(x, y) = sum_k A_k sin(kx_k x + ky_k y + phi_k), drawn from a seed:
    A_k ~ U[0.05, 0.3] mm, |k| ~ U[0.2, 2] /mm, direction and phase ~ U[0, 2 pi)."""

    def __init__(self, seed: int, terms: int = 4):
        rng = np.random.default_rng(seed)
        self.A = rng.uniform(0.05, 0.3, terms)
        kmag = rng.uniform(0.2, 2.0, terms)
        ang = rng.uniform(0.0, 2 * math.pi, terms)
        self.kx = kmag * np.cos(ang)
        self.ky = kmag * np.sin(ang)
        self.phi = rng.uniform(0.0, 2 * math.pi, terms)

    def __call__(self, x, y):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        out = np.zeros(np.broadcast(x, y).shape)
        for a, kx, ky, p in zip(self.A, self.kx, self.ky, self.phi):
            out = out + a * np.sin(kx * x + ky * y + p)
        return out


def heightfield(z0: float, height: SineHeight, cells: int = 128, half: float = 20.0):
    xs = np.linspace(-half, half, cells + 1)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    Z = z0 + height(X, Y)
    v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float64)
    return v, grid_triangles(cells)


def icosphere(level: int) -> Tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: the icosahedron with every face split into 4, `level`
    times (20 * 4^level triangles, 10 * 4^level + 2 vertices)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = len(v) + inv.reshape(3, -1)  # midpoint of edge (0,1), (1,2), (2,0) of every face
        v = np.concatenate([v, mid])
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = m[0], m[1], m[2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    return v, f.astype(np.int32)


class SphereNoise:
    """Smooth radial displacement on the unit sphere: a sum of `terms` seeded
    plane waves sin(k . u + phi) with |k| ~ U[2, 6], scaled so that its
    largest value over the mesh is `amplitude` mm."""

    def __init__(self, seed: int, amplitude: float = 1.0, terms: int = 8):
        rng = np.random.default_rng(seed)
        d = rng.standard_normal((terms, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        self.k = d * rng.uniform(2.0, 6.0, (terms, 1))
        self.phi = rng.uniform(0.0, 2 * math.pi, terms)
        self.w = rng.uniform(0.5, 1.0, terms)
        self.amplitude = amplitude

    def raw(self, u):
        return np.sin(u @ self.k.T + self.phi) @ self.w

    def __call__(self, u):
        r = self.raw(u)
        return self.amplitude * r / np.max(np.abs(r))


def ellipsoid(axes, unit_vertices, faces, noise: Optional[np.ndarray] = None):
    """The ellipsoid x^2/a^2 + y^2/b^2 + z^2/c^2 = 1 through the icosphere's
    directions, each vertex moved `noise` mm along its radial direction."""
    p = unit_vertices * np.asarray(axes, dtype=np.float64)
    if noise is not None:
        r = np.linalg.norm(p, axis=1, keepdims=True)
        p = p * (1.0 + noise[:, None] / r)
    return p, faces


def axis_hit_z(v: np.ndarray, f: np.ndarray) -> float:
    """z of the lowest point where the z axis crosses the mesh (Moller-Trumbore
    in double precision over every triangle)."""
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    e1, e2 = b - a, c - a
    d = np.array([0.0, 0.0, 1.0])
    o = np.array([0.0, 0.0, -1e4])
    pv = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, pv)
    ok = np.abs(det) > 1e-15
    inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
    tv = o - a
    uu = np.einsum("ij,ij->i", tv, pv) * inv
    qv = np.cross(tv, e1)
    vv = (qv @ d) * inv
    t = np.einsum("ij,ij->i", e2, qv) * inv
    hit = ok & (uu >= 0) & (vv >= 0) & (uu + vv <= 1) & (t > 0)
    if not hit.any():
        raise ValueError("the z axis misses the mesh")
    return float(o[2] + t[hit].min())


# The heavy-geometry workload: outer semi-axes and the thickness of each
# shell from the outside in (mm), and (n, mua, mus, g) of the five tissues.
HEAD_AXES = (90.0, 75.0, 65.0)
HEAD_SHELLS = [5.0, 7.0, 2.0, 3.0]
HEAD_TISSUES = [(1.4, 0.02, 19.0, 0.9), (1.4, 0.01, 16.0, 0.9), (1.33, 0.004, 0.3, 0.9), (1.4, 0.02, 22.0, 0.9),
                (1.4, 0.008, 91.0, 0.9)]


def head_scene(photons: int = 1_000_000_000, level: int = 7, seed: int = 5, time_bins: int = 256,
               time_ns: float = 5.0, detector_radii=(10.0, 20.0, 30.0, 40.0), ring_width: float = 2.0) -> SceneSpec:
    """Five nested closed ellipsoids (scalp, skull, CSF, grey, white matter
    from the outside in), each an icosphere of level `level` (327,680
    triangles at level 7) with the same seeded 1 mm radial noise on every
    shell, so the shells keep their thicknesses and never cross.  Pencil
    beam entering at the bottom pole along +z; ring detectors of
    `ring_width` mm centred at the given source distances, around the beam
    axis; a time-of-flight histogram of time_bins bins over 0..time_ns."""
    u, f = icosphere(level)
    nz = SphereNoise(seed)(u)
    surfaces = []
    axes = np.array(HEAD_AXES)
    for k in range(len(HEAD_SHELLS) + 1):
        surfaces.append(ellipsoid(axes, u, f, nz))
        if k < len(HEAD_SHELLS):
            axes = axes - HEAD_SHELLS[k]
    ns = len(surfaces)
    layers = [LayerSpec(1.0, 0.0, 0.0, 0.0, [0])]
    for i, (ri, mua, mus, g) in enumerate(HEAD_TISSUES):
        ids = [i, i + 1] if i + 1 < ns else [i]
        layers.append(LayerSpec(ri, mua, mus, g, ids))
    # entry point: where the -z axis meets the outer mesh (on its flat triangle, not on the
    # ideal ellipsoid, so that the launch offset puts the photon inside the scalp)
    start_z = axis_hit_z(*surfaces[0])
    half = [a + 3.0 for a in HEAD_AXES]
    dets = [RingSpec(r - ring_width / 2, r + ring_width / 2, (0.0, 0.0, 0.0)) for r in detector_radii]
    return SceneSpec(
        layers=layers, surfaces=surfaces, area_corner=tuple(-h for h in half), area_length=tuple(2 * h for h in half),
        area_partitions=(1, 1, 1), rings=dets, min_weight=1e-4, termination=TERMINATE_ROULETTE,
        roulette_chance=0.1, time_start=0.0, time_finish=SPEED_OF_LIGHT_MM_PER_NS * time_ns, time_bins=time_bins,
        start_position=(0.0, 0.0, start_z), photons=photons, name=f"head-5shell-l{level}")


def rings(count: int, width: float, center=(0.0, 0.0, 0.0)) -> List[RingSpec]:
    return [RingSpec(i * width, (i + 1) * width, center) for i in range(count)]


def _slab_layers(layers=SLAB_LAYERS) -> List[LayerSpec]:
    n = len(layers)
    out = [LayerSpec(1.0, 0.0, 0.0, 0.0, [0, n])]
    for i, (ri, mua, mus, g) in enumerate(layers):
        out.append(LayerSpec(ri, mua, mus, g, [i, i + 1]))
    return out


def slab_scene(photons: int = 1_000_000, time_bins: int = 50, time_finish: float = 100.0) -> SceneSpec:
    """3-layer planar slab, pencil beam, 20 rings of 0.5 mm, 100^3 grid over
    20 x 20 x 20 mm, roulette at 1e-4 with survival 0.1.  Time is optical
    path in mm (mcml_kernel.cpp:161)."""
    surfaces = [plane(z) for z in SLAB_DEPTHS]
    return SceneSpec(
        layers=_slab_layers(), surfaces=surfaces, area_corner=(-10.0, -10.0, 0.0),
        area_length=(20.0, 20.0, 20.0), area_partitions=(100, 100, 100), rings=rings(20, 0.5),
        min_weight=1e-4, termination=TERMINATE_ROULETTE, roulette_chance=0.1, time_start=0.0,
        time_finish=time_finish, time_bins=time_bins, photons=photons, name="slab-3layer")


def heightfield_scene(photons: int = 100_000_000, seed: int = 2024, cells: int = 128,
                      time_bins: int = 256, time_ns: float = 2.0) -> SceneSpec:
    """Three conformal seeded heightfield boundaries (z0 = 0, 0.1, 2.1 mm) on a
    cells x cells grid over 40 x 40 mm plus a flat bottom at 22.1 mm; 20 rings
    of 0.5 mm; a time-of-flight histogram of time_bins bins over 0..time_ns
    at speed c/n, i.e. over optical path 0..c*time_ns mm."""
    h = SineHeight(seed)
    surfaces = [heightfield(z, h, cells) for z in SLAB_DEPTHS[:3]]
    surfaces.append(plane(SLAB_DEPTHS[3], side=40.0))
    z_top = float(h(0.0, 0.0))
    return SceneSpec(
        layers=_slab_layers(), surfaces=surfaces, area_corner=(-20.0, -20.0, -2.0),
        area_length=(40.0, 40.0, 24.2), area_partitions=(1, 1, 1), rings=rings(20, 0.5),
        min_weight=1e-4, termination=TERMINATE_ROULETTE, roulette_chance=0.1, time_start=0.0,
        time_finish=SPEED_OF_LIGHT_MM_PER_NS * time_ns, time_bins=time_bins,
        start_position=(0.0, 0.0, z_top), photons=photons, name="heightfield-3layer")


def small_scene(kind: str, photons: int = 2000) -> SceneSpec:
    """Small cases that exercise rarely taken branches."""
    if kind == "glass":
        # a non-scattering glass layer on top of tissue (mcml_kernel.cpp:50-77)
        layers = [(1.5, 0.0, 0.0, 0.0), (1.4, 0.1, 10.0, 0.85), (1.4, 0.02, 5.0, 0.9)]
        s = slab_scene(photons)
        s.layers = _slab_layers(layers)
        s.surfaces = [plane(z) for z in [0.0, 1.0, 3.0, 10.0]]
        s.name = "glass-top"
        return s
    if kind == "mismatch":
        # strong index steps: total internal reflection inside the slab
        layers = [(1.6, 0.2, 8.0, 0.7), (1.1, 0.05, 5.0, 0.9), (1.45, 0.02, 10.0, 0.0)]
        s = slab_scene(photons)
        s.layers = _slab_layers(layers)
        s.cubes = [CubeSpec((0.0, 0.0, 0.0), (2.0, 2.0, 0.01), filter_layers=(2,), target_angle=60.0)]
        s.target_range_layers = (2,)
        s.termination = TERMINATE_KILL
        s.min_weight = 1e-3
        s.name = "index-mismatch"
        return s
    if kind == "heightfield":
        s = heightfield_scene(photons, cells=32, time_bins=64)
        s.area_partitions = (40, 40, 20)
        s.name = "heightfield-small"
        return s
    if kind == "heightfield-glass":
        # the small heightfield with a glass top layer: every step in it is unlimited, so
        # the scene's surface rule hands it to the BVH walk
        s = heightfield_scene(photons, cells=32, time_bins=64)
        s.layers = _slab_layers([(1.5, 0.0, 0.0, 0.0), (1.4, 0.1, 10.0, 0.85), (1.37, 0.02, 5.0, 0.9)])
        # tilted off the vertical: straight down from the lattice vertex the beam starts on, it
        # would meet the conformal surface below at its vertex too, where six triangles tie
        s.start_direction = (0.01, 0.02, math.sqrt(1.0 - 0.0005))
        s.area_partitions = (40, 40, 20)
        s.name = "heightfield-glass"
        return s
    raise ValueError(kind)


def layered_scene(layers, photons: int, rings_n: int = 32, ring_width: float = 0.5, time_bins: int = 64,
                  time_ns: float = 1.0, side: float = 200.0, name: str = "layered") -> SceneSpec:
    """Planar stack of tissue layers [(n, mua, mus, g, thickness), ...] under an
    ambient n = 1 half-space, pencil beam at the origin, `rings_n` rings of
    `ring_width` mm, a `time_bins`-bin time-of-flight histogram over
    0..time_ns ns per ring, roulette 1e-4 / 0.1."""
    z = [0.0]
    for l in layers:
        z.append(z[-1] + l[4])
    surfaces = [plane(zz, side=side) for zz in z]
    nl = len(layers)
    specs = [LayerSpec(1.0, 0.0, 0.0, 0.0, [0, nl])]
    for i, (ri, mua, mus, g, _) in enumerate(layers):
        specs.append(LayerSpec(ri, mua, mus, g, [i, i + 1]))
    half = side / 2.0
    return SceneSpec(layers=specs, surfaces=surfaces, area_corner=(-half, -half, 0.0),
                     area_length=(side, side, z[-1] * (1.0 + 1e-9) + 1e-6), area_partitions=(1, 1, 1),
                     rings=rings(rings_n, ring_width), min_weight=1e-4, termination=TERMINATE_ROULETTE,
                     roulette_chance=0.1, time_start=0.0, time_finish=SPEED_OF_LIGHT_MM_PER_NS * time_ns,
                     time_bins=time_bins, photons=photons, name=name)


def draw_layers(rng: np.random.Generator, n_layers: int):
    """Layer parameters of the batched workload: log-uniform mua in [0.001, 1]/mm,
    reduced scattering mus' log-uniform in [0.5, 5]/mm (mus = mus'/(1-g)),
    g uniform in [0.7, 0.95], n uniform in [1.33, 1.5], thickness log-uniform
    in [0.05, 10] mm."""
    out = []
    for _ in range(n_layers):
        mua = math.exp(rng.uniform(math.log(1e-3), math.log(1.0)))
        musp = math.exp(rng.uniform(math.log(0.5), math.log(5.0)))
        g = rng.uniform(0.7, 0.95)
        n = rng.uniform(1.33, 1.5)
        d = math.exp(rng.uniform(math.log(0.05), math.log(10.0)))
        out.append((n, mua, musp / (1.0 - g), g, d))
    return out


def batched_configs(count: int = 1024, seed: int = 3, photon_scale: float = 1.0,
                    min_photons: float = 1e5, max_photons: float = 1e7) -> List[SceneSpec]:
    """The batched workload: `count` layered configurations of 2-5 layers with
    parameters from `draw_layers` and photon budgets log-uniform in
    [min_photons, max_photons] (times photon_scale), all drawn from `seed`.
    Outputs per configuration: 32 rings of 0.5 mm and a 64-bin TOF over 0-1 ns."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(count):
        layers = draw_layers(rng, int(rng.integers(2, 6)))
        budget = math.exp(rng.uniform(math.log(min_photons), math.log(max_photons)))
        out.append(layered_scene(layers, max(1, int(budget * photon_scale)), name=f"batch-{seed}-{c}"))
    return out


def biased(spec: SceneSpec, target=(3.0, 0.0, 0.0), factor: float = 1.0, tables=None,
           integral_precision: int = 128, table_precision: int = 262144) -> SceneSpec:
    """The same input with attractive-vector biasing towards `target`
    (InputInfo.useBiasing..targetPoint); tables = [(anisotropy, elements)] or
    None for tables the library computes at the given precisions."""
    import copy
    s = copy.copy(spec)
    s.use_biasing = True
    s.attractive_factor = factor
    s.target_point = tuple(float(v) for v in target)
    s.weight_integral_precision = integral_precision
    s.weight_table_precision = table_precision
    s.weight_tables = list(tables or [])
    s.name = spec.name + "-biased"
    return s


def distinct_anisotropies(spec: SceneSpec):
    """GetDifferentAnisotropies (weight_integral.cpp:67-90): layers 1.., first seen first."""
    out = []
    for l in spec.layers[1:]:
        if l.g not in out:
            out.append(l.g)
    return out
