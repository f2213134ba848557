"""GPU engine against the double oracle under independent streams, and the
GPU's roulette against a walk without termination, at the slab
configuration's full 1e6 photons.

* Roulette (stable/plane_parallel/1.0/McmlGpu/kernel/kernel.cu:214-222) is
  unbiased: slab rings, total reflectance and absorption with roulette at
  1e-4 / 0.1 equal those of a walk that only stops below 1e-12 within three
  combined standard errors.
* The 100^3 absorption grid (the slab configuration's 8 MB tally), voxel by
  voxel: GPU photons of seed 1 against double-oracle photons of seed 2, 1e6
  on each side in 20 batches.  Over the ~4e5 populated voxels about a
  thousand |z| > 3 are expected by chance, so the voxels are tested as a family:
  the share beyond 3 is at most 1 % (0.3 % for a normal z, ~0.5 % for the t
  tail of 20-batch variances), the mean z^2 is within [0.8, 1.3], none is
  beyond 7 (a t tail over 4e5 voxels reaches ~5.6: measured with the 32- and
  64-bit oracles on the CPU), and the grid totals agree within 0.5 %.  A
  voxel index with two axes swapped (mean z^2 211 in that measurement), a
  deposit of the wrong fraction or a dropped deposit fails it.
"""
import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd import scenes
from xmcml_amd.abi import InputHandle, OutputHandle, TERMINATE_KILL, TERMINATE_ROULETTE

pytestmark = pytest.mark.gpu


def _zscores(A, B):
    nb_a, nb_b = len(A), len(B)
    se = np.sqrt(A.var(0, ddof=1) / nb_a + B.var(0, ddof=1) / nb_b)
    return np.abs(A.mean(0) - B.mean(0)) / np.where(se > 0, se, 1.0)


def test_gpu_roulette_matches_a_walk_without_termination(engine):
    nb, n = 10, 100_000

    def batches(term, min_w, seed):
        spec = scenes.slab_scene(0)
        spec.termination = term
        spec.min_weight = min_w
        rows = []
        with engine.Scene(spec, absorption=False) as sc:
            for b in range(nb):
                sc.run(seed, b * n, n)
                g = sc.collect()
                st = sc.stats()
                rows.append(np.concatenate([g.det_weight() / n, [g.det_weight().sum() / n, st["wAbsorbed"] / n]]))
        return np.array(rows), st

    A, sa = batches(TERMINATE_ROULETTE, 1e-4, 21)
    B, sb = batches(TERMINATE_KILL, 1e-12, 22)
    assert sa["rouletteEvents"] > 0 and sb["rouletteEvents"] == 0
    z = _zscores(A, B)
    pop = B.mean(0) > 1e-4
    print(f"roulette vs kill at 1e-12: max |z| {z[pop].max():.2f} over {pop.sum()} tallies; "
          f"R {A[:, -2].mean():.5f} vs {B[:, -2].mean():.5f}, A {A[:, -1].mean():.5f} vs {B[:, -1].mean():.5f}")
    assert pop.sum() >= 20
    assert np.all(z[pop] < 3.0), z[pop]
    assert abs(A[:, -2].mean() / B[:, -2].mean() - 1) < 5e-3
    assert abs(A[:, -1].mean() / B[:, -1].mean() - 1) < 5e-3


def test_absorption_grid_voxel_by_voxel_at_1e6_photons(engine):
    spec = scenes.slab_scene(0)
    inp = InputHandle(spec)
    nb, n = 20, 50_000
    print(f"absorption grid: GPU seed 1 vs double oracle seed 2, {nb} x {n} photons each", flush=True)
    G = np.empty((nb, spec.grid_size))
    with engine.Scene(spec) as sc:
        for b in range(nb):
            sc.run(1, b * n, n)
            G[b] = sc.collect().absorption / n
    C = np.empty((nb, spec.grid_size))
    for b in range(nb):
        o = OutputHandle(spec)
        ol.run_oracle(spec, inp, o, 2, b * n, n, 64, 0)
        C[b] = o.absorption / n
    populated = (G > 0).all(axis=0) & (C > 0).all(axis=0)
    z = _zscores(G[:, populated], C[:, populated])
    beyond3 = float(np.mean(z > 3.0))
    print(f"absorption grid: {populated.sum()} populated voxels, {beyond3 * 100:.2f} % with |z| > 3, "
          f"mean z^2 {np.mean(z ** 2):.3f}, max |z| {z.max():.2f}; totals {G.sum(1).mean():.6f} vs "
          f"{C.sum(1).mean():.6f}")
    assert populated.sum() >= 5000
    assert beyond3 <= 0.01
    assert 0.8 <= np.mean(z ** 2) <= 1.3
    assert z.max() < 7.0
    assert abs(G.sum(1).mean() / C.sum(1).mean() - 1) < 5e-3
