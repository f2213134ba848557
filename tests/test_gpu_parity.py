"""GPU engine (through the C-ABI) against the oracle and the reference.

* bit for bit against the 32-bit oracle: BVH hits, per-step traces, event
  counts -- the same float arithmetic on the same MCG59 streams;
* against the reference's own double walk driven by the same uniforms (a
  committed fixture written from the reference sources): integer decisions
  identical over whole lives, weights within 1e-5 relative; and every GPU
  step replayed in double from the GPU's state within 1e-5;
* statistically, under independent streams: every ring bin and the total
  reflectance within 3 combined standard errors;
* the weight ledger closes and the LaunchOMP-style entry point accumulates.
"""
import math
from pathlib import Path

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
    "heightfield-walk": lambda n: scenes.small_scene("heightfield", n),
    "heightfield-glass": lambda n: scenes.small_scene("heightfield-glass", n),
}
# the heightfield with its lattice rule switched off on both sides: the
# interleaved kernel's BVH walks with the walk-skipping grids
WALK = {"heightfield-walk"}


def opts(name):
    return dict(lattice_rule=False) if name in WALK else {}


def fixture_name(name):
    return "heightfield" if name in WALK else name


def _fields_equal(a, b):
    if len(a) != len(b):
        return False
    return all(np.array_equal(a[k], b[k]) for k in a.dtype.names)


def test_bvh_is_the_reference_tree(engine):
    for spec in (scenes.small_scene("heightfield", 10), scenes.heightfield_scene(10)):
        with engine.Scene(spec) as sc:
            A = sc.bvh()
        B = ol.bvh_arrays(ol.oracle().xmco_bvh, InputHandle(spec))
        for k in A:
            assert np.array_equal(A[k], B[k]), k


def _random_rays(n, rng, lo, hi):
    o = rng.uniform(lo, hi, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    steps = rng.exponential(0.3, size=n)
    steps[rng.random(n) < 0.2] = 0.0  # unlimited (glass-layer query)
    return o, d, steps


def test_bvh_hits_bit_exact_vs_brute_force(engine):
    spec = scenes.heightfield_scene(10)  # 98,306 triangles
    rng = np.random.default_rng(7)
    o, d, s = _random_rays(20000, rng, np.array([-19.0, -19.0, -1.5]), np.array([19.0, 19.0, 3.0]))
    with engine.Scene(spec) as sc:
        bv = sc.query(o, d, s, brute=False)
        bf = sc.query(o, d, s, brute=True)
    assert bv["found"].sum() > 2000
    assert np.array_equal(bv["found"], bf["found"])
    assert np.array_equal(bv["surface"], bf["surface"])
    assert np.array_equal(bv["distance"], bf["distance"])
    ties = bv["triangle"] != bf["triangle"]
    assert ties.sum() <= 2  # equal distances on a shared edge may pick either triangle
    # and the same walk in the 32-bit oracle
    inp = InputHandle(spec)
    found = np.zeros(len(s), np.int32)
    dist = np.zeros(len(s))
    surf = np.zeros(len(s), np.int32)
    tri = np.zeros(len(s), np.int32)
    nrm = np.zeros((len(s), 3))
    ol.oracle().xmco_query(inp.ptr, ol.dptr(o), ol.dptr(d), ol.dptr(s), len(s), 32, 0, ol.iptr(found),
                           ol.dptr(dist), ol.iptr(surf), ol.iptr(tri), ol.dptr(nrm))
    assert np.array_equal(found, bv["found"])
    assert np.array_equal(dist, bv["distance"])
    assert np.array_equal(tri, bv["triangle"])
    assert np.array_equal(nrm, bv["normal"])


@pytest.mark.parametrize("name", list(SMALL))
def test_traces_bit_identical_to_float_oracle(engine, name):
    spec = SMALL[name](64)
    inp = InputHandle(spec)
    with engine.Scene(spec, **opts(name)) as sc:
        traces, counts = sc.trace(2024, 0, 64, max_events=20000)
    assert counts.max() < 20000
    with ol.lattice_rule(name not in WALK):
        bad = [p for p in range(64) if not _fields_equal(traces[p], ol.trace(ol.oracle().xmco_trace, inp, 2024, p,
                                                                              20000, precision=32))]
    assert not bad, f"photons whose GPU trace differs from the float oracle: {bad[:10]}"


@pytest.mark.parametrize("name", list(SMALL))
def test_tallies_match_float_oracle(engine, name):
    n = 4000
    spec = SMALL[name](n)
    inp = InputHandle(spec)
    with engine.Scene(spec, **opts(name)) as sc:
        sc.run(5, 100, n)
        g = sc.collect()
        st = sc.stats()
    o = OutputHandle(spec)
    with ol.lattice_rule(name not in WALK):
        c = ol.run_oracle(spec, inp, o, 5, 100, n, 32, 8)
    if name in WALK:
        assert st["culledWalks"] > 0 and st["entryWalks"] > 0
    assert st["photons"] == n
    assert st["scatterEvents"] == c.scatterEvents
    assert st["boundaryEvents"] == c.boundaryEvents
    assert st["rouletteEvents"] == c.rouletteEvents
    assert np.array_equal(g.det_count(), o.det_count())
    assert np.array_equal(g.bin_count(), o.bin_count())
    np.testing.assert_allclose(g.det_weight(), o.det_weight(), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(g.bin_weight(), o.bin_weight(), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(g.det_target(), o.det_target(), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(g.absorption, o.absorption, rtol=1e-10, atol=1e-12)
    assert g.specular == o.specular
    assert st["wAbsorbed"] == pytest.approx(c.wAbsorbed, rel=1e-10)
    assert st["wEscaped"] == pytest.approx(c.wEscaped, rel=1e-10)


TAPE_FIXTURE = Path(__file__).parent / "golden" / "tape_traces.npz"


@pytest.mark.parametrize("name", [n for n in SMALL if n != "heightfield-glass"])  # the fixture's scenes
def test_traces_match_the_reference_on_the_same_uniforms(engine, name):
    """GPU float walk vs the reference's own double walk on the same uniforms:
    tests/golden/tape_traces.npz holds the reference sources' traces of photons
    0..7 of seed 99 while they draw the engine's 24-bit uniforms (written by
    tests/golden/make_tape_traces.py from oracle/_ref/libxmcml_ref_tape.so).
    Integer decisions -- event type, layer, uniforms drawn -- identical over
    every photon's whole life; weights within 1e-5 relative throughout;
    positions and directions within 2e-5 until the second boundary
    interaction and within 1e-3 over the whole life (the float walk's surface
    offset and accumulated rounding; test_reference_tape.py measures both on
    512 lives)."""
    z = np.load(TAPE_FIXTURE, allow_pickle=False)
    seed, nphot = int(z["seed"]), int(z["photons"])
    spec = SMALL[name](64)
    with engine.Scene(spec, **opts(name)) as sc:
        traces, _ = sc.trace(seed, 0, nphot, max_events=200000)
    for p in range(nphot):
        r = ol.compare_walks(z[f"{fixture_name(name)}_{p}"], traces[p])
        assert r["full"], (p, r)
        assert r["weight"] < 1e-5 and r["early"] < 2e-5 and r["whole"] < 1e-3, (p, r)


@pytest.mark.parametrize("name", list(SMALL))
def test_every_gpu_step_matches_double(engine, name):
    """Each interaction of the GPU walk replayed in double from the GPU's
    recorded state with the same recorded uniforms and inputs: identical
    decisions at every step of every photon, position / direction / weight
    within 1e-5 relative."""
    nphot = 64
    spec = SMALL[name](nphot)
    inp = InputHandle(spec)
    with engine.Scene(spec, **opts(name)) as sc:
        traces, _ = sc.trace(4242, 0, nphot, max_events=50000)
    worst = 0.0
    for p in range(nphot):
        ok, w = ol.shadow_compare(inp, 4242, p, traces[p], 65)
        assert ok, p
        worst = max(worst, w)
    assert worst < 1e-5, worst


def _batches(run, nb):
    return np.array([run(i) for i in range(nb)])


def test_statistics_within_three_sigma(engine):
    """Independent streams: GPU (seed 1) vs the double oracle (seed 2)."""
    spec = scenes.slab_scene(0)
    inp = InputHandle(spec)
    nb, n_gpu, n_cpu = 10, 100000, 20000

    def gpu_batch(i):
        with engine.Scene(spec) as sc:
            sc.run(1, i * n_gpu, n_gpu)
            g = sc.collect()
            st = sc.stats()
        return np.concatenate([g.det_weight() / n_gpu, [g.det_weight().sum() / n_gpu, st["wAbsorbed"] / n_gpu]])

    def cpu_batch(i):
        o = OutputHandle(spec)
        c = ol.run_oracle(spec, inp, o, 2, i * n_cpu, n_cpu, 64, 0)
        return np.concatenate([o.det_weight() / n_cpu, [o.det_weight().sum() / n_cpu, c.wAbsorbed / n_cpu]])

    G = _batches(gpu_batch, nb)
    Cc = _batches(cpu_batch, nb)
    mg, mc = G.mean(0), Cc.mean(0)
    se = np.sqrt(G.var(0, ddof=1) / nb + Cc.var(0, ddof=1) / nb)
    z = np.abs(mg - mc) / np.where(se > 0, se, 1.0)
    assert np.all(z[mg > 1e-4] < 3.0), z
    # total diffuse reflectance and absorption agree to 0.5 %
    assert abs(mg[-2] / mc[-2] - 1) < 5e-3
    assert abs(mg[-1] / mc[-1] - 1) < 5e-3


@pytest.mark.parametrize("name", list(SMALL))
def test_weight_ledger_closes(engine, name):
    n = 20000
    spec = SMALL[name](n)
    with engine.Scene(spec, **opts(name)) as sc:
        sc.run(3, 0, n)
        g = sc.collect()
        st = sc.stats()
    total = g.specular + (st["wAbsorbed"] + st["wEscaped"] + st["wLeftArea"] + st["wKilled"] + st["wRouletteLost"]
                          + st["wOther"] - st["wRouletteGain"]) / n
    assert abs(total - 1.0) < 1e-5
    assert st["wDetected"] == pytest.approx(g.det_weight().sum(), rel=1e-9)


def test_launch_entry_point_accumulates_and_advances(engine):
    spec = scenes.slab_scene(0)
    out = OutputHandle(spec)
    nxt = engine.launch(spec, out, 77, 0, 3000)
    assert nxt == 3000
    nxt = engine.launch(spec, out, 77, nxt, 2000)
    assert nxt == 5000
    with engine.Scene(spec) as sc:
        sc.run(77, 0, 5000)
        one = sc.collect()
    assert np.array_equal(out.det_count(), one.det_count())
    np.testing.assert_allclose(out.det_weight(), one.det_weight(), rtol=1e-11)
    np.testing.assert_allclose(out.absorption, one.absorption, rtol=1e-10, atol=1e-13)


def test_heightfield_full_size_ledger_and_symmetry(engine):
    """The 98k-triangle heightfield workload at 2e6 photons: ledger closes,
    reflectance is positive in every ring, counters are consistent."""
    n = 2_000_000
    spec = scenes.heightfield_scene(n)
    with engine.Scene(spec, absorption=False) as sc:
        sc.run(11, 0, n)
        g = sc.collect()
        st = sc.stats()
    total = g.specular + (st["wAbsorbed"] + st["wEscaped"] + st["wLeftArea"] + st["wKilled"] + st["wRouletteLost"]
                          + st["wOther"] - st["wRouletteGain"]) / n
    assert abs(total - 1.0) < 1e-4
    assert st["photons"] == n
    assert np.all(g.det_weight() > 0)
    assert 0.999 * g.det_count().sum() <= g.bin_count().sum() <= g.det_count().sum()


@pytest.mark.parametrize("grids", [(True, True), (False, True), (True, False)])
def test_walk_skipping_grids_change_nothing(engine, grids):
    """The distance bound and the entry nodes only skip work: tallies and
    event counts equal those of walks that always start at the root."""
    dist, entry = grids
    n = 6000
    spec = scenes.small_scene("heightfield", n)
    res = []
    for o in (dict(distance_grid=False, entry_grid=False), dict(distance_grid=dist, entry_grid=entry)):
        with engine.Scene(spec, lattice_rule=False, **o) as sc:
            sc.run(31, 0, n)
            out = sc.collect()
            res.append((out, sc.stats()))
    (a, sa), (b, sb) = res
    for k in ("scatterEvents", "boundaryEvents", "rouletteEvents"):
        assert sa[k] == sb[k], k
    assert sa["culledWalks"] == 0 and sa["entryWalks"] == 0
    if entry:
        assert sb["entryWalks"] > 0
    if dist:
        assert sb["culledWalks"] > 0
    assert np.array_equal(a.det_count(), b.det_count())
    assert np.array_equal(a.bin_count(), b.bin_count())
    np.testing.assert_allclose(a.absorption, b.absorption, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(a.det_weight(), b.det_weight(), rtol=1e-11, atol=1e-12)


def test_entry_trees_change_nothing_on_the_full_heightfield(engine):
    """The 98,306-triangle heightfield, where a cell's entry is a restricted
    copy of the BVH (chains of one-child nodes skipped, chain-ending leaves
    behind gate nodes): traces of 64 photons and the tallies of 20,000 equal
    those of walks from the root, event for event."""
    n = 20_000
    spec = scenes.heightfield_scene(n)
    res = []
    for o in (dict(distance_grid=False, entry_grid=False), dict()):
        with engine.Scene(spec, lattice_rule=False, **o) as sc:
            traces, counts = sc.trace(8, 0, 64, max_events=20000)
            assert counts.max() < 20000
            sc.run(9, 0, n)
            res.append((traces, sc.collect(), sc.stats()))
    (ta, a, sa), (tb, b, sb) = res
    assert sb["entryWalks"] > 0 and sb["culledWalks"] > 0
    for p in range(64):
        assert len(ta[p]) == len(tb[p]), p
        for k in ta[p].dtype.names:
            assert np.array_equal(ta[p][k], tb[p][k]), (p, k)
    for k in ("scatterEvents", "boundaryEvents", "rouletteEvents"):
        assert sa[k] == sb[k], k
    assert np.array_equal(a.det_count(), b.det_count())
    assert np.array_equal(a.bin_count(), b.bin_count())
    np.testing.assert_allclose(a.det_weight(), b.det_weight(), rtol=1e-11, atol=1e-12)


def test_lattice_rule_changes_nothing_on_the_full_heightfield(engine):
    """The 98,306-triangle heightfield under its lattice rule (plain kernel)
    and on the BVH walk (interleaved kernel with the walk-skipping grids):
    traces of 48 photons equal event for event, and the tallies and event
    counts of 20,000 photons bin for bin."""
    n = 20_000
    spec = scenes.heightfield_scene(n)
    res = []
    for rule in (True, False):
        with engine.Scene(spec, absorption=False, lattice_rule=rule) as sc:
            traces, counts = sc.trace(43, 0, 48, max_events=20000)
            assert counts.max() < 20000
            sc.collect()
            sc.stats()
            sc.run(44, 0, n)
            res.append((traces, sc.collect(), sc.stats()))
    (ta, a, sa), (tb, b, sb) = res
    assert sa["culledWalks"] == 0 and sb["culledWalks"] > 0  # the rule builds no grid
    for p in range(48):
        assert len(ta[p]) == len(tb[p]), p
        for k in ta[p].dtype.names:
            assert np.array_equal(ta[p][k], tb[p][k]), (p, k)
    for k in ("scatterEvents", "boundaryEvents", "rouletteEvents"):
        assert sa[k] == sb[k], k
    assert np.array_equal(a.det_count(), b.det_count())
    assert np.array_equal(a.bin_count(), b.bin_count())
    np.testing.assert_allclose(a.det_weight(), b.det_weight(), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(a.bin_weight(), b.bin_weight(), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("kind", ["heightfield", "head"])
def test_separating_slabs_change_nothing(engine, kind):
    """The separating slabs only skip walks that cannot hit: with and without
    them (distance bound and entry trees on in both), the tallies and event
    counts of 20,000 photons are equal, on the 98,306-triangle heightfield
    (and the traces of 48 photons, event for event) and on a five-shell head
    (level 5, 102,400 triangles); the slabs do skip walks."""
    n = 20_000
    spec = scenes.heightfield_scene(n) if kind == "heightfield" else scenes.head_scene(n, level=5)
    print(f"slabs on/off: {kind}, {spec.num_triangles} triangles", flush=True)
    res = []
    for slabs in (False, True):
        with engine.Scene(spec, absorption=False, slab_grid=slabs, lattice_rule=False) as sc:
            traces = None
            if kind == "heightfield":
                traces, counts = sc.trace(41, 0, 48, max_events=20000)
                assert counts.max() < 20000
                sc.collect()
                sc.stats()
            sc.run(42, 0, n)
            res.append((traces, sc.collect(), sc.stats()))
    (ta, a, sa), (tb, b, sb) = res
    assert sb["culledWalks"] > sa["culledWalks"]
    if ta is not None:
        for p in range(48):
            assert len(ta[p]) == len(tb[p]), p
            for k in ta[p].dtype.names:
                assert np.array_equal(ta[p][k], tb[p][k]), (p, k)
    for k in ("scatterEvents", "boundaryEvents", "rouletteEvents", "entryWalks"):
        if k == "entryWalks":
            assert sb[k] <= sa[k]
        else:
            assert sa[k] == sb[k], k
    assert np.array_equal(a.det_count(), b.det_count())
    assert np.array_equal(a.bin_count(), b.bin_count())
    np.testing.assert_allclose(a.det_weight(), b.det_weight(), rtol=1e-11, atol=1e-12)


def test_heightfield_statistics_at_1e8_photons(engine):
    """The 3-D surface configuration at its full 1e8 photons on the GPU (seed
    1, twenty launches of 5e6) against the double oracle (seed 2, forty
    samples of 40,000 photons: 1.6e6), independent streams.  R and A within
    three combined standard errors and within 0.5 %.  The 20 rings and the
    80-odd populated TOF bins (summed over rings) are tested as families,
    because a 3-sigma test per bin raises false alarms in proportion to the
    bins: P(all 20 rings within 3) = 0.9973^20 = 0.95 for a normal z, less for
    the t-like tail of 20- and 40-sample variances (the first run at this
    power had one ring of 20 at 3.26 while the paired test below bounds that
    ring's float-vs-double bias at -1.3e-4 +- 3.5e-4 relative).  Per family:
    at least 90 % (rings) / 97 % (TOF) within 3, none beyond 3.5 (rings) /
    4.5 (TOF) -- a family-wise level near one 3-sigma test's 0.27 % -- and a
    chi-square-like mean z^2 (< 1.6 / < 1.5)."""
    spec = scenes.heightfield_scene(0)
    inp = InputHandle(spec)
    nb_g, n_gpu, nb_c, n_cpu = 20, 5_000_000, 40, 40_000

    def vec(det_w, bins_w, absorbed, n):
        return np.concatenate([det_w / n, bins_w.sum(axis=0) / n, [det_w.sum() / n, absorbed / n]])

    G = []
    with engine.Scene(spec, absorption=False) as sc:
        for i in range(nb_g):
            sc.run(1, i * n_gpu, n_gpu)
            g = sc.collect()
            st = sc.stats()
            G.append(vec(g.det_weight(), g.bin_weight(), st["wAbsorbed"], n_gpu))
    print(f"heightfield: GPU {nb_g} x {n_gpu} photons done; double oracle {nb_c} x {n_cpu}", flush=True)
    Cc = []
    for i in range(nb_c):
        o = OutputHandle(spec, absorption=False)
        c = ol.run_oracle(spec, inp, o, 2, i * n_cpu, n_cpu, 64, 0)
        Cc.append(vec(o.det_weight(), o.bin_weight(), c.wAbsorbed, n_cpu))
    G, Cc = np.array(G), np.array(Cc)
    mg, mc = G.mean(0), Cc.mean(0)
    se = np.sqrt(G.var(0, ddof=1) / nb_g + Cc.var(0, ddof=1) / nb_c)
    populated = (mc > 1e-5) & ((Cc > 0).sum(axis=0) >= nb_c // 2)
    z = np.abs(mg - mc) / np.where(se > 0, se, 1.0)
    rings, tof = np.arange(20), np.arange(20, 20 + spec.time_bins)
    zr = z[rings][populated[rings]]
    zt = z[tof][populated[tof]]
    print(f"max |z|: rings {zr.max():.2f}, TOF {zt.max():.2f} over {zt.size} bins "
          f"({np.mean(zt < 3):.3f} within 3, mean z^2 {np.mean(zt ** 2):.2f}), R {z[-2]:.2f}, A {z[-1]:.2f}; "
          f"R {mg[-2]:.5f} vs {mc[-2]:.5f}, A {mg[-1]:.5f} vs {mc[-1]:.5f}")
    assert zr.size == 20 and np.mean(zr < 3.0) >= 0.9 and zr.max() < 3.5 and np.mean(zr ** 2) < 1.6
    assert z[-2] < 3.0 and z[-1] < 3.0
    assert zt.size > 40
    assert np.mean(zt < 3.0) >= 0.97 and zt.max() < 4.5 and np.mean(zt ** 2) < 1.5
    assert abs(mg[-2] / mc[-2] - 1) < 5e-3
    assert abs(mg[-1] / mc[-1] - 1) < 5e-3


def test_photon_indices_are_capped_at_2_33(engine):
    """A seed has 2^33 disjoint streams: a run that would reach photon 2^33
    fails with XMC_ERR_LIMIT instead of repeating photon 0's numbers."""
    spec = scenes.slab_scene(0)
    with engine.Scene(spec, absorption=False) as sc:
        sc.run(1, (1 << 33) - 100, 100)  # the last photons of the seed run
        sc.collect()
        with pytest.raises(engine.EngineError, match="2\\^33"):
            sc.run(1, (1 << 33) - 100, 101)


def test_heightfield_paired_with_the_double_walk(engine):
    """Float-transport bias, with the power that independent streams lack:
    the GPU's photons 0..4e5 of seed 7 on the 98k-triangle heightfield against
    the same photons walked in double arithmetic on the same 24-bit uniforms
    (the 64-bit oracle in its recorded-sequence mode, whose arithmetic is the
    reference's).  Most photons take identical decisions, so per-batch
    differences are small and their mean has a standard error far below that
    of an independent comparison.  Every ring, R and A: |mean difference| <
    3 paired standard errors, and < 2e-3 relative (measured at most 1.2e-3 on
    the outer rings, 1e-5 on R and A)."""
    spec = scenes.heightfield_scene(0)
    inp = InputHandle(spec)
    nb, n = 16, 25_000
    G, D = [], []
    with engine.Scene(spec, absorption=False) as sc:
        for b in range(nb):
            sc.run(7, b * n, n)
            g = sc.collect()
            st = sc.stats()
            G.append(np.concatenate([g.det_weight() / n, [g.det_weight().sum() / n, st["wAbsorbed"] / n]]))
    print(f"paired heightfield: GPU {nb} x {n} photons of seed 7 done; double walk of the same photons", flush=True)
    for b in range(nb):
        o = OutputHandle(spec, absorption=False)
        c = ol.run_oracle(spec, inp, o, 7, b * n, n, 65, 0)
        D.append(np.concatenate([o.det_weight() / n, [o.det_weight().sum() / n, c.wAbsorbed / n]]))
    d = np.array(G) - np.array(D)
    se = d.std(0, ddof=1) / math.sqrt(nb)
    z = np.abs(d.mean(0)) / np.where(se > 0, se, 1.0)
    rel = np.abs(d.mean(0)) / np.array(D).mean(0)
    print(f"paired: max |z| {z.max():.2f}, max relative difference {rel.max():.2e} (R {rel[-2]:.2e}, "
          f"A {rel[-1]:.2e})")
    assert np.all(z < 3.0), z
    assert np.all(rel < 2e-3), rel
