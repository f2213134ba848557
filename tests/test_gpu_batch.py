"""Batched configurations in one persistent launch (graph-replayed)."""
import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd import autotune, scenes
from xmcml_amd.abi import InputHandle, OutputHandle

pytestmark = pytest.mark.gpu


def small_batch(count=24, scale=2e-3, seed=5):
    return scenes.batched_configs(count, seed=seed, photon_scale=scale)


def test_batch_equals_separate_scenes(engine):
    specs = small_batch()
    prefix = np.concatenate([[0], np.cumsum([s.photons for s in specs])])
    with engine.Batch(specs) as b:
        b.run(41)
        outs = b.collect()
        st = b.stats()
    assert st["photons"] == prefix[-1]
    for c, s in enumerate(specs):
        with engine.Scene(s, absorption=False) as sc:
            sc.run(41, int(prefix[c]), s.photons)
            one = sc.collect()
        assert np.array_equal(outs[c].det_count(), one.det_count()), c
        assert np.array_equal(outs[c].bin_count(), one.bin_count()), c
        np.testing.assert_allclose(outs[c].det_weight(), one.det_weight(), rtol=1e-11, atol=1e-14)


def test_batch_matches_float_oracle(engine):
    specs = small_batch(8, 1e-3, seed=9)
    prefix = np.concatenate([[0], np.cumsum([s.photons for s in specs])])
    with engine.Batch(specs) as b:
        b.run(3)
        outs = b.collect()
    for c, s in enumerate(specs):
        o = OutputHandle(s, absorption=False)
        ol.run_oracle(s, InputHandle(s), o, 3, int(prefix[c]), s.photons, 32, 8)
        assert np.array_equal(outs[c].det_count(), o.det_count()), c
        np.testing.assert_allclose(outs[c].det_weight(), o.det_weight(), rtol=1e-11, atol=1e-14)


def test_graph_replays_match_eager(engine):
    specs = small_batch(16, 1e-3, seed=12)
    with engine.Batch(specs, use_graph=False) as eager:
        eager.run(8)
        ref = eager.collect()
    with engine.Batch(specs, use_graph=True) as g:
        for rep in range(3):
            g.run(8)
            outs = g.collect()
            # eager work and a synchronisation between replays
            with engine.Scene(specs[0], absorption=False) as sc:
                sc.run(1, 0, 1000)
                sc.collect()
            for c in range(len(specs)):
                assert np.array_equal(outs[c].det_count(), ref[c].det_count()), (rep, c)
                assert np.array_equal(outs[c].bin_count(), ref[c].bin_count()), (rep, c)
                np.testing.assert_allclose(outs[c].det_weight(), ref[c].det_weight(), rtol=1e-11, atol=1e-14)


def test_autotuner_picks_a_measured_shape(engine, tmp_path, monkeypatch):
    monkeypatch.setenv("XMC_AUTOTUNE_CACHE", str(tmp_path / "cache.json"))
    specs = small_batch(16, 1e-3, seed=2)
    with engine.Batch(specs) as b:
        res = autotune.tune_batch(b)
        assert not res["cached"] and len(res["timings"]) == len(autotune.CANDIDATES)
        assert res["ms"] == min(t["ms"] for t in res["timings"])
        block, blocks = b.get_launch()
        assert block == res["block"]
        res2 = autotune.tune_batch(b)
        assert res2["cached"] and res2["block"] == res["block"]


def test_tuned_batch_gives_the_same_answers(engine, tmp_path, monkeypatch):
    """Batch(tune=True) times the candidate shapes on a copy at 1 % of the
    budgets, keyed on the HIP device, and applies the winner; photon p of
    configuration c keeps its stream, so the tallies equal the default
    shape's."""
    monkeypatch.setenv("XMC_AUTOTUNE_CACHE", str(tmp_path / "cache.json"))
    specs = small_batch(24, 1e-2, seed=4)
    with engine.Batch(specs) as b:
        b.run(3)
        ref = b.collect()
    with engine.Batch(specs, tune=True) as b:
        t = b.tuning
        assert not t["cached"] and len(t["timings"]) == len(autotune.CANDIDATES)
        assert b.get_launch()[0] == t["block"]
        b.run(3)
        outs = b.collect()
    key = next(iter(__import__("json").loads((tmp_path / "cache.json").read_text())))
    assert engine.device_name() in key
    for c in range(len(specs)):
        assert np.array_equal(outs[c].det_count(), ref[c].det_count()), c
        np.testing.assert_allclose(outs[c].det_weight(), ref[c].det_weight(), rtol=1e-11, atol=1e-14)
