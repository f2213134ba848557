"""MCML Russian roulette (stable/plane_parallel/1.0/McmlGpu/kernel/kernel.cu:
214-222, applied as kernel.cu:449-452) pinned by what the mathematics fixes,
not by a second restatement of it:

* the step is unbiased: over 1e6 photons below the threshold the mean weight
  afterwards equals the weight before (within 4 standard errors), and the
  survival fraction equals the chance;
* it touches only live photons below the threshold: others draw nothing and
  keep their weight; a zero weight dies without a draw (kernel.cu:216-217);
* at the walk level, roulette at 1e-4 / 0.1 and a walk that only stops at
  1e-12 give the same ring, reflectance and absorption tallies within three
  combined standard errors (independent streams).

A flipped comparison (u > chance), a missing w /= chance or a draw taken for
a dead photon each fails one of these."""
import math

import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd import scenes
from xmcml_amd.abi import InputHandle, OutputHandle, TERMINATE_KILL, TERMINATE_ROULETTE

SKIP, ZERO, LIVE, DIE = 0, 1, 2, 3


def roulette(precision, weights, dead, min_weight, chance, seed):
    n = len(weights)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    d = np.ascontiguousarray(dead, dtype=np.uint8)
    out = np.zeros(n)
    outcome = np.zeros(n, np.int32)
    draws = np.zeros(n, np.int32)
    rc = ol.oracle().xmco_roulette_many(precision, w.ctypes.data, d.ctypes.data, min_weight, chance, seed, n,
                                        out.ctypes.data, outcome.ctypes.data, draws.ctypes.data)
    assert rc == 0
    return out, outcome, draws


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("chance", [0.1, 0.5])
def test_roulette_preserves_the_expected_weight(precision, chance):
    n, w0 = 1_000_000, 5e-5
    out, outcome, draws = roulette(precision, np.full(n, w0), np.zeros(n), 1e-4, chance, 321)
    assert np.all(draws == 1)
    assert set(np.unique(outcome)) <= {LIVE, DIE}
    live = outcome == LIVE
    # survival is Bernoulli(chance); a survivor carries w0 / chance
    assert abs(live.mean() - chance) < 4 * math.sqrt(chance * (1 - chance) / n)
    w_live = np.float32(w0) / np.float32(chance) if precision == 32 else w0 / chance
    assert np.all(out[live] == w_live) and np.all(out[~live] == 0.0)
    se = w0 * math.sqrt((1 - chance) / chance / n)
    assert abs(out.mean() - w0) < 4 * se, (out.mean(), w0, se)


@pytest.mark.parametrize("precision", [64, 32])
def test_roulette_only_touches_live_photons_below_the_threshold(precision):
    w = np.array([2e-4, 1e-4, 9.99e-5, 0.0, 5e-5, 5e-5, 1.0])
    dead = np.array([0, 0, 0, 0, 1, 0, 1])
    out, outcome, draws = roulette(precision, w, dead, 1e-4, 0.1, 7)
    # at or above the threshold, or already dead: no draw, weight kept
    for i in (0, 1, 4, 6):
        assert outcome[i] == SKIP and draws[i] == 0
    assert out[0] == pytest.approx(2e-4) and out[1] == pytest.approx(1e-4)
    # a zero weight dies without drawing (kernel.cu:216-217)
    assert outcome[3] == ZERO and draws[3] == 0 and out[3] == 0.0
    # below the threshold and alive: exactly one draw
    assert outcome[2] in (LIVE, DIE) and draws[2] == 1
    assert outcome[5] in (LIVE, DIE) and draws[5] == 1


def test_roulette_survival_follows_the_drawn_uniform():
    """Photon i draws the first uniform of its stream: it survives exactly when
    that uniform is below the chance (kernel.cu:218)."""
    n = 4096
    seed = 99
    _, outcome, _ = roulette(64, np.full(n, 1e-5), np.zeros(n), 1e-4, 0.3, seed)
    u = np.empty(n)
    for i in range(n):
        import ctypes as C
        v = C.c_uint64(ol.oracle().xmco_photon_stream(seed, i))
        u[i] = ol.oracle().xmco_mcg59_next(C.byref(v), 302875106592253)
    assert np.array_equal(outcome == LIVE, u < 0.3)


@pytest.mark.parametrize("precision", [64, 32])
def test_roulette_walk_matches_a_walk_without_termination(precision):
    """Slab configuration: roulette at 1e-4 / 0.1 against kill at 1e-12, which
    leaves < 1e-8 of the weight untracked.  Independent streams, 8 batches of
    5,000 photons each side; every populated ring, total reflectance and
    absorption within 3 combined standard errors."""
    nb, n = 8, 5000
    base = scenes.slab_scene(n)

    def batches(term, min_w, seed):
        spec = scenes.slab_scene(n)
        spec.termination = term
        spec.min_weight = min_w
        inp = InputHandle(spec)
        rows = []
        for b in range(nb):
            o = OutputHandle(spec, absorption=False)
            c = ol.run_oracle(spec, inp, o, seed, b * n, n, precision, 8)
            rows.append(np.concatenate([o.det_weight() / n, [o.det_weight().sum() / n, c.wAbsorbed / n]]))
        return np.array(rows)

    A = batches(TERMINATE_ROULETTE, 1e-4, 11)
    B = batches(TERMINATE_KILL, 1e-12, 12)
    ma, mb = A.mean(0), B.mean(0)
    se = np.sqrt(A.var(0, ddof=1) / nb + B.var(0, ddof=1) / nb)
    pop = mb > 1e-3
    z = np.abs(ma - mb) / np.where(se > 0, se, 1.0)
    assert pop.sum() >= 10
    assert np.all(z[pop] < 3.0), z[pop]
    assert base.rings  # the slab has its 20 rings
