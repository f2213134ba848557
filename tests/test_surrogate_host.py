"""CPU checks of the surrogate's host side and of its CPU reference network.

The reference network (tests/mlp_ref.py) is what the GPU kernels are held
to, so it is pinned here first: its backward pass in float64 against central
finite differences of its own loss, its bf16 rounding against hand-computed
bit patterns, and its loss against the closed form at the optimum."""
import math

import numpy as np
import pytest

import mlp_ref as R
from xmcml_amd import surrogate as S


def test_bf16_rounding_bits():
    one = np.float32(1.0)
    ulp = 2.0 ** -7  # bf16 spacing at 1
    cases = [
        (1.0, 1.0),
        (1.0 + ulp / 2, 1.0),             # tie -> even (1.0)
        (1.0 + 3 * ulp / 2, 1.0 + 2 * ulp),  # tie -> even (1 + 2 ulp)
        (1.0 + ulp / 2 + 2 ** -20, 1.0 + ulp),
        (-2.5, -2.5),
        (3.0e-3, None),
    ]
    for x, want in cases:
        got = float(R.bf16(np.float32(x)).item())
        if want is not None:
            assert got == want, (x, got, want)
        else:
            assert abs(got - x) <= abs(x) * 2 ** -8
    x = np.random.default_rng(0).standard_normal(10000).astype(np.float32)
    r = R.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 * (1 + 1e-7))
    assert np.all(r.view(np.uint32) & 0xFFFF == 0)
    assert one == R.bf16(one)


def _tiny(exact=True, seed=1):
    rng = np.random.default_rng(seed)
    inputs, hidden, depth, outputs = 5, 7, 2, 3
    shapes = R.layer_shapes(inputs, hidden, depth, outputs)
    params = np.concatenate([np.concatenate([rng.standard_normal(r * c) / math.sqrt(c), 0.1 * rng.standard_normal(r)])
                             for r, c in shapes])
    x = rng.standard_normal((6, inputs))
    y = rng.standard_normal((6, outputs))
    mask = (rng.random((6, outputs)) > 0.3).astype(np.float64)
    return params, (inputs, hidden, depth, outputs), x, y, mask


def test_reference_backward_matches_finite_differences():
    params, dims, x, y, mask = _tiny()
    scale = 1.0 / mask.sum()
    net = R.RefMLP(params, *dims, exact=True)
    loss, _, grads = net.loss_and_grads(x, y, mask, scale)
    g = R.pack([(gW, gb) for gW, gb in grads])
    h = 1e-6
    fd = np.zeros_like(params)
    for i in range(params.size):
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        l1 = R.RefMLP(p1, *dims, exact=True).loss_and_grads(x, y, mask, scale)[0]
        l2 = R.RefMLP(p2, *dims, exact=True).loss_and_grads(x, y, mask, scale)[0]
        fd[i] = (l1 - l2) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-6 * max(1.0, np.max(np.abs(g))), np.max(np.abs(fd - g))
    # the gradients are far from zero, so the comparison above is not vacuous
    assert np.max(np.abs(fd)) > 1e-3


def test_reference_loss_closed_form():
    # one output, mu = y exactly, s = log-variance: L = 0.5 * s per unmasked entry
    params, dims, x, y, mask = _tiny()
    net = R.RefMLP(params, *dims, exact=True)
    out, _ = net.forward(x)
    O = dims[3]
    y_eq = out[:, :O]
    loss, _, _ = net.loss_and_grads(x, y_eq, mask, 1.0)
    s = np.clip(out[:, O:], -10, 10)
    assert loss == pytest.approx(float(np.sum(mask * 0.5 * s)), rel=1e-12, abs=1e-12)


def test_reference_adamw_first_step():
    # after one step from zero moments, |update| = lr (sign of the gradient) for |g| >> eps
    g = np.array([0.5, -2.0, 1e-3], np.float32)
    p = np.array([1.0, 1.0, 1.0], np.float32)
    z = np.zeros(3, np.float32)
    p2, m1, m2 = R.adamw(p, g, z, z, 1, 1e-2, 0.9, 0.999, 1e-8, 0.0, np.ones(3, bool))
    assert np.allclose(p - p2, 1e-2 * np.sign(g), rtol=1e-4)
    p3, _, _ = R.adamw(p, g, z, z, 1, 1e-2, 0.9, 0.999, 1e-8, 0.1, np.ones(3, bool))
    assert np.allclose(p2 - p3, 1e-2 * 0.1 * p, rtol=2e-4)


def test_encode_maps_the_box_to_unit_interval():
    lo = [(1.33, 1e-3, 0.5 / (1 - 0.7), 0.7, 0.05)] * 2
    hi = [(1.5, 1.0, 5.0 / (1 - 0.95), 0.95, 10.0)] * 5
    x = S.encode(lo)
    assert np.allclose(x[:10], -1.0, atol=1e-6) and np.all(x[10:25] == 0)
    assert list(x[25:]) == [1, 1, 0, 0, 0]
    x = S.encode(hi)
    assert np.allclose(x[:25], 1.0, atol=1e-6) and np.all(x[25:] == 1)
    assert S.in_range(lo) and S.in_range(hi)
    assert not S.in_range(lo[:1])
    bad = [(1.4, 2.0, 10.0, 0.9, 1.0)] * 2  # mua beyond the box
    assert not S.in_range(bad)
    with pytest.raises(ValueError):
        S.encode([])


def test_draw_configs_in_range_and_seeded():
    a = S.draw_configs(200, 5)
    b = S.draw_configs(200, 5)
    assert a == b
    assert all(S.in_range(c) for c in a)
    assert {len(c) for c in a} == {2, 3, 4, 5}


def test_targets_from_output():
    from xmcml_amd.abi import OutputHandle
    spec = S.config_scene(S.draw_configs(1, 0)[0], 1000)
    out = OutputHandle(spec, absorption=False)
    for d in range(S.RINGS):
        out.dets[d].weight = 0.01 * (d + 1)
        out.dets[d].numberOfPhotons = 100 if d < 20 else 3
    out.bins[5].weight = 0.5
    out.bins[5].numberOfPhotons = 40
    out.bins[S.TOF_BINS + 5].weight = 0.25
    out.bins[S.TOF_BINS + 5].numberOfPhotons = 10
    y, m, cnt = S.targets(out, 1000)
    assert y.shape == (96,) and m.shape == (96,)
    assert y[0] == pytest.approx(math.log(0.01 / 1000), rel=1e-6)
    assert m[:20].all() and not m[20:32].any()
    assert m[32 + 5] == 1 and y[32 + 5] == pytest.approx(math.log(0.75 / 1000), rel=1e-6)
    assert cnt[32 + 5] == 50
    assert m[32:].sum() == 1


def test_standardisation_masked():
    Y = np.array([[1.0, 5.0], [3.0, 100.0]], np.float32)
    M = np.array([[1.0, 1.0], [1.0, 0.0]], np.float32)
    mean, std = S.Surrogate.standardisation(Y, M)
    assert mean[0] == pytest.approx(2.0) and mean[1] == pytest.approx(5.0)
    assert std[0] == pytest.approx(1.0)


def test_vectorised_encoding_and_sampling():
    """encode_arrays / in_range_arrays on (n, 5, 5) layer arrays give what the
    per-configuration calls give; draw_config_arrays samples the distribution
    of scenes.draw_layers (every configuration inside the box, marginals of
    20,000 draws within 4 standard errors of draw_configs' at the deciles)."""
    import math as m
    from xmcml_amd import surrogate as S
    L, cnt = S.draw_config_arrays(20000, 5)
    assert np.all(S.in_range_arrays(L, cnt))
    cfg = S.arrays_to_configs(L[:500], cnt[:500])
    X = S.encode_arrays(L[:500], cnt[:500])
    assert np.array_equal(X, np.stack([S.encode(c) for c in cfg]))
    assert [S.in_range(c) for c in cfg] == list(S.in_range_arrays(L[:500], cnt[:500]))
    bad = [list(c) for c in cfg[:3]]
    bad[0][0] = (1.4, 2.0, 10.0, 0.9, 1.0)  # mua 2 /mm is outside the box
    bad[1] = bad[1][:1]                     # one layer
    Lb, cb = S.configs_to_arrays(bad)
    assert list(S.in_range_arrays(Lb, cb)) == [False, False, True]
    ref = S.configs_to_arrays(S.draw_configs(20000, 6))
    qs = np.linspace(0.1, 0.9, 9)
    for k, name in enumerate(["n", "mua", "mus", "g", "d"]):
        a = L[:, :, k][np.arange(5)[None, :] < cnt[:, None]]
        b = ref[0][:, :, k][np.arange(5)[None, :] < ref[1][:, None]]
        qa, qb = np.quantile(np.log(a), qs), np.quantile(np.log(b), qs)
        # quantile standard error sqrt(p(1-p)/n) / density, density >= 0.35 / sd for these shapes;
        # two independent samples
        se = np.std(np.log(b)) * np.sqrt(qs * (1 - qs)) / m.sqrt(len(b)) / 0.35 * m.sqrt(2.0)
        assert np.all(np.abs(qa - qb) < 4 * se + 1e-12), name
    assert abs(cnt.mean() - ref[1].mean()) < 4 * ref[1].std() / m.sqrt(len(cnt))
