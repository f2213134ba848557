"""The surrogate's matrix-core kernels against the CPU reference network
(tests/mlp_ref.py, itself pinned by finite differences), and the trained
surrogate against plain Monte Carlo.

Tolerances: the kernels accumulate bf16 products in fp32 where the reference
accumulates in float64, so a value rounded to bf16 afterwards can land one
bf16 step (2^-8 relative) away on rare elements.  Element-wise checks are
therefore made on fp32 results before any rounding (GEMM, head output), and
gradients are compared by relative L2 norm per tensor (<= 5e-3) plus their
cosine (>= 0.9999)."""
import math
import time

import numpy as np
import pytest

import mlp_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev(engine):
    """engine.DeviceArray factory (the engine fixture skips without a GPU)."""
    return engine.DeviceArray


@pytest.mark.parametrize("tile", [0, 128])
def test_gemm_matches_reference(dev, tile):
    from xmcml_amd.surrogate import gemm_check
    rng = np.random.default_rng(0)
    shapes = [(64, 64, 64), (128, 192, 128), (256, 64, 512)] if tile == 0 else \
        [(128, 128, 64), (256, 384, 128), (128, 256, 512)]
    for (M, N, K) in shapes:
        a = rng.standard_normal((M, K)).astype(np.float32)
        b = rng.standard_normal((N, K)).astype(np.float32)
        c = gemm_check(a, b, tile)
        ref = R.bf16(a).astype(np.float64) @ R.bf16(b).astype(np.float64).T
        bound = (np.abs(R.bf16(a)).astype(np.float64) @ np.abs(R.bf16(b)).astype(np.float64).T) * K * 2.0 ** -24
        assert np.all(np.abs(c - ref) <= bound + 1e-30), (M, N, K, np.max(np.abs(c - ref) / bound))
        # exact on integer-valued operands (every partial sum is representable)
        ai = rng.integers(-8, 9, (M, K)).astype(np.float32)
        bi = rng.integers(-8, 9, (N, K)).astype(np.float32)
        assert np.array_equal(gemm_check(ai, bi, tile),
                              (ai.astype(np.float64) @ bi.T.astype(np.float64)).astype(np.float32))


def _model(hidden=128, depth=2, outputs=32, max_batch=256, seed=3):
    from xmcml_amd.surrogate import MLP
    m = MLP(30, hidden, depth, outputs, max_batch, seed=seed)
    rng = np.random.default_rng(seed)
    p = m.get_params()
    # nonzero biases so their gradients and the bias epilogue are exercised
    shapes = R.layer_shapes(30, hidden, depth, outputs)
    layers = R.unpack(p.copy(), shapes)
    layers = [(W, 0.1 * rng.standard_normal(W.shape[0]).astype(np.float32)) for W, _ in layers]
    p = R.pack(layers).astype(np.float32)
    m.set_params(p)
    return m, p, shapes


def _batch(dev, n, outputs, seed=4):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 30)).astype(np.float32)
    y = rng.standard_normal((n, outputs)).astype(np.float32)
    mask = (rng.random((n, outputs)) > 0.2).astype(np.float32)
    return x, y, mask, [dev(a.shape, a) for a in (x, y, mask)]


def test_forward_matches_reference(dev):
    m, p, shapes = _model()
    x, y, mask, (xt, _, _) = _batch(dev, 200, 32)  # 200: not a multiple of the 64-row tile
    out = m.forward(xt).numpy()
    ref, _ = R.RefMLP(p, 30, 128, 2, 32).forward(x)
    err = np.abs(out - ref)
    assert np.linalg.norm(out - ref) <= 5e-3 * np.linalg.norm(ref), np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert err.max() <= 0.05 * np.abs(ref).max()
    # more rows than maxBatch: processed in chunks, same answer per row
    big = np.concatenate([x, x, x[:100]])
    out2 = m.forward(dev(big.shape, big)).numpy()
    assert np.array_equal(out2[:200], out) and np.array_equal(out2[200:400], out)
    assert np.array_equal(out2[400:], out[:100])


@pytest.mark.parametrize("hidden,depth,outputs", [(128, 2, 32), (256, 3, 64), (512, 4, 96), (512, 1, 128)])
def test_fused_forward_equals_per_layer_launches(dev, hidden, depth, outputs):
    """The whole-network kernel (128-row tiles held in LDS) against one GEMM
    launch per layer: same k order and epilogue arithmetic, so the same bits,
    on batches that leave a ragged last tile, span several tiles, or exceed
    maxBatch (the per-layer path's chunking)."""
    from xmcml_amd.surrogate import MLP
    rng = np.random.default_rng(hidden + depth)
    m = MLP(30, hidden, depth, outputs, 384, seed=hidden)
    p = m.get_params()
    layers = R.unpack(p.copy(), R.layer_shapes(30, hidden, depth, outputs))
    layers = [(W, 0.1 * rng.standard_normal(W.shape[0]).astype(np.float32)) for W, _ in layers]
    p = R.pack(layers).astype(np.float32)
    m.set_params(p)
    for n in (1, 129, 1000):
        x = rng.uniform(-1, 1, (n, 30)).astype(np.float32)
        xt = dev(x.shape, x)
        fused = m.forward(xt).numpy()
        layered = m.forward(xt, per_layer=True).numpy()
        print(hidden, depth, outputs, n, "max |fused - per-layer|", float(np.max(np.abs(fused - layered))))
        assert np.array_equal(fused, layered), (n, float(np.max(np.abs(fused - layered))))
        if n == 129:
            ref, _ = R.RefMLP(p, 30, hidden, depth, outputs).forward(x)
            assert np.linalg.norm(fused - ref) <= 5e-3 * np.linalg.norm(ref)


def test_large_batch_forward_and_backward(dev):
    """Batches large enough for the 128 x 128 tile kernel (bias + GELU forward,
    GELU' backward with a transposed weight operand) against the reference."""
    from xmcml_amd.surrogate import MLP
    n, hidden = 16384, 256
    m = MLP(30, hidden, 2, 32, n, seed=5)
    p = m.get_params()
    x, y, mask, (xt, yt, mt) = _batch(dev, n, 32, seed=6)
    out = m.forward(xt).numpy()
    ref_net = R.RefMLP(p, 30, hidden, 2, 32)
    ref, _ = ref_net.forward(x)
    assert np.linalg.norm(out - ref) <= 5e-3 * np.linalg.norm(ref)
    scale = 1.0 / mask.sum()
    loss = m.backward(xt, yt, mt, scale)
    rloss, _, rgrads = ref_net.loss_and_grads(x, y, mask, scale)
    assert loss == pytest.approx(rloss, rel=2e-3)
    for l, ((gW, gb), (rW, rb)) in enumerate(zip(R.unpack(m.get_grads(), R.layer_shapes(30, hidden, 2, 32)),
                                                 rgrads)):
        for name, a, b in (("W", gW, rW), ("b", gb, rb)):
            rel, cos = _cmp(a, b)
            assert rel <= 5e-3 and cos >= 0.9999, (l, name, rel, cos)


def _cmp(g, r):
    g = np.asarray(g, np.float64).ravel()
    r = np.asarray(r, np.float64).ravel()
    rel = np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-30)
    cos = float(g @ r) / max(np.linalg.norm(g) * np.linalg.norm(r), 1e-30)
    return rel, cos


def test_backward_matches_reference(dev):
    m, p, shapes = _model()
    n = 256
    x, y, mask, (xt, yt, mt) = _batch(dev, n, 32)
    scale = 1.0 / mask.sum()
    loss = m.backward(xt, yt, mt, scale)
    grads = m.get_grads()
    rloss, _, rgrads = R.RefMLP(p, 30, 128, 2, 32).loss_and_grads(x, y, mask, scale)
    assert loss == pytest.approx(rloss, rel=2e-3)
    g_layers = R.unpack(grads, shapes)
    for l, ((gW, gb), (rW, rb)) in enumerate(zip(g_layers, rgrads)):
        for name, a, b in (("W", gW, rW), ("b", gb, rb)):
            rel, cos = _cmp(a, b)
            assert rel <= 5e-3 and cos >= 0.9999, (l, name, rel, cos)
    # the padded input columns carry no gradient into the caller's view
    assert np.isfinite(grads).all()


def test_adam_matches_reference(dev):
    m, p, shapes = _model()
    x, y, mask, (xt, yt, mt) = _batch(dev, 128, 32)
    flags = R.weight_flags(shapes)
    m1 = np.zeros_like(p)
    m2 = np.zeros_like(p)
    cur = p.copy()
    for step in (1, 2, 3):
        m.backward(xt, yt, mt, 1.0 / mask.sum(), want_loss=False)
        g = m.get_grads()
        m.adam(1e-3, 0.9, 0.999, 1e-8, 0.01)
        cur, m1, m2 = R.adamw(cur, g, m1, m2, step, 1e-3, 0.9, 0.999, 1e-8, 0.01, flags)
        got = m.get_params()
        assert np.allclose(got, cur, rtol=1e-5, atol=1e-7), (step, np.max(np.abs(got - cur)))


def test_training_fits_a_smooth_function(dev):
    from xmcml_amd.surrogate import Surrogate
    rng = np.random.default_rng(7)
    n = 8192
    X = rng.uniform(-1, 1, (n, 30)).astype(np.float32)
    A = rng.standard_normal((30, 96)).astype(np.float32) / math.sqrt(30)
    Y = (np.sin(X @ A) + 0.01 * rng.standard_normal((n, 96))).astype(np.float32)
    M = np.ones_like(Y)
    sur, log = Surrogate.fit(X[:7168], Y[:7168], M[:7168], steps=600, batch=1024, lr=2e-3, hidden=256, depth=3)
    mu, sig = sur.predict(X[7168:])
    rmse = float(np.sqrt(np.mean((mu - Y[7168:]) ** 2)))
    base = float(np.sqrt(np.mean((Y[7168:] - Y[:7168].mean(0)) ** 2)))
    print(f"fit {log.seconds:.2f} s, loss {log.first_loss:.3f} -> {log.last_loss:.3f}, rmse {rmse:.4f} vs {base:.4f}")
    assert log.last_loss < log.first_loss - 1.0
    assert rmse < 0.15 * base


def test_surrogate_beats_plain_mc_at_equal_cost(dev):
    """Trained on 4096 batched-MC configurations, the surrogate answers
    held-out configurations with a smaller ring-reflectance error than plain
    Monte Carlo given 1000x the surrogate's time per configuration."""
    from xmcml_amd import engine
    from xmcml_amd import surrogate as S
    train_cfg = S.draw_configs(4096, 11)
    test_cfg = S.draw_configs(64, 12)
    t0 = time.perf_counter()
    Y, M, _, _ = S.simulate(train_cfg, 20000, seed=100)
    t_train_mc = time.perf_counter() - t0
    _, Mref, Kref, Wref = S.simulate(test_cfg, 400000, seed=200)
    sur, log = S.Surrogate.fit(S.encode_many(train_cfg), Y, M, steps=1500, batch=1024, lr=2e-3)
    Xte = S.encode_many(test_cfg)
    # throughput on a query batch of realistic size (the test configurations, repeated)
    Xq = np.tile(Xte, (512, 1))
    xt = dev(Xq.shape, Xq)
    out = sur.predict_device(xt)
    engine.device_sync()
    t0 = time.perf_counter()
    for _ in range(5):
        sur.predict_device(xt, out)
    engine.device_sync()
    t_per_cfg = (time.perf_counter() - t0) / (5 * len(Xq))
    mu, _ = sur.predict(Xte)
    valid = (Kref[:, :S.RINGS] >= 400)
    err_s = np.median(S.ring_relative_error(np.exp(mu[:, :S.RINGS]), Wref[:, :S.RINGS], valid))
    # plain MC at 1000x the surrogate's time per configuration
    rate = len(train_cfg) * 20000 / t_train_mc  # photons/s of the batched MC, launch overheads included
    photons = max(1, int(1000 * t_per_cfg * rate))
    _, _, _, Wmc = S.simulate(test_cfg, photons, seed=300)
    err_mc = np.median(S.ring_relative_error(Wmc[:, :S.RINGS], Wref[:, :S.RINGS], valid))
    print(f"surrogate median ring error {err_s:.4f} ({t_per_cfg * 1e6:.3f} us/config); plain MC with "
          f"{photons} photons/config (1000x the time): {err_mc:.4f}; fit {log.seconds:.1f} s, MC data "
          f"{t_train_mc:.1f} s")
    assert err_s < 0.25
    assert err_s < err_mc


def test_answer_uses_the_network_or_monte_carlo(dev):
    """Surrogate.answer on in-range, out-of-range and uncertainty-flagged
    configurations: the flagged ones (outside the box, or above sigma_max)
    equal plain batched MC of the same configurations at the same seed, bit
    for bit; the others equal the network's prediction."""
    from pathlib import Path
    from xmcml_amd import surrogate as S
    model = Path(__file__).resolve().parents[1] / "xmcml_amd" / "data" / "surrogate_layered.npz"
    sur = S.Surrogate.load(model, max_batch=4096)
    cfg = S.draw_configs(40, 17)
    cfg[3] = [(1.4, 2.0, 10.0, 0.9, 1.0)] + list(cfg[3][1:])  # mua outside the box
    cfg[7] = cfg[7][:1]                                        # a single layer
    mu, sig = sur.predict(S.encode_many([c for i, c in enumerate(cfg) if i not in (3, 7)]))
    unc = sur.uncertainty(sig)
    sur.sigma_max = float(np.median(unc))  # flag about half of the in-range ones
    print(f"answer: 40 configurations, sigma_max {sur.sigma_max:.4f}", flush=True)
    rings, tof, mc = sur.answer(cfg, mc_photons=20_000, seed=5)
    assert mc[3] and mc[7]
    flagged_in_range = [i for i in range(40) if mc[i] and i not in (3, 7)]
    assert 10 <= len(flagged_in_range) <= 30
    fb = [cfg[i] for i in range(40) if mc[i]]
    _, _, _, W = S.simulate(fb, 20_000, 5)
    assert np.array_equal(np.concatenate([rings[mc], tof[mc]], axis=1), W)
    ok = ~mc
    mu_ok, _ = sur.predict(S.encode_many([cfg[i] for i in range(40) if ok[i]]))
    # exp on the device (the answer path) against numpy's float32 exp: within a few ulps
    np.testing.assert_allclose(np.concatenate([rings[ok], tof[ok]], axis=1), np.exp(mu_ok), rtol=1e-6, atol=0)
    # the same answers when the device path runs in chunks of 7 configurations
    chunk = S.ANSWER_CHUNK
    try:
        S.ANSWER_CHUNK = 7
        rings7, tof7, mc7 = sur.answer(cfg, mc_photons=20_000, seed=5)
    finally:
        S.ANSWER_CHUNK = chunk
    assert np.array_equal(mc7, mc) and np.array_equal(rings7, rings) and np.array_equal(tof7, tof)


def test_device_encoding_equals_the_host_encoding(dev):
    """xmc_surrogate_encode on 20,000 configurations of the sampling
    distribution with a few pushed out of the box, one-layer and six-layer
    counts: the range flags equal in_range_arrays and the feature rows equal
    encode_arrays (double arithmetic rounded to float; the device's log may
    differ from the host's in the last double bit, which moves at most one
    float ulp)."""
    import ctypes as C
    from xmcml_amd import engine as E, surrogate as S
    L, cnt = S.draw_config_arrays(20_000, 31)
    L[::97, 0, 1] = 3.0        # mua above the box
    L[5::113, 1, 4] = 0.01     # a thin layer below the box
    cnt[7::501] = 1
    n = len(cnt)
    Ld = E.DeviceArray(L.shape, L, dtype=np.float64)
    cd = E.DeviceArray(cnt.shape, cnt, dtype=np.int32)
    X = E.DeviceArray((n, S.N_INPUTS))
    ok = E.DeviceArray((n,), dtype=np.uint8)
    E._check(E.lib().xmc_surrogate_encode(Ld.ptr, cd.ptr, n, C.byref(S._layer_box()), X.ptr, ok.ptr, 0, None), "encode")
    Xg, okg = X.numpy(), ok.numpy() != 0
    okh = S.in_range_arrays(L, cnt)
    assert np.array_equal(okg, okh)
    assert 0.9 * n < okh.sum() < 0.99 * n
    Xh = S.encode_arrays(L, cnt)
    ulp = np.spacing(np.abs(Xh).astype(np.float32))
    assert np.all(np.abs(Xg - Xh) <= ulp)
    assert np.mean(Xg == Xh) > 0.999


def test_device_answers_equal_the_network(dev):
    """xmc_mlp_answer on 70,000 rows (two chunks): exp of the predicted means
    and the uncertainty flags equal Surrogate.predict / uncertainty on the host
    (float32 exp and sums: within a few ulps; flags equal away from sigma_max)."""
    from pathlib import Path
    import ctypes as C
    from xmcml_amd import engine as E, surrogate as S
    model = Path(__file__).resolve().parents[1] / "xmcml_amd" / "data" / "surrogate_layered.npz"
    sur = S.Surrogate.load(model, max_batch=4096)
    L, cnt = S.draw_config_arrays(70_000, 8)
    Xh = S.encode_arrays(L, cnt)
    mu, sig = sur.predict(Xh)
    unc = sur.uncertainty(sig)
    sur.sigma_max = float(np.quantile(unc, 0.9))
    n = len(cnt)
    X = E.DeviceArray(Xh.shape, Xh)
    V = E.DeviceArray((n, S.N_OUTPUTS))
    U = E.DeviceArray((n,), dtype=np.uint8)
    pf = C.POINTER(C.c_float)
    std, mean = np.ascontiguousarray(sur.y_std), np.ascontiguousarray(sur.y_mean)
    E._check(E.lib().xmc_mlp_answer(sur.mlp._h, X.ptr, n, std.ctypes.data_as(pf), mean.ctypes.data_as(pf),
                                    S.FLAG_RINGS, sur.sigma_max, V.ptr, U.ptr, None), "answer")
    np.testing.assert_allclose(V.numpy(), np.exp(mu), rtol=1e-6, atol=0)
    flag = U.numpy() != 0
    clear = np.abs(unc - sur.sigma_max) > 1e-5 * sur.sigma_max
    assert np.array_equal(flag[clear], (unc > sur.sigma_max)[clear])
    assert 0.05 * n < flag.sum() < 0.15 * n
