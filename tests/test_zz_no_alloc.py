"""xmc_scene_query_device, the allocation-free ray query (rays and results in
the caller's device memory, the BVH stack spill in the library's static
device array), against xmc_scene_query, which allocates its buffers: the same
walk, so the same hits bit for bit -- on two calls in a row, the second with
fewer rays, through the C-ABI."""
import ctypes as C

import numpy as np
import pytest

from xmcml_amd import scenes

pytestmark = pytest.mark.gpu


def _device_query(engine, sc, o, d, s, brute):
    n = len(s)
    print(f"no-alloc query: {n} rays, brute={brute}", flush=True)
    bufs = dict(o=engine.DeviceArray((n, 3), o.astype(np.float32)), d=engine.DeviceArray((n, 3), d.astype(np.float32)),
                s=engine.DeviceArray((n,), s.astype(np.float32)), found=engine.DeviceArray((n,)),
                dist=engine.DeviceArray((n,)), surf=engine.DeviceArray((n,)), tri=engine.DeviceArray((n,)),
                nrm=engine.DeviceArray((n, 3)))
    try:
        rc = engine.lib().xmc_scene_query_device(
            sc._h, bufs["o"].ptr, bufs["d"].ptr, bufs["s"].ptr, n, int(brute), bufs["found"].ptr,
            bufs["dist"].ptr, bufs["surf"].ptr, bufs["tri"].ptr, bufs["nrm"].ptr, None)
        engine._check(rc, "xmc_scene_query_device")
        engine.device_sync()
        return dict(found=bufs["found"].numpy().view(np.int32), distance=bufs["dist"].numpy().astype(np.float64),
                    surface=bufs["surf"].numpy().view(np.int32), triangle=bufs["tri"].numpy().view(np.int32),
                    normal=bufs["nrm"].numpy().astype(np.float64))
    finally:
        for b in bufs.values():
            b.close()


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform([-19.0, -19.0, -1.5], [19.0, 19.0, 3.0], size=(n, 3)).astype(np.float32).astype(np.float64)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32).astype(np.float64)
    s = rng.exponential(0.3, size=n).astype(np.float32).astype(np.float64)
    s[rng.random(n) < 0.2] = 0.0
    return o, d, s


@pytest.mark.parametrize("brute", [False, True])
def test_no_alloc_query_equals_the_allocating_query(engine, brute):
    spec = scenes.heightfield_scene(10) if not brute else scenes.small_scene("heightfield", 10)
    with engine.Scene(spec) as sc:
        for n, seed in ((300_000 if not brute else 20_000, 1), (5_000, 2)):  # the second call is smaller
            o, d, s = _rays(n, seed)
            a = sc.query(o, d, s, brute=brute)
            b = _device_query(engine, sc, o, d, s, brute)
            assert a["found"].sum() > 0.05 * n
            for k in ("found", "surface", "triangle"):
                assert np.array_equal(a[k], b[k]), k
            assert np.array_equal(a["distance"], b["distance"])
            assert np.array_equal(a["normal"], b["normal"])
