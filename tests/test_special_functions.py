"""The float log and sincos(2 pi u) of the 32-bit transport (oracle restatement,
mirrored by the GPU kernel) against libm in double, over every uniform the
engine can draw: u = k 2^-24, k = 1 .. 2^24 - 1."""
import numpy as np

import oracle_lib as ol


def _all_uniforms():
    return (np.arange(1, 1 << 24, dtype=np.float64) * 2.0 ** -24).astype(np.float32)


def _ulp(x):
    return np.spacing(np.abs(x).astype(np.float32)).astype(np.float64)


def test_log_every_uniform():
    u = _all_uniforms()
    y = np.empty_like(u)
    ol.oracle().xmco_logf_many(u.ctypes.data, y.ctypes.data, u.size)
    exact = np.log(u.astype(np.float64))
    err = np.abs(y.astype(np.float64) - exact) / _ulp(exact)
    assert err.max() <= 2.0, err.max()
    # the step sampler needs -log(u) > 0 for every u < 1
    assert np.all(y < 0)


def test_log_general_arguments():
    x = np.exp(np.linspace(-80.0, 80.0, 200001)).astype(np.float32)
    y = np.empty_like(x)
    ol.oracle().xmco_logf_many(x.ctypes.data, y.ctypes.data, x.size)
    exact = np.log(x.astype(np.float64))
    err = np.abs(y.astype(np.float64) - exact) / _ulp(exact)
    assert err.max() <= 2.0


def test_sincos_2pi_every_uniform():
    u = np.concatenate([[0.0], _all_uniforms()]).astype(np.float32)
    s = np.empty_like(u)
    c = np.empty_like(u)
    ol.oracle().xmco_sincos2pi_many(u.ctypes.data, s.ctypes.data, c.ctypes.data, u.size)
    ang = 2.0 * np.pi * u.astype(np.float64)
    es, ec = np.sin(ang), np.cos(ang)
    # absolute error in units of 2^-24 (the magnitude of an ulp near 1)
    assert np.abs(s - es).max() < 2.5 * 2.0 ** -24
    assert np.abs(c - ec).max() < 2.5 * 2.0 ** -24
    # exact special points and unit norm
    assert s[0] == 0.0 and c[0] == 1.0
    q = (u * 4).astype(np.float64)
    on_axis = q == np.round(q)
    assert np.all(np.abs(s[on_axis]) <= 1.0) and set(np.round(c[on_axis] ** 2 + s[on_axis] ** 2, 6)) == {1.0}
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1.0).max() < 4 * 2.0 ** -24


def test_hg_float_matches_the_reference_inversion():
    """The float scattering angle (1 - cos, 1 + cos without cancellation)
    against ComputeCosineTheta's closed form evaluated in double, over every
    uniform: cos within 4 * 2^-23 absolute and sin within 8 ulp of the true
    sin -- relative accuracy of sin(theta) also for the strongly forward and
    backward scatters where the cos-first formula loses it (measured 0.85-5.6 ulp)."""
    u = np.concatenate([[0.0], _all_uniforms()]).astype(np.float32)
    for g in (0.0, 0.5, 0.8, 0.9, 0.95, -0.3):
        c = np.empty_like(u)
        s = np.empty_like(u)
        ol.oracle().xmco_hg_float_many(np.float32(g), u.ctypes.data, c.ctypes.data, s.ctypes.data, u.size)
        gd = float(np.float32(g))
        ud = u.astype(np.float64)
        if gd != 0.0:
            t = (1 - gd * gd) / (1 - gd + 2 * gd * ud)
            exact = (1 + gd * gd - t * t) / (2 * gd)
            vers = (1 - gd) * (1 - ud) * (t + 1 - gd) / (1 - gd + 2 * gd * ud)
            cove = (1 + gd) * ud * (1 + gd + t) / (1 - gd + 2 * gd * ud)
        else:
            exact = 2 * ud - 1
            vers, cove = 2 * (1 - ud), 2 * ud
        sin_exact = np.sqrt(vers * cove)
        assert np.abs(c - np.clip(exact, -1, 1)).max() < 4 * 2.0 ** -23
        assert np.all(np.abs(s - sin_exact) <= 8 * np.spacing(sin_exact.astype(np.float32)) + 1e-30)
        assert np.all(np.abs(c) <= 1.0)
