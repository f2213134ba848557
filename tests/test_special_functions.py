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
