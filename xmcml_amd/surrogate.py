"""Learned surrogate of layered-tissue photon transport, trained on what the
batched GPU Monte Carlo produces, with plain Monte Carlo as its fallback.

Question: a planar stack of 2-5 tissue layers under an n = 1 half-space, a
pencil beam at the origin.  Answer: the diffuse reflectance of 32 rings of
0.5 mm (weight fraction per ring) and the 64-bin time-of-flight histogram
over 0-1 ns of all detected weight.  Both are learned in log space; the
network is the one of include/xmcml_surrogate.h (30 inputs, 4 x 512 GELU,
96 means + 96 log-variances) running on the hand-written matrix-core
kernels of xmcml_amd/csrc/xmc_mlp.hip.

Inputs (30): for each of 5 layer slots, ln(mua), ln(mus'), g, n, ln(d)
mapped linearly from the sampling box to [-1, 1], zeros for empty slots; then
5 slot-occupied flags.  mus' = mus (1 - g).  The sampling box (LAYER_BOX) is
the batched workload's distribution (scenes.draw_layers).

Fallback: a configuration whose layers leave the box, or whose predicted
standard deviation of ln R over the inner rings exceeds the model's
threshold, is answered by plain batched Monte Carlo instead.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import time
from dataclasses import dataclass
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import engine, scenes

MAX_LAYERS = 5
N_INPUTS = 6 * MAX_LAYERS
RINGS = 32
RING_WIDTH = 0.5
TOF_BINS = 64
TOF_NS = 1.0
N_OUTPUTS = RINGS + TOF_BINS
MIN_COUNT = 16  # detected packets below which a target is masked out

# (low, high, log-scaled) of mua [1/mm], mus' [1/mm], g, n, thickness [mm]
LAYER_BOX = [(1e-3, 1.0, True), (0.5, 5.0, True), (0.7, 0.95, False), (1.33, 1.5, False), (0.05, 10.0, True)]
FLAG_RINGS = 16  # rings whose predicted spread decides the fallback to Monte Carlo
ANSWER_CHUNK = 1 << 22  # configurations per pass of the device answer path (~2 GB of device buffers)
RANGE_TOL = 1e-9


class _LayerBox(C.Structure):
    """xmc_layer_box (include/xmcml_surrogate.h)."""
    _fields_ = [("slots", C.c_int), ("minLayers", C.c_int), ("lo", C.c_double * 5), ("hi", C.c_double * 5),
                ("logScale", C.c_int * 5), ("tol", C.c_double)]


def _layer_box() -> "_LayerBox":
    b = _LayerBox()
    b.slots, b.minLayers, b.tol = MAX_LAYERS, 2, RANGE_TOL
    for k, (lo, hi, lg) in enumerate(LAYER_BOX):
        b.lo[k], b.hi[k], b.logScale[k] = lo, hi, int(lg)
    return b

Layer = Tuple[float, float, float, float, float]  # (n, mua, mus, g, thickness), as scenes.layered_scene


def configs_to_arrays(configs: Sequence[Sequence[Layer]]) -> Tuple[np.ndarray, np.ndarray]:
    """(n, 5, 5) layer tuples (n, mua, mus, g, d), zero-padded, and the layer
    counts (n,) of a list of configurations."""
    n = len(configs)
    L = np.zeros((n, MAX_LAYERS, 5))
    cnt = np.zeros(n, np.int64)
    for i, c in enumerate(configs):
        if len(c) > MAX_LAYERS:
            raise ValueError(f"{len(c)} layers: the surrogate takes 1..{MAX_LAYERS}")
        cnt[i] = len(c)
        if c:
            L[i, :len(c)] = np.asarray(c, dtype=np.float64)
    return L, cnt


def _layer_values(L: np.ndarray) -> np.ndarray:
    """(..., 5) of (mua, mus', g, n, d) from layer tuples (n, mua, mus, g, d)."""
    return np.stack([L[..., 1], L[..., 2] * (1.0 - L[..., 3]), L[..., 3], L[..., 0], L[..., 4]], axis=-1)


def encode_arrays(L: np.ndarray, count: np.ndarray) -> np.ndarray:
    """Feature rows (n, 30) of configurations given as (n, 5, 5) layer tuples and
    layer counts: each used slot's (mua, mus', g, n, d) mapped from LAYER_BOX to
    [-1, 1] (log scale where the box says), then the 5 slot-occupied flags."""
    count = np.asarray(count, np.int64)
    if count.size and (count.min() < 1 or count.max() > MAX_LAYERS):
        raise ValueError(f"layer counts must be 1..{MAX_LAYERS}")
    used = np.arange(MAX_LAYERS)[None, :] < count[:, None]
    V = _layer_values(np.asarray(L, np.float64))
    F = np.zeros((len(count), MAX_LAYERS, 5))
    for k, (lo, hi, lg) in enumerate(LAYER_BOX):
        v = V[..., k]
        if lg:
            with np.errstate(divide="ignore", invalid="ignore"):
                u = 2.0 * (np.log(np.where(used, v, lo)) - math.log(lo)) / (math.log(hi) - math.log(lo)) - 1.0
        else:
            u = 2.0 * (v - lo) / (hi - lo) - 1.0
        F[..., k] = np.where(used, u, 0.0)
    X = np.zeros((len(count), N_INPUTS), np.float32)
    X[:, :5 * MAX_LAYERS] = F.reshape(len(count), -1)
    X[:, 5 * MAX_LAYERS:] = used
    return X


def encode(layers: Sequence[Layer]) -> np.ndarray:
    """Feature vector (30,) of one configuration."""
    if not 1 <= len(layers) <= MAX_LAYERS:
        raise ValueError(f"{len(layers)} layers: the surrogate takes 1..{MAX_LAYERS}")
    return encode_arrays(*configs_to_arrays([layers]))[0]


def encode_many(configs: Sequence[Sequence[Layer]]) -> np.ndarray:
    return encode_arrays(*configs_to_arrays(configs)) if configs else np.zeros((0, N_INPUTS), np.float32)


def in_range_arrays(L: np.ndarray, count: np.ndarray, tol: float = RANGE_TOL) -> np.ndarray:
    """Per configuration: 2..5 layers, every used layer inside the sampling box."""
    count = np.asarray(count, np.int64)
    used = np.arange(MAX_LAYERS)[None, :] < count[:, None]
    V = _layer_values(np.asarray(L, np.float64))
    ok = (count >= 2) & (count <= MAX_LAYERS)
    for k, (lo, hi, _) in enumerate(LAYER_BOX):
        inside = (V[..., k] >= lo * (1 - tol)) & (V[..., k] <= hi * (1 + tol))
        ok &= np.all(inside | ~used, axis=1)
    return ok


def in_range(layers: Sequence[Layer], tol: float = 1e-9) -> bool:
    """True when every layer lies in the sampling box (and 2..5 layers)."""
    if not 2 <= len(layers) <= MAX_LAYERS:
        return False
    return bool(in_range_arrays(*configs_to_arrays([layers]), tol=tol)[0])


def draw_config_arrays(count: int, seed: int) -> Tuple[np.ndarray, np.ndarray]:
    """`count` configurations of the batched workload's distribution (the one
    scenes.draw_layers samples), drawn vectorised: (n, 5, 5) layer tuples and
    layer counts uniform in 2..5."""
    rng = np.random.default_rng(seed)
    cnt = rng.integers(2, MAX_LAYERS + 1, size=count)
    shape = (count, MAX_LAYERS)
    mua = np.exp(rng.uniform(math.log(1e-3), math.log(1.0), shape))
    musp = np.exp(rng.uniform(math.log(0.5), math.log(5.0), shape))
    g = rng.uniform(0.7, 0.95, shape)
    n = rng.uniform(1.33, 1.5, shape)
    d = np.exp(rng.uniform(math.log(0.05), math.log(10.0), shape))
    L = np.stack([n, mua, musp / (1.0 - g), g, d], axis=-1)
    L[np.arange(MAX_LAYERS)[None, :] >= cnt[:, None]] = 0.0
    return L, cnt


def arrays_to_configs(L: np.ndarray, count: np.ndarray) -> List[List[Layer]]:
    return [[tuple(float(v) for v in L[i, j]) for j in range(int(count[i]))] for i in range(len(count))]


def draw_configs(count: int, seed: int) -> List[List[Layer]]:
    """`count` configurations of the batched workload's distribution."""
    rng = np.random.default_rng(seed)
    return [scenes.draw_layers(rng, int(rng.integers(2, MAX_LAYERS + 1))) for _ in range(count)]


def config_scene(layers: Sequence[Layer], photons: int) -> scenes.SceneSpec:
    return scenes.layered_scene(list(layers), photons, rings_n=RINGS, ring_width=RING_WIDTH, time_bins=TOF_BINS,
                                time_ns=TOF_NS)


def targets(out, photons: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(ln values, mask, packet counts), each (96,): 32 ring reflectances then
    64 TOF bins of the ring-summed histogram, as weight fractions."""
    w = np.concatenate([out.det_weight(), out.bin_weight().sum(axis=0)]) / float(photons)
    cnt = np.concatenate([out.det_count().astype(np.float64), out.bin_count().sum(axis=0).astype(np.float64)])
    mask = (cnt >= MIN_COUNT) & (w > 0)
    y = np.where(mask, np.log(np.where(w > 0, w, 1.0)), 0.0)
    return y.astype(np.float32), mask.astype(np.float32), cnt


def simulate(configs: Sequence[Sequence[Layer]], photons, seed: int, chunk: int = 1024,
             progress=None) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """Plain batched Monte Carlo of every configuration (photons: int or one
    per configuration), chunk configurations per persistent launch.
    Returns (ln targets, mask, counts, raw weight fractions), each (n, 96)."""
    n = len(configs)
    budgets = np.broadcast_to(np.asarray(photons, dtype=np.int64), (n,))
    Y = np.zeros((n, N_OUTPUTS), np.float32)
    M = np.zeros((n, N_OUTPUTS), np.float32)
    K = np.zeros((n, N_OUTPUTS), np.float64)
    W = np.zeros((n, N_OUTPUTS), np.float64)
    for c0 in range(0, n, chunk):
        specs = [config_scene(configs[i], int(budgets[i])) for i in range(c0, min(n, c0 + chunk))]
        with engine.Batch(specs) as b:
            b.run(seed + c0)
            outs = b.collect()
        for k, (o, s) in enumerate(zip(outs, specs)):
            y, m, cnt = targets(o, s.photons)
            Y[c0 + k], M[c0 + k], K[c0 + k] = y, m, cnt
            W[c0 + k] = np.where(m > 0, np.exp(y), 0.0)
        if progress:
            progress(min(n, c0 + chunk), n)
    return Y, M, K, W


class MLP:
    """The matrix-core MLP of include/xmcml_surrogate.h.  Device arrays are
    engine.DeviceArray (float32, row-major)."""

    def __init__(self, inputs: int = N_INPUTS, hidden: int = 512, depth: int = 4, outputs: int = N_OUTPUTS,
                 max_batch: int = 4096, seed: int = 0, device: int = 0):
        class Desc(C.Structure):
            _fields_ = [("inputs", C.c_int), ("hidden", C.c_int), ("depth", C.c_int), ("outputs", C.c_int),
                        ("maxBatch", C.c_int), ("device", C.c_int)]

        self.inputs, self.hidden, self.depth, self.outputs = inputs, hidden, depth, outputs
        self.max_batch, self.device = max_batch, device
        d = Desc(inputs, hidden, depth, outputs, max_batch, device)
        h = C.c_void_p()
        engine._check(engine.lib().xmc_mlp_create(C.byref(d), seed, C.byref(h)), "xmc_mlp_create")
        self._h = h
        n = C.c_int64()
        engine._check(engine.lib().xmc_mlp_param_count(h, C.byref(n)), "xmc_mlp_param_count")
        self.nparam = n.value

    def close(self):
        if getattr(self, "_h", None):
            engine.lib().xmc_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _vec(self, fn, what) -> np.ndarray:
        a = np.zeros(self.nparam, np.float32)
        engine._check(fn(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), self.nparam), what)
        return a

    def get_params(self) -> np.ndarray:
        return self._vec(engine.lib().xmc_mlp_get_params, "xmc_mlp_get_params")

    def get_grads(self) -> np.ndarray:
        return self._vec(engine.lib().xmc_mlp_get_grads, "xmc_mlp_get_grads")

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.nparam:
            raise ValueError(f"{p.size} parameters, the model has {self.nparam}")
        engine._check(engine.lib().xmc_mlp_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)),
                                                      self.nparam), "xmc_mlp_set_params")

    @staticmethod
    def _dev(a, cols, name):
        if not isinstance(a, engine.DeviceArray):
            raise TypeError(f"{name} must be an engine.DeviceArray")
        if len(a.shape) != 2 or a.shape[1] != cols:
            raise ValueError(f"{name} must have shape (batch, {cols})")
        return a.ptr

    def forward(self, x, out=None, per_layer: bool = False):
        """Head output (n, 2*outputs) of device rows x; returns the device array.
        per_layer: one GEMM launch per layer instead of the fused whole-network
        kernel (same bits)."""
        if out is None:
            out = engine.DeviceArray((x.shape[0], 2 * self.outputs), device=self.device)
        name = "xmc_mlp_forward_layers" if per_layer else "xmc_mlp_forward"
        engine._check(getattr(engine.lib(), name)(self._h, self._dev(x, self.inputs, "x"), x.shape[0],
                                                  self._dev(out, 2 * self.outputs, "out"), None), name)
        return out

    def backward(self, x, y, mask, scale: float, want_loss: bool = True) -> Optional[float]:
        loss = C.c_float()
        engine._check(engine.lib().xmc_mlp_backward(
            self._h, self._dev(x, self.inputs, "x"), self._dev(y, self.outputs, "y"),
            self._dev(mask, self.outputs, "mask"), x.shape[0], scale, C.byref(loss) if want_loss else None, None),
            "xmc_mlp_backward")
        return loss.value if want_loss else None

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        engine._check(engine.lib().xmc_mlp_adam(self._h, lr, beta1, beta2, eps, weight_decay, None), "xmc_mlp_adam")

    def train_step(self, x, y, mask, scale, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        engine._check(engine.lib().xmc_mlp_train_step(
            self._h, self._dev(x, self.inputs, "x"), self._dev(y, self.outputs, "y"),
            self._dev(mask, self.outputs, "mask"), x.shape[0], scale, lr, beta1, beta2, eps, weight_decay, None),
            "xmc_mlp_train_step")

    def train_sampled(self, X, Y, M, batch, seed, step, scale, lr, beta1=0.9, beta2=0.999, eps=1e-8,
                      weight_decay=0.0, want_loss=False) -> Optional[float]:
        loss = C.c_float()
        engine._check(engine.lib().xmc_mlp_train_sampled(
            self._h, self._dev(X, self.inputs, "X"), self._dev(Y, self.outputs, "Y"),
            self._dev(M, self.outputs, "M"), X.shape[0], batch, seed, step, scale, lr, beta1, beta2, eps,
            weight_decay, C.byref(loss) if want_loss else None, None), "xmc_mlp_train_sampled")
        return loss.value if want_loss else None


def gemm_check(a: np.ndarray, b: np.ndarray, tile: int = 0, device: int = 0) -> np.ndarray:
    """C = A B^T on the matrix-core kernel (bf16 operands, fp32 accumulate);
    tile = 128 forces the 128 x 128 tile kernel."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    M, K = a.shape
    N = b.shape[0]
    c = np.zeros((M, N), np.float32)
    pf = C.POINTER(C.c_float)
    engine._check(engine.lib().xmc_mlp_gemm_check(a.ctypes.data_as(pf), b.ctypes.data_as(pf), M, N, K,
                                                  c.ctypes.data_as(pf), tile, device), "xmc_mlp_gemm_check")
    return c


@dataclass
class FitLog:
    steps: int
    seconds: float
    first_loss: float
    last_loss: float


class Surrogate:
    """The trained network plus the target standardisation and the fallback
    threshold."""

    def __init__(self, mlp: MLP, y_mean: np.ndarray, y_std: np.ndarray, sigma_max: float = math.inf):
        self.mlp = mlp
        self.y_mean = np.asarray(y_mean, np.float32)
        self.y_std = np.asarray(y_std, np.float32)
        self.sigma_max = float(sigma_max)

    # ---- training -------------------------------------------------------
    @staticmethod
    def standardisation(Y: np.ndarray, M: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        cnt = np.maximum(M.sum(0), 1.0)
        mean = (Y * M).sum(0) / cnt
        var = (((Y - mean) ** 2) * M).sum(0) / cnt
        return mean.astype(np.float32), np.sqrt(np.maximum(var, 1e-6)).astype(np.float32)

    @classmethod
    def fit(cls, X: np.ndarray, Y: np.ndarray, M: np.ndarray, steps: int = 2000, batch: int = 4096,
            lr: float = 2e-3, lr_final: float = 2e-5, weight_decay: float = 0.0, seed: int = 0,
            hidden: int = 512, depth: int = 4, log_every: int = 0, device: int = 0) -> Tuple["Surrogate", FitLog]:
        """AdamW with a cosine learning rate from lr to lr_final; each step
        draws `batch` rows of the resident data set on the device."""
        mean, std = cls.standardisation(Y, M)
        Ys = np.where(M > 0, (Y - mean) / std, 0.0).astype(np.float32)
        n = X.shape[0]
        batch = min(batch, n)
        Xd = engine.DeviceArray(X.shape, X, device)
        Yd = engine.DeviceArray(Ys.shape, Ys, device)
        Md = engine.DeviceArray(M.shape, M, device)
        mlp = MLP(X.shape[1], hidden, depth, Y.shape[1], batch, seed=seed, device=device)
        scale = 1.0 / (batch * Y.shape[1])
        t0 = time.perf_counter()
        first = last = float("nan")
        for step in range(steps):
            frac = step / max(1, steps - 1)
            lr_t = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * frac))
            want = step == 0 or step == steps - 1 or bool(log_every and step % log_every == 0)
            loss = mlp.train_sampled(Xd, Yd, Md, batch, seed, step, scale, lr_t, weight_decay=weight_decay,
                                     want_loss=want)
            if want:
                if step == 0:
                    first = loss
                last = loss
                if log_every:
                    print(f"surrogate fit: step {step} loss {loss:.5f}", flush=True)
        engine.device_sync(device)
        seconds = time.perf_counter() - t0
        for a in (Xd, Yd, Md):
            a.close()
        return cls(mlp, mean, std), FitLog(steps, seconds, first, last)

    # ---- inference ------------------------------------------------------
    def predict_device(self, x, out=None):
        """Raw head output for feature rows already on the device."""
        return self.mlp.forward(x, out)

    def predict_raw(self, X: np.ndarray) -> np.ndarray:
        """Head output (n, 192) of host feature rows: copied to the device and
        run in chunks of the model's batch size on two reused device buffers."""
        n = len(X)
        O2 = 2 * self.y_mean.size
        out = np.empty((n, O2), np.float32)
        if n == 0:
            return out
        rows = min(self.mlp.max_batch, n)
        x = engine.DeviceArray((rows, self.mlp.inputs), device=self.mlp.device)
        o = engine.DeviceArray((rows, O2), device=self.mlp.device)
        try:
            for r0 in range(0, n, rows):
                m = min(rows, n - r0)
                if m != rows:  # the last, shorter chunk
                    x.close()
                    o.close()
                    x = engine.DeviceArray((m, self.mlp.inputs), device=self.mlp.device)
                    o = engine.DeviceArray((m, O2), device=self.mlp.device)
                x.upload(X[r0:r0 + m])
                self.mlp.forward(x, o)
                out[r0:r0 + m] = o.numpy()
        finally:
            x.close()
            o.close()
        return out

    def predict(self, X: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        """(ln-value means, ln-value standard deviations), each (n, 96)."""
        if len(X) == 0:
            return np.zeros((0, self.y_mean.size)), np.zeros((0, self.y_mean.size))
        out = self.predict_raw(X)
        O = self.y_mean.size
        mu = out[:, :O] * self.y_std + self.y_mean
        sig = np.exp(0.5 * np.clip(out[:, O:], -10, 10)) * self.y_std
        return mu, sig

    def uncertainty(self, sig: np.ndarray) -> np.ndarray:
        """Per-configuration score: mean predicted sd of ln R over the inner 16 rings."""
        return sig[:, :FLAG_RINGS].mean(axis=1)

    def answer_arrays(self, L: np.ndarray, count: np.ndarray, mc_photons: int = 1_000_000, seed: int = 1):
        """answer() on configurations given as (n, 5, 5) layer tuples and layer
        counts.  Encoding, the range check, the network, exp of its means and
        the uncertainty check run on the device (xmc_surrogate_encode,
        xmc_mlp_answer); flagged configurations are simulated.  Returns (rings
        (n, 32), TOF (n, 64), answered-by-MC flags (n,), seconds spent in each
        stage)."""
        t = {"encode_s": 0.0, "infer_s": 0.0}
        n = len(count)
        R = np.zeros((n, N_OUTPUTS))
        ok = np.zeros(n, bool)
        L = np.ascontiguousarray(L, np.float64).reshape(n, MAX_LAYERS, 5)
        count = np.asarray(count)
        dev = self.mlp.device
        std = np.ascontiguousarray(self.y_std, np.float32)
        mean = np.ascontiguousarray(self.y_mean, np.float32)
        pf = C.POINTER(C.c_float)
        for c0 in range(0, n, ANSWER_CHUNK):  # device buffers of at most ANSWER_CHUNK rows
            m = min(ANSWER_CHUNK, n - c0)
            t0 = time.perf_counter()
            Ld = engine.DeviceArray((m, MAX_LAYERS, 5), L[c0:c0 + m], device=dev, dtype=np.float64)
            cd = engine.DeviceArray((m,), count[c0:c0 + m], device=dev, dtype=np.int32)
            X = engine.DeviceArray((m, N_INPUTS), device=dev)
            inr = engine.DeviceArray((m,), device=dev, dtype=np.uint8)
            engine._check(engine.lib().xmc_surrogate_encode(Ld.ptr, cd.ptr, m, C.byref(_layer_box()), X.ptr, inr.ptr,
                                                            dev, None), "xmc_surrogate_encode")
            Ld.close()
            cd.close()
            engine.device_sync(dev)
            t["encode_s"] += time.perf_counter() - t0
            t0 = time.perf_counter()
            V = engine.DeviceArray((m, N_OUTPUTS), device=dev)
            U = engine.DeviceArray((m,), device=dev, dtype=np.uint8)
            engine._check(engine.lib().xmc_mlp_answer(self.mlp._h, X.ptr, m, std.ctypes.data_as(pf),
                                                      mean.ctypes.data_as(pf), FLAG_RINGS, self.sigma_max, V.ptr,
                                                      U.ptr, None), "xmc_mlp_answer")
            okc = (inr.numpy() != 0) & (U.numpy() == 0)
            vals = V.numpy()
            ok[c0:c0 + m] = okc
            R[c0:c0 + m][okc] = vals[okc]
            for a in (X, inr, V, U):
                a.close()
            t["infer_s"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        fb = np.nonzero(~ok)[0]
        if fb.size:
            _, _, _, W = simulate(arrays_to_configs(L[fb], count[fb]), mc_photons, seed)
            R[fb] = W
        t["fallback_s"] = time.perf_counter() - t0
        return R[:, :RINGS], R[:, RINGS:], ~ok, t

    def answer(self, configs: Sequence[Sequence[Layer]], mc_photons: int = 1_000_000, seed: int = 1):
        """Ring reflectance (n, 32) and TOF (n, 64) weight fractions for every
        configuration, with a boolean array of those answered by Monte Carlo:
        configurations outside the sampling box, or whose predicted uncertainty
        exceeds sigma_max, are simulated (plain batched MC, mc_photons each, in
        their order, from `seed`); the others come from the network."""
        L, cnt = configs_to_arrays(configs)
        rings, tof, mc, _ = self.answer_arrays(L, cnt, mc_photons, seed)
        return rings, tof, mc

    # ---- persistence ----------------------------------------------------
    def save(self, path: Path, meta: Optional[dict] = None):
        np.savez_compressed(path, params=self.mlp.get_params(), y_mean=self.y_mean, y_std=self.y_std,
                            sigma_max=np.float64(self.sigma_max),
                            shape=np.array([self.mlp.inputs, self.mlp.hidden, self.mlp.depth, self.mlp.outputs]),
                            meta=np.frombuffer(json.dumps(meta or {}).encode(), np.uint8))

    @classmethod
    def load(cls, path: Path, max_batch: int = 4096, device: int = 0) -> "Surrogate":
        z = np.load(path, allow_pickle=False)
        inputs, hidden, depth, outputs = (int(v) for v in z["shape"])
        mlp = MLP(inputs, hidden, depth, outputs, max_batch, device=device)
        mlp.set_params(z["params"])
        s = cls(mlp, z["y_mean"], z["y_std"], float(z["sigma_max"]))
        s.meta = json.loads(bytes(z["meta"]).decode()) if "meta" in z else {}
        return s


def ring_relative_error(R_pred: np.ndarray, R_ref: np.ndarray, valid: np.ndarray) -> np.ndarray:
    """|R_pred - R_ref| / R_ref over the valid (configuration, ring) entries."""
    return np.abs(R_pred[valid] - R_ref[valid]) / R_ref[valid]
