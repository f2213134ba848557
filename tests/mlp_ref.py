"""CPU reference of the surrogate MLP (test infrastructure only).

Restates, in numpy, the network and the training step that the HIP kernels
implement (xmcml_amd/csrc/xmc_mlp.hip): bfloat16 operands rounded to nearest
even at the same points as the kernels (input, weights, pre-activations,
activations, head gradient, hidden-layer gradients), products accumulated in
float64 (the kernels accumulate in fp32, so results differ by accumulation
order only), tanh-form GELU, masked Gaussian NLL with the log-variance
clamped to +-10, AdamW.

`exact=True` switches every bf16 rounding off and works in float64: that
variant is what finite differences check, which pins the backward formulas
themselves (a transposed product or a dropped GELU' term fails it).
"""
from __future__ import annotations

import numpy as np

GELU_C = 0.7978845608028654
GELU_A = 0.044715
LOGVAR_CLAMP = 10.0


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even); returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return (u.astype(np.uint32) << 16).view(np.float32)


def _tanh(y):
    return 1.0 - 2.0 / (1.0 + np.exp(2.0 * y))


def gelu(z):
    return 0.5 * z * (1.0 + _tanh(GELU_C * (z + GELU_A * z ** 3)))


def gelu_grad(z):
    t = _tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def layer_shapes(inputs, hidden, depth, outputs):
    shapes = []
    fin = inputs
    for _ in range(depth):
        shapes.append((hidden, fin))
        fin = hidden
    shapes.append((2 * outputs, fin))
    return shapes


def unpack(params, shapes):
    """Caller-order flat vector -> [(W [out][in], b [out])]."""
    out, o = [], 0
    for (r, c) in shapes:
        W = params[o:o + r * c].reshape(r, c)
        o += r * c
        b = params[o:o + r]
        o += r
        out.append((W, b))
    assert o == len(params)
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


class RefMLP:
    def __init__(self, params, inputs, hidden, depth, outputs, exact=False):
        self.shapes = layer_shapes(inputs, hidden, depth, outputs)
        self.outputs = outputs
        self.exact = exact
        dt = np.float64 if exact else np.float32
        self.layers = [(W.astype(dt), b.astype(dt)) for W, b in unpack(np.asarray(params), self.shapes)]

    def _r(self, x):
        return np.asarray(x, np.float64) if self.exact else bf16(x).astype(np.float64)

    def _f(self, x):
        """fp32 result of a kernel (float64 in exact mode)."""
        return np.asarray(x, np.float64) if self.exact else np.asarray(x, np.float32).astype(np.float64)

    def forward(self, x):
        a = self._r(x)
        cache = [(None, a)]
        L = len(self.layers)
        for l, (W, b) in enumerate(self.layers):
            acc = self._f(a @ self._r(W).T)
            if l < L - 1:
                z = self._f(acc + b)
                zr = self._r(z)
                a = self._r(self._f(gelu(z if self.exact else z.astype(np.float32))))
                cache.append((zr, a))
            else:
                out = self._f(acc + b)
        return out, cache

    def loss_and_grads(self, x, y, mask, scale):
        out, cache = self.forward(x)
        O = self.outputs
        mu, s = out[:, :O], out[:, O:]
        sc = np.clip(s, -LOGVAR_CLAMP, LOGVAR_CLAMP)
        e = np.exp(-sc)
        d = y - mu
        ed2 = e * d * d
        loss = scale * float(np.sum(mask * 0.5 * (ed2 + sc)))
        dmu = -scale * mask * e * d
        inside = (s >= -LOGVAR_CLAMP) & (s <= LOGVAR_CLAMP)
        ds = np.where(inside, scale * mask * 0.5 * (1.0 - ed2), 0.0)
        g = self._r(np.concatenate([dmu, ds], axis=1))
        grads = [None] * len(self.layers)
        for l in range(len(self.layers) - 1, -1, -1):
            W, _ = self.layers[l]
            a_in = cache[l][1]
            gW = self._f(g.T @ a_in)
            gb = self._f(np.sum(g, axis=0))
            grads[l] = (gW, gb)
            if l > 0:
                zprev = cache[l][0]
                da = self._f(g @ self._r(W))
                gz = gelu_grad(zprev if self.exact else zprev.astype(np.float32))
                g = self._r(self._f(da * gz))
        return loss, out, grads


def adamw(params, grads, m1, m2, step, lr, b1, b2, eps, decay, is_weight):
    """One AdamW update of caller-order flat arrays (float32 arithmetic)."""
    f = np.float32
    c1 = f(1) - f(b1) ** f(step)
    c2 = f(1) - f(b2) ** f(step)
    m1 = f(b1) * m1 + (f(1) - f(b1)) * grads
    m2 = f(b2) * m2 + (f(1) - f(b2)) * grads * grads
    dec = np.where(is_weight, f(decay), f(0))
    p = params - (f(lr) * (m1 / c1) / (np.sqrt(m2 / c2) + f(eps)) + f(lr) * dec * params)
    return p.astype(np.float32), m1.astype(np.float32), m2.astype(np.float32)


def weight_flags(shapes):
    """Caller-order boolean array: True for weights, False for biases."""
    parts = []
    for (r, c) in shapes:
        parts.append(np.ones(r * c, bool))
        parts.append(np.zeros(r, bool))
    return np.concatenate(parts)
