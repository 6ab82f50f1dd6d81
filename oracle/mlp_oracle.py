"""CPU oracle of the learned permutation allocator (test infrastructure only).

Only the tests, __graft_entry__.smoke() and the CPU-baseline leg of bench.py
use this module, as the checker; the product never imports it.  The
allocator has no counterpart in the reference: this restates, in FP64 numpy,
what the library's network and fitting steps are specified to compute.

- The network: n_in -> n_hidden (n_layers, exact GELU x Phi(x)) -> n_out,
  forward with BF16-rounded weights (round to nearest even) and FP32 biases,
  loss sum_j w_j (y - t)^2 averaged over the rows, gradients by
  backpropagation.  Pinned in tests/test_oracle_allocator.py by central
  finite differences of the same loss, by GELU against its closed forms and
  by the BF16 rounding against torch's float32 -> bfloat16 conversion.
- The GEV fit by probability-weighted moments (Hosking, Wallis & Wood 1985,
  Technometrics 27:251-261) and the GEV tail p = 1 - exp(-(1 + xi (x - mu) /
  sigma)^(-1/xi)).  Pinned by fits of large samples drawn from known GEVs
  (scipy.stats.genextreme, whose shape c = -xi) and the tail against scipy.
"""
import math

import numpy as np
from scipy.special import erfc, gamma


def bf16_round(a):
    """float32 -> bfloat16 (nearest even) -> float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * erfc(-z / math.sqrt(2.0))


def gelu_d(z):
    return 0.5 * erfc(-z / math.sqrt(2.0)) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


class Mlp:
    def __init__(self, n_in, n_hidden, n_layers, n_out):
        self.dims = []
        prev = n_in
        for _ in range(n_layers):
            self.dims.append((prev, n_hidden))
            prev = n_hidden
        self.dims.append((prev, n_out))
        self.n_params = sum(i * o + o for i, o in self.dims)

    def unpack(self, params, round_weights=True):
        p = np.asarray(params, dtype=np.float32)
        Ws, bs, off = [], [], 0
        for i, o in self.dims:
            W = p[off:off + i * o].reshape(o, i)
            Ws.append(bf16_round(W) if round_weights else W.astype(np.float64))
            off += i * o
            bs.append(p[off:off + o].astype(np.float64))
            off += o
        return Ws, bs

    def forward(self, params, x, keep=False, round_weights=True):
        Ws, bs = self.unpack(params, round_weights)
        h = np.asarray(x, dtype=np.float64)
        zs, hs = [], [h]
        L = len(self.dims)
        for l in range(L):
            z = h @ Ws[l].T + bs[l]
            h = gelu(z) if l + 1 < L else z
            zs.append(z)
            hs.append(h)
        return (h, zs, hs, Ws) if keep else h

    def loss_grad(self, params, x, t, w, round_weights=True):
        y, zs, hs, Ws = self.forward(params, x, keep=True, round_weights=round_weights)
        B = y.shape[0]
        w = np.asarray(w, dtype=np.float64)
        d = y - np.asarray(t, dtype=np.float64)
        loss = float(np.sum(w * d * d) / B)
        dz = 2.0 * w * d / B
        grads = []
        L = len(self.dims)
        for l in range(L - 1, -1, -1):
            dW = dz.T @ hs[l]
            db = dz.sum(axis=0)
            grads.append((dW, db))
            if l > 0:
                dz = (dz @ Ws[l]) * gelu_d(zs[l - 1])
        grads.reverse()
        flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
        return loss, flat


# ---- GEV ------------------------------------------------------------------------
def gev_fit_pwm(sample):
    """(mu, sigma, xi) of a GEV fitted by probability-weighted moments
    (Hosking et al. 1985, eqs. 14-15; their k is -xi)."""
    x = np.sort(np.asarray(sample, dtype=np.float64))
    n = x.shape[0]
    j = np.arange(n, dtype=np.float64)
    b0 = x.mean()
    b1 = np.sum(j / (n - 1) * x) / n
    b2 = np.sum(j * (j - 1) / ((n - 1) * (n - 2)) * x) / n
    c = (2 * b1 - b0) / (3 * b2 - b0) - math.log(2) / math.log(3)
    k = 7.8590 * c + 2.9554 * c * c
    if abs(k) < 1e-6:
        sigma = (2 * b1 - b0) / math.log(2)
        mu = b0 - 0.5772156649015329 * sigma
        return mu, sigma, 0.0
    g = gamma(1 + k)
    sigma = (2 * b1 - b0) * k / (g * (1 - 2.0 ** (-k)))
    mu = b0 + sigma * (g - 1) / k
    return mu, sigma, -k


def gev_sf(x, mu, sigma, xi):
    """P(X >= x) of a GEV (upper tail)."""
    z = (np.asarray(x, dtype=np.float64) - mu) / sigma
    if abs(xi) < 1e-9:
        return -np.expm1(-np.exp(-z))
    tt = 1 + xi * z
    with np.errstate(divide="ignore", invalid="ignore"):
        out = -np.expm1(-np.power(np.maximum(tt, 0.0), -1.0 / xi))
    return np.where(tt > 0, out, 1.0 if xi > 0 else 0.0)
