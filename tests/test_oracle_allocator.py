"""Pins of the allocator's CPU oracle (oracle/mlp_oracle.py): the network's
backpropagation against central finite differences, GELU against its closed
forms, the BF16 rounding against torch's float32 -> bfloat16 conversion, and
the GEV fit and tail against scipy.stats.genextreme."""
import numpy as np
import pytest
from scipy import stats

from oracle import mlp_oracle as mo


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=5000) * 10 ** rng.uniform(-8, 8, size=5000),
                        # exact ties between two bfloat16 values (round to even)
                        (np.arange(1, 200, dtype=np.float32).view(np.uint32) | 0x8000).view(np.float32)]
                       ).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(mo.bf16_round(x), ref)


def test_gelu_closed_forms():
    z = np.linspace(-8, 8, 2001)
    assert np.allclose(mo.gelu(z), z * stats.norm.cdf(z), rtol=1e-12, atol=1e-300)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.allclose(mo.gelu_d(z), fd, rtol=1e-7, atol=1e-8)
    assert mo.gelu(0.0) == 0.0 and mo.gelu_d(0.0) == 0.5


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(1)
    net = mo.Mlp(5, 7, 2, 3)
    params = (rng.normal(size=net.n_params) * 0.5).astype(np.float32)
    x = rng.normal(size=(11, 5))
    t = rng.normal(size=(11, 3))
    w = np.array([1.0, 0.5, 2.0])
    loss, g = net.loss_grad(params, x, t, w, round_weights=False)
    p64 = params.astype(np.float64)
    fd = np.zeros_like(g)
    for i in range(net.n_params):
        e = np.zeros_like(p64)
        e[i] = 1e-5

        def f(pp):
            y = _forward64(net, pp, x)
            return np.sum(w * (y - t) ** 2) / x.shape[0]
        fd[i] = (f(p64 + e) - f(p64 - e)) / 2e-5
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-8)
    assert loss == pytest.approx(np.sum(w * (_forward64(net, p64, x) - t) ** 2) / 11, rel=1e-6)


def _forward64(net, p, x):
    """FP64 forward without any float32 step (for the finite differences)."""
    h, off = x, 0
    for l, (i, o) in enumerate(net.dims):
        W = p[off:off + i * o].reshape(o, i)
        off += i * o
        b = p[off:off + o]
        off += o
        z = h @ W.T + b
        h = mo.gelu(z) if l + 1 < len(net.dims) else z
    return h


@pytest.mark.parametrize("xi", [-0.2, 0.0, 0.1, 0.25])
def test_gev_pwm_fit_recovers_parameters(xi):
    mu, sigma = 1.5, 0.7
    x = stats.genextreme.rvs(-xi, loc=mu, scale=sigma, size=400000, random_state=3)
    m, s, k = mo.gev_fit_pwm(x)
    assert m == pytest.approx(mu, abs=0.01)
    assert s == pytest.approx(sigma, rel=0.02)
    assert k == pytest.approx(xi, abs=0.02)


@pytest.mark.parametrize("xi", [-0.3, 0.0, 0.2])
def test_gev_tail_matches_scipy(xi):
    x = np.linspace(-2, 8, 101)
    ref = stats.genextreme.sf(x, -xi, loc=0.5, scale=1.2)
    assert np.allclose(mo.gev_sf(x, 0.5, 1.2, xi), ref, rtol=1e-10, atol=1e-15)
