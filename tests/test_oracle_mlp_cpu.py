"""Pins the learned-proposal oracle (oracle/mlp_oracle.c) without a GPU.

The network has no counterpart in the reference, so the oracle is pinned against
independent implementations: torch autograd (f64) with torch.distributions'
MixtureSameFamily for the loss, scipy's normal log-density for the mixture, torch's own
bfloat16 rounding for the operand rounding, central finite differences for the backward
pass, and torch.optim.Adam for the update rule.
"""
import numpy as np
import pytest
import torch
from scipy.special import logsumexp
from scipy.stats import norm

from oracle import OracleMlp, adam_step, mixture_nll

D = torch.distributions


def rand_params(dims, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    in_dim, hid, nh, nc = dims
    out = []
    for l in range(nh + 1):
        di = in_dim if l == 0 else hid
        do = 9 * nc if l == nh else hid
        b = scale / np.sqrt(di)
        out.append(rng.uniform(-b, b, do * di))
        out.append(rng.uniform(-b, b, do))
    return np.concatenate(out).astype(np.float32)


def unpack(p, dims):
    in_dim, hid, nh, nc = dims
    layers, o = [], 0
    for l in range(nh + 1):
        di = in_dim if l == 0 else hid
        do = 9 * nc if l == nh else hid
        W = p[o:o + do * di].reshape(do, di)
        o += do * di
        b = p[o:o + do]
        o += do
        layers.append((W, b))
    return layers


def rand_batch(dims, B, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.2, (B, dims[0])).astype(np.float32)
    th = rng.uniform(-np.pi, np.pi, B)
    y = np.stack([rng.uniform(-0.5, 0.5, B), rng.uniform(-0.5, 0.5, B), np.sin(th), np.cos(th)], 1).astype(np.float32)
    return x, y


def torch_loss(params, x, y, dims, bf16=False):
    """Network + mixture NLL in torch f64 (bf16=True rounds inputs, weights and hidden
    activations through torch's bfloat16 conversion, forward only)."""
    nc = dims[3]
    rb = (lambda t: t.to(torch.bfloat16).to(torch.float64)) if bf16 else (lambda t: t)
    P = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    h = rb(torch.tensor(x, dtype=torch.float64))
    layers = unpack(np.arange(params.size), dims)
    for l, (Wi, bi) in enumerate(layers):
        W = P[Wi.ravel()].reshape(Wi.shape)
        b = P[bi]
        z = h @ rb(W).T + b
        h = rb(torch.nn.functional.gelu(z)) if l + 1 < len(layers) else z
    out = h
    logits = out[:, :nc]
    mu = out[:, nc:5 * nc].reshape(-1, nc, 4)
    s = out[:, 5 * nc:].reshape(-1, nc, 4).clamp(-7.0, 3.0)
    mix = D.MixtureSameFamily(D.Categorical(logits=logits), D.Independent(D.Normal(mu, s.exp()), 1))
    loss = -mix.log_prob(torch.tensor(y, dtype=torch.float64)).mean()
    return P, out, loss


def test_mixture_nll_matches_scipy():
    rng = np.random.default_rng(1)
    nc, B = 5, 40
    out = np.concatenate([rng.normal(size=(B, nc)), rng.normal(size=(B, 4 * nc)) * 0.5,
                          rng.uniform(-2.0, 1.0, (B, 4 * nc))], 1)
    y = rng.normal(size=(B, 4)) * 0.5
    got = mixture_nll(out, y, nc)
    logw = out[:, :nc] - logsumexp(out[:, :nc], axis=1, keepdims=True)
    mu = out[:, nc:5 * nc].reshape(B, nc, 4)
    sd = np.exp(out[:, 5 * nc:].reshape(B, nc, 4))
    lp = norm.logpdf(y[:, None, :], mu, sd).sum(-1) + logw
    ref = -logsumexp(lp, axis=1).mean()
    assert abs(got - ref) < 1e-12 * max(1.0, abs(ref))


def test_mixture_clamps_log_sigma():
    """Log-sigmas outside [-7, 3] act as the bounds and get no gradient."""
    nc = 2
    out = np.zeros((1, 9 * nc))
    out[0, 5 * nc] = 5.0  # component 0, dimension 0
    out[0, 5 * nc + 4] = -9.0  # component 1, dimension 0
    y = np.array([[0.1, -0.2, 0.3, 0.4]])
    v, d = mixture_nll(out, y, nc, with_grad=True)
    out2 = out.copy()
    out2[0, 5 * nc] = 3.0
    out2[0, 5 * nc + 4] = -7.0
    assert v == mixture_nll(out2, y, nc)
    assert d[0, 5 * nc] == 0.0 and d[0, 5 * nc + 4] == 0.0 and d[0, 5 * nc + 1] != 0.0


@pytest.mark.parametrize("dims,B", [((7, 11, 2, 3), 5), ((20, 16, 3, 4), 9)])
def test_loss_grad_matches_torch_autograd(dims, B):
    p = rand_params(dims, 3)
    x, y = rand_batch(dims, B, 4)
    om = OracleMlp(*dims)
    assert om.param_count() == p.size
    loss, g = om.loss_grad(p, x, y, bf16=False)
    P, out, tl = torch_loss(p, x, y, dims)
    tl.backward()
    assert abs(loss - tl.item()) < 1e-12 * max(1.0, abs(loss))
    np.testing.assert_allclose(g, P.grad.numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(om.forward(p, x, bf16=False), out.detach().numpy(), rtol=1e-12, atol=1e-12)


def test_loss_grad_finite_differences():
    dims = (6, 8, 2, 2)
    p = rand_params(dims, 5, scale=2.0).astype(np.float64)
    x, y = rand_batch(dims, 4, 6)
    om = OracleMlp(*dims)
    # f32 parameters: p +- 2**-10 is exact for |p| < 2**13, so the central difference is clean
    _, g = om.loss_grad(p.astype(np.float32), x, y, bf16=False)
    rng = np.random.default_rng(7)
    for i in rng.choice(p.size, 12, replace=False):
        h = 2.0 ** -10
        pp = p.astype(np.float32).copy()
        pm = pp.copy()
        pp[i] += np.float32(h)
        pm[i] -= np.float32(h)
        fp = om.loss_grad(pp, x, y, bf16=False)[0]
        fm = om.loss_grad(pm, x, y, bf16=False)[0]
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[i]) < 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_forward_matches_torch_rounding():
    dims = (37, 24, 2, 3)
    p = rand_params(dims, 8)
    x, y = rand_batch(dims, 6, 9)
    om = OracleMlp(*dims)
    _, out, tl = torch_loss(p, x, y, dims, bf16=True)
    np.testing.assert_allclose(om.forward(p, x, bf16=True), out.detach().numpy(), rtol=1e-12, atol=1e-12)
    loss, _ = om.loss_grad(p, x, y, bf16=True)
    assert abs(loss - tl.item()) < 1e-12 * max(1.0, abs(loss))


def test_bf16_backward_close_to_exact():
    """The GPU rounds the upstream gradient to bf16 inside the products; the result stays
    within bf16 resolution of exact back-propagation."""
    dims = (40, 32, 2, 4)
    p = rand_params(dims, 10)
    x, y = rand_batch(dims, 16, 11)
    om = OracleMlp(*dims)
    _, ge = om.loss_grad(p, x, y, bf16=False)
    _, gb = om.loss_grad(p, x, y, bf16=True)
    assert np.linalg.norm(gb - ge) < 2e-2 * np.linalg.norm(ge)


def test_adam_matches_torch():
    rng = np.random.default_rng(12)
    p0 = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(3)]
    t = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([t], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for k, g in enumerate(grads, 1):
        opt.zero_grad()
        t.grad = torch.tensor(g)
        opt.step()
        p, m, v = adam_step(p, m, v, g, k, 1e-3)
    np.testing.assert_allclose(p, t.detach().numpy(), rtol=0, atol=1e-15)
