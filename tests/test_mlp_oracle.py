"""Pins of the numpy restatement of the bracket network (oracle/mlp_oracle.py)
and host-side pieces of the MLP ABI (CPU only)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402

import grace_amd as ga  # noqa: E402
from grace_amd import workloads as W  # noqa: E402
import _oracle as O  # noqa: E402


def random_problem(n=40, seed=3):
    rng = np.random.default_rng(seed)
    p = M.init_params(11).astype(np.float64)
    p[M.B1:M.W2] = rng.normal(0, 0.3, M.H)
    p[M.B2:M.W3] = rng.normal(0, 0.3, M.H)
    p[M.B3:M.NP] = rng.normal(0, 0.3, M.NO)
    feat = rng.normal(0, 1.5, (M.NF, n)).astype(np.float32)
    scale = np.concatenate([rng.normal(0, 0.2, 10), rng.uniform(0.5, 1.5, 10), [0.3, -0.1], [2.0, 0.5]])
    target = rng.normal(0, 1, (2, n)).astype(np.float32)
    target[1, ::7] = np.nan
    return p, scale.astype(np.float32), feat, target


def test_grad_matches_finite_differences():
    """The backward pass of the oracle is the derivative of its forward pass
    (rounding off, FP64), checked by central differences on every kind of
    parameter: a transposed weight, a dropped ReLU slope or a wrong sign fails."""
    p, scale, feat, target = random_problem()
    g, loss = M.grad(p, scale, feat, target, quant=False)
    assert np.all(g[M.W1 + M.NF * M.H:M.B1] == 0)
    idx = np.concatenate([np.arange(M.W1, M.W1 + 3 * M.H, 37), np.arange(M.B1, M.W2, 9), np.arange(M.W2, M.B2, 211),
                          np.arange(M.B2, M.W3, 13), np.arange(M.W3, M.B3, 17), np.arange(M.B3, M.NP)])
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = M.grad(pp, scale, feat, target, quant=False)[1]
        lm = M.grad(pm, scale, feat, target, quant=False)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_grad_loss_definition():
    """loss = (1/n) sum over finite targets of (y - t_normalised)^2."""
    p, scale, feat, target = random_problem()
    y, _ = M.forward(p, scale, feat, quant=True)
    sc = scale.astype(np.float64)
    tn = (target.astype(np.float64) - sc[20:22, None]) / sc[22:24, None]
    ok = np.isfinite(tn)
    ref = np.sum(np.where(ok, (y - np.where(ok, tn, 0)) ** 2, 0)) / feat.shape[1]
    assert abs(M.grad(p, scale, feat, target)[1] - ref) < 1e-12 * ref


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 10, 100000), rng.normal(0, 1e-3, 1000),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0])]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(x), ref)


def test_init_params_bitexact_with_library():
    for seed in (0, 20261202, 2 ** 63 + 5):
        assert np.array_equal(ga.Mlp.init_params(seed), M.init_params(seed))


def test_feature_bracket_contains_mu_plus():
    """The closed-form ends x_lo, x_hi of the features bound the oracle's
    Newton root mu+ h0 (the bracket of kastaun_c2p.hh:101-117)."""
    st = W.c2p_states(3000, seed=17)
    roots = np.zeros((4, 3000))
    O.c2p_batch(st["params"], st["cons"], st["met"], r=st["rad"], roots=roots)
    f, lxhi, xlo, xhi = M.features(st["cons"], st["met"], 1 + 1e-15)
    m = roots[2] > 0
    newton = m & (roots[2] < 1.0)
    assert newton.sum() > 2000
    assert np.all(roots[2][newton] >= xlo[newton] * (1 - 2e-10))
    assert np.all(roots[2][newton] <= xhi[newton] * (1 + 2e-10))
    assert np.all(np.isfinite(f))


def test_mlp_abi_symbols_present():
    L = ga.lib()
    for name in ga.ABI_SYMBOLS:
        assert hasattr(L, name)
