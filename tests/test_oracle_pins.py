"""Pins of the oracle's LayerNorm, attention, FermiNet aggregation and whole
log|psi| pipelines (CPU only).

Each check is an answer the mathematics fixes, chosen so that a plausible
mistake fails it:

* LayerNorm (flax nn.LayerNorm, epsilon 1e-5; wavefunction/backbone/psiformer.py:74,86):
  zero mean and mean square v/(v+eps) per row; a row of variance ~eps
  separates eps inside the rsqrt from eps outside; a constant row gives the
  bias exactly; torch.nn.functional.layer_norm agrees.
* Multi-head attention (flax MultiHeadDotProductAttention; psiformer.py:76-82):
  zero query/key weights and biases give uniform weights (output = mean_j v_j
  W_o + b_o); one electron gives v W_o + b_o; a hand-computed two-electron,
  one-head case fails a misplaced 1/sqrt(dh); torch's
  scaled_dot_product_attention agrees on random inputs.
* FermiNet aggregation (wavefunction/backbone/ferminet.py:82-90,
  utils/array.py:24-45): hand values for an asymmetric two-electron stream,
  whose mean over the first index differs from the mean over the second, in
  the two-channel and the polarized (one channel) case.
* Whole pipelines: value, gradient and Laplacian of the oracle's FermiNet and
  Psiformer log|psi| against an independent float64 torch evaluation built from
  torch library operations (layer_norm, softmax, linalg.slogdet) whose gradient
  comes from autograd and whose Laplacian is the trace of the autograd Hessian
  (double backward), for every LayerNorm placement and the module options.
"""

import numpy as np
import pytest
import torch

from jaqmc_amd import params as P
from jaqmc_amd import systems as S
from oracle import wavefunction as wf
from oracle.lap import Lap
from torch_ref import _t, torch_logpsi as _torch_logpsi

EPS = 1e-5


def _lap_const(v):
    v = np.asarray(v, np.float64)
    return Lap.const(v, 1, 0, np.float64)


# ---------------------------------------------------------------------------
# LayerNorm
# ---------------------------------------------------------------------------
def test_layernorm_moments():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((5, 16)) * rng.uniform(0.01, 3.0, (5, 1)) + rng.standard_normal((5, 1))
    y = wf.layer_norm(_lap_const(x), np.ones(16), np.zeros(16), EPS).v[0]
    v = x.var(axis=1)  # biased variance, as flax
    np.testing.assert_allclose(y.mean(axis=1), 0.0, atol=1e-12)
    np.testing.assert_allclose((y ** 2).mean(axis=1), v / (v + EPS), rtol=1e-12)


def test_layernorm_eps_inside_rsqrt():
    # a row of variance exactly eps: eps inside the rsqrt gives mean square 1/2,
    # eps added to the standard deviation would give (sqrt(eps)/(sqrt(eps)+eps))^2
    x = np.array([[-1.0, 1.0, -1.0, 1.0]]) * np.sqrt(EPS)
    y = wf.layer_norm(_lap_const(x), np.ones(4), np.zeros(4), EPS).v[0]
    assert np.mean(y ** 2) == pytest.approx(0.5, rel=1e-12)


def test_layernorm_constant_row_gives_bias():
    x = np.full((1, 8), 3.25)
    b = np.linspace(-1, 1, 8)
    y = wf.layer_norm(_lap_const(x), np.full(8, 7.0), b, EPS).v[0]
    np.testing.assert_array_equal(y, b[None])


def test_layernorm_matches_torch():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((6, 32))
    s, b = rng.standard_normal(32), rng.standard_normal(32)
    y = wf.layer_norm(_lap_const(x), s, b, EPS).v[0]
    t = torch.nn.functional.layer_norm(torch.from_numpy(x), (32,), torch.from_numpy(s), torch.from_numpy(b), EPS)
    np.testing.assert_allclose(y, t.numpy(), rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# attention
# ---------------------------------------------------------------------------
def _mha_params(rng, D, zero_qk=False):
    p = {}
    for nm in ("q", "k", "v", "o"):
        p["w" + nm] = rng.standard_normal((D, D)) / np.sqrt(D)
        p["b" + nm] = rng.standard_normal(D) * 0.1
    if zero_qk:
        for nm in ("q", "k"):
            p["w" + nm] = np.zeros((D, D))
            p["b" + nm] = np.zeros(D)
    return p


def test_mha_uniform_weights_when_scores_vanish():
    rng = np.random.default_rng(2)
    D, N = 8, 5
    p = _mha_params(rng, D, zero_qk=True)
    x = rng.standard_normal((1, N, D))
    out = wf.mha(Lap.const(x[0], 1, 0, np.float64), p, 2).v[0]
    v = x[0] @ p["wv"] + p["bv"]
    expect = v.mean(axis=0) @ p["wo"] + p["bo"]
    np.testing.assert_allclose(out, np.broadcast_to(expect, (N, D)), rtol=1e-12)


def test_mha_single_electron():
    rng = np.random.default_rng(3)
    D = 8
    p = _mha_params(rng, D)
    x = rng.standard_normal((1, D))
    out = wf.mha(Lap.const(x, 1, 0, np.float64), p, 4).v[0]
    np.testing.assert_allclose(out, (x @ p["wv"] + p["bv"]) @ p["wo"] + p["bo"], rtol=1e-12)


def test_mha_two_electrons_hand_softmax():
    # one head, dh = 4, identity projections, no biases: scores s_ij = x_i.x_j / 2
    D = 4
    eye = np.eye(D)
    p = {"wq": eye, "wk": eye, "wv": eye, "wo": eye}
    for nm in ("q", "k", "v", "o"):
        p["b" + nm] = np.zeros(D)
    x = np.array([[1.0, 0.0, 0.0, 0.0], [1.0, 2.0, 0.0, 0.0]])
    out = wf.mha(Lap.const(x, 1, 0, np.float64), p, 1).v[0]
    # row 0: s = (1/2, 1/2) -> weights (1/2, 1/2); row 1: s = (1/2, 5/2)
    w1 = np.exp([0.5, 2.5]) / np.exp([0.5, 2.5]).sum()
    np.testing.assert_allclose(out[0], 0.5 * x[0] + 0.5 * x[1], rtol=1e-12)
    np.testing.assert_allclose(out[1], w1[0] * x[0] + w1[1] * x[1], rtol=1e-12)
    # with the scale misplaced (none, or 1/dh) row 1 would differ
    w_bad = np.exp([1.0, 5.0]) / np.exp([1.0, 5.0]).sum()
    assert not np.allclose(out[1], w_bad[0] * x[0] + w_bad[1] * x[1], rtol=1e-3)


def test_mha_matches_torch_sdpa():
    rng = np.random.default_rng(4)
    D, H, N = 12, 3, 6
    p = _mha_params(rng, D)
    x = rng.standard_normal((N, D))
    out = wf.mha(Lap.const(x, 1, 0, np.float64), p, H).v[0]
    t = lambda a: torch.from_numpy(a)
    q = (t(x) @ t(p["wq"]) + t(p["bq"])).reshape(N, H, D // H).transpose(0, 1)
    k = (t(x) @ t(p["wk"]) + t(p["bk"])).reshape(N, H, D // H).transpose(0, 1)
    v = (t(x) @ t(p["wv"]) + t(p["bv"])).reshape(N, H, D // H).transpose(0, 1)
    o = torch.nn.functional.scaled_dot_product_attention(q, k, v).transpose(0, 1).reshape(N, D)
    ref = o @ t(p["wo"]) + t(p["bo"])
    np.testing.assert_allclose(out, ref.numpy(), rtol=1e-11, atol=1e-12)


# ---------------------------------------------------------------------------
# FermiNet aggregation
# ---------------------------------------------------------------------------
def test_aggregation_hand_values_two_channels():
    # 3 electrons (2 up, 1 down), h1 one feature, h2[i, j] = 10 i + j
    h1 = np.array([[1.0], [2.0], [4.0]])
    h2 = np.array([[10.0 * i + j for j in range(3)] for i in range(3)])[..., None]
    out = wf.aggregate(Lap.const(h1, 1, 0, np.float64), Lap.const(h2, 1, 0, np.float64), (2, 1)).v[0]
    # columns: h1 | mean_up h1 | mean_dn h1 | mean_{i in up} h2[i, j] | mean_{i in dn} h2[i, j]
    expect = np.array([[1.0, 1.5, 4.0, 5.0, 20.0],
                       [2.0, 1.5, 4.0, 6.0, 21.0],
                       [4.0, 1.5, 4.0, 7.0, 22.0]])
    np.testing.assert_array_equal(out, expect)
    # the mean over the second index would give 1, 11, 21 / 2, 12, 22 instead
    assert not np.allclose(out[:, 3], [1.0, 11.0, 21.0])


def test_aggregation_polarized_single_channel():
    h1 = np.array([[1.0], [2.0], [6.0]])
    h2 = np.array([[10.0 * i + j for j in range(3)] for i in range(3)])[..., None]
    out = wf.aggregate(Lap.const(h1, 1, 0, np.float64), Lap.const(h2, 1, 0, np.float64), (3, 0)).v[0]
    expect = np.array([[1.0, 3.0, 10.0], [2.0, 3.0, 11.0], [6.0, 3.0, 12.0]])
    np.testing.assert_array_equal(out, expect)


# ---------------------------------------------------------------------------
# whole pipelines against an independent torch float64 evaluation
# ---------------------------------------------------------------------------
def _torch_value_grad_lap(spec, p, x, atoms):
    xt = _t(x).requires_grad_(True)
    f = _torch_logpsi(spec, p, xt, _t(atoms))
    (g,) = torch.autograd.grad(f, xt, create_graph=True)
    g = g.reshape(-1)
    lap = 0.0
    for c in range(g.numel()):
        (h,) = torch.autograd.grad(g[c], xt, retain_graph=True)
        lap += float(h.reshape(-1)[c])
    return float(f.detach()), g.detach().numpy(), lap


PIPELINES = [
    ("ferminet", dict(hidden_single=[8, 8, 8], hidden_double=[4, 4, 4], ndets=2), (2, 1)),
    ("ferminet", dict(hidden_single=[6, 6], hidden_double=[3, 3], ndets=2, envelope="isotropic"), (2, 2)),
    ("ferminet", dict(hidden_single=[6, 6], hidden_double=[3, 3], ndets=2, orbitals_spin_split=False), (2, 2)),
    ("ferminet", dict(hidden_single=[6, 6], hidden_double=[3, 3], ndets=3), (3, 0)),
    ("psiformer", dict(num_layers=2, num_heads=2, heads_dim=4, mlp_hidden=[6], ndets=2), (2, 2)),
    ("psiformer", dict(num_layers=2, num_heads=2, heads_dim=4, mlp_hidden=[6], ndets=2, layer_norm_mode="post"),
     (2, 2)),
    ("psiformer", dict(num_layers=1, num_heads=1, heads_dim=4, mlp_hidden=[5], ndets=2, layer_norm_mode="null",
                       rescale=False, bias_orbitals=True, orbitals_spin_split=False), (2, 1)),
]


@pytest.mark.parametrize("kind,kw,nspins", PIPELINES)
def test_pipeline_matches_torch_autograd(kind, kw, nspins):
    N = sum(nspins)
    atoms = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]])
    sysm = S.make_system("pin", atoms, [float(N - N // 2), float(N // 2)], nspins)
    spec = P.ModelSpec(kind, nspins, 2)
    for k, v in kw.items():
        setattr(spec, k, v)
    p = P.cast_params(P.init_params(spec, seed=5, bias_scale=0.2, env_jitter=0.2), np.float64)
    x = S.init_walkers(sysm, 2, seed=6).astype(np.float64)
    _, lp = wf.logpsi(spec.as_oracle(), p, x, atoms, with_laplacian=True)
    for w in range(2):
        v, g, L = _torch_value_grad_lap(spec, p, x[w], atoms)
        assert lp.v[w] == pytest.approx(v, rel=1e-11, abs=1e-11)
        np.testing.assert_allclose(lp.J[w], g, rtol=1e-9, atol=1e-10)
        assert lp.lap[w] == pytest.approx(L, rel=1e-8, abs=1e-9)


@pytest.mark.parametrize("kind,kw", [
    ("ferminet", dict(hidden_single=[8, 8, 8], hidden_double=[4, 4, 4], ndets=2)),
    ("ferminet", dict(hidden_single=[6, 6], hidden_double=[3, 3], ndets=2, envelope="isotropic",
                      orbitals_spin_split=False)),
    ("psiformer", dict(num_layers=2, num_heads=2, heads_dim=4, mlp_hidden=[5], ndets=2)),
    ("psiformer", dict(num_layers=2, num_heads=2, heads_dim=4, mlp_hidden=[5], ndets=2, layer_norm_mode="post")),
    ("psiformer", dict(num_layers=1, num_heads=2, heads_dim=4, mlp_hidden=[], ndets=2, layer_norm_mode="null",
                       rescale=False, bias_orbitals=True, orbitals_spin_split=False)),
])
def test_torch_parameter_gradient_matches_oracle_directional_fd(kind, kw):
    """The torch evaluation's autograd parameter gradient (the reference for the
    GPU's jq_logpsi_vjp) against central finite differences of the numpy
    oracle's log|psi| along random parameter directions (Psiformer: every
    layer's parameters, the 1/sqrt(dh) placement and the Jastrow alphas)."""
    from torch_ref import flat_leaves, param_tree_to_torch

    atoms = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]])
    sysm = S.make_system("pin", atoms, [2.0, 2.0], (2, 2))
    spec = P.ModelSpec(kind, (2, 2), 2)
    for k, v in kw.items():
        setattr(spec, k, v)
    p = P.cast_params(P.init_params(spec, seed=7, bias_scale=0.2, env_jitter=0.3), np.float64)
    if kind == "psiformer":
        p["jastrow"] = np.array([0.8, 1.3])
    x = S.init_walkers(sysm, 3, seed=8).astype(np.float64)
    c = np.array([0.7, -1.3, 2.1])
    pt = param_tree_to_torch(p)
    f = sum(ci * _torch_logpsi(spec, pt, torch.from_numpy(x[w]), torch.from_numpy(atoms)) for w, ci in enumerate(c))
    leaves = flat_leaves(spec, pt)
    grads = torch.autograd.grad(f, leaves, allow_unused=True)  # the LayerNorm slots of 'null' are unused
    g = np.concatenate([(gi if gi is not None else torch.zeros_like(li)).reshape(-1).detach().numpy()
                        for gi, li in zip(grads, leaves)])
    flat = P.pack(spec, p).astype(np.float64)
    rng = np.random.default_rng(9)
    m = spec.as_oracle()

    def F(v):
        q = P.unpack(spec, v, p, dtype=np.float64)
        return float(c @ wf.logpsi(m, q, x, atoms, with_laplacian=False)[1].v)

    for _ in range(4):
        d = rng.standard_normal(flat.size)
        h = 1e-5
        fd = (F(flat + h * d) - F(flat - h * d)) / (2 * h)
        assert g @ d == pytest.approx(fd, rel=1e-6, abs=1e-8)
