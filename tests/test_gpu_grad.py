"""Parameter gradients of FermiNet and Psiformer log|psi| on the GPU
(jq_logpsi_vjp) against
float64 torch autograd of the test restatement (itself pinned against the
numpy oracle's log|psi| and its finite differences along parameter
directions, tests/test_oracle_pins.py), and the LossAndGrad assembly
(estimator/loss_grad.py:70-128).

Tolerance: the GPU evaluates the network in FP32 (FP64 determinant adjoint),
so the contracted gradient agrees with the FP64 one to |d| <= 1e-4 |g|_inf +
1e-3 |g| component-wise (relative FP32 error of the network's adjoint, a few
hundred roundings deep)."""

import numpy as np
import pytest

from jaqmc_amd import params as P
from jaqmc_amd import systems as S

pytestmark = pytest.mark.gpu

CASES = {
    "li": ("li", dict(hidden_single=[64, 64], hidden_double=[16, 16])),
    "n2_small": ("n2", dict(hidden_single=[32, 32, 32], hidden_double=[8, 16, 16], ndets=3)),
    "n2_shared_iso": ("n2", dict(hidden_single=[32, 64], hidden_double=[8, 8], ndets=2, envelope="isotropic",
                                 orbitals_spin_split=False)),
    "h_polarized": (None, dict(hidden_single=[32, 32], hidden_double=[8, 8], ndets=2)),
}


def _setup(case, B):
    name, kw = CASES[case]
    if name is None:
        sysm = S.make_system("Li(3,0)", [[0.0, 0.0, 0.0]], [3.0], (3, 0))
        _, spec, *_ = S.workload("li")
        spec.nspins, spec.n_atoms = sysm.nspins, 1
    else:
        sysm, spec, *_ = S.workload(name)
    for k, v in kw.items():
        setattr(spec, k, v)
    p = P.init_params(spec, seed=31, bias_scale=0.1, env_jitter=0.2)
    x = S.init_walkers(sysm, B, seed=32)
    return sysm, spec, p, x


def _torch_grad(spec, p, x, atoms, cot):
    import torch

    from torch_ref import flat_leaves, param_tree_to_torch, torch_logpsi

    pt = param_tree_to_torch(P.cast_params(p, np.float64))
    f = sum(float(c) * torch_logpsi(spec, pt, torch.from_numpy(x[w].astype(np.float64)),
                                    torch.from_numpy(np.asarray(atoms, np.float64)))
            for w, c in enumerate(cot))
    g = torch.autograd.grad(f, flat_leaves(spec, pt))
    return np.concatenate([gi.reshape(-1).numpy() for gi in g])


@pytest.mark.parametrize("case", list(CASES))
def test_logpsi_vjp_matches_torch_autograd(cuda, case):
    import torch

    from jaqmc_amd.device import DeviceModel

    print(f"[grad] {case}", flush=True)
    sysm, spec, p, x = _setup(case, 6)
    dm = DeviceModel(spec, sysm.atoms, sysm.charges, p, chunk_walkers=4, with_gradients=True)
    cot = np.array([0.5, -1.25, 2.0, 0.0, 0.75, -0.3])
    g = dm.logpsi_vjp(torch.from_numpy(x).to(cuda), torch.from_numpy(cot).to(cuda)).cpu().numpy()
    ref = _torch_grad(spec, p, x, sysm.atoms, cot)
    assert g.shape == ref.shape == (P.pack(spec, p).size,)
    tol = 1e-4 * np.abs(ref).max() + 1e-3 * np.abs(ref)
    bad = np.abs(g - ref) > tol
    print(f"[grad] {case}: max |d| {np.abs(g - ref).max():.2e}, |g|inf {np.abs(ref).max():.2e}, "
          f"violations {int(bad.sum())}", flush=True)
    assert not bad.any()
    # linearity in the weights: walker 3 has weight 0; one-hot weights give per-walker gradients
    one = np.zeros(6)
    one[2] = 1.0
    g2 = dm.logpsi_vjp(torch.from_numpy(x).to(cuda), torch.from_numpy(one).to(cuda)).cpu().numpy()
    r2 = _torch_grad(spec, p, x, sysm.atoms, one)
    assert np.all(np.abs(g2 - r2) <= 1e-4 * np.abs(r2).max() + 1e-3 * np.abs(r2))
    dm.close()


PSI_CASES = {
    # pre-LN, two layers, one hidden MLP width that is not a multiple of 32
    "psi_pre": (dict(num_layers=2, num_heads=2, heads_dim=32, mlp_hidden=[40], ndets=2), (2, 2)),
    # no LayerNorm, shared orbitals with biases, isotropic envelope, unscaled features, MLP without hidden layer
    "psi_null_shared": (dict(num_layers=1, num_heads=2, heads_dim=32, mlp_hidden=[], ndets=3, layer_norm_mode="null",
                             rescale=False, bias_orbitals=True, orbitals_spin_split=False, envelope="isotropic"),
                        (3, 1)),
    # spin-polarised, 64-wide heads, two hidden MLP layers
    "psi_polarized": (dict(num_layers=1, num_heads=2, heads_dim=64, mlp_hidden=[24, 48], ndets=2), (3, 0)),
    # post-LN, no attention biases
    "psi_post": (dict(num_layers=2, num_heads=2, heads_dim=32, mlp_hidden=[32], ndets=2, layer_norm_mode="post",
                      with_bias=False), (2, 2)),
}


def _psi_setup(case, B):
    kw, nspins = PSI_CASES[case]
    ne = nspins[0] + nspins[1]
    charges = [float(ne - 1), 1.0]
    sysm = S.make_system(case, [[0.0, 0.0, 0.0], [0.0, 0.3, 1.5]], charges, nspins)
    spec = P.ModelSpec("psiformer", nspins, 2, **kw)
    p = P.init_params(spec, seed=41, bias_scale=0.1, env_jitter=0.2)
    p["jastrow"] = np.array([0.8, 1.3], np.float32)
    x = S.init_walkers(sysm, B, seed=42)
    return sysm, spec, p, x


def _torch_grad_unused(spec, p, x, atoms, cot):
    """As _torch_grad, with zeros for parameters the model leaves unused (the
    LayerNorm slots of layer_norm_mode='null')."""
    import torch

    from torch_ref import flat_leaves, param_tree_to_torch, torch_logpsi

    pt = param_tree_to_torch(P.cast_params(p, np.float64))
    f = sum(float(c) * torch_logpsi(spec, pt, torch.from_numpy(x[w].astype(np.float64)),
                                    torch.from_numpy(np.asarray(atoms, np.float64)))
            for w, c in enumerate(cot))
    leaves = flat_leaves(spec, pt)
    g = torch.autograd.grad(f, leaves, allow_unused=True)
    return np.concatenate([(gi if gi is not None else torch.zeros_like(li)).reshape(-1).detach().numpy()
                           for gi, li in zip(g, leaves)])


@pytest.mark.parametrize("case", list(PSI_CASES))
def test_psiformer_logpsi_vjp_matches_torch_autograd(cuda, case):
    """Psiformer: input projection, LayerNorms, Q/K/V (with the 1/sqrt(dh)
    scale of the pack's q kernel), attention, output projection, MLP, orbital
    kernels and biases, envelopes and the Jastrow parameters."""
    import torch

    from jaqmc_amd.device import DeviceModel

    print(f"[grad] {case}", flush=True)
    sysm, spec, p, x = _psi_setup(case, 6)
    dm = DeviceModel(spec, sysm.atoms, sysm.charges, p, chunk_walkers=4, with_gradients=True)
    cot = np.array([0.5, -1.25, 2.0, 0.0, 0.75, -0.3])
    g = dm.logpsi_vjp(torch.from_numpy(x).to(cuda), torch.from_numpy(cot).to(cuda)).cpu().numpy()
    ref = _torch_grad_unused(spec, p, x, sysm.atoms, cot)
    assert g.shape == ref.shape == (P.pack(spec, p).size,)
    tol = 1e-4 * np.abs(ref).max() + 1e-3 * np.abs(ref)
    bad = np.abs(g - ref) > tol
    names = [n for n, a in P.param_arrays(spec, p) for _ in range(a.size)]
    worst = sorted({names[i] for i in np.flatnonzero(bad)})[:8]
    print(f"[grad] {case}: max |d| {np.abs(g - ref).max():.2e}, |g|inf {np.abs(ref).max():.2e}, "
          f"violations {int(bad.sum())} {worst}", flush=True)
    assert not bad.any()
    # every parameter group the model uses has a non-zero gradient somewhere
    off = 0
    for n, a in P.param_arrays(spec, p):
        if not (spec.layer_norm_mode == "null" and ".ln" in n):
            assert np.abs(ref[off:off + a.size]).max() > 0, n
        off += a.size
    dm.close()


def test_psiformer_gradient_options_rejected():
    """Psiformer gradients outside the fused attention's range (head width 128)
    are rejected at creation."""
    from jaqmc_amd import _lib
    from jaqmc_amd.device import DeviceModel

    sysm, spec, _, _ = _psi_setup("psi_pre", 2)
    spec.num_heads, spec.heads_dim = 1, 128
    p = P.init_params(spec, seed=1)
    with pytest.raises(_lib.JqError):
        DeviceModel(spec, sysm.atoms, sysm.charges, p, with_gradients=True)


def test_vjp_chunking_and_zero_weights(cuda):
    """Chunked evaluation equals one chunk; all-zero weights give an exact zero
    gradient (no walker's determinants are inverted)."""
    import torch

    from jaqmc_amd.device import DeviceModel

    sysm, spec, p, x = _setup("n2_small", 12)
    xe = torch.from_numpy(x).to(cuda)
    cot = torch.from_numpy(np.linspace(-1, 1, 12)).to(cuda)
    a = DeviceModel(spec, sysm.atoms, sysm.charges, p, chunk_walkers=5, with_gradients=True)
    b = DeviceModel(spec, sysm.atoms, sysm.charges, p, chunk_walkers=64, with_gradients=True)
    ga, gb = a.logpsi_vjp(xe, cot), b.logpsi_vjp(xe, cot)
    z = a.logpsi_vjp(xe, torch.zeros(12, dtype=torch.float64, device=cuda))
    torch.cuda.synchronize()
    np.testing.assert_allclose(ga.cpu().numpy(), gb.cpu().numpy(), rtol=1e-5, atol=1e-6 * ga.abs().max().item())
    assert torch.count_nonzero(z).item() == 0
    a.close()
    b.close()


def test_loss_and_grad_pipeline(cuda):
    """LossAndGrad on the device: E_L -> MAD clipping -> weights -> contracted
    gradient, against the same assembly from per-walker torch gradients."""
    import torch

    from jaqmc_amd.interface import EuclideanKinetic, FermiNetWavefunction, LossAndGrad, MoleculeData

    sysm, spec, p, x = _setup("li", 16)
    wf = FermiNetWavefunction(spec.nspins, 1, ndets=spec.ndets, hidden_dims_single=spec.hidden_single,
                              hidden_dims_double=spec.hidden_double, envelope=spec.envelope, with_gradients=True)
    data = MoleculeData(torch.from_numpy(x).to(cuda), sysm.atoms, sysm.charges)
    stats, _ = EuclideanKinetic(wf).evaluate_batch_walkers(p, data)
    e = (stats["energy:kinetic"] + stats["energy:potential"])
    lg = LossAndGrad(wf, clip_method="mad", clip_scale=5.0)
    ws, _ = lg.evaluate_batch_walkers(p, data, {"total_energy": e})
    out = lg.finalize_stats(lg.reduce(ws), None)
    ev = e.cpu().numpy()
    med = np.median(ev)
    mad = np.median(np.abs(ev - med))
    c = np.clip(ev, med - 5 * mad, med + 5 * mad)
    w = 2 * (c - c.mean()) / c.size
    ref = _torch_grad(spec, p, x, sysm.atoms, w)
    g = out["grads_flat"]
    assert np.all(np.abs(g - ref) <= 1e-4 * np.abs(ref).max() + 1e-3 * np.abs(ref))
    assert out["loss"] == pytest.approx(ev.mean(), rel=1e-12)
    # the unpacked tree has the parameter tree's shapes
    assert out["grads"]["single"][0][0].shape == p["single"][0][0].shape


def test_loss_and_grad_pipeline_psiformer(cuda):
    """LossAndGrad with a Psiformer wavefunction: E_L -> MAD clipping ->
    weights -> contracted gradient, against per-walker torch gradients."""
    import torch

    from jaqmc_amd.interface import EuclideanKinetic, LossAndGrad, MoleculeData, PsiformerWavefunction

    sysm, spec, p, x = _psi_setup("psi_pre", 12)
    wf = PsiformerWavefunction(spec.nspins, 2, ndets=spec.ndets, num_layers=spec.num_layers,
                               num_heads=spec.num_heads, heads_dim=spec.heads_dim, mlp_hidden_dims=spec.mlp_hidden,
                               with_gradients=True)
    assert wf.spec == spec
    data = MoleculeData(torch.from_numpy(x).to(cuda), sysm.atoms, sysm.charges)
    stats, _ = EuclideanKinetic(wf).evaluate_batch_walkers(p, data)
    e = (stats["energy:kinetic"] + stats["energy:potential"])
    lg = LossAndGrad(wf, clip_method="mad", clip_scale=5.0)
    ws, _ = lg.evaluate_batch_walkers(p, data, {"total_energy": e})
    out = lg.finalize_stats(lg.reduce(ws), None)
    ev = e.cpu().numpy()
    med = np.median(ev)
    mad = np.median(np.abs(ev - med))
    c = np.clip(ev, med - 5 * mad, med + 5 * mad)
    w = 2 * (c - c.mean()) / c.size
    ref = _torch_grad_unused(spec, p, x, sysm.atoms, w)
    g = out["grads_flat"]
    assert np.all(np.abs(g - ref) <= 1e-4 * np.abs(ref).max() + 1e-3 * np.abs(ref))
    assert out["grads"]["layers"][0]["wq"].shape == p["layers"][0]["wq"].shape


def test_psiformer_vjp_at_the_ring42_shape(cuda):
    """The 42-electron ring Psiformer (4 layers, 4 heads of 64, 16
    determinants) at its real shape, two walkers, against float64 torch
    autograd."""
    import torch

    from jaqmc_amd.device import DeviceModel

    sysm, spec, *_ = S.workload("ring42")
    p = P.init_params(spec, seed=43, bias_scale=0.05, env_jitter=0.1)
    x = S.init_walkers(sysm, 2, seed=44)
    cot = np.array([0.8, -0.6])
    dm = DeviceModel(spec, sysm.atoms, sysm.charges, p, with_gradients=True)
    g = dm.logpsi_vjp(torch.from_numpy(x).to(cuda), torch.from_numpy(cot).to(cuda)).cpu().numpy()
    ref = _torch_grad_unused(spec, p, x, sysm.atoms, cot)
    tol = 1e-4 * np.abs(ref).max() + 1e-3 * np.abs(ref)
    bad = np.abs(g - ref) > tol
    print(f"[grad] ring42: max |d| {np.abs(g - ref).max():.2e}, |g|inf {np.abs(ref).max():.2e}, "
          f"violations {int(bad.sum())} of {g.size}", flush=True)
    assert not bad.any()
    dm.close()
