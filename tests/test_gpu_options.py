"""GPU parity of the reference's module options, the orbitals export and the
device-side sampler (through the C ABI), against the CPU oracle.

Options: Psiformer post-LN and no-LN layers (wavefunction/backbone/psiformer.py:
13-26,70-98), unscaled features (rescale=False), attention / input layers
without biases, an orbital bias, orbitals shared between spin channels
(app/molecule/wavefunction/psiformer.py:80-87, ferminet.py:47-49), an MLP width
that is not a multiple of 32 (its GEMM zero-fills the padding columns the next
GEMM reads) and a head width without the fused attention kernel (the batched
GEMM + softmax path).  Tolerances as tests/test_gpu_parity.py.
"""

import numpy as np
import pytest

from jaqmc_amd import params as P
from jaqmc_amd import systems as S
from oracle import hamiltonian as ham
from oracle import wavefunction as owf
from parity_util import cancellation, el_report, logpsi_bound

pytestmark = pytest.mark.gpu


def _ring6():
    r = S.ring42()
    return S.make_system("ring6", r.atoms[:6] * 0.5, [1.0] * 6, (3, 3))


PSI = dict(num_layers=2, num_heads=2, heads_dim=32, mlp_hidden=[32], ndets=2)
CASES = {
    "post_ln": ("psiformer", dict(PSI, layer_norm_mode="post")),
    "null_ln": ("psiformer", dict(PSI, layer_norm_mode="null")),
    "unscaled_nobias_orbbias_shared": ("psiformer", dict(PSI, rescale=False, with_bias=False, input_bias=False,
                                                         bias_orbitals=True, orbitals_spin_split=False)),
    "mlp48": ("psiformer", dict(PSI, mlp_hidden=[48])),
    "heads128_unfused": ("psiformer", dict(num_layers=1, num_heads=1, heads_dim=128, mlp_hidden=[64], ndets=2)),
    "ferminet_shared_orbitals": ("ferminet", dict(hidden_single=[32, 32, 32], hidden_double=[8, 16, 16], ndets=3,
                                                   orbitals_spin_split=False)),
}

_cache = {}


def _setup(case, B=6):
    if case in _cache:
        return _cache[case]
    kind, kw = CASES[case]
    if kind == "psiformer":
        sysm = _ring6()
        _, spec, *_ = S.workload("ring42")
        spec.nspins, spec.n_atoms = sysm.nspins, sysm.atoms.shape[0]
    else:
        sysm, spec, *_ = S.workload("n2")
    for k, v in kw.items():
        setattr(spec, k, v)
    p = P.init_params(spec, seed=21, bias_scale=0.1, env_jitter=0.2)
    x = S.init_walkers(sysm, B, seed=22)
    p64 = P.cast_params(p, np.float64)
    at = sysm.atoms.astype(np.float64)
    ref = ham.local_energy(spec.as_oracle(), p64, x.astype(np.float64), at, sysm.charges.astype(np.float64))
    A = np.asarray(owf.orbital_matrices(spec.as_oracle(), p64, x.astype(np.float64), at))
    ref["A"] = A
    ref["kappa"] = np.linalg.cond(A).reshape(B, -1).max(axis=1)
    ref["cancel"] = cancellation(A)
    _cache[case] = (sysm, spec, p, x, ref)
    return _cache[case]


def _model(sysm, spec, p, prec="fp32", chunk=0):
    from jaqmc_amd.device import DeviceModel

    return DeviceModel(spec, sysm.atoms, sysm.charges, p, precision=prec, chunk_walkers=chunk)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec", ["fp32", "fp16x3"])
def test_module_options_match_oracle(cuda, case, prec):
    import torch

    print(f"[options] {case} {prec}", flush=True)
    sysm, spec, p, x, ref = _setup(case)
    dm = _model(sysm, spec, p, prec=prec, chunk=4)
    xe = torch.from_numpy(x).to(cuda)
    lp, sg = dm.logpsi(xe)
    out = dm.local_energy(xe)
    torch.cuda.synchronize()
    d_lp = np.abs(lp.cpu().numpy() - ref["logpsi"])
    assert np.all(d_lp <= logpsi_bound(ref["logpsi"], ref["kappa"], prec, ref["cancel"])), (d_lp, ref["kappa"])
    np.testing.assert_array_equal(sg.cpu().numpy(), ref["sign"])
    rel, bnd = el_report(case, out["e_loc"].cpu().numpy(), ref["e_loc"], ref["kappa"], prec, ref["cancel"])
    assert np.all(rel <= bnd)
    dm.close()


@pytest.mark.parametrize("case", ["ferminet_shared_orbitals", "post_ln"])
def test_orbitals_export_matches_oracle(cuda, case):
    """jq_orbitals: the (B, K, N, N) envelope-multiplied orbital matrices of
    MoleculeWavefunction.orbitals, chunked like every other entry point."""
    import torch

    sysm, spec, p, x, ref = _setup(case)
    dm = _model(sysm, spec, p, chunk=4)
    A = dm.orbitals(torch.from_numpy(x).to(cuda)).cpu().numpy()
    scale = np.abs(ref["A"]).max(axis=(2, 3), keepdims=True)
    np.testing.assert_allclose(A, ref["A"], rtol=2e-5, atol=2e-5 * scale.max())
    dm.close()


def test_full_n2_orbitals_export(cuda):
    import torch

    sysm, spec, *_ = S.workload("n2")
    p = P.init_params(spec, seed=3)
    x = S.init_walkers(sysm, 16, seed=4)
    dm = _model(sysm, spec, p)
    A = dm.orbitals(torch.from_numpy(x).to(cuda)).cpu().numpy()
    ref = owf.orbital_matrices(spec.as_oracle(), P.cast_params(p, np.float64), x.astype(np.float64),
                               sysm.atoms.astype(np.float64))
    np.testing.assert_allclose(A, ref, rtol=1e-4, atol=1e-5 * np.abs(ref).max())
    # log|det| of the exported matrices reproduces log|psi| (single-det combination)
    s, l = np.linalg.slogdet(A.astype(np.float64))
    m = l.max(axis=1, keepdims=True)
    lp = np.log(np.abs((s * np.exp(l - m)).sum(1))) + m[:, 0]
    np.testing.assert_allclose(lp, dm.logpsi(torch.from_numpy(x).to(cuda))[0].cpu().numpy(), rtol=1e-5, atol=1e-4)
    dm.close()


def test_device_sampler_equals_host_loop(cuda):
    """jq_sample (moves + MCMCSampler.step's adaptation on the device, no host
    synchronisation) against the same loop driven from the host with
    jq_logpsi / jq_mh_steps and the reference's adaptation rule
    (sampler/mcmc.py:139-197): identical walkers, accept counts and state."""
    import torch

    from jaqmc_amd.device import DeviceModel, reduce_accepts

    sysm, spec, *_ = S.workload("li")
    p = P.init_params(spec, seed=5)
    x0 = S.init_walkers(sysm, 96, seed=6)
    dm = DeviceModel(spec, sysm.atoms, sysm.charges, p, chunk_walkers=40)
    n_outer, steps, af, lo, hi = 7, 3, 2, 0.5, 0.55
    width0 = 0.3
    # device
    xd = torch.from_numpy(x0).to(cuda)
    st = torch.zeros(3 + af, dtype=torch.float64, device=cuda)
    st[0] = width0
    pm = torch.zeros(n_outer, dtype=torch.float64, device=cuda)
    logp_d, acc_d = dm.sample(xd, st, n_outer, steps, adapt_frequency=af, pmove_range=(lo, hi), seed=17, ctr=5,
                              pmove_out=pm)
    # host loop
    xh = torch.from_numpy(x0).to(cuda)
    acc_h = torch.zeros(96, dtype=torch.int32, device=cuda)
    width, pmoves, counter, prev = width0, np.zeros(af), 0, 0.0
    pms = []
    for o in range(n_outer):
        logp = 2 * dm.logpsi(xh)[0]
        dm.mh_steps(xh, logp, steps, float(np.float32(width)), seed=17, ctr=5 + o * steps, n_accept=acc_h)
        tot = float(reduce_accepts(acc_h).item())
        pmove = (tot - prev) / (steps * 96)
        prev = tot
        pms.append(pmove)
        counter += 1
        t = counter % af
        pmoves[t] = pmove
        if t == 0:
            if pmoves.mean() > hi:
                width *= 1.1
            elif pmoves.mean() < lo:
                width /= 1.1
    torch.cuda.synchronize()
    assert torch.equal(xd, xh)
    assert torch.equal(acc_d, acc_h)
    assert torch.equal(logp_d, logp)
    s = st.cpu().numpy()
    assert s[0] == pytest.approx(width, rel=1e-15) and s[1] == counter and s[2] == prev
    np.testing.assert_allclose(s[3:], pmoves, rtol=1e-15)
    np.testing.assert_allclose(pm.cpu().numpy(), pms, rtol=1e-15)
    assert width != width0  # the adaptation did act
    dm.close()


def test_burn_in_reaches_the_target_acceptance(cuda):
    """A burn-in of 60 device sampler steps (SamplingStage burn-in,
    workflow/stage/sampling.py:246-301) from a poor width moves the mean
    acceptance into the reference's adaptation window."""
    import torch

    from jaqmc_amd.device import DeviceModel

    sysm, spec, *_ = S.workload("n2")
    spec.hidden_single, spec.hidden_double, spec.ndets = [32, 32], [8, 8], 2
    p = P.init_params(spec, seed=8)
    x = torch.from_numpy(S.init_walkers(sysm, 512, seed=9)).to(cuda)
    dm = DeviceModel(spec, sysm.atoms, sysm.charges, p)
    af = 2
    st = torch.zeros(3 + af, dtype=torch.float64, device=cuda)
    st[0] = 1.5  # far too wide: most moves rejected
    pm = torch.zeros(60, dtype=torch.float64, device=cuda)
    dm.sample(x, st, 60, 5, adapt_frequency=af, seed=3, pmove_out=pm)
    torch.cuda.synchronize()
    pm = pm.cpu().numpy()
    assert pm[0] < 0.3
    assert 0.35 < pm[-10:].mean() < 0.7, pm
    assert st[0].item() < 1.5
    dm.close()
