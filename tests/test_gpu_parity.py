"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Tolerances (FP32 path vs the FP64 oracle on FP32 parameters and FP32 walker
positions; u = 2^-24, kappa = the walker's largest 2-norm condition number of
its orbital matrices, from the oracle; bounds in tests/parity_util.py):
  log|psi|  : |d| <= 1e-5 |log psi| + 1e-4 + 4 e kappa (forward path vs the
              oracle), e = 16 u the relative error of the orbital entries and
              4 e kappa the first-order perturbation bound of log|det A|; the
              Laplacian-mode value vs the forward path the same bound (the two
              paths round differently)
  E_L       : per walker, relative error <= 1e-5 + 2 e kappa (first-order
              amplification of the entry error by the inverse); the median
              and the fraction of walkers within 1e-5 are printed (1e-5 holds
              for the well-conditioned walkers, kappa below ~10^2)
  grad      : |d| <= 2e-4 max(1, |grad|) per component (median), 5e-3 (max)
Metropolis decisions with identical proposals and uniforms agree wherever the
oracle's log ratio is farther from log u than 1e-6 plus twice the measured
difference between the GPU's and the oracle's log ratio (a decision can only
flip when the two log ratios straddle log u); the count inside that band is
printed.
"""

import numpy as np
import pytest

import os

from jaqmc_amd import params as P
from jaqmc_amd import systems as S
from oracle import hamiltonian as ham
from oracle import wavefunction as owf
from parity_util import ENTRY_ERR, U32, cancellation, el_report, logpsi_bound

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu

CASES = {
    "li": ("li", {}),
    "n2_small": ("n2", dict(hidden_single=[32, 32, 32], hidden_double=[8, 16, 16], ndets=3)),
    "n2": ("n2", {}),
    "h60": ("h60", {}),
    "n2_iso": ("n2", dict(hidden_single=[64, 64], hidden_double=[16, 16], ndets=2, envelope="isotropic")),
}

_cache = {}


def _ring6():
    r = S.ring42()
    return S.make_system("ring6", r.atoms[:6] * 0.5, [1.0] * 6, (3, 3))


def _hydrogen():
    return S.make_system("H", [[0.0, 0.0, 0.0]], [1.0], (1, 0))


def _polarized_li():
    return S.make_system("Li(3,0)", [[0.0, 0.0, 0.0]], [3.0], (3, 0))


def _h17():
    return S.h_chain(17)


def _h24():
    return S.h_chain(24)


def _h20():
    return S.h_chain(20)


def _h26():
    return S.h_chain(26)


EDGE_CASES = {
    # 17 and 24 electrons: the FP64-MFMA determinant path at padded size 32
    "h17": (_h17, dict(hidden_single=[32, 32], hidden_double=[8, 8], ndets=2)),
    "h24": (_h24, dict(hidden_single=[32, 32], hidden_double=[8, 8], ndets=2)),
    # 20 and 26 electrons, 256-wide layers: the eight-wave FP16X3 layer GEMM
    # tiles at 64- and 96-row tangent units
    "h20_wide": (_h20, dict(hidden_single=[256, 256], hidden_double=[16, 16], ndets=2)),
    "h26_wide": (_h26, dict(hidden_single=[256, 256], hidden_double=[16, 16], ndets=2)),
    # the reference's hydrogen demo shape: one electron, one spin channel
    "h_atom": (_hydrogen, dict(hidden_single=[32, 32], hidden_double=[8, 8], ndets=2)),
    # fully polarized: a single non-empty spin channel (split_nonempty_channels)
    "li_polarized": (_polarized_li, dict(hidden_single=[32, 32], hidden_double=[8, 8], ndets=3)),
}

def _ring41():
    r = S.ring42()  # without its last hydrogen: 41 electrons
    return S.make_system("ring41", r.atoms[:11], r.charges[:11])


PSI_CASES = {
    # 6-electron piece of the ring, narrow Psiformer (heads_dim multiple of 32)
    "psi_small": (_ring6, dict(num_layers=2, num_heads=2, heads_dim=32, mlp_hidden=[32], ndets=2)),
    # full 42-electron ring with the 4-layer, 4-head, 256-wide Psiformer
    "ring42": (None, {}),
    # the ring less one hydrogen: 3N + 2 = 125 tangent rows in 128-row units (padding
    # rows), one MLP layer (the fused LayerNorm of the next layer written over
    # the MLP's own input)
    "psi41": (_ring41, dict(num_layers=2, mlp_hidden=[], ndets=2)),
}


def _setup(case, B):
    key = (case, B)
    if key in _cache:
        return _cache[key]
    if case in EDGE_CASES:
        mk, kw = EDGE_CASES[case]
        sysm, spec, *_ = S.workload("li")
        sysm = mk()
        spec.nspins = sysm.nspins
        spec.n_atoms = sysm.atoms.shape[0]
    elif case in PSI_CASES:
        mk, kw = PSI_CASES[case]
        sysm, spec, *_ = S.workload("ring42")
        if mk is not None:
            sysm = mk()
            spec.nspins = sysm.nspins
            spec.n_atoms = sysm.atoms.shape[0]
    else:
        name, kw = CASES[case]
        sysm, spec, *_ = S.workload(name)
    for k, v in kw.items():
        setattr(spec, k, v)
    p = P.init_params(spec, seed=11, bias_scale=0.1, env_jitter=0.2)
    x = S.init_walkers(sysm, B, seed=12)
    fx = os.path.join(GOLDEN, f"parity_{case}.npz")
    if case in CASES and os.path.exists(fx):
        # FP64 oracle references written by tests/golden/make_parity_fixtures.py
        # from the same seeds (the stored walkers are checked against them)
        with np.load(fx, allow_pickle=False) as f:
            ref = {k: f[k] for k in f.files}
        if ref["x"].shape[0] >= B:
            np.testing.assert_array_equal(ref["x"][:B], x)
            ref = {k: v[:B] for k, v in ref.items()}
            _cache[key] = (sysm, spec, p, x, ref)
            return _cache[key]
    p64 = P.cast_params(p, np.float64)
    ref = ham.local_energy(spec.as_oracle(), p64, x.astype(np.float64), sysm.atoms.astype(np.float64),
                           sysm.charges.astype(np.float64))
    A = np.asarray(owf.orbital_matrices(spec.as_oracle(), p64, x.astype(np.float64), sysm.atoms.astype(np.float64)))
    ref["kappa"] = np.linalg.cond(A).reshape(B, -1).max(axis=1)
    ref["cancel"] = cancellation(A)
    _cache[key] = (sysm, spec, p, x, ref)
    return _cache[key]


def _model(sysm, spec, p, prec="fp32", chunk=0):
    from jaqmc_amd.device import DeviceModel

    return DeviceModel(spec, sysm.atoms, sysm.charges, p, precision=prec, chunk_walkers=chunk)


@pytest.mark.parametrize("case,B", [("li", 16), ("n2_small", 16), ("n2", 32), ("h60", 8), ("n2_iso", 8),
                                    ("psi_small", 8), ("ring42", 8), ("h_atom", 16), ("li_polarized", 16),
                                    ("h17", 6), ("h24", 6)])
def test_logpsi_and_local_energy_fp32(cuda, case, B):
    import torch

    sysm, spec, p, x, ref = _setup(case, B)
    dm = _model(sysm, spec, p, chunk=5)
    xe = torch.from_numpy(x).to(cuda)
    lp, sg = dm.logpsi(xe)
    out = dm.local_energy(xe, with_grad=True)
    torch.cuda.synchronize()
    lp = lp.cpu().numpy()
    d_lp = np.abs(lp - ref["logpsi"])
    assert np.all(d_lp <= logpsi_bound(ref["logpsi"], ref["kappa"], "fp32", ref["cancel"])), (d_lp, ref["kappa"])
    np.testing.assert_array_equal(sg.cpu().numpy(), ref["sign"])
    d_lap = np.abs(out["logpsi"].cpu().numpy() - lp)
    assert np.all(d_lap <= logpsi_bound(lp, ref["kappa"], "fp32", ref["cancel"])), (d_lap, ref["kappa"])
    np.testing.assert_allclose(out["e_pot"].cpu().numpy(), ref["e_pot"], rtol=1e-6)
    rel, bnd = el_report(f"{case} B={B}", out["e_loc"].cpu().numpy(), ref["e_loc"], ref["kappa"], "fp32",
                         ref["cancel"])
    assert np.all(rel <= bnd), (rel, bnd)
    ek = out["e_kin"].cpu().numpy()
    np.testing.assert_allclose(ek + out["e_pot"].cpu().numpy(), out["e_loc"].cpu().numpy(), rtol=1e-12)
    g = out["grad"].cpu().numpy()
    gd = np.abs(g - ref["grad"]) / np.maximum(1.0, np.abs(ref["grad"]))
    gb = 5e-3 + 2 * ENTRY_ERR["fp32"] * U32 * (ref["kappa"] * ref["cancel"])[:, None]
    assert np.median(gd) <= 2e-4 and np.all(gd <= gb), (gd.max(), ref["kappa"])
    dm.close()


def test_chunking_is_bit_exact(cuda):
    import torch

    sysm, spec, p, x, _ = _setup("n2_small", 16)
    xe = torch.from_numpy(x).to(cuda)
    a = _model(sysm, spec, p, chunk=3)
    b = _model(sysm, spec, p, chunk=64)
    oa, ob = a.local_energy(xe), b.local_energy(xe)
    la, lb = a.logpsi(xe)[0], b.logpsi(xe)[0]
    torch.cuda.synchronize()
    assert torch.equal(oa["e_loc"], ob["e_loc"]) and torch.equal(la, lb)


@pytest.mark.parametrize("case", ["li", "n2_small", "n2"])
def test_metropolis_decisions_match_oracle(cuda, case):
    import torch

    sysm, spec, p, x, ref = _setup(case, 16)
    B, N = x.shape[:2]
    dm = _model(sysm, spec, p)
    rng = np.random.default_rng(21)
    steps, width = 4, 0.05
    normals = rng.standard_normal((steps, B, N, 3)).astype(np.float32)
    uniforms = rng.uniform(size=(steps, B)).astype(np.float32)
    xe = torch.from_numpy(x).to(cuda)
    logp = 2 * dm.logpsi(xe)[0]
    acc = dm.mh_steps(xe, logp, steps, width, normals=torch.from_numpy(normals).to(cuda),
                      uniforms=torch.from_numpy(uniforms).to(cuda))
    torch.cuda.synchronize()
    # oracle, step by step; the band of each decision is 1e-6 plus twice the
    # measured difference of the GPU's and the oracle's log ratio on the same
    # FP32 positions (GPU log|psi| from jq_logpsi, which the Metropolis
    # forward pass shares)
    xo = x.astype(np.float64)
    lpo = 2 * ref["logpsi"]
    p64 = P.cast_params(p, np.float64)
    at64 = sysm.atoms.astype(np.float64)
    nacc = np.zeros(B, int)
    ambiguous = np.zeros(B, bool)
    for s in range(steps):
        xp32 = (xo.astype(np.float32) + np.float32(width) * normals[s]).astype(np.float32)
        lg_x = 2 * dm.logpsi(torch.from_numpy(np.ascontiguousarray(xo.astype(np.float32))).to(cuda))[0].cpu().numpy()
        lg_p = 2 * dm.logpsi(torch.from_numpy(xp32).to(cuda))[0].cpu().numpy()
        xo, lpo_new, a, ratio, logu = ham.mh_step(spec.as_oracle(), p64, xo, lpo, at64, width,
                                                  normals[s].astype(np.float64), uniforms[s].astype(np.float64))
        delta = np.abs((lg_p - lg_x) - ratio)
        ambiguous |= np.abs(ratio - logu) < 1e-6 + 2 * delta
        lpo = lpo_new
        nacc += a
    print(f"[parity] metropolis {case}: {int(ambiguous.sum())} of {B} walkers with a decision inside the "
          f"FP32 band, max |ratio error| {delta.max():.2e}", flush=True)
    ok = ~ambiguous
    np.testing.assert_array_equal(acc.cpu().numpy()[ok], nacc[ok])
    np.testing.assert_allclose(xe.cpu().numpy()[ok], xo[ok], atol=1e-5)
    np.testing.assert_allclose(logp.cpu().numpy()[ok], lpo[ok], atol=2e-4 * np.abs(lpo).max())


def test_philox_sampler_is_keyed_by_walker(cuda):
    import torch

    sysm, spec, p, x, _ = _setup("li", 16)
    dm = _model(sysm, spec, p)
    xa = torch.from_numpy(x).to(cuda)
    xb = xa[:5].clone()
    la = 2 * dm.logpsi(xa)[0]
    lb = la[:5].clone()
    aa = dm.mh_steps(xa, la, 7, 0.1, seed=99, ctr=3)
    ab = dm.mh_steps(xb, lb, 7, 0.1, seed=99, ctr=3)
    torch.cuda.synchronize()
    assert torch.equal(xa[:5], xb) and torch.equal(aa[:5], ab)
    # a different seed gives a different trajectory
    xd = torch.from_numpy(x).to(cuda)[:5].clone()
    ld = la[:5].clone() * 0 + 2 * dm.logpsi(xd)[0]
    dm.mh_steps(xd, ld, 7, 0.1, seed=100, ctr=3)
    torch.cuda.synchronize()
    assert not torch.equal(xd, xb)


def test_reductions_match_numpy(cuda):
    import torch

    from jaqmc_amd.device import reduce_accepts, reduce_mean_var

    rng = np.random.default_rng(0)
    v = rng.standard_normal(10007) * 5 - 3
    v[[5, 77]] = np.nan
    v[100] = np.inf
    out = reduce_mean_var(torch.from_numpy(v).to(cuda)).cpu().numpy()
    ok = v[np.isfinite(v)]
    assert out[2] == ok.size
    assert out[0] == pytest.approx(ok.mean(), rel=1e-12)
    assert out[1] == pytest.approx(ok.var(), rel=1e-9)
    a = rng.integers(0, 20, 5000).astype(np.int32)
    assert reduce_accepts(torch.from_numpy(a).to(cuda)).item() == a.sum()


def test_full_size_n2_batch_properties(cuda):
    """C2 at its full size (4096 walkers): finite energies, swap antisymmetry
    on the GPU, and the batch mean consistent with the FP64 oracle on a sample."""
    import torch

    sysm, spec, B, _, _ = S.workload("n2")
    p = P.init_params(spec, seed=0)
    x = S.init_walkers(sysm, B, seed=0)
    dm = _model(sysm, spec, p)
    xe = torch.from_numpy(x).to(cuda)
    out = dm.local_energy(xe)
    lp, sg = dm.logpsi(xe)
    xs = xe.clone()
    xs[:, [0, 1]] = xs[:, [1, 0]]
    lps, sgs = dm.logpsi(xs)
    torch.cuda.synchronize()
    e = out["e_loc"].cpu().numpy()
    assert np.isfinite(e).mean() > 0.999
    np.testing.assert_allclose(lps.cpu().numpy(), lp.cpu().numpy(), rtol=1e-5, atol=1e-5)
    assert torch.all(sgs == -sg)
    ref = ham.local_energy(spec.as_oracle(), P.cast_params(p, np.float64), x[:32].astype(np.float64),
                           sysm.atoms.astype(np.float64), sysm.charges.astype(np.float64))["e_loc"]
    rel = np.abs(e[:32] - ref) / np.maximum(1, np.abs(ref))
    assert np.median(rel) < 2e-5


@pytest.mark.parametrize("prec,lp_tol,med_tol", [("bf16x6", 2e-3, 5e-5), ("bf16x3", 1e-2, 1e-3),
                                               ("fp16x3", 2e-3, 5e-5)])
def test_mixed_precision_modes(cuda, prec, lp_tol, med_tol):
    """Split 16-bit MFMA modes against the FP64 oracle on N2: bf16x6 (three
    BF16 terms, six products) stays at FP32-level error; bf16x3 (two terms) is
    ~2^-16; fp16x3 (two FP16 terms, three products) ~2^-22."""
    import torch

    sysm, spec, p, x, ref = _setup("n2", 32)
    dm = _model(sysm, spec, p, prec=prec)
    xe = torch.from_numpy(x).to(cuda)
    lp = dm.logpsi(xe)[0]
    out = dm.local_energy(xe)
    torch.cuda.synchronize()
    assert np.abs(lp.cpu().numpy() - ref["logpsi"]).max() < lp_tol  # kappa up to 10^5 on these walkers
    rel, bnd = el_report("n2 B=32", out["e_loc"].cpu().numpy(), ref["e_loc"], ref["kappa"],
                         prec if prec != "bf16x3" else "fp16x3", ref["cancel"])
    assert np.median(rel) < med_tol
    if prec != "bf16x3":
        assert np.all(rel <= bnd), (rel, bnd)


@pytest.mark.parametrize("case,B", [("ring42", 8), ("h60", 8), ("h20_wide", 4), ("h26_wide", 4)])
@pytest.mark.parametrize("prec", ["bf16x6", "fp16x3"])
def test_mixed_precision_large_configs(cuda, case, B, prec):
    """The mixed paths of the 42-electron ring (Psiformer) and H60 (FermiNet,
    60x60 FP64 determinants) against the FP64 oracle, with the kappa-scaled
    per-walker bounds of the FP32 path (entry error of the mode)."""
    import torch

    sysm, spec, p, x, ref = _setup(case, B)
    dm = _model(sysm, spec, p, prec=prec)
    xe = torch.from_numpy(x).to(cuda)
    lp = dm.logpsi(xe)[0]
    out = dm.local_energy(xe, with_grad=True)
    torch.cuda.synchronize()
    d_lp = np.abs(lp.cpu().numpy() - ref["logpsi"])
    assert np.all(d_lp <= logpsi_bound(ref["logpsi"], ref["kappa"], prec, ref["cancel"])), (d_lp, ref["kappa"])
    rel, bnd = el_report(f"{case} B={B}", out["e_loc"].cpu().numpy(), ref["e_loc"], ref["kappa"], prec,
                         ref["cancel"])
    assert np.all(rel <= bnd), (rel, bnd)


@pytest.mark.parametrize("prec", ["fp32", "fp16x3"])
def test_psiformer_padded_units_and_fused_layernorm(cuda, prec):
    """A 41-electron Psiformer (125 tangent rows in 128-row units, 256
    features, one MLP layer): in FP16X3 every LayerNorm runs in the epilogue
    of the preceding GEMM, including the one written over the MLP GEMM's own
    input, and the padding rows stay zero; log|psi| and E_L against the FP64
    oracle with the kappa-scaled bounds of the mode."""
    import torch

    sysm, spec, p, x, ref = _setup("psi41", 4)
    dm = _model(sysm, spec, p, prec=prec)
    xe = torch.from_numpy(x).to(cuda)
    lp = dm.logpsi(xe)[0]
    out = dm.local_energy(xe, with_grad=True)
    torch.cuda.synchronize()
    d_lp = np.abs(lp.cpu().numpy() - ref["logpsi"])
    assert np.all(d_lp <= logpsi_bound(ref["logpsi"], ref["kappa"], prec, ref["cancel"])), (d_lp, ref["kappa"])
    d_lap = np.abs(out["logpsi"].cpu().numpy() - ref["logpsi"])
    assert np.all(d_lap <= logpsi_bound(ref["logpsi"], ref["kappa"], prec, ref["cancel"])), (d_lap, ref["kappa"])
    rel, bnd = el_report(f"psi41 B=4", out["e_loc"].cpu().numpy(), ref["e_loc"], ref["kappa"], prec,
                         ref["cancel"])
    assert np.all(rel <= bnd), (rel, bnd)
    dm.close()


def test_empty_batch_is_a_no_op(cuda):
    import torch

    sysm, spec, p, x, _ = _setup("li", 16)
    dm = _model(sysm, spec, p)
    e = torch.empty(0, spec.n_elec, 3, dtype=torch.float32, device=cuda)
    lp, sg = dm.logpsi(e)
    out = dm.local_energy(e)
    acc = dm.mh_steps(e, torch.empty(0, device=cuda), 3, 0.1)
    torch.cuda.synchronize()
    assert lp.numel() == 0 and out["e_loc"].numel() == 0 and acc.numel() == 0


@pytest.mark.parametrize("case", ["n2_small", "h24"])
def test_coincident_electrons_are_data(cuda, case):
    """Two spin-up electrons at one point: that walker's orbital matrices are
    singular and its pair features divide by r = 0.  The walker's results are
    data (log|psi| far below its regular value or non-finite, no error) and
    every other walker of the batch is bit-identical to the regular run
    (estimator/base.py:44 treats non-finite E_L as data).  Covers the N <= 16
    and 16 < N <= 64 determinant kernels."""
    import torch

    sysm, spec, p, x, _ = _setup(case, 6)
    xs = x.copy()
    xs[2, 1] = xs[2, 0]
    dm = _model(sysm, spec, p)
    lp, _ = dm.logpsi(torch.from_numpy(xs).to(cuda))
    out = dm.local_energy(torch.from_numpy(xs).to(cuda))
    lp0, _ = dm.logpsi(torch.from_numpy(x).to(cuda))
    out0 = dm.local_energy(torch.from_numpy(x).to(cuda))
    torch.cuda.synchronize()
    keep = [0, 1, 3, 4, 5]
    assert torch.equal(lp[keep], lp0[keep])
    assert torch.equal(out["e_loc"][keep], out0["e_loc"][keep])
    v = lp[2].item()
    assert (not np.isfinite(v)) or v < lp0[2].item() - 5.0, (v, lp0[2].item())


@pytest.mark.parametrize("n", [20, 31, 33, 40, 47, 52, 57, 63])
def test_forward_logdet_every_padded_size(cuda, n):
    """log|psi| and sign on the forward (Metropolis) path for hydrogen chains
    whose electron counts reach every padded size of the register LU kernel
    (24 ... 64 in steps of 8), one determinant so log|psi| and sign are those
    of det A itself; against the FP64 oracle (jnp.linalg.slogdet,
    wavefunction/output/logdet.py:65) within the log|det| perturbation bound."""
    import torch

    sysm = S.h_chain(n)
    _, spec, *_ = S.workload("li")
    spec.nspins = sysm.nspins
    spec.n_atoms = n
    spec.hidden_single, spec.hidden_double, spec.ndets = [32, 32], [8, 8], 1
    p = P.init_params(spec, seed=5, bias_scale=0.1, env_jitter=0.2)
    x = S.init_walkers(sysm, 6, seed=6)
    p64 = P.cast_params(p, np.float64)
    at = sysm.atoms.astype(np.float64)
    sg_ref, lp_ref = owf.logpsi(spec.as_oracle(), p64, x.astype(np.float64), at, with_laplacian=False)
    A = np.asarray(owf.orbital_matrices(spec.as_oracle(), p64, x.astype(np.float64), at))
    kappa = np.linalg.cond(A).reshape(x.shape[0], -1).max(axis=1)
    dm = _model(sysm, spec, p)
    lp, sg = dm.logpsi(torch.from_numpy(x).to(cuda))
    torch.cuda.synchronize()
    d = np.abs(lp.cpu().numpy() - lp_ref.v)
    assert np.all(d <= logpsi_bound(lp_ref.v, kappa, "fp32")), (d, kappa)
    np.testing.assert_array_equal(sg.cpu().numpy(), np.asarray(sg_ref).reshape(-1))
    dm.close()


def test_psiformer_chunking_and_bf16x6(cuda):
    """Psiformer: chunked evaluation is bit-exact, and bf16x6 stays at FP32-level error."""
    import torch

    sysm, spec, p, x, ref = _setup("psi_small", 8)
    xe = torch.from_numpy(x).to(cuda)
    a = _model(sysm, spec, p, chunk=3)
    b = _model(sysm, spec, p, chunk=8)
    c = _model(sysm, spec, p, prec="bf16x6", chunk=8)
    oa, ob, oc = a.local_energy(xe), b.local_energy(xe), c.local_energy(xe)
    torch.cuda.synchronize()
    assert torch.equal(oa["e_loc"], ob["e_loc"])
    rel, bnd = el_report("psi_small B=8", oc["e_loc"].cpu().numpy(), ref["e_loc"], ref["kappa"], "bf16x6",
                         ref["cancel"])
    assert np.median(rel) < 2e-5 and np.all(rel <= bnd)
