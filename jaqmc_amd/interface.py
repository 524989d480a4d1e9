"""Host-side mirror of the reference's component interfaces for the hot path.

Each class keeps the reference's names, argument meaning and return shapes so
that code written against JaQMC's wavefunction / sampler / estimator
protocols drives the HIP kernels unchanged:

* :class:`FermiNetWavefunction`, :class:`PsiformerWavefunction` --
  ``init_params(data, rngs)``, ``evaluate(params, data)``, ``logpsi``,
  ``phase_logpsi`` (wavefunction/base.py:45-53,
  app/molecule/wavefunction/ferminet.py:21-127, .../psiformer.py:38-197);
* :class:`MCMCSampler` -- ``init(data, rngs)``, ``step(batch_log_prob, data,
  state, rngs)`` with width adaptation (sampler/mcmc.py:57-197);
* :class:`EuclideanKinetic`, :func:`potential_energy`, :class:`TotalEnergy`,
  :class:`EstimatorPipeline`, :func:`mean_reduce` (estimator/base.py:27-427,
  estimator/kinetic/euclidean.py:29-135, estimator/total_energy.py:16-70,
  app/molecule/hamiltonian.py:9-22);
* :func:`evaluation_step` -- EvaluationWorkStage.compute_step
  (workflow/stage/evaluation.py:179-199);
* :class:`LossAndGrad` -- the VMC loss and parameter gradient
  (estimator/loss_grad.py:24-128): per-walker value-and-grad, clipping,
  and the baseline-subtracted mean, here contracted on the device.

Differences forced by the device boundary: the wavefunction methods are
*batched* (the reference's are single-walker and run under ``jax.vmap``;
here the batch axis is explicit), parameters are the numpy tree of
``jaqmc_amd.params`` uploaded once per (params object, precision), and random
keys are integers (seed, counter) feeding the kernels' Philox streams.
"""

from __future__ import annotations

import dataclasses
from typing import Any, Dict, NamedTuple, Optional

import numpy as np
import torch

from .device import DeviceModel, loss_clip, reduce_accepts, reduce_mean_var
from .params import ModelSpec, init_params, unpack


@dataclasses.dataclass
class MoleculeData:
    """Walker batch (app/molecule/data.py:13-29): electrons (B, N, 3) on the
    GPU (float32); atoms (M, 3) and charges (M,) shared."""

    electrons: torch.Tensor
    atoms: np.ndarray
    charges: np.ndarray

    def merge(self, other: Dict[str, Any]):
        return dataclasses.replace(self, **other)


class _MoleculeWavefunction:
    """Shared implementation of the molecule wavefunction protocol."""

    spec: ModelSpec

    def __init__(self, precision: str = "fp32", chunk_walkers: int = 0, with_gradients: bool = False):
        self.precision = precision
        self.chunk_walkers = chunk_walkers
        self.with_gradients = with_gradients
        self._models: Dict[int, DeviceModel] = {}

    def init_params(self, data: MoleculeData, rngs: int = 0, bias_scale: float = 0.0):
        """Seeded random parameters (Glorot-normal kernels, envelope pi = sigma = 1)."""
        return init_params(self.spec, seed=int(rngs), bias_scale=bias_scale)

    def model(self, params, data: MoleculeData) -> DeviceModel:
        key = id(params)
        m = self._models.get(key)
        if m is None:
            m = DeviceModel(self.spec, data.atoms, data.charges, params, precision=self.precision,
                            chunk_walkers=self.chunk_walkers, with_gradients=self.with_gradients)
            self._models = {key: m}  # one resident parameter set at a time
        return m

    def evaluate(self, params, data: MoleculeData):
        lp, sg = self.model(params, data).logpsi(data.electrons)
        return {"logpsi": lp, "sign_logpsi": sg}

    def logpsi(self, params, data: MoleculeData):
        return self.evaluate(params, data)["logpsi"]

    def phase_logpsi(self, params, data: MoleculeData):
        out = self.evaluate(params, data)
        return out["sign_logpsi"], out["logpsi"]

    def orbitals(self, params, data: MoleculeData):
        """(B, K, N, N) orbital matrices, envelope applied
        (app/molecule/wavefunction/ferminet.py:128-138)."""
        return self.model(params, data).orbitals(data.electrons)


class FermiNetWavefunction(_MoleculeWavefunction):
    """FermiNet (app/molecule/wavefunction/ferminet.py:21-74); full determinants,
    abs_isotropic envelope and spin-split orbitals by default."""

    def __init__(self, nspins, n_atoms, ndets=16, hidden_dims_single=(256,) * 4, hidden_dims_double=(32,) * 4,
                 envelope="abs_isotropic", precision="fp32", chunk_walkers=0, orbitals_spin_split=True,
                 with_gradients=False):
        super().__init__(precision, chunk_walkers, with_gradients)
        self.spec = ModelSpec("ferminet", tuple(nspins), n_atoms, ndets=ndets,
                              hidden_single=list(hidden_dims_single), hidden_double=list(hidden_dims_double),
                              envelope=envelope, orbitals_spin_split=orbitals_spin_split)


class PsiformerWavefunction(_MoleculeWavefunction):
    """Psiformer (app/molecule/wavefunction/psiformer.py:38-134): pre-LN layers,
    SimpleEE Jastrow, rescaled features."""

    def __init__(self, nspins, n_atoms, ndets=16, num_layers=4, num_heads=4, heads_dim=64, mlp_hidden_dims=(256,),
                 jastrow=True, envelope="abs_isotropic", precision="fp32", chunk_walkers=0, layer_norm_mode="pre",
                 with_bias=True, input_bias=True, bias_orbitals=False, rescale=True, orbitals_spin_split=True,
                 with_gradients=False):
        super().__init__(precision, chunk_walkers, with_gradients)
        self.spec = ModelSpec("psiformer", tuple(nspins), n_atoms, ndets=ndets, num_layers=num_layers,
                              num_heads=num_heads, heads_dim=heads_dim, mlp_hidden=list(mlp_hidden_dims),
                              jastrow=jastrow, envelope=envelope, layer_norm_mode=layer_norm_mode,
                              with_bias=with_bias, input_bias=input_bias, bias_orbitals=bias_orbitals,
                              rescale=rescale, orbitals_spin_split=orbitals_spin_split)


class BatchLogProb:
    """``lambda x: 2 * batch_log_amplitude(params, data.merge(x))``
    (sampler/base.py:174) as an object, so the sampler can hand the whole
    Metropolis loop to the device model."""

    def __init__(self, wf: _MoleculeWavefunction, params, data: MoleculeData):
        self.wf, self.params, self.data = wf, params, data

    def __call__(self, electrons: torch.Tensor):
        return 2.0 * self.wf.logpsi(self.params, self.data.merge({"electrons": electrons}))

    def model(self):
        return self.wf.model(self.params, self.data)


class MCMCState(NamedTuple):
    """sampler/mcmc.py:39-50."""

    stddev: float
    pmoves: np.ndarray
    counter: int


class MCMCSampler:
    """All-electron Metropolis-Hastings with width adaptation
    (sampler/mcmc.py:57-197).  ``rngs`` is an integer seed; step ``t`` of the
    loop uses Philox counter ``counter * steps + t``."""

    def __init__(self, steps=10, initial_width=0.1, adapt_frequency=100, pmove_range=(0.5, 0.55)):
        self.steps = steps
        self.initial_width = initial_width
        self.adapt_frequency = adapt_frequency
        self.pmove_range = pmove_range

    def init(self, data, rngs=0):
        return MCMCState(float(self.initial_width), np.zeros(self.adapt_frequency), 0)

    def step(self, batch_log_prob: BatchLogProb, data: MoleculeData, state: MCMCState, rngs: int = 0):
        if self.steps == 0:
            return data, {"pmove": 0.0}, state
        m = batch_log_prob.model()
        x = data.electrons.clone()
        logp = batch_log_prob(x)  # recomputed at the start, as the reference does
        acc = torch.zeros(x.shape[0], dtype=torch.int32, device=x.device)
        m.mh_steps(x, logp, self.steps, state.stddev, seed=int(rngs), ctr=state.counter * self.steps, n_accept=acc)
        pmove = float(reduce_accepts(acc).item()) / (self.steps * x.shape[0])
        stddev, pmoves, counter = state
        counter += 1
        t = counter % self.adapt_frequency
        pmoves = pmoves.copy()
        pmoves[t] = pmove
        if t == 0:
            mp = float(np.mean(pmoves))
            if mp > self.pmove_range[1]:
                stddev = stddev * 1.1
            elif mp < self.pmove_range[0]:
                stddev = stddev / 1.1
        return data.merge({"electrons": x}), {"pmove": pmove}, MCMCState(stddev, pmoves, counter)


def mean_reduce(walker_stats: Dict[str, torch.Tensor], include_variance=True):
    """nanmean over walkers (+ ``{key}_var`` = E[x^2] - E[x]^2), FP64 on the GPU
    (estimator/base.py:27-53)."""
    out = {}
    for k, v in walker_stats.items():
        mv = reduce_mean_var(v.double().contiguous()).cpu().numpy()
        out[k] = float(mv[0])
        if include_variance:
            out[f"{k}_var"] = float(mv[1])
    return out


class EuclideanKinetic:
    """Kinetic energy by the forward Laplacian (estimator/kinetic/euclidean.py:114-135).
    ``evaluate_batch_walkers`` is overridden directly, as the reference's
    batched estimators may."""

    def __init__(self, wf: _MoleculeWavefunction):
        self.wf = wf
        self._last = None

    def evaluate_batch_walkers(self, params, batched_data: MoleculeData, prev_walker_stats=None, state=None,
                               rngs=None):
        out = self.wf.model(params, batched_data).local_energy(batched_data.electrons)
        self._last = out
        return {"energy:kinetic": out["e_kin"], "energy:potential": out["e_pot"]}, state


def potential_energy(params, data: MoleculeData, prev_walker_stats=None, state=None, rngs=None):
    """Coulomb energy (app/molecule/hamiltonian.py:9-22); computed by the
    local-energy kernel, so when the kinetic estimator ran first its value is
    reused (``prev_walker_stats``)."""
    if prev_walker_stats and "energy:potential" in prev_walker_stats:
        return {"energy:potential": prev_walker_stats["energy:potential"]}, state
    raise ValueError("potential_energy is produced together with the kinetic term by EuclideanKinetic")


class TotalEnergy:
    """Sum of the ``energy:*`` keys (estimator/total_energy.py:36-60)."""

    def evaluate_batch_walkers(self, params, batched_data, prev_walker_stats, state=None, rngs=None):
        tot = sum(v for k, v in prev_walker_stats.items() if k.startswith("energy:"))
        return {"total_energy": tot}, state


class EstimatorPipeline:
    """Ordered estimators, each seeing the previous ones' walker stats
    (estimator/base.py:287-362); reduce = mean_reduce."""

    def __init__(self, estimators):
        self.estimators = list(estimators)

    def evaluate(self, params, batched_data, state=None, rngs=None):
        walker_stats: Dict[str, torch.Tensor] = {}
        for est in self.estimators:
            if callable(est) and not hasattr(est, "evaluate_batch_walkers"):
                s, _ = est(params, batched_data, walker_stats, None, rngs)
            else:
                s, _ = est.evaluate_batch_walkers(params, batched_data, walker_stats, None, rngs)
            walker_stats.update(s)
        return walker_stats, mean_reduce(walker_stats)


class LossAndGrad:
    """VMC loss and parameter gradient (estimator/loss_grad.py:24-128).

    ``evaluate_batch_walkers`` keeps the per-walker loss (``loss_key`` of the
    previous estimators) and log|psi|; ``reduce`` clips the losses on the
    device (jq_loss_clip: "mad" / "iqr" / "none", FP64) and forms the walker
    weights 2 (c_w - mean c) / n; ``finalize_stats`` contracts the weights with
    the per-walker parameter gradients in one device pass (jq_logpsi_vjp),
    which equals the reference's 2 (mean(grad c) - mean(c) mean(grad)).  The
    gradient is returned flat (pack order) and as the parameter tree.
    """

    def __init__(self, wf: _MoleculeWavefunction, loss_key="total_energy", clip_method="mad", clip_scale=5.0):
        if not wf.with_gradients:
            raise ValueError("LossAndGrad needs a wavefunction created with with_gradients=True")
        self.wf, self.loss_key = wf, loss_key
        self.clip_method, self.clip_scale = clip_method, clip_scale
        self._ctx = None

    def evaluate_batch_walkers(self, params, batched_data: MoleculeData, prev_walker_stats, state=None, rngs=None):
        loss = prev_walker_stats[self.loss_key].double().contiguous()
        self._ctx = (params, batched_data)
        return {"logpsi": self.wf.logpsi(params, batched_data), "loss": loss}, state

    def reduce(self, walker_stats):
        clipped, w, st = loss_clip(walker_stats["loss"], self.clip_method, self.clip_scale)
        return {"weights": w, "clipped_loss": clipped, "stats": st}

    def finalize_stats(self, batch_stats, state=None):
        params, data = self._ctx
        m = self.wf.model(params, data)
        g = m.logpsi_vjp(data.electrons, batch_stats["weights"]).cpu().numpy()
        st = batch_stats["stats"].cpu().numpy()
        return {"loss": float(st[0]), "clipped_loss": float(st[1]), "grads_flat": g,
                "grads": unpack(self.wf.spec, g, params)}


def evaluation_step(wf, sampler: MCMCSampler, pipeline: EstimatorPipeline, params, data: MoleculeData,
                    sampler_state: MCMCState, rngs: int = 0):
    """One EvaluationWorkStage.compute_step (workflow/stage/evaluation.py:179-199):
    sample, then evaluate the estimators at the new walkers."""
    blp = BatchLogProb(wf, params, data)
    data, sstats, sampler_state = sampler.step(blp, data, sampler_state, rngs)
    walker_stats, stats = pipeline.evaluate(params, data)
    stats.update(sstats)
    return data, stats, sampler_state, walker_stats
