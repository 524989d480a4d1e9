"""CPU oracle: Coulomb potential, local energy, walker reductions, Metropolis.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""

from __future__ import annotations

import numpy as np

from . import wavefunction as wf


def potential_energy(x, atoms, charges):
    """V = -sum Z/r_iI + sum_{i<j} 1/r_ij + sum_{I<J} Z_I Z_J / R_IJ.

    Restates app/molecule/hamiltonian.py:9-22 (batched over walkers).
    """
    x = np.asarray(x)
    atoms = np.asarray(atoms, x.dtype)
    Z = np.asarray(charges, x.dtype)
    B, N, _ = x.shape
    M = atoms.shape[0]
    r_ae = np.linalg.norm(x[:, :, None, :] - atoms[None, None], axis=-1)
    v = -np.sum(Z[None, None, :] / r_ae, axis=(1, 2))
    d = x[:, :, None, :] - x[:, None, :, :]
    r_ee = np.linalg.norm(d, axis=-1) + np.eye(N, dtype=x.dtype)
    v = v + np.sum(np.triu(1.0 / r_ee, 1), axis=(1, 2))
    if M > 1:
        R = np.linalg.norm(atoms[:, None] - atoms[None], axis=-1) + np.eye(M, dtype=x.dtype)
        v = v + np.sum(np.triu(Z[:, None] * Z[None, :] / R, 1))
    return v


def local_energy(model, params, x, atoms, charges):
    """Per-walker E_L = -1/2 (lap logpsi + |grad logpsi|^2) + V.

    estimator/kinetic/_common.py:61-73 + estimator/kinetic/euclidean.py:114-135
    + estimator/total_energy.py:36-60.  Returns dict of per-walker arrays.
    """
    sign, lp = wf.logpsi(model, params, x, atoms, with_laplacian=True)
    grad2 = np.sum(lp.J * lp.J, axis=1)
    e_kin = -0.5 * lp.lap - 0.5 * grad2
    e_pot = potential_energy(x, atoms, charges)
    return dict(logpsi=lp.v, sign=sign, grad=lp.J, lap=lp.lap, e_kin=e_kin, e_pot=e_pot,
                e_loc=e_kin + e_pot)


def mean_var(x):
    """nanmean and E[x^2] - E[x]^2 in float64 (estimator/base.py:27-53)."""
    x = np.asarray(x, np.float64)
    m = np.nanmean(x)
    return m, np.nanmean(x * x) - m * m


def mh_step(model, params, x, logp, atoms, width, normals, uniforms):
    """One all-electron Metropolis-Hastings move with explicit randomness.

    Restates sampler/mcmc.py:96-137 with the normal draws ``normals`` (B, N, 3)
    and uniforms ``uniforms`` (B,) supplied by the caller: x' = x + width*xi,
    logp = 2 log|psi| (sampler/base.py:174), accept iff logp' - logp > log u
    (strict).  Returns (x_new, logp_new, accept mask, ratio, log u).
    """
    x2 = x + width * normals
    _, lp2 = wf.logpsi(model, params, x2, atoms, with_laplacian=False)
    logp2 = 2.0 * lp2.v
    ratio = logp2 - logp
    logu = np.log(uniforms)
    acc = ratio > logu
    x_new = np.where(acc[:, None, None], x2, x)
    logp_new = np.where(acc, logp2, logp)
    return x_new, logp_new, acc, ratio, logu


def adapt_width(width, pmoves, counter, pmove, adapt_frequency=100, pmove_range=(0.5, 0.55)):
    """Width adaptation of MCMCSampler.step (sampler/mcmc.py:182-195)."""
    counter += 1
    t = counter % adapt_frequency
    pmoves = pmoves.copy()
    pmoves[t] = pmove
    if t == 0:
        mp = np.mean(pmoves)
        if mp > pmove_range[1]:
            width = width * 1.1
        elif mp < pmove_range[0]:
            width = width / 1.1
    return width, pmoves, counter
