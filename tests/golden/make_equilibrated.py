"""Write oracle-equilibrated walkers with their FP64 local energies.

Run in the build container (CPU):
    python tests/golden/make_equilibrated.py [workload ...]
For each benchmark workload (jaqmc_amd.systems.workload) with the benchmark's
parameters (params.init_params(spec, seed=0)): the first `n` walkers of
systems.init_walkers(seed=0) are moved by the float64 oracle's Metropolis
step (all-electron Gaussian proposals, the reference's width adaptation every
5 steps toward acceptance 0.5-0.55, seeded numpy randomness) for `steps`
steps, then the oracle's float64 local energy is evaluated at the final
positions.  The positions are stored as float32 (what the GPU path reads) and
the energies are evaluated at those float32 positions.  One further
Metropolis move is prepared for the accept/reject comparison: seeded normals
and uniforms, the proposal x' = x + width * xi in float32 with the workload's
own proposal width, and the oracle's float64 log ratio 2 log|psi(x')| -
2 log|psi(x)| and decision (ratio > log u).  bench.py reports the GPU paths'
paired errors and decision agreement on these walkers.
"""

import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from jaqmc_amd import params as P  # noqa: E402
from jaqmc_amd import systems as S  # noqa: E402

# workload: (walkers, equilibration steps, starting width)
PLAN = {"li": (256, 300, 0.3), "n2": (256, 300, 0.1), "ring42": (256, 200, 0.08), "h60": (128, 200, 0.15)}


def _setup(name):
    sysm, spec, *_ = S.workload(name)
    p = P.cast_params(P.init_params(spec, seed=0), np.float64)
    return sysm, spec, p


def _equilibrate(args):
    from oracle import hamiltonian as ham
    from oracle import wavefunction as owf

    name, x, seed = args
    sysm, spec, p = _setup(name)
    n, steps, width = PLAN[name]
    m = spec.as_oracle()
    at = sysm.atoms.astype(np.float64)
    rng = np.random.default_rng(seed)
    x = x.astype(np.float64)
    logp = 2 * owf.logpsi(m, p, x, at, with_laplacian=False)[1].v
    pmoves = np.zeros(5)
    counter = 0
    for _ in range(steps):
        x, logp, acc, _, _ = ham.mh_step(m, p, x, logp, at, width, rng.standard_normal(x.shape),
                                         rng.uniform(size=x.shape[0]))
        width, pmoves, counter = ham.adapt_width(width, pmoves, counter, float(acc.mean()), adapt_frequency=5)
    return x.astype(np.float32), width


def _energy(args):
    from oracle import hamiltonian as ham
    from oracle import wavefunction as owf

    name, x, xi, u = args
    sysm, spec, p = _setup(name)
    _, _, _, width = S.workload(name)[1:]
    at, ch = sysm.atoms.astype(np.float64), sysm.charges.astype(np.float64)
    ref = ham.local_energy(spec.as_oracle(), p, x.astype(np.float64), at, ch)
    A = np.asarray(owf.orbital_matrices(spec.as_oracle(), p, x.astype(np.float64), at))
    ref["kappa"] = np.linalg.cond(A).reshape(x.shape[0], -1).max(axis=1)
    xp = (x + np.float32(width) * xi).astype(np.float32)
    lpp = owf.logpsi(spec.as_oracle(), p, xp.astype(np.float64), at, with_laplacian=False)[1].v
    ref["mh_ratio"] = 2.0 * lpp - 2.0 * ref["logpsi"]
    ref["mh_accept"] = ref["mh_ratio"] > np.log(u.astype(np.float64))
    return {k: np.asarray(ref[k]) for k in ("logpsi", "e_loc", "e_kin", "e_pot", "kappa", "mh_ratio", "mh_accept")}


def main():
    names = sys.argv[1:] or list(PLAN)
    nproc = min(8, os.cpu_count() or 1)
    with mp.Pool(nproc) as pool:
        for name in names:
            n, steps, _ = PLAN[name]
            sysm, spec, _ = _setup(name)
            x0 = S.init_walkers(sysm, n, seed=0)
            per = (n + nproc - 1) // nproc
            chunks = [(name, x0[i:i + per], 1000 + i) for i in range(0, n, per)]
            res = pool.map(_equilibrate, chunks)
            x = np.concatenate([r[0] for r in res])
            rng = np.random.default_rng(77)
            xi = rng.standard_normal(x.shape).astype(np.float32)
            u = rng.uniform(size=n).astype(np.float32)
            parts = pool.map(_energy, [(name, x[i:i + 1], xi[i:i + 1], u[i:i + 1]) for i in range(n)])
            out = {k: np.concatenate([q[k] for q in parts]) for k in parts[0]}
            out["x"] = x
            out["mh_normals"] = xi
            out["mh_uniforms"] = u
            out["steps"] = np.int64(steps)
            out["final_width"] = np.array([r[1] for r in res])
            np.savez_compressed(os.path.join(HERE, f"equilibrated_{name}.npz"), **out)
            print(name, n, "done; mean E_L", float(out["e_loc"].mean()), flush=True)


if __name__ == "__main__":
    main()
