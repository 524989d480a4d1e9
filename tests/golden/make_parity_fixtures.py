"""Write the FP64 oracle references of the large GPU parity cases.

Run in the build container (CPU):  python tests/golden/make_parity_fixtures.py
Each case is a workload (jaqmc_amd.systems.workload) with parameters from
params.init_params(seed=11, bias_scale=0.1, env_jitter=0.2) and walkers from
systems.init_walkers(seed=12); the file holds the walkers and the float64
oracle's log|psi|, sign, local energy (kinetic, potential, total), gradient
each walker's largest orbital-matrix 2-norm condition number and the
cancellation factor of its determinant sum (sum_k |det| / |sum_k det|).  The GPU
tests regenerate parameters and walkers from the same seeds and check them
against the stored copy before comparing.
"""

import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from jaqmc_amd import params as P  # noqa: E402
from jaqmc_amd import systems as S  # noqa: E402

CASES = {"ring42": 8, "h60": 8, "n2": 32}


def _case(name):
    sysm, spec, *_ = S.workload(name)
    p = P.init_params(spec, seed=11, bias_scale=0.1, env_jitter=0.2)
    return sysm, spec, p


def _one(args):
    from oracle import hamiltonian as ham
    from oracle import wavefunction as owf

    name, x = args
    sysm, spec, p = _case(name)
    p64 = P.cast_params(p, np.float64)
    at, ch = sysm.atoms.astype(np.float64), sysm.charges.astype(np.float64)
    ref = ham.local_energy(spec.as_oracle(), p64, x.astype(np.float64), at, ch)
    A = np.asarray(owf.orbital_matrices(spec.as_oracle(), p64, x.astype(np.float64), at))
    ref["kappa"] = np.linalg.cond(A).reshape(x.shape[0], -1).max(axis=1)
    s, l = np.linalg.slogdet(A)  # cancellation factor of the determinant sum
    m = l.max(axis=1, keepdims=True)
    ref["cancel"] = np.exp(l - m).sum(1) / np.abs((s * np.exp(l - m)).sum(1))
    return {k: np.asarray(ref[k]) for k in ("logpsi", "sign", "e_loc", "e_kin", "e_pot", "grad", "kappa", "cancel")}


def main():
    names = sys.argv[1:] or list(CASES)
    with mp.Pool(min(8, os.cpu_count() or 1)) as pool:
        for name in names:
            B = CASES[name]
            sysm, spec, p = _case(name)
            x = S.init_walkers(sysm, B, seed=12)
            parts = pool.map(_one, [(name, x[i:i + 1]) for i in range(B)])
            out = {k: np.concatenate([q[k] for q in parts]) for k in parts[0]}
            out["x"] = x
            np.savez_compressed(os.path.join(HERE, f"parity_{name}.npz"), **out)
            print(name, B, "done", flush=True)


if __name__ == "__main__":
    main()
