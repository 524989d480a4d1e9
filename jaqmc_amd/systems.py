"""Molecular systems of the benchmark workloads and Gaussian walker initialisation.

Walker initialisation restates utils/atomic/initialization.py:124-163: each
electron starts on an assigned nucleus plus ``width * N(0, 1)`` noise, with the
electron order all spin-up electrons (atom 0, atom 1, ...) then all spin-down
electrons (atom 0, atom 1, ...).  The per-atom spins follow ``distribute_spins``
(initialization.py:39-122) without per-atom hints: every atom starts from its
chemical default (n_alpha, n_beta) = ((Z + u) / 2, (Z - u) / 2) with u the
ground-state unpaired-electron count of the element (utils/atomic/atom.py:57-79),
then single electrons move between the channels of randomly chosen atoms until
the target (n_up, n_dn) is met.  The random choice uses a seeded numpy
generator (JAX's threefry stream is not reproduced); it only affects where
walkers start.
"""

from __future__ import annotations

import dataclasses
from typing import List, Tuple

import numpy as np

from .params import ModelSpec


@dataclasses.dataclass
class System:
    name: str
    atoms: np.ndarray  # (M, 3) float32
    charges: np.ndarray  # (M,) float32
    nspins: Tuple[int, int]
    spins_per_atom: List[Tuple[int, int]]


# Ground-state unpaired electrons of the neutral atoms H .. Ne (Hund's rule)
UNPAIRED = {1: 1, 2: 0, 3: 1, 4: 0, 5: 1, 6: 2, 7: 3, 8: 2, 9: 1, 10: 0}


def _spins_per_atom(charges, nspins, seed=0):
    """distribute_spins (initialization.py:39-122) for neutral atoms without
    hints: chemical defaults, then random single-electron moves between the
    channels of candidate atoms until the target (n_up, n_dn) is met."""
    out = []
    for z in charges:
        z = int(round(float(z)))
        u = UNPAIRED[z]
        out.append(((z + u) // 2, (z - u) // 2))
    if sum(a + b for a, b in out) != sum(nspins):
        raise ValueError(f"{sum(a + b for a, b in out)} electrons on the atoms, nspins {nspins}")
    rng = np.random.default_rng(seed)
    cu = sum(a for a, _ in out)
    while cu != nspins[0]:
        up = nspins[0] > cu  # move an electron from beta to alpha
        cand = [i for i, (a, b) in enumerate(out) if (b if up else a) > 0]
        if not cand:
            raise ValueError(f"cannot reach spins {nspins} from {out}")
        i = cand[int(rng.integers(0, len(cand)))]
        a, b = out[i]
        out[i] = (a + 1, b - 1) if up else (a - 1, b + 1)
        cu += 1 if up else -1
    return out


def make_system(name, atoms, charges, nspins=None):
    atoms = np.asarray(atoms, np.float32).reshape(-1, 3)
    charges = np.asarray(charges, np.float32)
    ne = int(round(float(charges.sum())))
    if nspins is None:
        nspins = ((ne + 1) // 2, ne // 2)
    return System(name, atoms, charges, tuple(nspins), _spins_per_atom(charges, nspins))


def li_atom():
    return make_system("Li", [[0.0, 0.0, 0.0]], [3.0], (2, 1))


def n2_molecule():
    return make_system("N2", [[0, 0, 1.0375], [0, 0, -1.0375]], [7.0, 7.0], (7, 7))


def ring42():
    """6 C on a radius-2.64 hexagon, 6 H radially outward at 4.68 bohr, z = 0."""
    ang = np.arange(6) * (np.pi / 3)
    c = np.stack([2.64 * np.cos(ang), 2.64 * np.sin(ang), np.zeros(6)], 1)
    h = np.stack([4.68 * np.cos(ang), 4.68 * np.sin(ang), np.zeros(6)], 1)
    return make_system("ring42", np.concatenate([c, h]), [6.0] * 6 + [1.0] * 6, (21, 21))


def h_chain(n=60, spacing=1.8):
    x = np.arange(n) * spacing
    return make_system(f"H{n}", np.stack([x, np.zeros(n), np.zeros(n)], 1), [1.0] * n, (n // 2, n - n // 2))


def random_chain(rng, n_elec):
    """Seeded random neutral, spin-balanced chain with total charge n_elec:
    Z in 1..8, consecutive nuclei 1.5-3 bohr apart along a random-walk chain."""
    zs = []
    left = n_elec
    while left > 0:
        z = int(rng.integers(1, min(8, left) + 1))
        zs.append(z)
        left -= z
    pos = [np.zeros(3)]
    for _ in range(len(zs) - 1):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        pos.append(pos[-1] + d * rng.uniform(1.5, 3.0))
    return make_system(f"chain{n_elec}", np.array(pos), zs, ((n_elec + 1) // 2, n_elec // 2))


def init_walkers(system: System, batch: int, seed: int = 0, width: float = 1.0, dtype=np.float32):
    """Gaussian walkers around the nuclei, ordered up-then-down by atom
    (utils/atomic/initialization.py:150-163)."""
    centers = []
    for s in range(2):
        for j, sp in enumerate(system.spins_per_atom):
            centers += [system.atoms[j]] * sp[s]
    centers = np.asarray(centers, np.float64)
    rng = np.random.default_rng(seed)
    x = centers[None] + width * rng.standard_normal((batch,) + centers.shape)
    return x.astype(dtype)


# --- benchmark workloads -----------------------------------------------------
def workload(name: str):
    """(system, ModelSpec, batch, mh_steps, width) for a named workload."""
    if name == "li":
        s = li_atom()
        spec = ModelSpec("ferminet", s.nspins, 1, ndets=4, hidden_single=[64, 64], hidden_double=[16, 16],
                         envelope="isotropic")
        return s, spec, 1024, 20, 0.05
    if name == "n2":
        s = n2_molecule()
        return s, ModelSpec("ferminet", s.nspins, 2, ndets=16), 4096, 100, 0.02
    if name == "ring42":
        s = ring42()
        return s, ModelSpec("psiformer", s.nspins, 12, ndets=16, num_layers=4, num_heads=4, heads_dim=64), 2048, 50, 0.02
    if name == "h60":
        s = h_chain(60)
        return s, ModelSpec("ferminet", s.nspins, 60, ndets=8), 1024, 30, 0.02
    raise KeyError(name)
