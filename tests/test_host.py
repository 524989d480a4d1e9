"""Host logic and the C-ABI library's exports (CPU only)."""

import ctypes
import re

import numpy as np
import pytest

from jaqmc_amd import _lib
from jaqmc_amd import params as P
from jaqmc_amd import systems as S


def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    import pathlib

    root = pathlib.Path(__file__).resolve().parent.parent
    names = set()
    for h in (root / "include").glob("*.h"):
        names |= set(re.findall(r"\b(jq_[a-z_]+)\s*\(", h.read_text()))
    assert "jq_local_energy" in names and "jq_mh_steps" in names
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_typed_binding_loads():
    lib = _lib.load()
    assert lib.jq_last_error() is not None


@pytest.mark.parametrize("name", ["li", "n2", "ring42", "h60"])
def test_param_pack_roundtrip(name):
    sysm, spec, *_ = S.workload(name)
    p = P.init_params(spec, seed=3, bias_scale=0.1)
    flat = P.pack(spec, p)
    q = P.unpack(spec, flat, p)
    np.testing.assert_array_equal(P.pack(spec, q), flat)
    assert flat.dtype == np.float32


def test_ferminet_layer_widths():
    # FermiLayers.aggregate_features: in_l = f1 (1 + nch) + f2 nch
    _, spec, *_ = S.workload("n2")
    ins, ins2 = spec.ferminet_dims()
    assert ins == [4 * 2 * 3 + 4 * 2, 256 * 3 + 32 * 2, 832, 832]
    assert ins2 == [4, 32, 32]
    _, li, *_ = S.workload("li")
    assert li.ferminet_dims()[0] == [4 * 3 + 8, 64 * 3 + 16 * 2]


def test_walker_init_order_and_geometry():
    s = S.n2_molecule()
    assert s.spins_per_atom == [(4, 3), (3, 4)] or sum(a for a, _ in s.spins_per_atom) == 7
    x = S.init_walkers(s, 64, seed=0, width=1e-3)
    # electrons ordered all-up (atom 0, atom 1) then all-down
    ups = [0] * s.spins_per_atom[0][0] + [1] * s.spins_per_atom[1][0]
    for i, a in enumerate(ups):
        np.testing.assert_allclose(x[:, i].mean(0), s.atoms[a], atol=1e-2)
    r = S.ring42()
    assert r.atoms.shape == (12, 3) and r.nspins == (21, 21)
    np.testing.assert_allclose(np.linalg.norm(r.atoms[:6], axis=1), 2.64, rtol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(r.atoms[6:], axis=1), 4.68, rtol=1e-6)
    h = S.h_chain()
    assert h.nspins == (30, 30) and h.atoms[1, 0] == pytest.approx(1.8)


def test_random_chain_is_neutral_and_spin_balanced():
    rng = np.random.default_rng(0)
    for n in range(2, 43):
        s = S.random_chain(rng, n)
        assert int(s.charges.sum()) == n
        assert s.nspins == ((n + 1) // 2, n // 2)
        assert s.charges.max() <= 8 and s.charges.min() >= 1
        if len(s.atoms) > 1:
            d = np.linalg.norm(np.diff(s.atoms, axis=0), axis=1)
            assert d.min() >= 1.5 - 1e-6 and d.max() <= 3.0 + 1e-6


def test_device_model_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from jaqmc_amd.device import DeviceModel

    sysm, spec, *_ = S.workload("li")
    with pytest.raises(_lib.JqError):
        DeviceModel(spec, sysm.atoms, sysm.charges, P.init_params(spec))
