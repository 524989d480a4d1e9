"""Reference parameter trees and checkpoint files -> the device pack (CPU only).

The tree is built to the reference's Flax naming (module attribute names from
``setup``, ``Dense_i`` / ``DenseGeneral_i`` / ``PsiformerLayer_l`` numbering
in creation order, app/molecule/wavefunction/{ferminet,psiformer}.py,
wavefunction/backbone/*.py, wavefunction/output/*.py, wavefunction/jastrow.py)
and the checkpoint keys follow utils/checkpoint.py:165-217
(``params/params/backbone_layer/Dense_0/kernel``).  No reference checkpoint
file exists in the reference tree, so the naming is checked on files written
here to that layout.
"""

import numpy as np
import pytest

from jaqmc_amd import checkpoint as C
from jaqmc_amd import params as P
from jaqmc_amd import systems as S

SPECS = [
    ("n2", {}),
    ("li", {}),
    ("ring42", dict(num_layers=2, num_heads=2, heads_dim=32, mlp_hidden=[32])),
    ("ring42", dict(num_layers=1, num_heads=2, heads_dim=32, mlp_hidden=[48], layer_norm_mode="post",
                    with_bias=False, input_bias=False, bias_orbitals=True)),
    ("ring42", dict(num_layers=1, num_heads=1, heads_dim=32, mlp_hidden=[32], layer_norm_mode="null",
                    orbitals_spin_split=False, jastrow=False)),
    ("n2", dict(hidden_single=[32, 32], hidden_double=[8, 8], orbitals_spin_split=False)),
]


def _spec(name, kw):
    _, spec, *_ = S.workload(name)
    for k, v in kw.items():
        setattr(spec, k, v)
    return spec


@pytest.mark.parametrize("name,kw", SPECS)
def test_flax_tree_roundtrip(name, kw):
    spec = _spec(name, kw)
    p = P.init_params(spec, seed=2, bias_scale=0.1, env_jitter=0.1)
    tree = C.to_flax_tree(spec, p)
    q = C.from_flax_tree(spec, tree)
    np.testing.assert_array_equal(P.pack(spec, q), P.pack(spec, p))
    # the bare parameter dict (no outer 'params') is accepted too
    np.testing.assert_array_equal(P.pack(spec, C.from_flax_tree(spec, tree["params"])), P.pack(spec, p))


def test_flax_naming_of_the_trees():
    spec = _spec("n2", {})
    t = C.to_flax_tree(spec, P.init_params(spec))["params"]
    assert t["backbone_layer"]["Dense_0"]["kernel"].shape == (32, 256)   # layer-0 one-electron
    assert t["backbone_layer"]["Dense_1"]["kernel"].shape == (4, 32)     # layer-0 two-electron
    assert t["backbone_layer"]["Dense_6"]["kernel"].shape == (832, 256)  # last layer, single only
    assert "Dense_7" not in t["backbone_layer"]
    dg = t["orbital_layer"]["SplitChannelDense_0"]["DenseGeneral_1"]["kernel"]
    assert dg.shape == (256, 16, 14)
    assert t["envelope_layer"]["_env_down"]["sigma"].shape == (14, 2, 16)
    spec = _spec("ring42", {})
    t = C.to_flax_tree(spec, P.init_params(spec))["params"]
    mha = t["backbone_layer"]["PsiformerLayer_3"]["MultiHeadDotProductAttention_0"]
    assert mha["query"]["kernel"].shape == (256, 4, 64) and mha["query"]["bias"].shape == (4, 64)
    assert mha["out"]["kernel"].shape == (4, 64, 256)
    assert t["backbone_layer"]["PsiformerLayer_0"]["LayerNorm_1"]["scale"].shape == (256,)
    assert t["jastrow_layer"]["alpha_par"].shape == (1,)


def test_checkpoint_file_roundtrip(tmp_path):
    spec = _spec("n2", dict(hidden_single=[32, 32], hidden_double=[8, 8], ndets=4))
    sysm = S.n2_molecule()
    p = P.init_params(spec, seed=9, bias_scale=0.1)
    x = S.init_walkers(sysm, 8, seed=1)
    flat = C.flat_from_tree(C.to_flax_tree(spec, p), "params")
    flat["batched_data/data/electrons"] = x
    flat["batched_data/data/atoms"] = sysm.atoms
    flat["sampler_state/stddev"] = np.float32(0.02)
    assert "params/params/backbone_layer/Dense_0/kernel" in flat
    path = tmp_path / "train_ckpt_000010.npz"
    np.savez_compressed(path, allow_pickle=False, step=10, **flat)
    q, elec = C.load_checkpoint(path, spec)
    np.testing.assert_array_equal(P.pack(spec, q), P.pack(spec, p))
    np.testing.assert_array_equal(elec, x)


def test_orbital_bias_mismatch_is_an_error():
    spec = _spec("ring42", dict(num_layers=1, num_heads=1, heads_dim=32, mlp_hidden=[32], bias_orbitals=True))
    tree = C.to_flax_tree(spec, P.init_params(spec))
    spec.bias_orbitals = False
    with pytest.raises(ValueError):
        C.from_flax_tree(spec, tree)
    spec.use_last_layer = True
    with pytest.raises(ValueError):
        C.from_flax_tree(spec, tree)
