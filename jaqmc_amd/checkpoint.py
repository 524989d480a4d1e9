"""Reference parameter trees and checkpoints -> the device parameter pack.

The reference keeps wavefunction parameters as a Flax variable tree and saves
training state with ``np.savez_compressed(..., allow_pickle=False)``, one array
per leaf under its slash-joined key path (utils/checkpoint.py:165-176,
:204-217); a VMC checkpoint ``train_ckpt_%06d.npz`` holds ``params/...``,
``batched_data/...`` and ``sampler_state/...`` (workflow/evaluation.py:93-116).

Module names in the tree are Flax's: attributes assigned in ``setup`` keep
their attribute name, and submodules created inside a compact method are
numbered per type in creation order (``Dense_0``, ``Dense_1``, ...).  For the
two ansatzes of the hot path this gives:

FermiNet (app/molecule/wavefunction/ferminet.py:51-74)
    backbone_layer/Dense_{2l}          one-electron layer l  (wavefunction/backbone/ferminet.py:36-58:
    backbone_layer/Dense_{2l+1}        two-electron layer l   single then double per layer, the last
                                                              layer single only)
    orbital_layer/SplitChannelDense_0/DenseGeneral_{c}/kernel (d, K, N)  (output/orbital.py:23-30,73-76)
    orbital_layer/DenseGeneral_0/kernel                        (orbitals_spin_split=False)
    envelope_layer/_env_up|_env_down/{pi, sigma} (N, M, K)      (output/envelope.py:82-96,130-135)
    envelope_layer/_env/{pi, sigma}                             (shared)

Psiformer (app/molecule/wavefunction/psiformer.py:93-134, backbone/psiformer.py:64-187)
    backbone_layer/Dense_0                             input projection
    backbone_layer/PsiformerLayer_l/LayerNorm_{0,1}    pre / post LayerNorms (none for "null")
    backbone_layer/PsiformerLayer_l/MultiHeadDotProductAttention_0/{query,key,value}
                                                       kernel (D, H, dh), bias (H, dh)
    .../MultiHeadDotProductAttention_0/out             kernel (H, dh, D), bias (D,)
    backbone_layer/PsiformerLayer_l/Dense_{j}          MLP layers in order
    orbital_layer/..., envelope_layer/...              as above (optional orbital bias (K, N))
    jastrow_layer/{alpha_par, alpha_anti}              (1,) each  (wavefunction/jastrow.py:60-63)

Checkpoint files are read with ``numpy.load(..., allow_pickle=False)`` only.
"""

from __future__ import annotations

from typing import Dict, Optional, Tuple

import numpy as np

from .params import ModelSpec

Tree = Dict[str, object]


def _leaf(tree: Tree, *path: str) -> np.ndarray:
    node = tree
    for p in path:
        if not isinstance(node, dict) or p not in node:
            raise KeyError("missing parameter " + "/".join(path))
        node = node[p]
    return np.asarray(node, np.float32)


def _opt_leaf(tree: Tree, *path: str) -> Optional[np.ndarray]:
    try:
        return _leaf(tree, *path)
    except KeyError:
        return None


def _variables(tree: Tree) -> Tree:
    """Accept the variable dict ({'params': {...}}) or the bare parameter dict."""
    if isinstance(tree, dict) and set(tree) == {"params"} and isinstance(tree["params"], dict):
        return tree["params"]
    return tree


def _outputs(spec: ModelSpec, t: Tree, p: dict):
    K, N = spec.ndets, spec.n_elec
    orb = t["orbital_layer"]
    if spec.n_orbital_sets == 2:
        dg = [orb["SplitChannelDense_0"][f"DenseGeneral_{c}"] for c in range(2)]
    else:
        dg = [orb["DenseGeneral_0"]]
    p["orbital"] = [np.asarray(d["kernel"], np.float32).reshape(-1, K * N) for d in dg]
    biases = [d.get("bias") for d in dg]
    if spec.kind == "psiformer" and spec.bias_orbitals:
        if any(b is None for b in biases):
            raise KeyError("bias_orbitals=True but the orbital layer has no bias")
        p["orbital_bias"] = [np.asarray(b, np.float32).reshape(K * N) for b in biases]
    elif any(b is not None for b in biases):
        raise ValueError("the orbital layer has a bias: set bias_orbitals=True")
    if spec.envelope == "null":
        M = spec.n_atoms
        one = np.ones((N, M, K), np.float32)
        p["envelope"] = [(one, one) for _ in range(spec.n_orbital_sets)]
    else:
        names = ["_env_up", "_env_down"] if spec.n_orbital_sets == 2 else ["_env"]
        env = t["envelope_layer"]
        p["envelope"] = [(_leaf(env, n, "pi"), _leaf(env, n, "sigma")) for n in names]


def from_flax_tree(spec: ModelSpec, tree: Tree) -> dict:
    """The ``jaqmc_amd.params`` tree of a reference Flax parameter tree."""
    if spec.use_last_layer:
        raise ValueError("use_last_layer=True is not supported by the device path")
    t = _variables(tree)
    p: dict = {}
    if spec.kind == "ferminet":
        bb = t["backbone_layer"]
        L = len(spec.hidden_single)
        p["single"] = [(_leaf(bb, f"Dense_{2 * l}", "kernel"), _leaf(bb, f"Dense_{2 * l}", "bias"))
                       for l in range(L)]
        p["double"] = [(_leaf(bb, f"Dense_{2 * l + 1}", "kernel"), _leaf(bb, f"Dense_{2 * l + 1}", "bias"))
                       for l in range(L - 1)]
    else:
        bb = t["backbone_layer"]
        D = spec.num_heads * spec.heads_dim
        zero = np.zeros(D, np.float32)
        b_in = _opt_leaf(bb, "Dense_0", "bias")
        p["input"] = (_leaf(bb, "Dense_0", "kernel"), b_in if b_in is not None else zero)
        layers = []
        for l in range(spec.num_layers):
            ly = bb[f"PsiformerLayer_{l}"]
            mha = ly["MultiHeadDotProductAttention_0"]
            d = {}
            if spec.layer_norm_mode == "null":
                d["ln1_scale"], d["ln1_bias"] = np.ones(D, np.float32), zero
                d["ln2_scale"], d["ln2_bias"] = np.ones(D, np.float32), zero
            else:
                for i in (1, 2):
                    d[f"ln{i}_scale"] = _leaf(ly, f"LayerNorm_{i - 1}", "scale")
                    d[f"ln{i}_bias"] = _leaf(ly, f"LayerNorm_{i - 1}", "bias")
            for nm, key in (("q", "query"), ("k", "key"), ("v", "value")):
                d["w" + nm] = _leaf(mha, key, "kernel").reshape(D, D)
                b = _opt_leaf(mha, key, "bias")
                d["b" + nm] = b.reshape(D) if b is not None else zero
            d["wo"] = _leaf(mha, "out", "kernel").reshape(D, D)
            b = _opt_leaf(mha, "out", "bias")
            d["bo"] = b if b is not None else zero
            n_mlp = len(spec.mlp_hidden) + 1
            d["mlp"] = [(_leaf(ly, f"Dense_{j}", "kernel"), _leaf(ly, f"Dense_{j}", "bias")) for j in range(n_mlp)]
            layers.append(d)
        p["layers"] = layers
        if spec.jastrow:
            jas = t["jastrow_layer"]
            p["jastrow"] = np.array([_leaf(jas, "alpha_par").reshape(-1)[0], _leaf(jas, "alpha_anti").reshape(-1)[0]],
                                    np.float32)
        else:
            p["jastrow"] = np.array([1.0, 1.0], np.float32)
    _outputs(spec, t, p)
    return p


def to_flax_tree(spec: ModelSpec, p: dict) -> Tree:
    """Inverse of :func:`from_flax_tree`: the reference's variable dict
    ({'params': {...}}) for a ``jaqmc_amd.params`` tree (Flax shapes)."""
    K, N = spec.ndets, spec.n_elec
    t: Tree = {}
    if spec.kind == "ferminet":
        bb = {}
        for l, (W, b) in enumerate(p["single"]):
            bb[f"Dense_{2 * l}"] = {"kernel": W, "bias": b}
        for l, (W, b) in enumerate(p["double"]):
            bb[f"Dense_{2 * l + 1}"] = {"kernel": W, "bias": b}
    else:
        D = spec.num_heads * spec.heads_dim
        H, dh = spec.num_heads, spec.heads_dim
        bb = {"Dense_0": {"kernel": p["input"][0]}}
        if spec.input_bias:
            bb["Dense_0"]["bias"] = p["input"][1]
        for l, d in enumerate(p["layers"]):
            ly: Tree = {}
            if spec.layer_norm_mode != "null":
                for i in (1, 2):
                    ly[f"LayerNorm_{i - 1}"] = {"scale": d[f"ln{i}_scale"], "bias": d[f"ln{i}_bias"]}
            mha: Tree = {}
            for nm, key in (("q", "query"), ("k", "key"), ("v", "value")):
                mha[key] = {"kernel": d["w" + nm].reshape(D, H, dh)}
                if spec.with_bias:
                    mha[key]["bias"] = d["b" + nm].reshape(H, dh)
            mha["out"] = {"kernel": d["wo"].reshape(H, dh, D)}
            if spec.with_bias:
                mha["out"]["bias"] = d["bo"]
            ly["MultiHeadDotProductAttention_0"] = mha
            for j, (W, b) in enumerate(d["mlp"]):
                ly[f"Dense_{j}"] = {"kernel": W, "bias": b}
            bb[f"PsiformerLayer_{l}"] = ly
        if spec.jastrow:
            t["jastrow_layer"] = {"alpha_par": np.asarray(p["jastrow"][:1], np.float32),
                                  "alpha_anti": np.asarray(p["jastrow"][1:2], np.float32)}
    t["backbone_layer"] = bb
    dgs = []
    for c, W in enumerate(p["orbital"]):
        dg = {"kernel": np.asarray(W).reshape(-1, K, N)}
        if "orbital_bias" in p:
            dg["bias"] = np.asarray(p["orbital_bias"][c]).reshape(K, N)
        dgs.append(dg)
    if spec.n_orbital_sets == 2:
        t["orbital_layer"] = {"SplitChannelDense_0": {"DenseGeneral_0": dgs[0], "DenseGeneral_1": dgs[1]}}
    else:
        t["orbital_layer"] = {"DenseGeneral_0": dgs[0]}
    if spec.envelope != "null":
        names = ["_env_up", "_env_down"] if spec.n_orbital_sets == 2 else ["_env"]
        t["envelope_layer"] = {n: {"pi": pi, "sigma": sg} for n, (pi, sg) in zip(names, p["envelope"])}
    return {"params": t}


def tree_from_flat(flat: Dict[str, np.ndarray], prefix: str) -> Tree:
    """Nest the slash-joined key paths under ``prefix`` of a checkpoint file."""
    out: Tree = {}
    pre = prefix.rstrip("/") + "/"
    for key, val in flat.items():
        if not key.startswith(pre):
            continue
        node = out
        parts = key[len(pre):].split("/")
        for q in parts[:-1]:
            node = node.setdefault(q, {})
        node[parts[-1]] = np.asarray(val)
    return out


def flat_from_tree(tree: Tree, prefix: str) -> Dict[str, np.ndarray]:
    """Slash-joined key paths of a nested dict (the checkpoint's key format)."""
    out: Dict[str, np.ndarray] = {}

    def walk(node, path):
        if isinstance(node, dict):
            for k, v in node.items():
                walk(v, path + [str(k)])
        else:
            out["/".join(path)] = np.asarray(node)

    walk(tree, [prefix.rstrip("/")])
    return out


def load_checkpoint(path, spec: ModelSpec) -> Tuple[dict, Optional[np.ndarray]]:
    """(parameters, electrons or None) of a reference ``*ckpt_*.npz`` file.

    Parameters come from the ``params/`` subtree (the VMC state's Flax variable
    dict, so keys read ``params/params/backbone_layer/...``), the walkers from
    the ``batched_data/`` leaf named ``electrons`` when present.  The file is
    opened with ``allow_pickle=False``: object arrays are refused.
    """
    with np.load(path, allow_pickle=False) as f:
        flat = {k: f[k] for k in f.files}
    params = from_flax_tree(spec, tree_from_flat(flat, "params"))
    elec = None
    for k, v in flat.items():
        if k.startswith("batched_data/") and k.rsplit("/", 1)[-1] == "electrons":
            elec = np.asarray(v, np.float32)
    return params, elec
