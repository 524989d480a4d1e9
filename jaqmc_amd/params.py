"""Model descriptions, seeded parameter initialisation and the flat parameter pack.

The parameter tree mirrors the reference's Flax modules:

FermiNet (app/molecule/wavefunction/ferminet.py:44-74)
    single[l] = (kernel (in_l, d1_l), bias (d1_l,))       FermiLayers Dense, one-electron stream
    double[l] = (kernel (in2_l, d2_l), bias (d2_l,))      l < L-1 (use_last_layer=False)
    orbital[c] = kernel (d1_L, K*N)                       SplitChannelDense, no bias
    envelope[c] = (pi (N, M, K), sigma (N, M, K))         one per spin channel
    (orbitals_spin_split=False, or one occupied channel: a single orbital
    kernel and a single envelope shared by all electrons)

Psiformer (app/molecule/wavefunction/psiformer.py:93-134)
    input = (kernel (4M+1, D), bias (D,))
    layers[l] = {ln1_scale, ln1_bias, wq (D, D), bq, wk, bk, wv, bv, wo (D, D), bo,
                 ln2_scale, ln2_bias, mlp = [(kernel, bias), ...]}
    orbital[c], [orbital_bias[c] (K*N,) when bias_orbitals], envelope[c] as
    above; jastrow = (alpha_par, alpha_anti).  A module built without a bias
    (with_bias / input_bias False) keeps its bias slots in the pack as zeros.

``pack`` flattens the tree into one float32 buffer in exactly the order listed
above (row-major arrays); that buffer is what ``jq_model_create`` consumes.
``from_flax_tree`` maps a reference Flax parameter tree onto this layout.
"""

from __future__ import annotations

import dataclasses
from typing import List, Tuple

import numpy as np

ENVELOPES = {"isotropic": 0, "abs_isotropic": 1, "null": 2}


@dataclasses.dataclass
class ModelSpec:
    kind: str  # "ferminet" | "psiformer"
    nspins: Tuple[int, int]
    n_atoms: int
    ndets: int = 16
    # FermiNet
    hidden_single: List[int] = dataclasses.field(default_factory=lambda: [256] * 4)
    hidden_double: List[int] = dataclasses.field(default_factory=lambda: [32] * 4)
    # Psiformer
    num_layers: int = 4
    num_heads: int = 4
    heads_dim: int = 64
    mlp_hidden: List[int] = dataclasses.field(default_factory=lambda: [256])
    jastrow: bool = True
    envelope: str = "abs_isotropic"
    # module options of the reference (defaults = the reference's defaults)
    layer_norm_mode: str = "pre"      # Psiformer: "pre" | "post" | "null"
    with_bias: bool = True            # Psiformer attention biases
    input_bias: bool = True           # Psiformer input projection bias
    bias_orbitals: bool = False       # Psiformer orbital projection bias
    rescale: bool = True              # Psiformer log-scaled features (FermiNet: always unscaled)
    orbitals_spin_split: bool = True  # separate orbital kernel / envelope per spin channel
    use_last_layer: bool = False      # FermiNet (not supported by the device path)

    @property
    def n_elec(self):
        return self.nspins[0] + self.nspins[1]

    @property
    def n_channels(self):
        return 2 if (self.nspins[0] > 0 and self.nspins[1] > 0) else 1

    @property
    def n_orbital_sets(self):
        """Orbital kernels / envelopes in the tree (orbital.py:73-76, envelope.py:82-96)."""
        return 2 if (self.n_channels == 2 and self.orbitals_spin_split) else 1

    def options(self):
        """jq_model_desc.options bits for this spec."""
        from . import _lib

        o = 0
        if self.kind == "psiformer":
            o |= 0 if self.with_bias else _lib.OPT_NO_ATTN_BIAS
            o |= 0 if self.input_bias else _lib.OPT_NO_INPUT_BIAS
            o |= _lib.OPT_ORBITAL_BIAS if self.bias_orbitals else 0
            o |= 0 if self.rescale else _lib.OPT_NO_RESCALE
        o |= 0 if self.orbitals_spin_split else _lib.OPT_SHARED_ORBITALS
        o |= _lib.OPT_USE_LAST_LAYER if self.use_last_layer else 0
        return o

    def as_oracle(self):
        return dict(kind=self.kind, nspins=tuple(self.nspins), ndets=self.ndets,
                    envelope=self.envelope, num_heads=self.num_heads, jastrow=self.jastrow,
                    layer_norm_mode=self.layer_norm_mode, rescale=self.rescale if self.kind == "psiformer" else False,
                    orbitals_spin_split=self.orbitals_spin_split)

    # FermiNet layer input widths (FermiLayers.aggregate_features, ferminet.py:65-90)
    def ferminet_dims(self):
        nch = self.n_channels
        M = self.n_atoms
        ins, ins2 = [], []
        f1, f2 = 4 * M, 4
        for l, (d1, d2) in enumerate(zip(self.hidden_single, self.hidden_double)):
            ins.append(f1 * (1 + nch) + f2 * nch)
            if l < len(self.hidden_single) - 1:
                ins2.append(f2)
                f2 = d2
            f1 = d1
        return ins, ins2


def _glorot(rng, fan_in, fan_out, shape=None):
    std = np.sqrt(2.0 / (fan_in + fan_out))
    return (rng.standard_normal(shape or (fan_in, fan_out)) * std).astype(np.float32)


def init_params(spec: ModelSpec, seed: int = 0, bias_scale: float = 0.0, env_jitter: float = 0.0):
    """Seeded random initialisation (Glorot-normal kernels, biases ~ bias_scale*N(0,1),
    envelope pi = sigma = 1 (+ env_jitter*N(0,1)), LayerNorm scale 1 / bias 0,
    Jastrow alpha = 1).  Matches the reference's initialiser *families*; the JAX
    PRNG stream itself is not reproduced."""
    rng = np.random.default_rng(seed)
    N, M, K = spec.n_elec, spec.n_atoms, spec.ndets
    nch = spec.n_channels
    bias = lambda n: (rng.standard_normal(n) * bias_scale).astype(np.float32)
    p = {}
    if spec.kind == "ferminet":
        ins, ins2 = spec.ferminet_dims()
        p["single"] = [(_glorot(rng, i, d), bias(d)) for i, d in zip(ins, spec.hidden_single)]
        p["double"] = [(_glorot(rng, i, d), bias(d)) for i, d in zip(ins2, spec.hidden_double[:len(ins2)])]
        d_last = spec.hidden_single[-1]
    else:
        D = spec.num_heads * spec.heads_dim
        zero = lambda n: np.zeros(n, np.float32)
        p["input"] = (_glorot(rng, 4 * M + 1, D), bias(D) if spec.input_bias else zero(D))
        abias = bias if spec.with_bias else zero
        layers = []
        for _ in range(spec.num_layers):
            lp = dict(ln1_scale=(1.0 + rng.standard_normal(D) * bias_scale).astype(np.float32),
                      ln1_bias=bias(D))
            for nm in ("q", "k", "v"):
                lp["w" + nm] = _glorot(rng, D, D)
                lp["b" + nm] = abias(D)
            lp["wo"] = _glorot(rng, D, D)
            lp["bo"] = abias(D)
            lp["ln2_scale"] = (1.0 + rng.standard_normal(D) * bias_scale).astype(np.float32)
            lp["ln2_bias"] = bias(D)
            if spec.layer_norm_mode == "null":  # no LayerNorm modules: unused unit slots
                for nm in ("ln1", "ln2"):
                    lp[nm + "_scale"] = np.ones(D, np.float32)
                    lp[nm + "_bias"] = zero(D)
            dims = [D] + list(spec.mlp_hidden) + [D]
            lp["mlp"] = [(_glorot(rng, a, b), bias(b)) for a, b in zip(dims[:-1], dims[1:])]
            layers.append(lp)
        p["layers"] = layers
        d_last = D
        p["jastrow"] = np.array([1.0, 1.0], np.float32)
    no = spec.n_orbital_sets
    p["orbital"] = [_glorot(rng, d_last, K * N) for _ in range(no)]
    if spec.kind == "psiformer" and spec.bias_orbitals:
        p["orbital_bias"] = [bias(K * N) for _ in range(no)]
    p["envelope"] = [((1.0 + env_jitter * rng.standard_normal((N, M, K))).astype(np.float32),
                      (1.0 + env_jitter * rng.standard_normal((N, M, K))).astype(np.float32))
                     for _ in range(no)]
    return p


def param_arrays(spec: ModelSpec, p):
    """The flat pack order, as a list of (name, array)."""
    out = []
    if spec.kind == "ferminet":
        for l, (W, b) in enumerate(p["single"]):
            out += [(f"single{l}.w", W), (f"single{l}.b", b)]
        for l, (W, b) in enumerate(p["double"]):
            out += [(f"double{l}.w", W), (f"double{l}.b", b)]
    else:
        out += [("input.w", p["input"][0]), ("input.b", p["input"][1])]
        for l, lp in enumerate(p["layers"]):
            for nm in ("ln1_scale", "ln1_bias", "wq", "bq", "wk", "bk", "wv", "bv", "wo", "bo",
                       "ln2_scale", "ln2_bias"):
                out.append((f"layer{l}.{nm}", lp[nm]))
            for j, (W, b) in enumerate(lp["mlp"]):
                out += [(f"layer{l}.mlp{j}.w", W), (f"layer{l}.mlp{j}.b", b)]
    for c, W in enumerate(p["orbital"]):
        out.append((f"orbital{c}.w", W))
    for c, b in enumerate(p.get("orbital_bias", [])):
        out.append((f"orbital{c}.b", b))
    for c, (pi, sg) in enumerate(p["envelope"]):
        out += [(f"envelope{c}.pi", pi), (f"envelope{c}.sigma", sg)]
    if spec.kind == "psiformer":
        j = p["jastrow"]
        out.append(("jastrow", j if hasattr(j, "shape") else np.asarray(j, np.float32)))
    return out


def pack(spec: ModelSpec, p) -> np.ndarray:
    return np.concatenate([np.ascontiguousarray(a, np.float32).ravel() for _, a in param_arrays(spec, p)])


def unpack(spec: ModelSpec, flat: np.ndarray, like, dtype=np.float32):
    """Inverse of ``pack`` given a tree with the right shapes (``like``)."""
    flat = np.asarray(flat, dtype)
    pos = 0
    arrays = []
    for _, a in param_arrays(spec, like):
        n = a.size
        arrays.append(flat[pos:pos + n].reshape(a.shape))
        pos += n
    it = iter(arrays)
    q = {}
    if spec.kind == "ferminet":
        q["single"] = [(next(it), next(it)) for _ in like["single"]]
        q["double"] = [(next(it), next(it)) for _ in like["double"]]
    else:
        q["input"] = (next(it), next(it))
        q["layers"] = []
        for lp in like["layers"]:
            d = {nm: next(it) for nm in ("ln1_scale", "ln1_bias", "wq", "bq", "wk", "bk", "wv", "bv",
                                         "wo", "bo", "ln2_scale", "ln2_bias")}
            d["mlp"] = [(next(it), next(it)) for _ in lp["mlp"]]
            q["layers"].append(d)
    q["orbital"] = [next(it) for _ in like["orbital"]]
    if "orbital_bias" in like:
        q["orbital_bias"] = [next(it) for _ in like["orbital_bias"]]
    q["envelope"] = [(next(it), next(it)) for _ in like["envelope"]]
    if spec.kind == "psiformer":
        q["jastrow"] = next(it)
    return q


def cast_params(p, dtype):
    """Same tree with every array cast (used to evaluate FP32 params in FP64)."""
    if isinstance(p, dict):
        return {k: cast_params(v, dtype) for k, v in p.items()}
    if isinstance(p, (list, tuple)):
        return type(p)(cast_params(v, dtype) for v in p)
    return np.asarray(p, dtype)
