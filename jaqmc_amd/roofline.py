"""Algorithmic work per walker for each launch class (the canonical,
structure-aware counts used for roofline fractions).

C = 3N + 2 tangent columns (value, Laplacian, 3N Jacobian rows).  The
one-electron features of layer 0 are Local1 (5 non-zero rows per electron),
the two-electron stream is Local2 (8 rows per pair).  Padding rows and columns
that the kernels compute are overhead, not algorithmic work.
"""

from __future__ import annotations

from .params import ModelSpec

PEAKS = {  # MI355X published dense peaks
    "fp32": 157.3e12,
    "bf16": 2.5e15,
    "bf16x3": 2.5e15,
    "fp16": 2.5e15,
    "fp64": 78.6e12,
    "hbm": 8.0e12,
}


def ferminet_flops(spec: ModelSpec, n_atoms: int):
    """Per-walker FLOPs of the Laplacian pass, by launch class."""
    N = spec.n_elec
    C = 3 * N + 2
    nch = spec.n_channels
    K = spec.ndets
    ins, _ = spec.ferminet_dims()
    f1, f2 = 4 * n_atoms, 4
    layer = spin = pair = 0.0
    L = len(spec.hidden_single)
    for l, (d1, d2) in enumerate(zip(spec.hidden_single, spec.hidden_double)):
        ch = 5 if l == 0 else C
        layer += 2.0 * N * f1 * d1 * ch + 2.0 * N * nch * f2 * d1 * C
        spin += 2.0 * nch * f1 * d1 * C
        if l < L - 1:
            pair += 2.0 * N * N * f2 * d2 * 8
            f2 = d2
        f1 = d1
    orb = 2.0 * N * f1 * K * N * C
    det = K * (6.0 * N ** 4 + (2.0 / 3.0) * N ** 3 + 2.0 * N ** 3)
    return {"gemm_layer_lap": layer, "gemm_spin_lap": spin, "pair_lap": pair, "gemm_orb_lap": orb,
            "det_lap": det}


def ferminet_forward_flops(spec: ModelSpec, n_atoms: int):
    """Per-walker FLOPs of one log|psi| evaluation (C = 1)."""
    N = spec.n_elec
    nch = spec.n_channels
    K = spec.ndets
    f1, f2 = 4 * n_atoms, 4
    tot = 0.0
    L = len(spec.hidden_single)
    for l, (d1, d2) in enumerate(zip(spec.hidden_single, spec.hidden_double)):
        tot += 2.0 * N * (f1 + nch * f2) * d1 + 2.0 * nch * f1 * d1
        if l < L - 1:
            tot += 2.0 * N * N * f2 * d2
            f2 = d2
        f1 = d1
    tot += 2.0 * N * f1 * K * N
    return {"mfma": tot, "fp64": K * (2.0 / 3.0) * N ** 3}


def psiformer_flops(spec: ModelSpec, n_atoms: int):
    """Per-walker FLOPs of the Psiformer Laplacian pass, by launch class.

    Dense layers on C tangent rows (input layer Local1: 5 rows); attention
    contractions 12 N^2 C D per layer (scores, their Jacobian/Laplacian cross
    terms, and a.v with its product-rule terms)."""
    N = spec.n_elec
    C = 3 * N + 2
    D = spec.num_heads * spec.heads_dim
    K = spec.ndets
    layer = 2.0 * N * (4 * n_atoms + 1) * D * 5
    attn = 0.0
    for _ in range(spec.num_layers):
        layer += 2.0 * N * C * D * 3 * D + 2.0 * N * C * D * D
        dims = [D] + list(spec.mlp_hidden) + [D]
        for a, b in zip(dims[:-1], dims[1:]):
            layer += 2.0 * N * C * a * b
        attn += 12.0 * N * N * C * D
    orb = 2.0 * N * D * K * N * C
    det = K * (6.0 * N ** 4 + (2.0 / 3.0) * N ** 3 + 2.0 * N ** 3)
    return {"gemm_layer_lap": layer, "attn_lap": attn, "gemm_orb_lap": orb, "det_lap": det}


def flops(spec: ModelSpec, n_atoms: int):
    return ferminet_flops(spec, n_atoms) if spec.kind == "ferminet" else psiformer_flops(spec, n_atoms)


def ferminet_layer_bytes(spec: ModelSpec, n_atoms: int):
    """Per-walker HBM bytes the Laplacian-mode layer GEMMs must move at least:
    each operand read once and each output written once, FP32, on the C
    algorithmic tangent rows (the residual is a slice of the input, so it adds
    nothing; weights are per launch, not per walker)."""
    N = spec.n_elec
    C = 3 * N + 2
    nch = spec.n_channels
    f1, f2 = 4 * n_atoms, 4
    tot = 0.0
    L = len(spec.hidden_single)
    for l, (d1, d2) in enumerate(zip(spec.hidden_single, spec.hidden_double)):
        rows1 = 5 if l == 0 else C
        tot += 4.0 * N * (rows1 * f1 + C * nch * f2)  # one-electron input and pair means
        tot += 4.0 * N * C * d1 + 4.0 * C * d1  # output rows and the spin-mean broadcast term
        if l < L - 1:
            f2 = d2
        f1 = d1
    return tot


def ln_fused(spec: ModelSpec, precision: str) -> bool:
    """Whether the Laplacian-mode LayerNorms of a pre-LN Psiformer run in the
    layer GEMMs' epilogue: FP16X3, 3N + 2 rows per unit rounded to 128 and 256
    features (the library's gemm_ln_fusable)."""
    if spec.kind != "psiformer" or precision != "fp16x3" or spec.layer_norm_mode != "pre":
        return False
    R = 3 * spec.n_elec + 2
    rp = next((o for o in (16, 32, 48, 64, 96, 128, 192) if R <= o), -1)
    return rp == 128 and spec.num_heads * spec.heads_dim == 256


def psiformer_layer_bytes(spec: ModelSpec, n_atoms: int, fused_ln: bool = False):
    """Per-walker minimum HBM bytes of the Psiformer Laplacian-mode projections:
    Q/K/V (input, 3D output), output projection and last MLP layer (input,
    residual stream read, residual stream written), hidden MLP layers; with
    fused_ln also the LayerNorm outputs those GEMMs write (LN1 of every layer,
    LN2 of every layer)."""
    N = spec.n_elec
    C = 3 * N + 2
    D = spec.num_heads * spec.heads_dim
    tot = 4.0 * N * (5 * (4 * n_atoms + 1) + C * D)
    for _ in range(spec.num_layers):
        tot += 4.0 * N * C * (D + 3 * D)
        tot += 4.0 * N * C * 3 * D
        dims = [D] + list(spec.mlp_hidden) + [D]
        for j, (a, b) in enumerate(zip(dims[:-1], dims[1:])):
            tot += 4.0 * N * C * (a + b + (D if j == len(dims) - 2 else 0))
        if fused_ln:
            tot += 2 * 4.0 * N * C * D
    return tot


def layer_bytes(spec: ModelSpec, n_atoms: int, precision: str = "fp32"):
    return (ferminet_layer_bytes(spec, n_atoms) if spec.kind == "ferminet"
            else psiformer_layer_bytes(spec, n_atoms, ln_fused(spec, precision)))


MFMA_PER_PRODUCT = {"fp32": 1, "bf16": 1, "bf16x3": 3, "bf16x6": 6, "fp16x3": 3}
# dense MFMA peak of each mode's operand type
MODE_PEAK = {"fp32": "fp32", "bf16": "bf16", "bf16x3": "bf16", "bf16x6": "bf16", "fp16x3": "fp16"}
