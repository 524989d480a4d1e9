"""Shared tolerances of the GPU parity tests (see test_gpu_parity.py)."""

import numpy as np

U32 = 2.0 ** -24  # FP32 unit roundoff

# Relative error the orbital-matrix entries may carry after the network, in
# units of u = 2^-24, per GEMM arithmetic: FP32 and BF16X6 (products to
# 2^-24) accumulate a few tens of roundings over four layers; FP16X3's operand
# terms are exact to 2^-22 (four times u).
ENTRY_ERR = {"fp32": 64, "bf16x6": 64, "fp16x3": 256}


def logpsi_bound(logpsi, kappa, prec="fp32"):
    """|log psi_gpu - log psi_oracle| bound: 1e-5 |log psi| + 1e-4 for the
    network's own rounding, plus the first-order perturbation bound of
    log|det A| under relative entry errors e: N e kappa(A) <= 4 e kappa for the
    determinant blocks here (kappa: the walker's largest 2-norm condition
    number, from the oracle)."""
    return 1e-5 * np.abs(logpsi) + 1e-4 + 4 * ENTRY_ERR[prec] * U32 * kappa


def el_bound(e_ref, kappa, prec="fp32"):
    """Per-walker |E_L gpu - E_L oracle| bound relative to max(1, |E_L|):
    1e-5 (the stated FP32 target) plus the first-order amplification of the
    orbital-entry error e by the inverse: tr(A^-1 L) and tr((A^-1 J)^2) move
    by about e kappa relative, so the kinetic energy by 2 e kappa |terms|."""
    return 1e-5 + 2 * ENTRY_ERR[prec] * U32 * kappa


def el_report(tag, e_gpu, e_ref, kappa, prec="fp32"):
    rel = np.abs(e_gpu - e_ref) / np.maximum(1.0, np.abs(e_ref))
    bnd = el_bound(e_ref, kappa, prec)
    worst = int(np.argmax(rel / bnd))
    print(f"[parity] {tag} {prec}: median rel {np.median(rel):.2e}, max rel {rel.max():.2e}, "
          f"within 1e-5: {np.mean(rel <= 1e-5):.3f}, worst rel/bound {rel[worst] / bnd[worst]:.3f} "
          f"(kappa {kappa[worst]:.2e})", flush=True)
    return rel, bnd
