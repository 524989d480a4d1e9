"""Shared tolerances of the GPU parity tests (see test_gpu_parity.py)."""

import numpy as np

U32 = 2.0 ** -24  # FP32 unit roundoff

# Relative error the orbital-matrix entries may carry after the network, in
# units of u = 2^-24, per GEMM arithmetic.  On the parity cases here the
# largest measured E_L error is 0.13 of the bound at 16 u in every mode (about
# 2 u of entry error); the bound uses 16 u.
ENTRY_ERR = {"fp32": 16, "bf16x6": 16, "fp16x3": 16}


def cancellation(A):
    """Per-walker cancellation factor of the determinant sum, sum_k |det A_k| /
    |sum_k det A_k| (>= 1), from the oracle's (B, K, N, N) orbital matrices:
    relative errors of the determinants reach log|psi| and its derivatives
    multiplied by it (the combination weights w_k sum to 1 but their absolute
    values sum to this factor)."""
    s, l = np.linalg.slogdet(np.asarray(A, np.float64))
    m = l.max(axis=1, keepdims=True)
    return np.exp(l - m).sum(1) / np.abs((s * np.exp(l - m)).sum(1))


def logpsi_bound(logpsi, kappa, prec="fp32", cancel=1.0):
    """|log psi_gpu - log psi_oracle| bound: 1e-5 |log psi| for the network's
    own rounding, plus (1e-4 + 4 e kappa) C: the first-order perturbation bound
    of log|det A| under relative entry errors e (N e kappa(A) <= 4 e kappa for
    the determinant blocks here; kappa the walker's largest 2-norm condition
    number) amplified by the determinant sum's cancellation factor C."""
    return 1e-5 * np.abs(logpsi) + (1e-4 + 4 * ENTRY_ERR[prec] * U32 * kappa) * cancel


def el_bound(e_ref, kappa, prec="fp32", cancel=1.0):
    """Per-walker |E_L gpu - E_L oracle| bound relative to max(1, |E_L|):
    1e-5 (the stated FP32 target) plus the first-order amplification of the
    orbital-entry error e by the inverse, 2 e kappa, times the cancellation
    factor C of the determinant sum."""
    return 1e-5 + 2 * ENTRY_ERR[prec] * U32 * kappa * cancel


def el_report(tag, e_gpu, e_ref, kappa, prec="fp32", cancel=1.0):
    rel = np.abs(e_gpu - e_ref) / np.maximum(1.0, np.abs(e_ref))
    bnd = el_bound(e_ref, kappa, prec, cancel)
    worst = int(np.argmax(rel / bnd))
    print(f"[parity] {tag} {prec}: median rel {np.median(rel):.2e}, max rel {rel.max():.2e}, "
          f"within 1e-5: {np.mean(rel <= 1e-5):.3f}, worst rel/bound {rel[worst] / bnd[worst]:.3f} "
          f"(kappa {kappa[worst]:.2e}, cancellation {np.broadcast_to(cancel, rel.shape)[worst]:.1f})", flush=True)
    return rel, bnd
