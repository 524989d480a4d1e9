"""LossAndGrad's clipping and gradient weights on the GPU against numpy
(estimator/loss_grad.py:91-128, utils/clip.py:11-60): nanmedian / nanquantile
with linear interpolation and jnp.clip semantics (NaN stays NaN).  The order
statistics are exact; the interpolated window agrees to FP64 rounding (numpy
and jnp write the same interpolation in different FP64 forms), and the
clipped values equal numpy's clip with the GPU's window exactly."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np_clip(x, method, scale):
    if method == "mad":
        m = np.nanmedian(x)
        d = np.nanmedian(np.abs(x - m))
        lo, hi = m - scale * d, m + scale * d
    elif method == "iqr":
        q1, q3 = np.nanquantile(x, 0.25), np.nanquantile(x, 0.75)
        lo, hi = q1 - scale * (q3 - q1), q3 + scale * (q3 - q1)
    else:
        lo, hi = -np.inf, np.inf
    c = np.where(x < lo, lo, np.where(x > hi, hi, x))
    ok = ~np.isnan(c)
    w = np.where(ok, 2.0 * (c - np.nanmean(c)) / ok.sum(), 0.0)
    return c, w, (np.nanmean(x), np.nanmean(c), lo, hi, ok.sum())


CASES = {
    "gauss_4096": lambda r: r.standard_normal(4096) * 3 - 40,
    "heavy_tails_nan": lambda r: np.concatenate([r.standard_cauchy(3001), [np.nan] * 7]),
    "ties_and_inf": lambda r: np.concatenate([np.round(r.standard_normal(500)), [np.inf, -np.inf, np.nan]]),
    "one": lambda r: np.array([-1.5]),
    "two": lambda r: np.array([2.0, -7.25]),
    "three_with_nan": lambda r: np.array([np.nan, 1.0, 4.0, -2.0]),
    "large_1M": lambda r: r.standard_normal(1 << 20) * 0.1 - 100.0,
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("method", ["mad", "iqr", "none"])
def test_loss_clip_matches_numpy(cuda, case, method):
    import torch

    from jaqmc_amd.device import loss_clip

    x = CASES[case](np.random.default_rng(3)).astype(np.float64)
    c, w, st = loss_clip(torch.from_numpy(x).to(cuda), method, 5.0)
    rc, rw, rst = _np_clip(x, method, 5.0)
    st = st.cpu().numpy()
    np.testing.assert_allclose(st[2:4], rst[2:4], rtol=1e-14, atol=0)
    lo, hi = st[2], st[3]
    np.testing.assert_array_equal(c.cpu().numpy(), np.where(x < lo, lo, np.where(x > hi, hi, x)))
    assert st[4] == rst[4]
    # the means: FP64 sums in a different order (Cauchy samples cancel strongly)
    scale = np.nanmean(np.abs(np.where(np.isfinite(x), x, 0.0))) + 1.0
    np.testing.assert_allclose(st[:2], rst[:2], rtol=1e-12, atol=1e-12 * scale)
    np.testing.assert_allclose(w.cpu().numpy(), rw, rtol=1e-9, atol=1e-15)


def test_weights_give_the_reference_gradient_combination(cuda):
    """sum_w weights_w g_w equals 2 (mean(g c) - mean(c) mean(g)) for any
    per-walker vectors g (LossAndGrad.finalize_stats)."""
    import torch

    from jaqmc_amd.device import loss_clip

    rng = np.random.default_rng(5)
    e = rng.standard_normal(1000) - 10
    e[[3, 17]] = [500.0, -900.0]  # outliers the MAD window clips
    g = rng.standard_normal((1000, 7))
    c, w, _ = loss_clip(torch.from_numpy(e).to(cuda), "mad", 5.0)
    c, w = c.cpu().numpy(), w.cpu().numpy()
    ref = 2 * ((g * c[:, None]).mean(0) - c.mean() * g.mean(0))
    np.testing.assert_allclose(w @ g, ref, rtol=1e-10)
