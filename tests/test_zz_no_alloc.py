"""jq_loss_clip_noalloc (static per-GPU scratch, no allocation) against
jq_loss_clip (caller workspace) through the C ABI: same kernels, so the
outputs agree bit for bit.  Two calls in a row, the second on a smaller batch,
show that the scratch state left by one call does not reach the next."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(rng, B):
    x = rng.standard_cauchy(B) * 0.5 - 75.0
    x[rng.integers(0, B, size=max(1, B // 500))] = np.nan
    return x


@pytest.mark.parametrize("method", ["mad", "iqr", "none"])
def test_noalloc_matches_workspace_entry(cuda, method):
    import torch

    from jaqmc_amd.device import loss_clip, loss_clip_noalloc

    rng = np.random.default_rng(11)
    for B in (200_003, 777):  # the second call is smaller than the first
        x = torch.from_numpy(_inputs(rng, B)).to(cuda)
        out = tuple(torch.full((n,), -1.0, dtype=torch.float64, device=cuda) for n in (B, B, 5))
        c1, w1, s1 = loss_clip_noalloc(x, method, 5.0, out=out)
        c0, w0, s0 = loss_clip(x, method, 5.0)
        torch.cuda.synchronize()
        print(f"[noalloc] {method} B={B} window {s1[2].item():.6f} {s1[3].item():.6f} n {s1[4].item():.0f}",
              flush=True)
        np.testing.assert_array_equal(c1.cpu().numpy(), c0.cpu().numpy())
        np.testing.assert_array_equal(w1.cpu().numpy(), w0.cpu().numpy())
        np.testing.assert_array_equal(s1.cpu().numpy(), s0.cpu().numpy())
        assert s1[4].item() == float(np.sum(~np.isnan(x.cpu().numpy())))


def test_noalloc_rejects_oversize_batch(cuda):
    import torch

    from jaqmc_amd import _lib

    lib = _lib.load()
    x = torch.zeros(4, dtype=torch.float64, device=cuda)
    w = torch.zeros(4, dtype=torch.float64, device=cuda)
    # rejected on the host before any kernel is launched
    rc = lib.jq_loss_clip_noalloc(x.data_ptr(), (1 << 28) + 1, 0, 5.0, None, w.data_ptr(), None, None)
    assert rc == -1
    assert b"JQ_LOSS_CLIP" in lib.jq_last_error()
