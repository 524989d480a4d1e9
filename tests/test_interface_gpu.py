"""The reference-shaped component interfaces driving the HIP path."""

import numpy as np
import pytest

from jaqmc_amd import interface as I
from jaqmc_amd import systems as S

pytestmark = pytest.mark.gpu


def test_evaluation_step_pipeline(cuda):
    import torch

    sysm = S.li_atom()
    wf = I.FermiNetWavefunction(sysm.nspins, 1, ndets=4, hidden_dims_single=(64, 64), hidden_dims_double=(16, 16),
                                envelope="isotropic")
    x = torch.from_numpy(S.init_walkers(sysm, 256, seed=0)).to(cuda)
    data = I.MoleculeData(x, sysm.atoms, sysm.charges)
    params = wf.init_params(data, rngs=0)
    sampler = I.MCMCSampler(steps=3, initial_width=0.05, adapt_frequency=2)
    pipe = I.EstimatorPipeline([I.EuclideanKinetic(wf), I.potential_energy, I.TotalEnergy()])
    st = sampler.init(data, 0)
    widths = []
    for t in range(4):
        data, stats, st, ws = I.evaluation_step(wf, sampler, pipe, params, data, st, rngs=5)
        widths.append(st.stddev)
        tot = (ws["energy:kinetic"] + ws["energy:potential"]).cpu().numpy()
        assert stats["total_energy"] == pytest.approx(np.nanmean(tot), rel=1e-10)
        assert stats["total_energy_var"] >= 0 and 0 < stats["pmove"] <= 1
    assert st.counter == 4
    # counter 2 and 4 adapt the width (sampler/mcmc.py:183-195): pmove at width
    # 0.05 on Li is far above 0.55, so the width grows by 1.1 each time
    assert widths[1] == pytest.approx(0.05 * 1.1) and widths[3] == pytest.approx(0.05 * 1.21)
    lp = wf.logpsi(params, data)
    sg, lp2 = wf.phase_logpsi(params, data)
    assert torch.equal(lp, lp2) and torch.all(sg.abs() == 1)
