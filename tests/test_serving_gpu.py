"""Continuous batching on the GPU: a job's result is identical alone and in a
mix of concurrently running jobs with other electron counts."""

import numpy as np
import pytest

from jaqmc_amd import serving

pytestmark = pytest.mark.gpu


def _run(jobs, budget=40 << 30, max_active=8):
    srv = serving.Server(lambda j: serving.GpuRunner(j), mem_budget=budget, max_active=max_active)
    for j in jobs:
        srv.submit(j)
    while srv.tick():
        pass
    return srv


def test_job_result_identical_alone_and_in_mix(cuda):
    jobs = serving.make_job_stream(4, rate=100.0, seed=7, steps=4, walkers=(64, 128), n_range=(3, 16))
    alone = _run([jobs[0]]).results[0]
    mixed_srv = _run(jobs)
    mixed = mixed_srv.results[0]
    assert mixed_srv.peak_active >= 2  # the jobs really overlapped
    assert alone.energy == mixed.energy and alone.variance == mixed.variance
    assert alone.count == mixed.count == jobs[0].walkers * jobs[0].steps
    assert alone.pmove == mixed.pmove
    for r in mixed_srv.results.values():
        assert np.isfinite(r.energy) and 0.0 < r.pmove <= 1.0


def test_server_matches_direct_device_model(cuda):
    import torch

    from jaqmc_amd.device import DeviceModel, reduce_mean_var

    job = serving.make_job_stream(1, rate=1.0, seed=11, steps=3, walkers=(96,), n_range=(6, 6))[0]
    res = _run([job]).results[0]
    dm = DeviceModel(job.spec, job.atoms, job.charges, job.params, chunk_walkers=job.walkers)
    x = torch.from_numpy(job.x0).cuda()
    logp = 2 * dm.logpsi(x)[0]
    acc = torch.zeros(job.walkers, dtype=torch.int32, device="cuda")
    es = []
    for i in range(job.steps):
        dm.mh_steps(x, logp, 1, job.width, seed=job.seed, ctr=i, n_accept=acc)
        es.append(dm.local_energy(x)["e_loc"].cpu().numpy())
    e = np.concatenate(es)
    assert res.energy == pytest.approx(e.mean(), rel=1e-12)
    assert res.variance == pytest.approx(e.var(), rel=1e-9)
