"""Scheduler logic of the continuous-batching server (CPU, fake runners)."""

import numpy as np
import pytest

from jaqmc_amd import serving
from jaqmc_amd.params import ModelSpec


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeRunner:
    """Completes one enqueued step per poll; energy = job_id + step count."""

    log = []

    def __init__(self, job, clock):
        self.job = job
        self.clock = clock
        self.enq = 0
        self.done = 0
        FakeRunner.log.append(("admit", job.job_id, clock()))

    def enqueue_step(self, i):
        assert i == self.enq
        self.enq += 1

    def poll(self):
        self.done = self.enq
        return True

    def result(self):
        return float(self.job.job_id), 1.0, self.job.walkers * self.job.steps, 0.5

    def close(self):
        FakeRunner.log.append(("close", self.job.job_id, self.clock()))


def _job(i, walkers, steps, arrival=0.0, n=4):
    spec = ModelSpec("ferminet", (n - n // 2, n // 2), 1, ndets=2, hidden_single=[8, 8], hidden_double=[4, 4])
    return serving.Job(i, np.zeros((1, 3), np.float32), np.array([float(n)], np.float32), spec, {}, walkers, steps,
                       arrival=arrival)


def test_admission_respects_memory_budget_and_fifo():
    clock = FakeClock()
    FakeRunner.log = []
    fp = lambda job: job.walkers  # bytes == walkers for the test
    srv = serving.Server(lambda j: FakeRunner(j, clock), mem_budget=5000, max_active=8, footprint=fp, clock=clock,
                         async_build=False)
    for i, (w, s) in enumerate([(4096, 3), (1024, 1), (256, 1), (4096, 2)]):
        srv.submit(_job(i, w, s))
    ticks = 0
    while srv.tick():
        clock.t += 1.0
        ticks += 1
        assert srv.mem_used <= 5000
        assert ticks < 50
    # 4096 + 1024 > 5000: job 1 waits for job 0 to finish (FIFO, no overtaking)
    admits = [e for e in FakeRunner.log if e[0] == "admit"]
    assert [a[1] for a in admits] == [0, 1, 2, 3]
    t_admit = {a[1]: a[2] for a in admits}
    t_close = {e[1]: e[2] for e in FakeRunner.log if e[0] == "close"}
    assert t_admit[1] >= t_close[0]
    assert t_admit[2] == t_admit[1]  # 1024 + 256 fit together
    assert set(srv.results) == {0, 1, 2, 3}
    assert srv.mem_used == 0


def test_oversized_job_rejected():
    srv = serving.Server(lambda j: None, mem_budget=100, footprint=lambda j: j.walkers)
    with pytest.raises(ValueError):
        srv.submit(_job(0, 256, 1))


def test_timed_arrivals_are_admitted_when_they_arrive():
    clock = FakeClock()
    FakeRunner.log = []

    class TickClock(FakeClock):
        def __call__(self):
            self.t += 0.25
            return self.t

    tc = TickClock()
    srv = serving.Server(lambda j: FakeRunner(j, tc), mem_budget=10 ** 9, footprint=lambda j: 1, clock=tc,
                         async_build=False)
    jobs = [_job(i, 256, 2, arrival=float(i)) for i in range(4)]
    res = srv.run(jobs, idle_sleep=0)
    assert set(res) == {0, 1, 2, 3}
    for r in res.values():
        assert r.admitted >= r.arrival and r.finished >= r.admitted
    summ = serving.latency_summary(res)
    assert summ["jobs"] == 4 and summ["p99_s"] >= summ["p50_s"] > 0


def test_job_stream_is_seeded_and_in_range():
    a = serving.make_job_stream(12, rate=5.0, seed=3, steps=2, walkers=(256,))
    b = serving.make_job_stream(12, rate=5.0, seed=3, steps=2, walkers=(256,))
    assert [j.spec.n_elec for j in a] == [j.spec.n_elec for j in b]
    assert all(2 <= j.spec.n_elec <= 42 for j in a)
    assert all(np.array_equal(x.x0, y.x0) for x, y in zip(a, b))
    arr = [j.arrival for j in a]
    assert arr == sorted(arr)


def test_footprint_grows_with_walkers_and_electrons():
    s1 = ModelSpec("ferminet", (3, 3), 2, ndets=16)
    s2 = ModelSpec("ferminet", (21, 21), 8, ndets=16)
    assert serving.footprint_bytes(s1, 2, 1024) < serving.footprint_bytes(s1, 2, 4096)
    assert serving.footprint_bytes(s1, 2, 1024) < serving.footprint_bytes(s2, 8, 1024)


def test_builder_thread_gives_the_same_results():
    """Runners built by the builder threads join the batch when ready; the
    results are those of synchronous construction."""
    out = {}
    for async_build in (False, True):
        clock = FakeClock()
        srv = serving.Server(lambda j: FakeRunner(j, clock), mem_budget=10 ** 9, max_active=3,
                             footprint=lambda j: 1, clock=clock, async_build=async_build)
        for i in range(7):
            srv.submit(_job(i, 64, 1 + i % 3))
        while srv.tick():
            clock.t += 1.0
        srv.close()
        out[async_build] = {k: (r.energy, r.count) for k, r in srv.results.items()}
        assert srv.peak_active <= 3
    assert out[False] == out[True] and len(out[True]) == 7
