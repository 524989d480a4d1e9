"""Continuous-batching server for VMC energy-evaluation jobs on one GPU.

A job is (parameters, geometry, walker count, step budget): ``steps``
evaluation iterations, each one all-electron Metropolis move of every walker
followed by its local energy -- the reference's EvaluationWorkStage.compute_step
(workflow/stage/evaluation.py:179-199) -- and the FP64 mean / variance of E_L
over all iterations and walkers (estimator/base.py:27-53).

Scheduling is iteration-level continuous batching:

* admission by memory footprint: a queued job is admitted as soon as its
  device footprint (weights + walker state + workspace, estimated before any
  allocation) fits next to the running jobs and a slot is free;
* every running job keeps its parameters and walker state resident in HBM and
  owns a HIP stream; each scheduler tick enqueues one iteration of every
  running job without host synchronisation, so iterations of different jobs
  (different electron counts, different kernels) run concurrently on the
  device;
* a job leaves the batch at its last iteration; its slot and memory are
  reused by the next queued job on the following tick;
* device state is built off the scheduler thread: an admitted job's runner
  (weight upload, workspace, initial log|psi|) is constructed by a builder
  thread while the scheduler keeps enqueueing the running jobs' iterations,
  and the job joins the batch on the first tick after its runner is ready;
  teardown waits only for the job's own stream (jq_model_destroy synchronises
  nothing device-wide and returns memory to the stream-ordered pool, where the
  next job's workspace is carved from it).

Results do not depend on the mix: each job's Metropolis randomness is Philox
keyed by (job seed, iteration, walker), its kernels read only its own buffers,
and its reductions are per job, so a job evaluated alone and in any mix gives
bit-identical energies.
"""

from __future__ import annotations

import collections
import concurrent.futures as cf
import dataclasses
import time
from typing import Callable, Deque, Dict, List, Optional

import numpy as np

from .params import ModelSpec


@dataclasses.dataclass
class Job:
    job_id: int
    atoms: np.ndarray
    charges: np.ndarray
    spec: ModelSpec
    params: dict
    walkers: int
    steps: int
    seed: int = 0
    width: float = 0.02
    x0: Optional[np.ndarray] = None  # initial walkers (B, N, 3); required
    arrival: float = 0.0             # submit time (seconds, scheduler clock)


@dataclasses.dataclass
class JobResult:
    job_id: int
    energy: float
    variance: float
    count: int
    pmove: float
    arrival: float
    admitted: float
    finished: float
    n_elec: int
    walkers: int
    steps: int

    @property
    def latency(self):
        return self.finished - self.arrival


JOB_WORKSPACE = 4 << 30  # bytes of walker-chunk workspace per job


def job_chunk(spec: ModelSpec, n_atoms: int, walkers: int) -> int:
    """Walkers per workspace chunk of a job: the whole job when it fits
    JOB_WORKSPACE, else as many as fit."""
    per = _per_walker_bytes(spec, n_atoms)
    return int(max(1, min(walkers, JOB_WORKSPACE // per)))


def footprint_bytes(spec: ModelSpec, n_atoms: int, walkers: int, precision: str = "fp32") -> int:
    """Device bytes of a job: weights + its chunk workspace + walker state.

    Mirrors the library's FermiNet workspace sizing (jq_model::init_ferminet):
    two tangent buffers of N*Rp rows, the spin-mean and broadcast terms, the
    pair stream, the orbital tensor, the envelope and the determinant records.
    """
    from .params import init_params, pack

    N = spec.n_elec
    T = 3 * N
    K = spec.ndets
    weights = pack(spec, init_params(spec, seed=0)).size * 4 * (4 if precision != "fp32" else 1)
    state = walkers * (T * 4 * 2 + 4 * 2 + 4 + 8 * 4)
    per = _per_walker_bytes(spec, n_atoms)
    return int(weights + per * job_chunk(spec, n_atoms, walkers) + state + (32 << 20))


def _per_walker_bytes(spec: ModelSpec, n_atoms: int) -> int:
    r32 = lambda v: (v + 31) // 32 * 32
    N = spec.n_elec
    T = 3 * N
    Rp = next(o for o in (16, 32, 48, 64, 96, 128, 192) if T + 2 <= o)
    K = spec.ndets
    ldphi = (K * N + 3) // 4 * 4
    if spec.kind == "ferminet":
        nch = spec.n_channels
        f1s = [4 * n_atoms] + list(spec.hidden_single[:-1])
        f2s = [4] + list(spec.hidden_double[:len(spec.hidden_single) - 1])
        kx = max(max(r32(f1 + nch * f2) for f1, f2 in zip(f1s, f2s)), r32(spec.hidden_single[-1]))
        kg = max(r32(nch * f1) for f1 in f1s)
        d1 = max(spec.hidden_single)
        f2m = max([4] + list(spec.hidden_double))
        per = 4 * (2 * N * Rp * kx + Rp * kg + Rp * d1 + 2 * N * N * 8 * f2m + N * Rp * ldphi + N * K * N * 5) \
            + 8 * K * (3 + T) + 64
    else:
        D = spec.num_heads * spec.heads_dim
        u = N * Rp
        H = spec.num_heads
        per = 4 * (7 * u * D + 2 * H * u * ((N + 3) // 4 * 4) + H * u * r32(N) + H * N * r32(u) + u * ldphi
                   + N * K * N * 5) + 8 * K * (3 + T) + 64
    return int(per)


class Server:
    """Iteration-level continuous batching over per-job runners.

    ``runner_factory(job) -> runner`` builds the device state of a job; a
    runner exposes ``enqueue_step(i)`` (asynchronous), ``poll()`` (True when
    every enqueued step has completed), ``result()`` and ``close()``.
    """

    def __init__(self, runner_factory: Callable, mem_budget: int, max_active: int = 8,
                 footprint: Callable = None, clock: Callable = time.perf_counter, async_build: bool = True,
                 builders: int = 2, on_done: Optional[Callable] = None):
        self.factory = runner_factory
        self.on_done = on_done
        self.async_build = async_build
        self._pool = cf.ThreadPoolExecutor(builders) if async_build else None
        self.starting: List[dict] = []
        self.mem_budget = mem_budget
        self.max_active = max_active
        self.footprint = footprint or (lambda job: footprint_bytes(job.spec, job.atoms.shape[0], job.walkers))
        self.clock = clock
        self.queue: Deque[Job] = collections.deque()
        self.active: List[dict] = []
        self.results: Dict[int, JobResult] = {}
        self.mem_used = 0
        self.peak_active = 0
        self.admit_log: List[tuple] = []

    def submit(self, job: Job):
        if self.footprint(job) > self.mem_budget:
            raise ValueError(f"job {job.job_id} needs {self.footprint(job)} bytes > budget {self.mem_budget}")
        self.queue.append(job)

    def _admit(self):
        # FIFO admission; a large job at the head waits for memory rather than
        # being overtaken (no starvation).  Memory and the slot are reserved at
        # admission; the runner is built by a builder thread (async_build) or
        # here.
        while self.queue and len(self.active) + len(self.starting) < self.max_active:
            job = self.queue[0]
            need = self.footprint(job)
            if self.mem_used + need > self.mem_budget:
                break
            self.queue.popleft()
            self.mem_used += need
            ent = dict(job=job, next=0, need=need, admitted=self.clock())
            if self.async_build:
                ent["future"] = self._pool.submit(self.factory, job)
                self.starting.append(ent)
            else:
                ent["runner"] = self.factory(job)
                self.active.append(ent)
            self.admit_log.append((job.job_id, len(self.active) + len(self.starting), self.mem_used))
        # runners that finished building join the batch
        still = []
        for ent in self.starting:
            if ent["future"].done():
                ent["runner"] = ent.pop("future").result()
                self.active.append(ent)
            else:
                still.append(ent)
        self.starting = still
        self.peak_active = max(self.peak_active, len(self.active))

    def tick(self):
        """Admit what fits, enqueue one iteration of every running job that has
        iterations left, retire finished jobs.  Returns True while work remains."""
        self._admit()
        for a in self.active:
            if a["next"] < a["job"].steps:
                a["runner"].enqueue_step(a["next"])
                a["next"] += 1
        still = []
        for a in self.active:
            job = a["job"]
            if a["next"] >= job.steps and a["runner"].poll():
                e, v, n, pm = a["runner"].result()
                self.results[job.job_id] = JobResult(job.job_id, e, v, n, pm, job.arrival, a["admitted"],
                                                     self.clock(), job.spec.n_elec, job.walkers, job.steps)
                a["runner"].close()
                self.mem_used -= a["need"]
                if self.on_done:
                    self.on_done(self.results[job.job_id])
            else:
                still.append(a)
        self.active = still
        return bool(self.queue or self.active or self.starting)

    def run(self, arrivals: Optional[List[Job]] = None, idle_sleep: float = 0.0005):
        """Serve until every submitted job (and every timed arrival) is done.
        ``arrivals`` are submitted when the clock passes their ``arrival`` time."""
        pending = collections.deque(sorted(arrivals or [], key=lambda j: j.arrival))
        while True:
            now = self.clock()
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            busy = self.tick()
            if not busy and not pending:
                break
            if not self.active and not self.starting and pending:
                wait = pending[0].arrival - self.clock()
                if wait > 0:
                    time.sleep(min(wait, 0.01))
            elif idle_sleep:
                time.sleep(idle_sleep)
        return self.results

    def close(self):
        if self._pool is not None:
            self._pool.shutdown(wait=True)


class GpuRunner:
    """Device state of one job: model, walkers, Metropolis state, per-step
    FP64 statistics, on the job's own HIP stream."""

    def __init__(self, job: Job, precision: str = "fp32"):
        import torch

        from .device import DeviceModel

        self.torch = torch
        self.job = job
        self.stream = torch.cuda.Stream()
        with torch.cuda.stream(self.stream):
            self.model = DeviceModel(job.spec, job.atoms, job.charges, job.params, precision=precision,
                                     chunk_walkers=job_chunk(job.spec, job.atoms.shape[0], job.walkers))
            self.x = torch.from_numpy(np.ascontiguousarray(job.x0, np.float32)).to("cuda", non_blocking=False)
            self.logp = 2.0 * self.model.logpsi(self.x)[0]
            self.acc = torch.zeros(job.walkers, dtype=torch.int32, device="cuda")
            self.stats = torch.zeros(job.steps, 3, dtype=torch.float64, device="cuda")
        self.done_event = torch.cuda.Event()

    def enqueue_step(self, i: int):
        from .device import reduce_mean_var

        with self.torch.cuda.stream(self.stream):
            self.model.mh_steps(self.x, self.logp, 1, self.job.width, seed=self.job.seed, ctr=i, n_accept=self.acc)
            out = self.model.local_energy(self.x)
            self.stats[i] = reduce_mean_var(out["e_loc"])
            self.done_event.record(self.stream)

    def poll(self) -> bool:
        return self.done_event.query()

    def result(self):
        s = self.stats.cpu().numpy()
        n = s[:, 2]
        tot = n.sum()
        mean = float((s[:, 0] * n).sum() / tot)
        ex2 = float(((s[:, 1] + s[:, 0] ** 2) * n).sum() / tot)
        pmove = float(self.acc.sum().item()) / (self.job.walkers * self.job.steps)
        return mean, ex2 - mean * mean, int(tot), pmove

    def close(self):
        self.stream.synchronize()
        self.model.close()
        del self.x, self.logp, self.acc, self.stats


def make_job_stream(n_jobs: int, rate: float, seed: int = 0, steps: int = 100, walkers=(256, 1024, 4096),
                    n_range=(2, 42), ndets: int = 16, start: float = 0.0) -> List[Job]:
    """Seeded Poisson stream of random-molecule jobs: neutral spin-balanced
    chains with Z in 1..8 and 1.5-3 bohr spacing, FermiNet 4x256/32, 16
    determinants, walkers uniform in ``walkers``."""
    from .params import init_params
    from .systems import init_walkers, random_chain

    rng = np.random.default_rng(seed)
    t = start
    jobs = []
    for j in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        sysm = random_chain(rng, n)
        spec = ModelSpec("ferminet", sysm.nspins, sysm.atoms.shape[0], ndets=ndets)
        B = int(rng.choice(walkers))
        params = init_params(spec, seed=int(rng.integers(1 << 30)))
        x0 = init_walkers(sysm, B, seed=int(rng.integers(1 << 30)))
        jobs.append(Job(j, sysm.atoms, sysm.charges, spec, params, B, steps, seed=int(rng.integers(1 << 30)),
                        x0=x0, arrival=t))
    return jobs


def latency_summary(results: Dict[int, JobResult]):
    lat = np.array([r.latency for r in results.values()])
    ws = sum(r.walkers * r.steps for r in results.values())
    t0 = min(r.arrival for r in results.values())
    t1 = max(r.finished for r in results.values())
    return {"jobs": len(lat), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "walker_steps_per_s": ws / (t1 - t0), "makespan_s": t1 - t0}


def sweep(rates, n_jobs, steps=100, walkers=(256, 1024, 4096), precision="fp16x3", seed=0, max_active=8,
          mem_budget=200 << 30, log=None, progress=None):
    """Serve a seeded Poisson stream of random-molecule jobs at each arrival
    rate and summarise latency and throughput (the serving workload of the
    benchmark: 2-42 electrons, FermiNet 4x256/32, 16 determinants).  With
    fewer than 100 jobs per rate the p99 is reported together with the
    maximum it approaches."""
    out = []
    for r in rates:
        jobs = make_job_stream(n_jobs, rate=r, seed=seed, steps=steps, walkers=walkers)
        t0 = time.perf_counter()
        for j in jobs:
            j.arrival += t0
        done = (lambda x: progress(f"rate {r}: job {x.job_id} ({x.n_elec} e, {x.walkers} walkers) "
                                   f"latency {x.latency:.2f} s")) if progress else None
        srv = Server(lambda j: GpuRunner(j, precision=precision), mem_budget=mem_budget, max_active=max_active,
                     on_done=done)
        res = srv.run(jobs)
        srv.close()
        s = latency_summary(res)
        lat = np.array([x.latency for x in res.values()])
        s.update({"rate_jobs_per_s": r, "p90_s": float(np.percentile(lat, 90)), "max_s": float(lat.max()),
                  "p99_is_near_max": n_jobs < 100, "peak_active": srv.peak_active,
                  "walker_steps": int(sum(x.walkers * x.steps for x in res.values()))})
        out.append(s)
        if log:
            log(s)
    return out


if __name__ == "__main__":
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description="serving sweep over Poisson arrival rates")
    ap.add_argument("--rates", default="1,5,50")
    ap.add_argument("--jobs", type=int, default=24)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--precision", default="fp16x3")
    ap.add_argument("--max-active", type=int, default=8)
    a = ap.parse_args()
    res = sweep([float(r) for r in a.rates.split(",")], a.jobs, steps=a.steps, precision=a.precision,
                max_active=a.max_active, log=lambda s: print(json.dumps(s), file=sys.stderr, flush=True),
                progress=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(res))
