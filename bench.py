#!/usr/bin/env python
"""Walker-steps/s of the per-walker VMC hot path on one MI355X.

One *walker-step* = one all-electron Metropolis move of a walker (Gaussian
proposal, log|psi| at the proposal, accept/reject) followed by its local energy
E_L = -1/2 (lap log|psi| + |grad log|psi||^2) + V at the post-move position,
plus the FP64 mean / variance of E_L over the batch -- the reference's
EvaluationWorkStage.compute_step with MCMCSampler(steps=1).  A timed *step*
is one walker-step of the whole batch; throughput = B * steps / wall.

Default workload: N2-like diatomic, 14 electrons (7/7), FermiNet 4 layers
256/32, 16 determinants, 4096 walkers, proposal width 0.02, seeded random
parameters (synthetic data; no checkpoint exists for this system).

Prints ONE JSON line on stdout; diagnostics go to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def measured_traffic(workload: str):
    """HBM bytes per dispatch of the mixed-precision layer GEMM of this workload,
    from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated;
    see jaqmc_amd/profiling.py), or None when no summary is committed."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"gemm_layer_lap_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_dispatch:"):
                return float(line.split(":", 1)[1]), "profiles/" + os.path.basename(path)
    return None, None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_path(spec, system, B, width, params, x0, precision, steps, warmup):
    import torch

    from jaqmc_amd import _lib
    from jaqmc_amd.device import DeviceModel, reduce_mean_var

    dm = DeviceModel(spec, system.atoms, system.charges, params, precision=precision)
    x = torch.from_numpy(x0).cuda()
    lp, _ = dm.logpsi(x)
    logp = 2.0 * lp
    acc = torch.zeros(B, dtype=torch.int32, device=x.device)
    ctr = [0]

    def step():
        dm.mh_steps(x, logp, 1, width, seed=1234, ctr=ctr[0], n_accept=acc)
        ctr[0] += 1
        out = dm.local_energy(x)
        mv = reduce_mean_var(out["e_loc"])
        return out, mv

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        out, mv = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _lib.profile_enable(False)
    prof = {k: _lib.profile_read(k) for k in ("gemm_layer_lap", "det_lap", "gemm_orb_lap", "pair_lap", "attn_lap",
                                               "forward", "local_energy")}
    wall = t1 - t0
    mv = mv.cpu().numpy()
    res = dict(wall=wall, ms_per_step=1e3 * wall / steps, value=B * steps / wall, prof=prof,
               mean=float(mv[0]), var=float(mv[1]), pmove=float(acc.sum().item()) / (B * (steps + warmup)))
    res["x_final"] = x.cpu().numpy()
    dm.close()
    return res


def paired_stats(a, b):
    """Per-walker differences a - b (Hartree in, mHa out): the mean with its
    standard error, a 1%-trimmed mean, and the spread of |a - b|.  E_L has heavy
    tails (walkers near nodes or coalescences), so the mean alone is dominated by
    a few walkers."""
    d = 1e3 * (np.asarray(a, np.float64) - np.asarray(b, np.float64))
    n = d.size
    keep = np.argsort(np.abs(d))[: max(1, int(np.floor(0.99 * n)))]
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1.0)
    return {"n": int(n), "mean_mHa": float(d.mean()),
            "stderr_mHa": float(d.std(ddof=1) / np.sqrt(n)) if n > 1 else None,
            "trimmed1pct_mean_mHa": float(d[keep].mean()),
            "median_abs_mHa": float(np.median(np.abs(d))), "p99_abs_mHa": float(np.quantile(np.abs(d), 0.99)),
            "max_abs_mHa": float(np.abs(d).max()), "median_rel": float(np.median(rel)),
            "p99_rel": float(np.quantile(rel, 0.99))}


def energies_on(spec, system, params, x, precision):
    import torch

    from jaqmc_amd.device import DeviceModel

    dm = DeviceModel(spec, system.atoms, system.charges, params, precision=precision)
    out = dm.local_energy(torch.from_numpy(x).cuda())
    e = out["e_loc"].cpu().numpy()
    dm.close()
    return e


def cpu_baseline(spec, system, params, x, width):
    """Oracle (numpy FP32 restatement) walker-steps/s on a bounded sample."""
    from oracle import hamiltonian as ham
    from oracle import wavefunction as owf

    m = spec.as_oracle()
    rng = np.random.default_rng(7)
    t0 = time.perf_counter()
    _, lp = owf.logpsi(m, params, x, system.atoms, with_laplacian=False)
    t_init = time.perf_counter() - t0
    normals = rng.standard_normal(x.shape).astype(np.float32)
    u = rng.uniform(size=x.shape[0]).astype(np.float32)
    t0 = time.perf_counter()
    xn, _, _, _, _ = ham.mh_step(m, params, x, 2 * lp.v, system.atoms, width, normals, u)
    ham.local_energy(m, params, xn, system.atoms, system.charges)
    t = time.perf_counter() - t0
    del t_init
    return x.shape[0] / t, t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="n2", choices=["li", "n2", "ring42", "h60"])
    ap.add_argument("--mixed", default="bf16x6", choices=["bf16x6", "bf16x3", "bf16"])
    ap.add_argument("--cpu-sample", type=int, default=256, help="walkers in the CPU-baseline sample")
    ap.add_argument("--err-sample", type=int, default=64, help="walkers checked against the FP64 oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-jobs", type=int, default=12, help="jobs in the serving sample (0: skip)")
    ap.add_argument("--serving-rate", type=float, default=5.0, help="Poisson arrival rate of the serving sample")
    ap.add_argument("--serving-steps", type=int, default=20, help="evaluation iterations per serving job")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="counter-calibration run for rocprofv3 --pmc: one 1 GiB device copy, then one local-energy "
                         "pass of the workload in the mixed precision; prints nothing on stdout")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); walkers shard with no exchange but "
              "a single one-GPU figure is what this build reports.", file=sys.stderr)
        sys.exit(2)

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import torch

    from jaqmc_amd import params as P
    from jaqmc_amd import roofline as RF
    from jaqmc_amd import systems as S

    system, spec, B, S_steps, width = S.workload(args.workload)
    params = P.init_params(spec, seed=0)
    x0 = S.init_walkers(system, B, seed=0)
    if args.traffic_probe:
        from jaqmc_amd.device import DeviceModel

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB read + 1 GiB written
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        dm = DeviceModel(spec, system.atoms, system.charges, params, precision=args.mixed)
        dm.local_energy(torch.from_numpy(x0).cuda())
        torch.cuda.synchronize()
        log("traffic probe done: copy 1 GiB, one local-energy pass")
        return
    log(f"workload {args.workload}: N={spec.n_elec} M={system.atoms.shape[0]} K={spec.ndets} B={B} "
        f"device={torch.cuda.get_device_name(0)}")

    res = {}
    for prec in ("fp32", args.mixed):
        res[prec] = run_path(spec, system, B, width, params, x0, prec, args.steps, args.warmup)
        log(prec, {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res[prec].items()
                   if k not in ("prof", "x_final")})
        log("  prof (ms total, launches):", res[prec]["prof"])

    # --- accuracy: FP32 and mixed E_L against the FP64 oracle on a sample, and
    # mixed against FP32 on the full batch; on the initial walkers and on the
    # walkers after the timed Metropolis steps ---
    from oracle import hamiltonian as ham

    ns = min(args.err_sample, B)
    p64 = P.cast_params(params, np.float64)
    acc_info = {"note": "paired per-walker differences; E_L is heavy-tailed, see stderr and trimmed mean"}
    for tag, xs in (("initial_walkers", x0), (f"after_{args.warmup + args.steps}_mh_steps",
                                               res[args.mixed]["x_final"])):
        e32 = energies_on(spec, system, params, xs, "fp32")
        emx = energies_on(spec, system, params, xs, args.mixed)
        t0 = time.perf_counter()
        ref = ham.local_energy(spec.as_oracle(), p64, xs[:ns].astype(np.float64), system.atoms.astype(np.float64),
                               system.charges.astype(np.float64))["e_loc"]
        log(f"FP64 oracle on {ns} walkers ({tag}): {time.perf_counter() - t0:.1f}s")
        acc_info[tag] = {
            "fp32_vs_fp64_sample": paired_stats(e32[:ns], ref),
            "mixed_vs_fp64_sample": paired_stats(emx[:ns], ref),
            "mixed_vs_fp32_full_batch": paired_stats(emx, e32),
            "full_batch_mean_Ha": {"fp32": float(e32.mean()), "mixed": float(emx.mean()),
                                   "mc_stderr": float(e32.std(ddof=1) / np.sqrt(B))},
        }
    log("accuracy:", json.dumps(acc_info))

    # --- roofline of the dominant kernel class (layer GEMMs in Laplacian mode) ---
    # achieved = algorithmic FLOPs of all layer-GEMM launches in the timed region
    # / their summed device time (HIP events on the launching stream)
    fl = RF.flops(spec, system.atoms.shape[0])
    mixed = res[args.mixed]
    tot_ms, n = mixed["prof"]["gemm_layer_lap"]
    achieved = fl["gemm_layer_lap"] * B * args.steps / (tot_ms * 1e-3) if n else 0.0
    peak = RF.PEAKS["bf16"]
    # Context for the fraction: each FP32-accurate product costs `mult` BF16
    # MFMA products, so issued MFMA work is mult x the algorithmic FLOPs; the
    # minimum HBM bytes (operands once, outputs once) give the memory side.
    # Algorithmic intensity is ~60 FLOP/B, above the emulated ridge
    # (2.5 PFLOP/s / mult / 8 TB/s = 52 FLOP/B for bf16x6), so MFMA bounds it.
    mult = RF.MFMA_PER_PRODUCT[args.mixed]
    alg_bytes = RF.layer_bytes(spec, system.atoms.shape[0]) * B * args.steps
    traffic, traffic_src = measured_traffic(args.workload)
    avg_s = tot_ms * 1e-3 / max(n, 1)
    roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic,
            "traffic_unit": "bytes/launch", "traffic_source": traffic_src,
            "mfma_per_product": mult, "issued_mfma_frac": achieved * mult / peak,
            "alg_bytes_per_launch": alg_bytes / max(n, 1),
            "hbm_frac_alg_bytes": (alg_bytes / (tot_ms * 1e-3) / RF.PEAKS["hbm"]) if n else 0.0,
            "hbm_frac_measured_traffic": (traffic / avg_s / RF.PEAKS["hbm"]) if (n and traffic) else None,
            "kernel": f"gemm_kernel ({spec.kind} dense layers, Laplacian mode, {args.mixed})",
            "avg_launch_ms": tot_ms / max(n, 1), "launches": n,
            "per_class_ms_per_step": {k: v[0] / args.steps for k, v in mixed["prof"].items()},
            "fp32_per_class_ms_per_step": {k: v[0] / args.steps for k, v in res["fp32"]["prof"].items()}}

    cpu = None
    if not args.no_cpu:
        nc = min(args.cpu_sample, B)
        v, t = cpu_baseline(spec, system, params, x0[:nc], width)
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": v, "unit": "walker-steps/s", "cores": cores, "kind": "port",
               "sample": f"{nc} walkers x 1 walker-step (FP32 numpy oracle, MH move + E_L), {t:.1f}s"}

    arch = (f"FermiNet {len(spec.hidden_single)}x{spec.hidden_single[0]}/{spec.hidden_double[0]}"
            if spec.kind == "ferminet" else
            f"Psiformer L{spec.num_layers} H{spec.num_heads} D{spec.num_heads * spec.heads_dim}")
    serving_info = None
    if args.serving_jobs > 0:
        from jaqmc_amd import serving

        jobs = serving.make_job_stream(args.serving_jobs, rate=args.serving_rate, seed=1, steps=args.serving_steps,
                                       walkers=(256, 1024))
        t0 = time.perf_counter()
        for j in jobs:
            j.arrival += t0
        srv = serving.Server(lambda j: serving.GpuRunner(j, precision=args.mixed), mem_budget=200 << 30,
                             max_active=8)
        res_s = srv.run(jobs)
        serving_info = serving.latency_summary(res_s)
        serving_info.update({"rate_jobs_per_s": args.serving_rate, "steps_per_job": args.serving_steps,
                             "electrons": "uniform 2..42", "walkers": [256, 1024], "peak_active": srv.peak_active,
                             "model": f"FermiNet 4x256/32 K=16, {args.mixed}"})
        log("serving:", json.dumps(serving_info))

    line = {
        "metric": f"walker-steps/s incl. E_L ({system.name} {arch}, BF16-mixed; FP32 beside it)",
        "value": mixed["value"],
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.mixed}+fp64",
        "data": "synthetic (seeded random-init params, Gaussian walkers)",
        "config": {"workload": f"{system.name} {arch} K={spec.ndets}", "walkers": B,
                   "electrons": spec.n_elec, "mh_width": width},
        "fp32_value": res["fp32"]["value"],
        "fp32_ms_per_step": res["fp32"]["ms_per_step"],
        "mixed_over_fp32": mixed["value"] / res["fp32"]["value"],
        "accuracy": acc_info,
        "roofline": roof,
        "cpu_baseline": cpu,
        "serving": serving_info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
