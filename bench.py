#!/usr/bin/env python
"""Walker-steps/s of the per-walker VMC hot path on one MI355X.

One *walker-step* = one all-electron Metropolis move of a walker (Gaussian
proposal, log|psi| at the proposal, accept/reject) followed by its local energy
E_L = -1/2 (lap log|psi| + |grad log|psi||^2) + V at the post-move position,
plus the FP64 mean / variance of E_L over the batch -- the reference's
EvaluationWorkStage.compute_step with MCMCSampler(steps=1).  A timed *step*
is one walker-step of the whole batch; throughput = B * steps / wall.

Default workload: the 42-electron planar ring (6 C on a radius-2.64 hexagon,
6 H at 4.68 bohr), Psiformer with 4 layers, 4 heads, model width 256, MLP
[256], 16 determinants, Jastrow, 2048 walkers, proposal width 0.02; seeded
random parameters (synthetic data; no checkpoint exists for this system).
Both the plain FP32 path and the mixed path are timed; `value` is the mixed
path.  Beside them: the configuration's protocol time (S Metropolis steps then
one local energy), the paths' paired errors against FP64 on oracle-equilibrated
walkers with their accept/reject agreement (committed fixtures), the roofline
of the dominant kernel, and the CPU baseline.

Prints ONE JSON line on stdout; diagnostics go to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))


def measured_traffic(workload: str, mixed: str):
    """HBM bytes per dispatch of the mixed-precision layer GEMM of this workload,
    from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated;
    see jaqmc_amd/profiling.py), or None when no summary is committed."""
    for name in (f"gemm_layer_lap_{workload}_{mixed}_traffic.txt", f"gemm_layer_lap_{workload}_traffic.txt"):
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            continue
        with open(path) as f:
            txt = f.read()
        if mixed not in txt and name.endswith(f"{workload}_traffic.txt") and mixed != "bf16x6":
            continue  # the generic summary is the bf16x6 kernel's
        for line in txt.splitlines():
            if line.startswith("traffic_bytes_per_dispatch:"):
                return float(line.split(":", 1)[1]), "profiles/" + name
    return None, None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_path(spec, system, B, width, params, x0, precision, steps, warmup, S_steps, chunk=0):
    import torch

    from jaqmc_amd import _lib
    from jaqmc_amd.device import DeviceModel, reduce_mean_var

    dm = DeviceModel(spec, system.atoms, system.charges, params, precision=precision, chunk_walkers=chunk)
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
    # the configuration's protocol: S Metropolis steps, then one local energy
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dm.mh_steps(x, logp, S_steps, width, seed=99, ctr=0, n_accept=acc)
    torch.cuda.synchronize()
    t_mh = time.perf_counter() - t0
    t0 = time.perf_counter()
    out = dm.local_energy(x)
    reduce_mean_var(out["e_loc"])
    torch.cuda.synchronize()
    t_el = time.perf_counter() - t0
    res["protocol"] = {"mh_steps": S_steps, "mh_s": t_mh, "local_energy_s": t_el,
                       "walker_steps_per_s": B * S_steps / (t_mh + t_el)}
    dm.close()
    return res


def paired_stats(a, b):
    """Per-walker differences a - b (Hartree in, mHa out): the mean with its
    standard error, a 1%-trimmed mean, and the spread of |a - b|."""
    d = 1e3 * (np.asarray(a, np.float64) - np.asarray(b, np.float64))
    n = d.size
    keep = np.argsort(np.abs(d))[: max(1, int(np.floor(0.99 * n)))]
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1.0)
    return {"n": int(n), "mean_mHa": float(d.mean()),
            "stderr_mHa": float(d.std(ddof=1) / np.sqrt(n)) if n > 1 else None,
            "trimmed1pct_mean_mHa": float(d[keep].mean()),
            "median_abs_mHa": float(np.median(np.abs(d))), "p99_abs_mHa": float(np.quantile(np.abs(d), 0.99)),
            "max_abs_mHa": float(np.abs(d).max()), "median_rel": float(np.median(rel)),
            "p99_rel": float(np.quantile(rel, 0.99)), "max_rel": float(rel.max()),
            "frac_within_1e-5_rel": float(np.mean(rel <= 1e-5))}


def fixture_accuracy(spec, system, params, workload, precisions, width):
    """Paired E_L errors against FP64 and accept/reject agreement on the
    committed oracle-equilibrated walkers (tests/golden/equilibrated_<w>.npz,
    written by tests/golden/make_equilibrated.py in the build container)."""
    import torch

    from jaqmc_amd.device import DeviceModel

    path = os.path.join(ROOT, "tests", "golden", f"equilibrated_{workload}.npz")
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as f:
        fx = {k: f[k] for k in f.files}
    x = torch.from_numpy(np.ascontiguousarray(fx["x"])).cuda()
    n = x.shape[0]
    info = {"walkers": int(n), "source": "tests/golden/" + os.path.basename(path),
            "equilibration": f"{int(fx['steps'])} FP64 oracle Metropolis steps",
            "fp64_mean_Ha": float(fx["e_loc"].mean())}
    for prec in precisions:
        dm = DeviceModel(spec, system.atoms, system.charges, params, precision=prec)
        e = dm.local_energy(x)["e_loc"].cpu().numpy()
        st = paired_stats(e, fx["e_loc"])
        # one Metropolis move with the fixture's normals and uniforms
        xm = x.clone()
        logp = 2.0 * dm.logpsi(xm)[0]
        acc = dm.mh_steps(xm, logp, 1, width, normals=torch.from_numpy(fx["mh_normals"]).cuda(),
                          uniforms=torch.from_numpy(fx["mh_uniforms"]).cuda())
        a = acc.cpu().numpy().astype(bool)
        amb = np.abs(fx["mh_ratio"] - np.log(fx["mh_uniforms"].astype(np.float64))) < 1e-6
        st["accept_agreement"] = {"decisions": int(n), "excluded_within_1e-6": int(amb.sum()),
                                  "agree": int(np.sum((a == fx["mh_accept"]) & ~amb)),
                                  "disagree": int(np.sum((a != fx["mh_accept"]) & ~amb))}
        info[f"{prec}_vs_fp64"] = st
        dm.close()
    return info


def cpu_baseline(spec, system, params, x, width):
    """Oracle (numpy FP32 restatement) walker-steps/s on a bounded sample,
    with the BLAS thread count set explicitly."""
    from threadpoolctl import threadpool_limits

    from oracle import hamiltonian as ham
    from oracle import wavefunction as owf

    try:
        import psutil

        phys = psutil.cpu_count(logical=False) or os.cpu_count() or 1
    except Exception:
        phys = os.cpu_count() or 1
    threads = int(min(phys, 16))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    m = spec.as_oracle()
    rng = np.random.default_rng(7)
    with threadpool_limits(limits=threads):
        _, lp = owf.logpsi(m, params, x, system.atoms, with_laplacian=False)
        normals = rng.standard_normal(x.shape).astype(np.float32)
        u = rng.uniform(size=x.shape[0]).astype(np.float32)
        t0 = time.perf_counter()
        xn, _, _, _, _ = ham.mh_step(m, params, x, 2 * lp.v, system.atoms, width, normals, u)
        ham.local_energy(m, params, xn, system.atoms, system.charges)
        t = time.perf_counter() - t0
    return x.shape[0] / t, t, threads, phys, model


CPU_SAMPLE = {"li": 131072, "n2": 512, "ring42": 12, "h60": 4}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ring42", choices=["li", "n2", "ring42", "h60"])
    ap.add_argument("--mixed", default="fp16x3", choices=["fp16x3", "bf16x6", "bf16x3", "bf16"])
    ap.add_argument("--cpu-sample", type=int, default=0, help="walkers in the CPU-baseline sample (0: per workload)")
    ap.add_argument("--workspace-gb", type=float, default=128.0,
                    help="device memory per model for walker-chunk workspace (sets the chunk size; 0: library default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--serving-jobs", type=int, default=12, help="jobs in the serving sample (0: skip)")
    ap.add_argument("--serving-rate", type=float, default=5.0, help="Poisson arrival rate of the serving sample")
    ap.add_argument("--serving-steps", type=int, default=20, help="evaluation iterations per serving job")
    ap.add_argument("--vjp-walkers", type=int, default=0,
                    help="time the parameter-gradient pass (jq_logpsi_vjp, FP32) on this many walkers of the "
                         "workload instead of the evaluation step; prints its own JSON line")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="counter-calibration run for rocprofv3 --pmc: one 1 GiB device copy, then one local-energy "
                         "pass of the workload in the mixed precision; prints nothing on stdout")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); walkers shard with no exchange but "
              "a single one-GPU figure is what this build reports.", file=sys.stderr)
        sys.exit(2)

    sys.path.insert(0, ROOT)
    import torch

    from jaqmc_amd import params as P
    from jaqmc_amd import roofline as RF
    from jaqmc_amd import systems as S

    system, spec, B, S_steps, width = S.workload(args.workload)
    params = P.init_params(spec, seed=0)
    x0 = S.init_walkers(system, B, seed=0)
    # walkers per workspace chunk: the batch in equal chunks within the
    # workspace budget (HBM is large enough to hold the whole 42-electron
    # batch's tangent rows at once; fewer chunks leave fewer partial rounds)
    chunk = 0
    if args.workspace_gb > 0:
        from jaqmc_amd.device import chunk_for_budget

        chunk = chunk_for_budget(spec, system.atoms, system.charges, params, args.mixed, B,
                                 args.workspace_gb * (1 << 30))
    log(f"walkers per workspace chunk: {chunk or 'library default'}")
    if args.vjp_walkers > 0:
        from jaqmc_amd.device import DeviceModel

        nv = args.vjp_walkers
        xv = torch.from_numpy(S.init_walkers(system, nv, seed=3)).cuda()
        cot = torch.linspace(-1.0, 1.0, nv, dtype=torch.float64, device="cuda")
        dm = DeviceModel(spec, system.atoms, system.charges, params, with_gradients=True)
        for _ in range(args.warmup):
            g = dm.logpsi_vjp(xv, cot)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g = dm.logpsi_vjp(xv, cot)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        ok = bool(torch.isfinite(g).all())
        dm.close()
        print(json.dumps({"metric": "parameter_gradient_walkers_per_s", "value": nv / dt, "unit": "walkers/s",
                          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
                          "higher_is_better": True, "dtype": "fp32", "data": "synthetic",
                          "config": {"workload": args.workload, "walkers": nv, "parameters": int(g.numel())},
                          "finite": ok}))
        return
    if args.traffic_probe:
        from jaqmc_amd.device import DeviceModel

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB read + 1 GiB written
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        dm = DeviceModel(spec, system.atoms, system.charges, params, precision=args.mixed, chunk_walkers=chunk)
        dm.local_energy(torch.from_numpy(x0).cuda())
        torch.cuda.synchronize()
        log("traffic probe done: copy 1 GiB, one local-energy pass")
        return
    log(f"workload {args.workload}: N={spec.n_elec} M={system.atoms.shape[0]} K={spec.ndets} B={B} "
        f"device={torch.cuda.get_device_name(0)}")

    res = {}
    for prec in ("fp32", args.mixed):
        res[prec] = run_path(spec, system, B, width, params, x0, prec, args.steps, args.warmup, S_steps, chunk)
        log(prec, {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res[prec].items() if k != "prof"})
        log("  prof (ms total, launches):", res[prec]["prof"])

    acc_info = None
    if not args.no_accuracy:
        acc_info = fixture_accuracy(spec, system, params, args.workload, ("fp32", args.mixed), width)
        log("accuracy:", json.dumps(acc_info))

    # --- roofline of the dominant kernel class (layer GEMMs in Laplacian mode) ---
    # achieved = algorithmic FLOPs of all layer-GEMM launches in the timed region
    # / their summed device time (HIP events on the launching stream)
    fl = RF.flops(spec, system.atoms.shape[0])
    mixed = res[args.mixed]
    tot_ms, n = mixed["prof"]["gemm_layer_lap"]
    achieved = fl["gemm_layer_lap"] * B * args.steps / (tot_ms * 1e-3) if n else 0.0
    peak = RF.PEAKS[RF.MODE_PEAK[args.mixed]]
    mult = RF.MFMA_PER_PRODUCT[args.mixed]
    alg_bytes = RF.layer_bytes(spec, system.atoms.shape[0], args.mixed) * B * args.steps
    traffic, traffic_src = measured_traffic(args.workload, args.mixed)
    avg_s = tot_ms * 1e-3 / max(n, 1)
    # the binding roof: the larger of the minimum MFMA time (algorithmic
    # FLOPs x MFMAs per product at the mode's dense peak) and the minimum HBM
    # time (algorithmic bytes at 8 TB/s)
    t_mfma = fl["gemm_layer_lap"] * B * args.steps * mult / peak
    t_hbm = alg_bytes / RF.PEAKS["hbm"]
    if t_hbm >= t_mfma and n:
        bw = alg_bytes / (tot_ms * 1e-3)
        head = {"bound": "hbm", "achieved": bw / 1e9, "peak": RF.PEAKS["hbm"] / 1e9, "unit": "GB/s",
                "frac": bw / RF.PEAKS["hbm"]}
    else:
        head = {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": achieved / peak}
    roof = {**head, "traffic": traffic,
            "algorithmic_tflops": achieved / 1e12, "mfma_peak_tflops": peak / 1e12, "mfma_frac": achieved / peak,
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
        nc = args.cpu_sample or CPU_SAMPLE[args.workload]
        # a sample larger than the workload's walker count (Li) draws more
        # walkers from the same initialisation
        xs = x0[:nc] if nc <= B else S.init_walkers(system, nc, seed=11)
        v, t, threads, phys, model = cpu_baseline(spec, system, params, xs, width)
        cpu = {"value": v, "unit": "walker-steps/s", "cores": threads, "kind": "port",
               "sample": f"{nc} walkers x 1 walker-step (FP32 numpy oracle: MH move + dense-Jacobian E_L), "
                         f"{t:.1f}s, BLAS threads {threads}",
               "cpu_model": model, "physical_cores": phys, "machine": platform.machine()}

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

    names = {"li": "C1 Li atom", "n2": "C2 N2-like diatomic", "ring42": "C3 42-electron planar ring",
             "h60": "C4 linear H60 chain"}
    line = {
        "metric": f"walker-steps/s incl. E_L ({names[args.workload]}, {arch}; {args.mixed}-mixed, FP32 beside it)",
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
        "config": {"workload": f"{names[args.workload]}: {system.name} {arch} K={spec.ndets}", "walkers": B,
                   "electrons": spec.n_elec, "mh_width": width, "protocol_mh_steps": S_steps,
                   "walkers_per_chunk": chunk or None},
        "fp32_value": res["fp32"]["value"],
        "fp32_ms_per_step": res["fp32"]["ms_per_step"],
        "mixed_over_fp32": mixed["value"] / res["fp32"]["value"],
        "protocol": {"mixed": mixed["protocol"], "fp32": res["fp32"]["protocol"],
                     "note": "the configuration's literal sequence: S Metropolis steps, then one local energy"},
        "accuracy": acc_info,
        "roofline": roof,
        "cpu_baseline": cpu,
        "serving": serving_info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
