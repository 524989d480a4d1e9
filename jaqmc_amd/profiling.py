"""Summaries of rocprofv3 result databases (SQLite, ``-o run`` output).

``kernel_stats`` turns a ``--kernel-trace --stats`` database into the per-kernel
table kept under ``profiles/``; ``traffic`` turns two ``--pmc FETCH_SIZE`` /
``--pmc WRITE_SIZE`` databases into HBM bytes per dispatch of one kernel,
calibrated against a plain device copy of known size recorded in the same runs
(the counters are reported in KiB; FETCH_SIZE on gfx950 reads half the bytes a
1 GiB copy moves, so both counters are scaled by copy bytes / counter value).

    python -m jaqmc_amd.profiling stats  <run_results.db> "<command>" "<title>"
    python -m jaqmc_amd.profiling traffic <fetch.db> <write.db> <kernel-substring> <copy-bytes> "<commands>" "<title>"
"""

from __future__ import annotations

import collections
import sqlite3
import sys


def _short(name: str) -> str:
    s = name.replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]
    return s if len(s) <= 90 else s[:90] + "..."


def kernel_stats(db: str):
    con = sqlite3.connect(db)
    rows = list(con.execute("select name,total_calls,total_duration,average,percentage from top_kernels"))
    # the top_kernels view reports microseconds
    return [(_short(n), int(c), t / 1e3, a / 1e3, p) for n, c, t, a, p in rows]


def _per_kernel(db: str, counter: str):
    con = sqlite3.connect(db)
    d = collections.defaultdict(list)
    for name, val in con.execute("select kernel_name, value from counters_collection where counter_name = ?",
                                 (counter,)):
        d[_short(name)].append(float(val))
    return d


def traffic(fetch_db: str, write_db: str, kernel: str, copy_bytes: int):
    """HBM bytes per dispatch of the kernels whose name contains ``kernel``.

    The calibration dispatch is the largest ``__amd_rocclr_copyBuffer`` in each
    run (a device-to-device copy of ``copy_bytes`` that reads and writes
    ``copy_bytes`` each)."""
    out = {}
    for counter, db in (("FETCH_SIZE", fetch_db), ("WRITE_SIZE", write_db)):
        d = _per_kernel(db, counter)
        cal = max(d.get("__amd_rocclr_copyBuffer", [0.0]))
        if cal <= 0:
            raise RuntimeError(f"{db}: no copy dispatch to calibrate {counter}")
        scale = copy_bytes / (cal * 1024.0)
        vals = [v for k, vs in d.items() if kernel in k for v in vs]
        if not vals:
            raise RuntimeError(f"{db}: no dispatch of {kernel}")
        out[counter] = {"kib_per_dispatch": sum(vals) / len(vals), "dispatches": len(vals), "scale": scale,
                        "bytes_per_dispatch": sum(vals) / len(vals) * 1024.0 * scale}
    out["bytes_per_dispatch"] = out["FETCH_SIZE"]["bytes_per_dispatch"] + out["WRITE_SIZE"]["bytes_per_dispatch"]
    return out


def main(argv):
    if argv[0] == "stats":
        db, cmd, title = argv[1:4]
        print(f"# {title}\n")
        print(f"Command (from the repo root, on one MI355X):\n\n    {cmd}\n")
        print("Kernel-trace statistics (durations in milliseconds; rocprofv3 --stats).\n")
        print(f"{'calls':>6} {'total_ms':>12} {'avg_ms':>10} {'pct':>6}  kernel")
        for n, c, t, a, p in kernel_stats(db):
            print(f"{c:6d} {t:12.1f} {a:10.3f} {p:6.2f}  {n}")
    elif argv[0] == "traffic":
        fdb, wdb, kernel, copy_bytes, cmds, title = argv[1:7]
        t = traffic(fdb, wdb, kernel, int(copy_bytes))
        print(f"# {title}\n")
        print(f"Commands (from the repo root, on one MI355X):\n\n{cmds}\n")
        print(f"Kernel: dispatches whose name contains `{kernel}`.\n")
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            r = t[c]
            print(f"{c}: {r['kib_per_dispatch']:.1f} KiB per dispatch (mean of {r['dispatches']}), copy-calibrated "
                  f"scale {r['scale']:.4f} -> {r['bytes_per_dispatch'] / 1e9:.4f} GB")
        print(f"\ntraffic_bytes_per_dispatch: {t['bytes_per_dispatch']:.6e}")
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
