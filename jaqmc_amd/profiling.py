"""Summaries of rocprofv3 result databases (SQLite, ``-o run`` output).

``kernel_stats`` turns a ``--kernel-trace --stats`` database into the per-kernel
table kept under ``profiles/``; ``traffic`` turns two ``--pmc FETCH_SIZE`` /
``--pmc WRITE_SIZE`` databases into HBM bytes per dispatch of one kernel,
calibrated against a plain device copy of known size recorded in the same runs
(the counters are reported in KiB; FETCH_SIZE on gfx950 reads half the bytes a
1 GiB copy moves, so both counters are scaled by copy bytes / counter value).

    python -m jaqmc_amd.profiling stats  <run_results.db> "<command>" "<title>"
    python -m jaqmc_amd.profiling traffic <fetch.db> <write.db> <kernel-substring> <copy-bytes> "<commands>" "<title>" [<exclusions>]

Exclusions (comma-separated) name dispatches of another launch class of the
same kernel instantiation, such as the orbital projection beside the layer
GEMMs: `g<grid size>` or `d<ordinal>` (0-based among the matching
dispatches).  They are listed but not averaged.
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


def _dispatches(db: str, counter: str, kernel: str):
    """(grid size, counter value) of every dispatch of the matching kernels, in dispatch order."""
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, grid_size, value from counters_collection where counter_name = ? "
                       "order by dispatch_id", (counter,))
    return [(int(g), float(v)) for n, g, v in rows if kernel in _short(n)]


def _excluded(i: int, grid: int, exclude) -> bool:
    return f"g{grid}" in exclude or f"d{i}" in exclude


def traffic(fetch_db: str, write_db: str, kernel: str, copy_bytes: int, exclude=()):
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
        each = _dispatches(db, counter, kernel)
        vals = [v for i, (g, v) in enumerate(each) if not _excluded(i, g, exclude)]
        if not vals:
            raise RuntimeError(f"{db}: no dispatch of {kernel}")
        out[counter] = {"kib_per_dispatch": sum(vals) / len(vals), "dispatches": len(vals), "scale": scale,
                        "bytes_per_dispatch": sum(vals) / len(vals) * 1024.0 * scale,
                        "each": [(g, v * 1024.0 * scale) for g, v in each]}
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
        exclude = tuple(argv[7].split(",")) if len(argv) > 7 and argv[7] else ()
        t = traffic(fdb, wdb, kernel, int(copy_bytes), exclude)
        print(f"# {title}\n")
        print(f"Commands (from the repo root, on one MI355X):\n\n{cmds}\n")
        print(f"Kernel: dispatches whose name contains `{kernel}`"
              + (f"; dispatches {', '.join(exclude)} (g: grid size, d: ordinal) listed but not averaged"
                 if exclude else "") + ".\n")
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            r = t[c]
            print(f"{c}: {r['kib_per_dispatch']:.1f} KiB per dispatch (mean of {r['dispatches']}), copy-calibrated "
                  f"scale {r['scale']:.4f} -> {r['bytes_per_dispatch'] / 1e9:.4f} GB")
        f, w = t["FETCH_SIZE"]["each"], t["WRITE_SIZE"]["each"]
        if len(f) == len(w):
            print("\nPer dispatch (grid size in work-items; GB):\n")
            print(f"{'grid':>10} {'fetch':>8} {'write':>8}")
            for i, ((g, fb), (_, wb)) in enumerate(zip(f, w)):
                print(f"{g:10d} {fb / 1e9:8.3f} {wb / 1e9:8.3f}" + ("  (not averaged)" if _excluded(i, g, exclude) else ""))
        print(f"\ntraffic_bytes_per_dispatch: {t['bytes_per_dispatch']:.6e}")
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
