"""Summaries of rocprofv3 result databases (rocpd SQLite) as short text tables.

    python -m hamiltonfastmarching_amd.profiling trace  <results.db> [--filter SUBSTRING]
    python -m hamiltonfastmarching_amd.profiling counters <fetch_results.db> <write_results.db> \
        [--kernel k_fim] [--copy-bytes N]

`trace` prints per-kernel call counts and durations (from --kernel-trace).  `counters` prints
the mean FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of the named kernel and of the plain copy
kernel, the scale factors copy_bytes / counter for each counter, and the scaled HBM bytes per
dispatch of the named kernel (emitted as JSON on the last line).
"""
from __future__ import annotations

import json
import sqlite3
import sys

import numpy as np


def kernel_trace(db: str, substring: str = ""):
    c = sqlite3.connect(db)
    rows = c.execute("select name, duration from kernels").fetchall()
    by = {}
    for name, d in rows:
        if substring and substring not in name:
            continue
        by.setdefault(name, []).append(d)
    out = []
    for name, ds in by.items():
        a = np.array(ds, dtype=float) / 1e3  # ns -> us
        out.append((a.sum(), name, len(a), a.mean(), a.min(), np.median(a), a.max()))
    out.sort(reverse=True)
    lines = ["| kernel | calls | total us | avg us | min us | median us | max us |", "|---|---|---|---|---|---|---|"]
    for tot, name, n, avg, mn, med, mx in out:
        short = name.replace("(anonymous namespace)::", "").split("(")[0]
        lines.append(f"| `{short}` | {n} | {tot:.1f} | {avg:.2f} | {mn:.2f} | {med:.2f} | {mx:.2f} |")
    return "\n".join(lines)


def counter_values(db: str, counter: str):
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, value from counters_collection where counter_name = ?", (counter,)).fetchall()
    by = {}
    for name, v in rows:
        by.setdefault(name, []).append(v)
    return by


def counters(fetch_db: str, write_db: str, kernel: str = "k_fim", copy_bytes: float = float(1 << 30)):
    f = counter_values(fetch_db, "FETCH_SIZE")
    w = counter_values(write_db, "WRITE_SIZE")

    def pick(by, key):
        vals = [v for name, vs in by.items() if key in name for v in vs]
        return np.array(vals, dtype=float)

    kf, kw = pick(f, kernel), pick(w, kernel)
    # the plain copy: the largest copyBuffer / elementwise dispatch
    cf = max((max(vs) for name, vs in f.items() if "copy" in name.lower() or "elementwise" in name.lower()), default=0)
    cw = max((max(vs) for name, vs in w.items() if "copy" in name.lower() or "elementwise" in name.lower()), default=0)
    fs = copy_bytes / (cf * 1024) if cf else float("nan")
    ws = copy_bytes / (cw * 1024) if cw else float("nan")
    per_launch = (kf.mean() * fs + kw.mean() * ws) * 1024 if len(kf) and len(kw) else float("nan")
    res = {
        "kernel": kernel, "dispatches_fetch": int(len(kf)), "dispatches_write": int(len(kw)),
        "fetch_kib_mean": float(kf.mean()) if len(kf) else None, "write_kib_mean": float(kw.mean()) if len(kw) else None,
        "copy_fetch_kib": float(cf), "copy_write_kib": float(cw), "copy_bytes": copy_bytes,
        "fetch_scale": fs, "write_scale": ws, "hbm_bytes_per_launch": per_launch,
    }
    return res


def main(argv):
    if len(argv) >= 2 and argv[0] == "trace":
        sub = argv[argv.index("--filter") + 1] if "--filter" in argv else ""
        print(kernel_trace(argv[1], sub))
    elif len(argv) >= 3 and argv[0] == "counters":
        kernel = "k_fim"
        copy_bytes = float(1 << 30)
        if "--kernel" in argv:
            kernel = argv[argv.index("--kernel") + 1]
        if "--copy-bytes" in argv:
            copy_bytes = float(argv[argv.index("--copy-bytes") + 1])
        r = counters(argv[1], argv[2], kernel, copy_bytes)
        print(json.dumps(r))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
