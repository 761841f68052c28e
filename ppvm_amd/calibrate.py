"""HBM traffic counters: calibration workload and counter summaries.

``python -m ppvm_amd.calibrate`` runs one elementwise kernel of known traffic
(b = a * 1 over 2^28 float32: 1 GiB read, 1 GiB written) three times.
Profiled with ``rocprofv3 --pmc FETCH_SIZE`` and then ``--pmc WRITE_SIZE``, the
per-dispatch counter values of this kernel divided by 1 GiB give the scale
factors that turn the same counters' readings for the propagation kernels into
bytes.

``python -m ppvm_amd.calibrate summarize DIR COUNTER`` reads the
``*counter_collection.csv`` files rocprofv3 wrote under DIR and prints, per
kernel, the dispatch count and the mean / total counter value.

``python -m ppvm_amd.calibrate table CAL_FETCH CAL_WRITE FETCH WRITE`` turns
four such summaries (calibration and workload, FETCH_SIZE and WRITE_SIZE) into
HBM bytes per dispatch: each counter is scaled by 1 GiB over what it read for
the calibration kernel.
"""

from __future__ import annotations

import csv
import glob
import json
import os
import re
import sys
from collections import defaultdict

N = 1 << 28


def run_copy() -> int:
    import torch
    a = torch.ones(N, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    assert float(b[-1]) == 1.0
    print(json.dumps({"kernel": "elementwise mul (b = a * 1)", "elements": N,
                      "bytes_read": 4 * N, "bytes_written": 4 * N, "dispatches": 3}))
    return 0


def short_name(name: str) -> str:
    m = re.search(r"\b(k_[A-Za-z0-9_]+)", name)
    if m:
        return m.group(1)
    return name.split("(")[0][-80:]


def summarize(root: str, counter: str) -> int:
    files = sorted(glob.glob(os.path.join(root, "**", "*counter_collection.csv"), recursive=True))
    if not files:
        print(f"no counter_collection.csv under {root}", file=sys.stderr)
        return 1
    acc = defaultdict(lambda: [0, 0.0])
    for f in files:
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                a = acc[short_name(row["Kernel_Name"])]
                a[0] += 1
                a[1] += float(row["Counter_Value"])
    print(f"# {counter} per kernel (rocprofv3 counter values, summed over the dispatch)")
    print(f"{'kernel':<40} {'dispatches':>10} {'mean':>16} {'total':>18}")
    for k, (n, tot) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        print(f"{k:<40} {n:>10d} {tot / n:>16.1f} {tot:>18.1f}")
    return 0


def read_summary(path: str) -> dict:
    """kernel -> (dispatches, mean counter value) from a `summarize` output."""
    out = {}
    with open(path) as f:
        for line in f:
            if line.startswith("#") or line.startswith("kernel"):
                continue
            parts = line.split()
            if len(parts) < 4:
                continue
            out[" ".join(parts[:-3])] = (int(parts[-3]), float(parts[-2]))
    return out


def _calib(path: str) -> float:
    rows = read_summary(path)
    name = next(k for k in rows if "MulFunctor" in k)
    return (4 * N / 1024.0) / rows[name][1]  # KiB moved per counter unit


def table(cal_fetch: str, cal_write: str, fetch: str, write: str) -> int:
    kf, kw = _calib(cal_fetch), _calib(cal_write)
    f, w = read_summary(fetch), read_summary(write)
    print(f"# HBM traffic per dispatch: FETCH_SIZE x {kf:.4f} + WRITE_SIZE x {kw:.4f} (KiB -> bytes x 1024);")
    print("# scale factors = 1 GiB / counter value of the calibration kernel (b = a * 1 over 2^28 float32)")
    print(f"{'kernel':<24} {'dispatches':>10} {'fetch_KiB':>12} {'write_KiB':>12} {'hbm_bytes':>14}")
    for k in sorted(f, key=lambda k: -f[k][1] * f[k][0]):
        if not k.startswith("k_") or k not in w:
            continue
        n, fk = f[k]
        wk = w[k][1]
        print(f"{k:<24} {n:>10d} {fk:>12.1f} {wk:>12.1f} {(fk * kf + wk * kw) * 1024:>14.0f}")
    return 0


def main(argv) -> int:
    if len(argv) >= 3 and argv[0] == "summarize":
        return summarize(argv[1], argv[2])
    if len(argv) >= 5 and argv[0] == "table":
        return table(*argv[1:5])
    return run_copy()


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
