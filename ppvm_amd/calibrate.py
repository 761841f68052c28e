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


def main(argv) -> int:
    if len(argv) >= 3 and argv[0] == "summarize":
        return summarize(argv[1], argv[2])
    return run_copy()


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
