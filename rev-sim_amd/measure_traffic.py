#!/usr/bin/env python3
"""HBM traffic per launch of the library's kernels, from rocprofv3 hardware counters.

Runs a short bench.py workload (20 + 20 steps of the 98,304-atom box after a 300-step
equilibration, plus five launches of the library's 256 MiB copy kernel) under rocprofv3, once
counting FETCH_SIZE and once WRITE_SIZE (each pass holds one TCC counter group), for FP32 and for
FP64 pair arithmetic, averages the counters per kernel over its dispatches and scales them by what
the same counters read for the copy kernel in that run, whose traffic is known exactly (256 MiB
read and 256 MiB written per launch). Writes
the summary (default gpurun_out/traffic_water32768.json; kept as profiles/traffic_water32768.json),
which bench.py reads for the roofline's `traffic` when the HIP sources still hash to the value
recorded there.

Run from the repository root on a machine with the GPU: python3 rev-sim_amd/measure_traffic.py [out]
"""
from __future__ import annotations

import csv
import glob
import hashlib
import json
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "gpurun_out", "traffic_water32768.json")
WORK = os.path.join(ROOT, "gpurun_out", "traffic")
COPY_MB = 256
BENCH_ARGS = ["bench.py", "--workload", "water32768", "--n-steps", "20", "--steps", "1", "--warmup", "0",
              "--equil", "300", "--no-cpu-baseline", "--no-fp64-check", "--no-mlp", "--no-fp64", "--no-long",
              "--calibrate-copy-mb", str(COPY_MB)]


def sources_digest() -> str:
    h = hashlib.sha256()
    for p in sorted(glob.glob(os.path.join(ROOT, "rev-sim_amd", "csrc", "*"))):
        with open(p, "rb") as f:
            h.update(os.path.basename(p).encode() + b"\0" + f.read())
    return h.hexdigest()


def short_name(name: str) -> str:
    """'void (anonymous namespace)::k_pair<(anonymous namespace)::P32, 16, true, false, false>(...)'
    -> 'k_pair<P32, 16, true, false, false>'."""
    n = name.replace("(anonymous namespace)::", "")
    n = re.sub(r"^void\s+", "", n)
    depth, cut = 0, len(n)
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    return n[:cut].strip()


def run_pass(counter: str, precision: str) -> dict:
    d = os.path.join(WORK, f"{counter.lower()}_{precision}")
    cmd = ["rocprofv3", "--pmc", counter, "--kernel-trace", "-f", "csv", "-d", d, "-o", "run", "--",
           sys.executable, *BENCH_ARGS, "--precision", precision]
    print("[traffic]", " ".join(cmd), flush=True)
    subprocess.run(["timeout", "-s", "KILL", "300", *cmd], cwd=ROOT, check=True,
                   stdout=subprocess.DEVNULL)
    sums: dict = {}
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"no counter CSV under {d}")
    for fn in files:
        with open(fn) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                k = short_name(row["Kernel_Name"])
                s = sums.setdefault(k, [0.0, 0])
                s[0] += float(row["Counter_Value"])
                s[1] += 1
    return {k: (v[0] / v[1], v[1]) for k, v in sums.items()}


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else OUT
    os.makedirs(WORK, exist_ok=True)
    nb = COPY_MB << 20
    ck = "k_copy16"
    kernels, scale = {}, {}
    for prec in ("fp32", "fp64"):
        fetch = run_pass("FETCH_SIZE", prec)
        write = run_pass("WRITE_SIZE", prec)
        sf = nb / (fetch[ck][0] * 1024.0)
        sw = nb / (write[ck][0] * 1024.0)
        scale[prec] = {"fetch": sf, "write": sw}
        for k in sorted(set(fetch) & set(write)):
            if k in kernels and prec == "fp64":
                continue  # kernels both runs launch: keep the FP32 run's figures
            kernels[k] = {"fetch_bytes": fetch[k][0] * 1024.0 * sf, "write_bytes": write[k][0] * 1024.0 * sw,
                          "dispatches": fetch[k][1], "fetch_kib_raw": fetch[k][0], "write_kib_raw": write[k][0],
                          "run": prec}
    rec = {
        "command": "python3 rev-sim_amd/measure_traffic.py  (rocprofv3 --pmc FETCH_SIZE | WRITE_SIZE "
                   "--kernel-trace -f csv -- python3 " + " ".join(BENCH_ARGS) + " --precision fp32 | fp64)",
        "sources_sha256": sources_digest(),
        "scale": dict(scale, calibration_kernel=ck, bytes_each_way=nb),
        "kernels": kernels,
    }
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps({k: round((v["fetch_bytes"] + v["write_bytes"]) / 1e6, 2) for k, v in kernels.items()}))


if __name__ == "__main__":
    main()
