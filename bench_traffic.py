"""HBM traffic per kernel launch from two rocprofv3 counter passes.

bench.py's roofline.traffic field comes from here.  Run rocprofv3 twice over
the same command (`bench.py --counters-run`), once with `--pmc FETCH_SIZE` and
once with `--pmc WRITE_SIZE`, each into its own output folder, then

    python bench_traffic.py --fetch DIR1 --write DIR2 --workload "..." --out profiles/NAME.json

Both counters are read in KiB per dispatch.  Each is scaled by what it reads
for the calibration dispatch (the largest device-to-device copy of the
command, `__amd_rocclr_copyBuffer`, which moves `--copy-bytes` bytes each way):
scale = bytes / (1024 * counter).  A kernel's traffic is the mean over its
dispatches of fetch * fetch_scale + write * write_scale, in bytes.
"""
import argparse
import csv
import glob
import json
import os
import sys
from collections import defaultdict


def read_counter(folder, counter):
    """{kernel name: [values per dispatch]} for one counter of one pass."""
    files = glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        sys.exit(f"no counter_collection.csv under {folder}")
    per = defaultdict(lambda: defaultdict(float))  # kernel -> dispatch -> value
    for path in files:
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                key = (path, row.get("Dispatch_Id") or row.get("Correlation_Id"))
                per[row["Kernel_Name"]][key] += float(row["Counter_Value"])
    return {k: list(v.values()) for k, v in per.items()}


def short_name(name):
    """Kernel name without "void ", the hmx:: namespace and the argument list."""
    if name.startswith("void "):
        name = name[5:]
    if name.startswith("hmx::"):
        name = name[5:]
    depth, out = 0, []
    for ch in name:  # stop at the argument list, keep template arguments
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out).strip()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fetch", required=True, help="rocprofv3 output folder of the FETCH_SIZE pass")
    ap.add_argument("--write", required=True, help="rocprofv3 output folder of the WRITE_SIZE pass")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--workload", required=True)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    fetch = read_counter(a.fetch, "FETCH_SIZE")
    write = read_counter(a.write, "WRITE_SIZE")
    cal = [k for k in fetch if "copyBuffer" in k]
    if not cal:
        sys.exit("no copy dispatch to calibrate against")
    fk = max(v for k in cal for v in fetch[k])
    wk = max(v for k in cal if k in write for v in write[k])
    fs, ws = a.copy_bytes / (1024.0 * fk), a.copy_bytes / (1024.0 * wk)
    out = {"workload": a.workload,
           "calibration": {"copy_bytes": a.copy_bytes, "copy_fetch_kib": fk, "copy_write_kib": wk,
                           "fetch_scale": fs, "write_scale": ws}}
    for k, fv in sorted(fetch.items()):
        if "copyBuffer" in k or k not in write:
            continue
        wv = write[k]
        f = sum(fv) / len(fv)
        w = sum(wv) / len(wv)
        out[short_name(k)] = {"fetch_kib": f, "write_kib": w, "traffic_bytes": 1024.0 * (f * fs + w * ws),
                              "dispatches": len(fv)}
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    for k, v in out.items():
        if isinstance(v, dict) and "traffic_bytes" in v:
            print(f"{k}: {v['traffic_bytes'] / 1e9:.2f} GB per launch over {v['dispatches']} dispatches")


if __name__ == "__main__":
    main()
