"""Stage timers and the times.csv row of the reference (src/counters.h:9-55).

The reference appends one row per run to results/times.csv:
    N, TOTAL_TIME, GENERATE_DATASET, KDTREE, TLR_MATRIX, CMTOMO, HMATRIX,
with times in milliseconds (printCountersInFile).  The same columns come first
here; STRUCTURE (timed inside KDTREE in the reference's main) and MATVEC
follow.  Without the TLR matrix (direct assembly) TLR_MATRIX and CMTOMO are 0
and HMATRIX holds the whole assembly.
"""
import os

COLUMNS = ["TOTAL_TIME", "GENERATE_DATASET", "KDTREE", "TLR_MATRIX", "CMTOMO", "HMATRIX", "STRUCTURE", "MATVEC"]


def row(n, timings, matvec_s=None):
    """Counter values in ms, reference column order, from a Pipeline timings dict (seconds)."""
    t = dict(timings)
    if "HMATRIX_DIRECT" in t:
        t["HMATRIX"] = t.pop("HMATRIX_DIRECT")
    ms = {k: 1e3 * t.get(k, 0.0) for k in COLUMNS}
    ms["MATVEC"] = 1e3 * matvec_s if matvec_s is not None else 0.0
    ms["TOTAL_TIME"] = sum(ms[k] for k in COLUMNS[1:7])
    return [n] + [ms[k] for k in COLUMNS]


def append_csv(path, n, timings, matvec_s=None):
    """Append one row (printCountersInFile format: values followed by ', ')."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    r = row(n, timings, matvec_s)
    with open(path, "a") as f:
        f.write(f"{r[0]}, " + "".join(f"{v:f}, " for v in r[1:]) + "\n")
    return r
