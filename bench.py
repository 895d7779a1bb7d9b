"""H-matvec benchmark on one MI355X.

Workload: N=65536 points uniform in [0,1]^2 (Philox, seed 1234), exponential
kernel exp(-r/0.1), KD-tree leaves of 32 points, weak admissibility, TLR tiles
of 32 (the reference pipeline's tile = leaf), ARA tolerance 1e-6 with 16
samples per block, uncapped ranks, FP64 factors.  This is the largest single-GPU
configuration of the benchmark set that fits: at N=2^20 in 3D the HODLR ranks
grow to ~1e4 and the FP64 factors alone exceed the 288 GB of one GPU.

A step is one H-matvec with one FP64 right-hand side, inputs resident in HBM.
value = algorithmic bytes (factors + x + y) / time, in GB/s.  The 16-RHS
matvec, build time and accuracy on 1024 sampled rows are reported alongside.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "kaust-hierarchical-matrices_amd"))

PEAK_HBM_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--d", type=int, default=2)
    ap.add_argument("--b", type=int, default=32)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--policy", choices=["ref", "full"], default="full")
    ap.add_argument("--rows", type=int, default=1024, help="sampled rows for the dense check")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline (0 = skip)")
    ap.add_argument("--mixed-target", type=float, default=3e-6, help="storage error budget of the mixed factors")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); sharding across GPUs is not part of "
              "this build.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import hmx

    torch.cuda.init()
    n, d, b = args.n, args.d, args.b
    ctx = hmx.Context()
    timings = {}
    t0 = time.perf_counter()
    pipe = hmx.Pipeline(n, d, b, args.tol, adm="weak",
                        max_rank_policy=hmx.MAXRANK_FULL if args.policy == "full" else hmx.MAXRANK_REF,
                        seed=1234, ctx=ctx, timings=timings)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    print(f"[bench] build {build_s:.2f}s stages {timings}", file=sys.stderr, flush=True)
    H = pipe.H
    nb = H.bytes()
    # matvec plan (built on first call) is outside the timed region
    x1 = torch.empty(n, dtype=torch.float64, device="cuda")
    hmx.generateRandomVector(ctx, 1, n, 4321, x1)
    y1 = torch.empty_like(x1)
    pipe.matvec(x1, y1)
    torch.cuda.synchronize()

    def timed(x, y, steps, warmup):
        for _ in range(warmup):
            pipe.matvec(x, y)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            pipe.matvec(x, y)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    sec1 = timed(x1, y1, args.steps, args.warmup)
    print(f"[bench] 1-RHS matvec {sec1 * 1e3:.4f} ms", file=sys.stderr, flush=True)
    # per-kernel times with hipEvents on the library's stream
    ms = [float(v) for v in hmx.matvec_profile(ctx, H, 1, x1, y1, iters=max(10, args.steps))]
    # algorithmic bytes: every factor entry once, x read, y written
    fac = nb["lowRankBytes"] + nb["denseBytes"]
    bytes1 = fac + 16 * n
    gbs1 = float(bytes1 / sec1 / 1e9)
    # dominant kernel: phase 1 reads V and x; phase 2 reads U, D, x and writes y
    ph1_bytes = nb["lowRankBytes"] / 2 + 8 * n
    ph2_bytes = nb["lowRankBytes"] / 2 + nb["denseBytes"] + 16 * n
    dom = 1 if ms[0] >= ms[2] else 2
    dom_ms = ms[0] if dom == 1 else ms[2]
    dom_bytes = ph1_bytes if dom == 1 else ph2_bytes
    achieved = float(dom_bytes / (dom_ms / 1e3) / 1e9)

    # 16 right-hand sides
    x16 = torch.empty(16, n, dtype=torch.float64, device="cuda")
    hmx.generateRandomVector(ctx, 16, n, 999, x16)
    y16 = torch.empty_like(x16)
    sec16 = timed(x16, y16, max(5, args.steps // 4), 2)
    gbs16 = (fac + 16 * n * 16) / sec16 / 1e9

    # accuracy against the dense FP64 kernel on sampled rows (EXPAND_MATRIX check)
    g = torch.Generator().manual_seed(1235)
    rows = torch.randperm(n, generator=g)[:args.rows].to(torch.int32).cuda()
    yd = torch.empty(1, args.rows, dtype=torch.float64, device="cuda")
    hmx.denseRowsXVector(ctx, n, d, pipe.points, pipe.tree.device_perm(), pipe.ell, rows, 1, x1, yd)
    pipe.matvec(x1, y1)
    yh = y1[rows.long()]
    rel_err = float(torch.linalg.norm(yh - yd[0]) / torch.linalg.norm(yd[0]))
    yd16 = torch.empty(16, args.rows, dtype=torch.float64, device="cuda")
    hmx.denseRowsXVector(ctx, n, d, pipe.points, pipe.tree.device_perm(), pipe.ell, rows, 16, x16, yd16)
    pipe.matvec(x16, y16)
    rel_err16 = float(torch.linalg.norm(y16[:, rows.long()] - yd16) / torch.linalg.norm(yd16))

    # 64 right-hand sides: FP64 matrix cores vs mixed FP32/FP16 matrix cores
    x64 = torch.empty(64, n, dtype=torch.float64, device="cuda")
    hmx.generateRandomVector(ctx, 64, n, 777, x64)
    y64 = torch.empty_like(x64)
    sec64 = timed(x64, y64, max(5, args.steps // 4), 2)
    ms64 = [float(v) for v in hmx.matvec_profile_multi(ctx, H, 64, x64, y64, iters=10)]
    t0 = time.perf_counter()
    minfo = pipe.make_mixed(args.mixed_target)
    torch.cuda.synchronize()
    mixed_build_s = time.perf_counter() - t0
    x32 = x64.float()
    y32 = torch.empty_like(x32)

    def timed_mixed(steps, warmup):
        for _ in range(warmup):
            pipe.matvec_mixed(x32, y32)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            pipe.matvec_mixed(x32, y32)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    secm = timed_mixed(max(5, args.steps // 4), 2)
    msm = [float(v) for v in hmx.matvec_profile_multi(ctx, H, 64, x32, y32, iters=10, mixed=True)]
    pipe.matvec(x64, y64)
    pipe.matvec_mixed(x32, y32)
    # error of the mixed product against the FP64 H-matvec and the dense kernel (sampled rows)
    rel_m_vs_fp64 = float(torch.linalg.norm(y32.double() - y64) / torch.linalg.norm(y64))
    yd64 = torch.empty(64, args.rows, dtype=torch.float64, device="cuda")
    hmx.denseRowsXVector(ctx, n, d, pipe.points, pipe.tree.device_perm(), pipe.ell, rows, 64, x32.double().contiguous(), yd64)
    rel_m_vs_dense = float(torch.linalg.norm(y32.double()[:, rows.long()] - yd64) / torch.linalg.norm(yd64))
    print(f"[bench] 64-RHS FP64 {sec64 * 1e3:.3f} ms, mixed {secm * 1e3:.3f} ms, rel {rel_m_vs_fp64:.2e}",
          file=sys.stderr, flush=True)

    ranks_per_level = {}
    for l in range(1, H.num_levels):
        li = H.level_info(l)
        if li["numTiles"]:
            ranks_per_level[l] = round(li["rankSum"] / li["numTiles"], 1)

    print(f"[bench] 16-RHS {sec16 * 1e3:.3f} ms, rel err {rel_err:.3e} / {rel_err16:.3e}", file=sys.stderr, flush=True)
    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(pipe, ctx, x1, bytes1, args.cpu_seconds)

    out = {
        "metric": "H-matvec ms/RHS & GB/s + build s, N=2^20 3D, mixed vs FP64, rel err \u22641e-5",
        "value": round(gbs1, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec1 * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox uniform points, N(0,1) right-hand side)",
        "config": {"workload": f"H-matvec, N={n} {d}D uniform, exp kernel l={pipe.ell}, leaf {b}, TLR tile {b}, "
                               f"weak admissibility, ARA eps={args.tol}, {args.policy} rank policy, FP64, 1 RHS",
                   "n": n, "d": d, "bucket": b, "tlr_tile": b, "tol": args.tol, "rhs": 1},
        "ms_per_rhs": round(sec1 * 1e3, 4),
        "rel_err_vs_dense": rel_err,
        "rhs16": {"ms_per_rhs": round(sec16 * 1e3 / 16, 4), "GBps": round(gbs16, 2), "rel_err_vs_dense": rel_err16},
        "rhs64": {"fp64_ms_per_rhs": round(sec64 * 1e3 / 64, 5), "fp64_call_ms": round(sec64 * 1e3, 4),
                  "fp64_kernel_ms": ms64,
                  "mixed_ms_per_rhs": round(secm * 1e3 / 64, 5), "mixed_call_ms": round(secm * 1e3, 4),
                  "mixed_kernel_ms": msm,
                  "speedup_mixed_vs_fp64_64rhs": round(sec64 / secm, 3),
                  "speedup_mixed_per_rhs_vs_fp64_1rhs": round(sec1 / (secm / 64), 2),
                  "mixed_rel_err_vs_fp64": rel_m_vs_fp64, "mixed_rel_err_vs_dense": rel_m_vs_dense,
                  "mixed_target": args.mixed_target, "mixed_bytes": int(minfo["bytes"]),
                  "mixed_fp32_cols": int(minfo["hiCols"]), "mixed_fp16_cols": int(minfo["loCols"]),
                  "mixed_build_s": round(mixed_build_s, 3)},
        "build_s": round(build_s, 3),
        "build_stages_s": {k: round(v, 3) for k, v in timings.items()},
        "factor_bytes": int(fac),
        "mean_rank_per_level": ranks_per_level,
        "kernel_ms": {"phase1_VTx": float(ms[0]), "reduce": float(ms[1]), "phase2_Uz_Dx": float(ms[2]),
                      "call": float(ms[3])},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                     "kernel": "k_mv_phase1" if dom == 1 else "k_mv_phase2", "kernel_bytes": int(dom_bytes)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def cpu_baseline(pipe, ctx, x1, bytes1, budget_s):
    """Oracle OpenMP H-matvec on the GPU-built factors copied to the host."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    H = pipe.H
    L = H.num_levels
    tiles, ranks, offs, Us, Vs = [], [], [], [], []
    for l in range(1, L):
        r, o, U, V = H.level_raw(ctx, l)
        tiles.append(pipe.structure.tiles(l))
        ranks.append(r)
        offs.append(o)
        Us.append(U)
        Vs.append(V)
    S = pipe.tree.info()[0]
    D = H.diag(ctx, S)
    X = x1.cpu().numpy()
    reps, t = 0, 0.0
    while t < budget_s and reps < 50:
        t0 = time.perf_counter()
        Y = O.matvec_factors(pipe.n, pipe.b, L, tiles, ranks, offs, Us, Vs, D, X)
        t += time.perf_counter() - t0
        reps += 1
    yg = pipe.matvec(x1).cpu().numpy()
    agree = float(np.linalg.norm(Y[:, 0] - yg) / np.linalg.norm(yg))
    return {"value": round(bytes1 * reps / t / 1e9, 3), "unit": "GB/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"full 1-RHS H-matvec on the GPU-built factors copied to host, {reps} reps, "
                      f"{t / reps * 1e3:.1f} ms each (oracle or_matvec_factors, OpenMP)",
            "ms_per_rhs": round(t / reps * 1e3, 3), "rel_diff_vs_gpu": agree}


if __name__ == "__main__":
    main()
