"""H-matvec benchmark on one MI355X.

Workload: N=2^20 points uniform in [0,1]^3 (Philox, seed 1234), exponential
kernel exp(-r/0.2), KD-tree leaves of 64 points, bounding-box admissibility
max(diam) <= 1.0*dist, ARA tolerance 1e-5 with 16 samples per block, uncapped
ranks, FP64 factors.  The S x S TLR matrix of the reference pipeline does not
fit at this size (S = 16384 leaves), so the H-matrix is assembled directly:
dense diagonal blocks and leaf-level tiles by the TLR tile kernel, upper levels
by the same batched ARA sampling the kernel on the fly.

A step is one FP64 H-matvec with one right-hand side, inputs resident in HBM.
value = algorithmic bytes of that product (the factors and dense blocks its
plan reads once each, + x + y) / time, in GB/s.
Alongside: build seconds, accuracy on 2048 sampled rows of the dense kernel,
and the 64-RHS product in FP64 (matrix cores) against the mixed FP32/FP16
product (matrix cores, per-column scales), in ms per right-hand side.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "kaust-hierarchical-matrices_amd"))

PEAK_HBM_GBS = 8000.0
PEAK_FP64_MFMA_TFLOPS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--d", type=int, default=3)
    ap.add_argument("--b", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--adm", default="maxDiam", choices=["weak", "boxCenter", "maxDiam"])
    ap.add_argument("--policy", choices=["ref", "full"], default="full")
    ap.add_argument("--rows", type=int, default=2048, help="sampled rows for the dense check")
    ap.add_argument("--rhs", type=int, default=64, help="right-hand sides of the multi-RHS products")
    ap.add_argument("--cpu-seconds", type=float, default=8.0,
                    help="budget of the CPU baseline matvec loop (0 = skip the CPU baseline)")
    ap.add_argument("--cpu-stride", type=int, default=32, help="CPU baseline samples every n-th tile")
    ap.add_argument("--mixed-target", type=float, default=1e-5, help="error budget of the mixed factors")
    ap.add_argument("--assembly", choices=["direct", "tlr"], default="direct",
                    help="direct: every level sampled from the kernel; tlr: the reference's path (S x S TLR "
                         "matrix of leaf tiles, Morton order, H-matrix from the structure)")
    ap.add_argument("--symmetric", type=int, choices=[0, 1], default=1,
                    help="direct assembly: compress only the tiles (u, v) with u < v and mirror (v, u) (the kernel "
                         "and the admissibility are symmetric); 0: every tile on its own, as the reference")
    ap.add_argument("--batch", action="store_true",
                    help="many point clouds instead: batched build, autotuned hipGraph matvec, per-cloud dense check")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-nmin", type=int, default=2048)
    ap.add_argument("--batch-nmax", type=int, default=16384,
                    help="largest cloud (sizes up to 65536 need ~280 GB of FP64 factors at eps 1e-6: more than one GPU)")
    ap.add_argument("--counters-run", action="store_true",
                    help="build, one 1 GiB device copy (HBM counter calibration), 3 FP64 1-RHS matvecs, the FP64 "
                         "ceiling probes, one FP64 and one mixed --rhs product, exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): the library runs on one GPU; how the "
              "path shards across GPUs is a design, not an implementation.", file=sys.stderr)
        sys.exit(2)

    import torch
    import hmx

    torch.cuda.init()
    if args.batch:
        batch_main(args)
        return
    n, d, b, k = args.n, args.d, args.b, args.rhs
    ctx = hmx.Context()
    timings = {}
    log(f"building N={n} d={d} b={b} eps={args.tol} {args.adm}")
    t0 = time.perf_counter()
    pipe = hmx.Pipeline(n, d, b, args.tol, adm=args.adm,
                        max_rank_policy=hmx.MAXRANK_FULL if args.policy == "full" else hmx.MAXRANK_REF,
                        seed=1234, ctx=ctx, timings=timings, direct=args.assembly == "direct",
                        symmetric=bool(args.symmetric) and args.assembly == "direct")
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    log(f"build {build_s:.2f}s stages {timings}")
    H = pipe.H
    nb = H.bytes()
    fac = nb["lowRankBytes"] + nb["denseBytes"]
    # the build's dominant kernels: the upper-level samplers on the FP64 matrix cores
    bi = H.build_info()
    build_roofline = None
    if bi["samplerMs"] > 0:
        ach = bi["samplerFlops"] / (bi["samplerMs"] * 1e-3) / 1e12
        build_roofline = {"bound": "mfma", "kernel": "k_dir_sample_mf + k_dir_sampleT_mf",
                          "achieved": round(ach, 3), "peak": PEAK_FP64_MFMA_TFLOPS, "unit": "TFLOP/s",
                          "frac": round(ach / PEAK_FP64_MFMA_TFLOPS, 4), "flops": bi["samplerFlops"],
                          "kernel_s": round(bi["samplerMs"] * 1e-3, 3),
                          "kernel_evals_per_s": round(bi["kernelEvals"] / (bi["samplerMs"] * 1e-3), 1)}
        log(f"samplers {bi['samplerMs'] / 1e3:.2f} s, {ach:.1f} TFLOP/s FP64 MFMA "
            f"({bi['kernelEvals'] / (bi['samplerMs'] * 1e-3):.3g} kernel evaluations/s)")

    # ---- FP64, one right-hand side (the plan is built on the first call, outside the timed region)
    x1 = torch.empty(n, dtype=torch.float64, device="cuda")
    hmx.generateRandomVector(ctx, 1, n, 4321, x1)
    y1 = torch.empty_like(x1)
    pipe.matvec(x1, y1)
    torch.cuda.synchronize()

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    if args.counters_run:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of its dispatch
        # against 2^30 bytes gives the scale of the counters
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        src.fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        for _ in range(3):
            pipe.matvec(x1, y1)
        torch.cuda.synchronize()
        # the FP64 ceiling probes: the matrix-core busy counters' reference
        hmx.fp64_peak_probe(ctx, 0, iters=4000)
        hmx.fp64_peak_probe(ctx, 1, iters=4000)
        if args.rhs:  # one FP64 and one mixed multi-RHS product as well
            X = torch.empty(args.rhs, n, dtype=torch.float64, device="cuda")
            hmx.generateRandomVector(ctx, args.rhs, n, 777, X)
            Y = torch.empty_like(X)
            pipe.matvec(X, Y)
            pipe.make_mixed(args.mixed_target)
            X32 = X.float()
            Y32 = torch.empty_like(X32)
            pipe.matvec_mixed(X32, Y32)
            torch.cuda.synchronize()
        log("counters run done")
        return
    sec1 = timed(lambda: pipe.matvec(x1, y1), args.steps, args.warmup)
    # the streaming plan's bytes: U and V of the tiles applied as factors,
    # the diagonal blocks and the leaf tiles applied as dense blocks
    sp = H.plan_info(2)
    S = pipe.tree.info()[0]
    dense1 = 8 * (S + sp["denseBlocks"]) * b * b
    lr1 = sp["bytes"] - dense1
    bytes1 = sp["bytes"] + 16 * n
    gbs1 = float(bytes1 / sec1 / 1e9)
    log(f"FP64 1-RHS {sec1 * 1e3:.3f} ms, {gbs1:.0f} GB/s over the {bytes1 / 1e9:.2f} GB its plan reads "
        f"({sp['denseBlocks']} leaf tiles as dense blocks; the factors hold {fac / 1e9:.2f} GB)")
    # per-kernel times: hipEvents on the library's stream around each launch
    ms = [float(v) for v in hmx.matvec_profile(ctx, H, 1, x1, y1, iters=max(10, args.steps // 2))]
    ph1_bytes = lr1 / 2 + 8 * n           # V, x
    ph2_bytes = lr1 / 2 + dense1 + 16 * n  # U, dense blocks, x, y
    dom = 1 if ms[0] >= ms[2] else 2
    dom_ms = ms[0] if dom == 1 else ms[2]
    dom_bytes = ph1_bytes if dom == 1 else ph2_bytes
    achieved = float(dom_bytes / (dom_ms / 1e3) / 1e9)
    # measured streaming-read ceiling of this GPU (8 GiB read once per pass)
    probe = torch.empty(1 << 33, dtype=torch.uint8, device="cuda")
    read_ceiling = hmx.stream_read_probe(ctx, probe, iters=5)
    del probe
    torch.cuda.empty_cache()
    log(f"streaming-read ceiling {read_ceiling:.0f} GB/s; dominant kernel {achieved:.0f} GB/s")
    # measured FP64 arithmetic ceilings (register-only chains, 8 waves per SIMD)
    mfma_ceiling = hmx.fp64_peak_probe(ctx, 0)
    valu_ceiling = hmx.fp64_peak_probe(ctx, 1)
    log(f"FP64 ceilings: matrix cores {mfma_ceiling:.1f} TFLOP/s, vector FMA {valu_ceiling:.1f} TFLOP/s")
    if build_roofline:
        build_roofline["measured_mfma_ceiling"] = round(mfma_ceiling, 2)
        build_roofline["frac_of_measured_ceiling"] = round(build_roofline["achieved"] / mfma_ceiling, 4)

    # ---- accuracy on sampled rows of the dense FP64 kernel (EXPAND_MATRIX check)
    g = torch.Generator().manual_seed(1235)
    rows = torch.randperm(n, generator=g)[:args.rows].to(torch.int32).cuda()
    rl = rows.long()
    perm = pipe.tree.device_perm()

    def dense_rows(X, kk):
        out = torch.empty(kk, args.rows, dtype=torch.float64, device="cuda")
        hmx.denseRowsXVector(ctx, n, d, pipe.points, perm, pipe.ell, rows, kk, X, out)
        return out

    yd1 = dense_rows(x1, 1)[0]
    pipe.matvec(x1, y1)
    rel_err1 = float(torch.linalg.norm(y1[rl] - yd1) / torch.linalg.norm(yd1))
    log(f"FP64 1-RHS rel err on {args.rows} rows {rel_err1:.3e}")

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(pipe, ctx, x1, args.cpu_seconds, args.cpu_stride, args)
        log(f"CPU baseline {cpu['value']} GB/s on {cpu['cores']} threads; CPU build ~{cpu['build_s_extrapolated']} s "
            f"({cpu['build_sample']})")

    # the library allocates its matrix-core plans with hipMalloc: hand back
    # what torch's caching allocator still holds from the staging copies
    torch.cuda.empty_cache()

    rhs_block = None
    if k:
        # ---- k right-hand sides: FP64 on the matrix cores
        X = torch.empty(k, n, dtype=torch.float64, device="cuda")
        hmx.generateRandomVector(ctx, k, n, 777, X)
        Y = torch.empty_like(X)
        pipe.matvec(X, Y)
        secK = timed(lambda: pipe.matvec(X, Y), 5, 1)
        msK = [float(v) for v in hmx.matvec_profile_multi(ctx, H, k, X, Y, iters=5)]
        pipe.matvec(X, Y)
        ydK = dense_rows(X, k)
        rel_errK = float(torch.linalg.norm(Y[:, rl] - ydK) / torch.linalg.norm(ydK))
        pinfo = H.plan_info(0)
        log(f"FP64 {k}-RHS {secK * 1e3:.2f} ms, rel err {rel_errK:.3e}, plan {pinfo}")

        # ---- k right-hand sides: mixed FP32/FP16 on the matrix cores (FP32 inputs and outputs)
        torch.cuda.empty_cache()
        t0 = time.perf_counter()
        minfo = pipe.make_mixed(args.mixed_target)
        torch.cuda.synchronize()
        mixed_build_s = time.perf_counter() - t0
        X32 = X.float()
        Y32 = torch.empty_like(X32)
        pipe.matvec_mixed(X32, Y32)
        secM = timed(lambda: pipe.matvec_mixed(X32, Y32), 5, 1)
        msM = [float(v) for v in hmx.matvec_profile_multi(ctx, H, k, X32, Y32, iters=5, mixed=True)]
        pipe.matvec_mixed(X32, Y32)
        pipe.matvec(X, Y)
        rel_m_vs_fp64 = float(torch.linalg.norm(Y32.double() - Y) / torch.linalg.norm(Y))
        ydM = dense_rows(X32.double().contiguous(), k)
        rel_m_vs_dense = float(torch.linalg.norm(Y32.double()[:, rl] - ydM) / torch.linalg.norm(ydM))
        bytesM = int(minfo["bytes"]) + 8 * n * k
        log(f"mixed {k}-RHS {secM * 1e3:.2f} ms, rel vs FP64 {rel_m_vs_fp64:.2e}, vs dense {rel_m_vs_dense:.2e}")
        # per-phase rooflines.  FP64: matrix-core flops 2 k per factor entry
        # (phase 1 V^T X, phase 2 U Z; the dense blocks 2 k b^2 each) against
        # the FP64 matrix-core peak.  Mixed: algorithmic bytes of the phase
        # (its factors; X read once as FP16; z written once (phase 1) or read
        # once (phase 2) as FP16/FP32; Y written once as FP32) against HBM.
        f1, f2 = H.plan_phase_bytes(0)
        fl1, fl2 = 2.0 * k * f1 / 8, 2.0 * k * f2 / 8
        mfma_fp64 = {"bound": "mfma", "unit": "TFLOP/s", "peak": PEAK_FP64_MFMA_TFLOPS,
                     "phase1": round(fl1 / (msK[0] * 1e-3) / 1e12, 3),
                     "phase2": round(fl2 / (msK[2] * 1e-3) / 1e12, 3),
                     "call": round((fl1 + fl2) / secK / 1e12, 3)}
        mfma_fp64["frac"] = round(mfma_fp64["call"] / PEAK_FP64_MFMA_TFLOPS, 4)
        mfma_fp64["measured_ceiling"] = round(mfma_ceiling, 2)
        mfma_fp64["measured_vector_fma_ceiling"] = round(valu_ceiling, 2)
        mfma_fp64["frac_of_measured_ceiling"] = round(mfma_fp64["call"] / mfma_ceiling, 4)
        m1, m2 = H.plan_phase_bytes(1)
        zbytes = k * (2 * int(minfo["loCols"]) + 4 * int(minfo["hiCols"]))
        mb1 = m1 + 2 * n * k + zbytes
        mb2 = m2 + 2 * n * k + zbytes + 4 * n * k
        mdom = 1 if msM[0] >= msM[2] else 2
        mach = (mb1 if mdom == 1 else mb2) / ((msM[0] if mdom == 1 else msM[2]) * 1e-3) / 1e9
        mixed_roofline = {"bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS,
                          "kernel": "k_mf_phase1 (f16, f32)" if mdom == 1 else "k_mf_phase2g (mixed)",
                          "achieved": round(mach, 1), "frac": round(mach / PEAK_HBM_GBS, 4),
                          "phase1_bytes": int(mb1), "phase2_bytes": int(mb2),
                          "phase1_GBps": round(mb1 / (msM[0] * 1e-3) / 1e9, 1),
                          "phase2_GBps": round(mb2 / (msM[2] * 1e-3) / 1e9, 1),
                          "call_GBps": round((mb1 + mb2) / secM / 1e9, 1)}
        tm = traffic_bytes(args, "mixed")
        if tm is not None:  # counter traffic of the same kernels (committed rocprofv3 passes)
            mixed_roofline["traffic_phase1"], mixed_roofline["traffic_phase2"] = tm
            mixed_roofline["traffic_over_algorithmic"] = round((tm[0] + tm[1]) / (mb1 + mb2), 3)
        rhs_block = {
                "fp64_ms_per_rhs": round(secK * 1e3 / k, 5), "fp64_call_ms": round(secK * 1e3, 3),
                "fp64_kernel_ms": msK, "fp64_rel_err_vs_dense": rel_errK,
                "fp64_GBps": round((pinfo["bytes"] + 16 * n * k) / secK / 1e9, 1),
                "fp64_plan_bytes": int(pinfo["bytes"]), "dense_near_blocks": int(pinfo["denseBlocks"]),
                "mixed_ms_per_rhs": round(secM * 1e3 / k, 5), "mixed_call_ms": round(secM * 1e3, 3),
                "mixed_kernel_ms": msM, "mixed_GBps": round(bytesM / secM / 1e9, 1),
                "mixed_rel_err_vs_fp64": rel_m_vs_fp64, "mixed_rel_err_vs_dense": rel_m_vs_dense,
                "speedup_mixed_vs_fp64_same_rhs": round(secK / secM, 3),
                "speedup_mixed_per_rhs_vs_fp64_1rhs": round(sec1 / (secM / k), 2),
                "mixed_target": args.mixed_target, "mixed_bytes": int(minfo["bytes"]),
                "mixed_fp32_cols": int(minfo["hiCols"]), "mixed_fp16_cols": int(minfo["loCols"]),
                "mixed_build_s": round(mixed_build_s, 3),
                "fp64_roofline": mfma_fp64, "mixed_roofline": mixed_roofline}

    ranks_per_level = {}
    for l in range(1, H.num_levels):
        li = H.level_info(l)
        if li["numTiles"]:
            ranks_per_level[l] = {"tiles": li["numTiles"], "mean_rank": round(li["rankSum"] / li["numTiles"], 1)}

    out = {
        "metric": "H-matvec ms/RHS & GB/s + build s, N=2^20 3D, mixed vs FP64, rel err ≤1e-5",
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
        "data": "synthetic (Philox uniform points, N(0,1) right-hand sides)",
        "config": {"workload": f"H-matvec, N={n} points uniform in [0,1]^{d}, exp kernel l={pipe.ell}, leaf {b}, "
                               f"{args.adm} admissibility (eta 1), ARA eps={args.tol} (16-sample blocks), "
                               f"{args.policy} rank policy, FP64 factors, 1 RHS",
                   "n": n, "d": d, "bucket": b, "tol": args.tol, "admissibility": args.adm, "rhs": 1,
                   "assembly": "direct (no S x S TLR matrix)" if args.assembly == "direct" else
                               "TLR (S x S leaf tiles) then H-matrix from the structure",
                   "tlr_tile": ("none: every level is sampled from the kernel; the reference's TLR tile is the "
                                "leaf (src/kDTree.cu:14), and its S x S TLR matrix at S = N/64 does not fit")
                               if args.assembly == "direct" else f"the leaf ({b} points)",
                   "symmetric_tiles": ("tiles (u, v) with u < v compressed, (v, u) mirrored with U and V exchanged"
                                       if args.symmetric and args.assembly == "direct"
                                       else "every tile compressed on its own")},
        "ms_per_rhs": round(sec1 * 1e3, 4),
        "rel_err_vs_dense": rel_err1,
        "sampled_rows": args.rows,
        "build_s": round(build_s, 3),
        "build_stages_s": {k2: round(v, 3) for k2, v in timings.items()},
        "build_roofline": build_roofline,
        "factor_bytes": int(fac),
        "plan_bytes": int(sp["bytes"]),
        "dense_near_blocks": int(sp["denseBlocks"]),
        "levels": ranks_per_level,
        "kernel_ms": {"phase1_VTx": ms[0], "reduce": ms[1], "phase2_Uz_Dx": ms[2], "call": ms[3]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic_bytes(args, dom),
                     "kernel": "k_mv_phase1" if dom == 1 else "k_mv_phase2 + k_mv_phase2b",
                     "kernel_bytes": int(dom_bytes),
                     "kernel_ms": dom_ms, "measured_read_ceiling": round(read_ceiling, 1),
                     "frac_of_measured_ceiling": round(achieved / read_ceiling, 4)},
        "cpu_baseline": cpu,
    }
    if rhs_block:
        out[f"rhs{k}"] = rhs_block
    print(json.dumps(out), flush=True)


def batch_main(args):
    """Point clouds of uneven sizes: sizes log-uniform in [nmin, nmax], the
    first half in [0,1]^2 and the second in [0,1]^3, correlation length
    uniform in [0.05, 0.3] per cloud, leaf 32, eps 1e-6, bounding-box
    admissibility, 8 FP64 right-hand sides."""
    import numpy as np
    import torch
    import hmx
    rng = np.random.default_rng(1234)
    cnt = args.batch_count
    ns = np.floor(np.exp(rng.uniform(np.log(args.batch_nmin), np.log(args.batch_nmax), cnt))).astype(int)
    ells = rng.uniform(0.05, 0.3, cnt)
    clouds = [(int(ns[i]), 2 if i < cnt // 2 else 3, float(ells[i])) for i in range(cnt)]
    tot = int(ns.sum())
    ctx = hmx.Context()
    log(f"batch: {cnt} clouds, {tot} points")
    t0 = time.perf_counter()
    bp = hmx.BatchPipeline(clouds, b=32, tol=1e-6, adm="maxDiam", ctx=ctx, symmetric=bool(args.symmetric))
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    log(f"batch build {build_s:.2f}s {bp.timings}")
    k = 8
    chunk, tune_ms = bp.batch.tune(k)
    X = torch.empty(k, tot, dtype=torch.float64, device="cuda")
    hmx.generateRandomVector(ctx, k, tot, 99, X)
    Y = torch.empty_like(X)
    bp.matvec(X, Y, graph=True)
    torch.cuda.synchronize()
    steps = max(10, args.steps)
    t = time.perf_counter()
    for _ in range(steps):
        bp.matvec(X, Y, graph=True)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t) / steps
    fac = sum(h.bytes()["lowRankBytes"] + h.bytes()["denseBytes"] for h in bp.H)
    # per-cloud accuracy on 128 sampled rows of the dense kernel
    errs = []
    g = torch.Generator().manual_seed(7)
    for i, (n, d, ell) in enumerate(clouds):
        a, e = int(bp.batch.offsets[i]), int(bp.batch.offsets[i + 1])
        rows = torch.randperm(n, generator=g)[:128].to(torch.int32).cuda()
        yd = torch.empty(k, 128, dtype=torch.float64, device="cuda")
        hmx.denseRowsXVector(ctx, n, d, bp.points[i], bp.trees[i].device_perm(), ell, rows, k,
                             X[:, a:e].contiguous(), yd)
        errs.append(float(torch.linalg.norm(Y[:, a:e][:, rows.long()] - yd) / torch.linalg.norm(yd)))
    out = {"metric": f"batched H-matvec throughput, {cnt} clouds", "value": round((fac + 16 * tot * k) / sec / 1e9, 2),
           "unit": "GB/s", "n_gpus": 1, "steps": steps, "ms_per_step": round(sec * 1e3, 4), "higher_is_better": True,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{cnt} clouds, n log-uniform in [{args.batch_nmin}, {args.batch_nmax}], "
                                  "half 2D half 3D, l ~ U[0.05, 0.3], leaf 32, eps 1e-6, maxDiam, 8 RHS FP64",
                      "total_points": tot,
                      "symmetric_tiles": "mirrored" if args.symmetric else "every tile compressed"},
           "build_s": round(build_s, 3), "build_stages_s": {k2: round(v, 3) for k2, v in bp.timings.items()},
           "factor_bytes": int(fac), "tuned_xcd_chunk" if k >= 2 else "tuned_chunk": chunk, "tuned_ms": round(tune_ms, 4),
           "rel_err_vs_dense": {"max": max(errs), "median": float(np.median(errs))}}
    print(json.dumps(out), flush=True)


def traffic_bytes(args, dom):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a 1 GiB copy) of this
    workload; None for other workloads or when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "hmatvec_fp64_n1048576_3d_hbm_traffic.json")
    if not os.path.exists(path) or (args.n, args.d, args.b, args.tol, args.adm) != (1 << 20, 3, 64, 1e-5, "maxDiam"):
        return None
    with open(path) as f:
        t = json.load(f)
    def total(prefixes):
        found = [next((v for k, v in t.items() if k.startswith(p) and isinstance(v, dict)), None) for p in prefixes]
        return None if any(v is None for v in found) else int(sum(v["traffic_bytes"] for v in found))
    if dom == "mixed":  # (phase 1: the conversion and both phase-1 kernels, phase 2)
        # the FP32 group's V as FP16 hi + lo (the plan's default): the split
        # phase-1 kernel (FP32 z) beside the FP16 one
        p1 = total(["_ZN3hmx9k_x_mixed", "_ZN3hmx11k_mf_phase1IDF16_Li4ELi4ELb1ELb1EfE",
                    "_ZN3hmx11k_mf_phase1IDF16_Li4ELi4ELb1ELb0EDF16_"])
        p2 = total(["_ZN3hmx12k_mf_phase2gILi64ELb0EffDF16_Li4EfLb0E"])
        return None if p1 is None or p2 is None else (p1, p2)
    # phase 2 is the per-leaf pass plus the tall-tile (128-row block) pass
    return total(["k_mv_phase1<1>"] if dom == 1 else ["k_mv_phase2<64, 1>", "k_mv_phase2b<128, 1"])


def cpu_baseline(pipe, ctx, x1, budget_s, stride, args):
    """The same H-matvec on the host cores: every block product in OpenBLAS
    dgemm (numpy's bundled OpenBLAS, one thread per call) with OpenMP over
    tiles and leaves (oracle/cpu_blas.c), on a sample of the GPU-built factors
    copied to the host: every stride-th tile of every level plus all dense
    blocks.  Beside it, the CPU build: the oracle's whole pipeline (TLR path)
    at N = 2^13 and 2^14 of the same workload, extrapolated to N with the
    growth exponent the two sizes show."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    H = pipe.H
    L = H.num_levels
    n, b = pipe.n, pipe.b
    tiles, ranks, offs, Us, Vs = [], [], [], [], []
    fbytes = 0
    for l in range(1, L):
        li = H.level_info(l)
        m = (1 << (L - 1 - l)) * b
        all_tiles = pipe.structure.tiles(l)
        if li["numTiles"] == 0:
            tiles.append(np.zeros(0, np.int32)); ranks.append(np.zeros(0, np.int32))
            offs.append(np.zeros(1, np.int64)); Us.append(np.zeros(1)); Vs.append(np.zeros(1))
            continue
        U, V = H.level_device(ctx, l)
        r_all = H.level_ranks(ctx, l)
        o_all = np.concatenate([[0], np.cumsum(r_all)]).astype(np.int64)
        pick = np.arange(0, li["numTiles"], stride)
        idx = np.concatenate([np.arange(o_all[t] * m, (o_all[t] + r_all[t]) * m) for t in pick]) \
            if r_all[pick].sum() else np.zeros(0, np.int64)
        it = torch.from_numpy(idx).cuda()
        Us.append(U[it].cpu().numpy() if len(idx) else np.zeros(1))
        Vs.append(V[it].cpu().numpy() if len(idx) else np.zeros(1))
        del U, V, it
        tiles.append(all_tiles[pick].astype(np.int32))
        ranks.append(r_all[pick].astype(np.int32))
        offs.append(np.concatenate([[0], np.cumsum(r_all[pick])]).astype(np.int64))
        fbytes += 16 * int(r_all[pick].sum()) * m
    S = pipe.tree.info()[0]
    D = H.diag(ctx, S)
    X = x1.cpu().numpy()
    sbytes = fbytes + D.nbytes + 16 * n
    reps, t = 0, 0.0
    while t < budget_s and reps < 200:
        t0 = time.perf_counter()
        O.blas_matvec_factors(n, b, L, tiles, ranks, offs, Us, Vs, D, X)
        t += time.perf_counter() - t0
        reps += 1
    out = {"value": round(sbytes * reps / t / 1e9, 3), "unit": "GB/s", "cores": O.num_threads(), "kind": "port",
           "sample": f"every {stride}-th tile of every level ({fbytes / 1e9:.2f} GB of factors) plus all "
                     f"{S} dense blocks, 1 RHS, {reps} reps of {t / reps * 1e3:.1f} ms (OpenBLAS dgemm per block, "
                     "OpenMP over tiles and leaves)",
           "sample_bytes": int(sbytes)}
    # CPU build of the same workload at two small sizes, extrapolated
    sizes, secs = [1 << 13, 1 << 14], []
    adm = {"weak": 0, "boxCenter": 1, "maxDiam": 2}[args.adm]
    for nn in sizes:
        p = O.generate_points(nn, args.d, 1234)
        t0 = time.perf_counter()
        O.OracleHMatrix(p, nn, args.d, b, adm=adm, tol=args.tol,
                        max_rank_policy=1 if args.policy == "full" else 0, seed=1234)
        secs.append(time.perf_counter() - t0)
    expo = float(np.log(secs[1] / secs[0]) / np.log(2.0))
    out["build_s_extrapolated"] = round(secs[1] * (n / sizes[1]) ** expo, 1)
    out["build_sample"] = (f"oracle pipeline (KD-tree, structure, TLR ARA, H from TLR) at N={sizes[0]}: "
                           f"{secs[0]:.2f} s, N={sizes[1]}: {secs[1]:.2f} s on {O.num_threads()} threads; "
                           f"extrapolated to N={n} as N^{expo:.2f}")
    return out


if __name__ == "__main__":
    main()
