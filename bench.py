"""Benchmark: solves/sec and p50/p99 latency at 1e-6 relative residual,
mixed-precision vs plain FP32 PCG, on the n=100,000 / d=10 attention-kernel
system (one GPU).

A step is one complete solve of (lambda I + exp(E E^T)) alpha = y with E and
y already resident in HBM: operator packing, CCCP build (probe matvec, QR,
FP64 CCCP loop, inverse square root) and the PCG solve.
  mixed: BF16 probe matvec, split-Gram inner PCG (FP16 2-way split for the k=1 products), FP64 iterative
         refinement until the FP64 true residual ||y - A alpha|| / ||y|| <= 1e-6
  fp32:  FP32 probe matvec and FP32-MFMA-Gram PCG until the recurrence
         residual <= 1e-6 (the reference's stopping rule)
Prints one JSON line (rank 0).  Use --quick for a small configuration.
"""

import argparse
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.join(ROOT, "learning-based-attention-kernel-regression_amd")
sys.path.insert(0, PKG_ROOT)

MFMA_BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0
# VALU issue on MI355X, measured with dependency-free streams of 16 chains per
# lane, 12 waves per CU: v_fma_f32 4.5 and v_exp_f32 8.6 cycles per wave64
# instruction at 2.4 GHz, no co-issue of the two (exp + fma: 12.3 cycles).
# Pure v_exp_f32 stream: 64 lanes * 1024 SIMDs * 2.4e9 / 8.6 = 18.3e12 exp/s.
V_EXP_F32_RATE_T = 18.3
FMA_CYCLES = 4.5
EXP_CYCLES = 8.6
CLOCK_HZ = 2.4e9
N_SIMD = 1024


def read_traffic(n, d):
    """HBM bytes per launch of the k=1 product, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary committed for this workload (None when absent)."""
    path = os.path.join(ROOT, "profiles", f"traffic_matvec_mixed_n{n}_d{d}.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    if "hbm_bytes_per_launch" not in vals:
        return None
    return vals["hbm_bytes_per_launch"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--d", type=int, default=10)
    ap.add_argument("--lambda-reg", type=float, default=1e-2)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--inner-tol", type=float, default=5e-2)
    ap.add_argument("--mixed-probes", default="mixed", choices=["bf16", "mixed"],
                    help="precision of the CCCP probe product in the mixed solve")
    ap.add_argument("--oracle-n", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--matvec-only", type=int, default=0,
                    help="only launch the k=1 matvec this many times per precision (profiling)")
    ap.add_argument("--serving-requests", type=int, default=24,
                    help="requests in the bounded serving sample (0 disables it)")
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--serving-rate", type=float, default=0.0,
                    help="Poisson arrival rate (requests/s, seed 0) of the serving sample, honoured against "
                         "the wall clock; 0 makes every request available at t=0 (saturating)")
    ap.add_argument("--serving-seq", type=int, default=0,
                    help="requests of the sequential fp32 comparison (0: the same count as the sample)")
    ap.add_argument("--quick", action="store_true")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solve shards by output rows across GPUs by design, "
              "but only the one-GPU path is implemented; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.quick:
        args.n, args.oracle_n = 4000, 2000
    import torch

    import laker
    from laker import _native as N

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n, d = args.n, args.d
    x_loc, y = laker.pathloss_problem(n, seed=0)
    emb = laker.PositionEmbedding(2, d, device=dev)
    with torch.no_grad():
        e = emb(x_loc.to(dev)).contiguous()
    y_d = y.to(dev).contiguous()
    nr = max(200, int(2 * n ** 0.5))
    stream = torch.cuda.current_stream()

    if args.matvec_only:
        # profiling mode: k=1 matvecs of both precisions, the preconditioner
        # apply, and a plain 256 MiB device copy (the reference point for
        # scaling hardware byte counters)
        x1 = torch.randn(n, device=dev)
        for prec in ("mixed", "fp32"):
            op = laker.AttentionKernelOperator(e, lambda_reg=args.lambda_reg, precision=prec)
            for _ in range(args.matvec_only):
                op.matvec(x1)
            torch.cuda.synchronize()
            op.close()
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(args.matvec_only):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"matvec_only": args.matvec_only, "n": n, "copy_bytes": src.numel() * 4}))
        return

    def one_solve(precision):
        op = laker.AttentionKernelOperator(e, lambda_reg=args.lambda_reg, precision=precision)
        pre = laker.CCCPPreconditioner(gamma=0.1, verbose=False, probe_source="device",
                                       probe_precision=args.mixed_probes if precision == "mixed" else "fp32")
        pre.build(op.matvec, n, seed=0)
        pcg = laker.PreconditionedConjugateGradient(tol=args.tol, max_iter=1000, verbose=False,
                                                    precision=precision, inner_tol=args.inner_tol,
                                                    check_every=16)
        alpha = pcg.solve(op.matvec, pre.apply, y_d)
        info = {"iterations": pcg.iterations, "outer": pcg.outer_iterations, "cccp_iterations": pre.iterations,
                "converged": pcg.converged, "residual_norm": pcg.residual_norm}
        sol = pcg.solution_fp64 if precision == "mixed" else alpha
        return op, pre, sol, info

    def timed(precision, steps, warmup):
        for _ in range(warmup):
            op, pre, _, _ = one_solve(precision)
            op.close()
            pre.close()
        lat, infos = [], []
        torch.cuda.synchronize()
        t_all = time.perf_counter()
        for _ in range(steps):
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            op, pre, sol, info = one_solve(precision)
            ev1.record(stream)
            ev1.synchronize()
            lat.append(ev0.elapsed_time(ev1))
            infos.append(info)
            last = (op, pre, sol)
            print(f"[bench] {precision} solve {len(lat)}/{steps}: {lat[-1]:.1f} ms, "
                  f"iters {info['iterations']} outer {info['outer']} cccp {info['cccp_iterations']}",
                  file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_all
        return lat, infos, wall, last

    lat_m, inf_m, wall_m, (op_m, pre_m, sol_m) = timed("mixed", args.steps, args.warmup)
    # true FP64 residual of the mixed solution (device FP64 residual kernel)
    r = op_m.residual_fp64(y_d.double(), sol_m)
    relres_m = float(torch.linalg.norm(r) / torch.linalg.norm(y_d.double()))
    res = {}
    if not args.no_fp32:
        lat_f, inf_f, wall_f, (op_f, pre_f, sol_f) = timed("fp32", max(1, min(args.steps, 3)), 0)
        rf = op_m.residual_fp64(y_d.double(), sol_f.double())
        res["fp32"] = {
            "solves_per_sec": 1000.0 / statistics.mean(lat_f),
            "p50_ms": statistics.median(lat_f), "p99_ms": max(lat_f),
            "iterations": inf_f[-1]["iterations"], "cccp_iterations": inf_f[-1]["cccp_iterations"],
            "recurrence_relres": inf_f[-1]["residual_norm"] / float(torch.linalg.norm(y_d)),
            "fp64_true_relres": float(torch.linalg.norm(rf) / torch.linalg.norm(y_d.double())),
        }
        op_f.close()
        pre_f.close()

    # dominant kernel: the k=1 mixed matvec of the inner PCG; time with HIP events on the stream
    x1 = torch.randn(n, device=dev)
    reps = 20
    mv = {}
    y1 = torch.empty_like(x1)
    for prec, code in (("mixed", N.PREC_MIXED), ("fp32", N.PREC_FP32)):
        opx = laker.AttentionKernelOperator(e, lambda_reg=args.lambda_reg, precision=prec)
        # straight C-ABI calls: the launches stay ahead of the GPU, so the events
        # bracket back-to-back kernels (product + reduction) and no host gaps
        h, xp, yp, sp = opx._h.h, N.ptr(x1), N.ptr(y1), N.stream_ptr()
        N.check(N.lib.laker_matvec(h, xp, 1, 1, yp, 1, code, sp))
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(reps):
            N.lib.laker_matvec(h, xp, 1, 1, yp, 1, code, sp)
        ev1.record(stream)
        ev1.synchronize()
        mv[prec] = ev0.elapsed_time(ev1) / reps
        opx.close()
    # preconditioner apply (HBM-bound): two passes over Q (n x N_r fp32)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    pre_m.apply(x1)
    ev0.record(stream)
    for _ in range(reps):
        pre_m.apply(x1)
    ev1.record(stream)
    ev1.synchronize()
    t_apply = ev0.elapsed_time(ev1) / reps
    apply_bytes = 2 * n * nr * 4 + 3 * n * 4

    flops_mv = 2.0 * n * n * d + 2.0 * n * n + 2.0 * n
    bytes_mv = n * d * 4 + 2 * n * 4
    t_mv = mv["mixed"] * 1e-3
    traffic = read_traffic(n, d)
    exp_peak = V_EXP_F32_RATE_T
    # issue-time floor of the symmetric k=1 product: each pair (i, j), (j, i)
    # costs one exp and two FMAs (row sum with x_j, column sum with x_i)
    valu_floor_ms = (n * n / 2) / 64 * (EXP_CYCLES + 2 * FMA_CYCLES) / N_SIMD / CLOCK_HZ * 1e3
    roofline = {
        "kernel": "mv_sym1_kernel<GRAM_F16S> + mv_sym_reduce_kernel (laker_matvec, k=1, "
                  "FP16 2-way split Gram on MFMA, each exp once per (i,j),(j,i) pair)",
        "bound": "mfma", "unit": "TFLOP/s",
        "achieved": flops_mv / t_mv / 1e12, "peak": MFMA_BF16_PEAK_TFLOPS,
        "frac": flops_mv / t_mv / 1e12 / MFMA_BF16_PEAK_TFLOPS,
        "traffic": traffic,
        "algorithmic_flops_per_launch": flops_mv,
        "algorithmic_bytes_per_launch": bytes_mv,
        "avg_launch_ms": mv["mixed"],
        "binding_resource": "VALU issue (v_exp_f32 plus the row and column FMAs of every exp), "
                            "not MFMA or HBM: see valu_floor and exp_rate",
        "valu_floor": {"ms": valu_floor_ms, "frac": valu_floor_ms / mv["mixed"],
                       "model": "n^2/2 pairs x (one v_exp_f32 + two v_fma_f32) / 64 lanes at the measured "
                                f"{EXP_CYCLES} + 2 x {FMA_CYCLES} cycles per wave64 instruction, "
                                f"{N_SIMD} SIMDs at {CLOCK_HZ / 1e9} GHz"},
        "exp_rate": {"achieved_Texp_s": n * n / t_mv / 1e12, "measured_peak_Texp_s": exp_peak,
                     "frac": n * n / t_mv / 1e12 / exp_peak,
                     "note": "n^2 algorithmic exps per product; the symmetric kernel evaluates n^2/2"},
        "fp32_matvec_ms": mv["fp32"],
        "precond_apply": {"bound": "hbm", "ms": t_apply, "achieved_GBs": apply_bytes / (t_apply * 1e-3) / 1e9,
                          "peak_GBs": HBM_PEAK_GBS,
                          "frac": apply_bytes / (t_apply * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }

    # FP64-oracle error on the matched instance (dense FP64 Cholesky on the GPU, untimed)
    oracle = None
    if not args.no_oracle:
        no = args.oracle_n
        xo, yo = laker.pathloss_problem(no, seed=0)
        with torch.no_grad():
            eo = emb(xo.to(dev)).contiguous()
        opo = laker.AttentionKernelOperator(eo, lambda_reg=args.lambda_reg, precision="mixed")
        preo = laker.CCCPPreconditioner(gamma=0.1, verbose=False, probe_source="device",
                                        probe_precision=args.mixed_probes).build(opo.matvec, no, seed=0)
        pm = laker.PreconditionedConjugateGradient(tol=args.tol, max_iter=1000, verbose=False, precision="mixed",
                                                   inner_tol=args.inner_tol)
        pm.solve(opo.matvec, preo.apply, yo.to(dev))
        pf = laker.PreconditionedConjugateGradient(tol=args.tol, max_iter=1000, verbose=False)
        af = pf.solve(opo.matvec, preo.apply, yo.to(dev))
        e64 = eo.double()
        a = torch.exp(e64 @ e64.T)
        a.diagonal().add_(args.lambda_reg)
        l_fac = torch.linalg.cholesky(a)
        a64 = torch.cholesky_solve(yo.to(dev).double().unsqueeze(1), l_fac).squeeze(1)
        del a, l_fac
        nrm = float(torch.linalg.norm(a64))
        oracle = {"n": no, "mixed_alpha_relerr": float(torch.linalg.norm(pm.solution_fp64 - a64)) / nrm,
                  "fp32_alpha_relerr": float(torch.linalg.norm(af.double() - a64)) / nrm,
                  "mixed_iterations": pm.iterations, "fp32_iterations": pf.iterations}
        opo.close()
        preo.close()
        torch.cuda.empty_cache()

    # CPU baseline: the oracle (torch-CPU restatement of the reference path) on host cores
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import laker_oracle as O

        threads = len(os.sched_getaffinity(0))
        torch.set_num_threads(threads)
        e_cpu = e.float().cpu()
        y_cpu = y.float()
        rows = min(n, 4096)
        chunk = O.auto_chunk(n)
        t0 = time.perf_counter()
        # one 1-D matvec row block (the reference's 2-D tile loop over all columns)
        acc = torch.zeros(rows)
        for j0 in range(0, n, chunk or n):
            j1 = min(j0 + (chunk or n), n)
            g = O.exp_safe(e_cpu[:rows] @ e_cpu[j0:j1].T, True)
            acc += g @ y_cpu[j0:j1]
        t_rows = time.perf_counter() - t0
        t_mv_cpu = t_rows * n / rows
        probe_rows = min(n, 256)
        pr = torch.randn(n, nr)
        t0 = time.perf_counter()
        for j0 in range(0, n, chunk or n):
            j1 = min(j0 + (chunk or n), n)
            g = O.exp_safe(e_cpu[:probe_rows] @ e_cpu[j0:j1].T, True)
            _ = g @ pr[j0:j1]
        t_probe_cpu = (time.perf_counter() - t0) * n / probe_rows
        iters = res.get("fp32", {}).get("iterations", inf_m[-1]["iterations"])
        t_solve_cpu = t_probe_cpu + iters * t_mv_cpu
        cpu_model = "unknown"
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": 1.0 / t_solve_cpu, "unit": "solves/sec", "cores": torch.get_num_threads(),
               "kind": "port",
               "sample": (f"oracle fp32 matvec on {rows} of {n} rows (all columns, reference tiling) and the "
                          f"{nr}-RHS probe matvec on {probe_rows} rows, extrapolated to one solve = probe + "
                          f"{iters} PCG matvecs; {cpu_model}"),
               "matvec_s": t_mv_cpu, "probe_matvec_s": t_probe_cpu}

    # Bounded sample of the serving stream: requests r = 0..R-1 (n log-uniform in
    # [1000, 65536], d in {10, 16}, 1..32 RHS, all seeded by r), all available at
    # t=0 (saturating) or arriving as a Poisson stream (--serving-rate),
    # continuous batching over the slots; against the same requests (or the
    # first --serving-seq of them) solved one after another in plain fp32.
    serving = None
    if args.serving_requests > 0:
        from laker.serving import SolveServer, make_request, poisson_arrivals

        realtime = args.serving_rate > 0
        arr = poisson_arrivals(args.serving_requests, args.serving_rate, seed=0) if realtime else \
            [0.0] * args.serving_requests
        reqs = [make_request(r, arrival=arr[r]) for r in range(args.serving_requests)]
        srv = SolveServer(slots=args.serving_slots, precision="mixed", tol=args.tol, inner_tol=args.inner_tol)
        srv.run(reqs[:2], realtime=False)  # warm-up (embeddings, code objects)
        res_s = srv.run(reqs, realtime=realtime)
        seq = SolveServer(slots=1, precision="fp32", tol=args.tol)
        seq.run(reqs[:2], realtime=False)
        res_q = seq.run(reqs[:args.serving_seq or len(reqs)], realtime=False)
        sm, sq = SolveServer.summary(res_s, srv.elapsed_s), SolveServer.summary(res_q, seq.elapsed_s)
        worst = max((r.relres for r in res_s if r.error is None), default=float("nan"))
        serving = {"requests": len(reqs), "slots": args.serving_slots,
                   "arrivals": (f"Poisson, {args.serving_rate}/s, seed 0, wall clock" if realtime
                                else "all at t=0 (saturating)"),
                   "n_range": [min(r.n for r in reqs), max(r.n for r in reqs)],
                   "rhs_total": sum(r.k for r in reqs), "max_inflight": srv.max_inflight,
                   "mixed_continuous": sm, "max_fp64_relres": worst,
                   "fp32_sequential": sq, "fp32_sequential_requests": len(res_q)}
        print(f"[bench] serving: {sm.get('solves_per_sec', 0):.2f} solves/s p50 {sm.get('p50_ms', 0):.0f} ms "
              f"p99 {sm.get('p99_ms', 0):.0f} ms; sequential fp32 {sq.get('solves_per_sec', 0):.2f} solves/s",
              file=sys.stderr, flush=True)

    out = {
        "metric": "solves/sec + p50/p99 latency at 1e-6 rel. residual, mixed vs plain fp32 PCG",
        "value": 1000.0 / statistics.mean(lat_m),
        "unit": "solves/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": statistics.mean(lat_m),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16-split-gram/fp32-pcg/fp64-refine",
        "data": "synthetic (path-loss field, PositionEmbedding default init, seed 0)",
        "config": {"workload": f"large matrix-free solve n={n} d={d} k=1 lambda={args.lambda_reg} "
                               f"CCCP N_r={nr} tol={args.tol}",
                   "n": n, "embedding_dim": d, "rhs": 1, "num_probes": nr, "parallelism": "single GPU"},
        "p50_ms": statistics.median(lat_m),
        "p99_ms": max(lat_m),
        "mixed": {"iterations_inner": inf_m[-1]["iterations"], "outer": inf_m[-1]["outer"],
                  "cccp_iterations": inf_m[-1]["cccp_iterations"], "fp64_true_relres": relres_m},
        "fp32": res.get("fp32"),
        "speedup_mixed_vs_fp32": (res["fp32"]["p50_ms"] / statistics.median(lat_m)) if "fp32" in res else None,
        "fp64_oracle": oracle,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "serving": serving,
        "host": platform.node(),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
