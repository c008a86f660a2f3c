"""Benchmark: LU solve at FP64 accuracy on one MI355X, plain FP64 and mixed-precision IR.

Headline workload (one step, the `value`): DGETRF + DGETRS (laswp, trsm(L,L,N,U),
trsm(L,U,N,N)) of a column-major n x n matrix with entries uniform in [-1, 1) and nrhs = 64
right-hand sides, n = 8192, inputs resident in HBM before the timed region (each step
restores A and B from HBM copies first; the copy is outside the timed interval of that
step).  Throughput counts 2n^3/3 + 2n^2*nrhs flops per solve.  Per-step latencies give
p50/p99.

Beside it, in the same JSON line:
  mixed_ir  -- A X = b at n = 16384, nrhs = 1, A = U diag(sigma) V^T with kappa = 1e2
               (seeded Householder products), solved to backward error <= n * 2^-52 by the
               mixed-precision solver (FP32 or FP16 factors, FP64 refinement / GMRES-IR,
               FP64 fallback), each mode timed once after a warm-up, next to the plain FP64
               GETRF + GETRS of the same system; the forward error is stated against the
               plain FP64 solution (which is bit-identical to the oracle's).
  batched   -- 20,000 independent systems, orders uniform in [8, 512], half LU and half
               SPD (Cholesky), nrhs uniform in [1, 16], one persistent launch each for the
               FP64 and the FP32 + FP64-refinement variants, with the autotuned launch
               configuration.

The dominant kernel of the headline step is the LU panel (phase GETRF2_IT, the persistent
wavefront panel kernel); its algorithmic traffic is the panel read and written once per
launch (2 * rows * 32 * 8 bytes), measured with HIP events on the library's stream, against
the HBM peak.  The events run in a second pass of as many steps as the timed ones: inside
the timed steps they would add 10 ms or more per step.  The trailing-update DGEMM is
reported beside it against the FP64 matrix peak.

CPU baseline: the oracle (a C restatement of the reference's own algorithm, OpenMP
standing in for rayon), timed on this host on one factorization + solve of the same
size and inputs; its result is also compared bit for bit with the GPU's.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def lu_flops(n, nrhs):
    return 2.0 * n ** 3 / 3.0 + 2.0 * n * n * nrhs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-n", type=int, default=0, help="size of the CPU baseline sample (default: n)")
    ap.add_argument("--no-extra", action="store_true", help="skip the mixed-IR and batched sections")
    ap.add_argument("--only", choices=["mixed", "batched", "dgemm", "serving"], default=None,
                    help="run one section alone (for profiling) and print it")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many bytes device to device (calibrates the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        src = torch.ones(args.copy_probe // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; it has no multi-GPU mode (across "
                                   "GPUs the factorization would shard by block-cyclic column blocks)"}))
        sys.exit(2)

    import torch
    import bola_amd as B

    if args.only:
        torch.cuda.set_device(0)
        sec = {"mixed": mixed_section, "batched": batched_section, "dgemm": dgemm_section,
               "serving": serving_section}[args.only](B, torch)
        print(json.dumps({args.only: sec}))
        return

    n, nrhs = args.n, args.nrhs
    torch.cuda.set_device(0)
    rng = np.random.default_rng(1)
    A0 = rng.uniform(-1.0, 1.0, n * n)
    B0 = rng.uniform(-1.0, 1.0, n * nrhs)
    dA0 = torch.from_numpy(A0).cuda()
    dB0 = torch.from_numpy(B0).cuda()
    dA = torch.empty_like(dA0)
    dB = torch.empty_like(dB0)
    dp = torch.zeros(n, dtype=torch.int32, device="cuda")
    h = B.default_handle()

    def step():
        B.getrf(n, n, dA, n, dp)
        B.getrs(n, nrhs, dA, n, dp, dB, n)

    for _ in range(args.warmup):
        dA.copy_(dA0)
        dB.copy_(dB0)
        step()
    torch.cuda.synchronize()

    # the timed steps run without per-phase events: recording them (two per kernel call,
    # thousands per factorization at n = 8192) adds 10 ms or more to a step
    times = []
    t_all = 0.0
    for _ in range(args.steps):
        dA.copy_(dA0)
        dB.copy_(dB0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        times.append(dt)
        t_all += dt
        print(f"# step {len(times)}: {dt * 1e3:.2f} ms", file=sys.stderr, flush=True)
    # the same number of steps again with HIP events around every kernel call on the
    # library's streams: the per-kernel averages of the roofline and gemm objects
    h.profile(True)
    h.profile_reset()
    t_prof = 0.0
    for _ in range(args.steps):
        dA.copy_(dA0)
        dB.copy_(dB0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        t_prof += time.perf_counter() - t0
    print(f"# instrumented steps: {1e3 * t_prof / args.steps:.2f} ms each", file=sys.stderr, flush=True)
    gemm_ms, gemm_calls, gemm_flops = h.profile_get("GEMM")
    panel_ms, panel_calls, _ = h.profile_get("GETRF2_IT")
    # algorithmic bytes of the panel launches: each panel of w <= 32 columns and the rows at
    # and below its diagonal, read once and written once
    panel_bytes = 0.0
    for j in range(0, n, 32):
        panel_bytes += 2.0 * (n - j) * min(32, n - j) * 8.0
    report = h.profile_report(args.steps)
    h.profile(False)

    ms = 1e3 * t_all / args.steps
    value = lu_flops(n, nrhs) / (t_all / args.steps) / 1e9
    lat = np.array(times) * 1e3
    gemm_tflops = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0

    # residual check of the last solve (the GPU path is also compared bit for bit with the
    # oracle in the CPU leg below)
    x = dB.cpu().numpy().reshape(nrhs, n).T
    Am = A0.reshape(n, n).T
    bm = B0.reshape(nrhs, n).T
    r = bm - Am @ x
    berr = float(np.max(np.linalg.norm(r, 1, axis=0) /
                        (np.linalg.norm(Am, 1) * np.linalg.norm(x, 1, axis=0) + np.linalg.norm(bm, 1, axis=0))))

    cpu = None
    if not args.no_cpu:
        import oracle as O
        cn = args.cpu_n or n
        cA = A0 if cn == n else rng.uniform(-1.0, 1.0, cn * cn)
        cB = B0 if cn == n else rng.uniform(-1.0, 1.0, cn * nrhs)
        a, b, p = cA.copy(), cB.copy(), np.zeros(cn, np.int32)
        t0 = time.perf_counter()
        O.getrf(cn, cn, a, cn, p)
        O.getrs(cn, nrhs, a, cn, p, b, cn)
        ct = time.perf_counter() - t0
        same = None
        if cn == n:
            same = bool(np.array_equal(p, dp.cpu().numpy()) and np.array_equal(b, dB.cpu().numpy()))
        cpu = {"value": lu_flops(cn, nrhs) / ct / 1e9, "unit": "GFLOP/s", "cores": O.num_threads(),
               "kind": "port", "sample": f"one DGETRF+DGETRS n={cn} nrhs={nrhs} of the same seeded inputs, "
                                         f"{ct:.2f} s; bit-identical to the GPU result: {same}"}

    extra_mixed = extra_batched = extra_dgemm = extra_serving = None
    if not args.no_extra:
        extra_dgemm = dgemm_section(B, torch)
        extra_mixed = mixed_section(B, torch)
        extra_batched = batched_section(B, torch)
        extra_serving = serving_section(B, torch)

    panel_gbs = panel_bytes * args.steps / (panel_ms * 1e-3) / 1e9 if panel_ms > 0 else 0.0
    out = {
        "metric": "LU solve GFLOP/s at FP64 accuracy (plain FP64 DGETRF+DGETRS)",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (uniform [-1,1), seed 1)",
        "config": {"workload": f"DGETRF+DGETRS n={n} nrhs={nrhs} column-major f64", "n": n, "nrhs": nrhs},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "backward_error": berr,
        "roofline": {"bound": "hbm", "kernel": "getf2_wave (LU panel, 32 columns, persistent)",
                     "achieved": round(panel_gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(panel_gbs / HBM_PEAK_GBS, 6),
                     "traffic": panel_traffic() if n == 8192 else None,
                     "panel_ms_per_step": round(panel_ms / args.steps, 3),
                     "panel_launches_per_step": panel_calls // max(args.steps, 1),
                     "panel_avg_launch_us": round(1e3 * panel_ms / max(panel_calls, 1), 2),
                     "algorithmic_bytes_per_step": panel_bytes,
                     "measured": "HIP events around each panel launch on the library stream, over "
                                 f"{args.steps} instrumented steps of the same workload that follow the "
                                 f"timed ones ({1e3 * t_prof / args.steps:.2f} ms per instrumented step)"},
        "gemm": {"bound": "mfma", "kernel": "dgemm_mfma (trailing update)", "achieved": round(gemm_tflops, 3),
                 "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(gemm_tflops / FP64_PEAK_TFLOPS, 4),
                 "gemm_ms_per_step": round(gemm_ms / args.steps, 3),
                 "gemm_launches_per_step": gemm_calls // max(args.steps, 1)},
        "dgemm": extra_dgemm,
        "mixed_ir": extra_mixed,
        "batched": extra_batched,
        "serving": extra_serving,
        "cpu_baseline": cpu,
    }
    print(report, file=sys.stderr)
    print(json.dumps(out))


def panel_traffic():
    """HBM bytes per panel launch at n = 8192, from the FETCH_SIZE / WRITE_SIZE counter runs
    (scaled by a plain copy's counts) recorded in profiles/getf2_wave_lu_n8192_hbm_traffic.md."""
    path = os.path.join(ROOT, "profiles", "getf2_wave_lu_n8192_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def dgemm_section(B, torch):
    """Plain DGEMM 'N','N', alpha = beta = 1, A, B, C uniform [0, 1) (the crate's gemm
    bench inputs, drawn here with numpy from seed 1): m = n = k = 1024 (the crate's bench
    size) and 8192, each the mean of repeated calls between events."""
    out = {}
    rng = np.random.default_rng(1)
    for n, reps in [(1024, 50), (8192, 5)]:
        a = torch.from_numpy(rng.random(n * n)).cuda()
        b = torch.from_numpy(rng.random(n * n)).cuda()
        c0 = torch.from_numpy(rng.random(n * n)).cuda()
        c = c0.clone()
        B.gemm("N", "N", n, n, n, 1.0, a, n, b, n, 1.0, c, n)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            B.gemm("N", "N", n, n, n, 1.0, a, n, b, n, 1.0, c, n)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        tf = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
        out[f"n{n}"] = {"ms": round(ms, 4), "tflops": round(tf, 2), "frac_fp64_peak": round(tf / FP64_PEAK_TFLOPS, 4)}
        print(f"# dgemm {n}: {ms:.3f} ms, {tf:.2f} TF/s", file=sys.stderr, flush=True)
        del a, b, c, c0
    torch.cuda.empty_cache()
    return out


def mixed_section(B, torch, n=16384, kappa=1e2):
    """Config: mixed-precision IR at n = 16384, nrhs = 1, kappa = 1e2."""
    from bola_amd import testmat
    A = testmat.conditioned(n, kappa, 7, device="cuda")
    b = torch.from_numpy(np.random.default_rng(8).uniform(-1, 1, n)).cuda()
    flops = lu_flops(n, 1)
    # plain FP64 on the same system
    a = A.clone()
    p = torch.zeros(n, dtype=torch.int32, device="cuda")
    xp = b.clone()
    B.getrf(n, n, a, n, p)
    B.getrs(n, 1, a, n, p, xp, n)
    torch.cuda.synchronize()
    a.copy_(A)
    xp.copy_(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B.getrf(n, n, a, n, p)
    B.getrs(n, 1, a, n, p, xp, n)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    del a
    out = {"workload": f"A X = b, n={n} nrhs=1, A = U diag(sigma) V^T, kappa={kappa:g}, b uniform [-1,1)",
           "plain_fp64": {"ms": round(plain_s * 1e3, 2), "gflops": round(flops / plain_s / 1e9, 1)},
           "modes": {}}
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    xpn = float(xp.abs().max())
    for name, prec, rbt in [("fp32_butterfly", "fp32", True), ("fp32_tournament", "fp32", False),
                            ("fp16_tournament", "fp16", False)]:
        B.dgesv_mixed(n, 1, A, n, b, n, x, n, prec, rbt=rbt)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = B.dgesv_mixed(n, 1, A, n, b, n, x, n, prec, rbt=rbt)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ferr = float((x - xp).abs().max()) / xpn
        out["modes"][name] = {"ms": round(dt * 1e3, 2), "gflops": round(flops / dt / 1e9, 1),
                              "speedup_vs_plain_fp64": round(plain_s / dt, 3),
                              "final_berr": st["final_berr"], "tolerance": st["tolerance"],
                              "converged": st["converged"], "ir_iterations": st["ir_iterations"],
                              "gmres_iterations": st["gmres_iterations"], "fallback": st["fallback"],
                              "factor_ms": round(st["factor_ms"], 2), "refine_ms": round(st["refine_ms"], 2),
                              "forward_error_vs_fp64": ferr}
        print(f"# mixed {name}: {dt * 1e3:.1f} ms, berr {st['final_berr']:.2e}", file=sys.stderr, flush=True)
    best = max(out["modes"].values(), key=lambda m: m["gflops"] if m["converged"] else 0.0)
    out["best_gflops"] = best["gflops"]
    out["best_speedup_vs_plain_fp64"] = best["speedup_vs_plain_fp64"]
    del A
    torch.cuda.empty_cache()
    return out


def serving_section(B, torch, nmax=8192, loads=(("latency", 10.0, 10.0), ("throughput", 60.0, 5.0))):
    """Solve server under Poisson arrivals (the serving configuration): orders log-uniform
    in [32, nmax] (quantised to 16 sizes), 70 % LU (A = U diag(sigma) V^T, kappa
    log-uniform in [1e2, 1e6]) / 30 % SPD (G G^T + n I, Cholesky), nrhs uniform in [1, 64],
    30 % of requests reuse a resident factorization with a new right-hand side; host
    matrices (uploaded by the server), 4 slots, 16 GiB budget.  Two loads: a light one
    (latency) and one near the server's capacity (throughput)."""
    import math
    from bola_amd import testmat
    rng = np.random.default_rng(1)
    sizes = np.unique(np.round(np.exp(np.linspace(math.log(32), math.log(nmax), 16))).astype(int))
    pool = []
    for n in sizes:
        for kind in (B.SERVE_LU, B.SERVE_CHOL):
            if kind == B.SERVE_CHOL:
                g = torch.Generator(device="cuda")
                g.manual_seed(int(n))
                G = torch.rand(int(n), int(n), dtype=torch.float64, device="cuda", generator=g) * 2.0 - 1.0
                A = (G @ G.t() + n * torch.eye(int(n), dtype=torch.float64, device="cuda")).t().contiguous()
                A = A.reshape(-1).cpu().numpy()
                del G
            else:
                kappa = float(np.exp(rng.uniform(math.log(1e2), math.log(1e6))))
                A = testmat.conditioned(int(n), kappa, int(n), device="cuda").cpu().numpy()
            pool.append((int(n), kind, A))
    torch.cuda.empty_cache()
    out = {"workload": f"Poisson arrivals, n log-uniform [32,{nmax}], 70% LU (kappa log-uniform [1e2,1e6]) / "
                       f"30% SPD, nrhs uniform [1,64], "
                       f"30% reuse a resident factorization, FP64, 4 slots, 16 GiB budget, host operands"}
    with B.Server(16 << 30, 4) as s:
        for (n, kind, A) in pool:  # warm-up, one at a time
            s.wait(s.submit(kind, n, 1, A, np.ones(n), np.zeros(n), 0))
        for name, qps, dur in loads:
            nreq = int(qps * dur)
            arr = np.cumsum(rng.exponential(1.0 / qps, nreq))
            ln = np.exp(rng.uniform(math.log(32), math.log(nmax), nreq))
            reqs = []
            for k in range(nreq):
                n = sizes[np.argmin(np.abs(np.log(sizes) - math.log(ln[k])))]
                kind = B.SERVE_LU if rng.random() < 0.7 else B.SERVE_CHOL
                idx = [j for j, p in enumerate(pool) if p[0] == n and p[1] == kind][0]
                nrhs = int(rng.integers(1, 65))
                reqs.append((arr[k], idx, nrhs, rng.random() < 0.3, rng.uniform(-1, 1, n * nrhs)))
            s.reset_stats()
            live = []
            t0 = time.perf_counter()
            for (ta, idx, nrhs, reuse, b) in reqs:
                n, kind, A = pool[idx]
                while time.perf_counter() - t0 < ta:
                    time.sleep(0.0002)
                x = np.zeros(n * nrhs)
                live.append((s.submit(kind, n, nrhs, A, b, x, (idx + 1) if reuse else 0), n, kind, nrhs))
            res = [s.wait(t[0]) for t in live]
            span = time.perf_counter() - t0
            st = s.stats()
            flops = sum(((2 * n ** 3 / 3 if kind == B.SERVE_LU else n ** 3 / 3) if not r["cache_hit"] else 0.0)
                        + 2.0 * n * n * nrhs for (t, n, kind, nrhs), r in zip(live, res))
            out[name] = {"offered_per_s": qps, "requests": nreq, "p50_ms": round(st["p50_us"] / 1e3, 3),
                         "p99_ms": round(st["p99_us"] / 1e3, 3), "solves_per_s": round(st["solves_per_s"], 1),
                         "gflops": round(flops / span / 1e9, 2), "cache_hits": st["cache_hits"],
                         "evictions": st["evictions"], "hbm_peak_gib": round(st["hbm_peak"] / 2 ** 30, 2),
                         "failed": sum(1 for r in res if r["status"] != 0 or r["info"] != 0)}
            print(f"# serving {name}: p50 {out[name]['p50_ms']} ms p99 {out[name]['p99_ms']} ms, "
                  f"{out[name]['solves_per_s']} solves/s, failed {out[name]['failed']}", file=sys.stderr, flush=True)
    return out


def batched_section(B, torch, count=20000):
    """Config: 20,000 uneven systems (orders 8..512, half LU / half SPD, nrhs 1..16)."""
    from bola_amd import batchgen
    d, arena, ptot = batchgen.make_batch(count, 8, 512, 16, seed=1)
    flops = batchgen.cost_flops(d)
    piv = torch.zeros(ptot, dtype=torch.int32, device="cuda")
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    iters = torch.zeros(count, dtype=torch.int32, device="cuda")
    out = {"workload": f"{count} systems, n uniform [8,512], half LU / half SPD, nrhs uniform [1,16], "
                       f"{arena.numel() * 8 / 2**30:.2f} GiB in HBM", "fp64_equiv_gflop": round(flops / 1e9, 1)}
    for variant, name in [(B.BATCH_FP64, "fp64"), (B.BATCH_FP32_IR, "fp32_ir")]:
        cfg = B.batch_tune(variant, d, arena)
        work = arena.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        B.batch_solve(variant, d, work, piv, info, iters, cfg)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        it = iters.cpu().numpy()
        out[name] = {"ms": round(dt * 1e3, 2), "gflops": round(flops / dt / 1e9, 1), "config": list(cfg),
                     "failed": int((info != 0).sum()), "refined_fallbacks": int((it == -1).sum())}
        print(f"# batched {name}: {dt * 1e3:.1f} ms", file=sys.stderr, flush=True)
        del work
    del arena
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
