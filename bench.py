"""Benchmark: LU solve at FP64 accuracy on one MI355X, plain FP64 and mixed-precision IR.

Headline workload (one step, the `value`): DGETRF + DGETRS (laswp, trsm(L,L,N,U),
trsm(L,U,N,N)) of a column-major n x n matrix with entries uniform in [-1, 1) and nrhs = 64
right-hand sides, n = 8192, inputs resident in HBM before the timed region (each step
restores A and B from HBM copies first; the copy is outside the timed interval of that
step).  Throughput counts 2n^3/3 + 2n^2*nrhs flops per solve.  Per-step latencies give
p50/p99.

Beside it, in the same JSON line:
  dgemm     -- plain DGEMM 'N','N' at 1024^3 (the reference's bench size) and 8192^3.
  lp_gemm   -- the mixed LU's trailing-update product C -= A B at m = n = 16384, k = 512 in
               fp32, fp16, bf16 and split-fp16 products.
  cholesky  -- DPOTRF + DPOTRS ('L' and 'U') at n = 4096 and 8192, nrhs = 64, on
               A = G G^T + n I, median of 5 solves.
  mixed_ir  -- the mixed-precision grid: A X = B at n in {8192, 16384}, kappa in {1e2, 1e5,
               1e8}, nrhs in {1, 16}, A = U diag(sigma) V^T (seeded Householder products),
               solved to backward error <= n * 2^-52 by the mixed-precision solver in eight
               modes (fp32 / fp16 / bf16 / split-fp16 products, tournament pivoting or butterflies;
               FP64 refinement, GMRES-IR, FP64 fallback), each mode and the plain FP64
               GETRF + GETRS the median of 3 solves after a warm-up; the forward error is
               stated against the plain FP64 solution (which is bit-identical to the
               oracle's).
  batched   -- 20,000 independent systems, orders uniform in [8, 512], half LU and half
               SPD (Cholesky), nrhs uniform in [1, 16], one persistent launch each for the
               FP64 and the FP32 + FP64-refinement variants, with the autotuned launch
               configuration.
  serving   -- the solve server under Poisson arrivals, orders log-uniform in [32, 8192].

The dominant kernel of the headline step is the LU panel (phase GETRF2_IT, the persistent
panel kernel with two wavefronts a workgroup); its algorithmic traffic is the panel read and written once per
launch (2 * rows * 32 * 8 bytes), measured with HIP events on the library's stream, against
the HBM peak.  The events run in a second pass of as many steps as the timed ones: inside
the timed steps they would add 10 ms or more per step.  The trailing-update DGEMM is
reported beside it against the FP64 matrix peak.

CPU baseline (a reported baseline, not the target): the oracle (a C restatement of the
reference's own algorithm, OpenMP standing in for rayon) on this host's cores, one warm-up
then the median and p10/p90 of 10 runs of the headline solve on the same inputs (its first
timed result compared bit for bit with the GPU's), plus rows for DGEMM 1024^3, DGETRF +
DGETRS at 4096 and DPOTRF + DPOTRS 'L' / 'U' at 4096; the CPU model and core counts are
recorded with it.
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
    ap.add_argument("--cpu-runs", type=int, default=10, help="timed CPU runs per row, after one warm-up")
    ap.add_argument("--no-cpu-rows", action="store_true",
                    help="time only the headline row on the CPU (not DGEMM 1024, LU 4096, Cholesky L/U 4096)")
    ap.add_argument("--no-extra", action="store_true", help="skip the mixed-IR and batched sections")
    ap.add_argument("--only", choices=["mixed", "batched", "dgemm", "lp_gemm", "serving", "cholesky", "cpu"],
                    default=None,
                    help="run one section alone (for profiling) and print it")
    ap.add_argument("--mixed-n", type=int, default=0, help="mixed section: this order only (default 8192 and 16384)")
    ap.add_argument("--mixed-kappa", type=float, default=0.0, help="mixed section: this kappa only")
    ap.add_argument("--mixed-nrhs", type=int, default=0, help="mixed section: this nrhs only")
    ap.add_argument("--mixed-modes", default="", help="mixed section: these modes only (comma list, e.g. fp16x3_butterfly)")
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
        mixed_kw = {}
        if args.mixed_n:
            mixed_kw["sizes"] = (args.mixed_n,)
        if args.mixed_kappa:
            mixed_kw["kappas"] = (args.mixed_kappa,)
        if args.mixed_nrhs:
            mixed_kw["nrhss"] = (args.mixed_nrhs,)
        if args.mixed_modes:
            mixed_kw["only_modes"] = tuple(args.mixed_modes.split(","))
        sec = {"mixed": lambda B, torch: mixed_section(B, torch, **mixed_kw), "batched": batched_section, "dgemm": dgemm_section, "lp_gemm": lp_gemm_section,
               "serving": serving_section, "cholesky": chol_section, "cpu": lambda B, torch: dict(cpu_rows(), **cpu_info())}[
            args.only](B, torch)
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

        def cpu_lu():
            O.getrf(cn, cn, a, cn, p)
            O.getrs(cn, nrhs, a, cn, p, b, cn)
            return bool(cn == n and np.array_equal(p, dp.cpu().numpy()) and np.array_equal(b, dB.cpu().numpy()))

        ts, same = cpu_timed(lambda: (np.copyto(a, cA), np.copyto(b, cB)), cpu_lu, args.cpu_runs)
        cpu = cpu_stats(ts, lu_flops(cn, nrhs))
        cpu.update({"value": round(cpu["value"], 2), "unit": "GFLOP/s", "cores": O.num_threads(), "kind": "port",
                    "sample": f"DGETRF+DGETRS n={cn} nrhs={nrhs} of the same seeded inputs, one warm-up then "
                              f"{args.cpu_runs} runs (median {cpu['median_s']:.2f} s); first timed run "
                              f"bit-identical to the GPU result: {same}",
                    "label": CPU_LABEL})
        cpu.update(cpu_info())
        print(f"# cpu headline: median {cpu['median_s']:.3f} s, {cpu['value']} GFLOP/s", file=sys.stderr, flush=True)
        if not args.no_cpu_rows:
            cpu["rows"] = cpu_rows(args.cpu_runs)

    extra_mixed = extra_batched = extra_dgemm = extra_serving = extra_chol = extra_lp = None
    if not args.no_extra:
        extra_dgemm = dgemm_section(B, torch)
        extra_lp = lp_gemm_section(B, torch)
        extra_chol = chol_section(B, torch)
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
        "roofline": {"bound": "hbm", "kernel": "getf2_duo (LU panel, 32 columns, two wavefronts a workgroup, persistent)",
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
        "lp_gemm": extra_lp,
        "cholesky": extra_chol,
        "mixed_ir": extra_mixed,
        "batched": extra_batched,
        "serving": extra_serving,
        "cpu_baseline": cpu,
    }
    print(report, file=sys.stderr)
    print(json.dumps(out))


CPU_LABEL = "BOLA CPU algorithm, C restatement (OpenMP for rayon) - reported baseline, not target"


def cpu_info():
    """CPU model, physical cores and logical CPUs of this host, and the CPUs this process may use."""
    model, cores, cur = None, set(), {}
    try:
        for line in list(open("/proc/cpuinfo")) + [""]:
            if ":" not in line:
                if cur:
                    cores.add((cur.get("physical id"), cur.get("core id")))
                cur = {}
                continue
            k, v = (t.strip() for t in line.split(":", 1))
            if k == "model name" and model is None:
                model = v
            if k in ("physical id", "core id"):
                cur[k] = v
    except OSError:
        pass
    return {"cpu_model": model, "physical_cores": len(cores) or None, "logical_cpus": os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0))}


def cpu_timed(prep, run, runs):
    """One warm-up, then `runs` timed calls, each after an untimed prep (the inputs copied
    fresh, as criterion's iter_batched does).  Returns the times in seconds and the result
    of the first timed call."""
    prep()
    run()
    ts, first = [], None
    for r in range(runs):
        prep()
        t0 = time.perf_counter()
        res = run()
        ts.append(time.perf_counter() - t0)
        if r == 0:
            first = res
    return ts, first


def cpu_stats(ts, work, unit_scale=1e9):
    med = float(np.median(ts))
    return {"value": work / med / unit_scale, "median_s": round(med, 4),
            "p10_s": round(float(np.percentile(ts, 10)), 4), "p90_s": round(float(np.percentile(ts, 90)), 4),
            "runs": len(ts)}


def cpu_rows(runs=10, n_lu=4096, n_chol=4096, nrhs=64):
    """The oracle's rows for the plain configurations beside the headline one: DGEMM 'N','N'
    1024^3 (alpha = beta = 1, uniform [0, 1)), DGETRF + DGETRS at n_lu, DPOTRF + DPOTRS for
    'L' (single-threaded in the reference) and 'U' at n_chol, each one warm-up then the
    median and p10/p90 of `runs` calls.  Cholesky at 8192 is left out: its 'L' factorization
    alone takes more than ten seconds a call on one core."""
    import oracle as O
    rng = np.random.default_rng(1)
    rows = {}
    n = 1024
    a, b, c0 = rng.random(n * n), rng.random(n * n), rng.random(n * n)
    c = c0.copy()
    ts, _ = cpu_timed(lambda: np.copyto(c, c0), lambda: O.gemm("N", "N", n, n, n, 1.0, a, n, b, n, 1.0, c, n), runs)
    rows["dgemm_n1024"] = cpu_stats(ts, 2.0 * n ** 3)
    n = n_lu
    A0, B0 = rng.uniform(-1, 1, n * n), rng.uniform(-1, 1, n * nrhs)
    a, x, p = A0.copy(), B0.copy(), np.zeros(n, np.int32)

    def lu():
        O.getrf(n, n, a, n, p)
        O.getrs(n, nrhs, a, n, p, x, n)

    ts, _ = cpu_timed(lambda: (np.copyto(a, A0), np.copyto(x, B0)), lu, runs)
    rows[f"getrf_getrs_n{n}_nrhs{nrhs}"] = cpu_stats(ts, lu_flops(n, nrhs))
    n = n_chol
    G = rng.uniform(-1, 1, (n, n))
    S0 = (G @ G.T + n * np.eye(n)).T.reshape(-1).copy()
    del G
    B0 = rng.uniform(-1, 1, n * nrhs)
    a, x = S0.copy(), B0.copy()
    for uplo in ("L", "U"):
        def chol():
            O.potrf(uplo, n, a, n)
            O.potrs(uplo, n, nrhs, a, n, x, n)

        ts, _ = cpu_timed(lambda: (np.copyto(a, S0), np.copyto(x, B0)), chol, runs)
        rows[f"potrf_potrs_{uplo}_n{n}_nrhs{nrhs}"] = cpu_stats(ts, n ** 3 / 3.0 + 2.0 * n * n * nrhs)
    for k, v in rows.items():
        v["value"] = round(v["value"], 2)
        v["unit"] = "GFLOP/s"
        print(f"# cpu {k}: median {v['median_s'] * 1e3:.1f} ms, {v['value']} GFLOP/s", file=sys.stderr, flush=True)
    rows["threads"] = O.num_threads()
    rows["label"] = CPU_LABEL
    return rows


def panel_traffic():
    """HBM bytes per panel launch at n = 8192, from the FETCH_SIZE / WRITE_SIZE counter runs
    (scaled by a plain copy's counts) recorded in profiles/getf2_duo_lu_n8192_hbm_traffic.md.
    The record names the SHA-256 of the panel source it was measured on; when getrf.hip has
    changed since, the figure no longer describes the kernel that runs and None is returned."""
    import hashlib
    path = os.path.join(ROOT, "profiles", "getf2_duo_lu_n8192_hbm_traffic.md")
    src = os.path.join(ROOT, "bola_amd", "csrc", "getrf.hip")
    fields = {}
    try:
        for line in open(path):
            key, sep, val = line.partition(":")
            if sep and key in ("traffic_bytes_per_launch", "panel_source_sha256"):
                fields[key] = val.strip()
        with open(src, "rb") as f:
            digest = hashlib.sha256(f.read()).hexdigest()
    except OSError:
        return None
    if fields.get("panel_source_sha256") != digest:
        print("# panel traffic record is from another panel source: roofline.traffic null", file=sys.stderr)
        return None
    return float(fields["traffic_bytes_per_launch"]) if "traffic_bytes_per_launch" in fields else None


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


def lp_gemm_section(B, torch, m=16384, n=16384, k=512, reps=5):
    """The mixed LU's trailing-update product C -= A B (fp32 storage) at the size of its
    first bulk update (m = n = 16384, k = 512: two 256-column panels) in each product
    precision, the operand conversions included; rates in fp32-equivalent TFLOP/s (2 m n k
    flops) and as a fraction of the matrix peak the products run on (fp32 157.3 TF/s;
    fp16 / bf16 2.5 PF/s, three fp16 products per fp32-equivalent one for the split mode)."""
    peaks = {"fp32": 157.3, "fp16": 2500.0, "bf16": 2500.0, "fp16x3": 2500.0 / 3.0}
    rng = torch.Generator(device="cuda")
    rng.manual_seed(1)
    a = torch.rand(m * k, dtype=torch.float32, device="cuda", generator=rng) * 2.0 - 1.0
    b = torch.rand(k * n, dtype=torch.float32, device="cuda", generator=rng) * 2.0 - 1.0
    c = torch.zeros(m * n, dtype=torch.float32, device="cuda")
    out = {"workload": f"C -= A B, m={m} n={n} k={k}, fp32 storage, mean of {reps} calls after a warm-up"}
    for prec in ("fp32", "fp16", "bf16", "fp16x3"):
        B.sgemm_lp(m, n, k, a, m, b, k, c, m, prec)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            B.sgemm_lp(m, n, k, a, m, b, k, c, m, prec)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        tf = 2.0 * m * n * k / (ms * 1e-3) / 1e12
        out[prec] = {"ms": round(ms, 4), "tflops_fp32_equiv": round(tf, 1), "frac_peak": round(tf / peaks[prec], 4)}
        print(f"# lp_gemm {prec}: {ms:.3f} ms, {tf:.1f} TF/s", file=sys.stderr, flush=True)
    del a, b, c
    torch.cuda.empty_cache()
    return out


def chol_section(B, torch, sizes=(4096, 8192), nrhs=64, runs=5):
    """Plain FP64 Cholesky solve: DPOTRF + DPOTRS ('L' and 'U') of A = G G^T + n I with G
    uniform [-1, 1) (seeded, built in FP64 on the device), nrhs = 64 right-hand sides
    uniform [-1, 1); one warm-up, then the median of `runs` solves, each from a fresh HBM
    copy of A and B made outside its timed interval.  Throughput counts n^3/3 + 2 n^2 nrhs
    flops per solve."""
    out = {"workload": f"DPOTRF+DPOTRS, A = G G^T + n I (G uniform [-1,1)), nrhs={nrhs}, column-major f64, "
                       f"median of {runs}"}
    for n in sizes:
        g = torch.Generator(device="cuda")
        g.manual_seed(n)
        G = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2.0 - 1.0
        A0 = (G @ G.t() + n * torch.eye(n, dtype=torch.float64, device="cuda")).reshape(-1)
        del G
        B0 = torch.rand(n * nrhs, dtype=torch.float64, device="cuda", generator=g) * 2.0 - 1.0
        a = torch.empty_like(A0)
        x = torch.empty_like(B0)
        flops = n ** 3 / 3.0 + 2.0 * n * n * nrhs
        for uplo in ("L", "U"):
            ts = []
            for r in range(runs + 1):
                a.copy_(A0)
                x.copy_(B0)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                B.potrf(uplo, n, a, n)
                B.potrs(uplo, n, nrhs, a, n, x, n)
                torch.cuda.synchronize()
                if r:
                    ts.append(time.perf_counter() - t0)
            t = float(np.median(ts))
            # backward error of the last solve, on the device in FP64
            Am = A0.view(n, n).t()
            X = x.view(nrhs, n).t()
            R = B0.view(nrhs, n).t() - Am @ X
            an = float(Am.abs().sum(dim=1).max())
            berr = float((R.abs().amax(dim=0) / (an * X.abs().amax(dim=0) + B0.view(nrhs, n).t().abs().amax(dim=0))).max())
            out[f"n{n}_{uplo}"] = {"ms": round(t * 1e3, 3), "gflops": round(flops / t / 1e9, 1),
                                   "frac_fp64_peak": round(flops / t / 1e12 / FP64_PEAK_TFLOPS, 4),
                                   "p10_ms": round(float(np.percentile(ts, 10)) * 1e3, 3),
                                   "p90_ms": round(float(np.percentile(ts, 90)) * 1e3, 3),
                                   "backward_error": berr}
            print(f"# cholesky {n} {uplo}: {t * 1e3:.2f} ms, {flops / t / 1e9:.0f} GFLOP/s, berr {berr:.2e}",
                  file=sys.stderr, flush=True)
        del A0, B0, a, x
        torch.cuda.empty_cache()
    return out


def mixed_section(B, torch, sizes=(8192, 16384), kappas=(1e2, 1e5, 1e8), nrhss=(1, 16), runs=3, only_modes=()):
    """Mixed-precision IR grid: n in {8192, 16384}, kappa in {1e2, 1e5, 1e8}, nrhs in {1, 16},
    A = U diag(sigma) V^T (seeded Householder products, sigma geometric from 1 to 1/kappa),
    B uniform [-1, 1).  For every point: the plain FP64 GETRF + GETRS and each mode
    (fp32 / fp16 / bf16 products, tournament pivoting or random butterflies without
    interchanges), one warm-up then the median of `runs` timed solves each.  The forward
    error of each mode is stated against the plain FP64 solution, which equals the oracle's
    bit for bit (the GPU parity tests check that at these sizes)."""
    from bola_amd import testmat
    modes = [(f"{p}_{'butterfly' if r else 'tournament'}", p, r) for p in ("fp32", "fp16", "bf16", "fp16x3")
             for r in (True, False)]
    if only_modes:
        modes = [md for md in modes if md[0] in only_modes]
    out = {"workload": "A X = B, A = U diag(sigma) V^T (Householder products, sigma geometric 1..1/kappa), "
                       "B uniform [-1,1); stop at backward error <= n * eps64; median of "
                       f"{runs} solves after a warm-up",
           "points": {}}

    def timed(f, prep=None):
        if prep:
            prep()
        f()
        torch.cuda.synchronize()
        ts, res = [], None
        for _ in range(runs):
            if prep:
                prep()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = f()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), res

    for n in sizes:
        for kappa in kappas:
            A = testmat.conditioned(n, kappa, 7, device="cuda")
            a = torch.empty_like(A)
            p = torch.zeros(n, dtype=torch.int32, device="cuda")
            for nrhs in nrhss:
                b = torch.from_numpy(np.random.default_rng(8).uniform(-1, 1, n * nrhs)).cuda()
                xp = torch.empty_like(b)
                flops = lu_flops(n, nrhs)

                def restore():
                    a.copy_(A)
                    xp.copy_(b)

                def plain():
                    B.getrf(n, n, a, n, p)
                    B.getrs(n, nrhs, a, n, p, xp, n)

                plain_s, _ = timed(plain, restore)
                pt = {"plain_fp64": {"ms": round(plain_s * 1e3, 2), "gflops": round(flops / plain_s / 1e9, 1)},
                      "modes": {}}
                xpn = xp.view(nrhs, n).abs().amax(dim=1)
                x = torch.zeros_like(b)
                for name, prec, rbt in modes:
                    dt, st = timed(lambda: B.dgesv_mixed(n, nrhs, A, n, b, n, x, n, prec, rbt=rbt))
                    ferr = float(((x - xp).view(nrhs, n).abs().amax(dim=1) / xpn).max())
                    pt["modes"][name] = {
                        "ms": round(dt * 1e3, 2), "gflops": round(flops / dt / 1e9, 1),
                        "speedup_vs_plain_fp64": round(plain_s / dt, 3), "final_berr": st["final_berr"],
                        "tolerance": st["tolerance"], "converged": st["converged"],
                        "ir_iterations": st["ir_iterations"], "gmres_iterations": st["gmres_iterations"],
                        "fallback": st["fallback"], "factor_ms": round(st["factor_ms"], 2),
                        "refine_ms": round(st["refine_ms"], 2), "forward_error_vs_fp64": ferr}
                ok = {k: v for k, v in pt["modes"].items() if v["converged"] and not v["fallback"]}
                if ok:
                    bk = max(ok, key=lambda k: ok[k]["gflops"])
                    pt["best"] = {"mode": bk, "speedup_vs_plain_fp64": ok[bk]["speedup_vs_plain_fp64"]}
                key = f"n{n}_kappa{kappa:g}_nrhs{nrhs}"
                out["points"][key] = pt
                print(f"# mixed {key}: plain {plain_s * 1e3:.1f} ms; " +
                      ", ".join(f"{k} {v['ms']:.1f} ms{' (fallback)' if v['fallback'] else ''}"
                                for k, v in pt["modes"].items()), file=sys.stderr, flush=True)
                del b, xp, x
            del A, a
            torch.cuda.empty_cache()
    head = out["points"].get("n16384_kappa100_nrhs1")
    if head and "best" in head:
        out["best_mode_n16384_kappa1e2_nrhs1"] = head["best"]["mode"]
        out["best_speedup_n16384_kappa1e2_nrhs1"] = head["best"]["speedup_vs_plain_fp64"]
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
