#!/usr/bin/env python3
"""Benchmark: NMF factorizations/s on one MI355X.

Workload (the single-large-job configuration): V 20000 x 1000 FP32, planted rank 16
(V = W*.H* (.) exp(N(0, 0.3^2)), W*, H* ~ Gamma(1,1)), k = 16, W0/H0 ~ U(0,1] from Philox,
KL multiplicative updates for a fixed 500 iterations with the convergence test (clamp +
classification + connectivity difference) every 10 iterations.  One step = one whole
factorization (initialisation + 500 iterations) with V already resident in HBM.

Prints one JSON line with the throughput, the HBM roofline of the dominant kernel (HIP
events on the job's stream) and a CPU baseline (the oracle's scalar C rendering of the
reference updates on the host cores, on a bounded sample of iterations).  The same line
carries three more workloads: "sweep" (rank sweep V 5000 x 200, k = 2..10 x 20 restarts,
batched against one job at a time), "dense_rank" (V 50000 x 2000, k = 64, BF16 against FP32
contractions) and "served" (a Poisson stream of factorization requests, p50/p99 latency).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--no-extra]
    python bench.py --served-only --served-requests 2000 --served-rate 4   # the full stream
    python bench.py --sweep-only   # the batched rank sweep alone
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import pathlib
import subprocess
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0

N, M, K, KSTAR, SIGMA, ITERS, T = 20000, 1000, 16, 16, 0.3, 500, 10
SEED_V, SEED_INIT = 20250101, 7
KERNEL_PATHS = {"auto": 0, "valu": 1, "mfma32": 2}   # nmfk_job_set_kernel_path


def load_package():
    spec = importlib.util.spec_from_file_location("bionmf_gpu_amd", ROOT / "bionmf-gpu_amd" / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bionmf_gpu_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def host_cpu() -> dict:
    """The host the CPU baselines ran on: logical CPUs and the processor model."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_baseline(sample_iters: int) -> dict:
    """Oracle's FP32 scalar rendering of the reference update sequence on the host cores
    (TEST INFRASTRUCTURE, timed as a reported baseline only)."""
    exe = ROOT / "oracle" / "nmf_cpu_baseline"
    if not exe.exists():
        return None
    threads = min(16, os.cpu_count() or 1)
    cmd = [str(exe), "--n", str(N), "--m", str(M), "--k", str(K), "--iters", str(sample_iters),
           "--T", str(T), "--S", "1000000000", "--gen", "planted", "--kstar", str(KSTAR),
           "--sigma", str(SIGMA), "--threads", str(threads)]
    out = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=600).stdout
    r = json.loads(out.strip().splitlines()[-1])
    s_iter = r["s_per_iter"]
    return {"value": 1.0 / (ITERS * s_iter), "unit": "factorizations/s", "cores": r["threads"],
            "kind": "port", "host": host_cpu(),
            "sample": f"{sample_iters} iterations of the {N}x{M} k={K} FP32 update loop "
                      f"(s/iter {s_iter:.4f}) extrapolated to {ITERS} iterations"}


def cpu_configs(sweep: dict | None) -> dict:
    """The oracle's FP32 scalar rendering of the reference update sequence on the host cores
    for the other configurations (TEST INFRASTRUCTURE, timed as reported baselines only):
    the 5000 x 38 k = 3 factorization in full on 1 and on all threads (the seeds of
    bench_small_job, so the same factorization); the 20000 x 1000 k = 16
    loop on 1 thread (3 iterations); the rank sweep serially, from each rank's s/iteration
    (20 iterations) times the sweep's own iteration counts; the 50000 x 2000 k = 64 loop
    (2 iterations)."""
    exe = ROOT / "oracle" / "nmf_cpu_baseline"
    if not exe.exists():
        return None
    threads = min(16, os.cpu_count() or 1)

    def run(n, m, k, iters, th, gen, S=1_000_000_000, **kw):
        cmd = [str(exe), "--n", str(n), "--m", str(m), "--k", str(k), "--iters", str(iters), "--T", "10",
               "--S", str(S), "--gen", gen, "--threads", str(th)]
        for a, v in kw.items():
            cmd += ["--" + a.replace("_", "-"), str(v)]
        out = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=600).stdout
        return json.loads(out.strip().splitlines()[-1])

    out = {"kind": "port", "threads_all": threads}
    small = {}
    for th in (1, threads):
        r = run(5000, 38, 3, 2000, th, "lognormal", S=40, mu=5.0, sigma=1.5, seed_v=5, seed_init=7)
        small[f"threads_{th}"] = {"factorizations_per_s": 1.0 / r["seconds"], "iterations": r["iters_done"],
                                  "converged": bool(r["converged"])}
    out["small_5000x38_k3_full_run"] = small
    r = run(N, M, K, 3, 1, "planted", kstar=KSTAR, sigma=SIGMA)
    out["single_large_20000x1000_k16_one_thread"] = {"s_per_iteration": r["s_per_iter"], "sample": "3 iterations"}
    if sweep and "mean_iterations_by_k" in sweep:
        total = 0.0
        per_k = {}
        for k_str, iters in sweep["mean_iterations_by_k"].items():
            r = run(5000, 200, int(k_str), 20, threads, "planted", kstar=6, sigma=0.3)
            per_k[k_str] = r["s_per_iter"]
            total += 20 * iters * r["s_per_iter"]
        out["rank_sweep_serial"] = {"jobs_per_s": 180.0 / total, "s_per_iteration_by_k": per_k,
                                       "sample": "20 iterations per rank on all threads, times the sweep's "
                                                 "mean iterations per rank and 20 restarts"}
    r = run(50000, 2000, 64, 2, threads, "planted", kstar=48, sigma=0.3)
    out["dense_rank_50000x2000_k64"] = {"s_per_iteration": r["s_per_iter"], "sample": "2 iterations on all threads"}
    return out


def bench_small_job(pkg, runs: int = 5) -> dict:
    """The reference's CPU-runnable case: V 5000 x 38 lognormal(5, 1.5^2), k = 3, stop rule
    T = 10, 40 unchanged tests, 2000 iterations at most; whole factorizations (median of
    <runs>), V resident."""
    job = pkg.Job(5000, 38, 3)
    job.gen_V_lognormal(5, 5.0, 1.5)
    job.init_random(7)
    job.run(2000, 10, 40)            # warm-up (captures the iteration graph)
    times, iters = [], 0
    for r in range(runs):
        job.init_random(7)
        t0 = time.perf_counter()
        res = job.run(2000, 10, 40)
        times.append(time.perf_counter() - t0)
        iters = res["iterations"]
    job.close()
    t = float(np.median(times))
    return {"workload": "V 5000x38 lognormal(5, 1.5^2), k=3, T=10, stop_threshold=40, nIters=2000",
            "factorizations_per_s": 1.0 / t, "iterations": iters, "ms_per_iteration": 1e3 * t / max(iters, 1)}


def bench_sweep(pkg) -> dict:
    """Rank sweep: V 5000 x 200 planted rank 6, k = 2..10 x 20 restarts = 180 jobs, stop rule
    (T=10, 40, 2000 iterations); batched (one launch per phase for all jobs, launch shape
    from the autotuner) against the same jobs run one at a time with nmfk_job_run."""
    n, m = 5000, 200
    ks, R, base = list(range(2, 11)), 20, 4242
    src = pkg.Job(n, m, 1)
    src.gen_V_planted(777, 6, 0.3)
    a = pkg.sweep(src, ks, R, base, autotune=True, autotune_iters=10)   # plan and launch shape autotuned
    b = pkg.sweep(src, ks, R, base, plan_target=a["plan_target"], mode=pkg.SWEEP_ONE_AT_A_TIME)
    c = pkg.sweep(src, ks, R, base, mode=pkg.SWEEP_ONE_AT_A_TIME)   # default single-job plans
    src.close()
    nj = len(ks) * R
    same = all(x["iterations"] == y["iterations"] for x, y in zip(a["results"], b["results"]))
    best_one = max(nj / b["seconds"], nj / c["seconds"])
    return {"workload": f"V {n}x{m} planted rank 6, k=2..10 x {R} restarts = {nj} jobs, T=10, "
                        "stop_threshold=40, nIters=2000",
            "batched_jobs_per_s": nj / a["seconds"], "one_at_a_time_jobs_per_s": best_one,
            "speedup": (nj / a["seconds"]) / best_one,
            "one_at_a_time_same_plan_jobs_per_s": nj / b["seconds"],
            "one_at_a_time_default_plan_jobs_per_s": nj / c["seconds"],
            "identical_class_vectors": bool(np.array_equal(a["cls"], b["cls"])) and same,
            "mean_iterations": float(np.mean([r["iterations"] for r in a["results"]])),
            "mean_iterations_by_k": {str(k): float(np.mean([r["iterations"] for r in a["results"][q * R:(q + 1) * R]]))
                                     for q, k in enumerate(ks)},
            "converged_jobs": int(sum(r["converged"] for r in a["results"])),
            "autotune_seconds": a["tune_seconds"], "tuning": a["tuning"], "plan_target": a["plan_target"]}


def bench_sweep_batched(pkg) -> dict:
    """The rank sweep's batched form alone (the workload of bench_sweep), twice, each with its
    own autotuning of plan target and launch shape."""
    n, m = 5000, 200
    ks, R, base = list(range(2, 11)), 20, 4242
    src = pkg.Job(n, m, 1)
    src.gen_V_planted(777, 6, 0.3)
    a = pkg.sweep(src, ks, R, base, autotune=True, autotune_iters=10)
    b = pkg.sweep(src, ks, R, base, autotune=True, autotune_iters=10)
    src.close()
    nj = len(ks) * R
    same = all(x["iterations"] == y["iterations"] for x, y in zip(a["results"], b["results"]))
    return {"workload": f"V {n}x{m} planted rank 6, k=2..10 x {R} restarts = {nj} jobs, T=10, "
                        "stop_threshold=40, nIters=2000",
            "batched_jobs_per_s": [nj / a["seconds"], nj / b["seconds"]],
            "identical_runs": bool(np.array_equal(a["cls"], b["cls"])) and same,
            "plan_target": a["plan_target"], "tuning": a["tuning"]}


def fp64_errors(g, W, H, cls, kl, dv, dvwh) -> dict:
    """Errors of a device run against the FP64 oracle's run of the same workload from the
    same W0, H0 (tests/golden/config*_fp64.npz, written by tests/golden/make_golden.py)."""
    step = int(g["params"][7])
    rows = np.arange(0, W.shape[0], step)
    WH = W[rows].astype(np.float64) @ H.T.astype(np.float64)
    WHr = g["W_rows"].astype(np.float64) @ g["H"].T.astype(np.float64)
    d_ref = np.sqrt(float(g["dot_VWH"]) / float(g["dot_V"]))
    return {"rel_kl": abs(kl - float(g["kl"])) / float(g["kl"]),
            "rel_frobenius_WH_sample": float(np.linalg.norm(WH - WHr) / np.linalg.norm(WHr)),
            "WH_sample_rows": f"every {step}th row",
            "rel_distance": abs(np.sqrt(dvwh / dv) - d_ref) / d_ref,
            "class_agreement": float(np.mean(cls == g["cls"])),
            "identical_class_vector": bool(np.array_equal(cls, g["cls"]))}


def load_fixture(name):
    f = ROOT / "tests" / "golden" / f"{name}.npz"
    return np.load(f) if f.exists() else None


def bench_dense_rank(pkg) -> dict:
    """Dense rank: V 50000 x 2000 planted rank 48, k = 64, 300 iterations: BF16 contractions
    (FP32 accumulation) against the FP32 matrix-core path on the same start, and both
    against the FP64 oracle's run of the same workload."""
    n, m, k, iters = 50000, 2000, 64, 300
    job = pkg.Job(n, m, k)
    job.gen_V_planted(4848, 48, 0.3)
    out = {}
    for path in (3, 2):
        job.set_kernel_path(path)
        job.init_random(9)
        tm = job.time_iterations(10)
        job.init_random(9)
        t0 = time.perf_counter()
        job.run(iters, 10, 10**9)
        dt = time.perf_counter() - t0
        dv, dvwh, kl = job.dot_kl()
        out[path] = {"seconds": dt, "ms_per_iteration": tm["ms_per_iter"],
                     "ms_partial": 0.5 * (tm["ms_partial_H"] + tm["ms_partial_W"]),
                     "kl": kl, "W": job.get_W(), "H": job.get_H(), "cls": job.get_classification(),
                     "dv": dv, "dvwh": dvwh}
    job.close()
    g = load_fixture("config3_fp64")
    vs64 = None
    if g is not None:
        vs64 = {name: fp64_errors(g, o["W"], o["H"], o["cls"], o["kl"], o["dv"], o["dvwh"])
                for name, o in (("bf16", out[3]), ("fp32", out[2]))}
    b, f = out[3], out[2]
    rows = np.arange(0, n, 7)
    WHb = b["W"][rows].astype(np.float64) @ b["H"].T.astype(np.float64)
    WHf = f["W"][rows].astype(np.float64) @ f["H"].T.astype(np.float64)
    part_bytes = 4.0 * (n * m + n * k + m * k)
    return {"workload": f"V {n}x{m} planted rank 48, k={k}, {iters} iterations",
            "bf16_factorizations_per_s": 1.0 / b["seconds"], "fp32_factorizations_per_s": 1.0 / f["seconds"],
            "bf16_ms_per_iteration": b["ms_per_iteration"], "fp32_ms_per_iteration": f["ms_per_iteration"],
            "bf16_partial_GBps": part_bytes / (b["ms_partial"] * 1e-3) / 1e9,
            "fp32_partial_GBps": part_bytes / (f["ms_partial"] * 1e-3) / 1e9,
            "bf16_vs_fp32": {"rel_frobenius_WH_rows_step7": float(np.linalg.norm(WHb - WHf) / np.linalg.norm(WHf)),
                             "rel_kl_diff": abs(b["kl"] - f["kl"]) / f["kl"],
                             "class_agreement": float(np.mean(b["cls"] == f["cls"]))},
            "vs_fp64": vs64}


def bench_row_block_shards(pkg, gpus: int = 8) -> dict:
    """One GPU's share of a row-block split of V over <gpus> GPUs, at its real shape: the
    dense rank (50000 x 2000, k = 64, BF16 contractions) and the headline job (20000 x 1000,
    k = 16).  Each GPU would run these kernels on its rows and all-reduce the H half's
    partial sums (m x k floats) plus colsum(W) (k floats) once per iteration; the all-reduce
    is not run here (one GPU)."""
    out = {}
    for name, (n, m, k, path, kstar) in {"dense_rank": (50000, 2000, 64, 3, 48),
                                         "single_large": (20000, 1000, 16, 0, 16)}.items():
        rows = n // gpus
        job = pkg.Job(rows, m, k)
        job.gen_V_planted(SEED_V + rows, kstar, SIGMA)
        if path:
            job.set_kernel_path(path)
        job.init_random(SEED_INIT)
        tm = min((job.time_iterations(20) for _ in range(3)), key=lambda t: t["ms_per_iter"])
        job.close()
        part_bytes = 4.0 * (rows * m + rows * k + m * k)
        out[name] = {"shard": f"V {rows}x{m} (1/{gpus} of {n} rows), k={k}",
                     "ms_per_iteration": tm["ms_per_iter"],
                     "us_partial_H": 1e3 * tm["ms_partial_H"], "us_partial_W": 1e3 * tm["ms_partial_W"],
                     "partial_GBps": part_bytes / (0.5e-3 * (tm["ms_partial_H"] + tm["ms_partial_W"])) / 1e9,
                     "allreduce_bytes_per_iteration": 4 * (m * k + k)}
    return out


def bench_size_sweep(pkg) -> list:
    """The headline's kernels (k = 16, FP32 matrix cores) on planted inputs of growing height
    at m = 1000 and 2000: the partial launches' duration against their algorithmic bytes.
    The slope between sizes is the kernel's marginal streaming rate; the intercept is the
    fixed cost of one launch."""
    out = []
    for m in (1000, 2000):
        for n in (10000, 20000, 40000, 80000, 160000):
            job = pkg.Job(n, m, K)
            job.gen_V_planted(SEED_V + n + m, KSTAR, SIGMA)
            job.init_random(SEED_INIT)
            tm = min((job.time_iterations(20) for _ in range(3)), key=lambda t: t["ms_per_iter"])
            plan = job.plan()
            job.close()
            b = 4.0 * (n * m + n * K + m * K)
            out.append({"n": n, "m": m, "bytes_per_launch": b,
                        "us_partial_H": 1e3 * tm["ms_partial_H"], "us_partial_W": 1e3 * tm["ms_partial_W"],
                        "GBps_H": b / (tm["ms_partial_H"] * 1e-3) / 1e9, "GBps_W": b / (tm["ms_partial_W"] * 1e-3) / 1e9,
                        "splits_H": plan["H"]["S"], "splits_W": plan["W"]["S"]})
    return out


def bench_streamed(pkg, job, iters: int = 50) -> dict:
    """Out-of-core path: the headline V (20000 x 1000, k = 16) kept in host memory and
    streamed to the device in 4096-row blocks once per iteration, from pinned memory and
    from pageable memory (staged), against the resident job on the same start."""
    V = job.get_V()
    job.init_random(SEED_INIT)
    job.run(iters, T, 1_000_000_000)          # warm-up (captures the iteration graph)
    job.init_random(SEED_INIT)
    t0 = time.perf_counter()
    job.run(iters, T, 1_000_000_000)
    t_res = time.perf_counter() - t0
    kl_res = job.dot_kl()[2]
    out = {"workload": f"V {N}x{M} in host memory, k={K}, {iters} iterations, 4096-row blocks",
           "resident_ms_per_iteration": 1e3 * t_res / iters}
    for pinned in (True, False):
        with pkg.Job.streamed(V, K, 4096, pinned=pinned) as sj:
            sj.init_random(SEED_INIT)
            sj.run(2, T, 1_000_000_000)          # warm-up
            sj.init_random(SEED_INIT)
            t0 = time.perf_counter()
            sj.run(iters, T, 1_000_000_000)
            dt = time.perf_counter() - t0
            info = sj.streamed_info()
            kl = sj.dot_kl()[2]
        out["pinned" if pinned else "staged"] = {
            "ms_per_iteration": 1e3 * dt / iters, "host_to_device_GBps": info["bytes_copied"] / dt / 1e9,
            "nblocks": info["nblocks"], "rel_kl_diff_vs_resident": abs(kl - kl_res) / kl_res}
    return out


def bench_served(pkg, count: int, rate: float) -> dict:
    """Served stream: Poisson arrivals; n log-U[1000, 60000], m log-U[20, 3000], k U{2..32}
    (at most m), restarts U{1..10}, V lognormal(5, 1.5^2); HBM budget 200 GB; stop rule as
    the plain single factorization."""
    reqs = pkg.request_stream(count, rate, 2024)
    out, st = pkg.serve(reqs, progress_every=max(1, count // 20))
    return {"workload": f"{count} requests, Poisson arrivals at {rate}/s, n log-U[1000,60000], "
                        "m log-U[20,3000], k U{2..min(32,m)}, restarts U{1..10}, HBM budget 200 GB",
            "requests_per_s": st["requests_per_s"], "factorizations_per_s": st["factorizations_per_s"],
            "latency_s": {"p50": st["p50"], "p99": st["p99"], "mean": st["mean_latency"]},
            "seconds": st["seconds"], "peak_jobs": st["peak_jobs"], "peak_GB": st["peak_bytes"] / 1e9,
            "restarts": int(sum(r[3] for r in reqs)),
            "mean_iterations_per_restart": float(sum(o["iterations_sum"] for o in out) / sum(r[3] for r in reqs)),
            "serving_thread_s": {"step": st["step_s"], "admit": st["admit_s"], "wait_for_device": st["wait_s"],
                                 "retire": st["retire_s"], "chunks": st["chunks"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=30)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the rank-sweep, dense-rank and served-stream measurements")
    ap.add_argument("--served-requests", type=int, default=100)
    ap.add_argument("--served-rate", type=float, default=5.0)
    ap.add_argument("--served-only", action="store_true",
                    help="run only the served stream (e.g. the full 2000-request stream) and print its line")
    ap.add_argument("--streamed-only", action="store_true",
                    help="run only the out-of-core (V in host memory) workload and print its line")
    ap.add_argument("--size-sweep", action="store_true",
                    help="run only the headline kernels on inputs of growing height (m = 1000, 2000) and print their rates")
    ap.add_argument("--shards-only", action="store_true",
                    help="run only one GPU's row block of an 8-way split (dense rank and headline shapes)")
    ap.add_argument("--dense-only", action="store_true",
                    help="run only the dense-rank (k = 64, BF16 vs FP32) workload and print its line")
    ap.add_argument("--sweep-only", action="store_true",
                    help="run only the batched rank sweep (twice, autotuned) and print its line; "
                         "the one-at-a-time comparison stays in the full run")
    ap.add_argument("--kernel-path", choices=sorted(KERNEL_PATHS), default="auto",
                    help="contraction kernels of the single large job (auto: FP32 matrix cores at k = 16); "
                         "valu times the vector-ALU kernels on the same workload")
    args = ap.parse_args()
    if args.steps < 1 or args.warmup < 0:
        ap.error("--steps must be >= 1 and --warmup >= 0")
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (the multi-GPU row-block split "
                                   "of V is designed, not built); --gpus must be 1"}))
        sys.exit(2)

    pkg = load_package()
    pkg.load()
    if pkg.gpu_count() < 1:
        print(json.dumps({"error": "no HIP device visible"}))
        sys.exit(1)

    if args.dense_only:
        d = bench_dense_rank(pkg)
        print(json.dumps({"metric": "dense rank: factorizations/s (BF16 contractions)",
                          "value": d["bf16_factorizations_per_s"], "unit": "factorizations/s", "n_gpus": 1,
                          "higher_is_better": True, "dense_rank": d}))
        return
    if args.sweep_only:
        r = bench_sweep_batched(pkg)
        print(json.dumps({"metric": "rank sweep, batched: jobs/s", "value": r["batched_jobs_per_s"][-1],
                          "unit": "jobs/s", "n_gpus": 1, "higher_is_better": True, "sweep": r}))
        return
    if args.size_sweep:
        r = bench_size_sweep(pkg)
        best = max(max(x["GBps_H"], x["GBps_W"]) for x in r)
        print(json.dumps({"metric": "k = 16 partial kernel: GB/s by input height", "value": best, "unit": "GB/s",
                          "n_gpus": 1, "higher_is_better": True, "size_sweep": r}))
        return
    if args.shards_only:
        r = bench_row_block_shards(pkg)
        print(json.dumps({"metric": "one GPU's row block of an 8-way split: ms per iteration",
                          "value": r["dense_rank"]["ms_per_iteration"], "unit": "ms/iteration", "n_gpus": 1,
                          "higher_is_better": False, "row_block_shards": r}))
        return
    if args.served_only:
        s = bench_served(pkg, args.served_requests, args.served_rate)
        print(json.dumps({"metric": "served stream: requests/s, p50/p99 latency", "value": s["requests_per_s"],
                          "unit": "requests/s", "n_gpus": 1, "higher_is_better": True, "served": s}))
        return

    job = pkg.Job(N, M, K)
    job.gen_V_planted(SEED_V, KSTAR, SIGMA)
    job.set_kernel_path(KERNEL_PATHS[args.kernel_path])
    if args.streamed_only:
        o = bench_streamed(pkg, job)
        print(json.dumps({"metric": "out-of-core: ms per iteration with V streamed from host memory",
                          "value": o["pinned"]["ms_per_iteration"], "unit": "ms/iteration", "n_gpus": 1,
                          "higher_is_better": False, "streamed": o}))
        return

    def step(s):
        job.init_random(SEED_INIT + s)
        t0 = time.perf_counter()
        r = job.run(ITERS, T, 1_000_000_000)
        return time.perf_counter() - t0, r

    for s in range(args.warmup):
        step(s)
    job.sync()
    lat = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        dt, r = step(1000 + s)
        lat.append(dt)
    job.sync()
    elapsed = time.perf_counter() - t0
    assert r["iterations"] == ITERS

    dv, dvwh, kl = job.dot_kl()
    value = args.steps / elapsed

    # Accuracy of the measured path: one more factorization (untimed) from the FP64 fixture's
    # starting point, compared with the oracle's FP64 run of the same workload.
    accuracy = None
    g1 = load_fixture("config1_fp64")
    if g1 is not None and args.kernel_path == "auto":
        job.init_random(int(g1["params"][5]))
        ra = job.run(ITERS, T, 1_000_000_000)
        adv, advwh, akl = job.dot_kl()
        accuracy = fp64_errors(g1, job.get_W(), job.get_H(), job.get_classification(), akl, adv, advwh)
        accuracy["iterations"] = ra["iterations"]

    # Dominant-kernel roofline.  Both half-steps launch the same partial (contraction)
    # kernel, and each launch reads V (or its transposed copy) once plus both factors, so
    # they share one algorithmic byte count; the average launch duration comes from HIP
    # events around back-to-back launches on the job's stream.
    tm = job.time_iterations(50)
    bytes_partial = 4.0 * (N * M + N * K + M * K)
    ms = 0.5 * (tm["ms_partial_H"] + tm["ms_partial_W"])
    achieved = bytes_partial / (ms * 1e-3) / 1e9
    iter_bytes = 4.0 * (2 * N * M + 3 * N * K + 3 * M * K)
    read_gbs, copy_gbs = pkg.measure_stream(1 << 30, 10, 4096)
    traffic = None
    tfile = ROOT / "profiles" / "traffic_partial_mfma_20000x1000_k16.json"
    if tfile.exists() and args.kernel_path == "auto":
        traffic = json.loads(tfile.read_text()).get("bytes_per_launch")

    # The one-pass schedule (one pass over V per iteration: k_fused_mfma + k_fused_finish) on
    # the same job, timed the same way; reported beside the default two-pass schedule.
    one_pass = None
    if args.kernel_path in ("auto", "mfma32"):
        job.set_schedule(job.SCHED_ONE_PASS)
        job.init_random(SEED_INIT)
        job.time_one_pass(5)
        t1 = job.time_one_pass(50)
        job.init_random(SEED_INIT)
        r1 = job.run(ITERS, T, 1_000_000_000)
        pass_bytes = 4.0 * (N * M + 2 * N * K + 2 * M * K)
        one_pass = {"ms_per_iteration": t1["ms_per_iter"], "ms_pass": t1["ms_pass"], "ms_finish": t1["ms_finish"],
                    "workgroups": t1["workgroups"], "waves_per_workgroup": t1["waves_per_workgroup"],
                    "pass_bytes": pass_bytes, "pass_GBps": pass_bytes / (t1["ms_pass"] * 1e-3) / 1e9,
                    "pass_matrix_TFLOPs": 8.0 * N * M * K / (t1["ms_pass"] * 1e-3) / 1e12,
                    "factorization_s": r1["seconds"]}
        job.set_schedule(job.SCHED_AUTO)

    extra = {}
    if not args.no_extra:
        extra["small_job"] = bench_small_job(pkg)
        extra["sweep"] = bench_sweep(pkg)
        extra["dense_rank"] = bench_dense_rank(pkg)
        extra["served"] = bench_served(pkg, args.served_requests, args.served_rate)
        extra["streamed"] = bench_streamed(pkg, job)
        extra["row_block_shards"] = bench_row_block_shards(pkg)

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args.cpu_iters)
        except Exception as e:  # reported, never fatal for the GPU number
            cpu = {"error": str(e)}
        try:
            extra["cpu_configs"] = cpu_configs(extra.get("sweep"))
        except Exception as e:
            extra["cpu_configs"] = {"error": str(e)}

    line = {
        "metric": "NMF factorizations/sec at <=1e-3 rel KL err; p50/p99 job latency; % HBM roofline",
        "value": value,
        "unit": "factorizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (planted rank-16 V generated on the device from a Philox seed)",
        "config": {"workload": f"single large job: V {N}x{M} planted rank {KSTAR} (Gamma(1,1) factors, "
                               f"lognormal sigma={SIGMA} noise), k={K}, {ITERS} iterations, test every {T}",
                   "n": N, "m": M, "k": K, "iterations": ITERS, "niter_test_conv": T},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "ms_per_iteration": tm["ms_per_iter"],
        "kernels_ms": {k_: v for k_, v in tm.items() if k_ != "ms_per_iter"},
        "iteration_GBps": iter_bytes / (tm["ms_per_iter"] * 1e-3) / 1e9,
        "kl_final": kl,
        "frobenius_ratio": float(np.sqrt(dvwh / dv)),
        "vs_fp64": accuracy,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": ("k_partial_mfma<16,1,4,16> (H half) and k_partial_mfma<16,2,2,16> (W half)"
                                if args.kernel_path in ("auto", "mfma32")
                                else "k_partial<16,...> (vector ALUs, both half-steps)"),
                     "bytes_per_launch": bytes_partial, "ms_per_launch": ms,
                     "measured_read_GBps": read_gbs, "measured_copy_GBps": copy_gbs,
                     "frac_of_measured_read": achieved / read_gbs},
        "cpu_baseline": cpu,
        "one_pass_schedule": one_pass,
    }
    line.update(extra)
    job.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
