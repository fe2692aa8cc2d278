"""Benchmark of the MI355X multiplicative BLD solver (bld_mult).

Metric: S-element updates/s = sum over instances of x*y*z*iterations / wall
time, for the batched FP32 solver, beside the plain FP64 form, plus solve
latency percentiles under serving (--serving).

Headline workload (one "step" = one full batched solve from S0):
  256 independent instances, x, y ~ U{32..512}, z ~ U{4..32}, X sampled
  from the allocation model with a ~ U[10, 80], b = 1, alpha = 1, beta = 1,
  200 iterations each, FP32 bulk with psi_appr. Inputs are resident in HBM
  before the timed region; S restarts from the resident S0 every step.

Also measured (same JSON line):
  * fp64_plain_same_workload: the plain FP64 form (exact digamma) solving the
    same 256 instances one at a time, and the ratio to the batched FP32 rate;
    fp64_batched_same_workload: the same FP64 form solving the 256 instances
    as one batch (the ratio against it is reported too);
  * config1_fp64: one 1024x1024x32 instance, 500 iterations, exact digamma;
  * config2_fp32: one 2048x2048x64 instance, 300 iterations, psi_appr, with
    its normwise error and log-marginal gap against the FP64 form;
  * bld_add_batched_fp32: the sibling additive solver (bld::bld_add) on the
    headline's 256 instances, with the roofline of its pass;
  * bld_appr_batched_fp32: the sibling approximate solver (bld::bld_appr) on
    the same instances;
  * roofline of the fused pass kernel of the headline batch (HIP events on
    the batch stream): algorithmic bytes 2*x*y*z*4 + x*y*4 per launch, beside
    the rate plain streaming kernels reach on the same GPU (stream_ceiling:
    bnmf-algs_amd/lib/stream_bw, non-temporal copy of 1 GiB);
  * cpu_baseline: the FP64 C++/OpenMP oracle (a port of the reference CPU
    solver, which cannot be built here) on the reference's CPU config
    100x100x17, 1000 iterations, exact digamma.

  * serving: Poisson arrivals of config-4 requests (x, y log-uniform in
    [64, 2048], z log-uniform in [8, 64], FP32, 200 iterations) into the
    continuous-batching server at half its measured saturation rate, with
    p50/p99 latency and solves/s beside the one-at-a-time FP64 loop over the
    request pool (--requests, default 10000 as the configuration states).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--requests N]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bnmf-algs_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_c3(hip, n=256, seed=0xB1D0, iters=200):
    rng = np.random.default_rng(seed)
    insts = []
    for t in range(n):
        x, y = int(rng.integers(32, 513)), int(rng.integers(32, 513))
        z = int(rng.integers(4, 33))
        a = float(rng.uniform(10, 80))
        p = hip.Params(a, 1.0, [1.0] * x, [1.0] * z)
        X = hip.sample_X(seed + 17 * t, (x, y, z), p)
        S0 = hip.init_S(seed + 17 * t + 1, X, z)
        insts.append(hip.Instance(X=X, z=z, params=p, S0=S0))
    return insts


def single(hip, x, y, z, a, seed):
    p = hip.Params(a, 1.0, [1.0] * x, [1.0] * z)
    X = hip.sample_X(seed, (x, y, z), p)
    S0 = hip.init_S(seed + 1, X, z)
    return hip.Instance(X=X, z=z, params=p, S0=S0)


def as_dtype(inst, hip, dt):
    return hip.Instance(X=inst.X.astype(dt), z=inst.z, params=inst.params,
                        S0=inst.S0.astype(dt), seed=inst.seed)


def timed_solves(b, iters, steps, warmup):
    for _ in range(warmup):
        b.solve(iters)
    b.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.solve_async(iters)
    b.synchronize()
    return (time.perf_counter() - t0) / steps


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def make_pool(hip, n, seed=0xB1D4):
    """Config-4 request inputs: x, y log-uniform in [64, 2048], z log-uniform
    in [8, 64], allocation-model X with a = 40, b = 1, alpha = beta = 1."""
    rng = np.random.default_rng(seed)
    pool = []
    for t in range(n):
        x = int(round(np.exp(rng.uniform(np.log(64), np.log(2048)))))
        y = int(round(np.exp(rng.uniform(np.log(64), np.log(2048)))))
        z = int(round(np.exp(rng.uniform(np.log(8), np.log(64)))))
        p = hip.Params(40.0, 1.0, [1.0] * x, [1.0] * z)
        X = hip.sample_X(seed + 31 * t, (x, y, z), p).astype(np.float32)
        pool.append(hip.Instance(X=X, z=z, params=p, seed=seed + 31 * t + 1))
    return pool


def serving(hip, n_requests, rate_frac, log, iters=200, pool_size=48):
    """Poisson arrivals into the continuous-batching server (FP32, psi_appr,
    200 iterations each, S0 drawn on the GPU). Requests cycle through a
    seeded pool of distinct inputs. Saturation is measured first with a
    burst; the stream then arrives at rate_frac of it. Latency = completion
    minus submission (S resident in HBM), steady clock."""
    import threading
    log(f"[bench] serving: building a pool of {pool_size} request inputs")
    pool = make_pool(hip, pool_size)
    elems = [i.X.shape[0] * i.X.shape[1] * i.z for i in pool]
    budget = 96 << 30
    # saturation: a burst of pool_size requests, all queued at once
    with hip.Server(budget, dtype=np.float32, use_psi_appr=True) as srv:
        t0 = time.perf_counter()
        ts = [srv.submit(i, iters) for i in pool]
        for t in ts:
            srv.wait(t)
            srv.collect(t, want_S=False)
        t_burst = time.perf_counter() - t0
    sat = pool_size / t_burst
    rate = rate_frac * sat
    log(f"[bench] serving: saturation {sat:.1f} req/s ({sum(elems)*iters/t_burst:.3e} upd/s); "
        f"stream at {rate:.1f} req/s, {n_requests} requests")
    rng = np.random.default_rng(7)
    gaps = rng.exponential(1.0 / rate, n_requests)
    picks = rng.integers(0, pool_size, n_requests)
    lat, done_elems = [], 0
    with hip.Server(budget, dtype=np.float32, use_psi_appr=True) as srv:
        tickets = [None] * n_requests
        ready = threading.Semaphore(0)

        def collector():
            for k in range(n_requests):
                ready.acquire()
                srv.wait(tickets[k])
                _, _, st = srv.collect(tickets[k], want_S=False)
                lat.append(st["done_s"] - st["submit_s"])
                if (k + 1) % max(1, n_requests // 10) == 0:
                    log(f"[bench] serving: {k + 1}/{n_requests} done")

        th = threading.Thread(target=collector)
        th.start()
        t_start = time.perf_counter()
        next_t = t_start
        for k in range(n_requests):
            next_t += gaps[k]
            d = next_t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            inst = pool[picks[k]]
            tickets[k] = srv.submit(hip.Instance(X=inst.X, z=inst.z, params=inst.params,
                                                 seed=inst.seed + k), iters)
            done_elems += elems[picks[k]]
            ready.release()
        th.join()
        wall = time.perf_counter() - t_start
        st = srv.stats()
    lat = np.array(lat)
    # the plain one-at-a-time FP64 GPU loop over the whole request pool
    n64 = len(pool)
    t0 = time.perf_counter()
    lat64 = []
    for inst in pool:
        t1 = time.perf_counter()
        hip.bld_mult(inst.X.astype(np.float64), inst.z, inst.params, iters, False,
                     seed=inst.seed, dtype=np.float64)
        lat64.append(time.perf_counter() - t1)
    t64 = time.perf_counter() - t0
    res = {
        "requests": n_requests, "arrival": "poisson", "rate_req_per_s": rate,
        "rate_frac_of_saturation": rate_frac, "saturation_req_per_s": sat,
        "solves_per_s": n_requests / wall, "updates_per_s": done_elems * iters / wall,
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3),
        "peak_active": st["peak_active"], "peak_hbm_bytes": st["peak_bytes"],
        "hbm_budget_bytes": budget, "iterations": iters, "dtype": "f32", "pool": pool_size,
        "fp64_one_at_a_time": {"solves_per_s": n64 / t64,
                               "p50_ms": float(np.percentile(lat64, 50) * 1e3),
                               "requests": n64, "digamma": "exact"},
    }
    log(f"[bench] serving: {res['solves_per_s']:.1f} solves/s, p50 {res['p50_ms']:.1f} ms, "
        f"p99 {res['p99_ms']:.1f} ms; FP64 one-at-a-time {n64/t64:.2f} solves/s")
    return res


def measured_traffic():
    """HBM bytes per pass of the headline workload from the committed rocprof
    FETCH_SIZE/WRITE_SIZE runs (profiles/bld_pass_traffic.json) and where they
    come from, or (None, None). The figure is read from the file, not
    measured in this run: the source names the profile it was taken from."""
    path = os.path.join(ROOT, "profiles", "bld_pass_traffic.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
        src = {"file": "profiles/bld_pass_traffic.json", "profile": d.get("source"),
               "measured_in_this_run": False,
               "counters": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, scaled by a 1 GiB copy"}
        return float(d["traffic_bytes_per_pass"]), src
    except (OSError, KeyError, ValueError):
        return None, None


def stream_ceiling():
    """Bandwidth of plain streaming kernels on this GPU (bnmf-algs_amd/lib/
    stream_bw, run as a child process): the practical read + write ceiling
    that the update pass, one read and one write of S, is held against."""
    exe = os.path.join(ROOT, "bnmf-algs_amd", "lib", "stream_bw")
    if not os.path.exists(exe):
        return None
    try:
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120, check=True)
        return json.loads(r.stdout.strip().splitlines()[-1])
    except (subprocess.SubprocessError, ValueError, IndexError) as e:
        log(f"[bench] stream_bw failed: {e}")
        return None


def measure_add(hip, c3f, elems3, fibers3, iters3):
    """bld::bld_add batched FP32 on the headline's instances: solve rate and
    the roofline of its pass (same algorithmic bytes as bld_mult's)."""
    log("[bench] batched FP32 bld_add on the same instances")
    ba = hip.Batch(c3f, dtype=np.float32, algo=hip.ALGO_ADD)
    ta = timed_solves(ba, iters3, 1, 1)
    pa, ia = ba.time_pass(20)
    alg_bytes = 2 * elems3 * 4 + fibers3 * 4
    ach_a = alg_bytes / (pa * 1e-3) / 1e9
    ba.close()
    log(f"[bench] bld_add FP32: {ta*1e3:.1f} ms/solve, pass {pa*1e3:.1f} us ({ach_a:.0f} GB/s)")
    return {
        "value": elems3 * iters3 / ta, "unit": "S-element updates/s",
        "ms_per_solve": ta * 1e3, "pass_us": pa * 1e3, "iter_us": ia * 1e3,
        "roofline": {"bound": "hbm", "achieved": ach_a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_a / HBM_PEAK_GBS,
                     "kernel": "bld_pass_f32z_kernel additive step, all launches of one iteration"},
        "workload": "the headline's 256 instances, 200 iterations of bld::bld_add"}


def measure_appr(hip, c3f, elems3, iters3):
    """bld_appr (FP32) on the headline's 256 instances: one timed solve of
    iters3 iterations and the graph-timed iteration."""
    log("[bench] batched FP32 bld_appr on the same instances")
    bp = hip.ApprBatch(c3f, dtype=np.float32)
    bp.solve(2)
    t0 = time.perf_counter()
    bp.solve(iters3)
    tp = time.perf_counter() - t0
    ip = bp.time_iteration(10)
    bp.close()
    log(f"[bench] bld_appr FP32: {tp*1e3:.1f} ms/solve, iteration {ip*1e3:.1f} us")
    return {"value": elems3 * iters3 / tp, "unit": "S-element updates/s",
            "ms_per_solve": tp * 1e3, "iter_us": ip * 1e3,
            "bound": "latency and VALU (S is formed from the factors in registers; HBM traffic "
                     "is X and the factors only)",
            "workload": "the headline's 256 instances, 200 iterations of bld::bld_appr, "
                        "factors drawn from Dirichlet(1)"}


def measure_c1(hip, iters=500, solves=1, reps=20):
    """One 1024x1024x32 instance in the plain FP64 form (exact digamma):
    solve time and the graph-timed update pass (algorithmic bytes 2*x*y*z*8
    + x*y*8 per pass)."""
    log(f"[bench] config 1: 1024x1024x32 FP64 exact, {iters} iterations")
    c1 = single(hip, 1024, 1024, 32, 40.0, 0xB1D1)
    b1 = hip.Batch([c1], dtype=np.float64, use_psi_appr=False)
    t1 = timed_solves(b1, iters, solves, 1 if solves else 0) if solves else float("nan")
    p1, it1 = b1.time_pass(reps)
    b1.close()
    e1 = 1024 * 1024 * 32
    alg = 2 * e1 * 8 + 1024 * 1024 * 8
    res = {"value": e1 * iters / t1, "unit": "S-element updates/s", "ms_per_solve": t1 * 1e3,
           "iterations": iters, "pass_us": p1 * 1e3, "iter_us": it1 * 1e3,
           "pass_GBps": alg / (p1 * 1e-3) / 1e9, "pass_frac_of_peak": alg / (p1 * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "algorithmic_bytes_per_pass": alg}
    log(f"[bench] config1: {t1*1e3:.1f} ms, pass {p1*1e3:.1f} us ({res['pass_GBps']:.0f} GB/s)")
    return res


def measure_c2(hip, iters=300, reps=20, accuracy=True):
    """One 2048x2048x64 instance in the FP32-bulk form (psi_appr): solve time,
    the graph-timed pass and, with accuracy, the normwise error of S and the
    log-marginal gap against the FP64 form (psi_appr) on the same S0."""
    log(f"[bench] config 2: 2048x2048x64 FP32 psi_appr, {iters} iterations")
    c2 = single(hip, 2048, 2048, 64, 40.0, 0xB1D2)
    b2 = hip.Batch([as_dtype(c2, hip, np.float32)], dtype=np.float32, use_psi_appr=True)
    t2 = timed_solves(b2, iters, 1, 1)
    res = {}
    if accuracy:
        S2 = b2.S(0)
        lm2 = b2.log_marginal()[0]
        b2.solve(iters)
    p2, it2 = b2.time_pass(reps)
    b2.close()
    if accuracy:
        b2d = hip.Batch([c2], dtype=np.float64, use_psi_appr=True)
        b2d.solve(iters)
        S2d = b2d.S(0)
        lm2d = b2d.log_marginal()[0]
        b2d.close()
        res = {"rel_err_S_vs_fp64": rel_err(S2, S2d),
               "log_marginal_gap_vs_fp64": abs(lm2 - lm2d) / abs(lm2d)}
        del S2, S2d
    if accuracy:
        # the same solve through the plain call, which takes host buffers:
        # uploads X and S0, solves, downloads S (PCIe-inclusive; one call
        # after a warm-up call)
        c2f = as_dtype(c2, hip, np.float32)
        hip.bld_mult(c2f.X, 64, c2f.params, 2, True, S0=c2f.S0, dtype=np.float32)
        t0 = time.perf_counter()
        hip.bld_mult(c2f.X, 64, c2f.params, iters, True, S0=c2f.S0, dtype=np.float32)
        th = time.perf_counter() - t0
        res["host_buffers"] = {"ms_per_solve": th * 1e3, "updates_per_s": 2048 * 2048 * 64 * iters / th,
                               "mode": "bnmf_bld_mult_f32 on host X and S0, S returned to the host"}
        del c2f
    e2 = 2048 * 2048 * 64
    alg = 2 * e2 * 4 + 2048 * 2048 * 4
    res = {"value": e2 * iters / t2, "unit": "S-element updates/s", "ms_per_solve": t2 * 1e3,
           "iterations": iters, "pass_us": p2 * 1e3, "iter_us": it2 * 1e3,
           "pass_GBps": alg / (p2 * 1e-3) / 1e9, "pass_frac_of_peak": alg / (p2 * 1e-3) / 1e9 / HBM_PEAK_GBS,
           **res}
    log(f"[bench] config2: {t2*1e3:.1f} ms, pass {p2*1e3:.1f} us ({res['pass_GBps']:.0f} GB/s)")
    return res


def cpu_baseline(hip):
    from oracle import oracle
    inst = single(hip, 100, 100, 17, 40.0, 0xB1D0)
    iters = 1000
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.bld_mult(inst.X, inst.z, inst.params.alpha, inst.params.beta, inst.S0,
                        max_iter=iters)
        times.append(time.perf_counter() - t0)
    med = sorted(times)[1]
    upd = 100 * 100 * 17 * iters
    # the sibling solvers' oracles on the same input (bld_add's per-fiber
    # loop is OpenMP-parallel, bld_appr's restatement is sequential)
    t0 = time.perf_counter()
    oracle.bld_add(inst.X, inst.z, inst.params.alpha, inst.params.beta, inst.S0, max_iter=200)
    t_add = time.perf_counter() - t0
    rng = np.random.default_rng(0xB1D0)
    nu0 = rng.dirichlet(np.ones(17), size=100)
    mu0 = rng.dirichlet(np.ones(17), size=100).T.copy()
    t0 = time.perf_counter()
    oracle.bld_appr(inst.X, inst.z, inst.params.alpha, inst.params.beta, nu0, mu0, max_iter=50)
    t_appr = time.perf_counter() - t0
    return {"value": upd / med, "unit": "S-element updates/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": "reference CPU config: X 100x100 (allocation model a=40, b=1, alpha=1, "
                      "beta=1), z=17, 1000 iterations, exact digamma; median of 3 runs "
                      f"({med:.2f} s each); FP64 C++/OpenMP oracle",
            "siblings": {
                "bld_add": {"value": 100 * 100 * 17 * 200 / t_add, "unit": "S-element updates/s",
                            "cores": oracle.num_threads(),
                            "sample": f"same X and S0, 200 iterations ({t_add:.2f} s)"},
                "bld_appr": {"value": 100 * 100 * 17 * 50 / t_appr, "unit": "S-element updates/s",
                             "cores": 1,
                             "sample": f"same X, Dirichlet(1) factors, 50 iterations ({t_appr:.2f} s)"}}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving workload")
    ap.add_argument("--serving", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--requests", type=int, default=10000,
                    help="serving requests (the configuration's 10000 by default)")
    ap.add_argument("--rate-frac", type=float, default=0.5,
                    help="serving arrival rate as a fraction of measured saturation")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-loop", action="store_true")
    ap.add_argument("--no-add", action="store_true", help="skip the bld_add measurement")
    ap.add_argument("--no-appr", action="store_true", help="skip the bld_appr measurement")
    ap.add_argument("--add-only", action="store_true",
                    help="only the batched FP32 bld_add measurement (e.g. for profiling its pass)")
    ap.add_argument("--appr-only", action="store_true",
                    help="only the batched FP32 bld_appr measurement (e.g. for profiling its passes)")
    ap.add_argument("--only-headline", action="store_true")
    ap.add_argument("--only", choices=("c1", "c2"), default=None,
                    help="only one single-instance configuration (e.g. for profiling its pass): "
                         "c1 = 1024x1024x32 FP64 exact, c2 = 2048x2048x64 FP32")
    ap.add_argument("--iters", type=int, default=None,
                    help="with --only: iterations of the solve (default: the configuration's)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="plan the headline batch with the built-in tile rule")
    ap.add_argument("--tune-file", default=None,
                    help="load this tile table instead of autotuning (e.g. for profiling runs)")
    ap.add_argument("--save-tune", default=None, help="save the autotuned tile table here")
    ap.add_argument("--stream", action="store_true",
                    help="with --copy-calibration: also run the streaming-kernel ceiling")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 1 GiB (scales rocprof FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        # 5 device copies of 1 GiB (the rocprof traffic scale), then the
        # read + write bandwidth a plain copy reaches, timed with events
        import torch
        n = 256 << 20
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 20
        print(json.dumps({"copy_bytes_per_launch": 4 * n, "launches": 25,
                          "copy_GBps_read_plus_write": 2 * 4 * n / (ms * 1e-3) / 1e9}), flush=True)
        if args.stream:
            print(json.dumps({"stream_bw": stream_ceiling()}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(instances would shard across GPUs with no communication).", file=sys.stderr)
        sys.exit(2)

    import bnmf_hip as hip
    hip.init(0)
    out = {}
    if args.only == "c1":
        print(json.dumps({"config1_fp64": measure_c1(hip, iters=args.iters or 500)}), flush=True)
        return
    if args.only == "c2":
        print(json.dumps({"config2_fp32": measure_c2(hip, iters=args.iters or 300, accuracy=False)}),
              flush=True)
        return

    # ----------------------------------------------------- headline: config 3
    log("[bench] building 256-instance batch")
    c3 = make_c3(hip)
    iters3 = 200
    elems3 = sum(i.X.shape[0] * i.X.shape[1] * i.z for i in c3)
    fibers3 = sum(i.X.shape[0] * i.X.shape[1] for i in c3)
    c3f = [as_dtype(i, hip, np.float32) for i in c3]
    tuned = None
    if args.tune_file:
        hip.tune_load(args.tune_file)
        tuned = {"table": args.tune_file}
    elif not args.no_autotune:
        # tile heights per shape class, chosen by the autotuner on this
        # workload (untimed setup; the table applies to batches created after)
        log("[bench] autotuning tile heights on the batch (whole-iteration timings)")
        t0 = time.perf_counter()
        res = hip.autotune_workload(np.float32, c3f, candidates=(16, 32, 64), reps=10,
                                    trials=2, refine=True, log=log)
        tuned = {"seconds": round(time.perf_counter() - t0, 1), "classes": len(res["rows"]),
                 "evaluations": res["evaluations"], "iter_us": round(res["iter_ms"] * 1e3, 1),
                 "rows": {f"z{k[0]}f{k[1]}": v for k, v in sorted(res["rows"].items())}}
        if args.save_tune:
            hip.tune_save(args.save_tune)
    if args.add_only:
        print(json.dumps({"bld_add_batched_fp32": measure_add(hip, c3f, elems3, fibers3, iters3)}),
              flush=True)
        return
    if args.appr_only:
        print(json.dumps({"bld_appr_batched_fp32": measure_appr(hip, c3f, elems3, iters3)}), flush=True)
        return
    b32 = hip.Batch(c3f, dtype=np.float32, use_psi_appr=True)
    log("[bench] timing batched FP32 solves")
    t_step = timed_solves(b32, iters3, args.steps, args.warmup)
    value = elems3 * iters3 / t_step
    S32 = [b32.S(n) for n in range(len(c3f))]
    lm32 = b32.log_marginal()
    b32.solve(iters3)  # restore S (log_marginal reuses the partial buffers)
    pass_ms, iter_ms = b32.time_pass(50)
    alg_bytes = 2 * elems3 * 4 + fibers3 * 4
    achieved = alg_bytes / (pass_ms * 1e-3) / 1e9
    info = b32.info()
    log(f"[bench] batched FP32: {t_step*1e3:.2f} ms/solve, {value:.3e} upd/s, "
        f"pass {pass_ms*1e3:.1f} us ({achieved:.0f} GB/s), iter {iter_ms*1e3:.1f} us")

    # ------------------------------------- plain FP64, same instances, one by one
    fp64 = None
    if not args.no_fp64_loop and not args.only_headline:
        log("[bench] plain FP64 loop over the same 256 instances")
        c3d = [as_dtype(i, hip, np.float64) for i in c3]
        singles = [hip.Batch([i], dtype=np.float64, use_psi_appr=False) for i in c3d]
        for s in singles[:4]:
            s.solve(2)
        t0 = time.perf_counter()
        for s in singles:
            s.solve(iters3)
        t64 = time.perf_counter() - t0
        errs, gaps = [], []
        for n, s in enumerate(singles):
            S64 = s.S(0)
            errs.append(rel_err(S32[n], S64))
        lm64 = np.array([s.log_marginal()[0] for s in singles])
        gaps = np.abs(lm32 - lm64) / np.abs(lm64)
        for s in singles:
            s.close()
        fp64 = {"value": elems3 * iters3 / t64, "unit": "S-element updates/s",
                "seconds": t64, "dtype": "f64", "digamma": "exact",
                "mode": "one instance per solve call, 256 calls in sequence"}
        out["fp64_plain_same_workload"] = fp64
        out["ratio_vs_fp64_plain"] = value / fp64["value"]
        out["accuracy_vs_fp64"] = {"max_rel_err_S_normwise": float(max(errs)),
                                   "max_log_marginal_gap": float(gaps.max()),
                                   "instances": len(errs)}
        log(f"[bench] FP64 plain loop: {t64:.2f} s, {fp64['value']:.3e} upd/s; "
            f"max rel err {max(errs):.2e}, max lm gap {gaps.max():.2e}")
        # the same FP64 form with the 256 instances in one batch: the ratio
        # against a batched FP64 solve, not only against per-call launches
        log("[bench] FP64 batch of the same 256 instances")
        b64 = hip.Batch(c3d, dtype=np.float64, use_psi_appr=False)
        t64b = timed_solves(b64, iters3, 1, 1)
        p64, i64 = b64.time_pass(10)
        b64.close()
        out["fp64_batched_same_workload"] = {
            "value": elems3 * iters3 / t64b, "unit": "S-element updates/s",
            "ms_per_solve": t64b * 1e3, "pass_us": p64 * 1e3, "iter_us": i64 * 1e3,
            "pass_GBps": (2 * elems3 * 8 + fibers3 * 8) / (p64 * 1e-3) / 1e9,
            "dtype": "f64", "digamma": "exact", "mode": "the 256 instances as one batch"}
        out["ratio_vs_fp64_batched"] = value / out["fp64_batched_same_workload"]["value"]
        log(f"[bench] FP64 batch: {t64b*1e3:.1f} ms/solve, pass {p64*1e3:.1f} us")

    # ------------------------ sibling solver: bld_add on the same 256 instances
    if not args.only_headline and not args.no_add:
        out["bld_add_batched_fp32"] = measure_add(hip, c3f, elems3, fibers3, iters3)

    # ----------------------- sibling solver: bld_appr on the same 256 instances
    if not args.only_headline and not args.no_appr:
        out["bld_appr_batched_fp32"] = measure_appr(hip, c3f, elems3, iters3)

    # ------------------------------------------- configs 1 (FP64) and 2 (FP32)
    if not args.only_headline:
        out["config1_fp64"] = measure_c1(hip)
        out["config2_fp32"] = measure_c2(hip)

    if not args.no_serving and (args.serving or not args.only_headline):
        out["serving"] = serving(hip, args.requests, args.rate_frac, log)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (FP64 oracle, reference CPU config)")
        cpu = cpu_baseline(hip)

    traffic, traffic_src = measured_traffic()
    line = {
        "metric": "bld_mult S-elem updates/s (FP32 batched vs FP64 plain) + p50/p99 solve latency",
        "value": value,
        "unit": "S-element updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (allocation-model samples, seeded)",
        "config": {"workload": "batched FP32 bld_mult: 256 instances, x,y~U{32..512}, "
                               "z~U{4..32}, a~U[10,80], b=1, alpha=beta=1, 200 iterations, "
                               "psi_appr, S0 resident in HBM",
                   "instances": len(c3), "elements": elems3, "iterations": iters3,
                   "pass_items": info["pass_items"], "pass_launches": info["n_buckets"],
                   "tile_autotune": tuned, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "bld_pass_f32z_kernel (FP32 update), all launches of one iteration",
                     "pass_us": pass_ms * 1e3, "iter_us": iter_ms * 1e3,
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }
    ceil = stream_ceiling()
    if ceil:
        line["roofline"]["stream_copy_nt_GBps"] = ceil["copy_nt_GBps"]
        line["roofline"]["frac_of_stream_copy"] = achieved / ceil["copy_nt_GBps"]
        line["stream_ceiling"] = ceil
    line.update(out)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
