#!/usr/bin/env python3
"""Benchmark: seconds to a 1e-8 relative residual for the moving-mesh solve at 512^3 on one MI355X,
mixed precision against FP64, plus the smoother's HBM throughput.

Workload (synthetic, built on the device before timing): lognormal density, sigma = 2.0,
P(k) ~ k^-1.5 exp(-(k/k_s)^2), k_s = k_Nyquist/2, seed 42, periodic N^3 grid.  One reference outer
iteration (calcdefp + def += defp + relaxing) is run untimed to get a curvilinear mesh; the timed
step is the mesh solve of the second outer iteration: L(def, rho) phi = rhs with rhs from the grid
limiter, solved from phi = 0 by repeated V-cycles (the reference's 8-colour Gauss-Seidel and
sweep schedule) until max|rhs - L phi| / max|rhs| <= 1e-8.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512]

Prints one JSON line.  value = mixed-precision seconds per solve (lower is better).
"""
import argparse
import json
import os
import sys
import threading
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "moving-mesh-reconstruction_amd"))

PEAK_HBM_GBS = 8000.0
B_ALG_WORDS = 5  # phi read, rhs, def, rho-source, phi write per point-sweep


def cpu_baseline(n_gpu, cycles, seconds_budget=12.0):
    """Times the reference Fortran (FP32, OpenMP build under oracle/_ref) V-cycle on a 128^3 sample
    of the same field recipe and scales it to seconds-to-tolerance at n_gpu^3."""
    ng = 128
    so = os.path.join(REPO, "oracle", "_ref", "libmmhref_ng%d_f32omp.so" % ng)
    if not os.path.exists(so):
        return None
    out = {}

    def work():
        import ctypes as C
        import numpy as np
        lib = C.CDLL(so)
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import refbind as R
        rho = R.or_lognormal(ng, 2.0, 42).astype(np.float32)
        u = np.zeros((ng ** 3, 5), np.float32)
        u[:, 0] = rho
        u = u.ravel()
        rhs0 = (1 - rho).astype(np.float32)
        for _ in range(4):
            R.or_nsmooth(rhs0, ng)
        d = np.zeros(ng ** 3, np.float32)
        a = np.zeros(ng ** 3, np.float32)
        k, t0 = 0, time.perf_counter()
        while True:
            rhs = rhs0.copy()
            R.ref_multigrid(lib, a, rhs, d, u, 5, 1.0, ng)
            k += 1
            if time.perf_counter() - t0 > seconds_budget:
                break
        dt = (time.perf_counter() - t0) / k
        out["t"] = dt
        out["k"] = k

    threading.stack_size(1 << 30)
    th = threading.Thread(target=work)
    th.start()
    th.join()
    if "t" not in out:
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    scale = (n_gpu / ng) ** 3 * cycles
    return {"value": out["t"] * scale, "unit": "s", "cores": threads, "kind": "reference",
            "sample": "%d V-cycles of the reference multigrid_ (FP32, OpenMP, NG=128 build) on the sigma=2 lognormal "
                      "field at 128^3, %.3f s per cycle, scaled by (%d/128)^3 and by the %d cycles the FP64 GPU solve "
                      "needs" % (out["k"], out["t"], n_gpu, cycles)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--sigma", type=float, default=2.0)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--schedule", type=int, default=0, help="0 reference sweeps, 1 fixed 2/2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweeps", type=int, default=20, help="sweeps timed for the roofline")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it (--gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("OMP_STACKSIZE", "2000M")
    n = args.n

    cpu = None
    import torch
    import mmh_amd
    torch.cuda.init()
    ctx = mmh_amd.Context(0)
    n3 = n ** 3
    f64 = dict(dtype=torch.float64, device="cuda")
    rho = torch.zeros(n3, **f64)
    ctx.lognormal(rho, args.sigma, args.seed)
    u = torch.zeros(n3, 5, **f64)
    u[:, 0] = rho
    u = u.contiguous().view(-1)
    del rho
    defn = torch.zeros(n3, **f64)
    defp = torch.zeros(n3, **f64)
    tmp = torch.zeros(n3, **f64)
    tmp2 = torch.zeros((n + 2) * n * n, **f64)
    ls = ctx.calcdefp(defp, tmp, tmp2, defn, u)
    defn += defp
    ctx.relaxing(u, defn, defp)
    del tmp2
    rhs0 = torch.zeros(n3, **f64)
    ctx.limiter_rhs(rhs0, defn, u)
    torch.cuda.synchronize()
    print("bench: setup done (outer iteration 1: ncompress %d)" % ls.ncompress, file=sys.stderr, flush=True)
    ctx.release_workspace()

    def step(mixed):
        phi = torch.zeros(n3, **f64)
        rhs = rhs0.clone()
        st, ok = ctx.solve(phi, rhs, defn, u, tol=args.tol, max_cycles=200, schedule=args.schedule, nu_coarse=16,
                           mixed=mixed, nu=5)
        if not ok:
            raise RuntimeError("solve did not converge")
        return st, phi

    results = {}
    for mixed in (False, True):
        for w in range(args.warmup):
            st, _ = step(mixed)
            print("bench: warmup %d %s: %d cycles" % (w, "mixed" if mixed else "fp64", st.cycles), file=sys.stderr,
                  flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            st, phi = step(mixed)
            print("bench: step %d %s done (%.3f s elapsed)" % (k, "mixed" if mixed else "fp64",
                                                              time.perf_counter() - t0), file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        results[mixed] = (dt, st.cycles, st.relres[st.cycles - 1], phi)
    t64, c64, r64, phi64 = results[False]
    tmx, cmx, rmx, phimx = results[True]
    err = float((phimx - phi64).abs().max() / phi64.abs().max())
    del phi64, phimx

    # roofline of the dominant kernel (k_colour: one colour class per launch), FP64 and FP32
    G = torch.zeros(6 * n3, **f64)
    ctx.metric(G, defn, u, 1.0, 0, nu=5)
    phi = torch.zeros(n3, **f64)
    ctx.time_sweeps(phi, rhs0, G, 1.0, 2)
    ms64 = ctx.time_sweeps(phi, rhs0, G, 1.0, args.sweeps)
    G32, phi32, rhs32 = G.float(), phi.float(), rhs0.float()
    del G
    ctx.time_sweeps(phi32, rhs32, G32, 1.0, 2)
    ms32 = ctx.time_sweeps(phi32, rhs32, G32, 1.0, args.sweeps)
    launch_ms = ms64 / 8.0
    alg_bytes_launch = B_ALG_WORDS * 8 * n3 / 8.0
    achieved = alg_bytes_launch / (launch_ms * 1e-3) / 1e9

    # The reference prints per V-cycle through the Fortran runtime, which flushes at process exit:
    # from here on fd 1 goes to /dev/null and the result line is written to a saved descriptor.
    result_out = os.fdopen(os.dup(1), "w")
    sys.stdout.flush()
    dn = os.open(os.devnull, os.O_WRONLY)
    os.dup2(dn, 1)
    if not args.no_cpu:
        cpu = cpu_baseline(n, c64)

    line = {
        "metric": "s to 1e-8 rel residual at 512^3 (mixed vs FP64 GPU); smoother HBM GB/s",
        "value": round(tmx, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tmx * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle in f64 defect correction",
        "data": "synthetic lognormal density (sigma=%.1f, seed %d), generated on device" % (args.sigma, args.seed),
        "config": {"workload": "mesh solve of outer iteration 2 to %.0e relative residual, %d^3 periodic, "
                               "reference sweep schedule" % (args.tol, n) if args.schedule == 0 else
                               "mesh solve, fixed 2/2 schedule, %d^3" % n,
                   "n": n, "sigma": args.sigma, "seed": args.seed,
                   "fp64_s": round(t64, 4), "fp64_cycles": c64, "mixed_cycles": cmx,
                   "speedup_mixed_vs_fp64": round(t64 / tmx, 3),
                   "mixed_rel_err_vs_fp64": err, "final_relres_fp64": r64, "final_relres_mixed": rmx,
                   "limiter_ncompress_iter1": ls.ncompress,
                   "sweep_ms_fp64": round(ms64, 4), "sweep_ms_fp32": round(ms32, 4),
                   "smoother_alg_GBs_fp32": round(B_ALG_WORDS * 4 * n3 / (ms32 * 1e-3) / 1e9, 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                     "kernel": "k_colour<double> (one of 8 colour classes of a Gauss-Seidel sweep, %d^3)" % n,
                     "alg_bytes_per_launch": alg_bytes_launch, "avg_launch_ms": round(launch_ms, 5)},
        "cpu_baseline": cpu,
    }
    result_out.write(json.dumps(line) + "\n")
    result_out.flush()


if __name__ == "__main__":
    main()
