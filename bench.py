#!/usr/bin/env python3
"""Benchmark: seconds to a 1e-8 relative residual for the moving-mesh solve at 512^3 on one MI355X,
mixed precision against FP64, plus the smoother's HBM throughput.

Workload (synthetic, built on the device before timing): lognormal density, sigma = 2.0,
P(k) ~ k^-1.5 exp(-(k/k_s)^2), k_s = k_Nyquist/2, seed 42, periodic N^3 grid.  The timed step is
the mesh equation of the first outer iteration, L(def = 0, rho) phi = rhs, rhs from the grid
limiter (calcdefp's right-hand side), solved from phi = 0 until
max|rhs - L phi| / max|rhs| <= 1e-8.  Solver: flexible preconditioned CG whose preconditioner is one
symmetric V-cycle with the reference's 8-colour curvilinear Gauss-Seidel smoother on the finest level
(1+1 sweeps) and Galerkin coarse operators below (2+2 sweeps, 4+4 on 4^3); FP64, or FP32 with an FP64
coarsest level for the mixed path.  The plain reference V-cycle iteration is timed per cycle beside
it (on this input it does not reach 1e-8 at 512^3 within 120 cycles).

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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(n, rho_host, rhs_host, cycles=2):
    """Times the reference itself, built from its sources at NG = n (oracle/_ref, FP32 as shipped,
    -O3 -fopenmp -mcmodel=large), for `cycles` V-cycles of its multigrid_ on the same equation the GPU
    solves (the density and limiter right-hand side of the workload, rounded to FP32), from phi = 0.
    Returns seconds per reference V-cycle; the GPU's reference V-cycle is timed beside it."""
    import numpy as np
    so = os.path.join(REPO, "oracle", "_ref", "libmmhref_ng%d_f32omp.so" % n)
    if not os.path.exists(so):
        return None
    out = {}

    def work():
        import ctypes as C
        lib = C.CDLL(so)
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import refbind as R
        u = np.zeros((n ** 3, 5), np.float32)
        u[:, 0] = rho_host
        u = u.ravel()
        d = np.zeros(n ** 3, np.float32)
        a = np.zeros(n ** 3, np.float32)
        times = []
        for _ in range(cycles):
            rhs = rhs_host.astype(np.float32)
            t0 = time.perf_counter()
            R.ref_multigrid(lib, a, rhs, d, u, 5, 1.0, n)
            times.append(time.perf_counter() - t0)
        out["t"] = times

    threading.stack_size(1 << 30)  # mg512's level arrays are automatic (stack) arrays
    th = threading.Thread(target=work)
    th.start()
    th.join()
    if "t" not in out:
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t = sum(out["t"]) / len(out["t"])
    return {"value": round(t, 3), "unit": "s per reference V-cycle", "cores": threads, "kind": "reference",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": "%d V-cycles of the reference's multigrid_ built from its sources at NG=%d (FP32, flang -O3 "
                      "-fopenmp -mcmodel=large, OMP_NUM_THREADS=%d) on the workload's density and limiter rhs "
                      "(rounded to FP32), from phi = 0; per-cycle times %s s" % (
                          cycles, n, threads, [round(x, 3) for x in out["t"]])}


def traffic_from_profile(n, kind="f64"):
    """HBM bytes per k_colour_col launch (kind "f64": FP64 values and coefficients; "bf16": FP32 values,
    bfloat16 coefficients) from the committed counter summary, or None."""
    path = os.path.join(REPO, "profiles", "k_colour_col_%s_%d_hbm_traffic.txt" % (kind, n))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1])
    return None


def smoother_only(ctx, torch, defn, u, rhs0, n, sweeps, coef):
    """k_colour_col launches (8 per sweep) at n^3, FP64 (coef "f64") or FP32 with bfloat16 coefficients
    (coef "bf16"), then a plain 1 GiB device-to-device copy whose FETCH_SIZE / WRITE_SIZE give the
    counters' scale against known bytes."""
    n3 = n ** 3
    G = torch.zeros(6 * n3, dtype=torch.float64, device="cuda")
    ctx.metric(G, defn, u, 1.0, 0, nu=1)
    if coef == "bf16":
        phi = torch.zeros(n3, dtype=torch.float32, device="cuda")
        ms = ctx.time_sweeps_bf16(phi, rhs0.float(), G, 1.0, sweeps)
    else:
        phi = torch.zeros(n3, dtype=torch.float64, device="cuda")
        ms = ctx.time_sweeps(phi, rhs0, G, 1.0, sweeps)
    del G, phi
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("smoother-only: %d^3 %s %.4f ms/sweep (%d sweeps); copy 1 GiB x3 done" % (n, coef, ms, sweeps), flush=True)


def krylov_probe(ctx, torch, defn, u, rhs0, n, iters=2):
    """A few PCG iterations in FP64 and in mixed precision (no tolerance), for counter runs over the
    residual/operator kernels (k_op_march), then the same 1 GiB calibration copy."""
    n3 = n ** 3
    for mixed in (0, 1):
        phi = torch.zeros(n3, dtype=torch.float64, device="cuda")
        rhs = rhs0.clone()
        st, _ = ctx.solve(phi, rhs, defn, u, tol=1e-30, max_cycles=iters, schedule=1, nu_pre=2, nu_post=2,
                          nu_coarse=8, mixed=mixed, nu=1, krylov=True)
        del phi, rhs
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("krylov-probe: %d^3, %d PCG iterations per precision; copy 1 GiB x3 done" % (n, iters), flush=True)


def curved_leg(ctx, torch, rho, args, outer=2):
    """The mesh equation at a curved mesh: outer iteration outer + 1 of the same field, after `outer`
    reconstruction iterations (calcdefp, def += defp, relaxing) whose mesh solves ran to the bench
    tolerance (option mesh_solve = 2), so that the metric's cross terms G12, G13, G23 are live.  FP64
    and mixed-precision solves are timed as the first-iteration ones.  calcdefp's step is scaled by
    its dtold argument (the reference's main passes 1, mmh/main.f90:95); where a full step leaves a
    mesh whose discrete operator the solve reports as not definite, the record says so and the next
    step size (2, then 4) is used."""
    import mmh_amd
    n = args.n
    n3 = n ** 3
    f64 = dict(dtype=torch.float64, device="cuda")
    attempts = []
    for dtold in (1.0, 2.0, 4.0):
        u5 = torch.zeros(n3, 5, **f64)
        u5[:, 0] = rho
        u5 = u5.view(-1)
        d, p = torch.zeros(n3, **f64), torch.zeros(n3, **f64)
        ctx.set_option("mesh_solve", 2)
        ctx.set_option("mesh_tol", int(round(-__import__("math").log10(args.tol))))
        ctx.set_option("nu_top", args.nu_top)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        try:
            st = ctx.reconstruct(u5, d, p, outer, dtold=dtold)
        except mmh_amd.MmhError as e:
            attempts.append({"dtold": dtold, "failed": str(e)})
            print("bench: curved mesh, dtold %g: %s" % (dtold, e), file=sys.stderr, flush=True)
            del u5, d, p
            ctx.release_workspace()
            torch.cuda.empty_cache()
            continue
        finally:
            ctx.set_option("mesh_solve", 0)
        t_outer = time.perf_counter() - t0
        rhs0 = torch.zeros(n3, **f64)
        ctx.limiter_rhs(rhs0, d, u5, dtold1=dtold)
        u1 = u5.view(n3, 5)[:, 0].contiguous()
        del u5, p
        torch.cuda.synchronize()
        ctx.release_workspace()
        torch.cuda.empty_cache()
        print("bench: curved mesh after %d outer iterations, dtold %g (%.1f s)" % (outer, dtold, t_outer),
              file=sys.stderr, flush=True)
        res = {}
        for mixed in (False, True):
            for _ in range(args.warmup):
                phi, rhs = torch.zeros(n3, **f64), rhs0.clone()
                ctx.solve(phi, rhs, d, u1, tol=args.tol, max_cycles=200, schedule=1, nu_pre=args.nu,
                          nu_post=args.nu, nu_coarse=args.nu_coarse, mixed=mixed, nu=1, krylov=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                phi, rhs = torch.zeros(n3, **f64), rhs0.clone()
                s2, ok = ctx.solve(phi, rhs, d, u1, tol=args.tol, max_cycles=200, schedule=1, nu_pre=args.nu,
                                   nu_post=args.nu, nu_coarse=args.nu_coarse, mixed=mixed, nu=1, krylov=True)
                if not ok:
                    raise RuntimeError("curved-mesh solve did not converge")
            torch.cuda.synchronize()
            res[mixed] = ((time.perf_counter() - t0) / args.steps, s2.cycles, s2.relres[s2.cycles - 1], phi)
        err = float((res[True][3] - res[False][3]).abs().max() / res[False][3].abs().max())
        out = {"outer_iterations_before": outer, "dtold": dtold, "attempts": attempts,
               "def_sigma": round(st[-1].def_sigma, 4), "ncompress": [s.lim.ncompress for s in st],
               "detmin_detmax_last_limiter": [st[-1].lim.detmin, st[-1].lim.detmax],
               "fp64_s": round(res[False][0], 4), "fp64_cycles": res[False][1], "final_relres_fp64": res[False][2],
               "mixed_s": round(res[True][0], 4), "mixed_cycles": res[True][1], "final_relres_mixed": res[True][2],
               "speedup_mixed_vs_fp64": round(res[False][0] / res[True][0], 3), "mixed_rel_err_vs_fp64": err}
        del res, d, rhs0, u1
        ctx.set_option("nu_top", 0)
        ctx.release_workspace()
        torch.cuda.empty_cache()
        return out
    ctx.set_option("nu_top", 0)
    return {"outer_iterations_before": outer, "attempts": attempts}


def batched_leg(ctx, torch, nu, nu_top, nu_coarse):
    """Batched uneven realisations: 40 x 32^3, 40 x 64^3 and 20 x 128^3 lognormal fields, sigma
    uniform in [0.5, 2.0] (seeded), seeds 0..99, each solved from phi = 0 to 1e-8 (FP64 PCG).
    Realisations per second for one batched call against a host loop of single solves, both with
    the FP64 solve's preconditioner (Galerkin levels, nu_top sweeps on the finest level, nu below)."""
    import numpy as np
    f64 = dict(dtype=torch.float64, device="cuda")
    sizes = [32] * 40 + [64] * 40 + [128] * 20
    sig = np.random.default_rng(2024).uniform(0.5, 2.0, len(sizes))
    probs = []
    for i, (n, s) in enumerate(zip(sizes, sig)):
        rho = torch.zeros(n ** 3, **f64)
        ctx.lognormal(rho, float(s), i)
        u = torch.zeros(n ** 3, 5, **f64)
        u[:, 0] = rho
        u = u.contiguous().view(-1)
        d = torch.zeros(n ** 3, **f64)
        rhs = torch.zeros(n ** 3, **f64)
        ctx.limiter_rhs(rhs, d, u)
        probs.append((u, rhs, d))
    opts = dict(tol=1e-8, max_cycles=200, nu_pre=nu, nu_coarse=nu_coarse)
    ctx.set_option("nu_top", nu_top)

    def batched():
        items = [(torch.zeros_like(r), r.clone(), d, u) for (u, r, d) in probs]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, ok = ctx.solve_batch(items, **opts)
        torch.cuda.synchronize()
        if not ok:
            raise RuntimeError("batched solve did not converge")
        return time.perf_counter() - t0, st

    def looped():
        work = [(torch.zeros_like(r), r.clone(), d, u) for (u, r, d) in probs]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for phi, r, d, u in work:
            st, ok = ctx.solve(phi, r, d, u, tol=1e-8, max_cycles=200, schedule=1, nu_pre=nu, nu_post=nu,
                               nu_coarse=nu_coarse, nu=5, krylov=True)
            if not ok:
                raise RuntimeError("single solve did not converge")
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    batched()  # warm-up: workspaces and launch-shape tuning
    tb, st = batched()
    print("bench: batched 100 realisations %.3f s" % tb, file=sys.stderr, flush=True)
    # the preconditioner's coarse levels one workgroup per grid up to 4^3 and 8^3, and all by launches
    # (the default: up to 16^3)
    tfused = {}
    for fz in (0, 4, 8):
        ctx.set_option("batch_coarse_fused", fz)
        try:
            batched()
            tfused[fz] = batched()[0]
        finally:
            ctx.set_option("batch_coarse_fused", 16)
        print("bench: batched 100 realisations, coarse levels <= %d^3 one workgroup per grid: %.3f s" % (fz, tfused[fz]),
              file=sys.stderr, flush=True)
    # the small groups solved whole in one workgroup per grid (option batch_persistent): up to 32^3,
    # and up to 64^3
    tpers = {}
    for pers in (32, 64):
        ctx.set_option("batch_persistent", pers)
        try:
            batched()
            tpers[pers] = batched()
        finally:
            ctx.set_option("batch_persistent", 0)
        print("bench: batched 100 realisations, one workgroup per grid up to %d^3: %.3f s" % (pers, tpers[pers][0]),
              file=sys.stderr, flush=True)
    looped()
    tl = looped()
    print("bench: looped 100 realisations %.3f s" % tl, file=sys.stderr, flush=True)
    its = [s.cycles for s in st]
    return {"workload": "40x32^3 + 40x64^3 + 20x128^3 lognormal, sigma U[0.5,2], seeds 0..99, to 1e-8 (FP64 PCG, "
                        "Galerkin levels, %d+%d sweeps on the finest level, %d+%d below)" % (nu_top, nu_top, nu, nu),
            "batched_realizations_per_s": round(len(sizes) / tb, 2), "looped_realizations_per_s": round(len(sizes) / tl, 2),
            "batched_coarse_levels_fused_realizations_per_s": {str(k): round(len(sizes) / v, 2)
                                                               for k, v in sorted(tfused.items())},
            "batched_one_workgroup_per_grid_upto_32_realizations_per_s": round(len(sizes) / tpers[32][0], 2),
            "batched_one_workgroup_per_grid_upto_64_realizations_per_s": round(len(sizes) / tpers[64][0], 2),
            "pcg_iterations_min_max_one_workgroup_per_grid": [min(s.cycles for s in tpers[64][1]),
                                                              max(s.cycles for s in tpers[64][1])],
            "batched_speedup": round(tl / tb, 3), "pcg_iterations_min_max": [min(its), max(its)]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--sigma", type=float, default=2.0)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--nu", type=int, default=3, help="pre/post sweeps of the preconditioning V-cycle below the top")
    ap.add_argument("--nu-fp64", type=int, default=2,
                    help="the same for the FP64 solve timed at its own fastest schedule (0: skip)")
    ap.add_argument("--nu-top", type=int, default=1, help="pre/post sweeps on the finest level (0: same as --nu)")
    ap.add_argument("--nu-coarse", type=int, default=4, help="sweeps each way on the 4^3 level")
    ap.add_argument("--no-galerkin", action="store_true", help="coarse levels rebuilt from restricted fields")
    ap.add_argument("--ref-cycles", type=int, default=3, help="plain reference V-cycles timed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweeps", type=int, default=20, help="sweeps timed for the roofline")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-realisations leg")
    ap.add_argument("--no-curved", action="store_true", help="skip the curved-mesh leg")
    ap.add_argument("--krylov-probe", action="store_true",
                    help="only two PCG iterations per precision plus the 1 GiB copy (counter runs)")
    ap.add_argument("--smoother-only", action="store_true",
                    help="only the smoother timing plus a 1 GiB device copy (the counter-calibration run)")
    ap.add_argument("--solves-only", action="store_true",
                    help="only the warm-up and timed FP64 and mixed solves (kernel-trace runs)")
    ap.add_argument("--coef", choices=("f64", "bf16"), default="f64",
                    help="--smoother-only: FP64 sweeps, or FP32 sweeps with bfloat16 coefficients")
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
    defn = torch.zeros(n3, **f64)
    rhs0 = torch.zeros(n3, **f64)
    u5 = torch.zeros(n3, 5, **f64)
    u5[:, 0] = rho
    ctx.limiter_rhs(rhs0, defn, u5.contiguous().view(-1))
    del u5
    # the solves read only the density component of the fluid array: pass it alone (nu = 1)
    u = rho
    torch.cuda.synchronize()
    ctx.release_workspace()  # the generator's and limiter's buffers (tens of GiB at 1024^3)
    torch.cuda.empty_cache()
    print("bench: setup done", file=sys.stderr, flush=True)
    if args.smoother_only:
        smoother_only(ctx, torch, defn, u, rhs0, n, args.sweeps, args.coef)
        return
    if args.krylov_probe:
        krylov_probe(ctx, torch, defn, u, rhs0, n)
        return

    # preconditioner: Galerkin coarse operators, nu_top sweeps each way on the finest level
    ctx.set_option("galerkin", 0 if args.no_galerkin else 1)

    def step(mixed):
        ctx.set_option("nu_top", args.nu_top)
        phi = torch.zeros(n3, **f64)
        rhs = rhs0.clone()
        st, ok = ctx.solve(phi, rhs, defn, u, tol=args.tol, max_cycles=200, schedule=1, nu_pre=args.nu,
                           nu_post=args.nu, nu_coarse=args.nu_coarse, mixed=mixed, nu=1, krylov=True)
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
    if args.solves_only:
        print("solves-only: %d^3 fp64 %.4f s (%d iterations), mixed %.4f s (%d iterations), error %.2e" % (
            n, t64, c64, tmx, cmx, err), flush=True)
        return
    # The FP64 solve at its own fastest schedule (2+2 below the finest level; the mixed solve's is 3+3,
    # which costs the FP64 cycle more than its saved iterations give back), for a best-against-best ratio
    fp64_best = None
    if args.nu_fp64 and args.nu_fp64 != args.nu:
        def step64():
            phi = torch.zeros(n3, **f64)
            rhs = rhs0.clone()
            st, ok = ctx.solve(phi, rhs, defn, u, tol=args.tol, max_cycles=200, schedule=1, nu_pre=args.nu_fp64,
                               nu_post=args.nu_fp64, nu_coarse=args.nu_coarse, mixed=False, nu=1, krylov=True)
            if not ok:
                raise RuntimeError("solve did not converge")
            return st
        step64()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            st = step64()
        torch.cuda.synchronize()
        tb = (time.perf_counter() - t0) / args.steps
        fp64_best = {"nu": args.nu_fp64, "s": round(tb, 4), "cycles": st.cycles,
                     "speedup_mixed_vs_fp64_best": round(tb / tmx, 3)}
        print("bench: fp64 at %d+%d below the finest level: %.4f s (%d cycles)" % (args.nu_fp64, args.nu_fp64, tb,
                                                                                 st.cycles), file=sys.stderr, flush=True)

    # the plain reference algorithm: FP64 V-cycles with the reference sweep schedule
    ref_ms, ref_relres, ref32_ms = None, None, None
    if args.ref_cycles > 0:
        # one untimed cycle first (first launches of the schedule's kernels, its workspaces)
        phi = torch.zeros(n3, **f64)
        rhs = rhs0.clone()
        ctx.solve(phi, rhs, defn, u, tol=1e-30, max_cycles=1, schedule=0, nu=1)
        phi = torch.zeros(n3, **f64)
        rhs = rhs0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, _ = ctx.solve(phi, rhs, defn, u, tol=1e-30, max_cycles=args.ref_cycles, schedule=0, nu=1)
        torch.cuda.synchronize()
        ref_ms = (time.perf_counter() - t0) * 1e3 / st.cycles
        ref_relres = st.relres[st.cycles - 1]
        del phi, rhs
        print("bench: reference V-cycle %.1f ms" % ref_ms, file=sys.stderr, flush=True)
        # the same cycle as the Fortran drop-in runs it, beside the CPU leg: FP32, through multigrid's
        # own entry (rgzerosum of a fresh rhs, one V-cycle continuing from phi), 5-component u; the
        # FP64 solves' workspace goes first (at 1024^3 both would not fit)
        ctx.release_workspace()
        torch.cuda.empty_cache()
        try:
            u32 = torch.zeros(n3, 5, dtype=torch.float32, device="cuda")
            u32[:, 0] = rho.float()
            u32 = u32.view(-1)
            d32 = defn.float()
            a32 = torch.zeros(n3, dtype=torch.float32, device="cuda")
            ctx.multigrid(a32, rhs0.float(), d32, u32, dx=1.0, iopt=2, nu=5)  # untimed first call
            a32.zero_()
            ts = []
            for _ in range(2):
                r32 = rhs0.float()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                ctx.multigrid(a32, r32, d32, u32, dx=1.0, iopt=2, nu=5)
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            ref32_ms = sum(ts) / len(ts) * 1e3
            del u32, d32, a32, r32
            print("bench: reference V-cycle, FP32 drop-in entry %.1f ms" % ref32_ms, file=sys.stderr, flush=True)
        except (mmh_amd.MmhError, torch.cuda.OutOfMemoryError) as e:
            print("bench: reference V-cycle, FP32 drop-in entry not timed: %s" % e, file=sys.stderr, flush=True)
        ctx.release_workspace()
        torch.cuda.empty_cache()

    curved = None
    if not args.no_curved:
        curved = curved_leg(ctx, torch, rho, args)

    batch = None
    if not args.no_batch:
        # FP64 throughout, so at the FP64 solve's fastest schedule
        batch = batched_leg(ctx, torch, args.nu_fp64 or args.nu, args.nu_top, args.nu_coarse)

    ctx.release_workspace()
    torch.cuda.empty_cache()
    # host copies for the CPU leg (the reference's precision is FP32)
    rho_host = rho.cpu().numpy().astype("float32") if not args.no_cpu else None
    rhs_host = rhs0.cpu().numpy().astype("float32") if not args.no_cpu else None
    # roofline of the dominant kernel of the timed (mixed) solve: the finest level's smoother,
    # k_colour_col with FP32 values and bfloat16 coefficients, one colour class per launch; the
    # FP64 smoother of the FP64 solve beside it.  Sweeps timed with HIP events on the library's
    # stream.  Algorithmic bytes: 5 words per point update (phi read and written, rhs, def, rho),
    # the metric counted as recomputed from def and rho; the kernel reads a stored metric instead,
    # which the counter traffic (committed rocprofv3 --pmc summary) shows.
    G = torch.zeros(6 * n3, **f64)
    ctx.metric(G, defn, u, 1.0, 0, nu=1)
    phi = torch.zeros(n3, **f64)
    ctx.time_sweeps(phi, rhs0, G, 1.0, 2)
    ms64 = ctx.time_sweeps(phi, rhs0, G, 1.0, args.sweeps)
    phi32, rhs32 = phi.float(), rhs0.float()
    ctx.time_sweeps_bf16(phi32, rhs32, G, 1.0, 2)
    msb = ctx.time_sweeps_bf16(phi32, rhs32, G, 1.0, args.sweeps)
    G32 = G.float()
    del G
    ctx.time_sweeps(phi32, rhs32, G32, 1.0, 2)
    ms32 = ctx.time_sweeps(phi32, rhs32, G32, 1.0, args.sweeps)
    del G32, phi32, rhs32, phi

    def roof(ms_sweep, word, kind, kernel, launches=8):
        launch_ms = ms_sweep / launches
        alg = B_ALG_WORDS * word * n3 / launches
        ach = alg / (launch_ms * 1e-3) / 1e9
        tr = traffic_from_profile(n, kind)
        out = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
               "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": tr, "kernel": kernel,
               "alg_bytes_per_launch": alg, "avg_launch_ms": round(launch_ms, 5)}
        if tr:
            out["traffic_over_alg"] = round(tr / alg, 2)
            out["traffic_GBs"] = round(tr / (launch_ms * 1e-3) / 1e9, 1)
            out["traffic_frac"] = round(out["traffic_GBs"] / PEAK_HBM_GBS, 4)
        return out
    # the mixed solve's finest-level sweeps run slab by slab (option sweep_slabs, levels >= 512^3)
    slabs = ctx.get_option("sweep_slabs")
    slabs = slabs if slabs > 1 and n >= 512 and (n // 16) % slabs == 0 else 1
    roof_b = roof(msb, 4, "bf16", "k_colour_col<float,bf16> (finest-level smoother of the mixed solve: FP32 "
                                  "values, bfloat16 metric; one of 8 colour classes on one of %d k-slabs, %d^3, 8 "
                                  "points per thread along k; %d launches per sweep)" % (slabs, n, 8 * slabs),
                  launches=8 * slabs)
    roof_64 = roof(ms64, 8, "f64", "k_colour_col<double> (the FP64 solve's smoother, %d^3)" % n)

    # The reference prints per V-cycle through the Fortran runtime, which flushes at process exit:
    # from here on fd 1 goes to /dev/null and the result line is written to a saved descriptor.
    result_out = os.fdopen(os.dup(1), "w")
    sys.stdout.flush()
    dn = os.open(os.devnull, os.O_WRONLY)
    os.dup2(dn, 1)
    if not args.no_cpu:
        cpu = cpu_baseline(n, rho_host, rhs_host)
        if cpu and ref_ms:
            cpu["gpu_reference_vcycle_s"] = round(ref_ms / 1e3, 4)
            cpu["gpu_vs_cpu_reference_vcycle"] = round(cpu["value"] / (ref_ms / 1e3), 1)
        if cpu and ref32_ms:
            cpu["gpu_reference_vcycle_fp32_dropin_s"] = round(ref32_ms / 1e3, 4)
            cpu["gpu_fp32_dropin_vs_cpu_reference_vcycle"] = round(cpu["value"] / (ref32_ms / 1e3), 1)

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
        "dtype": "f32 multigrid preconditioner inside f64 PCG",
        "data": "synthetic lognormal density (sigma=%.1f, seed %d), generated on device" % (args.sigma, args.seed),
        "config": {"workload": "first-iteration mesh solve to %.0e relative residual, %d^3 periodic, "
                               "PCG with a symmetric V-cycle preconditioner (%s coarse operators; %d+%d sweeps on "
                               "the finest level, %d+%d below, %d+%d on 4^3)" % (
                                   args.tol, n, "rebuilt" if args.no_galerkin else "Galerkin",
                                   args.nu_top or args.nu, args.nu_top or args.nu, args.nu, args.nu,
                                   args.nu_coarse, args.nu_coarse),
                   "n": n, "sigma": args.sigma, "seed": args.seed,
                   "fp64_s": round(t64, 4), "fp64_cycles": c64, "mixed_cycles": cmx,
                   "speedup_mixed_vs_fp64": round(t64 / tmx, 3), "fp64_own_best_schedule": fp64_best,
                   "mixed_rel_err_vs_fp64": err, "final_relres_fp64": r64, "final_relres_mixed": rmx,
                   "reference_vcycle_ms_fp64": ref_ms, "reference_relres_after_%d_cycles" % args.ref_cycles: ref_relres,
                   "reference_vcycle_ms_fp32_dropin_entry": ref32_ms,
                   "sweep_ms_fp64": round(ms64, 4), "sweep_ms_fp32": round(ms32, 4),
                   "sweep_ms_fp32_bf16": round(msb, 4),
                   "smoother_alg_GBs_fp32": round(B_ALG_WORDS * 4 * n3 / (ms32 * 1e-3) / 1e9, 1),
                   "roofline_fp64_smoother": roof_64,
                   "curved_mesh": curved,
                   "batched_realizations": batch},
        "roofline": roof_b,
        "cpu_baseline": cpu,
    }
    result_out.write(json.dumps(line) + "\n")
    result_out.flush()


if __name__ == "__main__":
    main()
