"""Pressure-solve benchmark on one MI355X.

Workload: the two-phase sloshing tank, 256x128x96 (dx = 5e-3 m, rho
1000/1.2 blended over 1.5 dx, u,v,w ~ U(-0.05, 0.05), seed 1), solved from
prd1 = 0 to a relative residual ||b - A e|| / ||b|| <= 1e-8 on the Jacobi-
scaled system.  One step = one pressure solve (SuLi's graddin) with every
input already resident in HBM.  The mixed-precision solver (FP32 SpMV/axpy,
FP64 dots, scalars and true residual) is the headline; the FP64 solver runs
the same steps for the comparison.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-solves C]

Prints one JSON line.  `roofline` is for the dominant kernel, k_cg_m (the
whole FP32 CG step of the mixed solver in one launch: residual and iterate
updates of the previous iteration, direction update, SpMV, four dot-product
partials): algorithmic bytes 9 words/cell on average (every step reads
q and the last two p of the previous steps, from which it recovers r, and
the three face-coefficient arrays, and writes p and q; every fourth step
also reads the correction c and the two older p of the direction ring and
writes c) divided by its mean launch time, from HIP
events on the solver's stream around windows of 8 consecutive launches (one
window per 16 steps)
during two directly launched solves.  The FP64 solver's step (k_cg_a +
k_cg_b, 12 FP64 words/cell) is reported beside it.  `traffic` is the HBM
bytes per launch from rocprofv3's FETCH_SIZE and WRITE_SIZE counters
(profiles/hbm_traffic.json), calibrated on the copy kernel k_copy.
`cpu_baseline` runs the reference's own Fortran
graddin (oracle/_ref, compiled from the SuLi sources) with OpenMP on the
host cores, on one solve of the same case.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "pressure solves/s + p99 latency at 1e-8 residual, mixed vs FP64 CG"


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def run_solves(solver, grid, df, precision, steps, warmup, tol, graphs=True):
    """Timed solves (kernels replayed from captured HIP graphs, tile shape
    from the autotuner, which runs during the first warm-up solve), then two
    solves launched directly with HIP events around the CG step."""
    solver.configure(precision=precision, stop_rule="relative", tol=tol, verbose=False,
                     graphs=graphs, autotune=True)
    for _ in range(warmup):
        df.zero("prd1")
        solver.graddin_dev(grid, df)
    solver.synchronize()
    lat, iters, rel, refreshes = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        df.zero("prd1")
        st = solver.graddin_dev(grid, df)
        lat.append((time.perf_counter() - ts) * 1e3)
        iters.append(st.iters)
        rel.append(st.rel_true)
        refreshes.append(st.refreshes)
        if st.status != 0:
            raise RuntimeError(f"{precision} solve did not converge: {st.as_dict()}")
    solver.synchronize()
    total = time.perf_counter() - t0
    solver.configure(precision=precision, stop_rule="relative", tol=tol, verbose=False,
                     timing=True, autotune=True)
    spmv = []
    for _ in range(2):
        df.zero("prd1")
        st = solver.graddin_dev(grid, df)
        if st.spmv_timed:
            spmv.append(st.spmv_ms)
    # the plain SpMV q = A p of this solve's operator (k_spmv), 50 launches
    spmv_only = solver.time_spmv(precision, reps=50)
    kz, vx, _ = solver.autotune(grid, precision)
    return dict(solves_per_s=steps / total, ms_per_step=total * 1e3 / steps,
                p50_ms=percentile(lat, 50), p99_ms=percentile(lat, 99),
                iters=int(np.median(iters)), refreshes=int(np.median(refreshes)),
                rel_true_max=float(max(rel)), graphs=graphs, tiles={"kz": kz, "vx": vx},
                spmv_ms=float(np.mean(spmv)) if spmv else None, spmv_only_ms=spmv_only)


def rel_diff(a, b):
    """max |a - b| / max |b| over the interior and face ghosts (edge and
    corner ghosts of the halo box excluded)."""
    m = np.ones(a.shape, dtype=bool)
    for ax in range(3):
        for other in range(3):
            if other == ax:
                continue
            for e in (0, -1):
                for e2 in (0, -1):
                    idx = [slice(None)] * 3
                    idx[ax] = e
                    idx[other] = e2
                    m[tuple(idx)] = False
    return float(np.abs(a - b)[m].max() / np.abs(b[m]).max())


def bench_stretched_cylinder(steps):
    """384x192x128 stretched grid with the masked cylinder (not a reference
    mode): solves/s of both solvers, and the mixed result against the FP64
    one (the FP64 solver is pinned to the CPU oracle at small sizes)."""
    from suli_amd import cases
    from suli_amd.solver import PressureSolver
    c = cases.config3(seed=1)
    g = c.grid
    solver = PressureSolver(verbose=False)
    df = solver.upload(g, c.fields)
    out, prd = {}, {}
    for precision in ("mixed", "fp64"):
        solver.configure(precision=precision, stop_rule="relative", tol=1e-8, verbose=False,
                         autotune=True)
        df.zero("prd1")
        solver.graddin_dev(g, df)
        solver.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            df.zero("prd1")
            st = solver.graddin_dev(g, df)
        solver.synchronize()
        dt = (time.perf_counter() - t0) / steps
        if st.status != 0:
            raise RuntimeError(f"cylinder {precision} did not converge: {st.as_dict()}")
        prd[precision] = df.download("prd1")
        kz, vx, _ = solver.autotune(g, precision)
        out[precision] = {"solves_per_s": 1.0 / dt, "ms_per_solve": dt * 1e3, "iters": st.iters,
                          "rel_true": st.rel_true, "tiles": {"kz": kz, "vx": vx}}
    df.free()
    solver.close()
    out["mixed_vs_fp64_rel_pressure_diff"] = rel_diff(prd["mixed"], prd["fp64"])
    out["speedup_mixed_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["workload"] = (f"stretched dz (tanh, ratio 1.05, finest at 0.6H) {g.nx}x{g.ny}x{g.nz}, masked "
                       "full-height cylinder r = 0.1 Ly, rho 1000/1.2, 1e-8 relative residual")
    return out


def bench_constant_density(steps):
    """The two-phase tank's grid and velocities solved as IPA (t_press = 2:
    every face uses rho_f1 = 1000), whose operator the CG kernels read from
    per-plane tables: per-step time and solves/s of both solvers, and the
    plain SpMV at 2 words per cell."""
    from suli_amd import cases
    from suli_amd.solver import PressureSolver
    c = cases.config2(seed=1)
    g = c.grid
    g.t_press = 2
    g.rho_f1 = 1000.0
    solver = PressureSolver(verbose=False)
    df = solver.upload(g, c.fields)
    out = {}
    n = g.ncell
    for precision in ("mixed", "fp64"):
        solver.configure(precision=precision, stop_rule="relative", tol=1e-8, verbose=False,
                         autotune=True)
        df.zero("prd1")
        solver.graddin_dev(g, df)
        solver.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            df.zero("prd1")
            st = solver.graddin_dev(g, df)
        solver.synchronize()
        dt = (time.perf_counter() - t0) / steps
        if st.status != 0:
            raise RuntimeError(f"IPA {precision} did not converge: {st.as_dict()}")
        # one more solve launched directly, with the CG step timed
        solver.configure(precision=precision, stop_rule="relative", tol=1e-8, verbose=False,
                         autotune=True, timing=True)
        df.zero("prd1")
        st = solver.graddin_dev(g, df)
        w = 4 if precision == "mixed" else 8
        spmv_ms = solver.time_spmv(precision, reps=50)
        words = 6.0 if precision == "mixed" else 9   # mixed: 5 words, and 9 every fourth step
        out[precision] = {"solves_per_s": 1.0 / dt, "ms_per_solve": dt * 1e3, "iters": st.iters,
                          "rel_true": st.rel_true, "step_ms": st.spmv_ms, "words_per_cell_step": words,
                          "step_achieved_gbs": words * w * n / (st.spmv_ms * 1e-3) / 1e9,
                          "spmv_ms": spmv_ms, "spmv_achieved_gbs": 2 * w * n / (spmv_ms * 1e-3) / 1e9}
    df.free()
    solver.close()
    out["speedup_mixed_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["workload"] = (f"IPA (t_press=2, rho_f1=1000) on the tank grid {g.nx}x{g.ny}x{g.nz}, "
                       "1e-8 relative residual; operator read from per-plane tables")
    return out


def bench_batch(seed=1, count=64):
    """64 uneven problems (sizes log-uniform in [32, 128] per axis) solved
    in one batch, against solving them one after another."""
    from suli_amd import cases
    from suli_amd.solver import PressureSolver
    cs = cases.batch_cases(seed=seed, count=count)
    cells = sum(c.grid.ncell for c in cs)
    out = {"problems": count, "cells": cells,
           "workload": f"{count} two-phase problems, nx,ny,nz log-uniform in [32,128] (x8), "
                       "flat surface at U(0.3,0.7)H, density ratio in {1,10,1000}, 1e-8 relative"}
    solver = PressureSolver(verbose=False)
    dfs = [solver.upload(c.grid, c.fields) for c in cs]
    items = [(c.grid, df) for c, df in zip(cs, dfs)]
    for precision in ("mixed", "fp64"):
        solver.configure(precision=precision, stop_rule="relative", tol=1e-8, verbose=False,
                         autotune=True)
        for df in dfs:
            df.zero("prd1")
        solver.graddin_batch_dev(items)      # warm-up: tunes each size
        for df in dfs:
            df.zero("prd1")
        solver.synchronize()
        t0 = time.perf_counter()
        sts = solver.graddin_batch_dev(items)
        t_batch = time.perf_counter() - t0
        t0 = time.perf_counter()
        for (g, df) in items:
            df.zero("prd1")
            solver.graddin_dev(g, df)
        t_seq = time.perf_counter() - t0
        out[precision] = {"batch_s": t_batch, "solves_per_s": count / t_batch,
                          "mcells_per_s": cells / t_batch / 1e6,
                          "one_by_one_solves_per_s": count / t_seq,
                          "batching_speedup": t_seq / t_batch,
                          "max_iters": max(s.iters for s in sts),
                          "not_converged": sum(1 for s in sts if s.status != 0)}
        print(f"# batch {precision} {out[precision]}", file=sys.stderr, flush=True)
    for df in dfs:
        df.free()
    solver.close()
    out["speedup_mixed_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    return out


def bench_serving(seed=1, count=48, rates=(60.0, 240.0)):
    """Requests of uneven sizes arriving as a Poisson process, served with
    continuous batching; latency from arrival to the result on the host."""
    from suli_amd import cases
    from suli_amd.server import PressureServer, poisson_stream
    base = cases.serving_cases(seed=seed, count=count)
    out = {"requests": count, "rates_per_s": list(rates),
           "workload": f"{count} requests, nx log-uniform [32,256], ny,nz [32,128] (x8), "
                       "two-phase tank fields, Poisson arrivals, 1e-8 relative"}
    for precision in ("mixed", "fp64"):
        out[precision] = []
        for rate in rates:
            cs = [cases.Case(c.name, c.grid, c.fields.copy(), c.seed) for c in base]
            with PressureServer(precision=precision, stop_rule="relative", tol=1e-8) as srv:
                rep = poisson_stream(srv, cs, rate=rate, seed=seed)
            rep.pop("stats")
            info = rep.pop("info")
            rep["peak_resident_mb"] = info["peak_resident_bytes"] / 1e6
            rep["chunks"] = info["chunks"]
            out[precision].append(rep)
            print(f"# serving {precision} {rep}", file=sys.stderr, flush=True)
    return out


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    return d.get(name)


def cpu_baseline(case, tol, solves):
    """The reference's Fortran graddin (or, if it was not built, the C
    restatement) with OpenMP on the host cores.  Its stop rule is absolute,
    r.r <= 1e-4/N (SuLi-IC/9_grad.f90:393); inputs are scaled by lam so that
    it equals the relative 1e-8 rule (the solution scales with lam)."""
    import oracle
    from suli_amd import cases
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    _, st, _ = oracle.graddin(case.grid, case.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-2,
                              threads=threads)
    lam = float(np.sqrt((1e-4 / case.grid.ncell) / (tol * tol * st.bb)))
    f = cases.scale_velocity(case.fields, lam)
    kind = "reference" if oracle.ref_available("ic") else "port"
    times = []
    for _ in range(solves):
        t = time.perf_counter()
        if kind == "reference":
            oracle.ref_graddin(case.grid, f, omp_t=1)
        else:
            oracle.graddin(case.grid, f, omp_t=1, threads=threads)
        times.append(time.perf_counter() - t)
    t = float(np.median(times))
    what = ("SuLi graddin (amdflang -O3 -fopenmp, omp_t=1)" if kind == "reference"
            else "C restatement of SuLi graddin (gcc -O2 -fopenmp)")
    return {"value": 1.0 / t, "unit": "solves/s", "cores": threads, "kind": kind,
            "sample": f"{solves} FP64 solve(s) of the same 256x128x96 case to 1e-8 relative "
                      f"(absolute rule with lam={lam:.3e}); {what}; {t:.2f} s per solve"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-solves", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--nz", type=int, default=96)
    ap.add_argument("--only", choices=("both", "mixed", "fp64"), default="both",
                    help="solver(s) to time (profiling runs)")
    ap.add_argument("--copy-mib", type=int, default=256,
                    help="size of the plain copy kernel that calibrates bandwidth and counters")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel directly instead of replaying HIP graphs")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the stretched-cylinder, batch and serving workloads")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this solver runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent problems would shard across GPUs, one process per GPU, with no "
              "communication; this benchmark measures one)", file=sys.stderr)
        sys.exit(2)

    from suli_amd import cases
    from suli_amd.solver import PressureSolver

    case = cases.config2(seed=1, nx=args.nx, ny=args.ny, nz=args.nz)
    g = case.grid
    n = g.ncell
    solver = PressureSolver(verbose=False)
    df = solver.upload(g, case.fields)
    print(f"# case {case.name} n={n}", file=sys.stderr, flush=True)
    copy_bytes = args.copy_mib << 20
    copy_ms = solver.copy_bench(copy_bytes, reps=10)
    copy = {"kernel": "k_copy", "bytes_read": copy_bytes, "bytes_written": copy_bytes,
            "ms": copy_ms, "gbs": 2 * copy_bytes / (copy_ms * 1e-3) / 1e9}
    print(f"# copy {copy}", file=sys.stderr, flush=True)
    mixed = fp64 = None
    if args.only in ("both", "mixed"):
        mixed = run_solves(solver, g, df, "mixed", args.steps, args.warmup, args.tol,
                           graphs=not args.no_graphs)
        print(f"# mixed {mixed}", file=sys.stderr, flush=True)
    if args.only in ("both", "fp64"):
        fp64 = run_solves(solver, g, df, "fp64", args.steps, args.warmup, args.tol,
                          graphs=not args.no_graphs)
        print(f"# fp64 {fp64}", file=sys.stderr, flush=True)
    df.free()
    solver.close()
    if mixed is None or fp64 is None:
        print(json.dumps({"profile_run": args.only, "mixed": mixed, "fp64": fp64, "copy": copy}))
        return

    bytes32 = 36 * n              # k_cg_m: 9 FP32 words per cell on average (8, and 12 every 4th step)
    bytes64 = 12 * 8 * n          # FP64 step (k_cg_a + k_cg_b): 12 FP64 words per cell
    spmv_ms = mixed["spmv_ms"]
    achieved = bytes32 / (spmv_ms * 1e-3) / 1e9 if spmv_ms else None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None,
                "traffic": load_traffic(f"k_cg_m {g.nx}x{g.ny}x{g.nz}"),
                "kernel": "k_cg_m", "bytes_per_launch": bytes32,
                "launch_ms": spmv_ms,
                "spmv": {"kernel": "k_spmv<float>", "words_per_cell": 5,
                         "launch_ms": mixed["spmv_only_ms"],
                         "achieved": 5 * 4 * n / (mixed["spmv_only_ms"] * 1e-3) / 1e9,
                         "frac": 5 * 4 * n / (mixed["spmv_only_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "fp64_launch_ms": fp64["spmv_only_ms"],
                         "fp64_achieved": 5 * 8 * n / (fp64["spmv_only_ms"] * 1e-3) / 1e9},
                "fp64_step": {"kernels": "k_cg_a<double> + k_cg_b<double>",
                              "step_ms": fp64["spmv_ms"], "bytes_per_step": bytes64,
                              "achieved": (bytes64 / (fp64["spmv_ms"] * 1e-3) / 1e9)
                              if fp64["spmv_ms"] else None}}

    extra = None
    if not args.no_extra:
        extra = {}
        print("# stretched cylinder", file=sys.stderr, flush=True)
        extra["stretched_cylinder"] = bench_stretched_cylinder(2)
        print(f"# {extra['stretched_cylinder']}", file=sys.stderr, flush=True)
        extra["constant_density"] = bench_constant_density(3)
        print(f"# {extra['constant_density']}", file=sys.stderr, flush=True)
        print("# batch", file=sys.stderr, flush=True)
        extra["batch_uneven"] = bench_batch()
        print("# serving", file=sys.stderr, flush=True)
        extra["serving"] = bench_serving()

    cpu = None
    if not args.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(case, args.tol, args.cpu_solves)

    out = {
        "metric": METRIC,
        "value": mixed["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-inner/fp64-refined",
        "data": "synthetic (seeded two-phase sloshing-tank fields, seed 1)",
        "config": {"workload": f"two-phase sloshing tank {g.nx}x{g.ny}x{g.nz}, rho 1000/1.2, "
                               f"1e-8 relative residual from x0=0",
                   "grid": [g.nx, g.ny, g.nz], "tol": args.tol, "parallelism": "single-gpu"},
        "mixed": mixed,
        "fp64": fp64,
        "speedup_mixed_vs_fp64": mixed["solves_per_s"] / fp64["solves_per_s"],
        "latency_ms": {"p50": mixed["p50_ms"], "p99": mixed["p99_ms"]},
        "roofline": roofline,
        "copy_kernel": copy,
        "workloads": extra,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
