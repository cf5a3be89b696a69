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
partials): algorithmic bytes 11 words/cell (reads r, q, p of the previous
step, the correction c and the three face-coefficient arrays; writes r, p,
q and c) divided by its mean launch time, from HIP events on the solver's
stream around windows of 7 consecutive launches (one window per 16 steps)
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
    """Timed solves (kernels replayed from captured HIP graphs), then two
    solves launched directly with HIP events around the fused CG kernel."""
    solver.configure(precision=precision, stop_rule="relative", tol=tol, verbose=False,
                     graphs=graphs)
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
                     timing=True)
    spmv = []
    for _ in range(2):
        df.zero("prd1")
        st = solver.graddin_dev(grid, df)
        if st.spmv_timed:
            spmv.append(st.spmv_ms)
    # the plain SpMV q = A p of this solve's operator (k_spmv), 50 launches
    spmv_only = solver.time_spmv(precision, reps=50)
    return dict(solves_per_s=steps / total, ms_per_step=total * 1e3 / steps,
                p50_ms=percentile(lat, 50), p99_ms=percentile(lat, 99),
                iters=int(np.median(iters)), refreshes=int(np.median(refreshes)),
                rel_true_max=float(max(rel)), graphs=graphs,
                spmv_ms=float(np.mean(spmv)) if spmv else None, spmv_only_ms=spmv_only)


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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this solver runs on one GPU; --gpus {args.gpus} is not supported "
              "(problems shard across GPUs without communication, see DESIGN.md)", file=sys.stderr)
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
        mixed = run_solves(solver, g, df, "mixed", args.steps, args.warmup, args.tol)
        print(f"# mixed {mixed}", file=sys.stderr, flush=True)
    if args.only in ("both", "fp64"):
        fp64 = run_solves(solver, g, df, "fp64", args.steps, args.warmup, args.tol)
        print(f"# fp64 {fp64}", file=sys.stderr, flush=True)
    df.free()
    solver.close()
    if mixed is None or fp64 is None:
        print(json.dumps({"profile_run": args.only, "mixed": mixed, "fp64": fp64, "copy": copy}))
        return

    bytes32 = 11 * 4 * n          # k_cg_m: 11 FP32 words per cell
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
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
