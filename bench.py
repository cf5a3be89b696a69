"""Pressure-solve benchmark on one MI355X.

Workload: the two-phase sloshing tank, 256x128x96 (dx = 5e-3 m, rho
1000/1.2 blended over 1.5 dx, u,v,w ~ U(-0.05, 0.05), seed 1), solved from
prd1 = 0 to a relative residual ||b - A e|| / ||b|| <= 1e-8 on the Jacobi-
scaled system.  One step = one pressure solve (SuLi's graddin) with every
input already resident in HBM.  The mixed-precision solver (FP32 SpMV/axpy,
FP64 dots, scalars and true residual) is the headline; the FP64 solver runs
the same steps for the comparison.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-solves C] [--no-cpu]
                    [--no-extra] [--extra NAMES] [--no-graphs] [--only mixed|fp64|both|none]
                    [--tiles PRECISION=KZxVX]

Prints one JSON line.  Beside the headline it reports `workloads` (--extra,
all by default): the stretched masked cylinder, the tank as IPA, the
64-problem batch, the serving stream, the host-pointer drop-in call per RK
sub-stage, and `rows`, the device-resident entry points of the rows beside
the projection step (density from the level set, the hydrostatic step,
exp_press, SuLi-IT's solve and posdin), each with its time per call and the
bytes it moves.  With --only both it also times the device-resident posdin
(`posdin_dev`).

`roofline` is for the dominant kernel, k_cg_m, the whole FP32 CG iteration of
the mixed solver in one launch.  `achieved` counts the algorithmic bytes of
one CG iteration, 14 FP32 words per cell (SpMV 5: p and the three face
coefficients read, Ap written; iterate, residual and direction 3 each), i.e.
56 B/cell, divided by the kernel's mean launch time, which HIP events on the
solver's stream measure around windows of 8 consecutive launches (one window
per 16 steps) during two directly launched solves; a window's time includes
the gaps between its launches.  The fused kernel moves fewer bytes than
that: `traffic` is its HBM bytes per launch from rocprofv3's FETCH_SIZE and
WRITE_SIZE counters (profiles/hbm_traffic.json, calibrated on the copy
kernel k_copy), and `traffic_achieved` those bytes at the same launch rate.
`copy_kernel` and `read_kernel` give what a plain 256 MiB copy and a plain
read reach on the same GPU in the same run: the HBM rates a kernel that
mixes reads and writes can be held against.

The CPU leg (`cpu_baseline`, skipped with --no-cpu) times the reference's own
Fortran graddin (oracle/_ref, compiled from the SuLi sources) with OpenMP on
the host cores, median of C solves after one warm-up, and runs the CPU oracle
(the C restatement, FP64 CG to a 1e-12 relative residual from x0 = 0) on the
inputs of every GPU result above, to state each result's error against an
FP64 oracle (`oracle_rel_diff`: max |p - p_oracle| / max |p_oracle| over the
interior and face ghosts).
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
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# captured graphs replayed where the engine uses them (off: --no-graphs)
GRAPHS = True

METRIC = "pressure solves/s + p99 latency at 1e-8 residual, mixed vs FP64 CG"
CG_WORDS = 14          # algorithmic FP32/FP64 words per cell of one CG iteration
SPMV_WORDS = 5         # algorithmic words per cell of the SpMV
REF_WORDS = 17         # words per cell per iteration of the reference's sweeps


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def run_solves(solver, grid, df, precision, steps, warmup, tol, graphs=True):
    """Timed solves (tile shape from the autotuner, which runs during the
    first warm-up solve; captured graphs replayed where the engine uses
    them), then two solves launched directly with HIP events around the CG
    step.  Returns the figures and the last timed solve's prd1."""
    solver.configure(precision=precision, stop_rule="relative", tol=tol, verbose=False,
                     graphs=graphs, autotune=True)
    for _ in range(warmup):
        df.zero("prd1")
        solver.graddin_dev(grid, df)
    solver.synchronize()
    lat, iters, rel, refreshes, replays = [], [], [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        df.zero("prd1")
        st = solver.graddin_dev(grid, df)
        lat.append((time.perf_counter() - ts) * 1e3)
        iters.append(st.iters)
        rel.append(st.rel_true)
        refreshes.append(st.refreshes)
        replays.append(st.graph_replays)
        if st.status != 0:
            raise RuntimeError(f"{precision} solve did not converge: {st.as_dict()}")
    solver.synchronize()
    total = time.perf_counter() - t0
    prd1 = df.download("prd1")
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
    from suli_amd.solver import tiles_dict
    kz, vx, _ = solver.autotune(grid, precision)
    out = dict(solves_per_s=steps / total, ms_per_step=total * 1e3 / steps,
               p50_ms=percentile(lat, 50), p99_ms=percentile(lat, 99),
               iters=int(np.median(iters)), refreshes=int(np.median(refreshes)),
               rel_true_max=float(max(rel)), graphs=bool(max(replays) > 0),
               graph_replays_per_solve=float(np.mean(replays)), tiles=tiles_dict(kz, vx),
               step_ms=float(np.mean(spmv)) if spmv else None, spmv_only_ms=spmv_only)
    return out, prd1


def bench_stretched_cylinder(steps, keep):
    """384x192x128 stretched grid with the masked cylinder (not a reference
    mode): solves/s of both solvers."""
    from suli_amd import cases
    from suli_amd.solver import PressureSolver, tiles_dict
    c = cases.config3(seed=1)
    g = c.grid
    solver = PressureSolver(verbose=False)
    df = solver.upload(g, c.fields)
    out, prd = {}, {}
    for precision in ("mixed", "fp64"):
        solver.configure(precision=precision, stop_rule="relative", tol=1e-8, verbose=False,
                         autotune=True, graphs=GRAPHS)
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
                          "rel_true": st.rel_true, "tiles": tiles_dict(kz, vx)}
    df.free()
    solver.close()
    out["speedup_mixed_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["workload"] = (f"stretched dz (tanh, ratio 1.05, finest at 0.6H) {g.nx}x{g.ny}x{g.nz}, masked "
                       "full-height cylinder r = 0.1 Ly, rho 1000/1.2, 1e-8 relative residual")
    keep["stretched_cylinder"] = [(c, {"mixed": prd["mixed"], "fp64": prd["fp64"]})]
    return out


def bench_constant_density(steps, keep):
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
    out, prd = {}, {}
    n = g.ncell
    for precision in ("mixed", "fp64"):
        solver.configure(precision=precision, stop_rule="relative", tol=1e-8, verbose=False,
                         autotune=True, graphs=GRAPHS)
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
        prd[precision] = df.download("prd1")
        # one more solve launched directly, with the CG step timed
        solver.configure(precision=precision, stop_rule="relative", tol=1e-8, verbose=False,
                         autotune=True, timing=True, graphs=GRAPHS)
        df.zero("prd1")
        st = solver.graddin_dev(g, df)
        w = 4 if precision == "mixed" else 8
        spmv_ms = solver.time_spmv(precision, reps=50)
        out[precision] = {"solves_per_s": 1.0 / dt, "ms_per_solve": dt * 1e3, "iters": st.iters,
                          "rel_true": st.rel_true, "step_ms": st.spmv_ms,
                          "spmv_ms": spmv_ms, "spmv_achieved_gbs": 2 * w * n / (spmv_ms * 1e-3) / 1e9}
    df.free()
    solver.close()
    out["speedup_mixed_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["workload"] = (f"IPA (t_press=2, rho_f1=1000) on the tank grid {g.nx}x{g.ny}x{g.nz}, "
                       "1e-8 relative residual; operator read from per-plane tables")
    keep["constant_density"] = [(c, {"mixed": prd["mixed"], "fp64": prd["fp64"]})]
    return out


def bench_batch(keep, seed=1, count=64):
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
    results = [dict() for _ in cs]
    for precision in ("mixed", "fp64"):
        solver.configure(precision=precision, stop_rule="relative", tol=1e-8, verbose=False,
                         autotune=True, graphs=GRAPHS)
        for df in dfs:
            df.zero("prd1")
        solver.graddin_batch_dev(items)      # warm-up: tunes each size
        for df in dfs:
            df.zero("prd1")
        solver.synchronize()
        t0 = time.perf_counter()
        sts = solver.graddin_batch_dev(items)
        t_batch = time.perf_counter() - t0
        for r, df in zip(results, dfs):
            r[precision] = df.download("prd1")
        t0 = time.perf_counter()
        for (g, df) in items:
            df.zero("prd1")
            solver.graddin_dev(g, df)
        t_seq = time.perf_counter() - t0
        its = [s.iters for s in sts]
        cell_iters = sum(c.grid.ncell * s.iters for c, s in zip(cs, sts))
        out[precision] = {"batch_s": t_batch, "solves_per_s": count / t_batch,
                          "mcells_per_s": cells / t_batch / 1e6,
                          "cell_iterations_per_s": cell_iters / t_batch,
                          "one_by_one_solves_per_s": count / t_seq,
                          "batching_speedup": t_seq / t_batch,
                          "iters": {"min": min(its), "median": float(np.median(its)), "max": max(its),
                                    "per_problem": its},
                          "graph_replays": int(sts[0].graph_replays),
                          "not_converged": sum(1 for s in sts if s.status != 0)}
        log(f"batch {precision} {out[precision]['solves_per_s']:.1f} solves/s")
    for df in dfs:
        df.free()
    solver.close()
    out["speedup_mixed_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    keep["batch_uneven"] = list(zip(cs, results))
    return out


def bench_serving(keep, seed=1, count=48, rates=(60.0, 240.0)):
    """Requests of uneven sizes arriving as a Poisson process, served with
    continuous batching; latency from arrival to the result on the host."""
    from suli_amd import cases
    from suli_amd.server import PressureServer, poisson_stream
    base = cases.serving_cases(seed=seed, count=count)
    out = {"requests": count, "rates_per_s": list(rates),
           "workload": f"{count} requests, nx log-uniform [32,256], ny,nz [32,128] (x8), "
                       "two-phase tank fields, Poisson arrivals, 1e-8 relative"}
    results = [dict() for _ in base]
    for precision in ("mixed", "fp64"):
        out[precision] = []
        for ri, rate in enumerate(rates):
            cs = [cases.Case(c.name, c.grid, c.fields.copy(), c.seed) for c in base]
            with PressureServer(precision=precision, stop_rule="relative", tol=1e-8) as srv:
                rep = poisson_stream(srv, cs, rate=rate, seed=seed)
            rep.pop("stats")
            info = rep.pop("info")
            rep["peak_resident_mb"] = info["peak_resident_bytes"] / 1e6
            rep["chunks"] = info["chunks"]
            if ri == 0:
                for r, c in zip(results, cs):
                    r[precision] = c.fields.prd1
            else:
                # each result is the request's own, whatever else was resident
                rep["same_bits_as_first_rate"] = all(
                    np.array_equal(r[precision], c.fields.prd1) for r, c in zip(results, cs))
            out[precision].append(rep)
            log(f"serving {precision} rate {rate}: p50 {rep['p50_ms']:.1f} p99 {rep['p99_ms']:.1f} ms")
    keep["serving"] = list(zip(base, results))
    return out


def bench_dropin(stages=6):
    """The call SuLi makes: host-pointer graddin then posdin (the Fortran
    shim's path, 1_principal.f90:102-103) on the 256x128x96 tank, once per
    RK sub-stage, each solve warm-started from the previous prd1.  Between
    sub-stages the predicted velocity changes by 5 % (u* = (1 + 0.05 s) u0),
    as successive sub-stages of one time step do.  Pageable host arrays;
    times include the PCIe copies."""
    from suli_amd import cases
    from suli_amd.solver import PressureSolver
    c = cases.config2(seed=1)
    g = c.grid
    nx, ny, nz = g.nx, g.ny, g.nz
    n = nx * ny * nz
    nu, nv, nw = (nx + 3) * (ny + 2) * (nz + 2), (nx + 2) * (ny + 3) * (nz + 2), (nx + 2) * (ny + 2) * (nz + 3)
    nbox = (nx + 2) * (ny + 2) * (nz + 2)
    up = 8 * (nu + nv + nw + n + nbox)             # u, v, w, rho, prd1 to the device
    out = {"workload": f"host-pointer graddin + posdin per RK sub-stage, tank {nx}x{ny}x{nz}, "
                       f"{stages} sub-stages, warm start, 1e-8 relative",
           "bytes_per_substage": {"graddin_h2d": up, "graddin_d2h": 8 * nbox,
                                  "posdin_h2d": up, "posdin_h2d_reused": 0,
                                  "posdin_d2h": 8 * (nu + nv + nw)}}
    u0, v0, w0 = c.fields.u.copy(order="F"), c.fields.v.copy(order="F"), c.fields.w.copy(order="F")
    for precision, reuse in (("mixed", True), ("mixed", False), ("fp64", True)):
        f = c.fields.copy()
        f.prd1[...] = 0.0
        ms, iters = [], []
        with PressureSolver(precision=precision, stop_rule="relative", tol=1e-8, verbose=False,
                            autotune=True, reuse_staged=reuse) as s:
            for stage in range(stages + 1):    # the first is a cold start and a warm-up
                a = 1.0 + 0.05 * stage
                f.u[...] = a * u0
                f.v[...] = a * v0
                f.w[...] = a * w0
                t0 = time.perf_counter()
                st = s.graddin(g, f)
                s.posdin(g, f)
                dt = (time.perf_counter() - t0) * 1e3
                if st.status != 0:
                    raise RuntimeError(f"drop-in {precision} did not converge: {st.as_dict()}")
                if stage > 0:
                    ms.append(dt)
                    iters.append(st.iters)
        key = f"{precision}_{'reuse' if reuse else 'upload'}"
        out[key] = {"ms_per_substage": float(np.median(ms)), "iters": iters}
        log(f"drop-in {key}: {out[key]['ms_per_substage']:.2f} ms per sub-stage, iters {iters}")
    return out


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    return d.get(name)


def host_cores():
    """Physical cores of the host, from lscpu (None if it cannot tell)."""
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        return len({line for line in out.splitlines() if line and not line.startswith("#")})
    except (OSError, subprocess.CalledProcessError):
        return None


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


def cpu_timing(case, lam, solves, threads, out_npy):
    """The reference's Fortran graddin (or, if it was not built, the C
    restatement) with OpenMP on the host cores, run in a process of its own
    whose OpenMP runtime binds one thread per core (OMP_PROC_BIND=spread,
    OMP_PLACES=cores): one warm-up solve, then the median of `solves`, all
    on one thread that starts the runtime.  The reference's stop rule is
    absolute, r.r <= 1e-4/N (SuLi-IC/9_grad.f90:393); the inputs are scaled
    by lam so that it equals the relative rule (the solution scales with
    lam; prd1 / lam is saved to out_npy)."""
    import oracle
    from suli_amd import cases
    g = case.grid
    f = cases.scale_velocity(case.fields, lam)
    kind = "reference" if oracle.ref_available("ic") else "port"
    if kind == "reference":
        prd_ref, times = oracle.ref_graddin_timed(g, f, solves + 1, omp_t=1)
    else:
        times = []
        for _ in range(solves + 1):
            t = time.perf_counter()
            prd_ref, _, _ = oracle.graddin(g, f, omp_t=1, threads=threads)
            times.append(time.perf_counter() - t)
    np.save(out_npy, prd_ref / lam)
    t = float(np.median(times[1:]))
    what = ("SuLi graddin (amdflang -O3 -fopenmp, omp_t=1)" if kind == "reference"
            else "C restatement of SuLi graddin (gcc -O2 -fopenmp)")
    return {"value": 1.0 / t, "unit": "solves/s", "cores": threads, "kind": kind,
            "sample": f"{solves} FP64 solves (median, after 1 warm-up) of the same {g.nx}x{g.ny}x{g.nz} "
                      f"case to 1e-8 relative (absolute rule with lam={lam:.3e}); {what}; "
                      f"{t:.3f} s per solve",
            "ms_per_solve": t * 1e3, "all_ms": [x * 1e3 for x in times],
            "threads": threads, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}


def cpu_leg(args, case, headline, keep, avail):
    """The CPU baseline (cpu_timing, in a child process) and the CPU
    oracle's FP64 solution of every GPU result's problem: FP64 CG to 1e-12
    relative from x0 = 0, the constant-pressure component of the initial
    residual removed as the GPU solver does."""
    import tempfile

    import oracle
    from suli_amd import cases
    phys = host_cores()
    threads = int(os.environ.get("OMP_NUM_THREADS", min(16, avail)))
    g = case.grid
    n = g.ncell
    # lam from b.b of the scaled system; the reference rule's iteration count
    # on the scaled inputs from the oracle (the reference keeps it local)
    _, st, _ = oracle.graddin(g, case.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-2, threads=threads)
    lam = float(np.sqrt((1e-4 / n) / (args.tol * args.tol * st.bb)))
    _, st_ref, _ = oracle.graddin(g, cases.scale_velocity(case.fields, lam), stop_rule=oracle.STOP_REFERENCE,
                                  threads=threads, omp_t=1)
    with tempfile.TemporaryDirectory() as tmp:
        npy = os.path.join(tmp, "prd_ref.npy")
        env = dict(os.environ, OMP_PROC_BIND="spread", OMP_PLACES="cores", OMP_NUM_THREADS=str(threads))
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-timing-to", npy, "--cpu-lam", repr(lam),
               "--cpu-solves", str(args.cpu_solves), "--nx", str(args.nx), "--ny", str(args.ny),
               "--nz", str(args.nz)]
        r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError(f"CPU timing failed:\n{r.stderr[-2000:]}")
        cpu = json.loads(r.stdout.strip().splitlines()[-1])
        prd_ref = np.load(npy)
    t = cpu["ms_per_solve"] * 1e-3
    iters = int(st_ref.iters)
    cpu.update({"iters": iters, "iters_from": "oracle, same input and rule",
                "gbs_at_17_words": REF_WORDS * 8 * n * iters / t / 1e9,
                "gbs_at_14_words": CG_WORDS * 8 * n * iters / t / 1e9,
                "physical_cores_lscpu": phys, "cpus_available": avail})
    log(f"cpu baseline {cpu['ms_per_solve']:.0f} ms per solve, {iters} iterations")

    def tight(c):
        p, s, rc = oracle.graddin(c.grid, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-12,
                                  max_iter=50000, threads=threads, omp_t=1, project_null=True)
        if rc != 0:
            raise RuntimeError(f"oracle did not converge on {c.name}")
        return p
    t0 = time.perf_counter()
    checks = {}
    p_or = tight(case)
    checks["headline"] = {k: rel_diff(v, p_or) for k, v in headline.items()}
    # the reference itself at the same 1e-8 rule, for scale
    checks["headline"]["reference_fortran_1e-8"] = rel_diff(prd_ref, p_or)
    for name, items in keep.items():
        per = {}
        for c, res in items:
            p = tight(c)
            for k, v in res.items():
                per.setdefault(k, []).append(rel_diff(v, p))
        checks[name] = {k: (v[0] if len(v) == 1 else {"max": max(v), "median": float(np.median(v)),
                                                       "count": len(v)})
                        for k, v in per.items()}
        log(f"oracle check {name}: {checks[name]}")
    cpu["oracle_checks_s"] = time.perf_counter() - t0
    return cpu, checks


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-solves", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--nz", type=int, default=96)
    ap.add_argument("--only", choices=("both", "mixed", "fp64", "none"), default="both",
                    help="solver(s) to time on the headline workload (profiling runs)")
    ap.add_argument("--copy-mib", type=int, default=256,
                    help="size of the plain copy kernel that calibrates bandwidth and counters")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel directly instead of replaying HIP graphs")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the cylinder, IPA, batch, serving and drop-in workloads")
    ap.add_argument("--extra", default=",".join(EXTRAS),
                    help="comma-separated subset of the extra workloads to run (profiling runs)")
    ap.add_argument("--tiles", action="append", default=[], metavar="PRECISION=KZxVX",
                    help="fix the headline solves' tile shape instead of tuning it (profiling runs, "
                         "whose serialised launches would tune differently), e.g. mixed=6x4, or "
                         "mixed=6x4p for the pipelined FP32 step")
    ap.add_argument("--cpu-timing-to", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-lam", type=float, default=1.0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    global GRAPHS
    GRAPHS = not args.no_graphs
    # logical CPUs this process may use, read before any OpenMP runtime starts
    avail = len(os.sched_getaffinity(0))
    if args.cpu_timing_to:
        from suli_amd import cases
        case = cases.config2(seed=1, nx=args.nx, ny=args.ny, nz=args.nz)
        threads = int(os.environ.get("OMP_NUM_THREADS", min(16, avail)))
        print(json.dumps(cpu_timing(case, args.cpu_lam, args.cpu_solves, threads, args.cpu_timing_to)))
        return
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
    log(f"case {case.name} n={n}")
    copy_bytes = args.copy_mib << 20
    copy_ms = solver.copy_bench(copy_bytes, reps=10)
    copy = {"kernel": "k_copy", "bytes_read": copy_bytes, "bytes_written": copy_bytes,
            "ms": copy_ms, "gbs": 2 * copy_bytes / (copy_ms * 1e-3) / 1e9}
    read_ms = solver.read_bench(copy_bytes, reps=10)
    read = {"kernel": "k_read", "bytes_read": copy_bytes, "ms": read_ms,
            "gbs": copy_bytes / (read_ms * 1e-3) / 1e9}
    log(f"copy {copy['gbs']:.0f} GB/s, read {read['gbs']:.0f} GB/s")
    for spec in args.tiles:
        from suli_amd.solver import TILES_PIPELINED
        prec, _, shape = spec.partition("=")
        kz, _, vx = shape.partition("x")
        pipe = vx.endswith("p")
        solver.set_tiles(g, prec, int(kz), int(vx.rstrip("p")) | (TILES_PIPELINED if pipe else 0))
        log(f"tiles fixed: {prec} kz {kz} vx {vx}")
    mixed = fp64 = None
    headline = {}
    if args.only in ("both", "mixed"):
        mixed, headline["mixed"] = run_solves(solver, g, df, "mixed", args.steps, args.warmup, args.tol,
                                              graphs=not args.no_graphs)
        log(f"mixed {mixed['solves_per_s']:.2f} solves/s, {mixed['iters']} iterations")
    if args.only in ("both", "fp64"):
        fp64, headline["fp64"] = run_solves(solver, g, df, "fp64", args.steps, args.warmup, args.tol,
                                            graphs=not args.no_graphs)
        log(f"fp64 {fp64['solves_per_s']:.2f} solves/s, {fp64['iters']} iterations")
    posdin = None
    if args.only == "both":
        posdin = time_posdin(solver, g, df, reps=20)
        log(f"posdin {posdin['ms'] * 1e3:.1f} us per call, {posdin['achieved']:.0f} GB/s")
    df.free()
    solver.close()
    wanted = [] if args.no_extra else [e for e in args.extra.split(",") if e]
    unknown = sorted(set(wanted) - set(EXTRAS))
    if unknown:
        sys.exit(f"bench.py: unknown --extra workload(s) {unknown}; known: {list(EXTRAS)}")
    extra, keep = run_extras(wanted)
    if mixed is None or fp64 is None:
        # a profiling run: no headline ratio, roofline or CPU leg
        print(json.dumps({"profile_run": args.only, "mixed": mixed, "fp64": fp64, "copy": copy,
                          "workloads": extra}))
        return

    launch_ms = mixed["step_ms"]
    alg = CG_WORDS * 4 * n                 # one CG iteration in FP32: 56 B/cell
    achieved = alg / (launch_ms * 1e-3) / 1e9
    traffic = load_traffic(f"k_cg_m {g.nx}x{g.ny}x{g.nz}")
    spmv_gbs = SPMV_WORDS * 4 * n / (mixed["spmv_only_ms"] * 1e-3) / 1e9
    fp64_step = fp64["step_ms"]
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic,
                "kernel": "k_cg_m", "launch_ms": launch_ms,
                "algorithmic_bytes_per_launch": alg,
                "algorithmic_basis": "one FP32 CG iteration, 14 words/cell (SpMV 5 + x, r, p 3 each)",
                # the bytes the fused step actually moves (profiled HBM traffic) at its launch rate
                "traffic_achieved": traffic / (launch_ms * 1e-3) / 1e9 if traffic else None,
                "traffic_frac": traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                "launch_ms_basis": "HIP events around windows of 8 consecutive launches (gaps included)",
                "spmv": {"kernel": "k_spmv<float>", "words_per_cell": SPMV_WORDS,
                         "launch_ms": mixed["spmv_only_ms"], "achieved": spmv_gbs,
                         "frac": spmv_gbs / HBM_PEAK_GBS,
                         "fp64_launch_ms": fp64["spmv_only_ms"],
                         "fp64_achieved": SPMV_WORDS * 8 * n / (fp64["spmv_only_ms"] * 1e-3) / 1e9},
                "fp64_step": {"kernels": "k_cg_a<double> + k_cg_b<double>", "step_ms": fp64_step,
                              "achieved": CG_WORDS * 8 * n / (fp64_step * 1e-3) / 1e9 if fp64_step else None,
                              "kernel_achieved": 12 * 8 * n / (fp64_step * 1e-3) / 1e9 if fp64_step else None}}

    cpu = checks = None
    if not args.no_cpu:
        log("cpu baseline and oracle checks")
        cpu, checks = cpu_leg(args, case, headline, keep, avail)
        mixed["oracle_rel_diff"] = checks["headline"]["mixed"]
        fp64["oracle_rel_diff"] = checks["headline"]["fp64"]
        for name in keep:
            extra[name]["oracle_rel_diff"] = checks[name]

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
        "oracle_rel_diff": checks["headline"] if checks else None,
        "roofline": roofline,
        "posdin_dev": posdin,
        "copy_kernel": copy,
        "read_kernel": read,
        "workloads": extra,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    reap_children()


def time_posdin(solver, g, df, reps):
    """The device-resident velocity correction (suli_ps_posdin_dev) on the
    headline case, after its solves: mean ms per call (host clock around
    `reps` calls, after one), against its algorithmic bytes in FP64: every
    face of u, v and w read and written, rho and the prd1 halo box read."""
    solver.posdin_dev(g, df)
    solver.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        solver.posdin_dev(g, df)
    solver.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    faces = (g.nx + 1) * g.ny * g.nz + g.nx * (g.ny + 1) * g.nz + g.nx * g.ny * (g.nz + 1)
    nbytes = 8 * (2 * faces + g.ncell + (g.nx + 2) * (g.ny + 2) * (g.nz + 2))
    return {"kernel": "k_posdin", "ms": ms, "algorithmic_bytes": nbytes,
            "achieved": nbytes / (ms * 1e-3) / 1e9, "basis": "host clock around 20 calls"}


def time_calls(fn, reps):
    """Mean ms per call of fn (which returns once its work is done): the host
    clock around `reps` calls, after one untimed call."""
    fn()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    return (time.perf_counter() - t0) * 1e3 / reps


def bench_rows(reps=20, it_steps=3):
    """The rows widened beside the projection step, device-resident:
    rho from the level set (heaviside, t_tens 0 and 1), the hydrostatic
    step (pressh, then posdin's other branch) and IPA's exp_press on the
    256x128x96 tank grid, and SuLi-IT's solve and its whole posdin on a
    256x128x48 SuLi-IT state.  Each entry point: mean ms per call (host
    clock, launch and synchronisation included) and the bytes of the arrays
    it reads and writes, each counted once, at that rate."""
    from suli_amd import cases
    from suli_amd.solver import PressureSolver
    c = cases.config2(seed=1)
    g = c.grid
    nx, ny, nz = g.nx, g.ny, g.nz
    n = nx * ny * nz
    nbox = (nx + 2) * (ny + 2) * (nz + 2)
    nu, nv, nw = (nx + 3) * (ny + 2) * (nz + 2), (nx + 2) * (ny + 3) * (nz + 2), (nx + 2) * (ny + 2) * (nz + 3)
    faces = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)
    rate = lambda nbytes, ms: {"ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9}
    H = nz * g.dz
    ls = cases.sloshing_level_set(g, 0.25 * H / 0.48, 0.02 * H / 0.48)
    eps, rho_w, mu_a, mu_w = 1.5 * g.dx, 1000.0, 1.8e-5, 1.0e-3
    out = {}
    solver = PressureSolver(verbose=False)
    f = c.fields.copy()
    f.hs = np.zeros((nx, ny, nz), order="F")
    df = solver.upload(g, f)
    p = df.ptrs
    p_ls = solver.dev_array(ls)
    p_mu = solver.dev_array(np.zeros((nx, ny, nz)))
    grads = [solver.dev_array(np.zeros((nx, ny, nz))) for _ in range(3)]
    out["heaviside"] = rate(8 * 4 * n, time_calls(
        lambda: solver.heaviside_dev(g, p_ls, eps, rho_w, mu_a, mu_w, p["rho"], p["hs"], p_mu), reps))
    out["heaviside_tens"] = rate(8 * 7 * n, time_calls(
        lambda: solver.heaviside_tens_dev(g, p_ls, eps, rho_w, mu_a, mu_w, p["rho"], *grads, p["hs"], p_mu),
        reps))
    # the hydrostatic step: prd1 from rho, then u, v corrected, w from
    # continuity and the faces where hs < 1e-4 closed
    out["pressh"] = rate(8 * (n + nbox), time_calls(lambda: solver.pressh_dev(g, df, 9.81), reps))
    out["posdin_hydro"] = rate(8 * (2 * (nu + nv) + nw + 2 * n + nbox),
                               time_calls(lambda: solver.posdin_hydro_dev(g, df), reps))
    prd0 = solver.dev_array(np.zeros((nx + 2, ny + 2, nz + 2)))
    pxyz = [solver.dev_array(np.zeros(shape)) for shape in
            ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1))]
    out["exp_press"] = rate(8 * (n + 3 * nbox + faces), time_calls(
        lambda: solver.exp_press_dev(g, p["rho"], p["prd1"], prd0, *pxyz), reps))
    for ptr in (p_ls, p_mu, prd0, *grads, *pxyz):
        solver.dev_free(ptr)
    df.free()
    out["tank"] = f"{nx}x{ny}x{nz}"

    # SuLi-IT: the solve from prd1 = 0 to a 1e-8 relative residual, both
    # precisions, then posdin with its surface and layer update
    gi, fi = cases.it_case(256, 128, 48, seed=1)
    d = solver.it_upload(fi)
    box_bytes = 8 * (gi.nx + 2) * (gi.ny + 2) * (gi.nz + 2)
    it = {"state": f"{gi.nx}x{gi.ny}x{gi.nz}"}
    for precision in ("mixed", "fp64"):
        solver.configure(precision=precision, stop_rule="relative", tol=1e-8, verbose=False, autotune=True,
                         graphs=GRAPHS)
        ms = []
        for k in range(it_steps + 1):
            solver.dev_zero(d["prd1"], box_bytes)
            solver.synchronize()
            t0 = time.perf_counter()
            st = solver.it_graddin_dev(gi, d)
            solver.synchronize()
            if k:
                ms.append((time.perf_counter() - t0) * 1e3)
            if st.status != 0:
                raise RuntimeError(f"SuLi-IT {precision} did not converge: {st.as_dict()}")
        it[precision] = {"solves_per_s": 1e3 / float(np.mean(ms)), "ms_per_solve": float(np.mean(ms)),
                         "iters": st.iters, "rel_true": st.rel_true}
    it["speedup_mixed_vs_fp64"] = it["mixed"]["solves_per_s"] / it["fp64"]["solves_per_s"]
    rng = np.random.default_rng(2)
    dz = np.asfortranarray(float(np.median(fi.dzz)) *
                           (1.0 + 0.1 * rng.uniform(-1, 1, (gi.nx + 2, gi.ny + 2, gi.nz + 2))))
    eta0 = np.asfortranarray(rng.uniform(-0.01, 0.01, (gi.nx + 2, gi.ny + 2)))
    p_dz, p_e0 = solver.dev_array(dz), solver.dev_array(eta0)
    inu, inv, inw = ((gi.nx + 3) * (gi.ny + 2) * (gi.nz + 2), (gi.nx + 2) * (gi.ny + 3) * (gi.nz + 2),
                     (gi.nx + 2) * (gi.ny + 2) * (gi.nz + 3))
    # u, v, w and dzx, dzy, dzz read and written, dz read and written, prd1 read
    it["posdin_layers"] = rate(8 * 4 * (inu + inv + inw) + 3 * box_bytes, time_calls(
        lambda: solver.it_posdin_layers_dev(gi, d, p_dz, p_e0), reps))
    solver.dev_free(p_dz)
    solver.dev_free(p_e0)
    solver.it_free(d)
    solver.close()
    out["suli_it"] = it
    out["basis"] = "host clock around 20 calls after one (launch and synchronisation included)"
    return out


# the extra workloads, in the order they run: (log line, result key, runner)
EXTRAS = {
    "cylinder": ("stretched cylinder", "stretched_cylinder", lambda keep: bench_stretched_cylinder(2, keep)),
    "ipa": ("constant density", "constant_density", lambda keep: bench_constant_density(3, keep)),
    "batch": ("batch", "batch_uneven", lambda keep: bench_batch(keep)),
    "serving": ("serving", "serving", lambda keep: bench_serving(keep)),
    "dropin": ("drop-in call", "dropin_call", lambda keep: bench_dropin()),
    "rows": ("widened rows", "rows", lambda keep: bench_rows()),
}


def run_extras(wanted):
    """The extra workloads named in `wanted`; returns their results (None
    when none ran) and the inputs/outputs kept for the oracle checks."""
    if not wanted:
        return None, {}
    extra, keep = {}, {}
    for name, (title, key, run) in EXTRAS.items():
        if name in wanted:
            log(title)
            extra[key] = run(keep)
    return extra, keep


def reap_children():
    """End the run with no process of its own left behind: any child still
    alive is named on stderr, then terminated and waited for."""
    try:
        import psutil
    except ImportError:
        return
    kids = psutil.Process().children(recursive=True)
    for k in kids:
        try:
            log(f"child process left at exit: pid {k.pid} {' '.join(k.cmdline())}")
            k.terminate()
        except psutil.Error:
            pass
    psutil.wait_procs(kids, timeout=10)


if __name__ == "__main__":
    main()
