#!/usr/bin/env python3
"""Benchmark of the Burton-Miller Galerkin BEM hot path on one MI355X.

Workload (dense horn sweep): closed exponential horn r(x)=r_t exp(m x), throat
12.7 mm, mouth 150 mm, length 300 mm, 5 mm wall with closed rim and rear
closure; driven throat disk (unit RMS normal velocity), rigid elsewhere;
~20.7k triangles / ~10.4k P1 nodes (structured mesh, edge <= 8.3 mm, minimum
angle 35 deg); 16 log-spaced frequencies 200 Hz - 8 kHz; air rho=1.205, c=343.
Dense FP64 Galerkin assembly + device GMRES (restart 50, relative residual
1e-8), far-field directivity at 7 directions. A step is one full sweep
request (16 frequency points) against the device-resident mesh state; the
interaction plan is built in the first (untimed) warmup step and reported as
the cold cost. The LU cross-check of one frequency runs in warmup.

Prints ONE JSON line (rank 0). value = solved frequency points per second.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from hornsculpt_amd import bem, meshes  # noqa: E402

FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FLOPS_PER_SAMPLE = 150.0     # FP64 arithmetic of one Burton-Miller pair sample (sqrt/sincos/div not counted)


def device_synchronize():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()


def horn_request(frequencies, precision, solver, tolerance):
    directions = [(1, 0, 0)] + [(math.cos(math.radians(a)), math.sin(math.radians(a)), 0) for a in (15, 30, 45, 60, 90)] \
        + [(-1, 0, 0)]
    return bem.BemRequest(
        medium=bem.AcousticMedium(1.205, 343.0), frequencies=list(frequencies),
        excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(1, 1.0)])],
        observers=[(0.3 + 10 * d[0], 10 * d[1], 10 * d[2]) for d in directions],
        far_field=bem.FarFieldObservation((0.3, 0.0, 0.0), 1.0),
        numerics=bem.BemNumerics(solver=solver, relative_residual_tolerance=tolerance, gmres_restart=50,
                                 gmres_maximum_iterations=2000, precision=precision))


def cpu_baseline(mesh, frequencies, threads):
    """FP64 oracle on host cores: assembly of a 1/128 row sample at the lowest and highest
    frequency (extrapolated to all rows) plus a complex LU of n=1280 (extrapolated by n^3)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_bind as ob

    nt, nv = len(mesh.triangles), len(mesh.vertices)
    num = bem.BemNumerics()
    kmax = 2 * math.pi * max(frequencies) / 343.0
    plan = bem.interaction_plan(mesh, num, kmax)
    tests = np.arange(0, nt, 128, dtype=np.uint32)
    per_assembly = []
    for f in (min(frequencies), max(frequencies)):
        k = 2 * math.pi * f / 343.0
        t0 = time.perf_counter()
        ob.oracle_assemble(mesh, plan, k, tests=tests, threads=threads)
        per_assembly.append((time.perf_counter() - t0) * nt / len(tests))
    n = 1280
    rng = np.random.default_rng(0)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + n * np.eye(n)
    b = rng.standard_normal((n, 1)) + 0j
    t0 = time.perf_counter()
    ob.oracle_lu_solve(A, b, threads=threads)
    lu = (time.perf_counter() - t0) * (nv / n) ** 3
    per_point = float(np.mean(per_assembly)) + lu
    return {"value": 1.0 / per_point, "unit": "freq-points/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle: assembly of {len(tests)}/{nt} test-panel rows at {min(frequencies):.0f} and "
                      f"{max(frequencies):.0f} Hz scaled to all rows ({np.mean(per_assembly):.1f} s/point) + "
                      f"complex LU n={n} scaled by (nv/n)^3 to nv={nv} ({lu:.1f} s/point)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--edge", type=float, default=0.0052)
    ap.add_argument("--frequencies", type=int, default=16)
    ap.add_argument("--solver", choices=["gmres", "lu"], default="gmres")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--tolerance", type=float, default=1e-8)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cross-check", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: every GPU box here has one MI355X and the solver has no multi-GPU path "
              f"(requested --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    frequencies = [200.0 * 40.0 ** (i / (args.frequencies - 1)) for i in range(args.frequencies)]
    mesh = meshes.exponential_horn(args.edge)
    nt, nv = len(mesh.triangles), len(mesh.vertices)
    solver_kind = bem.SOLVER_GMRES if args.solver == "gmres" else bem.SOLVER_LU
    precision = bem.PRECISION_MIXED if args.precision == "mixed" else bem.PRECISION_FP64
    req = horn_request(frequencies, precision, solver_kind, args.tolerance)
    print(f"# horn: {nt} triangles, {nv} P1 nodes, {len(frequencies)} frequencies", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    solver = bem.HipBemSolver(mesh)
    construct_s = time.perf_counter() - t0
    plan_ms = None
    cross = None
    for w in range(args.warmup):
        t = time.perf_counter()
        res = solver.solve(req)
        if plan_ms is None:
            plan_ms = res.frequencies[0].interaction_plan_milliseconds
            cold_s = time.perf_counter() - t
        print(f"# warmup {w}: {time.perf_counter() - t:.2f} s", file=sys.stderr, flush=True)
    if not args.no_cross_check:
        # Dense LU cross-check of the GMRES sweep at the highest frequency (untimed).
        lu_req = horn_request([frequencies[-1]], precision, bem.SOLVER_LU, 1e-10)
        r_lu = solver.solve(lu_req)
        r_gm = solver.solve(horn_request([frequencies[-1]], precision, solver_kind, args.tolerance))
        a = r_lu.frequencies[0].excitations[0].boundary.surface_pressure
        b = r_gm.frequencies[0].excitations[0].boundary.surface_pressure
        cross = float(np.linalg.norm(a - b) / np.linalg.norm(a))
        print(f"# LU cross-check at {frequencies[-1]:.0f} Hz: relative L2 {cross:.2e}, LU factor "
              f"{r_lu.frequencies[0].factorization_milliseconds:.0f} ms", file=sys.stderr, flush=True)

    latencies = []
    kernel_ms, samples, iterations = 0.0, 0.0, []
    stamps = []

    def progress(done, total, f, phase):
        if phase == 6:
            stamps.append(time.perf_counter())

    device_synchronize()
    start = time.perf_counter()
    for _ in range(args.steps):
        stamps.clear()
        t_step = time.perf_counter()
        res = solver.solve(req, progress=progress)
        prev = t_step
        for s in stamps:
            latencies.append(s - prev)
            prev = s
        for fr in res.frequencies:
            kernel_ms += fr.kernel_milliseconds["regular"]
            samples += fr.kernel_milliseconds["regular_samples"]
            iterations.append(fr.excitations[0].iterations)
    device_synchronize()
    elapsed = time.perf_counter() - start
    points = args.steps * len(frequencies)
    value = points / elapsed
    achieved = FLOPS_PER_SAMPLE * samples / (kernel_ms / 1e3) / 1e12 if kernel_ms > 0 else None
    line = {
        "metric": "freq-points solved/s (plain FP64 dense form; p50/p99 per-point latency)",
        "value": value, "unit": "freq-points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if precision == bem.PRECISION_FP64 else "f32+f64",
        "data": "synthetic (parametric horn mesh generated in-process)",
        "config": {"workload": "dense horn sweep: exponential horn r_t=12.7 mm, r_m=150 mm, L=300 mm, 5 mm wall, "
                               f"{nt} triangles / {nv} P1 nodes, {len(frequencies)} log-spaced frequencies "
                               f"200 Hz-8 kHz, dense assembly + {args.solver.upper()} "
                               f"(restart 50, rel tol {args.tolerance:g}), 7 far-field directions",
                   "triangles": nt, "nodes": nv, "frequencies": len(frequencies), "solver": args.solver,
                   "precision": args.precision},
        "latency_ms": {"p50": 1e3 * float(np.percentile(latencies, 50)),
                       "p99": 1e3 * float(np.percentile(latencies, 99))},
        "cold": {"mesh_setup_s": construct_s, "interaction_plan_ms": plan_ms},
        "gmres_iterations": {"min": int(min(iterations)), "max": int(max(iterations))},
        "lu_cross_check_rel_l2": cross,
        "roofline": {"kernel": "regular_tiles (FP64 pair quadrature, VALU)", "bound": "mfma",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS if achieved else None, "traffic": None,
                     "launch_ms": kernel_ms / points, "flops_per_launch": FLOPS_PER_SAMPLE * samples / points},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(mesh, frequencies, args.cpu_threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
