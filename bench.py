#!/usr/bin/env python3
"""Benchmark of the Burton-Miller Galerkin BEM hot path on one MI355X.

Workload (dense horn sweep): closed exponential horn r(x)=r_t exp(m x), throat
12.7 mm, mouth 150 mm, length 300 mm, 5 mm wall with closed rim and rear
closure; driven throat disk (unit RMS normal velocity), rigid elsewhere;
~20.7k triangles / ~10.4k P1 nodes (structured mesh, edge <= 8.3 mm, minimum
angle 35 deg); 16 log-spaced frequencies 200 Hz - 8 kHz; air rho=1.205, c=343.
Dense Galerkin assembly + device GMRES (restart 50, relative residual 1e-8),
far-field directivity at 7 directions. Two forms are timed on the same
resident mesh: plain FP64 and mixed (FP32 regular/near samples and matrix,
FP64 singular pairs, Krylov vectors and residual). A step is one full sweep
request (16 frequency points); the interaction plan is built in the first
(untimed) warmup step and reported as the cold cost. The FP64 LU cross-check
of one frequency is untimed.

Forms: plain FP64 and mixed, each solved directly (one request at a time, one
stream) and through the serving loop (slots with their own streams taking the
frequency points of all queued requests; saturated throughput from `steps`
sweep requests submitted together, latency from an open-loop Poisson stream of
single-point requests at 70 % of that throughput).

Prints ONE JSON line. value = solved frequency points per second of the served
mixed form (or of the form asked for), with the plain FP64 form and the mixed
form's surface-pressure error against it beside it.
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

from hornsculpt_amd import bem, meshes, variants  # noqa: E402

FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 vector (published; half of it without packed FP32 instructions)
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FLOPS_PER_SAMPLE = 150.0     # FP64 arithmetic of one Burton-Miller pair sample (sqrt/sincos/div not counted)
# The device counts ordered-pair samples (the Galerkin method's count); the tile kernels
# integrate each unordered panel pair once for both orientations, so they evaluate
# half as many sample points. Achieved rates count 150 flops per evaluated point.
EVALUATED_PER_SAMPLE = 0.5


def evaluated_flops(samples):
    return FLOPS_PER_SAMPLE * EVALUATED_PER_SAMPLE * samples


# HBM bytes per leg launch (one frequency point of the 121 360-triangle pool mesh, mean of
# 2500 Hz and 20 kHz) of regular_tiles_sym<float>: rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE x 1
# (1 GiB-copy calibration), profiles/hbm_traffic_leg_121k.md. Mostly the node-tile flush:
# FP32 atomics into A, which the L2 forwards to memory as one request each.
LEG_TRAFFIC_BYTES = {121360: (142.72 + 523.46) * 1e9}
# HBM bytes per regular_tiles launch on the default horn mesh (edge 5.2 mm), from rocprofv3
# FETCH_SIZE + WRITE_SIZE runs scaled by a 1 GiB copy calibration (fetch x2.000, write x1.000):
# most of it is the node-tile flush through L2 atomics (VALU-bound kernel).


def device_synchronize():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()


def horn_request(frequencies, precision, solver, tolerance, planning=()):
    directions = [(1, 0, 0)] + [(math.cos(math.radians(a)), math.sin(math.radians(a)), 0) for a in (15, 30, 45, 60, 90)] \
        + [(-1, 0, 0)]
    return bem.BemRequest(
        medium=bem.AcousticMedium(1.205, 343.0), frequencies=list(frequencies),
        excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(1, 1.0)])],
        observers=[(0.3 + 10 * d[0], 10 * d[1], 10 * d[2]) for d in directions],
        far_field=bem.FarFieldObservation((0.3, 0.0, 0.0), 1.0), interaction_plan_frequencies=list(planning),
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


def run_form(solver, frequencies, precision, solver_kind, tolerance, steps, warmup, label):
    """Warmup, then EXACTLY `steps` timed sweeps bracketed by device synchronisations."""
    req = horn_request(frequencies, precision, solver_kind, tolerance)
    plan_ms = None
    for w in range(warmup):
        t = time.perf_counter()
        res = solver.solve(req)
        if plan_ms is None:
            plan_ms = res.frequencies[0].interaction_plan_milliseconds
        print(f"# {label} warmup {w}: {time.perf_counter() - t:.2f} s", file=sys.stderr, flush=True)
    latencies, stamps = [], []
    kernel_ms, samples, iterations = 0.0, 0.0, []
    phases = dict.fromkeys(("assembly", "regular", "touching", "near", "solve", "field"), 0.0)

    def progress(done, total, f, phase):
        if phase == 6:
            stamps.append(time.perf_counter())

    device_synchronize()
    start = time.perf_counter()
    for _ in range(steps):
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
            phases["assembly"] += fr.assembly_milliseconds
            phases["solve"] += fr.solve_milliseconds
            phases["field"] += fr.field_milliseconds
            for key in ("regular", "touching", "near"):
                phases[key] += fr.kernel_milliseconds[key]
    device_synchronize()
    elapsed = time.perf_counter() - start
    points = steps * len(frequencies)
    print(f"# {label}: {points / elapsed:.2f} points/s; per point ms: "
          + ", ".join(f"{k} {v / points:.1f}" for k, v in phases.items()), file=sys.stderr, flush=True)
    return {"value": points / elapsed, "elapsed": elapsed, "points": points, "plan_ms": plan_ms,
            "p50": 1e3 * float(np.percentile(latencies, 50)), "p99": 1e3 * float(np.percentile(latencies, 99)),
            "kernel_ms": kernel_ms, "samples": samples, "iterations": iterations,
            "traces": [fr.excitations[0].boundary.surface_pressure for fr in res.frequencies]}


def run_served(solver, frequencies, precision, tolerance, steps, slots, label, open_loop_points=0, seed=20261203):
    """Serving loop: `steps` sweep requests submitted together (saturated throughput),
    then optionally an open-loop Poisson stream of single-frequency requests at 70 %
    of that throughput (latency from submission to completion of each point)."""
    req = horn_request(frequencies, precision, bem.SOLVER_GMRES, tolerance)
    server = bem.HipBemServer(slots=slots)
    server.poll(server.submit(solver, horn_request(frequencies[:slots], precision, bem.SOLVER_GMRES, tolerance)))
    device_synchronize()
    start = time.perf_counter()
    tickets = [server.submit(solver, req) for _ in range(steps)]
    results = [server.poll(t) for t in tickets]
    device_synchronize()
    elapsed = time.perf_counter() - start
    points = steps * len(frequencies)
    kernel_ms = sum(fr.kernel_milliseconds["regular"] for r in results for fr in r.frequencies)
    samples = sum(fr.kernel_milliseconds["regular_samples"] for r in results for fr in r.frequencies)
    out = {"value": points / elapsed, "elapsed": elapsed, "points": points, "plan_ms": None,
           "kernel_ms": kernel_ms, "samples": samples,
           "iterations": [fr.excitations[0].iterations for r in results for fr in r.frequencies],
           "traces": [fr.excitations[0].boundary.surface_pressure for fr in results[-1].frequencies],
           "stats": server.stats()}
    print(f"# {label}: {points / elapsed:.2f} points/s ({slots} slots, {steps} sweep requests)",
          file=sys.stderr, flush=True)
    if open_loop_points:
        rng = np.random.default_rng(seed)
        rate = 0.7 * out["value"]
        gaps = rng.exponential(1.0 / rate, open_loop_points)
        picks = rng.integers(0, len(frequencies), open_loop_points)
        t0 = time.perf_counter()
        arrival = t0
        tickets = []
        for gap, i in zip(gaps, picks):
            arrival += gap
            delay = arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            # Points drawn from the design sweep share its interaction plan (the sweep is
            # their planning raster, interactionPlanFrequencies in the reference request).
            tickets.append(server.submit(solver, horn_request([frequencies[i]], precision, bem.SOLVER_GMRES,
                                                              tolerance, planning=frequencies)))
        lat = [server.poll(t).frequencies[0].served_milliseconds for t in tickets]
        out["p50"] = float(np.percentile(lat, 50))
        out["p99"] = float(np.percentile(lat, 99))
        out["open_loop"] = {"requests": open_loop_points, "rate_per_s": rate,
                            "achieved_per_s": open_loop_points / (time.perf_counter() - t0)}
        print(f"# {label} open loop: {open_loop_points} single-point requests at {rate:.2f}/s, latency p50 "
              f"{out['p50']:.0f} ms p99 {out['p99']:.0f} ms", file=sys.stderr, flush=True)
    server.close()
    return out


def _area_weighted_error(mesh, a, b):
    V, T = mesh.vertices, mesh.triangles
    area = 0.5 * np.linalg.norm(np.cross(V[T[:, 1]] - V[T[:, 0]], V[T[:, 2]] - V[T[:, 0]]), axis=1)
    return float(math.sqrt(np.sum(area * np.abs(a - b) ** 2) / np.sum(area * np.abs(b) ** 2)))


def run_sweep(variant_count, frequency_count, slots, tolerance):
    """Batched sweep: the first `variant_count` seeded horn variants (of the 32
    drawn from mt19937_64(20261202)), `frequency_count` frequencies spread over the
    64-point 100 Hz - 20 kHz raster (the raster is every request's planning raster),
    four throat right-hand sides per frequency. Every (variant, frequency) request
    is submitted to the serving loop at once, first in the plain FP64 form, then in
    the mixed form; a warm pass per variant builds the interaction plans (cold cost,
    reported apart). Error: area-weighted surface-pressure error of the mixed form
    against the FP64 form, worst over variants, frequencies and right-hand sides."""
    raster = variants.sweep_frequencies()
    picks = sorted({round(i * (len(raster) - 1) / max(frequency_count - 1, 1)) for i in range(frequency_count)})
    chosen = [raster[i] for i in picks]
    drawn = variants.horn_variants()
    sizes_all = [len(v.mesh().triangles) for v in drawn]
    work = []
    for v in drawn[:variant_count]:
        mesh, excitations = variants.throat_excitations(v.mesh())
        work.append((v, mesh, bem.HipBemSolver(mesh), excitations))
        print(f"# variant {v.index}: r_m {v.mouth_radius * 1e3:.1f} mm, L {v.length * 1e3:.1f} mm, edge "
              f"{v.edge * 1e3:.2f} mm, {len(mesh.triangles)} triangles", file=sys.stderr, flush=True)

    def request(excitations, frequencies, precision):
        return bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=list(frequencies),
                              excitations=excitations, interaction_plan_frequencies=raster,
                              observers=[(1.0, 0.0, 0.0), (0.866, 0.5, 0.0), (0.866, 0.0, 0.5)],
                              far_field=bem.FarFieldObservation((0.0, 0.0, 0.0), 1.0),
                              numerics=bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=tolerance,
                                                       gmres_restart=50, gmres_maximum_iterations=2000,
                                                       precision=precision))

    t0 = time.perf_counter()
    for v, _, solver, excitations in work:  # cold: interaction plans at the raster's planning wavenumber
        r = solver.solve(request(excitations, [chosen[0]], bem.PRECISION_MIXED))
        print(f"# variant {v.index}: interaction plan {r.frequencies[0].interaction_plan_milliseconds / 1e3:.1f} s "
              f"(cold, {time.perf_counter() - t0:.0f} s so far)", file=sys.stderr, flush=True)
    cold_s = time.perf_counter() - t0
    out = {"variants": [{"index": v.index, "mouth_radius": v.mouth_radius, "length": v.length, "edge": v.edge,
                         "triangles": len(m.triangles), "nodes": len(m.vertices)} for v, m, _, _ in work],
           "frequencies": chosen, "right_hand_sides": 4, "cold_s": cold_s,
           "triangles_all_32_variants": [min(sizes_all), max(sizes_all)]}
    results = {}
    for label, precision in (("fp64", bem.PRECISION_FP64), ("mixed", bem.PRECISION_MIXED)):
        server = bem.HipBemServer(slots=slots)
        device_synchronize()
        start = time.perf_counter()
        tickets = [server.submit(solver, request(excitations, [f], precision))
                   for _, _, solver, excitations in work for f in chosen]
        done = []
        for i, t in enumerate(tickets):
            done.append(server.poll(t))
            if (i + 1) % len(chosen) == 0:
                print(f"# {label}: variant {(i + 1) // len(chosen)} of {len(work)} done", file=sys.stderr, flush=True)
        device_synchronize()
        elapsed = time.perf_counter() - start
        stats = server.stats()
        server.close()
        results[label] = done
        points = len(tickets)
        out[label] = {"points_per_s": points / elapsed, "elapsed_s": elapsed, "points": points,
                      "p50_ms": float(np.percentile([r.frequencies[0].served_milliseconds for r in done], 50)),
                      "p99_ms": float(np.percentile([r.frequencies[0].served_milliseconds for r in done], 99)),
                      "iterations": [min(e.iterations for r in done for e in r.frequencies[0].excitations),
                                     max(e.iterations for r in done for e in r.frequencies[0].excitations)],
                      "max_running": stats["max_running"]}
        print(f"# sweep {label}: {points / elapsed:.2f} points/s ({points} points x 4 right-hand sides)",
              file=sys.stderr, flush=True)
    worst = 0.0
    meshes_per_ticket = [m for _, m, _, _ in work for _ in chosen]
    for mesh, rf, rm in zip(meshes_per_ticket, results["fp64"], results["mixed"]):
        for ef, em in zip(rf.frequencies[0].excitations, rm.frequencies[0].excitations):
            worst = max(worst, _area_weighted_error(mesh, em.boundary.surface_pressure, ef.boundary.surface_pressure))
    out["mixed_vs_fp64_surface_pressure"] = worst
    out["mixed_over_fp64"] = out["mixed"]["points_per_s"] / out["fp64"]["points_per_s"]
    return out


STREAM_SEED = 20261203   # arrivals and request draws (mt19937_64)
POOL_SEED = 20261204     # horn shapes of the resident mesh pool
DIRECTIVITY_DEG = (0.0, 15.0, 30.0, 45.0, 60.0, 90.0)


def cpu_description():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((line.split(":", 1)[1].strip() for line in f if line.startswith("model name")), model)
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "nproc": usable}


class StreamPool:
    """Resident horn meshes of the served stream: `size` seeded horn shapes
    (mt19937_64(POOL_SEED): mouth radius U[80, 200] mm, length U[150, 450] mm) meshed
    at triangle counts spread log-uniformly over [lo, hi] (stratified, so the pool
    spans the whole range), one solver per mesh hash, each with the highest frequency
    its longest edge admits (k h_max <= 2, kept 0.1 % inside)."""

    def __init__(self, size, lo, hi):
        targets = [lo * (hi / lo) ** (i / max(size - 1, 1)) for i in range(size)]
        self.entries = []
        for v in variants.horn_variants(size, seed=POOL_SEED, triangles=(lo, hi), max_edge=math.inf, targets=targets):
            mesh = v.mesh()
            cap = 0.999 * 2.0 * 343.0 / (2 * math.pi * variants.maximum_edge(mesh))
            self.entries.append({"mesh": mesh, "solver": bem.HipBemSolver(mesh), "cap": cap, "variant": v,
                                 "triangles": len(mesh.triangles), "nodes": len(mesh.vertices)})

    def request(self, index, frequencies, precision, tolerance):
        e = self.entries[index]
        length = e["variant"].length
        # far field about the mouth centre; the horn axis is +x
        directions = [(math.cos(math.radians(a)), math.sin(math.radians(a)), 0.0) for a in DIRECTIVITY_DEG]
        return bem.BemRequest(
            medium=bem.AcousticMedium(1.205, 343.0), frequencies=list(frequencies),
            excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(1, 1.0)])],
            observers=[(length + 10 * d[0], 10 * d[1], 10 * d[2]) for d in directions],
            far_field=bem.FarFieldObservation((length, 0.0, 0.0), 1.0),
            interaction_plan_frequencies=sorted(list(frequencies) + [e["cap"]]),
            numerics=bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=tolerance, gmres_restart=50,
                                     gmres_maximum_iterations=3000, precision=precision))


def draw_stream(pool, rate, duration_s):
    """Open-loop Poisson arrivals at `rate` per second for `duration_s`: per request a
    pool mesh (uniform over the log-spread pool, so log-uniform in size), a frequency
    count round(log-uniform[1, 32]) and that many frequencies log-uniform in
    [100 Hz, 20 kHz], sorted; a draw above the mesh's k h_max = 2 limit is redrawn."""
    g = variants.Mt19937_64(STREAM_SEED)
    draws, redraws, t = [], 0, 0.0
    n = len(pool.entries)
    while True:
        t += -math.log(1.0 - variants.uniform(g, 0.0, 1.0)) / rate
        if t > duration_s:
            break
        index = min(n - 1, int(variants.uniform(g, 0.0, n)))
        while True:
            count = int(round(math.exp(variants.uniform(g, 0.0, math.log(32.0)))))
            freqs = sorted(math.exp(variants.uniform(g, math.log(100.0), math.log(20000.0))) for _ in range(count))
            if freqs[-1] <= pool.entries[index]["cap"]:
                break
            redraws += 1
        draws.append((t, index, freqs))
    return draws, redraws / max(1, redraws + len(draws))


def serve_stream(pool, draws, precision, tolerance, slots, window_s, warmup_steps, steps, label):
    """One form over the stream: arrivals submitted on their schedule for the whole
    run; the first `warmup_steps` windows are untimed, the next `steps` windows are
    timed (points completed inside them / their duration). Requests still open at
    the end are cancelled; their latency is censored at the end of the run."""
    server = bem.HipBemServer(slots=slots)
    start = time.perf_counter()
    timed_from, end = start + warmup_steps * window_s, start + (warmup_steps + steps) * window_s
    tickets, base_points, base_time, i = [], None, None, 0
    next_report = start + 15.0
    while True:
        now = time.perf_counter()
        if base_points is None and now >= timed_from:
            base_points, base_time = server.stats()["points"], now
        if now >= end:
            break
        if now >= next_report:
            st = server.stats()
            print(f"# {label}: t={now - start:.0f} s, {i} requests submitted, {st['points']} points done, "
                  f"{st['queued']} queued", file=sys.stderr, flush=True)
            next_report += 15.0
        if i < len(draws) and start + draws[i][0] <= now:
            tickets.append(server.submit(pool.entries[draws[i][1]]["solver"],
                                         pool.request(draws[i][1], draws[i][2], precision, tolerance)))
            i += 1
            continue
        wake = min(t for t in (timed_from if base_points is None else end, end,
                               start + draws[i][0] if i < len(draws) else end, next_report))
        time.sleep(max(0.0, min(wake - time.perf_counter(), 0.05)))
    end_points, end_time = server.stats()["points"], time.perf_counter()
    results, latencies, censored, finished = {}, [], 0, {}
    kernel_ms, samples = 0.0, 0.0
    phases = {"assembly": 0.0, "solve": 0.0, "factorization": 0.0, "field": 0.0, "interaction_plan": 0.0}
    for n, ticket in enumerate(tickets):
        res = server.poll(ticket, 0)
        if res is None:
            server.cancel(ticket)
            censored += 1
            latencies.append(1e3 * (end_time - (start + draws[n][0])))
            continue
        results[n] = res
        latencies.append(max(fr.served_milliseconds for fr in res.frequencies))
        finished[n] = draws[n][0] + latencies[-1] / 1e3  # completion, seconds after the stream's start
        for fr in res.frequencies:
            kernel_ms += fr.kernel_milliseconds["regular"]
            samples += fr.kernel_milliseconds["regular_samples"]
            for phase in phases:
                phases[phase] += getattr(fr, f"{phase}_milliseconds")
    stats = server.stats()
    server.close()
    timed = end_time - base_time
    out = {"points_per_s": (end_points - base_points) / timed, "timed_s": timed,
           "points_timed": end_points - base_points, "requests_submitted": len(tickets),
           "requests_completed": len(results), "requests_censored": censored,
           "p50_ms": float(np.percentile(latencies, 50)), "p99_ms": float(np.percentile(latencies, 99)),
           "latency_censored_at_end": censored > 0, "max_running": stats["max_running"],
           "peak_held_bytes": stats["peak_held_bytes"], "budget_bytes": stats["budget_bytes"],
           "regular_kernel_ms_served": kernel_ms, "regular_samples_served": samples,
           "phase_ms_completed": phases, "finished_s": finished}
    print(f"# {label}: {out['points_per_s']:.2f} points/s over {timed:.0f} s timed, {len(results)}/{len(tickets)} "
          f"requests done, latency p50 {out['p50_ms']:.0f} ms p99 {out['p99_ms']:.0f} ms", file=sys.stderr, flush=True)
    return out, results


def common_prefix(draws, forms):
    """Both forms on identical work: the longest arrival-order prefix of requests that
    every form completed, its points, and per form the time from the stream's start
    to the last completion among them (the forms see the same arrivals and serve
    them in arrival order, so this is each form's time for the same work)."""
    k = 0
    while all(k in f["finished_s"] for f in forms.values()):
        k += 1
    if k == 0:
        return None
    points = sum(len(draws[n][2]) for n in range(k))
    times = {name: max(f["finished_s"][n] for n in range(k)) for name, f in forms.items()}
    return {"requests": k, "points": points, "seconds": times,
            "points_per_s": {name: points / t for name, t in times.items()}}


def _directivity_db(observers):
    p = np.abs(np.asarray(observers))
    return 20 * np.log10(np.maximum(p, 1e-300) / max(p[0], 1e-300))


def compare_forms(pool, draws, exact, mixed):
    """Per-result error of the mixed form against the plain FP64 form (which equals
    the FP64 oracle to ~1e-15 in the parity tests) on every request both completed:
    area-weighted relative L2 of the surface pressure, on-axis SPL difference (dB) and
    the largest directivity difference (dB, levels relative to on-axis, 0-90 deg)."""
    worst = {"surface_pressure": 0.0, "on_axis_spl_db": 0.0, "directivity_db": 0.0}
    points = 0
    for n in sorted(set(exact) & set(mixed)):
        mesh = pool.entries[draws[n][1]]["mesh"]
        for fe, fm in zip(exact[n].frequencies, mixed[n].frequencies):
            ee, em = fe.excitations[0], fm.excitations[0]
            worst["surface_pressure"] = max(worst["surface_pressure"], _area_weighted_error(
                mesh, em.boundary.surface_pressure, ee.boundary.surface_pressure))
            worst["on_axis_spl_db"] = max(worst["on_axis_spl_db"], abs(
                bem.spl_db(complex(em.observer_pressure[0])) - bem.spl_db(complex(ee.observer_pressure[0]))))
            worst["directivity_db"] = max(worst["directivity_db"], float(np.max(np.abs(
                _directivity_db(em.observer_pressure) - _directivity_db(ee.observer_pressure)))))
            points += 1
    worst["points_compared"] = points
    return worst


def tuning_summary():
    """The autotuner's decisions so far: per tunable, the chosen candidates with their
    counts over keys ("gemv_shape" decoded to "COLSxROWS"), and every decision."""
    report = bem.tuning_report()
    label = {"gemv_shape": lambda c: f"{c & 0xffff}x{max(c >> 16, 1)}"}
    chosen = {}
    for r in report:
        name = label.get(r["tunable"], str)(int(r["chosen"]))
        chosen.setdefault(r["tunable"], {}).setdefault(name, 0)
        chosen[r["tunable"]][name] += 1
    return {"chosen": chosen, "decisions": report}


def roofline_leg(pool, precision, tolerance):
    """The dominant kernel alone on the GPU: one direct solve (no other slot) of the
    largest pool mesh at two frequencies of its raster (its limit / 8 and its limit);
    FP32/FP64 sample flops over the HIP-event time of regular_tiles on the solver's
    stream."""
    e = pool.entries[-1]
    freqs = [min(e["cap"], 20000.0) / 8, min(e["cap"], 20000.0)]
    res = e["solver"].solve(pool.request(len(pool.entries) - 1, freqs, precision, tolerance))
    ms = sum(fr.kernel_milliseconds["regular"] for fr in res.frequencies)
    samples = sum(fr.kernel_milliseconds["regular_samples"] for fr in res.frequencies)
    launches = len(res.frequencies)
    return {"mesh_triangles": e["triangles"], "frequencies_hz": freqs, "launch_ms": ms / launches,
            "flops_per_launch": evaluated_flops(samples) / launches,
            "achieved_tflops": evaluated_flops(samples) / (ms / 1e3) / 1e12,
            "ordered_pair_tflops": FLOPS_PER_SAMPLE * samples / (ms / 1e3) / 1e12}


def cpu_stream_baseline(pool, threads):
    """FP64 oracle on host cores for one point of the smallest pool mesh (full dense
    assembly + LU) and C1 (icosphere L3, 1 kHz) in full; the per-point time on the
    stream's mix is scaled from the small mesh by the mix's mean squared node count
    (assembly cost grows with nv^2; the LU's nv^3 is reported apart)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_bind as ob

    def full_point(mesh, f):
        k = 2 * math.pi * f / 343.0
        t0 = time.perf_counter()
        plan = bem.interaction_plan(mesh, bem.BemNumerics(), k)
        t1 = time.perf_counter()
        A, B = ob.oracle_assemble(mesh, plan, k, threads=threads)
        t2 = time.perf_counter()
        q = np.array([-1j * 1.205 * 2 * math.pi * f * (1.0 if fid == 1 else 0.0) for fid in mesh.face_ids])
        ob.oracle_lu_solve(A, B @ q, threads=threads)
        t3 = time.perf_counter()
        return t1 - t0, t2 - t1, t3 - t2

    small = pool.entries[0]
    f_small = min(small["cap"], 20000.0) / 2
    plan_s, asm_s, lu_s = full_point(small["mesh"], f_small)
    nv2 = np.mean([e["nodes"] ** 2 for e in pool.entries])
    nv3 = np.mean([e["nodes"] ** 3 for e in pool.entries])
    per_point = asm_s * nv2 / small["nodes"] ** 2 + lu_s * nv3 / small["nodes"] ** 3
    c1_plan, c1_asm, c1_lu = full_point(meshes.icosphere(3), 1000.0)
    return {"value": 1.0 / per_point, "unit": "freq-points/s", "cores": threads, "kind": "port",
            "cpu": cpu_description(),
            "sample": f"FP64 oracle: one full point of the smallest pool mesh ({small['triangles']} triangles, "
                      f"{f_small:.0f} Hz: assembly {asm_s:.1f} s, LU {lu_s:.2f} s, host plan {plan_s:.1f} s) scaled "
                      f"to the stream's mesh mix by mean nv^2 (assembly) and nv^3 (LU): {per_point:.0f} s per point",
            "c1_full_solve_s": c1_plan + c1_asm + c1_lu,
            "c1_detail": {"plan_s": c1_plan, "assembly_s": c1_asm, "lu_s": c1_lu}}


def run_fmm(solver, frequencies, leaf_capacity, repetitions, check=True):
    """FMM pressure operator on the same mesh: build time, HIP-event apply time and
    the agreement of its action with the dense FP64 device assembly (a random vector)."""
    num = bem.BemNumerics()
    rng = np.random.default_rng(11)
    out = []
    for f in frequencies:
        k = 2 * math.pi * f / 343.0
        t0 = time.perf_counter()
        op = solver.fmm_operator(num, k, k, leaf_capacity=leaf_capacity)
        build_ms = 1e3 * (time.perf_counter() - t0)
        st = op.statistics()
        apply_ms = op.time_apply(repetitions)
        apply32_ms = op.time_apply(repetitions, fp32=True) if st["plane_wave_levels"] == 0 else None
        err = None
        if check:
            A, _ = solver.assemble(num, k, k)
            x = rng.standard_normal(A.shape[0]) + 1j * rng.standard_normal(A.shape[0])
            y = A @ x
            err = float(np.linalg.norm(op.apply(x) - y) / np.linalg.norm(y))
            del A
        out.append({"frequency_hz": f, "build_ms": round(build_ms, 1), "apply_ms": round(apply_ms, 3),
                    "apply_fp32_ms": round(apply32_ms, 3) if apply32_ms else None,
                    "apply_rel_err_vs_dense": err, "level_orders": st["level_orders"], "leaves": st["leaves"],
                    "m2l_links": st["m2l_links"], "near_nnz": st["near_nnz"],
                    "device_mb": round(st["device_bytes"] / 1e6, 1),
                    "build_phases_ms": {key: round(st[f"{key}_ms"], 1) for key in ("tree", "near", "plans", "moments")}})
        print(f"# FMM {f:.0f} Hz: build {build_ms:.0f} ms (tree + expansion plan {st['tree_ms']:.0f}, near "
              f"{st['near_ms']:.0f}, translation plans {st['plans_ms']:.0f}, moments {st['moments_ms']:.0f}), apply "
              f"{apply_ms:.3f} ms (FP32-bulk {apply32_ms}), vs dense {err}, orders {st['level_orders']}, "
              f"{st['m2l_links']} M2L links",
              file=sys.stderr, flush=True)
        op.close()
    return out


def run_flange(edge, frequencies, leaf_capacity, tolerance, repeats, block_width=32):
    """Horn on a 0.6 m square flange (5 mm slab in the mouth plane) through the fast
    multipole form of the request solve (FP64, GMRES restart 50, relative residual
    tolerance, accuracy seed tolerance/10, seed/10 audit): the first request builds
    the cached near structure (cold), the repeats reuse it (warm). Then the same
    request in the FP32-bulk form (precision mixed: FP32 translations in the Krylov
    iterations where the operator allows them, FP64 residuals, audits and corrections),
    warm, with its surface-pressure difference from the plain form."""
    mesh = meshes.exponential_horn(edge, flange=0.6)
    nt, nv = len(mesh.triangles), len(mesh.vertices)
    print(f"# flanged horn: {nt} triangles, {nv} P1 nodes, frequencies {frequencies}", file=sys.stderr, flush=True)
    solver = bem.HipBemSolver(mesh)
    req = bem.BemRequest(
        medium=bem.AcousticMedium(1.205, 343.0), frequencies=list(frequencies),
        excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(1, 1.0)])],
        observers=[(0.3 + 10.0, 0.0, 0.0), (0.3 + 10 * math.cos(math.radians(30)), 10 * math.sin(math.radians(30)), 0.0)],
        numerics=bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=tolerance,
                                 acceleration=bem.ACCELERATION_FAST_MULTIPOLE, fmm_leaf_capacity=leaf_capacity,
                                 fmm_separation_ratio=1.5, fmm_restart_length=50, fmm_maximum_iterations=500,
                                 fmm_block_width=block_width))
    runs = []
    traces = {}
    for r in range(2 + repeats):
        bulk = r == 1 + repeats
        if bulk:
            req.numerics.precision = bem.PRECISION_MIXED
        t0 = time.perf_counter()
        res = solver.solve(req, progress=lambda done, total, f, phase: print(
            f"# flange point {done}/{total} {f:.0f} Hz phase {phase} at {time.perf_counter() - t0:.1f} s",
            file=sys.stderr, flush=True))
        wall = time.perf_counter() - t0
        traces["fp32_bulk" if bulk else "plain"] = [fr.excitations[0].boundary.pressure_trace for fr in res.frequencies]
        points = []
        for fr in res.frequencies:
            ex = fr.excitations[0]
            points.append({"frequency_hz": fr.frequency, "build_ms": round(fr.assembly_milliseconds, 1),
                           "solve_ms": round(fr.solve_milliseconds, 1), "iterations": ex.iterations,
                           "residual": ex.boundary_residual, **fr.fmm,
                           "on_axis_spl_db": round(bem.spl_db(complex(ex.observer_pressure[0])), 3)})
            print(f"# flange {'FP32-bulk' if bulk else 'cold' if r == 0 else 'warm'} {fr.frequency:.0f} Hz: build "
                  f"{fr.assembly_milliseconds:.0f} ms, "
                  f"solve {fr.solve_milliseconds:.0f} ms, {ex.iterations} iterations, residual "
                  f"{ex.boundary_residual:.1e}, audit {fr.fmm['audit_discrepancy']:.1e}, apply "
                  f"{fr.fmm['apply_milliseconds']:.1f} ms, plane-wave levels {fr.fmm['plane_wave_levels']}, endpoint "
                  f"degree {fr.fmm['endpoint_degree']} (comparison {fr.fmm['quadrature_discrepancy']:.1e}, "
                  f"{fr.fmm['promotions']} promotions), direct audit {fr.fmm['direct_audit']}",
                  file=sys.stderr, flush=True)
        runs.append({"wall_s": round(wall, 3), "points": points, "form": "fp32_bulk" if bulk else "plain"})
    bulk_run = runs.pop()
    warm = runs[1:] or runs
    wall = min(run["wall_s"] for run in warm)
    diff = [float(np.linalg.norm(b - p) / np.linalg.norm(p)) for b, p in zip(traces["fp32_bulk"], traces["plain"])]
    solve_plain = sum(p["solve_ms"] for p in runs[-1]["points"])
    solve_bulk = sum(p["solve_ms"] for p in bulk_run["points"])
    print(f"# flange FP32-bulk form (run warm): solves {solve_bulk / 1e3:.2f} s (plain {solve_plain / 1e3:.2f} s), "
          f"wall {bulk_run['wall_s']:.1f} s, surface pressure vs plain {max(diff):.1e}", file=sys.stderr, flush=True)
    return {"triangles": nt, "nodes": nv, "block_width": block_width, "cold_wall_s": runs[0]["wall_s"], "warm_wall_s": wall,
            "points_per_s_warm": len(frequencies) / wall, "runs": runs,
            "fp32_bulk": {"wall_s_warm": bulk_run["wall_s"], "solve_s": solve_bulk / 1e3, "plain_solve_s": solve_plain / 1e3,
                          "rel_l2_surface_pressure_vs_plain": diff, "points": bulk_run["points"]}}


def main_stream(args):
    """The headline: the served stream, plain FP64 then mixed, on the same arrivals."""
    steps = 3 if args.steps is None else args.steps
    warmup = 1 if args.warmup is None else args.warmup
    t0 = time.perf_counter()
    pool = StreamPool(args.pool, args.pool_min, args.pool_max)
    sizes = [e["triangles"] for e in pool.entries]
    print(f"# stream pool: {sizes} triangles ({time.perf_counter() - t0:.1f} s to mesh)", file=sys.stderr, flush=True)
    if args.roofline_leg:
        leg = roofline_leg(pool, bem.PRECISION_MIXED, args.mixed_tolerance)  # first solve: plan + kernel
        leg = roofline_leg(pool, bem.PRECISION_MIXED, args.mixed_tolerance)
        print(json.dumps({"metric": "regular_tiles FP32 TFLOP/s (isolated leg of the served stream)",
                          "unit": "TFLOP/s", "value": leg["achieved_tflops"], "higher_is_better": True, "n_gpus": 1,
                          "dtype": "f32", "data": "synthetic", "config": {"workload": "roofline leg", **leg},
                          "tuning": tuning_summary()}),
              flush=True)
        return
    # Cold: one interaction plan per resident mesh, at its frequency limit, through a
    # throwaway server so no direct-solve workspace stays allocated.
    t0 = time.perf_counter()
    server = bem.HipBemServer(slots=1)
    for n, e in enumerate(pool.entries):
        server.poll(server.submit(e["solver"], pool.request(n, [100.0], bem.PRECISION_FP64, args.tolerance)))
        print(f"# stream pool mesh {n} ({e['triangles']} triangles) planned, {time.perf_counter() - t0:.0f} s",
              file=sys.stderr, flush=True)
    server.close()
    cold_s = time.perf_counter() - t0
    duration = (warmup + steps) * args.window
    draws, redraw_rate = draw_stream(pool, args.rate, duration)
    offered_points = sum(len(d[2]) for d in draws)
    print(f"# stream: {len(draws)} requests / {offered_points} points offered over {duration:.0f} s", file=sys.stderr,
          flush=True)
    forms, results = {}, {}
    for label, precision, tol in (("served_fp64", bem.PRECISION_FP64, args.tolerance),
                                  ("served_mixed", bem.PRECISION_MIXED, args.mixed_tolerance)):
        device_synchronize()
        forms[label], results[label] = serve_stream(pool, draws, precision, tol, args.slots, args.window, warmup,
                                                    steps, label)
        forms[label]["gmres_tolerance"] = tol
    prefix = common_prefix(draws, forms)
    for f in forms.values():
        f.pop("finished_s")
    if prefix:
        print(f"# common prefix: {prefix['requests']} requests / {prefix['points']} points: FP64 "
              f"{prefix['seconds']['served_fp64']:.1f} s, mixed {prefix['seconds']['served_mixed']:.1f} s",
              file=sys.stderr, flush=True)
    errors = compare_forms(pool, draws, results["served_fp64"], results["served_mixed"])
    print(f"# mixed vs FP64 on {errors['points_compared']} points: surface pressure {errors['surface_pressure']:.2e}, "
          f"on-axis SPL {errors['on_axis_spl_db']:.2e} dB, directivity {errors['directivity_db']:.2e} dB",
          file=sys.stderr, flush=True)
    leg = roofline_leg(pool, bem.PRECISION_MIXED, args.mixed_tolerance)
    print(f"# roofline leg: regular_tiles FP32 {leg['achieved_tflops']:.1f} TF/s, {leg['launch_ms']:.1f} ms per launch",
          file=sys.stderr, flush=True)
    head = forms["served_mixed"]
    line = {
        "metric": "freq-points solved/s + p50/p99 latency, served mixed vs plain FP64 @<=1e-4 err",
        "value": head["points_per_s"], "unit": "freq-points/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * head["timed_s"] / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded stream of horn-variant requests, meshes generated in-process)",
        "config": {"workload": f"served stream: open-loop Poisson {args.rate:g} requests/s, {len(sizes)} resident horn "
                               f"meshes {min(sizes)}-{max(sizes)} triangles (log-spread; request mesh uniform over the "
                               f"pool), round(log-uniform[1, 32]) frequencies per request log-uniform in 100 Hz-20 kHz "
                               f"(k h_max <= 2, redraw rate {redraw_rate:.2f}), dense Galerkin assembly + GMRES "
                               f"(restart 50; FP64 form rel tol {args.tolerance:g}, mixed form {args.mixed_tolerance:g}), "
                               f"{args.slots} slots, {warmup} untimed + {steps} timed windows of {args.window:g} s",
                   "pool_triangles": sizes, "rate_per_s": args.rate, "requests_offered": len(draws),
                   "points_offered": offered_points, "window_s": args.window, "slots": args.slots},
        "latency_ms": {"p50": head["p50_ms"], "p99": head["p99_ms"]},
        "forms": forms,
        "mixed_vs_fp64": {"speedup_same_work": (prefix["seconds"]["served_fp64"] / prefix["seconds"]["served_mixed"]
                                                if prefix else None),
                          "same_work": prefix,
                          "speedup_timed_windows": head["points_per_s"] / forms["served_fp64"]["points_per_s"],
                          "max_rel_l2_surface_pressure": errors["surface_pressure"],
                          "max_on_axis_spl_db": errors["on_axis_spl_db"],
                          "max_directivity_db": errors["directivity_db"],
                          "points_compared": errors["points_compared"]},
        "cold": {"plans_s": cold_s},
        "tuning": tuning_summary()["chosen"],
        "roofline": {"kernel": "regular_tiles (FP32 pair quadrature of the mixed form, VALU; isolated leg)",
                     "bound": "valu", "achieved": leg["achieved_tflops"], "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": leg["achieved_tflops"] / FP32_PEAK_TFLOPS,
                     "traffic": LEG_TRAFFIC_BYTES.get(leg["mesh_triangles"]),
                     "launch_ms": leg["launch_ms"], "flops_per_launch": leg["flops_per_launch"],
                     "leg": {"mesh_triangles": leg["mesh_triangles"], "frequencies_hz": leg["frequencies_hz"]},
                     "ordered_pair_tflops": leg["ordered_pair_tflops"],
                     "served_achieved": (evaluated_flops(head["regular_samples_served"]) /
                                         (head["regular_kernel_ms_served"] / 1e3) / 1e12
                                         if head["regular_kernel_ms_served"] > 0 else None)},
    }
    if not args.no_c2:
        mesh = meshes.exponential_horn(args.edge)
        solver = bem.HipBemSolver(mesh)
        frequencies = [200.0 * 40.0 ** (i / (args.frequencies - 1)) for i in range(args.frequencies)]
        c2 = {}
        for name, prec, kind, tol in (("fp64", bem.PRECISION_FP64, bem.SOLVER_GMRES, args.tolerance),
                                      ("mixed", bem.PRECISION_MIXED, bem.SOLVER_GMRES, args.mixed_tolerance)):
            f = run_form(solver, frequencies, prec, kind, tol, 1, 1, f"dense horn sweep {name}")
            c2[name] = {"value": f["value"], "regular_kernel_ms": f["kernel_ms"] / f["points"],
                        "gmres_iterations": [int(min(f["iterations"])), int(max(f["iterations"]))],
                        "traces": f["traces"]}
        err = max(_area_weighted_error(mesh, m, p) for m, p in zip(c2["mixed"].pop("traces"), c2["fp64"].pop("traces")))
        line["forms"]["dense_horn_sweep"] = {"triangles": len(mesh.triangles), "frequencies": len(frequencies), **c2,
                                             "mixed_over_fp64": c2["mixed"]["value"] / c2["fp64"]["value"],
                                             "mixed_vs_fp64_surface_pressure": err}
        del solver
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_stream_baseline(pool, args.cpu_threads)
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps: stream windows (default 3), or sweeps with --horn (default 2)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 1)")
    ap.add_argument("--edge", type=float, default=0.0052)
    ap.add_argument("--frequencies", type=int, default=16)
    ap.add_argument("--solver", choices=["gmres", "lu"], default="gmres")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both",
                    help="both: time the plain FP64 form and the mixed form; value is the served mixed form")
    ap.add_argument("--slots", type=int, default=3, help="serving-loop slots (0: no served runs)")
    ap.add_argument("--open-loop", type=int, default=48, help="open-loop single-point requests for latency")
    ap.add_argument("--tolerance", type=float, default=1e-8)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cross-check", action="store_true")
    ap.add_argument("--no-fmm", action="store_true",
                    help="skip the FMM pressure operator measurement (lowest and highest frequency, same mesh)")
    ap.add_argument("--fmm-only", action="store_true", help="only the FMM operator measurement (one JSON line)")
    ap.add_argument("--fmm-leaf", type=int, default=64)
    ap.add_argument("--fmm-repetitions", type=int, default=20)
    ap.add_argument("--flange", action="store_true",
                    help="only the flanged-horn fast multipole solve (118k triangles, 500 Hz and 16 kHz; one JSON line)")
    ap.add_argument("--flange-edge", type=float, default=0.0045)
    ap.add_argument("--flange-repeats", type=int, default=1)
    ap.add_argument("--flange-block-width", type=int, default=32, help="FMM block-Jacobi width (32, 64, ..., 2048)")
    ap.add_argument("--sweep", action="store_true",
                    help="only the batched horn sweep (seeded variants x frequencies x 4 right-hand sides; one JSON line)")
    ap.add_argument("--sweep-variants", type=int, default=4)
    ap.add_argument("--sweep-frequencies", type=int, default=8)
    ap.add_argument("--fixture", default="", metavar="JSON",
                    help="only a benchmark fixture file (the reference's solver_benchmark input format; one JSON line)")
    ap.add_argument("--lu", type=int, default=0, metavar="N",
                    help="only the dense complex LU solve of a seeded random N x N system (one JSON line)")
    ap.add_argument("--horn", action="store_true",
                    help="only the dense horn sweep (20.7k triangles, 16 frequencies; one JSON line)")
    ap.add_argument("--rate", type=float, default=20.0, help="served stream: Poisson arrivals per second")
    ap.add_argument("--window", type=float, default=30.0, help="served stream: seconds per step (window)")
    ap.add_argument("--pool", type=int, default=8, help="served stream: resident horn meshes")
    ap.add_argument("--pool-min", type=float, default=2000.0)
    ap.add_argument("--pool-max", type=float, default=120000.0)
    ap.add_argument("--mixed-tolerance", type=float, default=1e-6,
                    help="served stream: GMRES relative residual of the mixed form (FP64 form: --tolerance)")
    ap.add_argument("--no-c2", action="store_true", help="served stream: skip the dense horn sweep forms entry")
    ap.add_argument("--roofline-leg", action="store_true",
                    help="only the isolated roofline leg of the served stream (for kernel profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: every GPU box here has one MI355X and the solver has no multi-GPU path "
              f"(requested --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    if args.flange:
        fl = run_flange(args.flange_edge, [500.0, 16000.0], args.fmm_leaf, 1e-6, args.flange_repeats,
                        args.flange_block_width)
        print(json.dumps({"metric": "FMM flanged-horn freq-points solved/s (warm)", "unit": "freq-points/s",
                          "value": fl["points_per_s_warm"], "higher_is_better": True, "n_gpus": 1, "dtype": "f64",
                          "data": "synthetic (parametric flanged horn mesh generated in-process)",
                          "config": {"workload": f"exponential horn on a 0.6 m square flange, {fl['triangles']} "
                                                 f"triangles / {fl['nodes']} P1 nodes, 500 Hz and 16 kHz, fast "
                                                 f"multipole form (leaf {args.fmm_leaf}, separation 1.5, GMRES restart "
                                                 f"50, rel tol 1e-6, accuracy seed 1e-7, block-Jacobi width {args.flange_block_width})"},
                          "flange": fl, "tuning": tuning_summary()}), flush=True)
        return
    if args.fixture:
        mesh, req, digest = bem.read_benchmark_fixture(args.fixture)
        req.numerics = bem.BemNumerics(precision=bem.PRECISION_MIXED if args.precision == "mixed" else
                                       bem.PRECISION_FP64,
                                       solver=bem.SOLVER_GMRES if args.precision == "mixed" else bem.SOLVER_LU)
        solver = bem.HipBemSolver(mesh)
        walls = []
        for rep in range(2):  # cold (plans) and warm
            t0 = time.perf_counter()
            res = solver.solve(req)
            walls.append(time.perf_counter() - t0)
        print(json.dumps({"metric": "benchmark fixture freq-points solved/s (warm)", "unit": "freq-points/s",
                          "value": len(req.frequencies) / walls[1], "higher_is_better": True, "n_gpus": 1,
                          "dtype": "f32+f64" if args.precision == "mixed" else "f64", "data": "fixture",
                          "config": {"workload": f"fixture {os.path.basename(args.fixture)} (sha256 {digest[:16]}), "
                                                 f"{len(mesh.triangles)} triangles, {len(req.frequencies)} frequencies, "
                                                 f"{len(req.excitations)} excitations"},
                          "cold_s": walls[0], "warm_s": walls[1],
                          "residual": max(e.boundary_residual for f in res.frequencies for e in f.excitations)}),
              flush=True)
        return
    if args.lu:
        n = args.lu
        rng = np.random.default_rng(1)
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        b = rng.standard_normal((n, 1)) + 1j * rng.standard_normal((n, 1))
        walls = []
        for rep in range(2):
            t0 = time.perf_counter()
            x = bem.dense_lu_solve(A, b)
            walls.append(time.perf_counter() - t0)
            print(f"# dense LU n={n} run {rep}: {walls[-1]:.3f} s including the host transfers", file=sys.stderr,
                  flush=True)
        residual = float(np.linalg.norm(A @ x - b) / np.linalg.norm(b))
        print(json.dumps({"metric": "dense complex LU solve wall time (incl. host transfers)", "unit": "s",
                          "value": min(walls), "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
                          "data": "synthetic (seeded random complex matrix)",
                          "config": {"workload": f"dense complex FP64 LU with partial pivoting, n = {n}, 1 RHS"},
                          "flops": 8.0 / 3.0 * n ** 3, "relative_residual": residual}), flush=True)
        return
    if args.sweep:
        sw = run_sweep(args.sweep_variants, args.sweep_frequencies, max(args.slots, 1), args.tolerance)
        sizes = [v["triangles"] for v in sw["variants"]]
        print(json.dumps({"metric": "batched-sweep freq-points solved/s (4 right-hand sides per point), served mixed "
                                    "vs plain FP64", "unit": "freq-points/s", "value": sw["mixed"]["points_per_s"],
                          "higher_is_better": True, "n_gpus": 1, "dtype": "f32+f64",
                          "data": "synthetic (seeded horn variants generated in-process)",
                          "config": {"workload": f"first {len(sizes)} of 32 seeded horn variants ({min(sizes)}-"
                                                 f"{max(sizes)} triangles), {len(sw['frequencies'])} of 64 log-spaced "
                                                 f"frequencies 100 Hz-20 kHz, 4 throat right-hand sides, dense "
                                                 f"assembly + GMRES (restart 50, rel tol {args.tolerance:g})"},
                          "sweep": sw}), flush=True)
        return
    if not args.horn:
        main_stream(args)
        return
    if args.steps is None:
        args.steps = 2
    if args.warmup is None:
        args.warmup = 1
    frequencies = [200.0 * 40.0 ** (i / (args.frequencies - 1)) for i in range(args.frequencies)]
    mesh = meshes.exponential_horn(args.edge)
    nt, nv = len(mesh.triangles), len(mesh.vertices)
    solver_kind = bem.SOLVER_GMRES if args.solver == "gmres" else bem.SOLVER_LU
    print(f"# horn: {nt} triangles, {nv} P1 nodes, {len(frequencies)} frequencies", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    solver = bem.HipBemSolver(mesh)
    construct_s = time.perf_counter() - t0
    if args.fmm_only:
        fmm = run_fmm(solver, [frequencies[0], frequencies[-1]], args.fmm_leaf, args.fmm_repetitions,
                      check=not args.no_cross_check)
        print(json.dumps({"metric": "FMM pressure operator apply time", "unit": "ms", "value": fmm[-1]["apply_ms"],
                          "higher_is_better": False, "n_gpus": 1, "dtype": "f64", "data": "synthetic",
                          "config": {"workload": f"exponential horn {nt} triangles / {nv} P1 nodes, leaf "
                                                 f"{args.fmm_leaf}, accuracy seed 5e-7"}, "fmm": fmm,
                          "tuning": tuning_summary()}), flush=True)
        return
    cross, cold_plan_ms = None, None
    if not args.no_cross_check:
        # Dense FP64 LU cross-check of the FP64 GMRES solve at the highest frequency (untimed).
        r_lu = solver.solve(horn_request([frequencies[-1]], bem.PRECISION_FP64, bem.SOLVER_LU, 1e-10))
        cold_plan_ms = r_lu.frequencies[0].interaction_plan_milliseconds  # the mesh's plan is built here
        r_gm = solver.solve(horn_request([frequencies[-1]], bem.PRECISION_FP64, solver_kind, args.tolerance))
        a = r_lu.frequencies[0].excitations[0].boundary.surface_pressure
        b = r_gm.frequencies[0].excitations[0].boundary.surface_pressure
        cross = float(np.linalg.norm(a - b) / np.linalg.norm(a))
        print(f"# LU cross-check at {frequencies[-1]:.0f} Hz: relative L2 {cross:.2e}, LU factor "
              f"{r_lu.frequencies[0].factorization_milliseconds:.0f} ms", file=sys.stderr, flush=True)

    forms = {}
    if args.precision in ("both", "fp64"):
        forms["fp64"] = run_form(solver, frequencies, bem.PRECISION_FP64, solver_kind, args.tolerance, args.steps,
                                 args.warmup, "plain FP64")
    if args.precision in ("both", "mixed"):
        forms["mixed"] = run_form(solver, frequencies, bem.PRECISION_MIXED, bem.SOLVER_GMRES, args.tolerance,
                                  args.steps, args.warmup, "mixed")
    if args.slots:
        for name, prec in (("served_fp64", bem.PRECISION_FP64), ("served_mixed", bem.PRECISION_MIXED)):
            if name.split("_")[1] in forms:
                forms[name] = run_served(solver, frequencies, prec, args.tolerance, args.steps, args.slots,
                                         name.replace("_", " "), args.open_loop)
    head_name = next(n for n in ("served_mixed", "mixed", "served_fp64", "fp64") if n in forms)
    head = forms[head_name]
    mixed_error = None
    if head_name in ("mixed", "served_mixed") and "fp64" in forms:
        # The plain FP64 GPU form equals the FP64 oracle to ~1e-15 (parity tests), so
        # this is the mixed form's surface-pressure error against the FP64 oracle.
        mixed_error = max(float(np.linalg.norm(m - p) / np.linalg.norm(p))
                          for m, p in zip(head["traces"], forms["fp64"]["traces"]))
    mixed_head = head_name in ("mixed", "served_mixed")
    peak = FP32_PEAK_TFLOPS if mixed_head else FP64_PEAK_TFLOPS
    # Kernel roofline from the direct run of the same precision: in the served runs,
    # kernels of other slots overlap and inflate each launch's event time.
    roof = forms.get("mixed" if mixed_head else "fp64", head)
    achieved = evaluated_flops(roof["samples"]) / (roof["kernel_ms"] / 1e3) / 1e12 if roof["kernel_ms"] > 0 else None
    line = {
        "metric": "freq-points solved/s + p50/p99 latency, served mixed vs plain FP64 @<=1e-4 err",
        "value": head["value"], "unit": "freq-points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * head["elapsed"] / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if mixed_head else "f64",
        "data": "synthetic (parametric horn mesh generated in-process)",
        "config": {"workload": "dense horn sweep: exponential horn r_t=12.7 mm, r_m=150 mm, L=300 mm, 5 mm wall, "
                               f"{nt} triangles / {nv} P1 nodes, {len(frequencies)} log-spaced frequencies "
                               f"200 Hz-8 kHz, dense assembly + {args.solver.upper()} "
                               f"(restart 50, rel tol {args.tolerance:g}), 7 far-field directions",
                   "triangles": nt, "nodes": nv, "frequencies": len(frequencies), "solver": args.solver,
                   "form": head_name},
        "latency_ms": {"p50": head.get("p50"), "p99": head.get("p99")},
        "forms": {name: {"value": f["value"], "ms_per_step": 1e3 * f["elapsed"] / args.steps,
                         "latency_ms": {"p50": f.get("p50"), "p99": f.get("p99")},
                         "regular_kernel_ms": f["kernel_ms"] / f["points"],
                         "gmres_iterations": {"min": int(min(f["iterations"])), "max": int(max(f["iterations"]))},
                         **({"open_loop": f["open_loop"]} if "open_loop" in f else {}),
                         **({"server": {k: f["stats"][k] for k in ("max_running", "peak_held_bytes", "budget_bytes")}}
                            if "stats" in f else {})}
                  for name, f in forms.items()},
        "mixed_vs_fp64": ({"form": head_name, "speedup": head["value"] / forms["fp64"]["value"],
                           "max_rel_l2_surface_pressure": mixed_error} if mixed_error is not None else None),
        "cold": {"mesh_setup_s": construct_s,
                 "interaction_plan_ms": cold_plan_ms if cold_plan_ms is not None else
                 next((f["plan_ms"] for f in forms.values() if f["plan_ms"] is not None), None)},
        "lu_cross_check_rel_l2": cross,
        "roofline": {"kernel": f"regular_tiles ({'FP32' if mixed_head else 'FP64'} pair quadrature, VALU)",
                     "bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak if achieved else None,
                     "traffic": None,
                     "launch_ms": roof["kernel_ms"] / roof["points"],
                     "flops_per_launch": evaluated_flops(roof["samples"]) / roof["points"]},
    }
    if not args.no_fmm:
        try:
            line["fmm"] = run_fmm(solver, [frequencies[0], frequencies[-1]], args.fmm_leaf, args.fmm_repetitions,
                                  check=not args.no_cross_check)
        except Exception as exc:  # reported, never fatal for the headline line
            line["fmm"] = {"error": f"{type(exc).__name__}: {exc}"}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(mesh, frequencies, args.cpu_threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
