"""Serving loop on the device: continuous batching of frequency points from
several requests over concurrent slots (hsbem_submit / hsbem_poll).

Each served result must equal the same request solved alone. Tolerances: the
assembly sums with device atomics, so results are reproducible to rounding
only: plain FP64 points <= 1e-10 relative L2 (surface pressure), mixed points
(FP32 matrix) <= 2e-5.
"""
import math

import numpy as np
import pytest

from fixtures import BEAT_FREQUENCIES, sphere_request
from hornsculpt_amd import bem, meshes

pytestmark = pytest.mark.gpu


def _horn_request(frequencies, numerics):
    return bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=list(frequencies),
                          excitations=[bem.BemExcitation("throat", [bem.FaceVelocity(1, 1.0)])],
                          observers=[(1.3, 0.0, 0.0), (0.3, 1.0, 0.0)],
                          far_field=bem.FarFieldObservation((0.3, 0.0, 0.0), 1.0), numerics=numerics)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _compare(alone, served, tol):
    assert len(alone.frequencies) == len(served.frequencies)
    worst = 0.0
    for fa, fs in zip(alone.frequencies, served.frequencies):
        assert fa.frequency == fs.frequency
        assert fs.served_milliseconds > 0
        for ea, es in zip(fa.excitations, fs.excitations):
            worst = max(worst, _rel(es.boundary.surface_pressure, ea.boundary.surface_pressure))
            if len(ea.observer_pressure):
                worst = max(worst, _rel(es.observer_pressure, ea.observer_pressure))
    assert worst <= tol, worst
    return worst


@pytest.fixture(scope="module")
def solvers():
    return {"ico": bem.HipBemSolver(meshes.icosphere(3)), "octa": bem.HipBemSolver(meshes.octa_sphere(3)),
            "horn": bem.HipBemSolver(meshes.exponential_horn(0.02))}


def _workload():
    gm = bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-9)
    return [
        ("ico", sphere_request([500.0, 1000.0, 1500.0], observers=[(0, 0, 1), (1, 0, 0)]), 1e-10),
        ("octa", sphere_request(BEAT_FREQUENCIES), 1e-10),
        ("horn", _horn_request([1000.0, 3000.0], bem.BemNumerics(precision=bem.PRECISION_MIXED)), 2e-5),
        ("ico", sphere_request([700.0], numerics=gm), 1e-10),
        ("horn", _horn_request([2000.0], gm), 1e-10),
    ]


def test_served_results_equal_alone_results(solvers):
    work = _workload()
    alone = [solvers[m].solve(r) for m, r, _ in work]
    server = bem.HipBemServer(slots=3)
    tickets = [server.submit(solvers[m], r) for m, r, _ in work]
    served = [server.poll(t) for t in tickets]
    stats = server.stats()
    server.close()
    for (m, r, tol), a, s in zip(work, alone, served):
        worst = _compare(a, s, tol)
        print(f"{m} {r.frequencies}: served vs alone {worst:.2e}")
    assert stats["points"] == sum(len(r.frequencies) for _, r, _ in work)
    assert stats["completed"] == len(work) and stats["failed"] == 0
    assert stats["max_running"] >= 2  # points of different requests were on the device together


def test_admission_by_memory_footprint(solvers):
    horn = solvers["horn"]
    req = _horn_request([1500.0, 2500.0], bem.BemNumerics(solver=bem.SOLVER_GMRES))
    probe = bem.HipBemServer(slots=1)
    footprint = probe.poll(probe.submit(horn, req)).peak_device_bytes
    probe.close()
    assert footprint > 0
    # Room for one point at a time: three slots, but the budget admits one footprint.
    server = bem.HipBemServer(slots=3, memory_budget_bytes=int(1.5 * footprint))
    tickets = [server.submit(horn, req) for _ in range(3)]
    results = [server.poll(t) for t in tickets]
    stats = server.stats()
    assert stats["max_running"] == 1
    assert stats["peak_held_bytes"] <= int(1.5 * footprint)
    for r in results[1:]:
        _compare(results[0], r, 1e-10)
    # A point that can never fit is refused at submission.
    small = bem.HipBemServer(slots=2, memory_budget_bytes=footprint // 2)
    with pytest.raises(ValueError):
        small.submit(horn, req)
    small.close()
    server.close()


def test_invalid_request_refused_at_submission(solvers):
    server = bem.HipBemServer(slots=2)
    bad = sphere_request([1000.0])
    bad.numerics.quadrature_order = 3
    with pytest.raises(ValueError):
        server.submit(solvers["ico"], bad)
    server.close()


def test_cancel_and_poll_pending(solvers):
    server = bem.HipBemServer(slots=1)
    freqs = [200.0 * 1.2 ** i for i in range(12)]
    t = server.submit(solvers["horn"], _horn_request(freqs, bem.BemNumerics()))
    assert server.poll(t, 0) is None  # twelve points cannot be done yet
    server.cancel(t)
    with pytest.raises(RuntimeError, match="cancelled"):
        server.poll(t)
    # the slot keeps serving after a cancellation
    ok = server.poll(server.submit(solvers["ico"], sphere_request([1000.0])))
    assert ok.frequencies[0].excitations[0].boundary_residual <= 5e-5
    server.close()


def test_progressive_sweep_equals_sweep(solvers):
    """The realtime order (frequency_schedule, simulation_dock.cpp:2568-2576): the
    sweep's points as single-frequency requests submitted in schedule order share the
    sweep's planning raster, so each point equals the same point of the sweep solved
    at once."""
    horn = solvers["horn"]
    freqs = [300.0 * 1.6 ** i for i in range(6)]
    req = _horn_request(freqs, bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-9))
    req.interaction_plan_frequencies = list(freqs)
    whole = horn.solve(req)
    server = bem.HipBemServer(slots=1)
    tickets = [(i, server.submit(horn, r)) for i, r in bem.progressive_requests(req)]
    done = [(i, server.poll(t)) for i, t in tickets]
    server.close()
    worst = 0.0
    for i, res in done:
        assert res.frequencies[0].frequency == freqs[i]
        got = res.frequencies[0].excitations[0].boundary.pressure_trace
        want = whole.frequencies[i].excitations[0].boundary.pressure_trace
        worst = max(worst, _rel(got, want))
    print(f"progressive order {[i for i, _ in done]}: worst vs sweep {worst:.1e}")
    assert [i for i, _ in done] == [5, 0, 2, 3, 4, 1]
    assert worst <= 1e-10


def test_fmm_and_dense_lu_under_one_point_budget():
    """A fast-multipole request and a dense LU request on the same mesh, served under
    a budget that holds the resident mesh state and one point at a time: the slot
    that served the FMM point must not reuse its workspace for the LU point (its A
    and LU factors are larger), held + resident bytes stay within the budget, and
    both results equal the requests solved alone."""
    mesh = meshes.exponential_horn(0.012)
    driven = sorted(set(int(f) for f, b in zip(mesh.face_ids, mesh.boundary) if b == bem.DRIVEN))

    def request(numerics):
        return bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=[800.0],
                              excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(f, 1.0) for f in driven])],
                              numerics=numerics)

    fmm = request(bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-6,
                                  acceleration=bem.ACCELERATION_FAST_MULTIPOLE))
    lu = request(bem.BemNumerics(relative_residual_tolerance=1e-6))
    solver = bem.HipBemSolver(mesh)
    alone = [solver.solve(fmm), solver.solve(lu)]
    probe = bem.HipBemServer(slots=1)  # the served footprint of each point (FMM: as measured by the solves)
    footprints = [probe.poll(probe.submit(solver, r)).peak_device_bytes for r in (fmm, lu)]
    probe.close()
    resident = solver.cache_info()["resident_bytes"]
    budget = resident + max(footprints) + min(footprints) // 2
    server = bem.HipBemServer(slots=2, memory_budget_bytes=budget)
    tickets = [server.submit(solver, fmm), server.submit(solver, lu), server.submit(solver, fmm)]
    served = [server.poll(t) for t in tickets]
    stats = server.stats()
    server.close()
    print(f"budget {budget}, resident {stats['resident_bytes']}, peak held {stats['peak_held_bytes']}, "
          f"peak total {stats['peak_total_bytes']}, footprints {footprints}")
    assert stats["failed"] == 0 and stats["max_running"] == 1
    assert stats["peak_total_bytes"] <= budget
    _compare(alone[0], served[0], 1e-8)
    _compare(alone[1], served[1], 1e-10)
    _compare(alone[0], served[2], 1e-8)


def test_plan_cache_stays_bounded_over_many_rasters(solvers):
    """Each distinct planning raster builds an interaction plan; the mesh keeps at
    most four that no solve holds, so serving many rasters keeps the plan count and
    the resident bytes bounded."""
    ico = solvers["ico"]
    server = bem.HipBemServer(slots=2)
    counts, resident = [], []
    for i in range(7):
        req = sphere_request([400.0 + 10.0 * i])
        req.interaction_plan_frequencies = [400.0 + 10.0 * i, 1500.0 + 100.0 * i]
        server.poll(server.submit(ico, req))
        info = ico.cache_info()
        counts.append(info["cached_plans"])
        resident.append(info["resident_bytes"])
    stats = server.stats()
    server.close()
    print("cached plans", counts, "resident bytes", resident)
    assert max(counts) <= 4 and counts[-1] == 4
    assert resident[-1] <= max(resident[3:5]) * 1.05  # no growth once the cache is full
    assert stats["peak_total_bytes"] <= stats["budget_bytes"]
