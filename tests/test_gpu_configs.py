"""Parity at the real sizes of the benchmark configurations (dense horn sweep,
flanged horn on the fast multipole path, full-size batched-sweep variant).

Dense horn sweep (20 742 triangles / 10 373 P1 nodes, planning raster up to
8 kHz): the device's FP64 rows of A and B equal the oracle's rows to 1e-12
(relative max-norm over the sampled rows; the oracle integrates only the test
panels that touch them), and the plain FP64 GMRES(50, 1e-8) solution solves
the oracle's equations on those rows:
||Ao_rows p - Bo_rows q|| / ||Bo_rows q|| <= 2e-7 (GMRES stops at a mass-weighted
relative residual of 1e-8 over all 10 373 rows; a 48-row sample may hold up to
sqrt(10373 / 48) ~ 15 times that share), and agrees with the device LU solve of
the same point to 1e-6 (the Burton-Miller system is well conditioned).
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from fixtures import flux
from hornsculpt_amd import bem, meshes

pytestmark = pytest.mark.gpu

RHO, C = 1.205, 343.0


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.fixture(scope="module")
def horn20k():
    mesh = meshes.exponential_horn(0.0052)
    assert len(mesh.triangles) == 20742 and len(mesh.vertices) == 10373
    kmax = 2 * math.pi * 8000.0 / C
    num = bem.BemNumerics()
    plan = bem.interaction_plan(mesh, num, kmax)
    return mesh, num, kmax, plan, bem.HipBemSolver(mesh)


def _horn_request(frequency, solver, tolerance):
    return bem.BemRequest(
        medium=bem.AcousticMedium(RHO, C), frequencies=[frequency],
        excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(1, 1.0)])],
        interaction_plan_frequencies=sorted({frequency, 8000.0}),
        numerics=bem.BemNumerics(solver=solver, relative_residual_tolerance=tolerance, gmres_restart=50,
                                 gmres_maximum_iterations=2000))


@pytest.mark.parametrize("frequency", [200.0, 8000.0])
def test_dense_horn_20k_rows_and_solution_match_oracle(horn20k, frequency):
    mesh, num, kmax, plan, solver = horn20k
    k = 2 * math.pi * frequency / C
    nodes = np.sort(np.random.default_rng(int(frequency)).choice(len(mesh.vertices), 48, replace=False))
    print(f"dense horn 20k, {frequency} Hz: device assembly", flush=True)
    A, B = solver.assemble(num, kmax, k)
    A, B = A[nodes], B[nodes]
    print("oracle rows", flush=True)
    Ao, Bo = ob.oracle_assemble_rows(mesh, plan, k, nodes, threads=16)
    ea, eb = _rel(A, Ao), _rel(B, Bo)
    print(f"rows: A {ea:.2e}  B {eb:.2e}", flush=True)
    assert ea <= 1e-12 and eb <= 1e-12
    del A, B

    gm = solver.solve(_horn_request(frequency, bem.SOLVER_GMRES, 1e-8)).frequencies[0].excitations[0]
    lu = solver.solve(_horn_request(frequency, bem.SOLVER_LU, 1e-12)).frequencies[0].excitations[0]
    p = gm.boundary.pressure_trace
    q = flux(mesh, {1: 1.0}, frequency, RHO)
    b = Bo @ q
    row_residual = np.linalg.norm(Ao @ p - b) / np.linalg.norm(b)
    lu_gap = np.linalg.norm(p - lu.boundary.pressure_trace) / np.linalg.norm(lu.boundary.pressure_trace)
    print(f"GMRES residual {gm.boundary_residual:.2e} ({gm.iterations} its), oracle-row residual "
          f"{row_residual:.2e}, GMRES vs LU {lu_gap:.2e}", flush=True)
    assert gm.boundary_residual <= 1e-8
    assert row_residual <= 2e-7
    assert lu_gap <= 1e-6


# Flanged horn (fast multipole form): 118 424 triangles / 59 214 P1 nodes, 500 Hz
# and 16 kHz. The operator's action on a random vector equals the oracle's rows
# of the row-scaled A on 24 sampled nodes to 2e-6 (relative L2 over the sampled
# rows; accuracy seed 5e-7), and the fast multipole request (GMRES restart 50,
# relative residual 1e-6, seed 1e-7, audits) meets its residual.
@pytest.fixture(scope="module")
def flange118k():
    mesh = meshes.exponential_horn(0.0045, flange=0.6)
    assert len(mesh.triangles) == 118424 and len(mesh.vertices) == 59214
    kmax = 2 * math.pi * 16000.0 / C
    num = bem.BemNumerics()
    print("flanged horn: interaction plan at 16 kHz", flush=True)
    plan = bem.interaction_plan(mesh, num, kmax)
    return mesh, num, kmax, plan, bem.HipBemSolver(mesh)


@pytest.mark.parametrize("frequency", [500.0, 16000.0])
def test_flanged_horn_118k_fmm_rows_match_oracle(flange118k, frequency):
    mesh, num, kmax, plan, solver = flange118k
    k = 2 * math.pi * frequency / C
    nodes = np.sort(np.random.default_rng(int(frequency) + 1).choice(len(mesh.vertices), 24, replace=False))
    print(f"flanged horn {frequency} Hz: FMM operator", flush=True)
    op = solver.fmm_operator(num, kmax, k, leaf_capacity=64)
    x = np.random.default_rng(5).standard_normal(len(mesh.vertices)) * (1 + 0.5j)
    y = op.apply(x)
    st = op.statistics()
    op.close()
    print("oracle rows", flush=True)
    Ao, _ = ob.oracle_assemble_rows(mesh, plan, k, nodes, threads=16)
    yo = Ao @ x
    err = np.linalg.norm(y[nodes] - yo) / np.linalg.norm(yo)
    print(f"FMM apply rows vs oracle {err:.2e} (levels {st['levels']}, m2l links {st['m2l_links']})", flush=True)
    assert err <= 2e-6


def test_flanged_horn_118k_fmm_request_residual(flange118k):
    mesh, _, _, _, solver = flange118k
    req = bem.BemRequest(
        medium=bem.AcousticMedium(RHO, C), frequencies=[500.0, 16000.0],
        excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(1, 1.0)])],
        observers=[(10.3, 0.0, 0.0)],
        numerics=bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-6,
                                 acceleration=bem.ACCELERATION_FAST_MULTIPOLE, fmm_leaf_capacity=64,
                                 fmm_separation_ratio=1.5, fmm_restart_length=50, fmm_maximum_iterations=500))
    res = solver.solve(req, progress=lambda done, total, f, phase: print(
        f"flange point {done}/{total} {f:.0f} Hz phase {phase}", flush=True))
    for fr in res.frequencies:
        ex = fr.excitations[0]
        print(f"{fr.frequency:.0f} Hz: residual {ex.boundary_residual:.1e}, {ex.iterations} iterations, "
              f"audit {fr.fmm['audit_discrepancy']:.1e}, endpoint comparison {fr.fmm['quadrature_discrepancy']:.1e}",
              flush=True)
        assert ex.boundary_residual <= 1e-6
        assert fr.fmm["quadrature_discrepancy"] <= 0.025 * 1e-6
        assert fr.fmm["direct_audit"] or fr.fmm["audit_discrepancy"] <= 0.25e-6


# Batched sweep at a full-size seeded variant: variant 26 of the 32 (21 592
# triangles, longest edge 4.57 mm, k h_max = 1.67 at 20 kHz), four throat
# right-hand sides, two frequencies of the 64-point raster (which is also the
# planning raster), plain FP64 and mixed requests served together. The mixed form
# is within 1e-4 of FP64 (area-weighted surface pressure) on every right-hand
# side; the FP64 rocking modes obey the horn's quarter-turn symmetry to the
# quadrature tolerance (as the reduced-size sweep test states), are odd under the
# half turn and radiate nothing on the axis.
R90 = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, 1.0, 0.0]])


def _rotation_map(vertices, rotation):
    key = {tuple(np.round(v, 9)): i for i, v in enumerate(vertices)}
    perm = np.array([key.get(tuple(np.round(w, 9)), -1) for w in vertices @ rotation.T])
    assert (perm >= 0).all()
    return perm


def _area_weighted(mesh, a, b):
    V, T = mesh.vertices, mesh.triangles
    area = 0.5 * np.linalg.norm(np.cross(V[T[:, 1]] - V[T[:, 0]], V[T[:, 2]] - V[T[:, 0]]), axis=1)
    return math.sqrt(np.sum(area * np.abs(a - b) ** 2) / np.sum(area * np.abs(b) ** 2))


def test_sweep_full_size_variant_served_mixed_and_fp64():
    from hornsculpt_amd import variants

    v = variants.horn_variants(32)[26]
    mesh, excitations = variants.throat_excitations(v.mesh())
    assert len(mesh.triangles) == 21592
    raster = variants.sweep_frequencies(64)
    freqs = [raster[30], raster[63]]
    directions = [(1.0, 0.0, 0.0), (math.cos(0.5), math.sin(0.5), 0.0)]

    def request(precision):
        return bem.BemRequest(medium=bem.AcousticMedium(RHO, C), frequencies=freqs, excitations=excitations,
                              observers=directions, far_field=bem.FarFieldObservation((0.0, 0.0, 0.0), 1.0),
                              interaction_plan_frequencies=raster,
                              numerics=bem.BemNumerics(precision=precision, solver=bem.SOLVER_GMRES,
                                                       relative_residual_tolerance=1e-8, gmres_restart=50,
                                                       gmres_maximum_iterations=2000))

    solver = bem.HipBemSolver(mesh)
    print("variant 26: cold plan", flush=True)
    solver.solve(bem.BemRequest(medium=bem.AcousticMedium(RHO, C), frequencies=[100.0], excitations=excitations[:1],
                                interaction_plan_frequencies=raster,
                                numerics=bem.BemNumerics(solver=bem.SOLVER_GMRES)))
    server = bem.HipBemServer(slots=3)
    tickets = [server.submit(solver, request(p)) for p in (bem.PRECISION_FP64, bem.PRECISION_MIXED)]
    exact, mixed = (server.poll(t) for t in tickets)
    stats = server.stats()
    server.close()
    assert stats["failed"] == 0
    perm, half = _rotation_map(mesh.vertices, R90), _rotation_map(mesh.vertices, R90 @ R90)
    for pe, pm in zip(exact.frequencies, mixed.frequencies):
        for ee, em in zip(pe.excitations, pm.excitations):
            err = _area_weighted(mesh, em.boundary.surface_pressure, ee.boundary.surface_pressure)
            print(f"{pe.frequency:.0f} Hz {ee.name}: mixed vs FP64 {err:.1e}, residual {ee.boundary_residual:.1e}",
                  flush=True)
            assert err <= 1e-4 and ee.boundary_residual <= 1e-8
        piston, cos_m, sin_m, _ = pe.excitations
        p_cos, p_sin = cos_m.boundary.pressure_trace, sin_m.boundary.pressure_trace
        rot = np.linalg.norm(p_sin[perm] - p_cos) / np.linalg.norm(p_cos)
        odd = np.linalg.norm(p_cos[half] + p_cos) / np.linalg.norm(p_cos)
        axis = max(abs(cos_m.observer_pressure[0]), abs(sin_m.observer_pressure[0])) / abs(piston.observer_pressure[0])
        print(f"{pe.frequency:.0f} Hz: rotation {rot:.1e}, half turn {odd:.1e}, on-axis rocking / piston {axis:.1e}",
              flush=True)
        assert rot <= 5e-5 and odd <= 5e-5 and axis <= 1e-5
