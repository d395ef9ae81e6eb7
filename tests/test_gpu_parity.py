"""Device parity: the HIP path against the FP64 oracle and the reference goldens.

Every test drives the product through the C ABI (libhsbem.so); the oracle
(oracle/liboracle.so) is the checker. Tolerances:
  * assembled matrix entries: relative max-norm error <= 1e-12 (FP64 arithmetic,
    different summation order);
  * plain FP64 solution vs the oracle's FP64 solution on the same plan: relative
    L2 <= 1e-8 (the north-star parity bar);
  * BEAT fixtures: observer L2 <= 2e-6 (oracle itself reaches 7.4e-7), tetra
    ports <= 1e-8 at singular order 12 (reference gate 5e-5);
  * analytic pulsating sphere: the reference's mesh-dependent limits.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from fixtures import (BEAT_FREQUENCIES, SPHERE_OBSERVERS, TETRA_EXCITATIONS, analytic, beat_sphere, beat_tetra,
                      flux, oracle_observer_field, sphere_request)
from hornsculpt_amd import bem, meshes

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("level,frequency", [(2, 1000.0), (3, 1715.0)])
def test_assembly_matches_oracle_octa_sphere(level, frequency):
    mesh = meshes.octa_sphere(level)
    k = 2 * math.pi * frequency / 343.0
    num = bem.BemNumerics()
    solver = bem.HipBemSolver(mesh)
    A, B = solver.assemble(num, k, k)
    plan = bem.interaction_plan(mesh, num, k)
    Ao, Bo = ob.oracle_assemble(mesh, plan, k)
    print(f"octa{level} {frequency} Hz: A {_rel(A, Ao):.2e}  B {_rel(B, Bo):.2e}")
    assert _rel(A, Ao) <= 1e-12
    assert _rel(B, Bo) <= 1e-12


def test_assembly_matches_oracle_lower_frequency_than_plan():
    # Regular degrees are re-qualified at the solve k (cuda_solver.cu:340-351).
    mesh = meshes.icosphere(2)
    kmax, k = 2 * math.pi * 2500 / 343.0, 2 * math.pi * 300 / 343.0
    num = bem.BemNumerics()
    A, B = bem.HipBemSolver(mesh).assemble(num, kmax, k)
    Ao, Bo = ob.oracle_assemble(mesh, bem.interaction_plan(mesh, num, kmax), k)
    assert _rel(A, Ao) <= 1e-12 and _rel(B, Bo) <= 1e-12


def test_assembly_matches_oracle_tetra_singular12():
    mesh = meshes.scalene_tetra()
    k = 2 * math.pi * 400 / 343.0
    num = bem.BemNumerics(singular_order=12)
    A, B = bem.HipBemSolver(mesh).assemble(num, k, k)
    Ao, Bo = ob.oracle_assemble(mesh, bem.interaction_plan(mesh, num, k), k, singular_order=12)
    assert _rel(A, Ao) <= 1e-12 and _rel(B, Bo) <= 1e-12


@pytest.mark.parametrize("frequency", [2000.0, 500.0])  # 500 Hz: most cluster tiles take the uniform far path
def test_assembly_matches_oracle_horn_rows(frequency):
    mesh = meshes.exponential_horn(0.02)
    k = 2 * math.pi * frequency / 343.0
    num = bem.BemNumerics()
    A, B = bem.HipBemSolver(mesh).assemble(num, k, k)
    plan = bem.interaction_plan(mesh, num, k)
    rng = np.random.default_rng(7)
    nodes = rng.choice(len(mesh.vertices), 40, replace=False)
    tests = np.unique([t for t, tri in enumerate(mesh.triangles) if set(tri) & set(nodes.tolist())])
    Ao, Bo = ob.oracle_assemble(mesh, plan, k, tests=tests)
    assert _rel(A[nodes], Ao[nodes]) <= 1e-12
    assert _rel(B[nodes], Bo[nodes]) <= 1e-12


def test_assembly_matches_oracle_thin_box_projection_fans():
    """0.2 mm gap: projection-fan near patches (per-job kernel) and the all-pairs mode."""
    V = np.array([(x, y, z) for z in (0, 0.0002) for y in (0, 0.02) for x in (0, 0.02)], float)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    T = [t for a, b, c, d in quads for t in ((a, b, c), (a, c, d))]
    mesh = bem.SurfaceMesh(V, np.array(T, np.uint32), np.arange(12, dtype=np.uint32), np.ones(12, np.uint8))
    k = 2 * math.pi * 3000 / 343.0
    num = bem.BemNumerics()
    plan = bem.interaction_plan(mesh, num, k)
    assert np.any(plan["patches"][:, 20] == 1)
    A, B = bem.HipBemSolver(mesh).assemble(num, k, k)
    Ao, Bo = ob.oracle_assemble(mesh, plan, k)
    assert _rel(A, Ao) <= 1e-12 and _rel(B, Bo) <= 1e-12


def test_solve_matches_oracle_and_analytic_icosphere():
    """Config 1: icosphere L3 (642 v / 1280 t), 1 kHz, rho=1.205, c=343, dense FP64 + LU."""
    mesh = meshes.icosphere(3)
    f, rho, c = 1000.0, 1.205, 343.0
    k = 2 * math.pi * f / c
    req = bem.BemRequest(medium=bem.AcousticMedium(rho, c), frequencies=[f],
                         excitations=[bem.BemExcitation("unit", [bem.FaceVelocity(1, 1.0)])])
    res = bem.HipBemSolver(mesh).solve(req)
    ex = res.frequencies[0].excitations[0]
    assert res.triangle_count == 1280 and res.pressure_dof_count == 642
    Ao, Bo = ob.oracle_assemble(mesh, bem.interaction_plan(mesh, bem.BemNumerics(), k), k)
    q = flux(mesh, {1: 1.0}, f, rho)
    p = ob.oracle_lu_solve(Ao, Bo @ q)
    err = np.linalg.norm(ex.boundary.pressure_trace - p) / np.linalg.norm(p)
    expected = analytic(f, 0.1, rho, c)  # 318.42 + 173.83i Pa
    areas = np.array([np.linalg.norm(np.cross(mesh.vertices[t[1]] - mesh.vertices[t[0]],
                                              mesh.vertices[t[2]] - mesh.vertices[t[0]])) / 2 for t in mesh.triangles])
    mean = np.sum(ex.boundary.surface_pressure * areas) / areas.sum()
    print(f"C1: L2 vs oracle {err:.2e}, mean surface pressure {mean:.3f} vs analytic {expected:.3f}, "
          f"residual {ex.boundary_residual:.2e}")
    assert err <= 1e-8
    assert abs(mean / expected - 1) < 0.02
    assert ex.boundary_residual <= 1e-10


def test_beat_sphere_observers():
    """Reference gate (bem_cuda_tests.cpp:416-459) tightened to the oracle's agreement with BEAT."""
    beat = beat_sphere()
    res = bem.HipBemSolver(meshes.octa_sphere(3)).solve(sphere_request(BEAT_FREQUENCIES))
    for fr in res.frequencies:
        got = fr.excitations[0].observer_pressure
        want = beat[fr.frequency]
        l2 = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"BEAT {fr.frequency} Hz: L2 {l2:.2e}")
        assert l2 <= 2e-6
        ratio = fr.excitations[1].observer_pressure / got
        assert np.max(np.abs(ratio - 0.6 * np.exp(0.71j))) < 1e-9
        for j, r in enumerate((.3, 1.0, 3.0)):
            assert abs(got[j] / analytic(fr.frequency, r) - 1) < 0.035


def test_beat_tetra_ports():
    ref = beat_tetra()
    mesh = meshes.scalene_tetra()
    req = bem.BemRequest(medium=bem.AcousticMedium(1.21, 343.0), frequencies=[400.0],
                         excitations=[bem.BemExcitation(n, [bem.FaceVelocity(f, v) for f, v in vel.items()])
                                      for n, vel in TETRA_EXCITATIONS.items()],
                         numerics=bem.BemNumerics(singular_order=12))
    res = bem.HipBemSolver(mesh).solve(req)
    for ex in res.frequencies[0].excitations:
        for imp in ex.impedances:
            want = ref[(ex.name, imp.face_id)]
            assert abs(imp.volume_impedance / want["Z"] - 1) <= 1e-8
            assert abs(imp.area - want["area"]) <= 1e-15
            assert abs(imp.mean_pressure / imp.volume_velocity - imp.volume_impedance) <= 1e-12 * abs(imp.volume_impedance)


def test_far_field_pulsating_sphere():
    mesh = meshes.icosphere(3)
    f = 1000.0
    k = 2 * math.pi * f / 343.0
    req = sphere_request([f], observers=[(0, 0, 1), (1, 0, 0), (0, -1, 0)])
    req.far_field = bem.FarFieldObservation((0, 0, 0), 1.0)
    res = bem.HipBemSolver(mesh).solve(req)
    got = res.frequencies[0].excitations[0].observer_pressure
    # p(r) -> A a exp(ika) exp(-ikr)/r with A = p(a)
    want = analytic(f, 0.1) * 0.1 * np.exp(1j * k * 0.1) * np.exp(-1j * k * 1.0) / 1.0
    assert np.max(np.abs(got / want - 1)) < 0.01


def test_far_field_matches_oracle():
    mesh = meshes.icosphere(2)
    f, rho = 1500.0, 1.21
    k = 2 * math.pi * f / 343.0
    dirs = [(0, 0, 1), (0.3, 0.4, -0.5), (-1, 0, 0)]
    req = sphere_request([f], observers=dirs)
    req.far_field = bem.FarFieldObservation((0.01, 0.0, -0.02), 2.0)
    res = bem.HipBemSolver(mesh).solve(req)
    got = np.stack([e.observer_pressure for e in res.frequencies[0].excitations], axis=1)
    P = np.stack([e.boundary.pressure_trace for e in res.frequencies[0].excitations], axis=1)
    Q = np.stack([flux(mesh, {1: 1.0}, f, rho), flux(mesh, {1: 0.6 * np.exp(0.71j)}, f, rho)], axis=1)
    want = oracle_observer_field(ob, mesh, k, P, Q, dirs, far=True, origin=(0.01, 0.0, -0.02), R=2.0)
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) <= 1e-10


def test_near_observer_adaptive_field():
    mesh = meshes.octa_sphere(3)
    f = 1715.0
    res = bem.HipBemSolver(mesh).solve(sphere_request([f], observers=[(0, 0, .101)]))
    p = res.frequencies[0].excitations[0].observer_pressure[0]
    assert abs(p / analytic(f, .101) - 1) < .035


def test_gmres_manufactured_system():
    """The manufactured 17x17 non-normal system of bem_device_gmres_tests.cpp:100-123."""
    n = 17
    expected = np.array([complex(math.sin(.7 * i) + .1, math.cos(.3 * i) - .2) for i in range(n)])
    A = np.zeros((n, n), complex)
    for r in range(n):
        for c in range(n):
            v = complex(.006 * math.sin(r + 2. * c), .004 * math.cos(3. * r - c))
            if r == c:
                v += complex(3.5 + .03 * r, .2)
            if c == r + 1:
                v += complex(1.3, .7)
            if r == c + 1:
                v += complex(-.2, .6)
            A[r, c] = v
    b = A @ expected
    x, res, its = bem.dense_gmres(A, b, restart=4, maximum_iterations=120, tolerance=1e-9)
    err = np.linalg.norm(x - expected) / np.linalg.norm(expected)
    assert 4 < its < 120 and err < 2e-8 and res < 1e-8


@pytest.mark.parametrize("n,nrhs", [(1, 1), (31, 2), (64, 1), (65, 1), (97, 3), (129, 2), (300, 2), (1234, 2)])
def test_dense_lu_matches_numpy(n, nrhs):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    X = bem.dense_lu_solve(A, B)
    Xn = np.linalg.solve(A, B)
    assert np.max(np.abs(X - Xn)) / np.max(np.abs(Xn)) <= 1e-10


def test_dense_lu_reports_singular_pivot():
    """A zero column stays zero through every update, so its pivot search finds no
    nonzero entry: the factorization reports the first singular pivot (1-based)."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((300, 300)) + 1j * rng.standard_normal((300, 300))
    A[:, 40] = 0
    with pytest.raises(Exception, match="pivot 41"):
        bem.dense_lu_solve(A, np.ones((300, 1), complex))


@pytest.mark.parametrize("n", [700, 3000])
def test_gemv_column_chunks_agree(monkeypatch, n):
    """The one-vector product's column chunk (512, 1024, 2048: the autotuner's
    candidates, pinned by HSBEM_GEMV_COLS) changes only the split-K summation order:
    GMRES reaches the same solution, which solves the system."""
    rng = np.random.default_rng(n)
    A = 4 * np.eye(n) + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    xs = {}
    for cols in ("512", "1024", "2048"):
        monkeypatch.setenv("HSBEM_GEMV_COLS", cols)
        x, achieved, its = bem.dense_gmres(A, b, 50, 500, 1e-12)
        assert achieved <= 1e-12
        xs[cols] = x
    monkeypatch.delenv("HSBEM_GEMV_COLS")
    x_auto, _, _ = bem.dense_gmres(A, b, 50, 500, 1e-12)  # the tuned chunk
    ref = np.linalg.solve(A, b)
    for x in list(xs.values()) + [x_auto]:
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-10
        assert np.linalg.norm(x - xs["512"]) / np.linalg.norm(ref) <= 1e-11


def test_block_jacobi_gmres_matches_unpreconditioned():
    """Right block-Jacobi preconditioning changes the Krylov path, not the solution."""
    mesh = meshes.exponential_horn(0.02)
    solver = bem.HipBemSolver(mesh)
    req = lambda pre: bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=[3000.0],
                                     excitations=[bem.BemExcitation("throat", [bem.FaceVelocity(1, 1.0)])],
                                     numerics=bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-10,
                                                              preconditioner=pre))
    plain = solver.solve(req(bem.PRECONDITIONER_NONE)).frequencies[0].excitations[0]
    jac = solver.solve(req(bem.PRECONDITIONER_BLOCK_JACOBI)).frequencies[0].excitations[0]
    err = np.linalg.norm(jac.boundary.pressure_trace - plain.boundary.pressure_trace) / \
        np.linalg.norm(plain.boundary.pressure_trace)
    print(f"block-Jacobi: {plain.iterations} -> {jac.iterations} iterations, difference {err:.2e}")
    assert err <= 1e-8
    assert jac.boundary_residual <= 1e-10 and plain.boundary_residual <= 1e-10
    assert jac.iterations < plain.iterations


def test_gmres_solve_matches_lu():
    mesh = meshes.icosphere(3)
    num_lu = bem.BemNumerics()
    num_gm = bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-11)
    solver = bem.HipBemSolver(mesh)
    r_lu = solver.solve(sphere_request([700.0], numerics=num_lu))
    r_gm = solver.solve(sphere_request([700.0], numerics=num_gm))
    a = r_lu.frequencies[0].excitations[0].boundary.pressure_trace
    b = r_gm.frequencies[0].excitations[0].boundary.pressure_trace
    assert np.linalg.norm(a - b) / np.linalg.norm(a) <= 1e-9
    assert r_gm.frequencies[0].excitations[0].iterations > 0


# Mixed form: FP32 regular and near samples with FP32 accumulation, FP64 geometry
# differences, FP64 touching (Sauter-Schwab) pairs, FP32 matrix, FP64 Krylov
# vectors and FP64 residual. Gates are against the FP64 oracle; the reference's
# own mixed CUDA output is gated at 1 % L2 / 0.1 dB / 1 degree against BEAT
# (bem_cuda_tests.cpp:416-459).
MIXED = dict(precision=bem.PRECISION_MIXED)


@pytest.mark.parametrize("level,frequency", [(3, 1715.0), (3, 4000.0)])
def test_mixed_assembly_close_to_oracle(level, frequency):
    mesh = meshes.octa_sphere(level)
    k = 2 * math.pi * frequency / 343.0
    num = bem.BemNumerics(**MIXED)
    A, B = bem.HipBemSolver(mesh).assemble(num, k, k)
    Ao, Bo = ob.oracle_assemble(mesh, bem.interaction_plan(mesh, num, k), k)
    ea, eb = _rel(A, Ao), _rel(B, Bo)
    print(f"mixed octa{level} {frequency} Hz: A {ea:.2e}  B {eb:.2e}")
    assert ea <= 2e-6 and eb <= 2e-6


def test_mixed_solve_close_to_fp64_icosphere():
    mesh = meshes.icosphere(3)
    f, rho = 1000.0, 1.205
    k = 2 * math.pi * f / 343.0
    req = bem.BemRequest(medium=bem.AcousticMedium(rho, 343.0), frequencies=[f],
                         excitations=[bem.BemExcitation("unit", [bem.FaceVelocity(1, 1.0)])],
                         numerics=bem.BemNumerics(**MIXED))
    ex = bem.HipBemSolver(mesh).solve(req).frequencies[0].excitations[0]
    Ao, Bo = ob.oracle_assemble(mesh, bem.interaction_plan(mesh, bem.BemNumerics(), k), k)
    p = ob.oracle_lu_solve(Ao, Bo @ flux(mesh, {1: 1.0}, f, rho))
    err = np.linalg.norm(ex.boundary.pressure_trace - p) / np.linalg.norm(p)
    print(f"mixed C1: L2 vs FP64 oracle {err:.2e}, residual {ex.boundary_residual:.2e}, iterations {ex.iterations}")
    assert err <= 1e-5
    assert ex.boundary_residual <= bem.BemNumerics().relative_residual_tolerance
    assert ex.iterations > 0


def test_mixed_beat_sphere_observers():
    beat = beat_sphere()
    res = bem.HipBemSolver(meshes.octa_sphere(3)).solve(sphere_request(BEAT_FREQUENCIES, numerics=bem.BemNumerics(**MIXED)))
    for fr in res.frequencies:
        got = fr.excitations[0].observer_pressure
        want = beat[fr.frequency]
        l2 = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"mixed BEAT {fr.frequency} Hz: L2 {l2:.2e}")
        assert l2 <= 1e-5


@pytest.mark.parametrize("frequency", [2000.0, 5000.0])
def test_mixed_horn_close_to_fp64(frequency):
    mesh = meshes.exponential_horn(0.02)
    solver = bem.HipBemSolver(mesh)
    req = lambda num: bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=[frequency],
                                     excitations=[bem.BemExcitation("throat", [bem.FaceVelocity(1, 1.0)])],
                                     numerics=num)
    a = solver.solve(req(bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-9)))
    b = solver.solve(req(bem.BemNumerics(**MIXED)))
    pa = a.frequencies[0].excitations[0].boundary.pressure_trace
    pb = b.frequencies[0].excitations[0].boundary.pressure_trace
    za = a.frequencies[0].excitations[0].impedances[0].volume_impedance
    zb = b.frequencies[0].excitations[0].impedances[0].volume_impedance
    err = np.linalg.norm(pb - pa) / np.linalg.norm(pa)
    print(f"mixed horn {frequency} Hz: trace L2 {err:.2e}, throat impedance {abs(zb / za - 1):.2e}")
    assert err <= 1e-5
    assert abs(zb / za - 1) <= 1e-5
