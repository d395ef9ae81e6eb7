"""Device FMM pressure operator: A x = exact near CSR x + low-frequency far field
(leaf moments, rotate / coaxial / rotate translations over the support octree).

Parity: the operator's action equals the dense row-scaled Burton-Miller matrix
(FP64 oracle on the CPU for the sphere, the device dense assembly, itself
parity-tested against the oracle, for the horn) to the FMM accuracy, and GMRES
on the FMM operator reproduces the dense solution."""
import dataclasses
import math

import numpy as np
import pytest

import oracle_bind as ob
from hornsculpt_amd import bem, meshes

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def sphere_case():
    mesh = meshes.icosphere(3)
    k = 2 * math.pi * 1000.0 / 343.0
    num = bem.BemNumerics()
    Ao, _ = ob.oracle_assemble(mesh, bem.interaction_plan(mesh, num, k), k)
    return mesh, k, num, Ao


def test_fmm_apply_matches_dense_oracle_sphere(sphere_case):
    mesh, k, num, Ao = sphere_case
    solver = bem.HipBemSolver(mesh)
    op = solver.fmm_operator(num, k, k, leaf_capacity=32)
    st = op.statistics()
    print("sphere FMM", {key: st[key] for key in ("levels", "leaves", "far_pairs", "m2l_links", "near_nnz")},
          "orders", st["level_orders"])
    assert st["far_pairs"] > 0 and st["m2l_links"] > 0
    rng = np.random.default_rng(1)
    for _ in range(2):
        x = rng.standard_normal(len(mesh.vertices)) + 1j * rng.standard_normal(len(mesh.vertices))
        err = _rel(op.apply(x), Ao @ x)
        print(f"sphere FMM apply rel err {err:.2e}")
        assert err <= 2e-6
    op.close()


def test_fmm_operator_rejects_bad_seed(sphere_case):
    mesh, k, num, _ = sphere_case
    solver = bem.HipBemSolver(mesh)
    with pytest.raises(ValueError, match="accuracy seed"):
        solver.fmm_operator(num, k, k, accuracy_seed=2.0)


@pytest.mark.parametrize("frequency", [500.0, 3000.0])
def test_fmm_apply_matches_dense_horn(frequency):
    mesh = meshes.exponential_horn(0.012)
    k = 2 * math.pi * frequency / 343.0
    num = bem.BemNumerics()
    solver = bem.HipBemSolver(mesh)
    A, _ = solver.assemble(num, k, k)
    op = solver.fmm_operator(num, k, k, leaf_capacity=32)
    st = op.statistics()
    print(f"horn {len(mesh.triangles)} panels {frequency} Hz", {key: st[key] for key in
          ("levels", "leaves", "far_pairs", "m2l_links", "rotations", "axial_plans")}, "orders", st["level_orders"],
          f"build ms tree {st['tree_ms']:.1f} near {st['near_ms']:.1f} plans {st['plans_ms']:.1f} "
          f"moments {st['moments_ms']:.1f}")
    rng = np.random.default_rng(2)
    x = rng.standard_normal(A.shape[0]) + 1j * rng.standard_normal(A.shape[0])
    err = _rel(op.apply(x), A @ x)
    print(f"horn FMM apply rel err {err:.2e}, apply {op.time_apply(5):.3f} ms")
    assert err <= 2e-6
    op.close()


def test_direct_audit_action_matches_dense_horn():
    """The direct Galerkin audit (exact near CSR + far part integrated directly over
    every far DOF pair of the same leaf partition) equals the dense matrix and the
    FMM action to quadrature accuracy."""
    mesh = meshes.exponential_horn(0.012)
    k = 2 * math.pi * 2000.0 / 343.0
    num = bem.BemNumerics()
    solver = bem.HipBemSolver(mesh)
    A, _ = solver.assemble(num, k, k)
    op = solver.fmm_operator(num, k, k, leaf_capacity=32)
    x = np.random.default_rng(5).standard_normal(A.shape[0]) + 1j * np.random.default_rng(6).standard_normal(A.shape[0])
    yd, yf, yx = A @ x, op.apply(x), op.apply_direct(x)
    print(f"direct audit vs dense {_rel(yx, yd):.2e}, vs FMM {_rel(yx, yf):.2e}, FMM vs dense {_rel(yf, yd):.2e}")
    assert _rel(yx, yd) <= 1e-6 and _rel(yx, yf) <= 1e-6
    op.close()


@pytest.mark.parametrize("degree", [12, 14])
def test_fmm_apply_with_promoted_endpoint_rule(degree):
    """Leaf panel rules of degree 12 and 14 (49 and 64 points: one wavefront per
    (leaf, panel) item) give the same action as the minimum rule to the FMM accuracy
    and match the dense matrix; the direct audit at the same rule agrees."""
    mesh = meshes.exponential_horn(0.012)
    k = 2 * math.pi * 2500.0 / 343.0
    num = bem.BemNumerics()
    solver = bem.HipBemSolver(mesh)
    A, _ = solver.assemble(num, k, k)
    base = solver.fmm_operator(num, k, k, leaf_capacity=32)
    high = solver.fmm_operator(num, k, k, leaf_capacity=32, endpoint_degree=degree)
    assert high.statistics()["rule_points"] > 16
    # same mesh, wavenumber and seed: every translation plan comes from the mesh's plan cache
    assert high.statistics()["plans_built"] == 0 and high.statistics()["plan_cache_hits"] > 0
    x = np.random.default_rng(9).standard_normal(A.shape[0]) + 1j * np.random.default_rng(10).standard_normal(A.shape[0])
    yd, yb, yh, yx = A @ x, base.apply(x), high.apply(x), high.apply_direct(x)
    print(f"degree {degree} ({high.statistics()['rule_points']} points): vs dense {_rel(yh, yd):.2e}, vs minimum rule "
          f"{_rel(yh, yb):.2e}, direct audit {_rel(yx, yh):.2e}")
    assert _rel(yh, yd) <= 2e-6 and _rel(yh, yb) <= 1e-6 and _rel(yx, yh) <= 1e-6
    with pytest.raises(ValueError, match="endpoint degree"):
        solver.fmm_operator(num, k, k, endpoint_degree=10)
    base.close()
    high.close()


def test_fmm_gmres_reproduces_dense_solution():
    mesh = meshes.exponential_horn(0.012)
    k = 2 * math.pi * 1500.0 / 343.0
    num = bem.BemNumerics()
    solver = bem.HipBemSolver(mesh)
    A, B = solver.assemble(num, k, k)
    driven = np.nonzero(mesh.boundary == bem.DRIVEN)[0]
    q = np.zeros(B.shape[1], complex)
    q[driven] = 1.0
    b = B @ q
    x_ref = np.linalg.solve(A, b)
    op = solver.fmm_operator(num, k, k, leaf_capacity=32)
    x, achieved, its = op.solve(b, tolerance=1e-9)
    x_plain, _, its_plain = op.solve(b, tolerance=1e-9, precondition=False)
    err = _rel(x, x_ref)
    print(f"FMM GMRES: {its} iterations (unpreconditioned {its_plain}), residual {achieved:.1e}, "
          f"rel err vs dense {err:.2e}")
    assert achieved <= 1e-9 and err <= 5e-6
    assert _rel(x_plain, x_ref) <= 5e-6
    assert its <= its_plain
    op.close()


@pytest.fixture(scope="module")
def horn_near():
    """1950-node horn at 1.5 kHz: the near CSR of the FMM split (mesh order) and the
    support tree's pressure order for leaf capacity 32."""
    mesh = meshes.exponential_horn(0.012)
    k = 2 * math.pi * 1500.0 / 343.0
    num = bem.BemNumerics()
    solver = bem.HipBemSolver(mesh)
    near = solver.near_operator(num, k, k, [1], leaf_capacity=32)
    n = len(mesh.vertices)
    rows = np.repeat(np.arange(n), np.diff(near["row_offsets"]))
    N = np.zeros((n, n), complex)
    N[rows, near["columns"]] = near["values"]
    shifted = bem.SurfaceMesh(mesh.vertices - mesh.vertices[0], mesh.triangles, mesh.face_ids, mesh.boundary)
    order = bem.parse_support_tree(bem.support_tree(shifted, leaf_capacity=32))["pressure_order"]
    return mesh, k, solver, N, order


@pytest.mark.parametrize("width", [32, 64, 128, 256, 1024, 2048])
def test_fmm_block_jacobi_widths(horn_near, width):
    """preconditionerBlockWidth 32..2048 (support_preconditioner.cu:33-190): z = M^{-1} v
    equals, block by block, the numpy solve with the diagonal W x W blocks of the near
    matrix in tree order (blocks with tails: 1950 = 60 x 32 + 30; at 2048 the single
    block is the whole near matrix, independent of the order), and GMRES converges."""
    mesh, k, solver, N, order = horn_near
    num = dataclasses.replace(bem.BemNumerics(), fmm_block_width=width)
    op = solver.fmm_operator(num, k, k, leaf_capacity=32)
    rng = np.random.default_rng(width)
    v = rng.standard_normal(len(order)) + 1j * rng.standard_normal(len(order))
    z = op.precondition(v)
    want = np.zeros_like(v)
    for b0 in range(0, len(order), width):
        idx = order[b0:b0 + width]
        want[idx] = np.linalg.solve(N[np.ix_(idx, idx)], v[idx])
    err = _rel(z, want)
    b = N @ np.ones(len(order))
    _, achieved, its = op.solve(b, tolerance=1e-9)
    print(f"block width {width}: M^-1 v rel err {err:.2e}, GMRES {its} iterations, residual {achieved:.1e}")
    assert err <= 1e-10
    assert achieved <= 1e-9
    op.close()


def test_fmm_block_width_validation(horn_near):
    mesh, k, solver, _, _ = horn_near
    for width in (16, 48, 4096):
        with pytest.raises(ValueError):
            solver.fmm_operator(dataclasses.replace(bem.BemNumerics(), fmm_block_width=width), k, k, leaf_capacity=32)


def _horn_request(frequencies, numerics, observers=()):
    mesh = meshes.exponential_horn(0.012)
    driven = sorted(set(int(f) for f, b in zip(mesh.face_ids, mesh.boundary) if b == bem.DRIVEN))
    req = bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=list(frequencies),
                         excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(f, 1.0) for f in driven])],
                         observers=list(observers), numerics=numerics)
    return mesh, req


def test_fmm_request_matches_dense_request():
    """hsbem_solve with the fast multipole form reproduces the dense FP64 solve (same
    residual tolerance) at low and high frequency, observers included."""
    freqs = [400.0, 2500.0]
    observers = [(0.35, 0.0, 0.0), (0.2, 0.15, 0.05)]
    dense = bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-6)
    fmm = bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-6,
                          acceleration=bem.ACCELERATION_FAST_MULTIPOLE)
    mesh, rd = _horn_request(freqs, dense, observers)
    _, rf = _horn_request(freqs, fmm, observers)
    solver = bem.HipBemSolver(mesh)
    res_d, res_f = solver.solve(rd), solver.solve(rf)
    for fd, ff in zip(res_d.frequencies, res_f.frequencies):
        ed, ef = fd.excitations[0], ff.excitations[0]
        err = _rel(ef.boundary.pressure_trace, ed.boundary.pressure_trace)
        obs = _rel(ef.observer_pressure, ed.observer_pressure)
        print(f"{ff.frequency:.0f} Hz: FMM vs dense trace {err:.2e}, observers {obs:.2e}, residual "
              f"{ef.boundary_residual:.1e}, iterations {ef.iterations} (dense {ed.iterations}), "
              f"build {ff.assembly_milliseconds:.0f} ms, solve {ff.solve_milliseconds:.0f} ms")
        assert ef.boundary_residual <= 1e-6
        assert err <= 1e-5 and obs <= 1e-5
        assert ef.iterations > 0
        print("  FMM evidence", ff.fmm)
        assert ff.fmm is not None and ff.fmm["m2l_links"] > 0
        assert ff.fmm["quadrature_discrepancy"] <= 0.025 * 1e-6 and ff.fmm["endpoint_degree"] >= 4
        assert ff.fmm["direct_audit"] or (ff.fmm["verification_seed"] <= 1e-9 and ff.fmm["audit_discrepancy"] <= 0.25e-6)


def test_fmm_request_fp32_bulk_form():
    """precision=mixed with the fast multipole form: the Krylov iterations apply FP32
    translations, the residual (FP64 operator), audits and corrections are those of the
    plain form, so the solution holds the same tolerance and matches the plain FMM and
    the dense solves to it."""
    freqs = [400.0, 2500.0]
    plain = bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-6,
                            acceleration=bem.ACCELERATION_FAST_MULTIPOLE)
    bulk = bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-6,
                           acceleration=bem.ACCELERATION_FAST_MULTIPOLE, precision=bem.PRECISION_MIXED)
    dense = bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-8)
    mesh, rp = _horn_request(freqs, plain)
    _, rb = _horn_request(freqs, bulk)
    _, rd = _horn_request(freqs, dense)
    solver = bem.HipBemSolver(mesh)
    res_p, res_b, res_d = solver.solve(rp), solver.solve(rb), solver.solve(rd)
    for fp, fb, fd in zip(res_p.frequencies, res_b.frequencies, res_d.frequencies):
        ep, eb, ed = fp.excitations[0], fb.excitations[0], fd.excitations[0]
        vs_plain = _rel(eb.boundary.pressure_trace, ep.boundary.pressure_trace)
        vs_dense = _rel(eb.boundary.pressure_trace, ed.boundary.pressure_trace)
        print(f"{fb.frequency:.0f} Hz: FP32-bulk vs plain FMM {vs_plain:.2e}, vs dense {vs_dense:.2e}, residual "
              f"{eb.boundary_residual:.1e}, iterations {eb.iterations} (plain {ep.iterations}), corrected "
              f"{fb.fmm['corrected']}, solve {fb.solve_milliseconds:.0f} ms (plain {fp.solve_milliseconds:.0f} ms)")
        assert fb.fmm["fp32_translations"] and not fp.fmm["fp32_translations"]
        assert eb.boundary_residual <= 1e-6
        assert vs_plain <= 1e-5 and vs_dense <= 1e-5


def test_fmm_fp32_apply_matches_fp64_apply():
    """The FP32-bulk action agrees with the FP64 one to FP32 rounding of the
    expansions (1e-5 relative here)."""
    mesh = meshes.exponential_horn(0.012)
    k = 2 * math.pi * 2000.0 / 343.0
    solver = bem.HipBemSolver(mesh)
    op = solver.fmm_operator(bem.BemNumerics(), k, k, leaf_capacity=32)
    rng = np.random.default_rng(17)
    x = rng.standard_normal(op.size) + 1j * rng.standard_normal(op.size)
    y64, y32 = op.apply(x), op.apply_fp32(x)
    err = _rel(y32, y64)
    print(f"FP32-bulk apply vs FP64 apply {err:.2e}")
    assert err <= 1e-5
    op.close()


def test_fmm_request_validation():
    mesh, req = _horn_request([500.0], bem.BemNumerics(acceleration=bem.ACCELERATION_FAST_MULTIPOLE,
                                                       fmm_leaf_capacity=2))
    solver = bem.HipBemSolver(mesh)
    with pytest.raises(ValueError, match="Fast Multipole"):
        solver.solve(req)
    for width in (16, 96, 4096):  # preconditionerBlockWidth: 32, 64, ..., 2048 (cuda_solver.cu:1063-1066)
        _, req = _horn_request([500.0], bem.BemNumerics(acceleration=bem.ACCELERATION_FAST_MULTIPOLE,
                                                        fmm_block_width=width))
        with pytest.raises(ValueError, match="Fast Multipole"):
            solver.solve(req)


def test_fmm_graph_replays_match_eager_applies():
    """The captured apply, replayed several times with eager applies and host
    synchronisations in between, equals the eager apply every time; GMRES through
    the graph equals GMRES without it."""
    mesh = meshes.exponential_horn(0.012)
    k = 2 * math.pi * 1200.0 / 343.0
    solver = bem.HipBemSolver(mesh)
    op = solver.fmm_operator(bem.BemNumerics(), k, k, leaf_capacity=32)
    rng = np.random.default_rng(7)
    xs = [rng.standard_normal(op.size) + 1j * rng.standard_normal(op.size) for _ in range(4)]
    eager = [op.apply(x) for x in xs]
    eager_ms = op.time_apply(5)
    op.set_graph(True)
    for rep in range(3):
        for x, ye in zip(xs, eager):
            yg = op.apply(x)
            assert np.array_equal(yg, ye), f"replay {rep} differs"
            op.set_graph(False)
            assert np.array_equal(op.apply(x), ye)  # eager work between replays
            op.set_graph(True)
    graph_ms = op.time_apply(5)
    b = eager[0]
    xg, resg, itg = op.solve(b, tolerance=1e-9)
    op.set_graph(False)
    xe, rese, ite = op.solve(b, tolerance=1e-9)
    print(f"FMM apply eager {eager_ms:.3f} ms, graph {graph_ms:.3f} ms; GMRES {itg} / {ite} iterations")
    assert itg == ite and np.array_equal(xg, xe)
    op.close()
