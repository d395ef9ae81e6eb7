"""Plane-wave (high-frequency) levels and wide spherical orders of the device FMM.

Parity: each device angular transfer equals the host restatement of the
reference's CPU transfer (resampleSpectrum + parent/child shifts,
scaledMultipoleToSpectrum / spectrumToScaledLocal with the hand-off shifts), which
the CPU suite pins against the reference build and the addition theorem; an FMM
operator whose coarse level the plan puts in plane waves, and one whose
spherical orders exceed 40, apply the dense Burton-Miller matrix to the FMM
accuracy, and the request solve with a plane-wave level reproduces the dense solve."""
import math

import numpy as np
import pytest

from hornsculpt_amd import bem, meshes

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _random(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


K, RADIUS, SHIFT = 41.0, 0.06, np.array([0.021, -0.013, 0.017])


@pytest.mark.parametrize("source_order,target_order,bandlimit", [(14, 30, 14), (30, 61, 27)])
def test_device_outgoing_to_parent_matches_host(source_order, target_order, bandlimit):
    v = _random(2 * (source_order + 1) ** 2, 1)
    host = bem.angular_transfer(v, source_order, target_order, bandlimit, bem.TRANSFER_OUTGOING_TO_PARENT, K, SHIFT)
    dev = bem.angular_transfer_device(v, bem.DEVICE_OUTGOING_TO_PARENT, source_order, target_order, bandlimit, K, 1.0,
                                      SHIFT)
    err = _rel(dev, host)
    print(f"outgoing to parent {source_order}->{target_order} B={bandlimit}: rel {err:.1e}")
    assert err <= 1e-12


@pytest.mark.parametrize("source_order,target_order,bandlimit", [(30, 14, 14), (61, 30, 27)])
def test_device_incoming_to_child_matches_host(source_order, target_order, bandlimit):
    v = _random(2 * (source_order + 1) ** 2, 2)
    host = bem.angular_transfer(v, source_order, target_order, bandlimit, bem.TRANSFER_INCOMING_TO_CHILD, K, SHIFT)
    dev = bem.angular_transfer_device(v, bem.DEVICE_INCOMING_TO_CHILD, source_order, target_order, bandlimit, K, 1.0,
                                      SHIFT)
    err = _rel(dev, host)
    print(f"incoming to child {source_order}->{target_order} B={bandlimit}: rel {err:.1e}")
    assert err <= 1e-12


@pytest.mark.parametrize("order,rule_order", [(12, 30), (41, 126)])
def test_device_scaled_multipole_to_parent_matches_host(order, rule_order):
    c = _random((order + 1) ** 2, 3) / np.arange(1, (order + 1) ** 2 + 1)
    dirs, _ = bem.sphere_rule(rule_order)
    host = bem.scaled_multipole_to_spectrum(c, order, RADIUS, K, rule_order) * np.exp(1j * K * dirs @ SHIFT)
    dev = bem.angular_transfer_device(c, bem.DEVICE_SCALED_TO_PARENT, order, rule_order, order, K, RADIUS, SHIFT)
    err = _rel(dev, host)
    print(f"scaled multipole {order} -> rule {rule_order}: rel {err:.1e}")
    assert err <= 1e-12


@pytest.mark.parametrize("order,rule_order", [(12, 30), (41, 126)])
def test_device_incoming_to_scaled_child_matches_host(order, rule_order):
    v = _random(2 * (rule_order + 1) ** 2, 4)
    dirs, _ = bem.sphere_rule(rule_order)
    host = bem.spectrum_to_scaled_local(v * np.exp(-1j * K * dirs @ SHIFT), rule_order, order, K, RADIUS)
    dev = bem.angular_transfer_device(v, bem.DEVICE_INCOMING_TO_SCALED_CHILD, rule_order, order, order, K, RADIUS,
                                      SHIFT)
    err = _rel(dev, host)
    print(f"rule {rule_order} -> scaled local {order}: rel {err:.1e}")
    assert err <= 1e-12


def test_device_angular_transfer_validation():
    v = _random(2 * 11 ** 2, 5)
    with pytest.raises(ValueError):
        bem.angular_transfer_device(v, bem.DEVICE_OUTGOING_TO_PARENT, 10, 20, 15, K, 1.0, SHIFT)  # bandlimit > source


@pytest.fixture(scope="module")
def sphere_hf():
    """Icosphere (10242 nodes) at 20 kHz: the plan puts level 2 in plane waves."""
    mesh = meshes.icosphere(5)
    k = 2 * math.pi * 20000.0 / 343.0
    return mesh, k


def test_fmm_apply_with_plane_wave_level_matches_dense(sphere_hf):
    mesh, k = sphere_hf
    num = bem.BemNumerics()
    solver = bem.HipBemSolver(mesh)
    op = solver.fmm_operator(num, k, k, leaf_capacity=32)
    st = op.statistics()
    print("plane-wave FMM", {key: st[key] for key in ("levels", "leaves", "far_pairs", "m2l_links", "angular_links",
                                                       "diagonal_classes", "plane_wave_leaves")},
          "representation", st["level_representation"], "orders", st["level_orders"], "rules", st["level_rule_orders"])
    assert st["plane_wave_levels"] >= 1 and st["diagonal_classes"] > 0 and st["angular_links"] > 0
    A, _ = solver.assemble(num, k, k)
    for seed in (1, 2):
        x = _random(len(mesh.vertices), seed)
        err = _rel(op.apply(x), A @ x)
        print(f"plane-wave FMM apply rel err {err:.2e}, apply {op.time_apply(5):.3f} ms")
        assert err <= 2e-6
    op.close()


def test_fmm_apply_with_plane_wave_leaf_matches_dense():
    """A small sphere (42 nodes) 0.25 m from the 10242-node sphere: with leaf
    capacity 64 its leaf sits on the plane-wave level 2 at 20 kHz, so its signature
    is emitted and tested directly on the level's sphere rule."""
    big, small = meshes.icosphere(5), meshes.icosphere(1, radius=0.008)
    V = np.vstack([big.vertices, small.vertices + np.array([0.0, 0.0, 0.25])])
    T = np.vstack([big.triangles, small.triangles + len(big.vertices)]).astype(np.uint32)
    mesh = bem.SurfaceMesh(V, T, np.ones(len(T), np.uint32), np.full(len(T), bem.DRIVEN, np.uint8))
    k = 2 * math.pi * 20000.0 / 343.0
    num = bem.BemNumerics()
    solver = bem.HipBemSolver(mesh)
    op = solver.fmm_operator(num, k, k, leaf_capacity=64)
    st = op.statistics()
    print("plane-wave leaves", st["plane_wave_leaves"], "representation", st["level_representation"],
          "rules", st["level_rule_orders"])
    assert st["plane_wave_leaves"] >= 1
    A, _ = solver.assemble(num, k, k)
    x = _random(len(V), 7)
    err = _rel(op.apply(x), A @ x)
    small_rows = slice(len(big.vertices), None)
    err_small = _rel(op.apply(x)[small_rows], (A @ x)[small_rows])
    print(f"apply rel err {err:.2e} (small sphere rows {err_small:.2e}), apply {op.time_apply(5):.3f} ms")
    assert err <= 2e-6 and err_small <= 2e-6
    op.close()


def test_fmm_request_with_plane_wave_level_matches_dense(sphere_hf):
    """Pulsating sphere at 20 kHz (ka = 36.6): the FMM request (accuracy seed 5e-7,
    at which the plan puts the coarsest active level in plane waves) reproduces the
    dense request, and matches the analytic surface pressure rho c v ika / (1 + ika)
    to the discretisation error."""
    mesh, k = sphere_hf
    tol = 5e-6
    req = lambda acc: bem.BemRequest(
        medium=bem.AcousticMedium(1.21, 343.0), frequencies=[20000.0],
        excitations=[bem.BemExcitation("pulsating", [bem.FaceVelocity(1, 1.0)])],
        numerics=bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=tol, acceleration=acc))
    solver = bem.HipBemSolver(mesh)
    rd = solver.solve(req(bem.ACCELERATION_DENSE)).frequencies[0]
    rf = solver.solve(req(bem.ACCELERATION_FAST_MULTIPOLE)).frequencies[0]
    ed, ef = rd.excitations[0], rf.excitations[0]
    err = _rel(ef.boundary.pressure_trace, ed.boundary.pressure_trace)
    ika = 1j * k * 0.1
    exact = 1.21 * 343.0 * ika / (1 + ika)
    analytic = np.abs(ef.boundary.pressure_trace - exact).max() / abs(exact)
    print(f"20 kHz sphere: FMM vs dense trace {err:.2e}, vs analytic {analytic:.2e}, residual "
          f"{ef.boundary_residual:.1e}, iterations {ef.iterations} (dense {ed.iterations}), build "
          f"{rf.assembly_milliseconds:.0f} ms, solve {rf.solve_milliseconds:.0f} ms")
    assert ef.boundary_residual <= tol
    assert err <= 5e-5
    assert analytic <= 2e-2


def test_fmm_wide_spherical_orders_match_dense():
    """Accuracy seed 1e-11 puts every level above order 40 (tiled Wigner staging):
    the apply still equals the dense matrix and agrees with the seed-1e-9 operator
    (orders <= 40) to that operator's accuracy."""
    mesh = meshes.exponential_horn(0.012)
    k = 2 * math.pi * 3000.0 / 343.0
    num = bem.BemNumerics()
    solver = bem.HipBemSolver(mesh)
    A, _ = solver.assemble(num, k, k)
    wide = solver.fmm_operator(num, k, k, leaf_capacity=32, accuracy_seed=1e-11)
    narrow = solver.fmm_operator(num, k, k, leaf_capacity=32, accuracy_seed=1e-9)
    orders = wide.statistics()["level_orders"]
    print("wide orders", orders, "narrow orders", narrow.statistics()["level_orders"])
    assert max(orders) > 40
    x = _random(A.shape[0], 3)
    yw, yn, yd = wide.apply(x), narrow.apply(x), A @ x
    print(f"wide vs dense {_rel(yw, yd):.2e}, narrow vs dense {_rel(yn, yd):.2e}, wide vs narrow {_rel(yw, yn):.2e}, "
          f"wide apply {wide.time_apply(5):.3f} ms")
    assert _rel(yw, yd) <= 2e-6
    assert _rel(yw, yn) <= 1e-7
    wide.close()
    narrow.close()


def test_fmm_orders_above_64_match_dense():
    """Accuracy seed 1e-16 on a 2562-node sphere at 14 kHz puts the leaf level at
    order 65 (reduction-tiled rotations and coaxial products, the larger radial
    table of the leaf moments, the 49-point degree-12 rule for k h = 2.1): the apply
    still equals the dense matrix and the seed-1e-13 operator (orders <= 56)."""
    mesh = meshes.icosphere(4)
    k = 2 * math.pi * 14000.0 / 343.0
    num = bem.BemNumerics()
    solver = bem.HipBemSolver(mesh)
    A, _ = solver.assemble(num, k, k)
    big = solver.fmm_operator(num, k, k, leaf_capacity=32, accuracy_seed=1e-16)
    ref = solver.fmm_operator(num, k, k, leaf_capacity=32, accuracy_seed=1e-13)
    orders = big.statistics()["level_orders"]
    print("orders", orders, "reference orders", ref.statistics()["level_orders"])
    assert max(orders) > 64
    x = _random(A.shape[0], 4)
    yb, yr, yd = big.apply(x), ref.apply(x), A @ x
    print(f"big vs dense {_rel(yb, yd):.2e}, vs seed-1e-13 operator {_rel(yb, yr):.2e}, apply {big.time_apply(3):.2f} ms")
    assert _rel(yb, yd) <= 2e-6 and _rel(yb, yr) <= 1e-9
    big.close()
    ref.close()
