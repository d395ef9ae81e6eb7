"""Dense GMRES with device-resident restart cycles replayed as a captured HIP
graph (the default) against the host-driven cycles (HSBEM_GMRES_GRAPH=0): the
Arnoldi reductions are the same kernels in the same order and the rotations the
same formulas evaluated on the device, so both reach the same solution to 1e-12
(relative) with the same iteration count, over several restart cycles, with and
without the block-Jacobi preconditioner, and inside full Burton-Miller solves.
Between two full solves of the mixed form the FP32 matrix itself differs in its
last bits (the regular-pair tiles accumulate with FP32 atomics, in no fixed
order), so there the bar is that of two mixed solves, 2e-5."""
import math

import numpy as np
import pytest

from hornsculpt_amd import bem, meshes

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("n,restart", [(300, 10), (1000, 12), (3000, 20)])
def test_graph_gmres_equals_host_driven(monkeypatch, n, restart):
    rng = np.random.default_rng(n)
    A = 3 * np.eye(n) + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("HSBEM_GMRES_GRAPH", mode)
        out[mode] = bem.dense_gmres(A, b, restart, 2000, 1e-11)
    (x0, r0, i0), (x1, r1, i1) = out["0"], out["1"]
    print(f"n={n} restart={restart}: host-driven {i0} iterations ({r0:.1e}), graph {i1} ({r1:.1e}), "
          f"difference {_rel(x1, x0):.1e}")
    assert i0 == i1 > restart and r1 <= 1e-11
    assert _rel(x1, x0) <= 1e-12
    assert _rel(A @ x1, b) <= 1e-10


@pytest.mark.parametrize("precision", [bem.PRECISION_FP64, bem.PRECISION_MIXED])
def test_graph_gmres_in_burton_miller_solve(monkeypatch, precision):
    mesh = meshes.exponential_horn(0.012)
    req = bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=[700.0, 4000.0],
                         excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(1, 1.0)])],
                         numerics=bem.BemNumerics(solver=bem.SOLVER_GMRES, precision=precision,
                                                  relative_residual_tolerance=1e-9, gmres_restart=30,
                                                  gmres_maximum_iterations=2000))
    solver = bem.HipBemSolver(mesh)
    res = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("HSBEM_GMRES_GRAPH", mode)
        res[mode] = solver.solve(req)
    for f0, f1 in zip(res["0"].frequencies, res["1"].frequencies):
        e0, e1 = f0.excitations[0], f1.excitations[0]
        err = _rel(e1.boundary.pressure_trace, e0.boundary.pressure_trace)
        print(f"{f0.frequency:.0f} Hz: iterations {e0.iterations} / {e1.iterations}, graph vs host-driven {err:.1e}, "
              f"solve {f0.solve_milliseconds:.1f} / {f1.solve_milliseconds:.1f} ms")
        assert e0.iterations == e1.iterations
        assert err <= (1e-12 if precision == bem.PRECISION_FP64 else 2e-5)
        assert e1.boundary_residual <= 1e-9
