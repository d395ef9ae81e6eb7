"""Shapes of the dense GMRES pieces that the autotuner or a tuning variable can pick:
the FP32 product with one, two or four rows per lane (HSBEM_GEMV_COLS = "COLS" or
"COLSxROWS") and the dense block-Jacobi width (HSBEM_DENSE_JACOBI_WIDTH, 64 or a
gathered wider block). Each pinned shape must solve the same Burton-Miller system:
the solve's final FP64 residual (a separate kernel) meets the tolerance and the
traces agree with the default shape's to the bar of two mixed solves (the FP32
matrix of two mixed assemblies differs in its last bits: the general tiles add with
FP32 atomics in no fixed order) or, in the FP64 form, to the GMRES tolerance."""
import numpy as np
import pytest

from hornsculpt_amd import bem, meshes

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _request(precision, tolerance):
    return bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=[900.0, 5000.0],
                          excitations=[bem.BemExcitation("piston", [bem.FaceVelocity(1, 1.0)])],
                          numerics=bem.BemNumerics(solver=bem.SOLVER_GMRES, precision=precision,
                                                   relative_residual_tolerance=tolerance, gmres_restart=30,
                                                   gmres_maximum_iterations=2000))


@pytest.fixture(scope="module")
def horn():
    mesh = meshes.exponential_horn(0.015)
    assert len(mesh.vertices) % 4 == 0  # every rows-per-lane shape applies
    return bem.HipBemSolver(mesh)


def test_fp32_product_shapes(monkeypatch, horn):
    req = _request(bem.PRECISION_MIXED, 1e-7)
    base = None
    for shape in ("512", "2048x2", "1024x4", "512x4"):
        monkeypatch.setenv("HSBEM_GEMV_COLS", shape)
        res = horn.solve(req)
        traces = [f.excitations[0].boundary.pressure_trace for f in res.frequencies]
        for f in res.frequencies:
            ex = f.excitations[0]
            print(f"shape {shape}, {f.frequency:.0f} Hz: {ex.iterations} iterations, residual {ex.boundary_residual:.1e}")
            assert ex.boundary_residual <= 1e-7
        if base is None:
            base = traces
        else:
            for t, b in zip(traces, base):
                assert _rel(t, b) <= 2e-5


@pytest.mark.parametrize("precision,bar", [(bem.PRECISION_FP64, 1e-7), (bem.PRECISION_MIXED, 2e-5)])
def test_dense_jacobi_widths(monkeypatch, horn, precision, bar):
    tolerance = 1e-9 if precision == bem.PRECISION_FP64 else 1e-7
    req = _request(precision, tolerance)
    out = {}
    for width in ("64", "128", "256"):
        monkeypatch.setenv("HSBEM_DENSE_JACOBI_WIDTH", width)
        out[width] = horn.solve(req)
    for k in range(len(req.frequencies)):
        base = out["64"].frequencies[k].excitations[0]
        for width in ("128", "256"):
            ex = out[width].frequencies[k].excitations[0]
            err = _rel(ex.boundary.pressure_trace, base.boundary.pressure_trace)
            print(f"{req.frequencies[k]:.0f} Hz, width {width}: {ex.iterations} iterations (64: {base.iterations}), "
                  f"residual {ex.boundary_residual:.1e}, vs width 64 {err:.1e}")
            assert ex.boundary_residual <= tolerance
            assert err <= bar
