"""The C++ drop-in header (include/hornsculpt_amd/bem.hpp) driven by a compiled C++
caller (tests/facade/facade_caller.cpp, built by build()), against the ctypes
binding on the same requests: config 1's icosphere (1280 triangles, 1 kHz), one
dense FP64 LU request and one fast multipole request. Both bindings reach the same
library entry points, so the dense results agree to rounding (1e-13 relative); the
fast multipole solves stop at relative residual 1e-6 with leaf moments accumulated
by atomics, so they agree to 1e-8."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from hornsculpt_amd import bem, meshes

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
CALLER = ROOT / "tests" / "bin" / "facade_caller"
OBSERVERS = [(0.0, 0.0, 1.0), (1.0, 0.0, 0.0), (0.0, -1.0, 0.0)]


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _request(fmm):
    num = bem.BemNumerics(relative_residual_tolerance=1e-10)
    if fmm:
        num = bem.BemNumerics(relative_residual_tolerance=1e-6, solver=bem.SOLVER_GMRES,
                              acceleration=bem.ACCELERATION_FAST_MULTIPOLE, fmm_leaf_capacity=32,
                              fmm_restart_length=50)
    return bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=[1000.0],
                          excitations=[bem.BemExcitation("unit", [bem.FaceVelocity(1, 1.0)])],
                          observers=OBSERVERS, far_field=bem.FarFieldObservation((0.0, 0.0, 0.0), 1.0), numerics=num)


def test_cpp_facade_caller_matches_ctypes_binding(tmp_path):
    assert CALLER.exists(), "tests/bin/facade_caller is missing: run __graft_entry__.build()"
    mesh = meshes.icosphere(3)
    nv, nt = len(mesh.vertices), len(mesh.triangles)
    path_in, path_out = tmp_path / "mesh.bin", tmp_path / "out.bin"
    with open(path_in, "wb") as f:
        f.write(np.array([nv, nt], np.uint32).tobytes())
        f.write(np.ascontiguousarray(mesh.vertices, np.float64).tobytes())
        f.write(np.ascontiguousarray(mesh.triangles, np.uint32).tobytes())
        f.write(np.ascontiguousarray(mesh.face_ids, np.uint32).tobytes())
        f.write((np.asarray(mesh.boundary) == bem.DRIVEN).astype(np.uint8).tobytes())
    print("running the C++ facade caller", flush=True)
    proc = subprocess.run([str(CALLER), str(path_in), str(path_out)], capture_output=True, text=True, timeout=240)
    assert proc.returncode == 0, proc.stderr
    raw = path_out.read_bytes()
    solver = bem.HipBemSolver(mesh)
    pos = 0
    for fmm in (False, True):
        trace = np.frombuffer(raw, np.complex128, nv, pos)
        pos += 16 * nv
        observed = np.frombuffer(raw, np.complex128, len(OBSERVERS), pos)
        pos += 16 * len(OBSERVERS)
        residual = np.frombuffer(raw, np.float64, 1, pos)[0]
        its, callbacks = np.frombuffer(raw, np.uint32, 2, pos + 8)
        links, near = np.frombuffer(raw, np.uint64, 2, pos + 16)
        evaluation = raw[pos + 32]
        pos += 33
        res = solver.solve(_request(fmm))
        point = res.frequencies[0]
        ex = point.excitations[0]
        tol = 1e-8 if fmm else 1e-13
        print(f"{'fast multipole' if fmm else 'dense'}: trace {_rel(trace, ex.boundary.pressure_trace):.1e}, "
              f"observers {_rel(observed, ex.observer_pressure):.1e}, residual {residual:.1e}, iterations {its}, "
              f"links {links}, near pairs {near}", flush=True)
        assert _rel(trace, ex.boundary.pressure_trace) <= tol
        assert _rel(observed, ex.observer_pressure) <= tol
        assert callbacks > 0
        if fmm:
            assert residual <= 1e-6 and its > 0 and links == point.fmm["m2l_links"] > 0
            assert 0 < near < nv * nv
            assert evaluation == (0 if point.fmm["direct_audit"] else 1)
        else:
            assert residual <= 1e-10 and its == 0 and links == 0 and near == 0
