"""Host-side product logic on the CPU: the C ABI library, the interaction planner
(bit-exact against the reference planner compiled from its sources), input
validation and the synthetic mesh generators. No device calls."""
import math
import re
from pathlib import Path

import numpy as np
import pytest

import oracle_bind as ob
from hornsculpt_amd import bem, meshes

ROOT = Path(__file__).resolve().parents[1]
needs_ref = pytest.mark.skipif(not ob.REF.exists(), reason="reference planner (oracle/_ref) not built")


def test_library_exports_every_declared_symbol():
    lib = bem.library()
    declared = re.findall(r"HSBEM_API\s+[\w\s\*]+?\b(hsbem_\w+)\s*\(", (ROOT / "include" / "hsbem.h").read_text())
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(lib, name), name


def test_numerics_defaults_match_reference():
    lib = bem.library()
    n = bem._Numerics()
    lib.hsbem_numerics_default(n)
    # BemNumerics defaults (bem.hpp:113-123)
    assert (n.quadrature_order, n.singular_order, n.near_maximum_depth, n.maximum_refinement_steps) == (4, 4, 6, 3)
    assert (n.near_distance_ratio, n.near_relative_tolerance, n.relative_residual_tolerance) == (1.0, 1e-4, 5e-5)
    py = bem.BemNumerics()
    assert (py.quadrature_order, py.singular_order, py.near_relative_tolerance) == (4, 4, 1e-4)
    # MI355X additions: LU by default (reference behaviour); block-Jacobi when GMRES is chosen
    assert (n.solver, n.precision, n.preconditioner) == (bem.SOLVER_LU, bem.PRECISION_FP64,
                                                         bem.PRECONDITIONER_BLOCK_JACOBI)
    assert (py.solver, py.precision, py.preconditioner) == (n.solver, n.precision, n.preconditioner)
    # acceleration and FastMultipoleSettings defaults (bem.hpp:60-71, 121)
    fmm = ("acceleration", "fmm_leaf_capacity", "fmm_separation_ratio", "fmm_maximum_iterations", "fmm_restart_length",
           "fmm_block_width")
    assert tuple(getattr(n, f) for f in fmm) == (bem.ACCELERATION_DENSE, 16, 1.5, 300, 40, 32)
    assert tuple(getattr(py, f) for f in fmm) == tuple(getattr(n, f) for f in fmm)


def _box(lx, ly, lz):
    V = np.array([(x, y, z) for z in (0, lz) for y in (0, ly) for x in (0, lx)], float)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    T = []
    for a, b, c, d in quads:
        T += [(a, b, c), (a, c, d)]
    return bem.SurfaceMesh(V, np.array(T, np.uint32), np.arange(12, dtype=np.uint32), np.ones(12, np.uint8))


PLAN_CASES = [
    ("octa2", lambda: meshes.octa_sphere(2), 1000.0, {}),
    ("octa3", lambda: meshes.octa_sphere(3), 1715.0, {}),
    ("octa3-strict", lambda: meshes.octa_sphere(3), 500.0, {"tolerance": 1e-6}),
    ("octa3-ratio0", lambda: meshes.octa_sphere(3), 1000.0, {"near_ratio": 0.0}),
    ("ico3", lambda: meshes.icosphere(3), 1000.0, {}),
    ("tetra", lambda: meshes.scalene_tetra(), 400.0, {}),
    ("horn12", lambda: meshes.exponential_horn(0.012), 4000.0, {}),
    ("thin-box", lambda: _box(0.02, 0.02, 0.0002), 3000.0, {}),          # projection fans + all-pairs mode
    ("ico2-high-kh", lambda: meshes.icosphere(2), 2400.0, {"depth": 4}),
]


@needs_ref
@pytest.mark.parametrize("name,make,frequency,opts", PLAN_CASES, ids=[c[0] for c in PLAN_CASES])
def test_plan_bit_identical_to_reference(name, make, frequency, opts):
    mesh = make()
    k = 2 * math.pi * frequency / 343.0
    tol = opts.get("tolerance", 1e-4)
    ratio = opts.get("near_ratio", 1.0)
    depth = opts.get("depth", 6)
    ref = ob.ref_plan(mesh, tolerance=tol, near_ratio=ratio, depth=depth, k_max=k)
    num = bem.BemNumerics(near_relative_tolerance=tol, near_distance_ratio=ratio, near_maximum_depth=depth)
    mine = bem.interaction_plan(mesh, num, k)
    for key in ("touching", "near", "patches", "whole", "near_offsets", "near_trials"):
        assert np.array_equal(mine[key], ref[key]), key
    for key in ("candidate_pairs", "maximum_depth", "maximum_estimated_error"):
        assert mine[key] == ref[key], key


@needs_ref
def test_plan_covers_projection_and_all_pairs_paths():
    mesh = _box(0.02, 0.02, 0.0002)
    plan = bem.interaction_plan(mesh, bem.BemNumerics(), 2 * math.pi * 3000 / 343)
    assert len(plan["whole"]) == 12 * 11 // 2        # all-pairs mode (sliver side panels)
    assert np.any(plan["patches"][:, 20] == 1)       # projected fans for the 0.2 mm gap


def test_plan_is_independent_of_thread_count():
    mesh = meshes.icosphere(3)
    k = 2 * math.pi * 1500 / 343
    a = bem.interaction_plan(mesh, bem.BemNumerics(plan_threads=1), k)
    b = bem.interaction_plan(mesh, bem.BemNumerics(plan_threads=8), k)
    for key in ("touching", "near", "patches", "near_offsets", "near_trials"):
        assert np.array_equal(a[key], b[key])


@pytest.mark.parametrize("field,value", [("quadrature_order", 3), ("singular_order", 1), ("singular_order", 21),
                                         ("near_relative_tolerance", 0.0), ("near_maximum_depth", 17)])
def test_plan_rejects_invalid_settings(field, value):
    num = bem.BemNumerics()
    setattr(num, field, value)
    with pytest.raises(ValueError):
        bem.interaction_plan(meshes.octa_sphere(1), num, 10.0)


def test_plan_rejects_degenerate_panels():
    V = np.array([(0, 0, 0), (1, 0, 0), (2, 0, 0), (0, 1, 0)], float)
    mesh = bem.SurfaceMesh(V, np.array([(0, 1, 2), (0, 1, 3)], np.uint32), np.ones(2, np.uint32), np.ones(2, np.uint8))
    with pytest.raises(ValueError):
        bem.interaction_plan(mesh, bem.BemNumerics(), 1.0)


def _closed_oriented(mesh):
    edges = {}
    for t in mesh.triangles:
        for l in range(3):
            e = (int(t[l]), int(t[(l + 1) % 3]))
            edges[e] = edges.get(e, 0) + 1
    paired = all(c == 1 and (b, a) in edges for (a, b), c in edges.items())
    V = mesh.vertices
    volume = sum(np.dot(V[a], np.cross(V[b], V[c])) for a, b, c in mesh.triangles) / 6
    return paired, volume


def _min_angle(mesh):
    V = mesh.vertices
    worst = 180.0
    for t in mesh.triangles:
        P = V[list(t)]
        for i in range(3):
            u, w = P[(i + 1) % 3] - P[i], P[(i + 2) % 3] - P[i]
            worst = min(worst, math.degrees(math.acos(np.clip(u @ w / np.linalg.norm(u) / np.linalg.norm(w), -1, 1))))
    return worst


@pytest.mark.parametrize("make,nv,nt", [(lambda: meshes.icosphere(3), 642, 1280), (lambda: meshes.octa_sphere(3), 258, 512),
                                        (lambda: meshes.scalene_tetra(), 4, 4)])
def test_reference_meshes(make, nv, nt):
    mesh = make()
    assert (len(mesh.vertices), len(mesh.triangles)) == (nv, nt)
    paired, volume = _closed_oriented(mesh)
    assert paired and volume > 0


@pytest.mark.parametrize("edge", [0.02, 0.012, 0.0052])
def test_horn_mesh_closed_oriented_and_shape_qualified(edge):
    mesh = meshes.exponential_horn(edge)
    paired, volume = _closed_oriented(mesh)
    assert paired and volume > 0
    assert _min_angle(mesh) >= 20.0  # keeps the planner out of its all-pairs mode
    assert set(np.unique(mesh.face_ids).tolist()) == {1, 2}
    assert np.all(mesh.boundary[mesh.face_ids == 1] == bem.DRIVEN)


@pytest.mark.parametrize("edge", [0.01, 0.0045])
def test_flanged_horn_mesh_closed_oriented_and_shape_qualified(edge):
    """The horn on a 0.6 m square flange (5 mm slab in the mouth plane): closed,
    consistently oriented outward, angles >= 20 degrees and, at the 4.5 mm size
    (~118k triangles / ~59k nodes), every edge short enough for k h <= 2 at 16 kHz."""
    mesh = meshes.exponential_horn(edge, flange=0.6)
    T = mesh.triangles.astype(np.int64)
    V = mesh.vertices
    directed = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    keys = directed[:, 0] * len(V) + directed[:, 1]
    assert len(np.unique(keys)) == len(keys)  # no directed edge twice
    assert np.all(np.isin(directed[:, 1] * len(V) + directed[:, 0], keys))  # every edge paired, opposite direction
    volume = np.einsum("ij,ij->i", V[T[:, 0]], np.cross(V[T[:, 1]], V[T[:, 2]])).sum() / 6
    assert volume > 0
    P = V[T]
    cos = []
    for i in range(3):
        u, w = P[:, (i + 1) % 3] - P[:, i], P[:, (i + 2) % 3] - P[:, i]
        cos.append((u * w).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(w, axis=1))
    assert np.degrees(np.arccos(np.clip(np.max(cos), -1, 1))) >= 20.0
    assert np.abs(V[:, 1]).max() == pytest.approx(0.3) and np.abs(V[:, 2]).max() == pytest.approx(0.3)
    assert np.all(mesh.boundary[mesh.face_ids == 1] == bem.DRIVEN)
    if edge == 0.0045:
        longest = max(np.linalg.norm(P[:, i] - P[:, (i + 1) % 3], axis=1).max() for i in range(3))
        assert 2 * math.pi * 16000 / 343 * longest <= 2.0
        assert 110_000 <= len(T) <= 125_000


def test_cpp_facade_compiles_against_c_abi(tmp_path):
    """The header-only C++ drop-in (include/hornsculpt_amd/bem.hpp) compiles and links."""
    import subprocess
    src = tmp_path / "caller.cpp"
    src.write_text(
        "#include <hornsculpt_amd/bem.hpp>\n"
        "int main(int argc, char**) {\n"
        "  if (argc > 5) { hornsculpt::SurfaceMesh m; hornsculpt::CudaBemSolver s(m);\n"
        "    hornsculpt::BemRequest r; r.frequencies = {1000.0}; auto out = s.solve(r); return (int)out.triangleCount; }\n"
        "  return 0; }\n")
    exe = tmp_path / "caller"
    proc = subprocess.run(["g++", "-std=c++20", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
                           str(ROOT / "hornsculpt_amd" / "libhsbem.so"), f"-Wl,-rpath,{ROOT / 'hornsculpt_amd'}"],
                          capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr


def test_frequency_schedule_matches_reference_tests():
    # realtime_simulation_tests.cpp:3639-3647: coarse-to-fine with high-gap priority, a permutation
    assert bem.frequency_schedule(9) == [8, 0, 4, 6, 2, 7, 5, 3, 1]
    for count in range(83):
        assert sorted(bem.frequency_schedule(count)) == list(range(count))
    # by hand: (0,4) -> 2; equal gaps (0,2), (2,4): the higher first -> 3, then 1
    assert bem.frequency_schedule(5) == [4, 0, 2, 3, 1]
    assert bem.frequency_schedule(1) == [0] and bem.frequency_schedule(2) == [1, 0]


def test_progressive_requests_share_the_sweep_raster():
    freqs = [200.0 * 1.5 ** i for i in range(6)]
    req = bem.BemRequest(frequencies=freqs, excitations=[bem.BemExcitation("p", [bem.FaceVelocity(1, 1.0)])])
    prog = bem.progressive_requests(req)
    assert [i for i, _ in prog] == bem.frequency_schedule(6)
    for i, r in prog:
        assert r.frequencies == [freqs[i]] and r.interaction_plan_frequencies == freqs
    req.interaction_plan_frequencies = freqs + [9000.0]
    assert all(r.interaction_plan_frequencies == freqs + [9000.0] for _, r in bem.progressive_requests(req))


def test_tuning_report_before_any_tuning():
    """No launch-shape decision without a device: the report is an empty JSON array; a
    short buffer receives a truncated, terminated copy and `needed` the full size."""
    import ctypes
    assert bem.tuning_report() == []
    lib = bem.library()
    needed = ctypes.c_size_t(0)
    buf = ctypes.create_string_buffer(2)
    assert lib.hsbem_tuning_report(buf, 2, ctypes.byref(needed)) == 0
    assert needed.value == 3 and buf.value == b"["
    assert lib.hsbem_tuning_report(None, 0, None) != 0  # null buffer refused
