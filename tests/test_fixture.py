"""Benchmark fixture JSON reader (readFixture, tests/solver_benchmark.cpp:81-175):
keys, legacy excitations, the DP0 consistency check and every refusal."""
import hashlib
import json

import numpy as np
import pytest

from hornsculpt_amd import bem, meshes


def _fixture(**over):
    m = meshes.icosphere(1)
    tags = [1 if c[2] > 0 else 2 for c in m.vertices[m.triangles].mean(axis=1)]
    root = {"phasor": "exp(+iwt)", "rho": 1.205, "c": 343.0, "vertices_m": m.vertices.tolist(),
            "triangles0": m.triangles.tolist(), "tags": tags, "driven_tags": [1],
            "observations_m": [[0.0, 0.0, 0.5], [0.3, 0.0, 0.0]], "frequencies_hz": [500.0, 1000.0],
            "excitations": [{"name": "top", "face_velocities": [{"face_id": 1, "real": 1.0, "imag": -0.5}]}]}
    root.update(over)
    return m, root


def test_reads_mesh_request_and_hash():
    m, root = _fixture()
    raw = json.dumps(root).encode()
    mesh, req, digest = bem.read_benchmark_fixture(raw)
    assert digest == hashlib.sha256(raw).hexdigest()
    assert np.array_equal(mesh.vertices, m.vertices) and np.array_equal(mesh.triangles, m.triangles)
    assert list(mesh.face_ids) == root["tags"]
    assert all((b == bem.DRIVEN) == (t == 1) for b, t in zip(mesh.boundary, root["tags"]))
    assert (req.medium.density, req.medium.sound_speed) == (1.205, 343.0)
    assert req.frequencies == [500.0, 1000.0] and req.observers == [(0.0, 0.0, 0.5), (0.3, 0.0, 0.0)]
    assert [(e.name, [(f.face_id, f.normal_velocity) for f in e.faces]) for e in req.excitations] == \
        [("top", [(1, 1 - 0.5j)])]


def test_legacy_velocity_gives_uniform_and_proportional_phase():
    _, root = _fixture(driven_tags=[], driven_velocity=[0.5, 0.25])
    del root["excitations"]
    mesh, req, _ = bem.read_benchmark_fixture(json.dumps(root))
    assert all(b == bem.DRIVEN for b in mesh.boundary)  # no driven tags: every tag is driven
    names = [e.name for e in req.excitations]
    assert names == ["uniform", "proportional-phase"]
    u, p = req.excitations
    assert [f.face_id for f in u.faces] == [1, 2]
    assert all(abs(fp.normal_velocity - 0.6 * np.exp(0.71j) * fu.normal_velocity) < 1e-15
               for fu, fp in zip(u.faces, p.faces))


def test_dp0_velocities_are_checked():
    m, root = _fixture()
    good = [[[1.0, -0.5] if t == 1 else [0.0, 0.0] for t in root["tags"]]]
    bem.read_benchmark_fixture(json.dumps(dict(root, normal_velocities_m_s=good)))
    bad = [[[1.0, -0.4] if t == 1 else [0.0, 0.0] for t in root["tags"]]]
    with pytest.raises(ValueError, match="does not reproduce"):
        bem.read_benchmark_fixture(json.dumps(dict(root, normal_velocities_m_s=bad)))
    # a wall triangle with a velocity: the excitation names the wall tag
    root2 = dict(root, excitations=[{"name": "w", "face_velocities": [{"face_id": 2, "real": 1.0, "imag": 0.0}]}])
    wall = [[[1.0, 0.0] if t == 2 else [0.0, 0.0] for t in root["tags"]]]
    with pytest.raises(ValueError, match="Nonzero Wall velocity"):
        bem.read_benchmark_fixture(json.dumps(dict(root2, normal_velocities_m_s=wall)))
    with pytest.raises(ValueError, match="DP0 RHS count"):
        bem.read_benchmark_fixture(json.dumps(dict(root, normal_velocities_m_s=good + good)))


@pytest.mark.parametrize("change,message", [
    ({"phasor": "exp(-iwt)"}, "positive-time phasor"),
    ({"tags": [1]}, "Triangle and tag counts differ"),
    ({"frequencies_hz": [0.0]}, "Frequency must be positive"),
    ({"excitations": []}, "no excitation"),
])
def test_refusals(change, message):
    _, root = _fixture(**change)
    with pytest.raises(ValueError, match=message):
        bem.read_benchmark_fixture(json.dumps(root))


def _without(key):
    _, root = _fixture()
    del root[key]
    return root


@pytest.mark.parametrize("root,message", [
    (_fixture(rho=float("nan"))[1], "invalid scalar"),
    (_fixture(c=True)[1], "invalid scalar"),  # a JSON bool is not a number
    (_without("rho"), "invalid scalar"),
    (_without("observations_m"), "Missing fixture array"),  # the reference requires the array
    (_without("vertices_m"), "Missing fixture array"),
    (_fixture(tags="1")[1], "Missing fixture array"),
    (_fixture(observations_m=[[0.0, 0.0]])[1], "three coordinates"),
    (_fixture(observations_m=[[0.0, float("inf"), 1.0]])[1], "Invalid position coordinate"),
    (_fixture(vertices_m=[[0.0, 0.0, "x"]])[1], "Invalid position coordinate"),
    (_fixture(frequencies_hz=[True])[1], "Frequency must be positive"),
    (_fixture(frequencies_hz=[-5.0])[1], "Frequency must be positive"),
    (_fixture(driven_tags=[1.5])[1], "Invalid unsigned fixture index"),
    (_fixture(driven_tags=[-1])[1], "Invalid unsigned fixture index"),
    (_fixture(excitations=[{"name": "x", "face_velocities": [{"face_id": 1, "real": 1.0}]}])[1],
     "invalid scalar"),
    (_fixture(excitations=[{"name": "x", "face_velocities": [{"face_id": 2.5, "real": 1.0, "imag": 0.0}]}])[1],
     "Invalid unsigned fixture index"),
    (_fixture(excitations=[{"name": "x"}])[1], "Missing fixture array"),
])
def test_field_refusals(root, message):
    with pytest.raises(ValueError, match=message):
        bem.read_benchmark_fixture(json.dumps(root))


def test_fractional_and_negative_tags_refused():
    for bad in (1.5, -2):
        _, root = _fixture()
        root["tags"][0] = bad
        with pytest.raises(ValueError, match="Invalid unsigned fixture index"):
            bem.read_benchmark_fixture(json.dumps(root))
    _, root = _fixture()
    root["triangles0"][2] = [0, 1.25, 2]
    with pytest.raises(ValueError, match="Invalid unsigned fixture index"):
        bem.read_benchmark_fixture(json.dumps(root))


def test_out_of_range_index_refused():
    _, root = _fixture()
    root["triangles0"][3] = [0, 1, 10 ** 6]
    with pytest.raises(ValueError, match="out of range"):
        bem.read_benchmark_fixture(json.dumps(root))


@pytest.mark.gpu
def test_legacy_fixture_excitations_solve_linearly():
    """The legacy fixture's second excitation is 0.6 e^{0.71 i} times the first on the
    same faces, so its solution is that multiple of the first one's."""
    _, root = _fixture(driven_velocity=[1.0, 0.0])
    del root["excitations"]
    mesh, req, _ = bem.read_benchmark_fixture(json.dumps(root))
    res = bem.HipBemSolver(mesh).solve(req)
    for point in res.frequencies:
        u, p = (e.boundary.pressure_trace for e in point.excitations)
        assert np.linalg.norm(p - 0.6 * np.exp(0.71j) * u) <= 1e-12 * np.linalg.norm(p)
        ou, op = (e.observer_pressure for e in point.excitations)
        assert np.linalg.norm(op - 0.6 * np.exp(0.71j) * ou) <= 1e-12 * np.linalg.norm(op)
