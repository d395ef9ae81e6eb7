"""Seeded horn-variant sweep (host): the std::mt19937_64 restatement against the
standard's known answer and the libstdc++ draws of the sweep seed
(tests/golden/sweep_draws_20261202.txt, written by make_sweep_draws.cpp), the
4-fold symmetric mesher and the four throat right-hand sides."""
import math
import os

import numpy as np

from hornsculpt_amd import meshes, variants

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sweep_draws_20261202.txt")


def test_mt19937_64_known_answer():
    # C++ [rand.predef]: the 10000th invocation of a default-constructed mt19937_64
    g = variants.Mt19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_sweep_draws_match_libstdcxx():
    want = np.loadtxt(GOLDEN)
    g = variants.Mt19937_64(20261202)
    got = np.array([[variants.uniform(g, 0.080, 0.200), variants.uniform(g, 0.150, 0.450),
                     variants.uniform(g, math.log(4000.0), math.log(60000.0))] for _ in range(len(want))])
    assert np.array_equal(got, want)
    vs = variants.horn_variants()
    assert [v.mouth_radius for v in vs] == list(want[:, 0])
    assert [v.length for v in vs] == list(want[:, 1])
    assert all(v.edge <= variants.MAX_EDGE for v in vs)


def _rotation_map(vertices, rotation):
    key = {tuple(np.round(v, 9)): i for i, v in enumerate(vertices)}
    return np.array([key.get(tuple(np.round(w, 9)), -1) for w in vertices @ rotation.T])


def test_symmetric_horn_is_closed_and_invariant_under_quarter_turns():
    mesh = meshes.exponential_horn(0.008, mouth_radius=0.12, length=0.25, ring_multiple=4)
    T = mesh.triangles.astype(np.int64)
    # closed and consistently oriented: every directed edge once, its reverse once
    directed = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    key = directed[:, 0] * len(mesh.vertices) + directed[:, 1]
    rev = directed[:, 1] * len(mesh.vertices) + directed[:, 0]
    assert len(np.unique(key)) == len(key) and np.array_equal(np.sort(key), np.sort(rev))
    R = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, 1.0, 0.0]])
    perm = _rotation_map(mesh.vertices, R)
    assert (perm >= 0).all()
    canon = lambda t: tuple(np.roll(t, -int(np.argmin(t))))
    assert {canon(t) for t in T.tolist()} == {canon(perm[t]) for t in T.tolist()}
    # the default mesher (no multiple) is unchanged: ring counts round(2 pi r / h)
    assert len(meshes.exponential_horn(0.0052).triangles) == 20742


def test_throat_excitations():
    mesh = meshes.exponential_horn(0.008, mouth_radius=0.1, length=0.2, ring_multiple=4)
    relabelled, (piston, cos_m, sin_m, ring) = variants.throat_excitations(mesh)
    throat = np.nonzero(mesh.face_ids == 1)[0]
    assert len(piston.faces) == len(throat) and set(relabelled.face_ids[throat]) == {f.face_id for f in piston.faces}
    assert all(abs(f.normal_velocity - 1) == 0 for f in piston.faces)
    # cos^2 + sin^2 = 1 per triangle; the rocking modes sum to zero over the symmetric disk
    c = np.array([f.normal_velocity for f in cos_m.faces])
    s = np.array([f.normal_velocity for f in sin_m.faces])
    assert np.allclose(np.abs(c) ** 2 + np.abs(s) ** 2, 1.0) and abs(c.sum()) < 1e-12 and abs(s.sum()) < 1e-12
    assert 0 < len(ring.faces) < len(piston.faces)
    assert np.array_equal(np.delete(relabelled.face_ids, throat), np.delete(mesh.face_ids, throat))


def test_sweep_frequencies():
    f = variants.sweep_frequencies()
    assert len(f) == 64 and f[0] == 100.0 and abs(f[-1] - 20000.0) < 1e-9
    assert np.allclose(np.diff(np.log(f)), math.log(200.0) / 63)
