"""Shared test fixtures: golden CSVs, analytic references and request builders."""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

from hornsculpt_amd import bem, meshes

GOLDEN = Path(__file__).resolve().parent / "golden"

SPHERE_OBSERVERS = np.array([(0, 0, .3), (1, 0, 0), (0, 0, 3)], float)
BEAT_FREQUENCIES = [100.0, 500.0, 1000.0, 1715.0]


def beat_sphere():
    """BEAT CPU Galerkin P1/DP0 observer pressures on the 512-triangle octa-sphere
    (tests/data/beat-sphere-512.csv of the reference): frequency -> 3 complex values."""
    out = {}
    for line in (GOLDEN / "beat-sphere-512.csv").read_text().splitlines():
        if not line or line.startswith("#"):
            continue
        v = [float(x) for x in line.split(",")]
        out[v[0]] = np.array([complex(v[1 + 2 * j], v[2 + 2 * j]) for j in range(3)])
    return out


def beat_tetra():
    """BEAT port impedances on the scalene tetra (singular order 12), keyed by (excitation, face)."""
    out = {}
    for line in (GOLDEN / "beat-scalene-tetra-singular12.csv").read_text().splitlines():
        if not line or line.startswith("#"):
            continue
        f = line.split(",")
        out[(f[0], int(f[1]))] = {"area": float(f[2]), "Q": complex(float(f[3]), float(f[4])),
                                  "p": complex(float(f[5]), float(f[6])), "Z": complex(float(f[7]), float(f[8]))}
    return out


TETRA_EXCITATIONS = {"uniform": {50: 1.0, 51: 1.0}, "antisymmetric": {50: 1.0, 51: -1.0}}


def sphere_request(frequencies, observers=SPHERE_OBSERVERS, numerics=None, density=1.21, c=343.0):
    return bem.BemRequest(medium=bem.AcousticMedium(density, c), frequencies=list(frequencies),
                          excitations=[bem.BemExcitation("unit", [bem.FaceVelocity(1, 1.0)]),
                                       bem.BemExcitation("phase", [bem.FaceVelocity(1, 0.6 * np.exp(0.71j))])],
                          observers=[tuple(o) for o in observers], numerics=numerics or bem.BemNumerics())


def analytic(frequency, radius, density=1.21, c=343.0):
    return meshes.analytic_sphere_pressure(frequency, radius, density, c)


def endpoint_degree(max_kh, requested=4):
    electrical = 4 if max_kh <= .25 else 6 if max_kh <= 1 else 8 if max_kh <= 2 else 12
    return max(electrical, requested)


def max_edge(mesh):
    V, T = mesh.vertices, mesh.triangles
    e = [np.linalg.norm(V[t[i]] - V[t[(i + 1) % 3]]) for t in T for i in range(3)]
    return max(e)


def oracle_observer_field(ob, mesh, k, p, q, observers, residual_tolerance=5e-5, near_tolerance=1e-4, depth=6,
                          far=False, origin=(0, 0, 0), R=1.0):
    """Observer pressures with the reference endpoint-degree convergence loop."""
    base = np.asarray(mesh.vertices[0], float)
    org = np.asarray(origin, float) - base
    if far:
        d = np.asarray(observers, float) - np.asarray(origin, float)
        obs = org + d / np.linalg.norm(d, axis=1)[:, None]
    else:
        obs = np.asarray(observers, float) - base
    seq = [4, 6, 8, 12, 14]
    deg = endpoint_degree(k * max_edge(mesh))
    tol = min(near_tolerance, .025 * residual_tolerance)
    lo = ob.oracle_field(mesh, k, deg, p, q, obs, far=far, origin=org, R=R, near_tolerance=tol, depth=depth)
    while True:
        nxt = seq[seq.index(deg) + 1]
        hi = ob.oracle_field(mesh, k, nxt, p, q, obs, far=far, origin=org, R=R, near_tolerance=tol, depth=depth)
        ok = all(np.linalg.norm(hi[:, e] - lo[:, e]) <= .025 * residual_tolerance * np.linalg.norm(hi[:, e])
                 for e in range(hi.shape[1]))
        lo = hi
        if ok or deg == 12:
            return lo
        deg = nxt


def flux(mesh, velocities_by_face, frequency, density):
    om = 2 * math.pi * frequency
    return np.array([-1j * density * om * velocities_by_face.get(int(f), 0.0) for f in mesh.face_ids])
