"""Seeded horn-variant sweeps and the throat excitations of a batched sweep.

A sweep draws horn variants from ``std::mt19937_64`` (restated here, with
libstdc++'s ``uniform_real_distribution<double>``: one 64-bit draw x per value,
u = double(x) / 2^64 clamped below 1, a + (b - a) u): per variant the mouth
radius r_m ~ U[80, 200] mm, the length L ~ U[150, 450] mm and a target triangle
count N ~ log-uniform [4k, 60k]; the edge length h reaching N (from the area
scaling of a coarse mesh) is lowered until the mesh's longest edge is at most
5.4 mm, so k h_max <= 2 at 20 kHz and the interaction planner stays out of its
all-pairs mode (interaction_plan.cpp:185-190). Meshes have 4-fold rotational
symmetry about the axis.

Each frequency carries four right-hand sides on the driven throat disk, whose
triangles get one face id each: the uniform piston (v = 1), the two rocking
modes v = cos(phi) and v = sin(phi) of the triangle centroid's azimuth about the
axis, and the outer ring (v = 1 on the triangles whose centroid lies beyond
0.7 r_t). Azimuth phi = atan2(z, y) for the horn along +x.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List

import numpy as np

from . import bem, meshes

_MASK = (1 << 64) - 1


class Mt19937_64:
    """std::mt19937_64 (w 64, n 312, m 156, r 31; the standard's parameters)."""

    def __init__(self, seed: int = 5489):
        self.state = [0] * 312
        self.state[0] = seed & _MASK
        for i in range(1, 312):
            prev = self.state[i - 1]
            self.state[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & _MASK
        self.index = 312

    def _twist(self):
        st = self.state
        for i in range(312):
            y = (st[i] & 0xFFFFFFFF80000000) | (st[(i + 1) % 312] & 0x7FFFFFFF)
            v = st[(i + 156) % 312] ^ (y >> 1)
            if y & 1:
                v ^= 0xB5026F5AA96619E9
            st[i] = v
        self.index = 0

    def __call__(self) -> int:
        if self.index >= 312:
            self._twist()
        y = self.state[self.index]
        self.index += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & _MASK


def uniform(engine: Mt19937_64, a: float, b: float) -> float:
    """std::uniform_real_distribution<double>(a, b)(engine) as libstdc++ computes it."""
    u = float(engine()) / 18446744073709551616.0
    if u >= 1.0:
        u = math.nextafter(1.0, 0.0)
    return u * (b - a) + a


@dataclasses.dataclass
class HornVariant:
    index: int
    mouth_radius: float
    length: float
    target_triangles: float
    edge: float

    def mesh(self) -> bem.SurfaceMesh:
        return meshes.exponential_horn(self.edge, mouth_radius=self.mouth_radius, length=self.length,
                                       ring_multiple=4)


MAX_EDGE = 0.0054


def horn_variants(count: int = 32, seed: int = 20261202, triangles=(4000.0, 60000.0),
                  max_edge: float = MAX_EDGE, targets=None) -> List[HornVariant]:
    """The sweep's variants: draws (r_m, L, N) per variant in that order. `targets`
    (one triangle count per variant) replaces the drawn N; the draw is still made so
    the shapes are those of the same seed."""
    g = Mt19937_64(seed)
    out = []
    for i in range(count):
        r_m = uniform(g, 0.080, 0.200)
        length = uniform(g, 0.150, 0.450)
        target = math.exp(uniform(g, math.log(triangles[0]), math.log(triangles[1])))
        if targets is not None:
            target = float(targets[i])
        coarse = 0.02
        n0 = len(meshes.exponential_horn(coarse, mouth_radius=r_m, length=length, ring_multiple=4).triangles)
        edge = coarse * math.sqrt(n0 / target)
        for _ in range(8):
            longest = maximum_edge(meshes.exponential_horn(edge, mouth_radius=r_m, length=length, ring_multiple=4))
            if longest <= max_edge:
                break
            edge *= 0.995 * max_edge / longest
        out.append(HornVariant(i, r_m, length, target, edge))
    return out


def maximum_edge(mesh: bem.SurfaceMesh) -> float:
    V, T = mesh.vertices, mesh.triangles
    return float(max(np.linalg.norm(V[T[:, a]] - V[T[:, b]], axis=1).max() for a, b in ((0, 1), (1, 2), (2, 0))))


def sweep_frequencies(count: int = 64, low: float = 100.0, high: float = 20000.0) -> List[float]:
    """count log-spaced frequencies low * (high/low)^(i/(count-1))."""
    return [low * (high / low) ** (i / (count - 1)) for i in range(count)]


THROAT_FACE_BASE = 1000


def throat_excitations(mesh: bem.SurfaceMesh, throat_radius: float = 0.0127, throat_face: int = 1):
    """(mesh with one face id per throat triangle, [piston, rocking cos, rocking sin,
    outer ring] excitations). Throat triangle j gets face id THROAT_FACE_BASE + j."""
    throat = np.nonzero(mesh.face_ids == throat_face)[0]
    if len(throat) == 0:
        raise ValueError("mesh has no throat triangles")
    faces = mesh.face_ids.copy()
    ids = THROAT_FACE_BASE + np.arange(len(throat))
    faces[throat] = ids
    c = mesh.vertices[mesh.triangles[throat]].mean(axis=1)
    phi = np.arctan2(c[:, 2], c[:, 1])
    radius = np.hypot(c[:, 1], c[:, 2])
    piston = bem.BemExcitation("piston", [bem.FaceVelocity(int(f), 1.0) for f in ids])
    cos_mode = bem.BemExcitation("rocking_cos", [bem.FaceVelocity(int(f), complex(math.cos(a)))
                                                 for f, a in zip(ids, phi)])
    sin_mode = bem.BemExcitation("rocking_sin", [bem.FaceVelocity(int(f), complex(math.sin(a)))
                                                 for f, a in zip(ids, phi)])
    ring = bem.BemExcitation("outer_ring", [bem.FaceVelocity(int(f), 1.0)
                                            for f, r in zip(ids, radius) if r > 0.7 * throat_radius])
    if not ring.faces:
        raise ValueError("no throat triangle lies beyond 0.7 r_t: the throat disk needs a finer edge")
    relabelled = bem.SurfaceMesh(mesh.vertices, mesh.triangles, faces.astype(np.uint32), mesh.boundary)
    return relabelled, [piston, cos_mode, sin_mode, ring]
