"""Synthetic closed meshes for the benchmark configurations and tests.

* ``octa_sphere(level)``: octahedron refined ``level`` times with midpoints
  projected to r=0.1 m, 8*4^level outward triangles (the construction of the
  reference fixture tests/fixture_geometry.hpp:11-51; 512 triangles at level 3).
* ``icosphere(level)``: icosahedron refined the same way (642 vertices /
  1280 triangles at level 3).
* ``scalene_tetra()``: the closed outward tetrahedron of
  tests/bem_singular_port_tests.cpp:71-81 (faces 50/51 driven, 52/53 wall).
* ``exponential_horn(...)``: a closed, watertight horn r(x)=r_t exp(m x) with a
  wall of given thickness, closed rim annulus, rigid rear closure and a driven
  throat disk; optional square flange at the mouth plane. Structured rings of
  nearly isotropic triangles with edge ~h (all angles >= 20 degrees).
All meshes are returned as ``SurfaceMesh`` (vertices in metres, outward winding).
"""
from __future__ import annotations

import math

import numpy as np

from .bem import DRIVEN, WALL, SurfaceMesh


def _refine_sphere(vertices, triangles, levels, radius):
    verts = [np.asarray(v, float) for v in vertices]
    tris = [tuple(t) for t in triangles]
    for _ in range(levels):
        edges = {}
        new_tris = []

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key in edges:
                return edges[key]
            p = verts[i] + verts[j]
            p = p * (radius / math.sqrt(float(p @ p)))
            verts.append(p)
            edges[key] = len(verts) - 1
            return edges[key]

        for a, b, c in tris:
            ab, bc, ac = mid(a, b), mid(b, c), mid(a, c)
            new_tris += [(a, ab, ac), (ab, b, bc), (ac, bc, c), (ab, bc, ac)]
        tris = new_tris
    return np.array(verts), np.array(tris, dtype=np.uint32)


def octa_sphere(level: int, radius: float = 0.1, face_id: int = 1) -> SurfaceMesh:
    r = radius
    v = [(r, 0, 0), (0, r, 0), (-r, 0, 0), (0, -r, 0), (0, 0, r), (0, 0, -r)]
    t = [(4, 0, 1), (4, 1, 2), (4, 2, 3), (4, 3, 0), (5, 1, 0), (5, 2, 1), (5, 3, 2), (5, 0, 3)]
    V, T = _refine_sphere(v, t, level, radius)
    n = len(T)
    return SurfaceMesh(V, T, np.full(n, face_id, np.uint32), np.full(n, DRIVEN, np.uint8))


def icosphere(level: int, radius: float = 0.1, face_id: int = 1) -> SurfaceMesh:
    p = (1 + math.sqrt(5)) / 2
    raw = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p), (0, 1, -p),
           (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    s = radius / math.sqrt(1 + p * p)
    v = [(x * s, y * s, z * s) for x, y, z in raw]
    t = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    V, T = _refine_sphere(v, t, level, radius)
    n = len(T)
    return SurfaceMesh(V, T, np.full(n, face_id, np.uint32), np.full(n, DRIVEN, np.uint8))


def scalene_tetra() -> SurfaceMesh:
    V = np.array([(0, 0, 0), (.2, 0, 0), (.03, .14, 0), (.02, .025, .09)], float)
    T = np.array([(0, 2, 1), (0, 1, 3), (0, 3, 2), (1, 2, 3)], np.uint32)
    return SurfaceMesh(V, T, np.array([50, 51, 52, 53], np.uint32),
                       np.array([DRIVEN, DRIVEN, WALL, WALL], np.uint8))


# ----------------------------------------------------------------------------- horn


class _Builder:
    def __init__(self, multiple: int = 1):
        self.vertices = []
        self.triangles = []
        self.faces = []
        self.bc = []
        self.multiple = multiple  # ring point counts are multiples of this; strips repeat per sector

    def count(self, r, h):
        n = max(8, int(round(2 * math.pi * r / h)))
        return self.multiple * -(-n // self.multiple)

    def ring(self, x, r, n, phase=0.0):
        start = len(self.vertices)
        for j in range(n):
            a = 2 * math.pi * (j + phase) / n
            self.vertices.append((x, r * math.cos(a), r * math.sin(a)))
        return list(range(start, start + n))

    def curve_ring(self, x, radius_of, n, phase=0.0):
        """Ring of n points at axial x, equally spaced in arc length along the
        star-shaped closed curve r = radius_of(angle) in the (y, z) plane."""
        a = np.linspace(0.0, 2 * math.pi, 4097)
        r = radius_of(a)
        y, z = r * np.cos(a), r * np.sin(a)
        s = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(y), np.diff(z)))])
        at = (np.arange(n) + phase) / n * s[-1]
        ys, zs = np.interp(at, s, y), np.interp(at, s, z)
        start = len(self.vertices)
        self.vertices.extend((x, float(u), float(w)) for u, w in zip(ys, zs))
        return list(range(start, start + n))

    def point(self, p):
        self.vertices.append(tuple(p))
        return len(self.vertices) - 1

    def tri(self, a, b, c, face, bc):
        self.triangles.append((a, b, c))
        self.faces.append(face)
        self.bc.append(bc)

    def strip(self, r0, r1, face, bc, flip=False):
        """Triangulate between two closed rings (counts may differ), always
        advancing along the shorter new diagonal (keeps the triangles near isotropic).
        When both counts are multiples of ``multiple`` = q, the moves of the first
        1/q sector are replayed in every sector, so the strip is invariant under a
        rotation by 2 pi / q about the axis (no rounding-decided ties)."""
        n0, n1 = len(r0), len(r1)
        q = self.multiple if self.multiple > 1 and n0 % self.multiple == 0 and n1 % self.multiple == 0 else 1
        n0s, n1s = n0 // q, n1 // q
        V = self.vertices
        # start r1 at the vertex angularly closest to r0[0]
        p0 = V[r0[0]]
        j0 = min(range(n1), key=lambda j: (V[r1[j]][1] - p0[1]) ** 2 + (V[r1[j]][2] - p0[2]) ** 2)
        r1 = r1[j0:] + r1[:j0]
        i = j = 0

        def d2(a, b):
            pa, pb = V[a], V[b]
            return (pa[0] - pb[0]) ** 2 + (pa[1] - pb[1]) ** 2 + (pa[2] - pb[2]) ** 2

        moves = []
        while i < n0s or j < n1s:
            top = i < n0s and (j >= n1s or d2(r0[(i + 1) % n0], r1[j % n1]) <= d2(r0[i % n0], r1[(j + 1) % n1]))
            moves.append(top)
            i, j = (i + 1, j) if top else (i, j + 1)
        for sector in range(q):
            i, j = sector * n0s, sector * n1s
            for top in moves:
                if top:
                    t = (r0[i % n0], r0[(i + 1) % n0], r1[j % n1])
                    i += 1
                else:
                    t = (r0[i % n0], r1[(j + 1) % n1], r1[j % n1])
                    j += 1
                if flip:
                    t = (t[0], t[2], t[1])
                self.tri(*t, face, bc)

    def mesh(self) -> SurfaceMesh:
        return SurfaceMesh(np.array(self.vertices, float), np.array(self.triangles, np.uint32),
                           np.array(self.faces, np.uint32), np.array(self.bc, np.uint8))


def _disk(b: _Builder, x, ring, r, h, face, bc, facing_negative):
    """Fill a planar disk bounded by ``ring`` (radius r at axial x) with concentric rings."""
    layers = max(1, int(round(r / h)))
    outer = ring
    for k in range(layers - 1, 0, -1):
        rr = r * k / layers
        inner = b.ring(x, rr, b.count(rr, h), phase=0.5 * (k % 2))
        b.strip(outer, inner, face, bc, flip=facing_negative)
        outer = inner
    c = b.point((x, 0.0, 0.0))
    n = len(outer)
    for j in range(n):
        t = (outer[j], outer[(j + 1) % n], c)
        if facing_negative:
            t = (t[0], t[2], t[1])
        b.tri(*t, face, bc)


def exponential_horn(h: float, throat_radius: float = 0.0127, mouth_radius: float = 0.150, length: float = 0.300,
                     wall: float = 0.005, throat_face: int = 1, wall_face: int = 2, flange: float = 0.0,
                     rear_length: float = None, ring_multiple: int = 1) -> SurfaceMesh:
    """Closed exponential horn along +x: r(x) = r_t exp(m x), m = ln(r_m/r_t)/L.

    Surfaces (outward normals into the air): inner flare (face ``wall_face``),
    driven throat disk at x=0 facing +x (face ``throat_face``), rim annulus at
    the mouth, outer flare offset by ``wall`` along the normal, and a rear closure
    disk at x=-rear_length. ``flange`` > 0 adds a square rigid baffle of that
    side length in the mouth plane (closed slab of thickness ``wall``).
    ``ring_multiple`` = q rounds every ring's point count up to a multiple of q
    and repeats the strips per sector: the unflanged mesh is then invariant under
    rotations by 2 pi / q about the axis (q = 4: the rocking modes cos / sin of
    the azimuth map onto each other).
    """
    m = math.log(mouth_radius / throat_radius) / length
    rear = wall if rear_length is None else rear_length
    b = _Builder(ring_multiple)

    def rin(x):
        return throat_radius * math.exp(m * x)

    # inner flare rings placed by arc length ~h (the flanged variant: equal arc
    # lengths, so no interval exceeds h)
    if flange > 0:
        grid = np.linspace(0.0, length, 20001)
        arc = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(grid), np.diff(throat_radius * np.exp(m * grid))))])
        xs = [float(v) for v in np.interp(np.linspace(0.0, arc[-1], max(1, int(math.ceil(arc[-1] / h))) + 1),
                                          arc, grid)]
        xs[-1] = length
    else:
        xs = [0.0]
        while xs[-1] < length:
            x = xs[-1]
            slope = m * rin(x)
            dx = h / math.sqrt(1 + slope * slope)
            xs.append(min(length, x + dx))
        if len(xs) > 2 and xs[-1] - xs[-2] < 0.3 * h:
            xs.pop(-2)
    inner = []
    for k, x in enumerate(xs):
        inner.append(b.ring(x, rin(x), b.count(rin(x), h), phase=0.5 * (k % 2)))
    for k in range(len(inner) - 1):
        b.strip(inner[k], inner[k + 1], wall_face, WALL, flip=True)
    # driven throat disk facing +x (into the horn air)
    _disk(b, 0.0, inner[0], rin(0.0), h, throat_face, DRIVEN, facing_negative=False)
    # outer surface: offset radius r + wall (approximately normal offset), from mouth back to the rear
    rout = lambda x: rin(x) + wall * math.sqrt(1 + (m * rin(x)) ** 2)
    if flange > 0:
        if 0.5 * flange <= rout(length) + 2 * h:
            raise ValueError("The flange must extend beyond the horn's outer wall.")
        xb = length - wall
        back = b.ring(xb, rout(xb), b.count(rout(xb), h), phase=0.25)
        _flange(b, inner[-1], rin(length), back, rout(xb), length, wall, 0.5 * flange, h, wall_face)
        outer = [back]
        grid = np.linspace(0.0, xb, 20001)
        arc = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(grid), np.diff([rout(v) for v in grid])))])
        count = max(1, int(math.ceil(arc[-1] / h)))
        outer_x = [float(v) for v in np.interp(np.linspace(0.0, arc[-1], count + 1), arc, grid)][count - 1::-1]
    else:
        mouth_out = b.ring(length, rout(length), b.count(rout(length), h), phase=0.25)
        # rim annulus between inner mouth ring and outer mouth ring (facing +x)
        ann = [inner[-1]]
        layers = max(1, int(round((rout(length) - rin(length)) / h)))
        for k in range(1, layers):
            rr = rin(length) + (rout(length) - rin(length)) * k / layers
            ann.append(b.ring(length, rr, b.count(rr, h), phase=0.5 * (k % 2)))
        ann.append(mouth_out)
        for k in range(len(ann) - 1):
            b.strip(ann[k], ann[k + 1], wall_face, WALL, flip=True)
        outer = [mouth_out]
        outer_x = list(reversed(xs[:-1]))
    for k, x in enumerate(outer_x):
        outer.append(b.ring(x, rout(x), b.count(rout(x), h), phase=0.25 + 0.5 * (k % 2)))
    # rear extension of the outer wall to the closure
    xr = -rear
    steps = max(1, int(round(rear / h)))
    for k in range(1, steps + 1):
        x = -rear * k / steps
        outer.append(b.ring(x, rout(0.0), b.count(rout(0.0), h), phase=0.5 * (k % 2)))
    for k in range(len(outer) - 1):
        b.strip(outer[k], outer[k + 1], wall_face, WALL, flip=True)
    _disk(b, xr, outer[-1], rout(0.0), h, wall_face, WALL, facing_negative=True)
    return b.mesh()


def _flange(b: _Builder, mouth, r_mouth, back, r_back, x, wall, half, h, face):
    """Square rigid baffle (side 2*half, thickness wall) in the mouth plane x: the
    front face (facing +x) from the inner mouth ring out to the square, the four
    side faces, and the back face (facing -x, at x - wall) from the square in to
    the outer horn wall's ring ``back``. Face rings morph from the circle to the
    square, r(t, a) = (1 - t) r0 + t half / max(|cos a|, |sin a|), spaced ~h in arc
    length; the square rings hold 8q points, so the four corners are vertices."""

    def square(a):
        return half / np.maximum(np.abs(np.cos(a)), np.abs(np.sin(a)))

    def morph(r0, t):
        return lambda a: (1 - t) * r0 + t * square(a)

    def count(r0, t):
        return max(8, int(round((2 * math.pi * r0 * (1 - t) + 8 * half * t) / h)))

    n_square = 8 * max(1, int(round(half / h)))

    def annulus(x_face, circle, r0, edge, outward):
        # rings between the circle (t = 0) and the square edge (t = 1)
        layers = max(1, int(round((math.sqrt(2) * half - r0) / h)))  # ~h at the corners
        rings = [circle]
        for k in range(1, layers):
            t = k / layers
            rings.append(b.curve_ring(x_face, morph(r0, t), count(r0, t), phase=0.5 * (k % 2)))
        rings.append(edge)
        if not outward:
            rings.reverse()
        for k in range(len(rings) - 1):
            b.strip(rings[k], rings[k + 1], face, WALL, flip=True)

    edge_front = b.curve_ring(x, square, n_square)
    annulus(x, mouth, r_mouth, edge_front, outward=True)
    steps = max(1, int(round(wall / h)))
    side = [edge_front] + [b.curve_ring(x - wall * k / steps, square, n_square) for k in range(1, steps + 1)]
    for k in range(steps):
        b.strip(side[k], side[k + 1], face, WALL, flip=True)
    annulus(x - wall, back, r_back, side[-1], outward=False)


def analytic_sphere_pressure(frequency: float, radius_observer: float, density: float = 1.21, c: float = 343.0,
                             a: float = 0.1, velocity: complex = 1.0) -> complex:
    """Pulsating sphere p(r) = rho c v (ika)/(1+ika) (a/r) exp(-ik(r-a)), exp(+iwt) convention."""
    k = 2 * math.pi * frequency / c
    ika = 1j * k * a
    return density * c * velocity * ika / (1 + ika) * (a / radius_observer) * np.exp(-1j * k * (radius_observer - a))
