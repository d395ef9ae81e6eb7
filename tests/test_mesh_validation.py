"""Closed-exterior mesh validation on the host (hsbem_mesh_validate, the checks of
HipBemSolver(mesh) = CudaBemSolver(SurfaceMesh), bem.hpp:369, with the role of
validatedSpace / inspectMesh / validateExteriorComponents, cuda_solver.cu:782-789):
valid meshes pass; open, misoriented, inward, self-intersecting and nested
meshes are refused with a reason."""
import time

import numpy as np
import pytest

from hornsculpt_amd import bem, meshes


def _join(*parts, shift=None):
    vs, ts, fs, bs, offset = [], [], [], [], 0
    for i, m in enumerate(parts):
        v = np.array(m.vertices, dtype=np.float64)
        if shift is not None:
            v = v + np.asarray(shift[i])
        vs.append(v)
        ts.append(np.array(m.triangles) + offset)
        fs.append(np.array(m.face_ids))
        bs.append(np.array(m.boundary))
        offset += len(v)
    return bem.SurfaceMesh(np.vstack(vs), np.vstack(ts).astype(np.uint32), np.concatenate(fs).astype(np.uint32),
                           np.concatenate(bs).astype(np.uint8))


def _with(mesh, vertices=None, triangles=None):
    return bem.SurfaceMesh(mesh.vertices if vertices is None else vertices,
                           mesh.triangles if triangles is None else triangles, mesh.face_ids, mesh.boundary)


@pytest.mark.parametrize("mesh", [meshes.icosphere(3), meshes.octa_sphere(2), meshes.scalene_tetra(),
                                  meshes.exponential_horn(0.02), meshes.exponential_horn(0.02, flange=0.6)],
                         ids=["icosphere", "octa", "tetra", "horn", "flanged-horn"])
def test_valid_meshes_pass(mesh):
    bem.validate_mesh(mesh)


def test_separate_bodies_pass():
    s = meshes.icosphere(2)
    bem.validate_mesh(_join(s, s, shift=[(0, 0, 0), (0.25, 0, 0)]))


def test_open_mesh_refused():
    m = meshes.icosphere(2)
    with pytest.raises(ValueError, match="closed"):
        bem.validate_mesh(_with(m, triangles=np.array(m.triangles)[1:]))


def test_one_flipped_triangle_refused():
    m = meshes.icosphere(2)
    t = np.array(m.triangles)
    t[5] = t[5][::-1]
    with pytest.raises(ValueError, match="outward-oriented"):
        bem.validate_mesh(_with(m, triangles=t))


def test_inward_orientation_refused():
    m = meshes.icosphere(2)
    with pytest.raises(ValueError, match="non-positive enclosed volume"):
        bem.validate_mesh(_with(m, triangles=np.array(m.triangles)[:, ::-1].copy()))


def test_nested_components_refused():
    outer, inner = meshes.icosphere(2, radius=0.2), meshes.icosphere(2, radius=0.05)
    with pytest.raises(ValueError, match="nested"):
        bem.validate_mesh(_join(outer, inner))


def test_overlapping_bodies_refused():
    s = meshes.icosphere(2)
    with pytest.raises(ValueError, match="intersects itself"):
        bem.validate_mesh(_join(s, s, shift=[(0, 0, 0), (0.12, 0.01, 0.0)]))


def test_vertex_pushed_through_surface_refused():
    """One vertex of a sphere moved through the opposite side: the closed,
    outward-oriented surface now cuts through itself."""
    m = meshes.icosphere(2)
    v = np.array(m.vertices, dtype=np.float64)
    top = int(np.argmax(v[:, 2]))
    v[top] = (0.0, 0.0, -0.13)
    with pytest.raises(ValueError, match="intersects itself|non-positive"):
        bem.validate_mesh(_with(m, vertices=v))


def test_folded_coplanar_neighbours_refused():
    """Two triangles of a flat square folded onto each other across their common
    edge (a zero-volume fold) are found as an overlap."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.4, 0.4, 0.0], [0, 0, 1]], dtype=np.float64)
    # a tetra-like closed surface whose bottom face is split so that two of its pieces overlap
    t = np.array([[0, 2, 1], [0, 1, 4], [1, 2, 4], [2, 0, 4]], dtype=np.uint32)
    bem.validate_mesh(bem.SurfaceMesh(v[[0, 1, 2, 4]], np.array([[0, 2, 1], [0, 1, 3], [1, 2, 3], [2, 0, 3]], np.uint32),
                                      np.ones(4, np.uint32), np.zeros(4, np.uint8)))
    folded = np.array([[0, 2, 1], [0, 1, 3], [1, 3, 4], [1, 2, 4], [2, 0, 4], [0, 4, 3]], dtype=np.uint32)
    with pytest.raises(ValueError):
        bem.validate_mesh(bem.SurfaceMesh(v, folded, np.ones(len(folded), np.uint32), np.zeros(len(folded), np.uint8)))


def test_large_mesh_validation_is_fast():
    m = meshes.exponential_horn(0.0085, flange=0.6)
    t0 = time.perf_counter()
    bem.validate_mesh(m)
    dt = time.perf_counter() - t0
    print(f"{len(m.triangles)} triangles validated in {dt:.2f} s")
    assert dt < 10.0
