"""Device near operator of the FMM path against the dense assembly.

Every planned near entry (i, j) of A_near receives the contributions of all
panel pairs (t ∋ i, s ∋ j), so it must equal the dense A entry; B_near(i, face)
sums the dense B columns of the face's panels in the near flux leaves of i's
leaf. With the near pairs' patch estimates on host threads (the dense plan's), the
tolerance is 1e-12 relative to max |A| (FP64, different summation order); the device
SpMV matches a numpy CSR product to 1e-13. With the device estimates (the default for
the fast multipole form), a pair whose leaves tie in error may refine another leaf
first and end on another valid patch set, so entries agree to the near quadrature's
accuracy (1e-8 of max |A| here).
"""
import math

import numpy as np
import pytest

from hornsculpt_amd import bem, meshes

pytestmark = pytest.mark.gpu


def _banded(mesh):
    V, T = mesh.vertices, mesh.triangles
    z = V[T].mean(axis=1)[:, 2]
    span = np.ptp(V[:, 2])
    faces = np.where(z > 0.15 * span, 1, np.where(z < -0.15 * span, 2, 3)).astype(np.uint32)
    return bem.SurfaceMesh(V, T, faces, np.where(faces == 3, bem.WALL, bem.DRIVEN).astype(np.uint8))


CASES = [
    ("horn20-throat", lambda: meshes.exponential_horn(0.02), [1], 2000.0, 16),
    ("ico3-two-faces", lambda: _banded(meshes.icosphere(3)), [2, 1], 1500.0, 8),
]


@pytest.mark.parametrize("estimator,bar", [("host", 1e-12), ("device", 1e-8)])
@pytest.mark.parametrize("name,make,faces,f,leaf", CASES, ids=[c[0] for c in CASES])
def test_near_operator_matches_dense_assembly(monkeypatch, name, make, faces, f, leaf, estimator, bar):
    monkeypatch.setenv("HSBEM_PLAN_ESTIMATOR", estimator)
    mesh = make()
    k = 2 * math.pi * f / 343.0
    num = bem.BemNumerics()
    solver = bem.HipBemSolver(mesh)
    A, B = solver.assemble(num, k, k)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(len(mesh.vertices)) + 1j * rng.standard_normal(len(mesh.vertices))
    op = solver.near_operator(num, k, k, faces, leaf_capacity=leaf, x=x)
    ro, co, va = op["row_offsets"], op["columns"], op["values"]
    scale = np.max(np.abs(A))
    rows = np.repeat(np.arange(len(ro) - 1), np.diff(ro))
    err_a = np.max(np.abs(va - A[rows, co])) / scale
    # the near pattern is a strict subset of the dense one (there is a far part)
    assert len(co) < A.size
    # B: near flux panels of each row's leaf, on the solver's recentred coordinates
    shifted = bem.SurfaceMesh(mesh.vertices - mesh.vertices[0], mesh.triangles, mesh.face_ids, mesh.boundary)
    tree = bem.parse_support_tree(bem.support_tree(shifted, leaf_capacity=leaf))
    nodes = tree["nodes"]
    near_panels = {}
    for tgt, src in tree["pf_near"]:
        first, count = int(nodes[src, 19]), int(nodes[src, 20])
        near_panels.setdefault(int(tgt), []).extend(tree["flux_order"][first:first + count].tolist())
    fro, fco, fva = op["flux_offsets"], op["flux_columns"], op["flux_values"]
    err_b = 0.0
    for i in range(len(mesh.vertices)):
        panels = np.array(near_panels.get(int(tree["pressure_leaf"][i]), []), dtype=np.int64)
        for e in range(fro[i], fro[i + 1]):
            face = faces[fco[e]]
            sel = panels[mesh.face_ids[panels] == face] if len(panels) else panels
            want = B[i, sel].sum()
            err_b = max(err_b, abs(fva[e] - want) / scale)
    y = np.zeros(len(mesh.vertices), complex)
    np.add.at(y, rows, va * x[co])
    err_y = np.max(np.abs(op["applied"] - y)) / np.max(np.abs(y))
    print(f"{name} ({estimator} estimates): nnz {len(co)} of {A.size}, jobs {op['counts']}, kernel "
          f"{op['kernel_ms']:.2f} ms, A {err_a:.2e} B {err_b:.2e} SpMV {err_y:.2e}")
    assert err_a <= bar and err_b <= bar and err_y <= 1e-13
