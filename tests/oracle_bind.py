"""ctypes bindings of the test oracle (oracle/liboracle.so) and of the reference
planner built from the reference sources (oracle/_ref/libref_galerkin.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle" / "liboracle.so"
REF = ROOT / "oracle" / "_ref" / "libref_galerkin.so"

_D = ctypes.POINTER(ctypes.c_double)
_U = ctypes.POINTER(ctypes.c_uint32)
_U8 = ctypes.POINTER(ctypes.c_uint8)


def _d(a):
    return a.ctypes.data_as(_D)


def _u(a):
    return a.ctypes.data_as(_U)


def oracle():
    if not ORACLE.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "liboracle.so"], check=True, capture_output=True)
    lib = ctypes.CDLL(str(ORACLE))
    lib.oracle_last_error.restype = ctypes.c_char_p
    return lib


def reference():
    """The reference planner library, or None where it is not built. Loaded on the
    first call only (test modules mark their tests by REF.exists(), so collecting
    them maps nothing into the process)."""
    if not REF.exists():
        return None
    lib = ctypes.CDLL(str(REF))
    lib.ref_last_error.restype = ctypes.c_char_p
    lib.ref_plan_create.restype = ctypes.c_void_p
    lib.ref_plan_create.argtypes = [ctypes.c_uint32, _D, ctypes.c_uint32, _U, ctypes.c_uint, ctypes.c_uint,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_uint, ctypes.c_double, ctypes.c_uint]
    lib.ref_plan_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
    lib.ref_plan_copy.argtypes = [ctypes.c_void_p, _U, _U, _D, _U8, _U, _U, _D]
    lib.ref_plan_destroy.argtypes = [ctypes.c_void_p]
    lib.ref_support_near_plan.argtypes = [ctypes.c_uint32, _D, ctypes.c_uint32, _U, _U, _U8, _U, ctypes.c_uint32,
                                          ctypes.c_uint, ctypes.c_uint, ctypes.c_double, ctypes.c_double, _D,
                                          ctypes.POINTER(ctypes.c_uint64)]
    lib.ref_support_expansion_plan.argtypes = [ctypes.c_uint32, _D, ctypes.c_uint32, _U, _U, _U8, _U, ctypes.c_uint32,
                                               ctypes.c_uint, ctypes.c_uint, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_double, _D, ctypes.POINTER(ctypes.c_uint64)]
    lib.ref_rotation_plan.argtypes = [ctypes.c_uint, _D, _D, _D, _D, ctypes.c_int, _D]
    lib.ref_rcr_queue_policy.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint64)]
    lib.ref_sphere_rule.argtypes = [ctypes.c_uint, _D, _D]
    lib.ref_spherical_harmonics.argtypes = [ctypes.c_uint, _D, _D]
    lib.ref_regular_radial_table.argtypes = [ctypes.c_uint, ctypes.c_double, _D, _D]
    lib.ref_angular_transfer.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_int, ctypes.c_double,
                                         _D, _D, _D]
    lib.ref_harmonic_spectrum.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_int, _D, _D]
    lib.ref_diagonal_translation.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_double, _D, _D]
    lib.ref_scaled_waves.argtypes = [ctypes.c_uint, ctypes.c_double, ctypes.c_double, _D, ctypes.c_int, _D]
    lib.ref_scaled_expansion_evaluate.argtypes = [ctypes.c_uint, ctypes.c_double, ctypes.c_double, _D, ctypes.c_int,
                                                  _D, _D]
    lib.ref_scaled_spectrum.argtypes = [ctypes.c_uint, ctypes.c_double, ctypes.c_double, ctypes.c_uint, ctypes.c_int,
                                        _D, _D]
    lib.ref_axial_plan.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_uint, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_uint, ctypes.c_double,
                                   ctypes.POINTER(ctypes.c_uint64), _D, _D, _D, _D]
    _I = ctypes.POINTER(ctypes.c_int)
    lib.ref_spl_db.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]
    lib.ref_normalized_volume_impedance.argtypes = [ctypes.c_double] * 5 + [_D]
    lib.ref_angular_level_db.argtypes = [_D, _D, ctypes.c_uint, ctypes.c_double, _D, _I]
    lib.ref_beamwidth.argtypes = [_D, _D, ctypes.c_uint, ctypes.c_double, _D, _I]
    lib.ref_cta2034_cuts.argtypes = [_D, _D, ctypes.c_uint, _D, _D, ctypes.c_uint, _D, _I]
    lib.ref_angular_frequency_response.argtypes = [_D, _U, _D, _D, ctypes.c_uint, ctypes.c_double, ctypes.c_int,
                                                   ctypes.c_double, _D, _I]
    lib.ref_acoustic_frame.argtypes = [_D, _D, _D, _D]
    lib.ref_support_tree.argtypes = [ctypes.c_uint32, _D, ctypes.c_uint32, _U, ctypes.c_uint, ctypes.c_uint,
                                     ctypes.c_double, ctypes.c_double, _D, ctypes.POINTER(ctypes.c_uint64)]
    return lib


def _mesh_args(mesh):
    V = np.ascontiguousarray(mesh.vertices, dtype=np.float64)
    T = np.ascontiguousarray(mesh.triangles, dtype=np.uint32)
    return V, T


def recentred(mesh):
    """Vertices shifted so vertex 0 is the origin (what the solver does internally)."""
    V = np.asarray(mesh.vertices, float)
    return V - V[0]


def _check(lib, rc, name):
    if rc != 0:
        err = lib.oracle_last_error() if hasattr(lib, "oracle_last_error") else lib.ref_last_error()
        raise RuntimeError(f"{name}: {err.decode()}")


def ref_plan(mesh, quadrature_order=4, singular_order=4, near_ratio=1.0, tolerance=1e-4, depth=6, k_max=0.0,
             threads=8, vertices=None):
    lib = reference()
    if lib is None:
        return None
    V = np.ascontiguousarray(recentred(mesh) if vertices is None else vertices)
    T = np.ascontiguousarray(mesh.triangles, dtype=np.uint32)
    h = lib.ref_plan_create(len(V), _d(V), len(T), _u(T), quadrature_order, singular_order, near_ratio, tolerance,
                            depth, k_max, threads)
    if not h:
        raise RuntimeError(lib.ref_last_error().decode())
    try:
        sizes = np.zeros(6, np.uint64)
        lib.ref_plan_sizes(h, sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        nt_, nn, npatch, nwhole = (int(x) for x in sizes[:4])
        touching = np.zeros((max(nt_, 1), 9), np.uint32)
        near = np.zeros((max(nn, 1), 5), np.uint32)
        patches = np.zeros((max(npatch, 1), 22))
        whole = np.zeros(max(nwhole, 1), np.uint8)
        offsets = np.zeros(len(T) + 1, np.uint32)
        trials = np.zeros(max(nn, 1), np.uint32)
        err = ctypes.c_double()
        lib.ref_plan_copy(h, _u(touching), _u(near), _d(patches), whole.ctypes.data_as(_U8), _u(offsets), _u(trials),
                          ctypes.byref(err))
        return {"touching": touching[:nt_], "near": near[:nn], "patches": patches[:npatch], "whole": whole[:nwhole],
                "near_offsets": offsets, "near_trials": trials[:nn], "candidate_pairs": int(sizes[4]),
                "maximum_depth": int(sizes[5]), "maximum_estimated_error": err.value}
    finally:
        lib.ref_plan_destroy(h)


def ref_support_tree(mesh, leaf_capacity=32, maximum_depth=20, separation=1.5, near_ratio=2.0):
    """Serialised reference makeSupportTree on the raw coordinates (layout of hsbem_support_tree_build)."""
    lib = reference()
    V, T = _mesh_args(mesh)
    size = ctypes.c_uint64(0)
    args = (len(V), _d(V), len(T), _u(T), leaf_capacity, maximum_depth, separation, near_ratio)
    _check(lib, lib.ref_support_tree(*args, None, ctypes.byref(size)), "ref_support_tree")
    blob = np.zeros(size.value)
    _check(lib, lib.ref_support_tree(*args, _d(blob), ctypes.byref(size)), "ref_support_tree")
    return blob


def ref_support_near_plan(mesh, active_faces, leaf_capacity=32, maximum_depth=20, separation=1.5, near_ratio=2.0):
    """Serialised reference makeSupportNearPlan + job iterator (layout of hsbem_support_near_plan_build)."""
    lib = reference()
    V, T = _mesh_args(mesh)
    F = np.ascontiguousarray(mesh.face_ids, dtype=np.uint32)
    D = np.ascontiguousarray(mesh.boundary, dtype=np.uint8)
    A = np.ascontiguousarray(active_faces, dtype=np.uint32)
    size = ctypes.c_uint64(0)
    args = (len(V), _d(V), len(T), _u(T), _u(F), D.ctypes.data_as(_U8), _u(A), len(A), leaf_capacity, maximum_depth,
            separation, near_ratio)
    _check(lib, lib.ref_support_near_plan(*args, None, ctypes.byref(size)), "ref_support_near_plan")
    blob = np.zeros(size.value)
    _check(lib, lib.ref_support_near_plan(*args, _d(blob), ctypes.byref(size)), "ref_support_near_plan")
    return blob


def ref_support_expansion_plan(mesh, active_faces, k, seed, leaf_capacity=32, maximum_depth=20, separation=1.5,
                               near_ratio=2.0):
    """Serialised reference makeSupportExpansionPlan (layout of hsbem_support_expansion_plan_build)."""
    lib = reference()
    V, T = _mesh_args(mesh)
    F = np.ascontiguousarray(mesh.face_ids, dtype=np.uint32)
    D = np.ascontiguousarray(mesh.boundary, dtype=np.uint8)
    A = np.ascontiguousarray(active_faces if len(active_faces) else [0], dtype=np.uint32)
    size = ctypes.c_uint64(0)
    args = (len(V), _d(V), len(T), _u(T), _u(F), D.ctypes.data_as(_U8), _u(A), len(active_faces), leaf_capacity,
            maximum_depth, separation, near_ratio, k, seed)
    _check(lib, lib.ref_support_expansion_plan(*args, None, ctypes.byref(size)), "ref_support_expansion_plan")
    blob = np.zeros(size.value)
    _check(lib, lib.ref_support_expansion_plan(*args, _d(blob), ctypes.byref(size)), "ref_support_expansion_plan")
    return blob


def ref_rotation(order, axis, coefficients=None, direction=1):
    """Reference Wigner small-d blocks, azimuth phases and (optionally) rotated coefficients."""
    lib = reference()
    nd = sum((2 * l + 1) ** 2 for l in range(order + 1))
    d, az = np.zeros(nd), np.zeros(2 * (2 * order + 1))
    ax = np.ascontiguousarray(axis, dtype=np.float64)
    out = None
    if coefficients is not None:
        c = np.ascontiguousarray(np.asarray(coefficients, np.complex128)).view(np.float64)
        out = np.zeros_like(c)
        rc = lib.ref_rotation_plan(order, _d(ax), _d(d), _d(az), _d(c), direction, _d(out))
    else:
        rc = lib.ref_rotation_plan(order, _d(ax), _d(d), _d(az), None, 0, None)
    _check(lib, rc, "ref_rotation_plan")
    return d, az.view(np.complex128), (out.view(np.complex128) if out is not None else None)


def ref_sphere_rule(order):
    lib = reference()
    n = 2 * (order + 1) ** 2
    d, w = np.zeros(3 * n), np.zeros(n)
    _check(lib, lib.ref_sphere_rule(order, _d(d), _d(w)), "ref_sphere_rule")
    return d.reshape(n, 3), w


def ref_spherical_harmonics(order, direction):
    lib = reference()
    out = np.zeros(2 * (order + 1) ** 2)
    dd = np.ascontiguousarray(direction, dtype=np.float64)
    _check(lib, lib.ref_spherical_harmonics(order, _d(dd), _d(out)), "ref_spherical_harmonics")
    return out.view(np.complex128)


def ref_regular_radial_table(order, x):
    lib = reference()
    v, dv = np.zeros(order + 1), np.zeros(order + 1)
    _check(lib, lib.ref_regular_radial_table(order, float(x), _d(v), _d(dv)), "ref_regular_radial_table")
    return v, dv


def ref_angular_transfer(values, source_order, target_order, bandlimit, mode=0, k=0.0, shift=None):
    lib = reference()
    v = np.ascontiguousarray(np.asarray(values, np.complex128)).view(np.float64)
    out = np.zeros(4 * (target_order + 1) ** 2)
    sh = None if shift is None else np.ascontiguousarray(shift, dtype=np.float64)
    _check(lib, lib.ref_angular_transfer(source_order, target_order, bandlimit, mode, float(k),
                                         None if sh is None else _d(sh), _d(v), _d(out)), "ref_angular_transfer")
    return out.view(np.complex128)


def ref_harmonic_spectrum(values, order, rule_order, direction):
    lib = reference()
    v = np.ascontiguousarray(np.asarray(values, np.complex128)).view(np.float64)
    n = 2 * (rule_order + 1) ** 2 if direction > 0 else (order + 1) ** 2
    out = np.zeros(2 * n)
    _check(lib, lib.ref_harmonic_spectrum(order, rule_order, direction, _d(v), _d(out)), "ref_harmonic_spectrum")
    return out.view(np.complex128)


def ref_diagonal_translation(rule_order, expansion_order, k, r):
    lib = reference()
    rr = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(4 * (rule_order + 1) ** 2)
    _check(lib, lib.ref_diagonal_translation(rule_order, expansion_order, float(k), _d(rr), _d(out)),
           "ref_diagonal_translation")
    return out.view(np.complex128)


def ref_scaled_waves(order, k, radius, point, outgoing=False):
    lib = reference()
    out = np.zeros(8 * (order + 1) ** 2)
    p = np.ascontiguousarray(point, dtype=np.float64)
    _check(lib, lib.ref_scaled_waves(order, float(k), float(radius), _d(p), int(bool(outgoing)), _d(out)),
           "ref_scaled_waves")
    return out


def ref_scaled_expansion_evaluate(coefficients, order, radius, k, point, local):
    lib = reference()
    c = np.ascontiguousarray(np.asarray(coefficients, np.complex128)).view(np.float64)
    out = np.zeros(8)
    p = np.ascontiguousarray(point, dtype=np.float64)
    _check(lib, lib.ref_scaled_expansion_evaluate(order, float(radius), float(k), _d(c), int(bool(local)), _d(p),
                                                  _d(out)), "ref_scaled_expansion_evaluate")
    return out


def ref_scaled_spectrum(values, order, radius, k, rule_order, direction):
    lib = reference()
    v = np.ascontiguousarray(np.asarray(values, np.complex128)).view(np.float64)
    n = 2 * (rule_order + 1) ** 2 if direction > 0 else (order + 1) ** 2
    out = np.zeros(2 * n)
    _check(lib, lib.ref_scaled_spectrum(order, float(radius), float(k), rule_order, direction, _d(v), _d(out)),
           "ref_scaled_spectrum")
    return out.view(np.complex128)


def ref_axial_plan(kind, ps, pt, k, a_s, a_t, d, polar, projection=0.0, source=None):
    lib = reference()
    count = ctypes.c_uint64(0)
    args = (kind, ps, pt, float(k), float(a_s), float(a_t), float(d), polar, float(projection))
    _check(lib, lib.ref_axial_plan(*args, ctypes.byref(count), None, None, None, None), "ref_axial_plan")
    coeffs, probe = np.zeros(2 * count.value), np.zeros(1)
    src = None if source is None else np.ascontiguousarray(np.asarray(source, np.complex128)).view(np.float64)
    applied = np.zeros(2 * (pt + 1) ** 2)
    _check(lib, lib.ref_axial_plan(*args, ctypes.byref(count), _d(coeffs), _d(probe),
                                   None if src is None else _d(src), _d(applied)), "ref_axial_plan")
    return coeffs.view(np.complex128), probe[0], (applied.view(np.complex128) if src is not None else None)


def _ref_samples(samples):
    a = np.ascontiguousarray([float(x[0]) for x in samples] or [0.0])
    l = np.ascontiguousarray([float(x[1]) for x in samples] or [0.0])
    return a, l, len(samples)


def ref_spl_db(p, reference_pressure=20e-6):
    lib, out = reference(), np.zeros(1)
    _check(lib, lib.ref_spl_db(complex(p).real, complex(p).imag, reference_pressure, _d(out)), "ref_spl_db")
    return out[0]


def ref_normalized_volume_impedance(z, area, rho, c):
    lib, out = reference(), np.zeros(2)
    _check(lib, lib.ref_normalized_volume_impedance(complex(z).real, complex(z).imag, area, rho, c, _d(out)),
           "ref_normalized_volume_impedance")
    return complex(out[0], out[1])


def ref_angular_level_db(samples, angle):
    lib = reference()
    a, l, n = _ref_samples(samples)
    out, present = np.zeros(1), np.zeros(1, np.int32)
    _check(lib, lib.ref_angular_level_db(_d(a), _d(l), n, float(angle), _d(out),
                                         present.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "ref_angular_level_db")
    return float(out[0]) if present[0] else None


def ref_beamwidth(samples, attenuation=6.0):
    lib = reference()
    a, l, n = _ref_samples(samples)
    out, present = np.zeros(3), np.zeros(1, np.int32)
    _check(lib, lib.ref_beamwidth(_d(a), _d(l), n, float(attenuation), _d(out),
                                  present.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "ref_beamwidth")
    return tuple(out.tolist()) if present[0] else None


def ref_cta2034_cuts(h, v):
    lib = reference()
    ha, hl, hn = _ref_samples(h)
    va, vl, vn = _ref_samples(v)
    out, present = np.zeros(7), np.zeros(7, np.int32)
    _check(lib, lib.ref_cta2034_cuts(_d(ha), _d(hl), hn, _d(va), _d(vl), vn, _d(out),
                                     present.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "ref_cta2034_cuts")
    return [float(x) if p else None for x, p in zip(out, present)]


def ref_angular_frequency_response(rows, angle, relative=False, reference_angle=0.0):
    lib = reference()
    freqs = np.ascontiguousarray([float(r[0]) for r in rows] or [0.0])
    offsets, angles, levels = [0], [], []
    for _, smp in rows:
        angles += [float(x[0]) for x in smp]
        levels += [float(x[1]) for x in smp]
        offsets.append(len(angles))
    off = np.ascontiguousarray(offsets, dtype=np.uint32)
    a, l = np.ascontiguousarray(angles or [0.0]), np.ascontiguousarray(levels or [0.0])
    out, present = np.zeros(max(len(rows), 1)), np.zeros(max(len(rows), 1), np.int32)
    _check(lib, lib.ref_angular_frequency_response(_d(freqs), _u(off), _d(a), _d(l), len(rows), float(angle),
                                                   int(bool(relative)), float(reference_angle), _d(out),
                                                   present.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
           "ref_angular_frequency_response")
    return [float(x) if p else None for x, p in zip(out[:len(rows)], present[:len(rows)])]


def ref_acoustic_frame(origin, forward, horizontal=None):
    lib = reference()
    o, f = np.ascontiguousarray(origin, dtype=np.float64), np.ascontiguousarray(forward, dtype=np.float64)
    h = None if horizontal is None else np.ascontiguousarray(horizontal, dtype=np.float64)
    out = np.zeros(12)
    _check(lib, lib.ref_acoustic_frame(_d(o), _d(f), None if h is None else _d(h), _d(out)), "ref_acoustic_frame")
    return out


def ref_rcr_queue_policy(source_order, target_order, reserved):
    lib = reference()
    out = np.zeros(3, np.uint64)
    _check(lib, lib.ref_rcr_queue_policy(source_order, target_order, reserved,
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))), "ref_rcr_queue_policy")
    return out


def oracle_assemble(mesh, plan, k, tolerance=1e-4, regular_order=4, singular_order=4, tests=None, threads=8,
                    vertices=None):
    """Row-scaled A (nv x nv) and B (nv x nt) of the FP64 oracle for one wavenumber."""
    lib = oracle()
    V = np.ascontiguousarray(recentred(mesh) if vertices is None else vertices)
    T = np.ascontiguousarray(mesh.triangles, dtype=np.uint32)
    nv, nt = len(V), len(T)
    A = np.zeros(2 * nv * nv)
    B = np.zeros(2 * nv * nt)
    touching = np.ascontiguousarray(plan["touching"], dtype=np.uint32)
    near = np.ascontiguousarray(plan["near"], dtype=np.uint32)
    patches = np.ascontiguousarray(plan["patches"], dtype=np.float64)
    whole = plan["whole"]
    sub = None if tests is None else np.ascontiguousarray(tests, dtype=np.uint32)
    lib.oracle_assemble.argtypes = [ctypes.c_uint32, _D, ctypes.c_uint32, _U, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_uint, ctypes.c_uint, _U, ctypes.c_uint64, _U, ctypes.c_uint64, _D, _U8, _U,
                                    ctypes.c_uint64, ctypes.c_uint, _D, _D]
    rc = lib.oracle_assemble(nv, _d(V), nt, _u(T), k, tolerance, regular_order, singular_order,
                             _u(touching), len(plan["touching"]), _u(near), len(plan["near"]),
                             _d(patches) if len(plan["patches"]) else None,
                             whole.ctypes.data_as(_U8) if len(whole) else None,
                             _u(sub) if sub is not None else None, 0 if sub is None else len(sub), threads, _d(A), _d(B))
    _check(lib, rc, "oracle_assemble")
    return (A.view(np.complex128).reshape(nv, nv).T.copy(), B.view(np.complex128).reshape(nt, nv).T.copy())


def oracle_assemble_rows(mesh, plan, k, rows, tolerance=1e-4, regular_order=4, singular_order=4, threads=8,
                         vertices=None):
    """Rows ``rows`` of the oracle's row-scaled A (len(rows) x nv) and B (len(rows) x nt),
    integrating only the test panels that touch those rows."""
    lib = oracle()
    V = np.ascontiguousarray(recentred(mesh) if vertices is None else vertices)
    T = np.ascontiguousarray(mesh.triangles, dtype=np.uint32)
    nv, nt = len(V), len(T)
    R = np.ascontiguousarray(rows, dtype=np.uint32)
    m = len(R)
    A = np.zeros(2 * m * nv)
    B = np.zeros(2 * m * nt)
    touching = np.ascontiguousarray(plan["touching"], dtype=np.uint32)
    near = np.ascontiguousarray(plan["near"], dtype=np.uint32)
    patches = np.ascontiguousarray(plan["patches"], dtype=np.float64)
    whole = plan["whole"]
    lib.oracle_assemble_rows.argtypes = [ctypes.c_uint32, _D, ctypes.c_uint32, _U, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_uint, ctypes.c_uint, _U, ctypes.c_uint64, _U, ctypes.c_uint64, _D,
                                         _U8, _U, ctypes.c_uint32, ctypes.c_uint, _D, _D]
    rc = lib.oracle_assemble_rows(nv, _d(V), nt, _u(T), k, tolerance, regular_order, singular_order,
                                  _u(touching), len(plan["touching"]), _u(near), len(plan["near"]),
                                  _d(patches) if len(plan["patches"]) else None,
                                  whole.ctypes.data_as(_U8) if len(whole) else None, _u(R), m, threads, _d(A), _d(B))
    _check(lib, rc, "oracle_assemble_rows")
    return (A.view(np.complex128).reshape(nv, m).T.copy(), B.view(np.complex128).reshape(nt, m).T.copy())


def oracle_lu_solve(A, R, threads=8):
    lib = oracle()
    n = A.shape[0]
    Rm = np.asarray(R, np.complex128).reshape(n, -1)
    Af = np.ascontiguousarray(A.T).astype(np.complex128).view(np.float64).copy()
    Rf = np.ascontiguousarray(Rm.T).view(np.float64).copy()
    lib.oracle_lu_solve.argtypes = [ctypes.c_uint64, _D, ctypes.c_uint64, _D, ctypes.c_uint]
    _check(lib, lib.oracle_lu_solve(n, _d(Af), Rm.shape[1], _d(Rf), threads), "oracle_lu_solve")
    X = Rf.view(np.complex128).reshape(Rm.shape[1], n).T
    return X.reshape(np.asarray(R).shape).copy()


def oracle_field(mesh, k, degree, pressure, flux, observers, far=False, origin=(0, 0, 0), R=1.0, near_tolerance=1e-6,
                 depth=6, vertices=None):
    """Observer (or far-field) pressures; pressure (nv, nrhs) and flux (nt, nrhs) complex arrays."""
    lib = oracle()
    V = np.ascontiguousarray(recentred(mesh) if vertices is None else vertices)
    T = np.ascontiguousarray(mesh.triangles, dtype=np.uint32)
    P = np.ascontiguousarray(np.asarray(pressure, np.complex128).reshape(len(V), -1).T).view(np.float64).copy()
    Q = np.ascontiguousarray(np.asarray(flux, np.complex128).reshape(len(T), -1).T).view(np.float64).copy()
    nrhs = P.size // (2 * len(V))
    O = np.ascontiguousarray(np.asarray(observers, float).reshape(-1, 3))
    out = np.zeros(2 * len(O) * nrhs)
    org = np.asarray(origin, float)
    lib.oracle_field.argtypes = [ctypes.c_uint32, _D, ctypes.c_uint32, _U, ctypes.c_double, ctypes.c_uint, _D, _D,
                                 ctypes.c_uint32, _D, ctypes.c_uint32, ctypes.c_int, _D, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_uint, _D]
    _check(lib, lib.oracle_field(len(V), _d(V), len(T), _u(T), k, degree, _d(P), _d(Q), nrhs, _d(O), len(O),
                                 1 if far else 0, _d(org), R, near_tolerance, depth, _d(out)), "oracle_field")
    return out.view(np.complex128).reshape(nrhs, len(O)).T.copy()


def triangle_rule(lib, name, degree):
    fn = getattr(lib, name)
    count = ctypes.c_uint()
    fn.argtypes = [ctypes.c_uint, _D, ctypes.POINTER(ctypes.c_uint)]
    _check(lib, fn(degree, None, ctypes.byref(count)), name)
    out = np.zeros(4 * count.value)
    _check(lib, fn(degree, _d(out), ctypes.byref(count)), name)
    return out.reshape(-1, 4)


def gauss(lib, name, order):
    fn = getattr(lib, name)
    fn.argtypes = [ctypes.c_uint, _D]
    out = np.zeros(2 * order)
    _check(lib, fn(order, _d(out)), name)
    return out.reshape(-1, 2)


def touching_points(lib, name, mesh, a, b, order, vertices=None):
    V = np.ascontiguousarray(recentred(mesh) if vertices is None else vertices)
    T = np.ascontiguousarray(mesh.triangles, dtype=np.uint32)
    fn = getattr(lib, name)
    count = ctypes.c_uint()
    if name == "ref_pair_points":
        topo = ctypes.c_int()
        fn.argtypes = [ctypes.c_uint32, _D, ctypes.c_uint32, _U, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint, _D,
                       ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_int)]
        _check(lib, fn(len(V), _d(V), len(T), _u(T), a, b, order, None, ctypes.byref(count), ctypes.byref(topo)), name)
        out = np.zeros(7 * count.value)
        _check(lib, fn(len(V), _d(V), len(T), _u(T), a, b, order, _d(out), ctypes.byref(count), ctypes.byref(topo)),
               name)
    else:
        fn.argtypes = [ctypes.c_uint32, _D, ctypes.c_uint32, _U, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint, _D,
                       ctypes.POINTER(ctypes.c_uint)]
        _check(lib, fn(len(V), _d(V), len(T), _u(T), a, b, order, None, ctypes.byref(count)), name)
        out = np.zeros(7 * count.value)
        _check(lib, fn(len(V), _d(V), len(T), _u(T), a, b, order, _d(out), ctypes.byref(count)), name)
    return out.reshape(-1, 7)


def patch_moments(lib, name, mesh, a, b, patch, k, vertices=None):
    V = np.ascontiguousarray(recentred(mesh) if vertices is None else vertices)
    T = np.ascontiguousarray(mesh.triangles, dtype=np.uint32)
    P = np.ascontiguousarray(patch, dtype=np.float64)
    out = np.zeros(72)
    fn = getattr(lib, name)
    if name == "ref_estimate_patch":
        errs = np.zeros(3)
        fn.argtypes = [ctypes.c_uint32, _D, ctypes.c_uint32, _U, ctypes.c_uint32, ctypes.c_uint32, _D, ctypes.c_double,
                       _D, _D]
        _check(lib, fn(len(V), _d(V), len(T), _u(T), a, b, _d(P), k, _d(out), _d(errs)), name)
    else:
        fn.argtypes = [ctypes.c_uint32, _D, ctypes.c_uint32, _U, ctypes.c_uint32, ctypes.c_uint32, _D, ctypes.c_double,
                       _D]
        _check(lib, fn(len(V), _d(V), len(T), _u(T), a, b, _d(P), k, _d(out)), name)
    return out.view(np.complex128)
