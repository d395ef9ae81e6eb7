"""Directivity from device far-field solves through the result post-processing:
polar sweeps in the acoustic frame, SPL, beamwidth and CTA-2034 averages."""
import math

import numpy as np
import pytest

from hornsculpt_amd import bem, meshes

pytestmark = pytest.mark.gpu
ANGLES = np.arange(-180.0, 181.0, 5.0)


def _sweeps(mesh, frequency, face_velocity, origin, forward):
    h = bem.polar_directions(ANGLES, False, origin, forward)
    v = bem.polar_directions(ANGLES, True, origin, forward)
    req = bem.BemRequest(medium=bem.AcousticMedium(1.205, 343.0), frequencies=[frequency],
                         excitations=[bem.BemExcitation("drive", [bem.FaceVelocity(f, 1.0) for f in face_velocity])],
                         observers=[tuple(np.asarray(origin) + d) for d in np.vstack([h, v])],
                         numerics=bem.BemNumerics(solver=bem.SOLVER_GMRES, relative_residual_tolerance=1e-6))
    req.far_field = bem.FarFieldObservation(tuple(origin), 1.0)
    p = bem.HipBemSolver(mesh).solve(req).frequencies[0].excitations[0].observer_pressure
    spl = [bem.spl_db(x) for x in p]
    n = len(ANGLES)
    return list(zip(ANGLES, spl[:n])), list(zip(ANGLES, spl[n:]))


def test_pulsating_sphere_is_omnidirectional():
    h, v = _sweeps(meshes.icosphere(3), 1000.0, [1], (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    levels = np.array([x[1] for x in h + v])
    # analytic far field |p| = rho c k a^2 v / |1 + ika| at 1 m
    k, a = 2 * math.pi * 1000 / 343.0, 0.1
    want = bem.spl_db(1.205 * 343.0 * k * a * a / abs(1 + 1j * k * a))
    print(f"sphere SPL {levels.min():.3f} .. {levels.max():.3f} dB, analytic {want:.3f} dB")
    assert levels.max() - levels.min() < 0.02 and abs(levels.mean() - want) < 0.1
    assert bem.beamwidth(h) is None
    c = bem.cta2034_cuts(h, v)
    assert abs(c["listening_window"] - c["sound_power"]) < 0.02 and abs(c["sound_power_di"]) < 0.02


def test_horn_directivity_is_axisymmetric_and_beamed():
    mesh = meshes.exponential_horn(0.012)
    driven = sorted(set(int(f) for f, b in zip(mesh.face_ids, mesh.boundary) if b == bem.DRIVEN))
    h, v = _sweeps(mesh, 4000.0, driven, (0.3, 0.0, 0.0), (1.0, 0.0, 0.0))
    hl, vl = np.array([x[1] for x in h]), np.array([x[1] for x in v])
    bw_h, bw_v = bem.beamwidth(h), bem.beamwidth(v)
    c = bem.cta2034_cuts(h, v)
    print(f"horn 4 kHz: on-axis {c['on_axis']:.2f} dB, H/V max difference {np.max(np.abs(hl - vl)):.3f} dB, "
          f"beamwidth H {bw_h}, V {bw_v}, LW {c['listening_window']:.2f}, SP {c['sound_power']:.2f}, "
          f"SP DI {c['sound_power_di']:.2f} dB")
    # the mesh is revolved: the H and V cuts agree closely
    assert np.max(np.abs(hl - vl)) < 0.5
    assert bw_h is not None and bw_v is not None and 10 < bw_h[2] < 180 and abs(bw_h[2] - bw_v[2]) < 3
    assert c["sound_power_di"] > 3 and c["on_axis"] >= hl.max() - 3
