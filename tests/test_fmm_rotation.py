"""LF FMM rotation plans: Wigner small-d by spin-1/2 coupling and the pullback of
spherical-harmonic expansions, bit-identical to the reference
(fmm/spherical_rotation.cpp compiled in oracle/_ref) and pinned by the
mathematics: orthogonality, d(0) = I, the closed form of d^1, and f(Q r) against
scipy's orthonormal Y_lm (Condon-Shortley phase)."""
import numpy as np
import pytest
from scipy.special import sph_harm

import oracle_bind as ob
from hornsculpt_amd import bem

needs_ref = pytest.mark.skipif(not ob.REF.exists(), reason="reference planner (oracle/_ref) not built")
pytestmark = pytest.mark.filterwarnings("ignore::DeprecationWarning")
AXES = [(0.3, -0.4, 0.86), (0.0, 0.0, 1.0), (0.0, 0.0, -2.0), (1.0, 0.0, 0.0), (-0.2, 0.9, -0.1)]


def _blocks(d, order):
    out, pos = [], 0
    for l in range(order + 1):
        n = 2 * l + 1
        out.append(d[pos:pos + n * n].reshape(n, n))
        pos += n * n
    return out


@needs_ref
@pytest.mark.parametrize("order", [0, 1, 7, 30, 128])
@pytest.mark.parametrize("axis", AXES)
def test_rotation_plan_bit_identical_to_reference(order, axis):
    rng = np.random.default_rng(order)
    c = rng.standard_normal((order + 1) ** 2) + 1j * rng.standard_normal((order + 1) ** 2)
    for direction in (1, -1):
        d, az, r = bem.rotation_plan(order, axis, c, direction)
        rd, raz, rr = ob.ref_rotation(order, axis, c, direction)
        assert np.array_equal(d, rd) and np.array_equal(az, raz) and np.array_equal(r, rr)


@pytest.mark.parametrize("axis", AXES)
def test_wigner_blocks_orthogonal_and_closed_form(axis):
    order = 12
    d, az, _ = bem.rotation_plan(order, axis)
    theta = np.arctan2(np.hypot(axis[0], axis[1]), axis[2])
    for l, block in enumerate(_blocks(d, order)):
        assert np.allclose(block @ block.T, np.eye(2 * l + 1), atol=1e-13)
    c, s = np.cos(theta), np.sin(theta)
    # d^1 in the m' , m = -1, 0, 1 basis (sign convention fixed by the recursion)
    d1 = _blocks(d, order)[1]
    want = np.array([[(1 + c) / 2, -s / np.sqrt(2), (1 - c) / 2],
                     [s / np.sqrt(2), c, -s / np.sqrt(2)],
                     [(1 - c) / 2, s / np.sqrt(2), (1 + c) / 2]])
    assert np.allclose(np.abs(d1), np.abs(want), atol=1e-15)
    assert np.allclose(np.abs(az), 1, atol=1e-15)


def test_identity_rotation_and_round_trip():
    order = 20
    rng = np.random.default_rng(5)
    c = rng.standard_normal((order + 1) ** 2) + 1j * rng.standard_normal((order + 1) ** 2)
    d, _, r = bem.rotation_plan(order, (0, 0, 1), c, 1)
    for l, block in enumerate(_blocks(d, order)):
        assert np.allclose(block, np.eye(2 * l + 1), atol=1e-15)
    assert np.allclose(r, c, atol=1e-14)
    _, _, there = bem.rotation_plan(order, (0.3, -0.4, 0.86), c, 1)
    _, _, back = bem.rotation_plan(order, (0.3, -0.4, 0.86), there, -1)
    assert np.allclose(back, c, atol=1e-12)


def _field(coeffs, order, dirs):
    theta = np.arccos(np.clip(dirs[:, 2], -1, 1))
    phi = np.arctan2(dirs[:, 1], dirs[:, 0])
    f = np.zeros(len(dirs), complex)
    for l in range(order + 1):
        for m in range(-l, l + 1):
            f += coeffs[l * l + l + m] * sph_harm(m, l, phi, theta)
    return f


@pytest.mark.parametrize("axis", [(0.3, -0.4, 0.86), (-0.2, 0.9, -0.1)])
def test_pullback_matches_spherical_harmonics(axis):
    """sum c'_lm Y_lm(r) = sum c_lm Y_lm(Q r) with Q = Rz(phi) Ry(theta), Q e_z = axis."""
    order = 8
    rng = np.random.default_rng(11)
    c = rng.standard_normal((order + 1) ** 2) + 1j * rng.standard_normal((order + 1) ** 2)
    _, _, cr = bem.rotation_plan(order, axis, c, 1)
    a = np.asarray(axis, float)
    theta, phi = np.arctan2(np.hypot(a[0], a[1]), a[2]), np.arctan2(a[1], a[0])
    Rz = np.array([[np.cos(phi), -np.sin(phi), 0], [np.sin(phi), np.cos(phi), 0], [0, 0, 1]])
    Ry = np.array([[np.cos(theta), 0, np.sin(theta)], [0, 1, 0], [-np.sin(theta), 0, np.cos(theta)]])
    Q = Rz @ Ry
    assert np.allclose(Q @ [0, 0, 1], a / np.linalg.norm(a))
    dirs = rng.standard_normal((40, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    assert np.allclose(_field(cr, order, dirs), _field(c, order, dirs @ Q.T), atol=1e-11)


@needs_ref
@pytest.mark.parametrize("s,t,r", [(0, 0, 1), (10, 20, 32), (64, 64, 256), (128, 128, 8), (37, 3, 100)])
def test_rcr_queue_policy_matches_reference(s, t, r):
    assert bem.rcr_queue_policy(s, t, r) == tuple(int(v) for v in ob.ref_rcr_queue_policy(s, t, r))
