"""Result post-processing: SPL, normalised impedance, angular interpolation,
beamwidth, CTA-2034 directional averages, angular frequency responses and the
acoustic H/V frame.

Bit-identical to the reference (analysis/acoustic_analysis.cpp and
acoustic_frame.cpp compiled in oracle/_ref) on random and edge-case inputs, and
pinned by the known answers of the reference's own tests
(tests/acoustic_analysis_tests.cpp): analytic triangle lobes, SPL of 20 uPa and
1 Pa, interpolation values, asymmetric lobes, CTA-2034 Table 7 sound-power
golden values (one 80 dB direction against true zeros)."""
import math

import numpy as np
import pytest

import oracle_bind as ob
from hornsculpt_amd import bem

needs_ref = pytest.mark.skipif(not ob.REF.exists(), reason="reference planner (oracle/_ref) not built")
NAN, INF = float("nan"), float("inf")


def _same(a, b):
    if a is None or b is None:
        return a is None and b is None
    if isinstance(a, tuple):
        return all(_same(x, y) for x, y in zip(a, b))
    return (math.isnan(a) and math.isnan(b)) or a == b


def _random_samples(rng, n, nan_fraction=0.1):
    angles = np.sort(rng.choice(np.arange(-180, 181, 5), size=n, replace=False)).astype(float)
    levels = 80 - np.abs(angles) / rng.uniform(3, 12) + rng.normal(0, 2, n)
    levels[rng.random(n) < nan_fraction] = NAN
    return list(zip(angles.tolist(), levels.tolist()))


# ---------------------------------------------------------------- bit parity


@needs_ref
def test_spl_and_impedance_bit_identical():
    for p in (20e-6, 1j, 1.0, 3.5 - 2.25j, 1e-300, 1e300 + 1e300j, 0.0):
        assert _same(bem.spl_db(p), ob.ref_spl_db(p))
        assert _same(bem.spl_db(p, 1e-3), ob.ref_spl_db(p, 1e-3))
    for z in (1.0, 4e5 - 2e5j, -1e-9j):
        assert bem.normalized_volume_impedance(z, 5e-4, 1.21, 343.0) == ob.ref_normalized_volume_impedance(
            z, 5e-4, 1.21, 343.0)


@needs_ref
@pytest.mark.parametrize("seed", range(12))
def test_angular_functions_bit_identical(seed):
    rng = np.random.default_rng(seed)
    samples = _random_samples(rng, int(rng.integers(3, 40)))
    for angle in rng.uniform(-190, 190, 20).tolist() + [0.0, samples[0][0], samples[-1][0]]:
        assert _same(bem.angular_level_db(samples, angle), ob.ref_angular_level_db(samples, angle))
    for att in (3.0, 6.0, 10.0):
        assert _same(bem.beamwidth(samples, att), ob.ref_beamwidth(samples, att))
    rows = [(100.0 * (i + 1), _random_samples(rng, int(rng.integers(2, 30)))) for i in range(5)]
    for angle, rel, refa in ((2.5, False, 0.0), (2.5, True, 0.0), (-7.0, True, 12.5)):
        assert all(_same(a, b) for a, b in zip(bem.angular_frequency_response(rows, angle, rel, refa),
                                               ob.ref_angular_frequency_response(rows, angle, rel, refa)))


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_cta_cuts_bit_identical(seed):
    rng = np.random.default_rng(100 + seed)
    angles = np.arange(-180, 181, 5.0)
    h = list(zip(angles, 85 - np.abs(angles) / 8 + rng.normal(0, 1.5, angles.size)))
    v = list(zip(angles, 84 - np.abs(angles) / 6 + rng.normal(0, 1.5, angles.size)))
    if seed % 3 == 1:
        i = int(rng.integers(len(h)))
        h[i] = (h[i][0], NAN)
    if seed % 3 == 2:
        v = v[:-1]
    got, want = bem.cta2034_cuts(h, v), ob.ref_cta2034_cuts(h, v)
    assert all(_same(got[k], w) for k, w in zip(bem.CTA_FIELDS, want))


@needs_ref
def test_acoustic_frames_bit_identical():
    for origin, forward, horizontal in (((0, 0, 0), (0, 0, 1), None), ((.3, -.2, .1), (2, 0, 0), None),
                                        ((.3, -.2, .1), (-2, 0, 0), None), ((1, 2, 3), (0, 3, 4), (4, 3, 4)),
                                        ((0, 0, 0), (1e-300, 2e-300, 0), None), ((0, 0, 0), (0, -1e300, 0), (0, 0, 1e300))):
        f = bem.acoustic_frame(origin, forward, horizontal)
        got = np.concatenate([f["origin"], f["forward"], f["horizontal"], f["vertical"]])
        assert np.array_equal(got, ob.ref_acoustic_frame(origin, forward, horizontal))


# ---------------------------------------------------------------- known answers


def test_analytic_triangle_lobes_and_offsets():
    tri = [(float(a), 80.0 - abs(a) / 5.0) for a in range(-90, 91, 15)]
    for att in (3.0, 6.0, 10.0):
        left, right, width = bem.beamwidth(tri, att)
        assert left == pytest.approx(-5 * att, abs=1e-10) and right == pytest.approx(5 * att, abs=1e-10)
        assert width == pytest.approx(10 * att, abs=1e-10)
    shifted = [(a, l - 123.0) for a, l in tri]
    assert bem.beamwidth(shifted)[2] == pytest.approx(60.0, abs=1e-10)


def test_spl_and_normalisation_known_values():
    assert bem.spl_db(20e-6) == pytest.approx(0.0, abs=1e-10)
    assert bem.spl_db(20e-6j) == pytest.approx(0.0, abs=1e-10)
    assert bem.spl_db(1.0) == pytest.approx(93.97940008672037, abs=1e-10)
    assert bem.spl_db(0.0) == -INF
    ref = 1.21 * 343.0 / 5e-4
    z = bem.normalized_volume_impedance(complex(ref, -2 * ref), 5e-4, 1.21, 343.0)
    assert z.real == pytest.approx(1.0, abs=1e-10) and z.imag == pytest.approx(-2.0, abs=1e-10)
    with pytest.raises(ValueError):
        bem.spl_db(1.0, 0.0)
    with pytest.raises(ValueError):
        bem.normalized_volume_impedance(1.0, 0.0, 1.21, 343.0)


def test_angular_interpolation_known_values():
    s = [(-180, 40), (-30, 80), (0, 90), (10, 93), (30, 85), (180, 40)]
    for angle, want in ((10, 93), (5, 91.5), (-15, 85), (-180, 40), (180, 40)):
        assert bem.angular_level_db(s, angle) == pytest.approx(want, abs=1e-10)
    assert bem.angular_level_db(s, 181) is None
    missing = list(s)
    missing[3] = (10, NAN)
    assert bem.angular_level_db(missing, 5) is None and bem.angular_level_db(missing, 0) == 90
    with pytest.raises(ValueError):
        bem.angular_level_db(s, INF)
    with pytest.raises(ValueError):
        bem.angular_level_db([(5, 86), (0, 90)], 0)


def test_angular_frequency_response_known_values():
    low = [(-10, 80), (-5, 86), (0, 90), (5, 94), (10, 84)]
    mid = [(-10, 60), (-5, 70), (0, 85), (5, 81), (10, 70)]
    high = [(-10, 75), (-5, 80), (0, 89), (5, 80), (10, 70)]
    rows = [(500, low), (1000, mid), (2000, high)]
    assert bem.angular_frequency_response(rows, 2.5) == pytest.approx([92, 83, 84.5], abs=1e-10)
    assert bem.angular_frequency_response(rows, 2.5, True) == pytest.approx([2, -2, -4.5], abs=1e-10)
    assert bem.angular_frequency_response(rows, 2.5, True, -2.5) == pytest.approx([4, 5.5, 0], abs=1e-10)
    assert bem.angular_frequency_response([], 0) == []
    with pytest.raises(ValueError):
        bem.angular_frequency_response([(200, low), (100, low)], 0)


def test_asymmetric_lobes_and_gaps():
    s = [(-90, 3), (-60, -12), (-20, -4), (0, 0), (15, -3), (45, -12), (70, 2), (90, -12)]
    assert bem.beamwidth(s) == pytest.approx((-30.0, 25.0, 55.0), abs=1e-10)
    peak = [(-40, -12), (-10, -2), (0, 0), (10, 2), (40, -12)]
    left, right, _ = bem.beamwidth(peak)
    assert left == pytest.approx(-22.0, abs=1e-10) and right == pytest.approx(10 + 30 * 8 / 14, abs=1e-10)
    assert bem.beamwidth([(-60, -12), (-10, 0), (10, 0), (60, -12)])[2] == pytest.approx(70.0, abs=1e-10)
    assert bem.beamwidth([]) is None
    assert bem.beamwidth([(-20, -3), (0, 0), (20, -3)]) is None
    assert bem.beamwidth([(-90, NAN), (-30, -12), (0, 0), (30, -12), (90, NAN)])[2] == pytest.approx(30.0)
    for bad in (0.0, -6.0, NAN, INF):
        with pytest.raises(ValueError):
            bem.beamwidth([(-30, -12), (0, 0), (30, -12)], bad)


def test_cta_uniform_field_and_single_direction():
    angles = np.arange(-180, 181, 5.0)
    h = [(a, 80.0) for a in angles]
    v = [(a, 80.0) for a in angles]
    c = bem.cta2034_cuts(h, v)
    for key in ("on_axis", "listening_window", "early_reflections", "sound_power", "predicted_in_room"):
        assert c[key] == pytest.approx(80.0, abs=1e-10)
    assert c["early_reflections_di"] == pytest.approx(0.0, abs=1e-10)
    h[38] = (10.0, 90.0)
    c = bem.cta2034_cuts(h, v)
    assert c["listening_window"] == pytest.approx(80 + 10 * math.log10(2), abs=1e-10)
    assert c["early_reflections"] == pytest.approx(80 + 10 * math.log10(1 + 9 / 35), abs=1e-10)
    assert c["sound_power"] == pytest.approx(80.181704804004, abs=1e-10)
    c = bem.cta2034_cuts(h[:-1], v)
    assert c["listening_window"] is not None and c["early_reflections"] is None and c["sound_power"] is None


@pytest.mark.parametrize("ring", range(19))
def test_cta_table7_single_direction_golden(ring):
    """One 80 dB direction against true zeros: the published ring weights over the
    70-direction sum 0.996373088."""
    expected = [47.829642582, 56.764565051, 59.536449109, 60.945565601, 61.773683737, 62.286455686, 62.608364126,
                62.804857365, 62.911176235, 62.944844389, 62.911176235, 62.804857365, 62.608364126, 62.286455686,
                61.773683737, 60.945565601, 59.536449109, 56.764565051, 47.829642582]
    angle = 10 * ring
    angles = np.arange(-180, 181, 5.0)
    for vertical in (False, True):
        if vertical and angle in (0, 180):
            continue
        for sign in (-1, 1):
            if angle in (0, 180) and sign < 0:
                continue
            h = [[a, -INF] for a in angles]
            v = [[a, -INF] for a in angles]
            (v if vertical else h)[(sign * angle + 180) // 5][1] = 80.0
            got = bem.cta2034_cuts([tuple(x) for x in h], [tuple(x) for x in v])["sound_power"]
            assert got == pytest.approx(expected[ring], abs=1e-8)


def test_frames_and_polar_directions():
    f = bem.acoustic_frame()
    assert np.allclose(f["forward"], (0, 0, 1)) and np.allclose(f["horizontal"], (1, 0, 0))
    assert np.allclose(f["vertical"], (0, 1, 0))
    o = bem.acoustic_frame((1, 2, 3), (0, 3, 4), (4, 3, 4))
    assert np.allclose(o["forward"], (0, .6, .8)) and np.allclose(o["horizontal"], (1, 0, 0))
    assert np.allclose(o["vertical"], (0, .8, -.6))
    d = bem.polar_directions([0, 90, -90, 30], forward=(1, 0, 0))
    fr = bem.acoustic_frame(forward=(1, 0, 0))
    assert np.allclose(d[0], fr["forward"]) and np.allclose(d[1], fr["horizontal"])
    assert np.allclose(d[2], -fr["horizontal"]) and np.allclose(np.linalg.norm(d, axis=1), 1)
    dv = bem.polar_directions([90], vertical=True, forward=(1, 0, 0))
    assert np.allclose(dv[0], fr["vertical"])
    for bad in (((0, 0, 1), (0, 0, 2)), ((0, 0, 1), (1e-10, 0, -1))):
        with pytest.raises(ValueError):
            bem.acoustic_frame(forward=bad[0], horizontal=bad[1])
