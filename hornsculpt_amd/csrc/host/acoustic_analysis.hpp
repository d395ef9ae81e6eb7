// Post-processing of BEM results: SPL, normalised impedance, angular levels,
// CTA-2034 directional averages, beamwidth, angular frequency responses and the
// acoustic H/V frame of a directivity sweep.
//
// Restates modules/analysis/include/hornsculpt/acoustic_analysis.hpp:11-78 /
// acoustic_analysis.cpp:1-207 and acoustic_frame.hpp:8-26 / acoustic_frame.cpp:1-55:
//   * SPL = 20 log10(|p| / p_ref) with the amplitude formed from the scaled
//     components (a true zero is -infinity, no floor);
//   * angular levels interpolate linearly in dB between finite samples only;
//   * beamwidth = first threshold crossings of the lobe containing 0 degrees,
//     relative to the (interpolated) on-axis level;
//   * CTA-2034 (ANSI/CEA-2034-A sections 5.2, 13, appendices C and F): power
//     averages over the listening window (9 directions), early reflections
//     (5 equally weighted groups) and sound power (70 directions, Table 7
//     weights), the DIs and the predicted in-room curve 0.12 LW + 0.44 ER + 0.44 SP;
//   * frames: forward normalised, horizontal re-orthogonalised, V = forward x H;
//     positive H angles rotate forward towards horizontal, positive V towards V.
#pragma once

#include "vec.hpp"

#include <complex>
#include <optional>
#include <vector>

namespace hsbem {

struct AngularSample {
    double angle{}, level{}; // degrees, dB
};

struct CtaCuts {
    std::optional<double> on_axis, listening_window, early_reflections, early_reflections_di, sound_power,
        sound_power_di, predicted_in_room;
};

struct Beamwidth {
    double left{}, right{}, width{};
};

struct AngularRow {
    double frequency{};
    std::vector<AngularSample> samples;
};

struct AcousticFrame {
    Vec3 origin{};
    Vec3 forward{0, 0, 1};
    Vec3 horizontal{1, 0, 0};
};

double spl_db(std::complex<double> pressure, double reference = 20e-6);
std::complex<double> normalized_volume_impedance(std::complex<double> impedance, double area, double density,
                                                 double sound_speed);
CtaCuts cta2034_cuts(const std::vector<AngularSample> &horizontal, const std::vector<AngularSample> &vertical);
std::optional<double> angular_level_db(const std::vector<AngularSample> &samples, double angle);
std::optional<Beamwidth> beamwidth(const std::vector<AngularSample> &samples, double attenuation = 6.0);
// relative: subtract the level at reference_angle at each frequency.
std::vector<std::optional<double>> angular_frequency_response(const std::vector<AngularRow> &rows, double angle,
                                                              bool relative, double reference_angle = 0.0);

AcousticFrame normalized_acoustic_frame(AcousticFrame frame);
AcousticFrame default_acoustic_frame(Vec3 origin = {}, Vec3 forward = {0, 0, 1});
Vec3 vertical_direction(const AcousticFrame &frame);
// Unit directions cos(a) forward + sin(a) axis (axis = H, or V when vertical).
std::vector<Vec3> polar_directions(const AcousticFrame &frame, bool vertical, const std::vector<double> &degrees);

} // namespace hsbem
