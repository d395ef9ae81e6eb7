#include "acoustic_analysis.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numbers>
#include <stdexcept>

namespace hsbem {
namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

// (value - a) / (b - a), all three scaled by the largest magnitude first
double fraction(double value, double a, double b) {
    const double scale = std::max({std::abs(value), std::abs(a), std::abs(b)});
    return (value / scale - a / scale) / (b / scale - a / scale);
}

void check_samples(const std::vector<AngularSample> &s) {
    for (std::size_t i = 0; i < s.size(); ++i)
        if (!std::isfinite(s[i].angle) || (i != 0 && s[i].angle <= s[i - 1].angle))
            throw std::invalid_argument("angles must be finite and strictly increasing");
}

bool finite(Vec3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

Vec3 unit(Vec3 v) {
    if (!finite(v))
        throw std::invalid_argument("Acoustic direction must be finite.");
    const double scale = std::max({std::abs(v.x), std::abs(v.y), std::abs(v.z)});
    if (scale == 0)
        throw std::invalid_argument("Acoustic direction must not be zero.");
    v = {v.x / scale, v.y / scale, v.z / scale}; // scaled first: huge or tiny finite vectors stay valid
    const double length = std::hypot(v.x, v.y, v.z);
    return {v.x / length, v.y / length, v.z / length};
}

} // namespace

double spl_db(std::complex<double> p, double reference) {
    if (!std::isfinite(p.real()) || !std::isfinite(p.imag()) || !std::isfinite(reference) || reference <= 0.0)
        throw std::invalid_argument("sound pressure level needs a finite pressure and a positive reference pressure");
    const double scale = std::max(std::abs(p.real()), std::abs(p.imag()));
    if (scale == 0.0)
        return -kInf;
    const double amplitude = std::hypot(p.real() / scale, p.imag() / scale);
    return 20.0 * (std::log10(scale) + std::log10(amplitude) - std::log10(reference));
}

std::complex<double> normalized_volume_impedance(std::complex<double> z, double area, double density, double c) {
    if (!std::isfinite(z.real()) || !std::isfinite(z.imag()) || !std::isfinite(area) || area <= 0.0 ||
        !std::isfinite(density) || density <= 0.0 || !std::isfinite(c) || c <= 0.0)
        throw std::invalid_argument("impedance normalisation needs finite values and positive area, density and sound speed");
    const auto out = z * ((area / density) / c);
    if (!std::isfinite(out.real()) || !std::isfinite(out.imag()))
        throw std::invalid_argument("normalised impedance overflows");
    return out;
}

CtaCuts cta2034_cuts(const std::vector<AngularSample> &horizontal, const std::vector<AngularSample> &vertical) {
    // level at an exact integer angle; NaN and +infinity count as missing
    auto find = [](const std::vector<AngularSample> &s, int angle) -> std::optional<double> {
        for (const auto &x : s)
            if (x.angle == angle && !std::isnan(x.level) && x.level != kInf)
                return x.level;
        return {};
    };
    struct Direction {
        bool vertical;
        int angle;
        double weight;
    };
    // weighted power mean in dB (relative to the largest level); any missing direction -> none
    auto average = [&](const std::vector<Direction> &dirs) -> std::optional<double> {
        std::vector<std::pair<double, double>> levels;
        double total = 0;
        for (const auto &[vert, angle, weight] : dirs) {
            const auto level = find(vert ? vertical : horizontal, angle);
            if (!level)
                return {};
            levels.emplace_back(*level, weight);
            total += weight;
        }
        const double top = std::max_element(levels.begin(), levels.end(), [](const auto &a, const auto &b) {
                               return a.first < b.first;
                           })->first;
        if (!std::isfinite(top))
            return top;
        double sum = 0;
        for (const auto &[level, weight] : levels)
            sum += weight * std::pow(10., (level - top) / 10.);
        return top + 10 * std::log10(sum / total);
    };
    const std::vector<Direction> listening{{false, 0, 1},   {false, -10, 1}, {false, 10, 1},
                                           {false, -20, 1}, {false, 20, 1},  {false, -30, 1},
                                           {false, 30, 1},  {true, -10, 1},  {true, 10, 1}};
    // floor and ceiling bounces (vertical), then the front, side and rear walls (horizontal)
    std::vector<Direction> reflections{{true, -20, 1. / 3}, {true, -30, 1. / 3}, {true, -40, 1. / 3},
                                       {true, 40, 1. / 3},  {true, 50, 1. / 3},  {true, 60, 1. / 3},
                                       {false, 0, 1. / 7},  {false, 180, 1. / 3}};
    for (int angle = 10; angle <= 90; angle += 10) {
        const double weight = angle <= 30 ? 1. / 7 : angle <= 80 ? 1. / 10 : 1. / 3;
        reflections.push_back({false, -angle, weight});
        reflections.push_back({false, angle, weight});
    }
    // Table 7 of ANSI/CEA-2034-A: weight per 10-degree ring
    constexpr double ring[] = {.000604486, .004730189, .008955027, .012387354, .014989611, .016868154, .018165962,
                               .019006744, .019477787, .019629373, .019477787, .019006744, .018165962, .016868154,
                               .014989611, .012387354, .008955027, .004730189, .000604486};
    std::vector<Direction> power{{false, 0, ring[0]}, {false, 180, ring[18]}};
    for (int angle = 10; angle < 180; angle += 10)
        for (bool vert : {false, true})
            for (int sign : {-1, 1})
                power.push_back({vert, sign * angle, ring[angle / 10]});
    CtaCuts out{find(horizontal, 0), average(listening), average(reflections), {}, average(power), {}, {}};
    if (out.listening_window && out.early_reflections && std::isfinite(*out.listening_window) &&
        std::isfinite(*out.early_reflections))
        out.early_reflections_di = *out.listening_window - *out.early_reflections;
    if (out.listening_window && out.sound_power && std::isfinite(*out.listening_window) &&
        std::isfinite(*out.sound_power))
        out.sound_power_di = *out.listening_window - *out.sound_power;
    if (out.listening_window && out.early_reflections && out.sound_power) {
        const double top = std::max({*out.listening_window, *out.early_reflections, *out.sound_power});
        out.predicted_in_room = std::isfinite(top)
                                    ? top + 10 * std::log10(.12 * std::pow(10., (*out.listening_window - top) / 10.) +
                                                            .44 * std::pow(10., (*out.early_reflections - top) / 10.) +
                                                            .44 * std::pow(10., (*out.sound_power - top) / 10.))
                                    : top;
    }
    return out;
}

std::optional<double> angular_level_db(const std::vector<AngularSample> &s, double angle) {
    if (!std::isfinite(angle))
        throw std::invalid_argument("the reference angle must be finite");
    check_samples(s);
    const auto right = std::lower_bound(s.begin(), s.end(), angle,
                                        [](const AngularSample &x, double a) { return x.angle < a; });
    if (right == s.end() || !std::isfinite(right->level))
        return {};
    if (right->angle == angle)
        return right->level;
    if (right == s.begin() || !std::isfinite((right - 1)->level))
        return {};
    const AngularSample &left = *(right - 1);
    return std::lerp(left.level, right->level, fraction(angle, left.angle, right->angle));
}

std::optional<Beamwidth> beamwidth(const std::vector<AngularSample> &s, double attenuation) {
    if (!std::isfinite(attenuation) || attenuation <= 0.0)
        throw std::invalid_argument("beamwidth attenuation must be a finite positive number of dB");
    const auto axis_level = angular_level_db(s, 0.0);
    if (!axis_level)
        return std::nullopt;
    const auto right = std::lower_bound(s.begin(), s.end(), 0.0,
                                        [](const AngularSample &x, double a) { return x.angle < a; });
    if (right == s.end())
        return std::nullopt;
    const auto index = static_cast<std::ptrdiff_t>(right - s.begin());
    const bool exact_axis = right->angle == 0.0;
    const double axis = *axis_level, threshold = axis - attenuation;
    if (!std::isfinite(axis) || !std::isfinite(threshold) || threshold >= axis)
        return std::nullopt;
    // walk outwards from the axis; a non-finite level on the way ends the lobe without a crossing
    auto crossing = [&](std::ptrdiff_t first, std::ptrdiff_t step) -> std::optional<double> {
        AngularSample previous{0.0, axis};
        for (auto i = first; i >= 0 && i < static_cast<std::ptrdiff_t>(s.size()); i += step) {
            const AngularSample &current = s[static_cast<std::size_t>(i)];
            if (!std::isfinite(current.level))
                return std::nullopt;
            if (current.level <= threshold)
                return std::lerp(previous.angle, current.angle, fraction(threshold, previous.level, current.level));
            previous = current;
        }
        return std::nullopt;
    };
    const auto left_crossing = crossing(index - 1, -1);
    const auto right_crossing = crossing(exact_axis ? index + 1 : index, 1);
    if (!left_crossing || !right_crossing)
        return std::nullopt;
    const double width = *right_crossing - *left_crossing;
    if (!std::isfinite(width) || *left_crossing >= 0.0 || *right_crossing <= 0.0)
        return std::nullopt;
    return Beamwidth{*left_crossing, *right_crossing, width};
}

std::vector<std::optional<double>> angular_frequency_response(const std::vector<AngularRow> &rows, double angle,
                                                              bool relative, double reference_angle) {
    if (!std::isfinite(angle) || !std::isfinite(reference_angle))
        throw std::invalid_argument("response and reference angles must be finite");
    std::vector<std::optional<double>> out;
    out.reserve(rows.size());
    for (std::size_t i = 0; i < rows.size(); ++i) {
        const AngularRow &row = rows[i];
        if (!std::isfinite(row.frequency) || row.frequency <= 0 || (i != 0 && row.frequency <= rows[i - 1].frequency))
            throw std::invalid_argument("response frequencies must be positive and strictly increasing");
        auto level = angular_level_db(row.samples, angle);
        if (level && relative) {
            const auto reference = angular_level_db(row.samples, reference_angle);
            if (reference && std::isfinite(*level - *reference))
                level = *level - *reference;
            else
                level.reset();
        }
        out.push_back(level);
    }
    return out;
}

AcousticFrame normalized_acoustic_frame(AcousticFrame f) {
    if (!finite(f.origin))
        throw std::invalid_argument("Acoustic origin must be finite.");
    f.forward = unit(f.forward);
    const Vec3 v = vcross(f.forward, unit(f.horizontal));
    // |forward x horizontal| = sin(angle): below sqrt(eps) the roll is not determined
    if (std::hypot(v.x, v.y, v.z) <= std::sqrt(std::numeric_limits<double>::epsilon()))
        throw std::invalid_argument("Acoustic axes must not be parallel.");
    f.horizontal = unit(vcross(unit(v), f.forward));
    return f;
}

AcousticFrame default_acoustic_frame(Vec3 origin, Vec3 forward) {
    forward = unit(forward);
    return normalized_acoustic_frame({origin, forward, std::abs(forward.x) < .9 ? Vec3{1, 0, 0} : Vec3{0, 1, 0}});
}

Vec3 vertical_direction(const AcousticFrame &frame) {
    const AcousticFrame f = normalized_acoustic_frame(frame);
    return unit(vcross(f.forward, f.horizontal));
}

std::vector<Vec3> polar_directions(const AcousticFrame &frame, bool vertical, const std::vector<double> &degrees) {
    const AcousticFrame f = normalized_acoustic_frame(frame);
    const Vec3 axis = vertical ? vertical_direction(f) : f.horizontal;
    std::vector<Vec3> out;
    out.reserve(degrees.size());
    for (double a : degrees) {
        if (!std::isfinite(a))
            throw std::invalid_argument("Polar angles must be finite.");
        const double r = a * std::numbers::pi / 180;
        out.push_back(vadd(vscale(f.forward, std::cos(r)), vscale(axis, std::sin(r))));
    }
    return out;
}

} // namespace hsbem
