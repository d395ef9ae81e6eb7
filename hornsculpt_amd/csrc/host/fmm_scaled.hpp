// Scaled low-frequency spherical waves and coaxial translation plans of the
// wideband FMM.
//
// Restates modules/bem/fmm/scaled_translation.hpp:8-69 / scaled_translation.cpp:1-399
// (scaledRegularWaves, scaledOutgoingWaves, evaluateScaledMultipole/Local,
// scaledMultipoleToSpectrum, spectrumToScaledLocal, makeScaledAxialPlan,
// applyScaledAxialPlan):
//   * regular basis   R_lm(x) = Jhat_l(k r) (r/a)^l Y_lm(x/r), with
//                     Jhat_l(z) = (2l+1)!! j_l(z) / z^l (-> 1 as z -> 0; DLMF 10.53.1),
//                     built from solid harmonics r^l Y_lm as polynomials with exact
//                     gradients, and grad Jhat_l(kr) = -k^2 x Jhat_{l+1}(kr) / (2l+3);
//   * outgoing basis  O_lm(x) = (a/r)^(l+1) Hhat_l(kr) Y_lm / (a (2l+1)), the radial
//                     factor by the rescaled recurrence of DLMF 10.51.1-2 that never
//                     forms h_l(kr) ~ (kr)^(-l-1);
//   * coaxial plans   translation along +z as one block per |m| (rows l = m..p_t,
//                     columns n = m..p_s, row-major), shared by +m and -m; M2M by
//                     the Green's-identity flux over a sphere of 1.01 a_s, M2L and
//                     L2L by Cauchy projection of the value and radial derivative on
//                     a probe sphere (min(a_t, d/2) for M2L, a_t for L2L), using the
//                     analytic azimuth integral and Gauss-Legendre in the polar angle.
#pragma once

#include "fmm_transfer.hpp"

#include <array>
#include <complex>
#include <cstddef>
#include <vector>

namespace hsbem {

struct WaveValue {
    std::complex<double> value{};
    std::array<std::complex<double>, 3> gradient{};
};

// (order+1)^2 waves, index l*l + l + m, at point x relative to the expansion centre.
std::vector<WaveValue> scaled_regular_waves(unsigned order, double k, double radius, Vec3 point);
std::vector<WaveValue> scaled_outgoing_waves(unsigned order, double k, double radius, Vec3 point);

struct ScaledExpansion {
    unsigned order{};
    double radius{}; // basis normalisation scale a
    std::vector<std::complex<double>> coefficients;
};

WaveValue evaluate_scaled_multipole(const ScaledExpansion &source, double k, Vec3 point);
WaveValue evaluate_scaled_local(const ScaledExpansion &local, double k, Vec3 point);
std::vector<std::complex<double>> scaled_multipole_to_spectrum(const ScaledExpansion &source, double k,
                                                               const SphereRule &rule);
ScaledExpansion spectrum_to_scaled_local(const std::vector<std::complex<double>> &incoming, const SphereRule &rule,
                                         unsigned order, double k, double radius);

enum class AxialKind : int { multipole_to_parent = 0, multipole_to_local = 1, local_to_child = 2 };

struct AxialPlan {
    AxialKind kind{};
    unsigned source_order{}, target_order{}, polar_count{};
    std::vector<std::size_t> block_offsets; // min(p_s, p_t) + 2 entries
    std::vector<std::complex<double>> coefficients;
    double projection_radius{};
};

// distance along +z: child centre - parent centre for multipole_to_parent and
// local_to_child, target centre - source centre for multipole_to_local.
// polar_count >= 2 max(p_s, p_t) + 1; projection_radius 0 selects the default probe.
AxialPlan make_axial_plan(AxialKind kind, unsigned source_order, unsigned target_order, double k,
                          double source_radius, double target_radius, double distance, unsigned polar_count,
                          double projection_radius = 0);
std::vector<std::complex<double>> apply_axial_plan(const AxialPlan &plan,
                                                   const std::vector<std::complex<double>> &source);

} // namespace hsbem
