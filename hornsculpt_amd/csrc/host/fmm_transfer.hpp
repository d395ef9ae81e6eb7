// Angular and radial building blocks of the wideband FMM translations.
//
// Restates modules/bem/fmm/wideband_transfer.hpp:14-104 / wideband_transfer.cpp:1-498
// (makeSphereRule, sphericalHarmonics, regularRadialTable, makeAngularTransferPlan,
// resampleSpectrum, outgoingToParent, incomingToChild, multipoleToSpectrum,
// spectrumToLocal, diagonalTranslation):
//   * sphere rule of order L: Gauss-Legendre in cos(theta) (L+1 nodes by Newton
//     on P_{L+1}) times 2(L+1) equispaced azimuths, polar-major; integrates
//     products of harmonics through degree L;
//   * orthonormal complex Y_lm (Condon-Shortley phase), index l*l + l + m, by the
//     stable diagonal / three-term recursions in l for each m >= 0 and
//     Y_{l,-m} = (-1)^m conj(Y_lm);
//   * j_l(x) and j_l'(x): power series for x <= 1, upward recurrence far above the
//     turning point (x > L + 32), normalised Miller downward recurrence otherwise;
//   * separable bandlimited resampling between two sphere rules: azimuthal Fourier
//     sums per polar ring, then normalised associated Legendre functions (index
//     l(l+1)/2 + m) against the rings' Gauss weights; parent/child shifts multiply
//     by exp(+ik d.(c_child - c_parent)) before (outgoing) or exp(-ik d.(...))
//     after (incoming) the resampling;
//   * outgoing F(d) = 4 pi sum i^l M_lm Y_lm(d) and incoming
//     L_lm = 4 pi (-i)^l sum_q w_q a(d_q) conj(Y_lm(d_q));
//   * diagonal (plane-wave) Helmholtz translation of e^{-ikr}/(4 pi r):
//     (-ik)/(16 pi^2) sum_l (2l+1) (-i)^l h_l^(2)(kR) P_l(s.R^) per direction s.
#pragma once

#include "vec.hpp"

#include <complex>
#include <vector>

namespace hsbem {

struct SphereSample {
    Vec3 direction{};
    double weight{};
};

struct SphereRule {
    unsigned order{};
    std::vector<SphereSample> samples; // (order+1) polar rings x 2(order+1) azimuths
};

SphereRule make_sphere_rule(unsigned order);
std::vector<std::complex<double>> spherical_harmonics(unsigned order, Vec3 direction);

struct RegularRadialTable {
    std::vector<double> values, derivatives; // j_l(x), j_l'(x), l = 0..order
};
RegularRadialTable regular_radial_table(unsigned order, double x);

struct PolarSample {
    double cosine{}, weight{}; // ring weight = sample weight * azimuths / (2 pi)
};

struct AngularTransferPlan {
    unsigned source_azimuths{}, target_azimuths{}, bandlimit{};
    std::vector<PolarSample> source_polar, target_polar;
    std::vector<double> recurrence_first, recurrence_second, diagonal_scale; // index l(l+1)/2 + m
};

AngularTransferPlan make_angular_transfer_plan(const SphereRule &source, const SphereRule &target, unsigned bandlimit);
std::vector<std::complex<double>> resample_spectrum(const std::vector<std::complex<double>> &values,
                                                    const AngularTransferPlan &plan);
std::vector<std::complex<double>> resample_spectrum(const std::vector<std::complex<double>> &values,
                                                    const SphereRule &source, const SphereRule &target,
                                                    unsigned bandlimit);
std::vector<std::complex<double>> outgoing_to_parent(const std::vector<std::complex<double>> &child,
                                                     const SphereRule &child_rule, const SphereRule &parent_rule,
                                                     unsigned child_bandlimit, double k, Vec3 child_minus_parent);
std::vector<std::complex<double>> incoming_to_child(const std::vector<std::complex<double>> &parent,
                                                    const SphereRule &parent_rule, const SphereRule &child_rule,
                                                    unsigned child_bandlimit, double k, Vec3 child_minus_parent);
std::vector<std::complex<double>> multipole_to_spectrum(const std::vector<std::complex<double>> &multipoles,
                                                        unsigned order, const SphereRule &rule);
std::vector<std::complex<double>> spectrum_to_local(const std::vector<std::complex<double>> &incoming,
                                                    const SphereRule &rule, unsigned order);

std::vector<std::complex<double>> diagonal_translation(const SphereRule &rule, unsigned expansion_order, double k,
                                                       Vec3 target_minus_source);

} // namespace hsbem
