// Angular and radial building blocks of the wideband FMM translations (host tables).
//
// The quantities are those of modules/bem/fmm/wideband_transfer.hpp:14-104 (sphere rule,
// spherical harmonics, j_l tables, bandlimited resampling between rules with the
// parent/child plane-wave shifts, multipole/local <-> signature spectra, the diagonal
// plane-wave translation); the numerical methods here are this project's own:
//   * sphere rule of order L: Gauss-Legendre in mu = cos(theta) with L+1 nodes
//     (Newton from Tricomi's asymptotic guess, half the nodes mirrored so the rule is
//     exactly symmetric), descending mu, times 2(L+1) equispaced azimuths from 0,
//     polar-major; it integrates products of harmonics through degree L;
//   * orthonormal associated Legendre functions Pbar_l^m (Condon-Shortley phase) in a
//     triangular table (index l(l+1)/2 + m) by the sectoral / first off-diagonal /
//     three-term column recurrences; Y_lm = Pbar_l^|m| e^{i m phi} with
//     Y_{l,-m} = (-1)^m conj(Y_lm), index l*l + l + m;
//   * j_l(x), j_l'(x): upward recurrence where it is stable (x above the order by a
//     margin), otherwise Miller's downward recurrence normalised by the identity
//     sum_l (2l+1) j_l(x)^2 = 1 (no dependence on zeros of j_0 or j_1);
//     j_l' = j_{l-1} - (l+1) j_l / x;
//   * every transform between harmonics and rule samples is separable: an azimuthal
//     DFT per polar ring (table of twiddles) and a Legendre sum per azimuthal order;
//   * the diagonal translation of e^{-ikr}/(4 pi r) (Rokhlin's operator)
//     (-ik)/(16 pi^2) sum_l (2l+1) (-i)^l h_l^(2)(kR) P_l(s.R^) is summed by
//     Clenshaw's recurrence per direction s.
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
// n-point Gauss-Legendre nodes (descending, exactly symmetric) and weights on [-1, 1].
void gauss_legendre(unsigned n, std::vector<double> &nodes, std::vector<double> &weights);

// Pbar_l^m(mu) for 0 <= m <= l <= order, index l(l+1)/2 + m.
std::vector<double> normalized_legendre(unsigned order, double mu);
std::vector<std::complex<double>> spherical_harmonics(unsigned order, Vec3 direction);

struct RegularRadialTable {
    std::vector<double> values, derivatives; // j_l(x), j_l'(x), l = 0..order
};
RegularRadialTable regular_radial_table(unsigned order, double x);

std::vector<std::complex<double>> resample_spectrum(const std::vector<std::complex<double>> &values,
                                                    const SphereRule &source, const SphereRule &target,
                                                    unsigned bandlimit);
std::vector<std::complex<double>> outgoing_to_parent(const std::vector<std::complex<double>> &child,
                                                     const SphereRule &child_rule, const SphereRule &parent_rule,
                                                     unsigned child_bandlimit, double k, Vec3 child_minus_parent);
std::vector<std::complex<double>> incoming_to_child(const std::vector<std::complex<double>> &parent,
                                                    const SphereRule &parent_rule, const SphereRule &child_rule,
                                                    unsigned child_bandlimit, double k, Vec3 child_minus_parent);
// F(s_q) = 4 pi sum_lm i^l M_lm Y_lm(s_q) at every rule direction.
std::vector<std::complex<double>> multipole_to_spectrum(const std::vector<std::complex<double>> &multipoles,
                                                        unsigned order, const SphereRule &rule);
// L_lm = 4 pi (-i)^l sum_q w_q a(s_q) conj(Y_lm(s_q)), l <= order <= rule order.
std::vector<std::complex<double>> spectrum_to_local(const std::vector<std::complex<double>> &incoming,
                                                    const SphereRule &rule, unsigned order);

std::vector<std::complex<double>> diagonal_translation(const SphereRule &rule, unsigned expansion_order, double k,
                                                       Vec3 target_minus_source);

} // namespace hsbem
