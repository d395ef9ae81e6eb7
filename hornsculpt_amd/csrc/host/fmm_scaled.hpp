// Scaled low-frequency spherical waves and coaxial translation plans of the
// wideband FMM.
//
// The quantities are those of modules/bem/fmm/scaled_translation.hpp:8-69 (scaled
// regular / outgoing waves with gradients, expansion evaluation, spectra, coaxial
// m-block plans and their application); the derivation and methods are this
// project's own:
//   * regular basis   R_lm(x) = Jhat_l(k r) S_lm(x / a), S_lm(p) = |p|^l Y_lm the
//                     regular solid harmonics, Jhat_l(z) = (2l+1)!! j_l(z) / z^l
//                     (-> 1 as z -> 0) by Miller's downward recurrence in that
//                     normalisation; gradients of S_lm from the degree-lowering ladder
//                     identities, grad Jhat_l(k r) = -k^2 x Jhat_{l+1}(k r) / (2l+3);
//   * outgoing basis  O_lm(x) = (a/r)^(l+1) Hhat_l(kr) Y_lm / (a (2l+1)) with
//                     Hhat_l(z) = -i z^(l+1) h_l^(2)(z) / (2l-1)!! (-> 1 as z -> 0), by
//                     the upward recurrence of h_l in that normalisation; radial
//                     derivative from h_l' = h_{l-1} - (l+1) h_l / z;
//   * coaxial plans   translation along +z as one block per |m| (rows l = m..p_t,
//                     columns n = m..p_s, row-major), shared by +m and -m; each
//                     Gauss-Legendre ring of a projection sphere contributes a rank-2
//                     update (target weights x source values and radial slopes):
//                     Green's identity over a sphere of 1.01 a_s for the multipole
//                     shift, least-squares (Cauchy) projection of value and radial
//                     slope onto the target's regular basis on a sphere of radius
//                     min(a_t, d/2) (M2L) or a_t (L2L) otherwise; the azimuthal
//                     integral of a coaxial term is exactly 2 pi.
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
