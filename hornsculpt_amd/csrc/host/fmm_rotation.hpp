// Rotations of spherical-harmonic expansions for the low-frequency FMM
// translations (rotate, translate along z, rotate back).
//
// Restates modules/bem/fmm/spherical_rotation.hpp:8-27 / spherical_rotation.cpp:32-116
// (makeSphericalRotationPlan, rotateToAxis, rotateFromAxis) and the RCR work
// model of modules/bem/fmm/rcr_work.hpp:7-34 (makeRcrQueuePolicy):
//   * Wigner small-d^l(theta) for integer l <= order, built by coupling one
//     spin-1/2 rotation at a time (maximal-spin Clebsch-Gordan coefficients
//     sqrt(i/q), sqrt((q-i)/q)), so every step is an isometry and no factorial
//     or alternating sum appears; azimuth phases exp(+i m phi);
//   * pullback f(Q r) with Q = Rz(phi) Ry(theta), Q e_z = axis:
//     c' = d(theta)^T diag(exp(+i m phi)) c, and its inverse.
// On the device these are per-degree block products (2l+1)^2 per degree.
#pragma once

#include "vec.hpp"

#include <complex>
#include <cstddef>
#include <vector>

namespace hsbem {

inline constexpr unsigned kMaximumRotationOrder = 128;

struct RotationPlan {
    unsigned order{};
    std::vector<std::size_t> degree_offsets; // order + 2 entries into small_d
    std::vector<double> small_d;             // per degree l: (2l+1)^2, row-major, m = -l..l
    std::vector<std::complex<double>> azimuth; // exp(+i m phi), m = -order..order
};

RotationPlan make_rotation_plan(unsigned order, Vec3 axis);
// The two factors of a plan: the Wigner blocks depend on the polar angle
// theta = atan2(|axis_xy|, axis_z) only (no azimuth in the result), the phases
// exp(+i m phi) on the azimuth phi = atan2(axis_y, axis_x).
double rotation_polar_angle(Vec3 axis);
RotationPlan wigner_blocks(unsigned order, double theta);
std::vector<std::complex<double>> rotation_azimuth(unsigned order, Vec3 axis);
// Coefficients indexed l*l + l + m.
std::vector<std::complex<double>> rotate_to_axis(const RotationPlan &plan, const std::vector<std::complex<double>> &c);
std::vector<std::complex<double>> rotate_from_axis(const RotationPlan &plan,
                                                   const std::vector<std::complex<double>> &c);

// Bounded work queue of the LF rotation-coaxial-rotation links: work per link in
// complex coefficient operations (both rotations, the axial product and the
// phases), links per batch and batches per chunk.
struct RcrQueuePolicy {
    unsigned batch_capacity{}, maximum_chunk_batches{};
    std::size_t work_per_link{};
};
RcrQueuePolicy make_rcr_queue_policy(unsigned source_order, unsigned target_order, unsigned reserved_batches);

} // namespace hsbem
