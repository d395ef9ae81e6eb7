#include "fmm_rotation.hpp"

#include <algorithm>
#include <cmath>
#include <math.h>
#include <stdexcept>

namespace hsbem {
namespace {

// sin and cos of one argument through the C library's joint routine: the
// reference build evaluates its sin/cos pairs that way, and the joint routine
// can differ from separate calls in the last bit.
void sin_cos(double x, double *s, double *c) { ::sincos(x, s, c); }

void check(const RotationPlan &p, const std::vector<std::complex<double>> &c) {
    if (p.order > kMaximumRotationOrder || c.size() != static_cast<std::size_t>(p.order + 1) * (p.order + 1) ||
        p.degree_offsets.size() != p.order + 2 || p.degree_offsets.front() != 0 ||
        p.degree_offsets.back() != p.small_d.size() || p.azimuth.size() != 2 * p.order + 1)
        throw std::invalid_argument("rotation: coefficient count does not match the order");
    for (unsigned l = 0; l <= p.order; ++l)
        if (p.degree_offsets[l] > p.degree_offsets[l + 1] ||
            p.degree_offsets[l + 1] - p.degree_offsets[l] != static_cast<std::size_t>(2 * l + 1) * (2 * l + 1))
            throw std::invalid_argument("rotation: a degree block has the wrong size");
    for (const auto &v : c)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw std::invalid_argument("rotation: source coefficients must be finite");
    for (double v : p.small_d)
        if (!std::isfinite(v))
            throw std::invalid_argument("rotation: a Wigner coefficient is not finite");
    for (const auto &v : p.azimuth)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw std::invalid_argument("rotation: an azimuth phase is not finite");
}

} // namespace

double rotation_polar_angle(Vec3 axis) { return std::atan2(std::hypot(axis.x, axis.y), axis.z); }

std::vector<std::complex<double>> rotation_azimuth(unsigned order, Vec3 axis) {
    const double phi = std::atan2(axis.y, axis.x);
    std::vector<std::complex<double>> out;
    out.reserve(2 * order + 1);
    for (int m = -static_cast<int>(order); m <= static_cast<int>(order); ++m) {
        double sm, cm;
        sin_cos(m * phi, &sm, &cm);
        out.emplace_back(cm, sm);
    }
    return out;
}

RotationPlan make_rotation_plan(unsigned order, Vec3 axis) {
    if (order > kMaximumRotationOrder)
        throw std::invalid_argument("rotation order is beyond the supported range");
    const double len = std::hypot(axis.x, axis.y, axis.z);
    if (!std::isfinite(len) || len <= 0)
        throw std::invalid_argument("rotation axis must be finite and nonzero");
    RotationPlan p = wigner_blocks(order, rotation_polar_angle(axis));
    p.azimuth = rotation_azimuth(order, axis);
    return p;
}

RotationPlan wigner_blocks(unsigned order, double theta) {
    if (order > kMaximumRotationOrder || !std::isfinite(theta))
        throw std::invalid_argument("Spherical rotation order or polar angle is invalid.");
    double sh, ch;
    sin_cos(theta / 2, &sh, &ch);
    RotationPlan p;
    p.order = order;
    p.degree_offsets = {0, 1};
    p.small_d = {1};
    // d^{q/2} from d^{(q-1)/2} and the spin-1/2 rotation [[c, -s], [s, c]]:
    // d_q[i][j] = sum over the four neighbours with coupling weights
    // u_i = sqrt(i/q) (from index i-1) and w_i = sqrt((q-i)/q) (from index i).
    std::vector<double> prev{1};
    for (unsigned q = 1; q <= 2 * order; ++q) {
        const std::size_t dim = static_cast<std::size_t>(q) + 1;
        std::vector<double> cur(dim * dim), u(dim), w(dim);
        for (unsigned i = 0; i <= q; ++i) {
            u[i] = std::sqrt(static_cast<double>(i) / q);
            w[i] = std::sqrt(static_cast<double>(q - i) / q);
        }
        for (unsigned i = 0; i <= q; ++i)
            for (unsigned j = 0; j <= q; ++j) {
                double v = 0;
                if (i && j)
                    v += u[i] * u[j] * ch * prev[(static_cast<std::size_t>(i) - 1) * q + j - 1];
                if (i && j < q)
                    v -= u[i] * w[j] * sh * prev[(static_cast<std::size_t>(i) - 1) * q + j];
                if (i < q && j)
                    v += w[i] * u[j] * sh * prev[static_cast<std::size_t>(i) * q + j - 1];
                if (i < q && j < q)
                    v += w[i] * w[j] * ch * prev[static_cast<std::size_t>(i) * q + j];
                cur[static_cast<std::size_t>(i) * dim + j] = v;
            }
        if (q % 2 == 0) { // integer degree l = q/2
            p.small_d.insert(p.small_d.end(), cur.begin(), cur.end());
            p.degree_offsets.push_back(p.small_d.size());
        }
        prev = std::move(cur);
    }
    return p;
}

std::vector<std::complex<double>> rotate_to_axis(const RotationPlan &p, const std::vector<std::complex<double>> &c) {
    check(p, c);
    std::vector<std::complex<double>> out(c.size());
    for (unsigned l = 0; l <= p.order; ++l) {
        const unsigned count = 2 * l + 1;
        const std::size_t base = static_cast<std::size_t>(l) * l;
        for (unsigned t = 0; t < count; ++t)
            for (unsigned s = 0; s < count; ++s)
                out[base + t] += p.small_d[p.degree_offsets[l] + static_cast<std::size_t>(s) * count + t] *
                                 p.azimuth[p.order + s - l] * c[base + s];
    }
    return out;
}

std::vector<std::complex<double>> rotate_from_axis(const RotationPlan &p, const std::vector<std::complex<double>> &c) {
    check(p, c);
    std::vector<std::complex<double>> out(c.size());
    for (unsigned l = 0; l <= p.order; ++l) {
        const unsigned count = 2 * l + 1;
        const std::size_t base = static_cast<std::size_t>(l) * l;
        for (unsigned t = 0; t < count; ++t) {
            for (unsigned s = 0; s < count; ++s)
                out[base + t] += p.small_d[p.degree_offsets[l] + static_cast<std::size_t>(t) * count + s] * c[base + s];
            out[base + t] *= std::conj(p.azimuth[p.order + t - l]);
        }
    }
    return out;
}

RcrQueuePolicy make_rcr_queue_policy(unsigned source_order, unsigned target_order, unsigned reserved) {
    constexpr std::size_t kMaximumChunkWork = 64 * 1024 * 1024;
    constexpr unsigned kMaximumQueuedBatches = 32;
    if (source_order > 128 || target_order > 128 || !reserved || reserved > 256)
        throw std::invalid_argument("translation queue: orders or batch capacity out of range");
    const std::size_t src = static_cast<std::size_t>(source_order + 1) * (source_order + 1),
                      tgt = static_cast<std::size_t>(target_order + 1) * (target_order + 1),
                      work = src * (2 * source_order + 2) + tgt * (2 * target_order + 2 + source_order + 1);
    const auto capacity = static_cast<unsigned>(std::min<std::size_t>(reserved, kMaximumChunkWork / work));
    if (!capacity)
        throw std::runtime_error("translation queue: one link needs more work than a batch holds");
    const auto chunks =
        static_cast<unsigned>(std::min<std::size_t>(kMaximumQueuedBatches, kMaximumChunkWork / (capacity * work)));
    return {capacity, chunks, work};
}

} // namespace hsbem
