// Device patch estimates of the adaptive near-pair rule.
//
// Replaces AdaptivePatchEstimator::evaluate (modules/bem/cuda/adaptive_patch_estimator.cu:81-345):
// every pending patch of a batch of near pairs is integrated with the two embedded
// outer rules (degree 6 / 8, or 12 / 14) and, for a projection patch, the finer outer
// rule once more with the coarser inner Gauss order (14 instead of 20); the estimate is
// the finer integral with the outer and inner differences in the max norm
// (estimatePairPatchCpu, pair_quadrature.cpp:633-683). The 36 moments are those of
// G = exp(-ikr)/r and of the scaled gradient -(1 + ikr) G (h / r^2) d against the P1 x P1
// basis products.
//
// One wavefront per patch: lanes stride over the patch's sample points (tensor pairs
// or projected-fan points, the same maps as the near assembly), each accumulating the
// 36 complex moments in registers; a butterfly reduction closes each rule. The host
// takes the decisions (subdivision, degree promotion, the largest-error heap).
#include "../host/rules.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

namespace hsbem::dev {
namespace {

constexpr int kEstimatorWaves = 4; // wavefronts (patches) per workgroup

__device__ __forceinline__ Vec3 local_of(const EstimatorPanel &p, const double *b) {
    const Vec3 e1 = vsub(p.fan.v[1], p.fan.v[0]), e2 = vsub(p.fan.v[2], p.fan.v[0]);
    return vadd(vscale(e1, b[1]), vscale(e2, b[2]));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// PROJECTION: the launch's tasks are all projection patches (three rules, fan maps),
// else all tensor patches (two rules).
template <bool PROJECTION>
__global__ __launch_bounds__(64 * kEstimatorWaves) void estimate_patches(EstimatorLaunch L) {
    const int wave = static_cast<int>(threadIdx.x) / 64, lane = static_cast<int>(threadIdx.x) % 64;
    const std::uint32_t task_index = blockIdx.x * kEstimatorWaves + wave;
    if (task_index >= L.count)
        return; // whole wavefronts only: no workgroup barrier below
    const EstimatorTask task = L.tasks[task_index];
    const PairPatch &patch = task.patch;
    const EstimatorPanel test = L.panels[task.test], trial = L.panels[task.trial];
    const double scale = fmax(test.max_edge, trial.max_edge), k = L.k;
    const Vec3 base = vsub(test.fan.v[0], trial.fan.v[0]);
    const bool fine = patch.degree == 14;
    constexpr int roles = PROJECTION ? 3 : 2;
    bool overlap = false;
    // after the reductions, lane l < 36 keeps moment l of the lower (lo), higher (hi)
    // and coarse-inner (co) rules
    double lo_re = 0, lo_im = 0, hi_re = 0, hi_im = 0, co_re = 0, co_im = 0;
    for (int role = 0; role < roles; ++role) {
        // role 0: lower outer rule; 1: higher outer rule; 2: higher outer rule, inner order 14
        const TrianglePoint *rule = L.rules + (fine ? (role == 0 ? 28 : 77) : (role == 0 ? 0 : 12));
        const unsigned count = fine ? (role == 0 ? 49u : 64u) : (role == 0 ? 12u : 16u);
        const unsigned inner = role == 2 ? 14u : 20u;
        const unsigned points = PROJECTION ? patch_point_count(patch, count, inner) : count * count;
        double acc[72];
#pragma unroll
        for (int c = 0; c < 72; ++c)
            acc[c] = 0;
        for (unsigned index = lane; index < points; index += 64) {
            const PairPoint p = PROJECTION ? patch_point(rule, count, patch, index, test.fan, trial.fan, inner)
                                           : map_patch_point(rule[index / count], rule[index % count], patch);
            const Vec3 d = vadd(base, vsub(local_of(test, p.test), local_of(trial, p.trial)));
            const double r2 = vdot(d, d), r = sqrt(r2);
            if (!(r > 0)) {
                overlap = true;
                continue;
            }
            double sn, cs;
            sincos(k * r, &sn, &cs);
            const double gre = cs / r, gim = -sn / r, kr = k * r, f = -scale / r2;
            // -(1 + i kr) g (h / r^2)
            const double dre = (gre - gim * kr) * f, dim = (gim + gre * kr) * f;
            const double kre[4] = {gre, dre * d.x, dre * d.y, dre * d.z};
            const double kim[4] = {gim, dim * d.x, dim * d.y, dim * d.z};
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double w = p.weight * p.test[i] * p.trial[j];
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        acc[2 * (c * 9 + i * 3 + j)] += kre[c] * w;
                        acc[2 * (c * 9 + i * 3 + j) + 1] += kim[c] * w;
                    }
                }
        }
        // butterfly sums: every lane holds every total
        double re = 0, im = 0;
#pragma unroll
        for (int c = 0; c < 36; ++c) {
            const double vr = wave_sum(acc[2 * c]), vi = wave_sum(acc[2 * c + 1]);
            re = c == lane ? vr : re;
            im = c == lane ? vi : im;
        }
        if (role == 0) {
            lo_re = re;
            lo_im = im;
        } else if (role == 1) {
            hi_re = re;
            hi_im = im;
        } else {
            co_re = re;
            co_im = im;
        }
    }
    const bool any_overlap = __any(overlap);
    double outer = 0, inner = 0;
    bool finite = true;
    if (lane < 36) {
        finite = isfinite(hi_re) && isfinite(hi_im);
        outer = hypot(hi_re - lo_re, hi_im - lo_im);
        if (roles == 3)
            inner = hypot(hi_re - co_re, hi_im - co_im);
        L.out[task_index].integral[2 * lane] = hi_re;
        L.out[task_index].integral[2 * lane + 1] = hi_im;
    }
    outer = wave_max(outer);
    inner = wave_max(inner);
    const bool all_finite = __all(finite) && isfinite(outer) && isfinite(inner);
    if (lane == 0) {
        L.out[task_index].outer_error = outer;
        L.out[task_index].inner_error = inner;
        L.out[task_index].status = any_overlap ? 1 : (all_finite ? 0 : 2);
    }
}

} // namespace

void launch_estimate_patches(const EstimatorLaunch &launch, bool projection, hipStream_t stream) {
    if (!launch.count)
        return;
    const dim3 grid((launch.count + kEstimatorWaves - 1) / kEstimatorWaves), block(64 * kEstimatorWaves);
    if (projection)
        hipLaunchKernelGGL(estimate_patches<true>, grid, block, 0, stream, launch);
    else
        hipLaunchKernelGGL(estimate_patches<false>, grid, block, 0, stream, launch);
}

} // namespace hsbem::dev
