// Plane-wave (high-frequency) levels of the wideband FMM operator, FP64.
//
// Replaces the device side of AngularTransferOperator and the diagonal M2L of
// SupportFmmOperator (modules/bem/cuda/wideband_angular.cu:1-513,
// support_fmm_operator.cu:753-778 and :1103-1209). A plane-wave level stores per
// node the far-field signature on its sphere rule: (q+1) Gauss rings in cos(theta)
// x 2(q+1) equispaced azimuths, polar-major. Every transfer between two rules, or
// between a rule and a scaled spherical expansion, is separable:
//   fourier     per (link, ring): the azimuthal Fourier coefficients
//               A_m = sum_a v_a exp(-i m phi_a), |m| <= B, after the optional
//               centre shift exp(i s . d_a) of every sample (direct sums from an
//               LDS twiddle table: B is at most the level's order, so the DFT costs
//               what an FFT plus mode selection would at these lengths);
//   project     c_lm = sum_rings w P_l|m|(mu) A_m (x (-1)^m for m < 0): the
//               spherical-harmonic coefficients; for the hand-off to a scaled
//               spherical child they are scaled by 4 pi (ka)^l / (2l+1)!! (-i)^l and
//               added to its local expansion;
//   evaluate    per (link, output ring): G_m = sum_l c_lm P_l|m|(mu) (the hand-off
//               from a scaled spherical child first scales its multipole by
//               4 pi (ka)^l / (2l+1)!! i^l);
//   synthesize  per (destination, sample): sum over its links of
//               exp(i s . d) sum_m G_m exp(i m phi), written or added;
//   diagonal    the plane-wave M2L: L_t(d) = sum_links T_class(d) F_s(d).
#include "fmm_kernels.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace hsbem::dev {
namespace {

constexpr double kPi = 3.14159265358979323846;

struct C2 {
    double re, im;
};
__device__ __forceinline__ C2 operator+(C2 a, C2 b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ C2 operator*(C2 a, C2 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ C2 operator*(double s, C2 a) { return {s * a.re, s * a.im}; }

__device__ __forceinline__ std::size_t tri(int l, int m) { return static_cast<std::size_t>(l) * (l + 1) / 2 + m; }

// exp(2 pi i a / n), a = 0..n-1, into LDS
__device__ __forceinline__ void twiddles(C2 *tw, int n) {
    for (int a = threadIdx.x; a < n; a += blockDim.x) {
        double s, c;
        sincospi(2.0 * a / n, &s, &c);
        tw[a] = {c, s};
    }
}

// 4 pi (ka)^l / (2l+1)!! times i^l (outgoing) or (-i)^l (incoming)
__device__ __forceinline__ C2 handoff_scale(int l, double ka, bool outgoing) {
    double f = 4 * kPi;
    for (int j = 1; j <= l; ++j)
        f *= ka / (2.0 * j + 1);
    const int quadrant = outgoing ? l % 4 : (4 - l % 4) % 4;
    return quadrant == 0 ? C2{f, 0} : quadrant == 1 ? C2{0, f} : quadrant == 2 ? C2{-f, 0} : C2{0, -f};
}

__device__ __forceinline__ void split_lm(int c, int &l, int &m) {
    l = static_cast<int>(sqrt(static_cast<double>(c)));
    while (l * l > c)
        --l;
    while ((l + 1) * (l + 1) <= c)
        ++l;
    m = c - l * l - l;
}

// Block (ring, link): A[ring][m + B], |m| <= B, of the link's source signature.
__global__ __launch_bounds__(256) void hf_fourier(HfTransferLaunch L, bool shifted) {
    __shared__ C2 tw[kMaxHfAzimuths];
    __shared__ C2 v[kMaxHfAzimuths];
    const HfLink lk = L.links[blockIdx.y];
    const HfRule R = L.rules[lk.in_rule];
    const int r = blockIdx.x;
    if (r >= static_cast<int>(R.rings))
        return;
    const int n = static_cast<int>(R.azimuths), B = static_cast<int>(lk.order);
    twiddles(tw, n);
    __syncthreads();
    const double mu = L.polar[R.polar + 3 * r], rho = L.polar[R.polar + 3 * r + 1];
    const C2 *src = reinterpret_cast<const C2 *>(L.source_expansions) + lk.source + static_cast<std::size_t>(r) * n;
    for (int a = threadIdx.x; a < n; a += blockDim.x) {
        C2 x = src[a];
        if (shifted) {
            const double arg = rho * (tw[a].re * lk.shift[0] + tw[a].im * lk.shift[1]) + mu * lk.shift[2];
            double s, c;
            sincos(arg, &s, &c);
            x = C2{c, s} * x;
        }
        v[a] = x;
    }
    __syncthreads();
    C2 *out = reinterpret_cast<C2 *>(L.fourier_ws) + lk.fourier + static_cast<std::size_t>(r) * (2 * B + 1);
    for (int e = threadIdx.x; e < 2 * B + 1; e += blockDim.x) {
        const int m = e - B;
        const int step = ((-m) % n + n) % n; // exp(-i m phi_a) = tw[(-m a) mod n]
        double re = 0, im = 0;
        int idx = 0;
        for (int a = 0; a < n; ++a) {
            const C2 t = tw[idx], x = v[a];
            re += t.re * x.re - t.im * x.im;
            im += t.re * x.im + t.im * x.re;
            idx += step;
            if (idx >= n)
                idx -= n;
        }
        out[e] = {re, im};
    }
}

// Block per link: c_lm = sum_rings w P_l|m| A_m; hand-off: scaled and added to the
// target's local expansion, else written to the harmonic workspace.
__global__ __launch_bounds__(256) void hf_project(HfTransferLaunch L, bool handoff) {
    const HfLink lk = L.links[blockIdx.x];
    const HfRule R = L.rules[lk.in_rule];
    const int B = static_cast<int>(lk.order), rings = static_cast<int>(R.rings);
    const std::size_t row = tri(static_cast<int>(R.bmax), static_cast<int>(R.bmax)) + 1;
    const C2 *A = reinterpret_cast<const C2 *>(L.fourier_ws) + lk.fourier;
    for (int c = threadIdx.x; c < (B + 1) * (B + 1); c += blockDim.x) {
        int l, m;
        split_lm(c, l, m);
        const int am = m < 0 ? -m : m;
        const double sign = m < 0 && (am & 1) ? -1.0 : 1.0;
        double re = 0, im = 0;
        for (int r = 0; r < rings; ++r) {
            const double p = L.polar[R.polar + 3 * r + 2] * L.legendre[R.legendre + r * row + tri(l, am)];
            const C2 a = A[static_cast<std::size_t>(r) * (2 * B + 1) + m + B];
            re += p * a.re;
            im += p * a.im;
        }
        const C2 sum{sign * re, sign * im};
        if (handoff) {
            C2 *t = reinterpret_cast<C2 *>(L.target_expansions) + lk.target + c;
            *t = *t + handoff_scale(l, lk.ka, false) * sum;
        } else {
            reinterpret_cast<C2 *>(L.harmonic_ws)[lk.harmonic + c] = sum;
        }
    }
}

// Block (output ring, link): G[ring][m + B] = sum_l c_lm P_l|m|(mu_ring).
__global__ __launch_bounds__(256) void hf_evaluate(HfTransferLaunch L, bool handoff) {
    const HfLink lk = L.links[blockIdx.y];
    const HfRule R = L.rules[lk.out_rule];
    const int r = blockIdx.x;
    if (r >= static_cast<int>(R.rings))
        return;
    const int B = static_cast<int>(lk.order);
    const std::size_t row = tri(static_cast<int>(R.bmax), static_cast<int>(R.bmax)) + 1;
    const double *P = L.legendre + R.legendre + r * row;
    const C2 *c = handoff ? reinterpret_cast<const C2 *>(L.source_expansions) + lk.source
                          : reinterpret_cast<const C2 *>(L.harmonic_ws) + lk.harmonic;
    C2 *out = reinterpret_cast<C2 *>(L.fourier_ws) + lk.fourier + static_cast<std::size_t>(r) * (2 * B + 1);
    for (int e = threadIdx.x; e < 2 * B + 1; e += blockDim.x) {
        const int m = e - B, am = m < 0 ? -m : m;
        const double sign = m < 0 && (am & 1) ? -1.0 : 1.0;
        C2 g{0, 0};
        for (int l = am; l <= B; ++l) {
            C2 v = c[static_cast<std::size_t>(l) * l + l + m];
            if (handoff)
                v = handoff_scale(l, lk.ka, true) * v;
            g = g + (sign * P[tri(l, am)]) * v;
        }
        out[e] = g;
    }
}

// Block (sample tile, destination): sum over the destination's links of
// exp(i s . d) sum_m G_m exp(i m phi), written (add = false) or added.
__global__ __launch_bounds__(256) void hf_synthesize(HfTransferLaunch L, bool shifted, bool add) {
    __shared__ C2 tw[kMaxHfAzimuths];
    const FmmDestination dst = L.destinations[blockIdx.y];
    const HfRule R = L.rules[L.links[dst.first_link].out_rule];
    const int n = static_cast<int>(R.azimuths);
    const int samples = static_cast<int>(R.rings) * n;
    if (static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) >= samples)
        return;
    twiddles(tw, n);
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= samples)
        return;
    const int r = s / n, a = s % n;
    const double mu = L.polar[R.polar + 3 * r], rho = L.polar[R.polar + 3 * r + 1];
    C2 total{0, 0};
    for (std::uint32_t i = 0; i < dst.link_count; ++i) {
        const HfLink lk = L.links[dst.first_link + i];
        const int B = static_cast<int>(lk.order);
        const C2 *G = reinterpret_cast<const C2 *>(L.fourier_ws) + lk.fourier + static_cast<std::size_t>(r) * (2 * B + 1);
        int idx = ((-B * a) % n + n) % n; // exp(i m phi_a) = tw[(m a) mod n], from m = -B
        double re = 0, im = 0;
        for (int e = 0; e <= 2 * B; ++e) {
            const C2 t = tw[idx], g = G[e];
            re += t.re * g.re - t.im * g.im;
            im += t.re * g.im + t.im * g.re;
            idx += a;
            if (idx >= n)
                idx -= n;
        }
        C2 val{re, im};
        if (shifted) {
            const double arg = rho * (tw[a].re * lk.shift[0] + tw[a].im * lk.shift[1]) + mu * lk.shift[2];
            double sn, cs;
            sincos(arg, &sn, &cs);
            val = C2{cs, sn} * val;
        }
        total = total + val;
    }
    C2 *out = reinterpret_cast<C2 *>(L.target_expansions) + dst.target + s;
    *out = add ? *out + total : total;
}

// Plane-wave M2L: thread per sample of a destination, sum over its links.
__global__ __launch_bounds__(256) void hf_diagonal(HfDiagonalLaunch L) {
    const FmmDestination dst = L.destinations[blockIdx.y];
    const std::uint64_t s = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= L.samples)
        return;
    const C2 *F = reinterpret_cast<const C2 *>(L.source_expansions);
    const C2 *T = reinterpret_cast<const C2 *>(L.coefficients);
    C2 total{0, 0};
    for (std::uint32_t i = 0; i < dst.link_count; ++i) {
        const HfDiagonalLink lk = L.links[dst.first_link + i];
        total = total + T[lk.coefficients + s] * F[lk.source + s];
    }
    reinterpret_cast<C2 *>(L.target_expansions)[dst.target + s] = total;
}

// direction of sample s of a rule (ring r = s / n, azimuth a = s % n) and its weight
__device__ __forceinline__ void sample_direction(const double *polar, const HfRule &R, int s, double d[3], double &w) {
    const int n = static_cast<int>(R.azimuths), r = s / n, a = s % n;
    const double mu = polar[R.polar + 3 * r], rho = polar[R.polar + 3 * r + 1];
    double sn, cs;
    sincospi(2.0 * a / n, &sn, &cs);
    d[0] = rho * cs;
    d[1] = rho * sn;
    d[2] = mu;
    w = polar[R.polar + 3 * r + 2];
}

// Block (sample tile, leaf): thread per sample of the leaf's rule.
__global__ __launch_bounds__(256) void hf_leaf_emit(HfLeafLaunch L, const double *p) {
    const HfLeaf leaf = L.leaves[blockIdx.y];
    const HfRule R = L.rules[leaf.rule];
    const int samples = static_cast<int>(R.rings * R.azimuths);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= samples)
        return;
    double d[3], w;
    sample_direction(L.polar, R, s, d, w);
    const double k = L.k;
    double re = 0, im = 0;
    for (std::uint32_t i = 0; i < leaf.item_count; ++i) {
        const HfLeafItem it = L.items[leaf.first_item + i];
        C2 pc[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            pc[c] = {0, 0};
            if (it.corner[c] >= 0) {
                const std::uint32_t dof = L.leaf_dofs[leaf.first_dof + it.corner[c]];
                pc[c] = {p[2 * dof], p[2 * dof + 1]};
            }
        }
        const double nd = it.normal[0] * d[0] + it.normal[1] * d[1] + it.normal[2] * d[2];
        for (std::uint32_t q = 0; q < L.rule_points; ++q) {
            const double b0 = L.rule[4 * q], b1 = L.rule[4 * q + 1], b2 = L.rule[4 * q + 2];
            const double wq = L.rule[4 * q + 3] * it.area2 * nd;
            double arg = 0;
#pragma unroll
            for (int x = 0; x < 3; ++x)
                arg += d[x] * (it.v0[x] + b1 * it.e1[x] + b2 * it.e2[x] - leaf.center[x]);
            double sn, cs;
            sincos(k * arg, &sn, &cs);
            const C2 rho = b0 * pc[0] + b1 * pc[1] + b2 * pc[2];
            const C2 v = C2{cs, sn} * rho;
            re += wq * v.re;
            im += wq * v.im;
        }
    }
    // times i k
    reinterpret_cast<C2 *>(L.expansions)[leaf.multipole + s] = {-k * im, k * re};
}

// Block per item: g_q = sum_d w_d (n.d - 1) exp(-i k d.(y_q - c)) L(d) for every
// rule point q of the panel (16 points per pass), reduced over the block.
__global__ __launch_bounds__(256) void hf_leaf_field(HfLeafLaunch L) {
    constexpr int kPoints = 16;
    __shared__ double part[2][kPoints][4]; // per point, per wavefront partial (re, im)
    const HfLeafItem it = L.items[blockIdx.x];
    const HfLeaf leaf = L.leaves[it.leaf];
    const HfRule R = L.rules[leaf.rule];
    const int samples = static_cast<int>(R.rings * R.azimuths), Q = static_cast<int>(L.rule_points);
    const C2 *Lc = reinterpret_cast<const C2 *>(L.expansions) + leaf.local;
    const double k = L.k;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    for (int q0 = 0; q0 < Q; q0 += kPoints) {
        double y[kPoints][3];
#pragma unroll
        for (int q = 0; q < kPoints; ++q) {
            const int qq = q0 + q < Q ? q0 + q : 0;
            const double b1 = L.rule[4 * qq + 1], b2 = L.rule[4 * qq + 2];
#pragma unroll
            for (int x = 0; x < 3; ++x)
                y[q][x] = it.v0[x] + b1 * it.e1[x] + b2 * it.e2[x] - leaf.center[x];
        }
        C2 acc[kPoints];
#pragma unroll
        for (int q = 0; q < kPoints; ++q)
            acc[q] = {0, 0};
        for (int s = threadIdx.x; s < samples; s += blockDim.x) {
            double d[3], w;
            sample_direction(L.polar, R, s, d, w);
            const double nd = it.normal[0] * d[0] + it.normal[1] * d[1] + it.normal[2] * d[2];
            const C2 a = (w * (nd - 1)) * Lc[s];
#pragma unroll
            for (int q = 0; q < kPoints; ++q) {
                if (q0 + q < Q) {
                    double sn, cs;
                    sincos(-k * (d[0] * y[q][0] + d[1] * y[q][1] + d[2] * y[q][2]), &sn, &cs);
                    acc[q] = acc[q] + C2{cs, sn} * a;
                }
            }
        }
        __syncthreads(); // part[] of the previous pass consumed
#pragma unroll
        for (int q = 0; q < kPoints; ++q) {
            double re = acc[q].re, im = acc[q].im;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                re += __shfl_xor(re, off, 64);
                im += __shfl_xor(im, off, 64);
            }
            if (lane == 0) {
                part[0][q][wave] = re;
                part[1][q][wave] = im;
            }
        }
        __syncthreads();
        if (threadIdx.x < kPoints && q0 + static_cast<int>(threadIdx.x) < Q) {
            const int q = threadIdx.x;
            double re = 0, im = 0;
            for (int v = 0; v < static_cast<int>(blockDim.x) / 64; ++v) {
                re += part[0][q][v];
                im += part[1][q][v];
            }
            reinterpret_cast<C2 *>(L.field)[static_cast<std::size_t>(blockIdx.x) * Q + q0 + q] = {re, im};
        }
    }
}

// Block per leaf, thread per DOF: y_i += (1/m_i) sum over the leaf's panels holding i
// of sum_q 2A w_q phi_i(y_q) g_q (panels in item order: deterministic).
__global__ __launch_bounds__(256) void hf_leaf_gather(HfLeafLaunch L, double *y) {
    const HfLeaf leaf = L.leaves[blockIdx.x];
    const C2 *g = reinterpret_cast<const C2 *>(L.field);
    const int Q = static_cast<int>(L.rule_points);
    for (std::uint32_t j = threadIdx.x; j < leaf.dof_count; j += blockDim.x) {
        C2 sum{0, 0};
        for (std::uint32_t i = 0; i < leaf.item_count; ++i) {
            const std::uint32_t item = leaf.first_item + i;
            const HfLeafItem &it = L.items[item];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (it.corner[c] != static_cast<std::int32_t>(j))
                    continue;
                for (int q = 0; q < Q; ++q)
                    sum = sum + (L.rule[4 * q + 3] * it.area2 * L.rule[4 * q + c]) * g[static_cast<std::size_t>(item) * Q + q];
            }
        }
        const std::uint32_t dof = L.leaf_dofs[leaf.first_dof + j];
        const double scale = L.row_scale[leaf.first_dof + j];
        y[2 * dof] += scale * sum.re;
        y[2 * dof + 1] += scale * sum.im;
    }
}

} // namespace

void launch_hf_leaf_emit(const HfLeafLaunch &launch, const double *p, hipStream_t stream) {
    if (launch.leaf_count)
        hipLaunchKernelGGL(hf_leaf_emit, dim3((launch.max_samples + 255) / 256, launch.leaf_count), dim3(256), 0, stream,
                           launch, p);
}

void launch_hf_leaf_test(const HfLeafLaunch &launch, double *y, hipStream_t stream) {
    if (!launch.leaf_count)
        return;
    hipLaunchKernelGGL(hf_leaf_field, dim3(launch.item_count), dim3(256), 0, stream, launch);
    hipLaunchKernelGGL(hf_leaf_gather, dim3(launch.leaf_count), dim3(256), 0, stream, launch, y);
}

void launch_hf_transfer(const HfTransferLaunch &launch, HfTransferKind kind, std::uint32_t link_count,
                        std::uint32_t destination_count, std::uint32_t in_rings, std::uint32_t out_rings,
                        std::uint32_t out_samples, hipStream_t stream) {
    if (!link_count)
        return;
    const bool up = kind == HfTransferKind::up_handoff || kind == HfTransferKind::up_resample;
    if (kind != HfTransferKind::up_handoff)
        hipLaunchKernelGGL(hf_fourier, dim3(in_rings, link_count), dim3(256), 0, stream, launch, !up);
    if (kind == HfTransferKind::down_handoff) {
        hipLaunchKernelGGL(hf_project, dim3(link_count), dim3(256), 0, stream, launch, true);
        return;
    }
    if (kind != HfTransferKind::up_handoff)
        hipLaunchKernelGGL(hf_project, dim3(link_count), dim3(256), 0, stream, launch, false);
    hipLaunchKernelGGL(hf_evaluate, dim3(out_rings, link_count), dim3(256), 0, stream, launch,
                       kind == HfTransferKind::up_handoff);
    hipLaunchKernelGGL(hf_synthesize, dim3((out_samples + 255) / 256, destination_count), dim3(256), 0, stream, launch,
                       up, !up);
}

void launch_hf_diagonal(const HfDiagonalLaunch &launch, std::uint32_t destination_count, hipStream_t stream) {
    if (destination_count && launch.samples)
        hipLaunchKernelGGL(hf_diagonal, dim3(static_cast<std::uint32_t>((launch.samples + 255) / 256), destination_count),
                           dim3(256), 0, stream, launch);
}

} // namespace hsbem::dev
