// Device-side data layout and the Burton-Miller pair sample.
//
// HBM layout (all SoA, one entry per panel unless noted):
//   geo[22][nt]   v0 xyz | e1=v1-v0 xyz | e2=v2-v0 xyz | unit normal xyz |
//                 curl0 xyz | curl1 xyz | curl2 xyz | area
//   nodes[3][nt]  internal pressure-DOF ids of the panel vertices
//   bounds[5][nt] centroid xyz | outward-rounded ball radius | longest edge
//   shape[nt]     1 when every interior angle is at least 20 degrees
//   A             nv x nv complex, column-major (row + col*nv), rows scaled by 1/lumped mass
//   BD            nv x nd complex, one column per flux-carrying panel
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace hsbem::dev {

struct Cplx {
    double re, im;
};

__device__ __forceinline__ Cplx cadd(Cplx a, Cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cplx cmul(Cplx a, Cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cplx cscale(Cplx a, double s) { return {a.re * s, a.im * s}; }

// sin and cos of x >= 0 for arguments below 2^19 (k*r of any panel pair here):
// Cody-Waite reduction by pi/2 with a three-part 33/33/53-bit split (exact
// products for |n| < 2^20), then Taylor polynomials on [-pi/4, pi/4] truncated
// after the y^15 / y^16 terms (truncation < 5e-17 relative) and the quadrant
// selection. Cody & Waite, "Software Manual for the Elementary Functions" (1980).
__device__ __forceinline__ void sincos_reduced(double x, double *s, double *c) {
    constexpr double two_over_pi = 0.6366197723675814;
    constexpr double p1 = 0x1.921fb54400000p+0, p2 = 0x1.0b4611a600000p-34, p3 = 0x1.3198a2e037073p-69;
    const double n = rint(x * two_over_pi);
    double y = fma(-n, p1, x);
    y = fma(-n, p2, y);
    y = fma(-n, p3, y);
    const double y2 = y * y;
    double ps = 1.0 / 1307674368000.0;           // 1/15!
    ps = fma(ps, -y2, 1.0 / 6227020800.0);        // 1/13!
    ps = fma(ps, -y2, 1.0 / 39916800.0);          // 1/11!
    ps = fma(ps, -y2, 1.0 / 362880.0);            // 1/9!
    ps = fma(ps, -y2, 1.0 / 5040.0);              // 1/7!
    ps = fma(ps, -y2, 1.0 / 120.0);               // 1/5!
    ps = fma(ps, -y2, 1.0 / 6.0);                 // 1/3!
    const double sy = fma(-y * y2, ps, y);
    double pc = 1.0 / 20922789888000.0;           // 1/16!
    pc = fma(pc, -y2, 1.0 / 87178291200.0);       // 1/14!
    pc = fma(pc, -y2, 1.0 / 479001600.0);         // 1/12!
    pc = fma(pc, -y2, 1.0 / 3628800.0);           // 1/10!
    pc = fma(pc, -y2, 1.0 / 40320.0);             // 1/8!
    pc = fma(pc, -y2, 1.0 / 720.0);               // 1/6!
    pc = fma(pc, -y2, 1.0 / 24.0);                // 1/4!
    pc = fma(pc, -y2, 0.5);                       // 1/2!
    const double cy = fma(-y2, pc, 1.0);
    const int q = static_cast<int>(static_cast<long long>(n) & 3);
    const double sa = (q & 1) ? cy : sy, ca = (q & 1) ? sy : cy;
    *s = (q & 2) ? -sa : sa;
    *c = ((q + 1) & 2) ? -ca : ca;
}

// 1/sqrt(v) to full FP64 accuracy: hardware estimate and two Newton steps.
__device__ __forceinline__ double rsqrt_full(double v) {
    double y = __builtin_amdgcn_rsq(v);
    y = y * fma(-0.5 * v * y, y, 1.5);
    y = y * fma(-0.5 * v * y, y, 1.5);
    return y;
}

// One FP64 sample of the Burton-Miller integrands at displacement d = x - y:
//   g   = w exp(-ikr)/(4 pi r)                        (S kernel, weight folded in)
//   a   = -dG/dn_y + i k (n_x.n_y) G                  (-D + alpha * (-k^2 n.n) part of H)
//   b   = -G + (i/k) dG/dn_x                          (-S - alpha K')
// with alpha = -i/k (panel_integrals.cuh:166-198). Coincident pairs have zero
// normal derivatives on a flat panel. wq = w/(4 pi); knn = k (n_x.n_y).
// Branch free: a zero or non-finite distance sets `bad` (the caller reports
// ERR_DISTANCE once) and leaves a non-finite sample.
struct Sample {
    Cplx g, a, b;
};

__device__ __forceinline__ Sample bm_sample(double dx, double dy, double dz, double wq, double k, double inv_k,
                                            double knn, double nsx, double nsy, double nsz, double ntx, double nty,
                                            double ntz, bool coincident, bool &bad) {
    const double r2 = dx * dx + dy * dy + dz * dz;
    bad |= !(r2 > 0.0) || !isfinite(r2);
    Sample s;
    const double rinv = rsqrt_full(r2);
    const double r = r2 * rinv, kr = k * r;
    double sn, cs;
    sincos_reduced(kr, &sn, &cs);
    const double scale = wq * rinv;
    s.g = {cs * scale, -sn * scale};
    const Cplx radial = {s.g.re - s.g.im * kr, s.g.im + s.g.re * kr};
    const double ir2 = rinv * rinv;
    const double fy = coincident ? 0.0 : (dx * nsx + dy * nsy + dz * nsz) * ir2;
    const double fx = coincident ? 0.0 : -(dx * ntx + dy * nty + dz * ntz) * ir2;
    const Cplx dn = cscale(radial, fy), adj = cscale(radial, fx);
    s.a = {-dn.re - knn * s.g.im, -dn.im + knn * s.g.re};
    s.b = {-s.g.re - adj.im * inv_k, -s.g.im + adj.re * inv_k};
    return s;
}

// Pair moments: 9 A moments (P1 x P1), 3 B moments (P1 x DP0) and sum G.
struct Moments {
    Cplx a[9], b[3], g;
};

__device__ __forceinline__ void zero(Moments &m) {
#pragma unroll
    for (int i = 0; i < 9; ++i)
        m.a[i] = {0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i)
        m.b[i] = {0, 0};
    m.g = {0, 0};
}

__device__ __forceinline__ void add_sample(Moments &m, const Sample &s, const double *xt, const double *xs) {
    m.g = cadd(m.g, s.g);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        m.b[i] = cadd(m.b[i], cscale(s.b, xt[i]));
#pragma unroll
        for (int j = 0; j < 3; ++j)
            m.a[3 * i + j] = cadd(m.a[3 * i + j], cscale(s.a, xt[i] * xs[j]));
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void wave_reduce(Moments &m) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        m.a[i].re = wave_sum(m.a[i].re);
        m.a[i].im = wave_sum(m.a[i].im);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        m.b[i].re = wave_sum(m.b[i].re);
        m.b[i].im = wave_sum(m.b[i].im);
    }
    m.g.re = wave_sum(m.g.re);
    m.g.im = wave_sum(m.g.im);
}

__device__ __forceinline__ void atomic_add_c(double *p, Cplx v) {
    unsafeAtomicAdd(p, v.re);
    unsafeAtomicAdd(p + 1, v.im);
}

} // namespace hsbem::dev
