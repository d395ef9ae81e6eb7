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

// One FP64 sample of the Burton-Miller integrands at displacement d = x - y:
//   g   = w exp(-ikr)/(4 pi r)                        (S kernel, weight folded in)
//   a   = -dG/dn_y + i k (n_x.n_y) G                  (-D + alpha * (-k^2 n.n) part of H)
//   b   = -G + (i/k) dG/dn_x                          (-S - alpha K')
// with alpha = -i/k (panel_integrals.cuh:166-198). Coincident pairs have zero
// normal derivatives on a flat panel.
struct Sample {
    Cplx g, a, b;
};

__device__ __forceinline__ Sample bm_sample(double dx, double dy, double dz, double w, double k, double nsx, double nsy,
                                            double nsz, double ntx, double nty, double ntz, double nn, bool coincident,
                                            int *error) {
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    Sample s;
    if (!(r > 0.0) || !isfinite(r)) {
        atomicExch(error, ERR_DISTANCE);
        s.g = s.a = s.b = {0, 0};
        return s;
    }
    double sn, cs;
    sincos(k * r, &sn, &cs);
    const double scale = w / (12.56637061435917295385 * r);
    s.g = {cs * scale, -sn * scale};
    const double kr = k * r;
    const Cplx radial = {s.g.re - s.g.im * kr, s.g.im + s.g.re * kr};
    const double ir2 = 1.0 / r2;
    const double fy = coincident ? 0.0 : (dx * nsx + dy * nsy + dz * nsz) * ir2;
    const double fx = coincident ? 0.0 : -(dx * ntx + dy * nty + dz * ntz) * ir2;
    const Cplx dn = cscale(radial, fy), adj = cscale(radial, fx);
    s.a = {-dn.re - k * nn * s.g.im, -dn.im + k * nn * s.g.re};
    s.b = {-s.g.re - adj.im / k, -s.g.im + adj.re / k};
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
