// Burton-Miller Galerkin assembly kernels, gfx950.
//
// Replaces the reference kernels assembleRegular / assembleSingular /
// assembleNear / addMass (modules/bem/cuda/cuda_solver.cu:266-540) with:
//   regular_tiles_sym  one workgroup per unordered pair of clusters of up to 64
//                   spatially sorted panels; panel geometry in LDS, the rules as
//                   kernel arguments; each panel pair integrated once for both
//                   orientations, every 3x3 pair block accumulated in two LDS node
//                   tiles (ds_add) flushed once per tile to HBM, so global atomics
//                   drop from 18 per ordered pair to about one per tile entry.
//   touching_jobs   one wavefront per touching pair: Sauter-Schwab (u,v,w) tuples
//                   over lanes with the radial coordinate innermost (every sector map
//                   is linear in r), shuffle reduction, one lane scatters. Always FP64.
//   near_patches    one thread per tensor near patch (16x16 or 64x64 points of the
//                   patch's test and trial sub-triangles), two-stage contraction as
//                   for regular pairs with the affine sub-triangle maps applied on
//                   the fly; one scatter per patch.
//   near_jobs       one wavefront per near job holding a projection-fan patch,
//                   patch points over lanes.
//   mass            1/2 M_pp into A, -alpha 1/2 M_pq into the flux columns.
// Precision: the sample type R is double (plain form) or float (mixed form:
// FP32 samples and accumulation, FP64 geometry differences, FP32 matrix).
#include "../host/plan.hpp"
#include "../host/qualification.hpp"
#include "../host/rules.hpp"
#include "device.hpp"
#include "kernels.hpp"

#include <stdexcept>
#include <string>
#include <type_traits>

namespace hsbem::dev {
namespace {

constexpr int kTilePanels = 64;  // panels per cluster
constexpr int kTileNodes = 64;   // distinct pressure nodes per cluster
constexpr int kTileThreads = 512;

template <class R> struct CplxT {
    R re, im;
};

template <class R> __device__ __forceinline__ CplxT<R> cadd_t(CplxT<R> a, CplxT<R> b) {
    return {a.re + b.re, a.im + b.im};
}
template <class R> __device__ __forceinline__ CplxT<R> cscale_t(CplxT<R> a, R s) { return {a.re * s, a.im * s}; }

template <class R> struct MomentsT {
    CplxT<R> a[9], b[3], g;
};

template <class R> __device__ __forceinline__ void zero_t(MomentsT<R> &m) {
#pragma unroll
    for (int i = 0; i < 9; ++i)
        m.a[i] = {0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i)
        m.b[i] = {0, 0};
    m.g = {0, 0};
}

template <class R> struct SampleT {
    CplxT<R> g, a, b;
};

// Per-pair constants of the pair kernel in the sample precision.
template <class R> struct PairConst {
    R k, inv_k, k_rev, knn, nsx, nsy, nsz, ntx, nty, ntz;
};

// FP32 Burton-Miller sample (mixed form): hardware rsqrt, the phase reduced in
// revolutions (v_fract) and hardware sin/cos of 2 pi x. Same algebra and the same
// branch-free distance check as bm_sample.
__device__ __forceinline__ SampleT<float> sample(float dx, float dy, float dz, float wq, const PairConst<float> &c,
                                                 bool coincident, bool &bad) {
    const float r2 = dx * dx + dy * dy + dz * dz;
    bad |= !(r2 > 0.f) || !isfinite(r2);
    SampleT<float> s;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    const float f = __builtin_amdgcn_fractf(r * c.k_rev);
    const float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);
    const float kr = c.k * r, scale = wq * rinv;
    s.g = {cs * scale, -sn * scale};
    const CplxT<float> radial = {s.g.re - s.g.im * kr, s.g.im + s.g.re * kr};
    const float ir2 = rinv * rinv;
    const float fy = coincident ? 0.f : (dx * c.nsx + dy * c.nsy + dz * c.nsz) * ir2;
    const float fx = coincident ? 0.f : -(dx * c.ntx + dy * c.nty + dz * c.ntz) * ir2;
    s.a = {-radial.re * fy - c.knn * s.g.im, -radial.im * fy + c.knn * s.g.re};
    s.b = {-s.g.re - radial.im * fx * c.inv_k, -s.g.im + radial.re * fx * c.inv_k};
    return s;
}

__device__ __forceinline__ SampleT<double> sample(double dx, double dy, double dz, double wq,
                                                  const PairConst<double> &c, bool coincident, bool &bad) {
    const Sample s = bm_sample(dx, dy, dz, wq, c.k, c.inv_k, c.knn, c.nsx, c.nsy, c.nsz, c.ntx, c.nty, c.ntz,
                               coincident, bad);
    return {{s.g.re, s.g.im}, {s.a.re, s.a.im}, {s.b.re, s.b.im}};
}

template <class R>
__device__ __forceinline__ PairConst<R> pair_const(const double *gt, const double *gs, double k) {
    const double nn = gt[G_N] * gs[G_N] + gt[G_N + 1] * gs[G_N + 1] + gt[G_N + 2] * gs[G_N + 2];
    return {static_cast<R>(k), static_cast<R>(1.0 / k), static_cast<R>(k * 0.15915494309189533577),
            static_cast<R>(k * nn), static_cast<R>(gs[G_N]), static_cast<R>(gs[G_N + 1]), static_cast<R>(gs[G_N + 2]),
            static_cast<R>(gt[G_N]), static_cast<R>(gt[G_N + 1]), static_cast<R>(gt[G_N + 2])};
}

template <class R>
__device__ __forceinline__ void add_sample_t(MomentsT<R> &m, const SampleT<R> &s, const R *xt, const R *xs) {
    m.g = cadd_t(m.g, s.g);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        m.b[i] = cadd_t(m.b[i], cscale_t(s.b, xt[i]));
#pragma unroll
        for (int j = 0; j < 3; ++j)
            m.a[3 * i + j] = cadd_t(m.a[3 * i + j], cscale_t(s.a, xt[i] * xs[j]));
    }
}

template <class R> __device__ __forceinline__ R wave_sum_t(R v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

template <class R> __device__ __forceinline__ void wave_reduce_t(MomentsT<R> &m) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        m.a[i].re = wave_sum_t(m.a[i].re);
        m.a[i].im = wave_sum_t(m.a[i].im);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        m.b[i].re = wave_sum_t(m.b[i].re);
        m.b[i].im = wave_sum_t(m.b[i].im);
    }
    m.g.re = wave_sum_t(m.g.re);
    m.g.im = wave_sum_t(m.g.im);
}

__device__ __forceinline__ bool contains_sorted(const std::uint32_t *v, std::uint32_t n, std::uint32_t key) {
    std::uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const std::uint32_t mid = lo + (hi - lo) / 2;
        if (v[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && v[lo] == key;
}

__device__ __forceinline__ double g_at(const double *geo, int field, std::uint32_t nt, std::uint32_t t) {
    return geo[static_cast<std::size_t>(field) * nt + t];
}

// Sample core of a disjoint pair (regular pairs and tensor near patches), from the
// squared distance r2, the numerators fyn = d.n_y and dntk = (d.n_x)/k already
// formed: the same algebra as bm_sample for g and a; of b = -g + i radial fxk only
// radial fxk is returned, the caller sums it and forms sum b = -sum g + i sum(radial
// fxk) once per test point. A zero, negative or non-finite r2 yields non-finite
// moments, which the caller checks once per pair.
template <class R> struct CoreT {
    CplxT<R> g, a, rb;
};

__device__ __forceinline__ CoreT<float> pair_core(float r2, float wq, float fyn, float dntk, float k, float k_rev,
                                                  float knn) {
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    const float f = __builtin_amdgcn_fractf(r * k_rev);
    const float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);
    const float kr = k * r, scale = wq * rinv;
    CoreT<float> s;
    s.g = {cs * scale, -sn * scale};
    const CplxT<float> radial = {s.g.re - s.g.im * kr, s.g.im + s.g.re * kr};
    const float ir2 = rinv * rinv, fy = fyn * ir2, fxk = -dntk * ir2;
    s.a = {-radial.re * fy - knn * s.g.im, -radial.im * fy + knn * s.g.re};
    s.rb = {radial.re * fxk, radial.im * fxk};
    return s;
}

__device__ __forceinline__ CoreT<double> pair_core(double r2, double wq, double fyn, double dntk, double k, double,
                                                   double knn) {
    const double rinv = rsqrt_full(r2);
    const double r = r2 * rinv, kr = k * r;
    double sn, cs;
    sincos_reduced(kr, &sn, &cs);
    const double scale = wq * rinv;
    CoreT<double> s;
    s.g = {cs * scale, -sn * scale};
    const CplxT<double> radial = {s.g.re - s.g.im * kr, s.g.im + s.g.re * kr};
    const double ir2 = rinv * rinv, fy = fyn * ir2, fxk = -dntk * ir2;
    s.a = {-radial.re * fy - knn * s.g.im, -radial.im * fy + knn * s.g.re};
    s.rb = {radial.re * fxk, radial.im * fxk};
    return s;
}

// The regular rules of degree 4, 6 and 8 (6, 12 and 16 points: barycentrics and
// weight per point), passed by value as kernel arguments: with the trial loop
// unrolled over a compile-time count its points are scalar operands, read once per
// wavefront instead of once per lane and sample.
template <class R> struct TileRules {
    R d4[6][4], d6[12][4], d8[16][4];
};

// Flat panel frame: vertex 0, edges e1 = v1 - v0, e2 = v2 - v0, unit normal (FP64).
struct Frame {
    double v0[3], e1[3], e2[3], n[3];
};

__device__ __forceinline__ Frame frame_of(const double *g) {
    Frame f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        f.v0[i] = g[G_V0 + i];
        f.e1[i] = g[G_E1 + i];
        f.e2[i] = g[G_E2 + i];
        f.n[i] = g[G_N + i];
    }
    return f;
}

__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Two-stage tensor contraction of a disjoint pair over rule x rule (test point
// x, trial point y, barycentrics and weight per rule row): the inner sum over
// trial points carries the trial basis, the outer sum the test basis. With
// d = x - y = b + l_x - l_y (b = v0_x - v0_y, l in each panel's plane):
//   |d|^2 = |b + l_x|^2 - 2 (b + l_x).l_y + |l_y|^2, expanded in y1, y2;
//   d.n_y = (b + l_x).n_y (l_y is orthogonal to n_y): constant in the inner loop;
//   d.n_x = b.n_x - l_y.n_x (l_x is orthogonal to n_x).
// Pair constants are formed in FP64 and rounded to R; wscale = 4 A_x A_y / (4 pi)
// (times the area fraction of a near patch). N = 0 takes n points from `rule` for
// both panels; N > 0 takes the test points from `rule` and the trial points from
// `yrule` (N of them, unrolled: a kernel-argument array). bad is raised when the
// sum of G is not finite (a zero or invalid distance somewhere in the pair).
template <class R, int N = 0>
__device__ __forceinline__ void contract(MomentsT<R> &m, const Frame &t, const Frame &s, double wscale_d,
                                         const R *rule, int n_runtime, double k, bool &bad,
                                         const R (*yrule)[4] = nullptr) {
    const int n = N > 0 ? N : n_runtime;
    const double b[3] = {t.v0[0] - s.v0[0], t.v0[1] - s.v0[1], t.v0[2] - s.v0[2]};
    const double inv_k = 1.0 / k;
    const R Q11 = static_cast<R>(dot3(s.e1, s.e1)), Q12 = static_cast<R>(2.0 * dot3(s.e1, s.e2)),
            Q22 = static_cast<R>(dot3(s.e2, s.e2));
    const R A1 = static_cast<R>(dot3(s.e1, t.n) * inv_k), A2 = static_cast<R>(dot3(s.e2, t.n) * inv_k),
            C0 = static_cast<R>(dot3(b, t.n) * inv_k);
    const R T1 = static_cast<R>(dot3(t.e1, s.n)), T2 = static_cast<R>(dot3(t.e2, s.n)),
            Bn = static_cast<R>(dot3(b, s.n));
    const R kk = static_cast<R>(k), k_rev = static_cast<R>(k * 0.15915494309189533577),
            knn = static_cast<R>(k * dot3(t.n, s.n)), wscale = static_cast<R>(wscale_d);
    R bb[3], te1[3], te2[3], se1[3], se2[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        bb[i] = static_cast<R>(b[i]);
        te1[i] = static_cast<R>(t.e1[i]);
        te2[i] = static_cast<R>(t.e2[i]);
        se1[i] = static_cast<R>(s.e1[i]);
        se2[i] = static_cast<R>(s.e2[i]);
    }
#pragma unroll 1
    for (int xi = 0; xi < n; ++xi) {
        const R *xp = rule + 4 * xi;
        const R x1 = xp[1], x2 = xp[2], xw = xp[3] * wscale;
        R blt[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            blt[i] = bb[i] + (te1[i] * x1 + te2[i] * x2);
        const R P1 = R(-2) * (blt[0] * se1[0] + blt[1] * se1[1] + blt[2] * se1[2]);
        const R P2 = R(-2) * (blt[0] * se2[0] + blt[1] * se2[1] + blt[2] * se2[2]);
        const R B = blt[0] * blt[0] + blt[1] * blt[1] + blt[2] * blt[2];
        const R fyn = Bn + x1 * T1 + x2 * T2;
        CplxT<R> ga[3] = {{0, 0}, {0, 0}, {0, 0}}, gg = {0, 0}, rb = {0, 0};
        auto sample_at = [&](R y0, R y1, R y2, R yw) {
            const R t1 = y1 * Q11 + (y2 * Q12 + P1), t2 = y2 * Q22 + P2;
            const R r2 = y1 * t1 + (y2 * t2 + B);
            const R dntk = C0 - (y1 * A1 + y2 * A2);
            const CoreT<R> smp = pair_core(r2, xw * yw, fyn, dntk, kk, k_rev, knn);
            ga[0] = cadd_t(ga[0], cscale_t(smp.a, y0));
            ga[1] = cadd_t(ga[1], cscale_t(smp.a, y1));
            ga[2] = cadd_t(ga[2], cscale_t(smp.a, y2));
            gg = cadd_t(gg, smp.g);
            rb = cadd_t(rb, smp.rb);
        };
        if constexpr (N > 0) {
            // the 6-point rule fully unrolled (its x-invariant per-point terms stay in
            // registers); longer rules four points at a time, which keeps them in the loop
            constexpr int kUnroll = N <= 6 ? N : 4;
#pragma unroll kUnroll
            for (int yi = 0; yi < N; ++yi)
                sample_at(yrule[yi][0], yrule[yi][1], yrule[yi][2], yrule[yi][3]);
        } else {
#pragma unroll 2
            for (int yi = 0; yi < n; ++yi) {
                const R *yp = rule + 4 * yi;
                sample_at(yp[0], yp[1], yp[2], yp[3]);
            }
        }
        const CplxT<R> gb = {-gg.re - rb.im, -gg.im + rb.re};
        m.g = cadd_t(m.g, gg);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const R xb = xp[i];
            m.b[i] = cadd_t(m.b[i], cscale_t(gb, xb));
#pragma unroll
            for (int j = 0; j < 3; ++j)
                m.a[3 * i + j] = cadd_t(m.a[3 * i + j], cscale_t(ga[j], xb));
        }
    }
    bad |= !isfinite(m.g.re) || !isfinite(m.g.im);
}

// Sub-triangle of a panel with vertices at parent barycentrics T[l] (3 x 3,
// row l = vertex l): frame of the physical sub-triangle.
__device__ __forceinline__ Frame sub_frame(const Frame &p, const double *T) {
    Frame f;
    f.n[0] = p.n[0];
    f.n[1] = p.n[1];
    f.n[2] = p.n[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        f.v0[i] = p.v0[i] + p.e1[i] * T[1] + p.e2[i] * T[2];
        f.e1[i] = p.e1[i] * (T[4] - T[1]) + p.e2[i] * (T[5] - T[2]);
        f.e2[i] = p.e1[i] * (T[7] - T[1]) + p.e2[i] * (T[8] - T[2]);
    }
    return f;
}

// Moments over local sub-triangle barycentrics -> parent barycentrics:
// a[i][j] = sum_l,q Tt[l][i] Ts[q][j] a_local[l][q], b[i] = sum_l Tt[l][i] b_local[l]
// (map_patch_point: parent bary = sum_l local_l T[l]).
template <class R> __device__ __forceinline__ void to_parent(MomentsT<R> &m, const double *Tt, const double *Ts) {
    MomentsT<R> o;
    o.g = m.g;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        o.b[i] = {0, 0};
#pragma unroll
        for (int l = 0; l < 3; ++l)
            o.b[i] = cadd_t(o.b[i], cscale_t(m.b[l], static_cast<R>(Tt[3 * l + i])));
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            CplxT<R> v = {0, 0};
#pragma unroll
            for (int l = 0; l < 3; ++l)
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    v = cadd_t(v, cscale_t(m.a[3 * l + q], static_cast<R>(Tt[3 * l + i] * Ts[3 * q + j])));
            o.a[3 * i + j] = v;
        }
    }
    m = o;
}

template <class M> __device__ __forceinline__ void atomic_add_entry(M *p, double re, double im) {
    unsafeAtomicAdd(p, static_cast<M>(re));
    unsafeAtomicAdd(p + 1, static_cast<M>(im));
}

template <class R> __device__ __forceinline__ R *matrix_of(const AssemblyArgs &a);
template <> __device__ __forceinline__ double *matrix_of<double>(const AssemblyArgs &a) { return a.A; }
template <> __device__ __forceinline__ float *matrix_of<float>(const AssemblyArgs &a) { return a.A32; }

// ---- symmetric regular tiles ----
//
// One workgroup per unordered cluster pair (A, B), A <= B. Each regular panel pair
// (t in A, s in B) is integrated once on its rule x rule samples and both ordered
// pairs (t, s) and (s, t) are formed from them: distance, phase, g and the radial
// factor are shared, only the normal projections differ. With d = x - y (x on t, y
// on s), P = radial (d.n_s)/r^2 and Q = -radial (d.n_t)/(k r^2):
//   (t, s):  a = -P + i knn g,     b = -g + i Q
//   (s, t):  a = -k Q + i knn g,   b = -g + i P / k
// (for (s, t) the displacement is -d, its trial normal n_t and its test normal n_s).
// The sums TG, TP, TQ of g, P, Q weighted by the test basis x_i and the trial basis
// y_j (3 x 3 complex each) give all 2 x 9 A moments and 2 x 3 B moments. The (t, s)
// entries go to node tile AB ([B node][A node]), the (s, t) entries to node tile BA
// ([A node][B node]); on a diagonal tile (A = B) only pairs t < s are taken and both
// orientations go to tile AB. Geometry: vertex 0 in FP64 (the origin difference is
// formed in FP64 and rounded to R), edges, normals, curls and area in R.
constexpr int kSymFields = G_COUNT - G_E1;     // fields held in R: e1, e2, n, curls, area
// Node tiles: odd strides, so entries of one wavefront's update (8 test x 8 trial
// panels, see tile_pair) fall in distinct banks whichever index varies.
constexpr int kStrideAB = 2 * kTileNodes + 1;
constexpr int kStrideBA = 2 * kTileNodes + 1;
enum : int { F_E1 = 0, F_E2 = 3, F_N = 6, F_C0 = 9, F_AREA = 18 };

template <class R> struct SymSums {
    CplxT<R> g[3][3], p[3][3], q[3][3]; // [test basis i][trial basis j]
};

template <class R> struct SymPair {
    R b[3], te1[3], te2[3], se1[3], se2[3];
    R Q11, Q12, Q22, A1, A2, C0, T1, T2, Bn, wscale, knn;
};

// Shared part of a sample: g = wq exp(-ikr)/r and radial = g (1 + i k r), and 1/r^2.
// FP32: hardware rsqrt, the phase reduced in revolutions, hardware sin/cos; FP64:
// full-accuracy rsqrt and the Cody-Waite sin/cos of k r.
template <class R> struct SymCore {
    CplxT<R> g, radial;
    R ir2;
};

__device__ __forceinline__ SymCore<float> sym_core(float r2, float wq, float k, float k_rev) {
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    const float f = __builtin_amdgcn_fractf(r * k_rev);
    const float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);
    const float kr = k * r, scale = wq * rinv;
    SymCore<float> o;
    o.g = {cs * scale, -sn * scale};
    o.radial = {o.g.re - o.g.im * kr, o.g.im + o.g.re * kr};
    o.ir2 = rinv * rinv;
    return o;
}

__device__ __forceinline__ SymCore<double> sym_core(double r2, double wq, double k, double) {
    const double rinv = rsqrt_full(r2);
    const double r = r2 * rinv, kr = k * r;
    double sn, cs;
    sincos_reduced(kr, &sn, &cs);
    const double scale = wq * rinv;
    SymCore<double> o;
    o.g = {cs * scale, -sn * scale};
    o.radial = {o.g.re - o.g.im * kr, o.g.im + o.g.re * kr};
    o.ir2 = rinv * rinv;
    return o;
}

// Sums of one pair over rule x rule: N > 0 takes both rules' points from the
// kernel-argument array frule (N points, trial loop unrolled), N = 0 takes n points
// from the FP64 table drule.
template <class R, int N>
__device__ __forceinline__ void contract_sym(SymSums<R> &T, const SymPair<R> &c, R k, R k_rev,
                                             const R (*frule)[4], const double *drule, int n_runtime) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            T.g[i][j] = T.p[i][j] = T.q[i][j] = {R(0), R(0)};
    const int n = N > 0 ? N : n_runtime;
#pragma unroll 1
    for (int xi = 0; xi < n; ++xi) {
        R x0, x1, x2, xw;
        if constexpr (N > 0) {
            x0 = frule[xi][0];
            x1 = frule[xi][1];
            x2 = frule[xi][2];
            xw = frule[xi][3];
        } else {
            x0 = static_cast<R>(drule[4 * xi]);
            x1 = static_cast<R>(drule[4 * xi + 1]);
            x2 = static_cast<R>(drule[4 * xi + 2]);
            xw = static_cast<R>(drule[4 * xi + 3]);
        }
        xw *= c.wscale;
        R blt[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            blt[i] = c.b[i] + (c.te1[i] * x1 + c.te2[i] * x2);
        const R P1 = R(-2) * (blt[0] * c.se1[0] + blt[1] * c.se1[1] + blt[2] * c.se1[2]);
        const R P2 = R(-2) * (blt[0] * c.se2[0] + blt[1] * c.se2[1] + blt[2] * c.se2[2]);
        const R B = blt[0] * blt[0] + blt[1] * blt[1] + blt[2] * blt[2];
        const R fyn = c.Bn + x1 * c.T1 + x2 * c.T2;
        CplxT<R> sg[3] = {{0, 0}, {0, 0}, {0, 0}}, sp[3] = {{0, 0}, {0, 0}, {0, 0}}, sq[3] = {{0, 0}, {0, 0}, {0, 0}};
        auto sample_at = [&](R y0, R y1, R y2, R yw) {
            const R t1 = y1 * c.Q11 + (y2 * c.Q12 + P1), t2 = y2 * c.Q22 + P2;
            const R r2 = y1 * t1 + (y2 * t2 + B);
            const R dntk = c.C0 - (y1 * c.A1 + y2 * c.A2);
            const SymCore<R> o = sym_core(r2, xw * yw, k, k_rev);
            const CplxT<R> g = o.g, radial = o.radial;
            const R tau = fyn * o.ir2, fxk = -dntk * o.ir2;
            const CplxT<R> P = {radial.re * tau, radial.im * tau}, Q = {radial.re * fxk, radial.im * fxk};
            const R yb[3] = {y0, y1, y2};
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                sg[j] = cadd_t(sg[j], cscale_t(g, yb[j]));
                sp[j] = cadd_t(sp[j], cscale_t(P, yb[j]));
                sq[j] = cadd_t(sq[j], cscale_t(Q, yb[j]));
            }
        };
        if constexpr (N > 0) {
            constexpr int kUnroll = N <= 6 ? N : 4;
#pragma unroll kUnroll
            for (int yi = 0; yi < N; ++yi)
                sample_at(frule[yi][0], frule[yi][1], frule[yi][2], frule[yi][3]);
        } else {
#pragma unroll 2
            for (int yi = 0; yi < n; ++yi)
                sample_at(static_cast<R>(drule[4 * yi]), static_cast<R>(drule[4 * yi + 1]),
                          static_cast<R>(drule[4 * yi + 2]), static_cast<R>(drule[4 * yi + 3]));
        }
        const R xb[3] = {x0, x1, x2};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                T.g[i][j] = cadd_t(T.g[i][j], cscale_t(sg[j], xb[i]));
                T.p[i][j] = cadd_t(T.p[i][j], cscale_t(sp[j], xb[i]));
                T.q[i][j] = cadd_t(T.q[i][j], cscale_t(sq[j], xb[i]));
            }
    }
}

// Rule of a regular pair of a general tile:
// -1 for a touching pair, a pair of the near list or an escaped pair (which raises
// ERR_ESCAPED), else the rule table offset; count receives the point count.
__device__ __forceinline__ int general_pair_rule(const AssemblyArgs &args, const NearView &near, std::uint32_t ti,
                                                 std::uint32_t si, double k, int &count) {
    const std::uint32_t nb = near.offsets[ti];
    if (contains_sorted(near.trials + nb, near.offsets[ti + 1] - nb, si))
        return -1;
    if (near.whole) {
        const unsigned d = near.whole[unordered_pair_index(ti, si)];
        if (d) {
            count = d == 14 ? 64 : 16;
            return d == 14 ? kRule14 : kRule8;
        }
    }
    const PanelView &pv = args.panels;
    auto bounds_of = [&](std::uint32_t id) {
        PanelBounds b;
        b.center = {pv.bounds[static_cast<std::size_t>(B_C) * pv.nt + id],
                    pv.bounds[static_cast<std::size_t>(B_C + 1) * pv.nt + id],
                    pv.bounds[static_cast<std::size_t>(B_C + 2) * pv.nt + id]};
        b.radius = pv.bounds[static_cast<std::size_t>(B_R) * pv.nt + id];
        b.max_edge = pv.bounds[static_cast<std::size_t>(B_H) * pv.nt + id];
        b.qualified_shape = pv.shape[id];
        return b;
    };
    const unsigned d = qualified_regular_degree(bounds_of(ti), bounds_of(si), k, args.tolerance, args.regular_order);
    if (!d) {
        atomicExch(args.error, ERR_ESCAPED);
        return -1;
    }
    count = d == 4 ? 6 : d == 6 ? 12 : 16;
    return d == 4 ? kRule4 : d == 6 ? kRule6 : kRule8;
}

// Pair p of an nT x nS tile: blocks of W = 2^shift consecutive trial panels (W = 4, 8
// or 16, chosen per mesh by the tile autotuner), test-major inside a block (the last
// block holds the remaining nS mod W), so a wavefront spans 64 / W test x W trial
// panels and its node-tile updates hit fewer equal entries than with one trial panel
// per wavefront. A bijection of [0, nT nS).
__device__ __forceinline__ void tile_pair(std::uint32_t p, std::uint32_t nT, std::uint32_t nS, std::uint32_t shift,
                                          std::uint32_t &lt, std::uint32_t &ls) {
    const std::uint32_t W = 1u << shift, full = nS >> shift, span = nT << shift, blk = p / span;
    if (blk < full) {
        const std::uint32_t rem = p - blk * span;
        lt = rem >> shift;
        ls = (blk << shift) + (rem & (W - 1));
    } else {
        const std::uint32_t r = nS - (full << shift), rem = p - full * span;
        lt = rem / r;
        ls = (full << shift) + rem % r;
    }
}

// Far tiles accumulate their node tiles in fixed point with integer LDS atomics
// (ds_add_u32 / ds_add_u64; the floating-point LDS atomics are several times slower
// on gfx950 and bound the far tiles otherwise). The scale is a power of two chosen per
// tile from a bound on every contribution (see far_scale), so no sum can overflow;
// the resolution is about 2^-30 (FP32 form) or 2^-51 (FP64 form) of that bound.
template <class R> struct FixedOf;
template <> struct FixedOf<float> {
    using type = int;
    static constexpr double headroom = 1073741824.0; // 2^30
};
template <> struct FixedOf<double> {
    using type = long long;
    static constexpr double headroom = 2251799813685248.0; // 2^51: rint by the 1.5 * 2^52 shift
};

// rint(x) as a 64-bit integer for |x| < 2^51: adding 1.5 * 2^52 leaves the rounded
// integer in the low mantissa bits (round to nearest even), read back as an integer
// difference of the two bit patterns.
__device__ __forceinline__ long long rint_to_int64(double x) {
    constexpr double shift = 6755399441055744.0; // 1.5 * 2^52
    return __double_as_longlong(x + shift) - __double_as_longlong(shift);
}

// Scale of a far tile (clusters fa, fb: kClusterFar doubles each). Every sample
// point pair of the tile is at least r = |c_a - c_b| - reach_a - reach_b apart, and
// each entry contribution of a pair is bounded by
//   Wq [ (1 + k r)/r^2  (P and k Q terms)  +  k/r  (knn G)  +  |curl|^2/(k r)  (G cc / k) ]
// with Wq the sum of the pair's sample weights (4 A_t A_s/(4 pi) times the squared
// rule weight sum); at most val_a * val_b pairs add to one entry. Returns false
// when the bound is not usable (the caller raises ERR_RANGE).
__device__ __forceinline__ bool far_scale(const double *fa, const double *fb, double k, double weight_sum,
                                          double headroom, double &scale, double &limit) {
    const double dx = fa[0] - fb[0], dy = fa[1] - fb[1], dz = fa[2] - fb[2];
    const double r = (sqrt(dx * dx + dy * dy + dz * dz) - fa[3] - fb[3]) * (1 - 1e-9);
    const double wq = 4.0 * fa[4] * fb[4] * 0.07957747154594766788 * weight_sum * weight_sum;
    const double per = 1.0625 * wq * ((1 + k * r) / (r * r) + k / r + fa[5] * fb[5] / (k * r));
    const double total = per * fa[6] * fb[6];
    if (!(r > 0) || !(total > 0) || !isfinite(total))
        return false;
    int e = 0;
    frexp(headroom / total, &e);
    e = min(max(e - 1, -100), 100);
    scale = ldexp(1.0, e);
    limit = per * scale;
    return true;
}

// FAR: the tile list holds far tiles only (A < B, every pair disjoint, not near,
// degree 4 at this wavenumber). Occupancy: FP32, two 512-thread workgroups per CU
// (LDS <= 80 KiB, VGPRs <= 128); FP64, one (LDS <= 160 KiB, VGPRs <= 256).
template <class R, bool FAR>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 2)))
void regular_tiles_sym(AssemblyArgs args, ClusterView cl, NearView near, const double *rules,
                       const std::uint32_t *tiles, TileRules<R> fixed, std::uint32_t trial_shift) {
    __shared__ R tab[kTileNodes * kStrideAB];
    __shared__ R tba[kTileNodes * kStrideBA];
    __shared__ double sv0[2][3][kTilePanels];
    __shared__ R sgf[2][kSymFields][kTilePanels];
    __shared__ std::uint32_t snode[2][3][kTilePanels];
    __shared__ std::uint8_t slocal[2][3][kTilePanels];
    __shared__ std::int32_t sfc[2][kTilePanels];
    __shared__ R srs[2][kTileNodes];

    const std::uint32_t ca = tiles[2 * blockIdx.x], cb = tiles[2 * blockIdx.x + 1];
    const bool diag = !FAR && ca == cb;
    const std::uint32_t t0 = cl.panel_offsets[ca], nT = cl.panel_offsets[ca + 1] - t0;
    const std::uint32_t s0 = cl.panel_offsets[cb], nS = cl.panel_offsets[cb + 1] - s0;
    const std::uint32_t a0 = cl.node_offsets[ca], nA = cl.node_offsets[ca + 1] - a0;
    const std::uint32_t b0 = cl.node_offsets[cb], nB = cl.node_offsets[cb + 1] - b0;
    const PanelView &pv = args.panels;
    const int tid = threadIdx.x;

    for (int e = tid; e < kTileNodes * kStrideAB; e += kTileThreads)
        tab[e] = R(0);
    if (!diag)
        for (int e = tid; e < kTileNodes * kStrideBA; e += kTileThreads)
            tba[e] = R(0);
    for (int e = tid; e < 2 * G_COUNT * kTilePanels; e += kTileThreads) {
        const int c = e / (G_COUNT * kTilePanels), f = (e / kTilePanels) % G_COUNT, p = e % kTilePanels;
        if (p < static_cast<int>(c ? nS : nT)) {
            const double v = g_at(pv.geo, f, pv.nt, cl.panels[(c ? s0 : t0) + p]);
            if (f < G_E1)
                sv0[c][f][p] = v;
            else
                sgf[c][f - G_E1][p] = static_cast<R>(v);
        }
    }
    for (int e = tid; e < 2 * kTilePanels; e += kTileThreads) {
        const int c = e / kTilePanels, p = e % kTilePanels;
        if (p < static_cast<int>(c ? nS : nT)) {
            const std::uint32_t id = cl.panels[(c ? s0 : t0) + p];
            sfc[c][p] = args.flux_column[id];
            for (int l = 0; l < 3; ++l) {
                snode[c][l][p] = pv.nodes[static_cast<std::size_t>(l) * pv.nt + id];
                slocal[c][l][p] = cl.local[static_cast<std::size_t>(l) * pv.nt + id];
            }
        }
        if (p < static_cast<int>(c ? nB : nA))
            srs[c][p] = static_cast<R>(args.row_scale[cl.nodes[(c ? b0 : a0) + p]]);
    }
    __syncthreads();

    const double k = args.k;
    const R kf = static_cast<R>(k), inv_k = static_cast<R>(1.0 / k), k_rev = static_cast<R>(k * 0.15915494309189533577);
    const std::uint32_t pairs = nT * nS;
    using Fixed = typename FixedOf<R>::type;
    R fixed_scale = 1, fixed_limit = 0;
    bool out_of_range = false;
    if constexpr (FAR) {
        double weight_sum = 0, scale = 1, limit = 0;
#pragma unroll
        for (int q = 0; q < 6; ++q)
            weight_sum += static_cast<double>(fixed.d4[q][3]);
        if (far_scale(cl.far + kClusterFar * ca, cl.far + kClusterFar * cb, k, weight_sum, FixedOf<R>::headroom,
                      scale, limit)) {
            fixed_scale = static_cast<R>(scale);
            fixed_limit = static_cast<R>(limit);
        } else {
            out_of_range = true;
        }
    }
    // One node-tile update: fixed point on far tiles (|v| within the tile bound, else
    // ERR_RANGE; a non-finite v is reported through `bad`), a floating-point LDS
    // atomic otherwise.
    auto tile_add = [&](R *slot, R v) {
        if constexpr (FAR) {
            const R x = v * fixed_scale;
            out_of_range |= fabs(x) > fixed_limit;
            if constexpr (sizeof(R) == 4)
                atomicAdd(reinterpret_cast<int *>(slot), static_cast<int>(rintf(x)));
            else
                atomicAdd(reinterpret_cast<unsigned long long *>(slot),
                          static_cast<unsigned long long>(rint_to_int64(x)));
        } else {
            unsafeAtomicAdd(slot, v);
        }
    };
    unsigned long long counted = 0;
    bool bad = false;
    // Pair constants from the staged geometry (LDS reads only).
    auto make_pair = [&](std::uint32_t lt, std::uint32_t ls) {
        SymPair<R> c;
        const double bd[3] = {sv0[0][0][lt] - sv0[1][0][ls], sv0[0][1][lt] - sv0[1][1][ls],
                              sv0[0][2][lt] - sv0[1][2][ls]};
        R tn[3], sn[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            c.b[i] = static_cast<R>(bd[i]);
            c.te1[i] = sgf[0][F_E1 + i][lt];
            c.te2[i] = sgf[0][F_E2 + i][lt];
            c.se1[i] = sgf[1][F_E1 + i][ls];
            c.se2[i] = sgf[1][F_E2 + i][ls];
            tn[i] = sgf[0][F_N + i][lt];
            sn[i] = sgf[1][F_N + i][ls];
        }
        auto dot = [](const R *u, const R *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
        c.Q11 = dot(c.se1, c.se1);
        c.Q12 = R(2) * dot(c.se1, c.se2);
        c.Q22 = dot(c.se2, c.se2);
        c.A1 = dot(c.se1, tn) * inv_k;
        c.A2 = dot(c.se2, tn) * inv_k;
        c.C0 = static_cast<R>((bd[0] * tn[0] + bd[1] * tn[1] + bd[2] * tn[2]) * (1.0 / k));
        c.T1 = dot(c.te1, sn);
        c.T2 = dot(c.te2, sn);
        c.Bn = static_cast<R>(bd[0] * sn[0] + bd[1] * sn[1] + bd[2] * sn[2]);
        c.wscale = R(4) * sgf[0][F_AREA][lt] * sgf[1][F_AREA][ls] * R(0.07957747154594766788);
        c.knn = kf * dot(tn, sn);
        return c;
    };
    // Both orientations' entries into the node tiles and, for driven panels, B_D.
    auto scatter = [&](const SymSums<R> &T, R knn, std::uint32_t lt, std::uint32_t ls) {
        CplxT<R> G = {R(0), R(0)};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                G = cadd_t(G, T.g[i][j]);
        bad |= !isfinite(G.re) || !isfinite(G.im);
        const CplxT<R> Gk = {G.re * inv_k, G.im * inv_k};
        // Every LDS read of the scatter first, then one wait, then the node-tile updates
        // back to back: LDS operations complete in order, so a read issued between the
        // updates would wait for every update queued before it.
        int a_of[3], b_of[3];
        R ct[3][3], cs[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            a_of[i] = slocal[0][i][lt];
            b_of[i] = slocal[1][i][ls];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                ct[i][c] = sgf[0][F_C0 + 3 * i + c][lt];
                cs[i][c] = sgf[1][F_C0 + 3 * i + c][ls];
            }
        }
        const std::int32_t fs = sfc[1][ls], ft = sfc[0][lt];
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0); vmcnt and expcnt left open
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int a = a_of[i];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int b = b_of[j];
                const R cc = ct[i][0] * cs[j][0] + ct[i][1] * cs[j][1] + ct[i][2] * cs[j][2];
                const CplxT<R> tg = T.g[i][j], tp = T.p[i][j], tq = T.q[i][j];
                // (t, s): row = t's node i (A), column = s's node j (B)
                R *ts = tab + b * kStrideAB + 2 * a;
                tile_add(ts, -tp.re - knn * tg.im + Gk.im * cc);
                tile_add(ts + 1, -tp.im + knn * tg.re - Gk.re * cc);
                // (s, t): row = s's node j (B), column = t's node i (A)
                R *st = diag ? tab + a * kStrideAB + 2 * b : tba + a * kStrideBA + 2 * b;
                tile_add(st, -kf * tq.re - knn * tg.im + Gk.im * cc);
                tile_add(st + 1, -kf * tq.im + knn * tg.re - Gk.re * cc);
            }
        }
        if (fs >= 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                R re = 0, im = 0;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    re -= T.g[i][j].re + T.q[i][j].im;
                    im += T.q[i][j].re - T.g[i][j].im;
                }
                const std::uint32_t row = snode[0][i][lt];
                const double rs = args.row_scale[row];
                atomic_add_entry(args.BD + 2 * (static_cast<std::size_t>(fs) * pv.nv + row), re * rs, im * rs);
            }
        }
        if (ft >= 0) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                R re = 0, im = 0;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    re -= T.g[i][j].re + T.p[i][j].im * inv_k;
                    im += T.p[i][j].re * inv_k - T.g[i][j].im;
                }
                const std::uint32_t row = snode[1][j][ls];
                const double rs = args.row_scale[row];
                atomic_add_entry(args.BD + 2 * (static_cast<std::size_t>(ft) * pv.nv + row), re * rs, im * rs);
            }
        }
    };
    if constexpr (FAR) {
        for (std::uint32_t p = tid; p < pairs; p += kTileThreads) {
            std::uint32_t lt, ls;
            tile_pair(p, nT, nS, trial_shift, lt, ls);
            const SymPair<R> c = make_pair(lt, ls);
            SymSums<R> T;
            contract_sym<R, 6>(T, c, kf, k_rev, fixed.d4, nullptr, 6);
            counted += 2ull * 36ull;
            scatter(T, c.knn, lt, ls);
        }
    } else {
        for (std::uint32_t p = tid; p < pairs; p += kTileThreads) {
            std::uint32_t lt, ls;
            tile_pair(p, nT, nS, trial_shift, lt, ls);
            if (diag && lt >= ls)
                continue;
            bool touching = false;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    touching |= snode[0][i][lt] == snode[1][j][ls];
            if (touching)
                continue;
            int count = 0;
            const int offset = general_pair_rule(args, near, cl.panels[t0 + lt], cl.panels[s0 + ls], k, count);
            if (offset < 0)
                continue;
            const SymPair<R> c = make_pair(lt, ls);
            SymSums<R> T;
            // the unrolled fixed-rule forms when every active lane of the wavefront
            // takes the same rule (a uniform branch), else the counted loop
            const int first = __builtin_amdgcn_readfirstlane(offset);
            const bool uniform = __all(offset == first);
            if (uniform && first == kRule6)
                contract_sym<R, 12>(T, c, kf, k_rev, fixed.d6, nullptr, 12);
            else if (uniform && first == kRule4)
                contract_sym<R, 6>(T, c, kf, k_rev, fixed.d4, nullptr, 6);
            else if (uniform && first == kRule8)
                contract_sym<R, 16>(T, c, kf, k_rev, fixed.d8, nullptr, 16);
            else
                contract_sym<R, 0>(T, c, kf, k_rev, nullptr, rules + 4 * offset, count);
            counted += 2ull * static_cast<unsigned long long>(count) * count;
            scatter(T, c.knn, lt, ls);
        }
    }
    if (bad)
        atomicExch(args.error, ERR_DISTANCE);
    if (args.samples) {
        unsigned long long v = counted;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            v += __shfl_xor(v, off, 64);
        if ((tid & 63) == 0 && v)
            atomicAdd(args.samples, v);
    }
    __syncthreads();
    if (out_of_range)
        atomicExch(args.error, ERR_RANGE);
    R *A = matrix_of<R>(args);
    // a far tile's entries back from fixed point (exact power-of-two scaling)
    const R unscale = R(1) / fixed_scale;
    auto entry = [&](const R *slot) {
        if constexpr (FAR)
            return static_cast<R>(*reinterpret_cast<const Fixed *>(slot)) * unscale;
        else
            return *slot;
    };
    for (std::uint32_t e = tid; e < nA * nB; e += kTileThreads) {
        const std::uint32_t a = e % nA, b = e / nA;
        const R re = entry(tab + b * kStrideAB + 2 * a), im = entry(tab + b * kStrideAB + 2 * a + 1);
        if (re != R(0) || im != R(0)) {
            const std::uint32_t row = cl.nodes[a0 + a], col = cl.nodes[b0 + b];
            const R rs = srs[0][a];
            unsafeAtomicAdd(A + 2 * (static_cast<std::size_t>(col) * pv.nv + row), re * rs);
            unsafeAtomicAdd(A + 2 * (static_cast<std::size_t>(col) * pv.nv + row) + 1, im * rs);
        }
    }
    if (!diag)
        for (std::uint32_t e = tid; e < nA * nB; e += kTileThreads) {
            const std::uint32_t b = e % nB, a = e / nB;
            const R re = entry(tba + a * kStrideBA + 2 * b), im = entry(tba + a * kStrideBA + 2 * b + 1);
            if (re != R(0) || im != R(0)) {
                const std::uint32_t row = cl.nodes[b0 + b], col = cl.nodes[a0 + a];
                const R rs = srs[1][b];
                unsafeAtomicAdd(A + 2 * (static_cast<std::size_t>(col) * pv.nv + row), re * rs);
                unsafeAtomicAdd(A + 2 * (static_cast<std::size_t>(col) * pv.nv + row) + 1, im * rs);
            }
        }
}

template <class M, class R>
__device__ __forceinline__ void scatter_pair(const AssemblyArgs &args, M *A, const MomentsT<R> &m, std::uint32_t ti,
                                             std::uint32_t si) {
    const PanelView &pv = args.panels;
    const double inv_k = 1.0 / args.k;
    const std::int32_t fc = args.flux_column[si];
    for (int i = 0; i < 3; ++i) {
        const std::uint32_t row = pv.nodes[static_cast<std::size_t>(i) * pv.nt + ti];
        const double rs = args.row_scale[row];
        for (int j = 0; j < 3; ++j) {
            const std::uint32_t col = pv.nodes[static_cast<std::size_t>(j) * pv.nt + si];
            double cc = 0;
            for (int c = 0; c < 3; ++c)
                cc += g_at(pv.geo, G_C0 + 3 * i + c, pv.nt, ti) * g_at(pv.geo, G_C0 + 3 * j + c, pv.nt, si);
            const double re = static_cast<double>(m.a[3 * i + j].re) + static_cast<double>(m.g.im) * cc * inv_k;
            const double im = static_cast<double>(m.a[3 * i + j].im) - static_cast<double>(m.g.re) * cc * inv_k;
            atomic_add_entry(A + 2 * (static_cast<std::size_t>(col) * pv.nv + row), re * rs, im * rs);
        }
        if (fc >= 0)
            atomic_add_entry(args.BD + 2 * (static_cast<std::size_t>(fc) * pv.nv + row),
                             static_cast<double>(m.b[i].re) * rs, static_cast<double>(m.b[i].im) * rs);
    }
}

// Destination of a pair's moments in the dense path.
template <class M> struct DenseSink {
    template <class R>
    __device__ __forceinline__ void operator()(const AssemblyArgs &args, const MomentsT<R> &m, std::uint32_t,
                                               std::uint32_t ti, std::uint32_t si) const {
        scatter_pair(args, matrix_of<M>(args), m, ti, si);
    }
};

__device__ __forceinline__ std::uint32_t csr_slot(const std::uint32_t *offsets, const std::uint32_t *columns,
                                                  std::uint32_t row, std::uint32_t col) {
    std::uint32_t lo = offsets[row], hi = offsets[row + 1];
    while (lo < hi) {
        const std::uint32_t mid = lo + (hi - lo) / 2;
        if (columns[mid] < col)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo; // the planned entry exists (job masks select planned entries only)
}

// Near operator: only the planned (masked) local entries of a panel pair,
// accumulated into the pressure and flux CSR values.
struct CsrSink {
    NearCsrView v;
    template <class R>
    __device__ __forceinline__ void operator()(const AssemblyArgs &args, const MomentsT<R> &m, std::uint32_t job,
                                               std::uint32_t ti, std::uint32_t si) const {
        const PanelView &pv = args.panels;
        const double inv_k = 1.0 / args.k;
        const std::uint32_t mask = v.masks[job];
        const std::uint32_t fc = v.face_column[si];
        for (int i = 0; i < 3; ++i) {
            const std::uint32_t row = pv.nodes[static_cast<std::size_t>(i) * pv.nt + ti];
            const double rs = args.row_scale[row];
            for (int j = 0; j < 3; ++j) {
                if (!(mask >> (3 * i + j) & 1u))
                    continue;
                const std::uint32_t col = pv.nodes[static_cast<std::size_t>(j) * pv.nt + si];
                double cc = 0;
                for (int c = 0; c < 3; ++c)
                    cc += g_at(pv.geo, G_C0 + 3 * i + c, pv.nt, ti) * g_at(pv.geo, G_C0 + 3 * j + c, pv.nt, si);
                const double re = static_cast<double>(m.a[3 * i + j].re) + static_cast<double>(m.g.im) * cc * inv_k;
                const double im = static_cast<double>(m.a[3 * i + j].im) - static_cast<double>(m.g.re) * cc * inv_k;
                atomic_add_entry(v.values + 2 * csr_slot(v.row_offsets, v.columns, row, col), re * rs, im * rs);
            }
            if ((mask >> (9 + i) & 1u) && fc != UINT32_MAX)
                atomic_add_entry(v.flux_values + 2 * csr_slot(v.flux_offsets, v.flux_columns, row, fc),
                                 static_cast<double>(m.b[i].re) * rs, static_cast<double>(m.b[i].im) * rs);
        }
    }
};

__device__ __forceinline__ void load_geo(const PanelView &pv, std::uint32_t t, double *g) {
#pragma unroll
    for (int f = 0; f < G_COUNT; ++f)
        g[f] = g_at(pv.geo, f, pv.nt, t);
}

// Touching pairs are integrated in FP64 in both forms (the hypersingular weak
// form); M is the matrix element type.
template <class Sink>
__global__ __launch_bounds__(256) void touching_jobs(AssemblyArgs args, const TouchingJob *jobs, std::uint32_t njobs,
                                                     const double *gauss, unsigned order, Sink sink) {
    const std::uint32_t job = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (job >= njobs)
        return;
    const TouchingJob tj = jobs[job];
    const PanelView &pv = args.panels;
    double gt[G_COUNT], gs[G_COUNT];
    load_geo(pv, tj.test, gt);
    load_geo(pv, tj.trial, gs);
    const Topology topo = tj.permutation.topology;
    const bool coincident = topo == Topology::coincident;
    const PairConst<double> pc = pair_const<double>(gt, gs, args.k);
    const double wscale = 4.0 * gt[G_AREA] * gs[G_AREA] * 0.07957747154594766788;
    const double bx = gt[G_V0] - gs[G_V0], by = gt[G_V0 + 1] - gs[G_V0 + 1], bz = gt[G_V0 + 2] - gs[G_V0 + 2];
    const unsigned per = order * order * order;
    const unsigned total = sector_count(topo) * per;
    const int at = tj.permutation.test[0], as = tj.permutation.trial[0];
    MomentsT<double> m;
    zero_t(m);
    bool bad = false;
    for (unsigned idx = lane; idx < total; idx += 64) {
        unsigned tuple = idx % per;
        const unsigned wi = tuple % order;
        tuple /= order;
        const unsigned vi = tuple % order;
        const unsigned ui = tuple / order;
        PairPoint c = sauter_schwab_point(topo, idx / per, 1.0, gauss[2 * ui], gauss[2 * vi], gauss[2 * wi]);
        c.weight *= gauss[2 * ui + 1] * gauss[2 * vi + 1] * gauss[2 * wi + 1];
        const PairPoint p = unpermute(c, tj.permutation);
        const double ltx = gt[G_E1] * p.test[1] + gt[G_E2] * p.test[2],
                     lty = gt[G_E1 + 1] * p.test[1] + gt[G_E2 + 1] * p.test[2],
                     ltz = gt[G_E1 + 2] * p.test[1] + gt[G_E2 + 2] * p.test[2];
        const double lsx = gs[G_E1] * p.trial[1] + gs[G_E2] * p.trial[2],
                     lsy = gs[G_E1 + 1] * p.trial[1] + gs[G_E2 + 1] * p.trial[2],
                     lsz = gs[G_E1 + 2] * p.trial[1] + gs[G_E2 + 2] * p.trial[2];
        const double dx = bx + (ltx - lsx), dy = by + (lty - lsy), dz = bz + (ltz - lsz);
        for (unsigned ri = 0; ri < order; ++ri) {
            const double r = gauss[2 * ri];
            double xt[3], xs[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                xt[i] = r * p.test[i];
                xs[i] = r * p.trial[i];
            }
            xt[at] += 1 - r;
            xs[as] += 1 - r;
            const SampleT<double> s =
                sample(r * dx, r * dy, r * dz, c.weight * r * r * r * gauss[2 * ri + 1] * wscale, pc, coincident, bad);
            add_sample_t(m, s, xt, xs);
        }
    }
    if (bad)
        atomicExch(args.error, ERR_DISTANCE);
    wave_reduce_t(m);
    if (lane == 0)
        sink(args, m, job, tj.test, tj.trial);
}

__device__ __forceinline__ FanPanel fan_of(const double *g) {
    FanPanel f;
    f.v[0] = {g[G_V0], g[G_V0 + 1], g[G_V0 + 2]};
    f.v[1] = {g[G_V0] + g[G_E1], g[G_V0 + 1] + g[G_E1 + 1], g[G_V0 + 2] + g[G_E1 + 2]};
    f.v[2] = {g[G_V0] + g[G_E2], g[G_V0 + 1] + g[G_E2 + 1], g[G_V0 + 2] + g[G_E2 + 2]};
    f.normal = {g[G_N], g[G_N + 1], g[G_N + 2]};
    f.area = g[G_AREA];
    return f;
}

// Patch maps in FP64; samples and accumulation in R; matrix element type R.
template <class R, class Sink>
__global__ __launch_bounds__(256) void near_jobs(AssemblyArgs args, const NearJob *jobs, std::uint32_t njobs,
                                                 const PairPatch *patches, const double *rules, Sink sink) {
    const std::uint32_t job = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (job >= njobs)
        return;
    const NearJob nj = jobs[job];
    const PanelView &pv = args.panels;
    double gt[G_COUNT], gs[G_COUNT];
    load_geo(pv, nj.test, gt);
    load_geo(pv, nj.trial, gs);
    const PairConst<R> pc = pair_const<R>(gt, gs, args.k);
    const double wscale = 4.0 * gt[G_AREA] * gs[G_AREA] * 0.07957747154594766788;
    const double bx = gt[G_V0] - gs[G_V0], by = gt[G_V0 + 1] - gs[G_V0 + 1], bz = gt[G_V0 + 2] - gs[G_V0 + 2];
    const bool swap = nj.swap != 0;
    const FanPanel ft = fan_of(gt), fs = fan_of(gs);
    const TrianglePoint *r8 = reinterpret_cast<const TrianglePoint *>(rules + 4 * kRule8);
    const TrianglePoint *r14 = reinterpret_cast<const TrianglePoint *>(rules + 4 * kRule14);
    MomentsT<R> m;
    zero_t(m);
    bool bad = false;
    for (std::uint32_t q = 0; q < nj.patch_count; ++q) {
        const PairPatch patch = patches[nj.first_patch + q];
        const bool high = patch.degree == 14;
        const TrianglePoint *rule = high ? r14 : r8;
        const unsigned count = high ? 64 : 16;
        const unsigned points = patch_point_count(patch, count);
        for (unsigned idx = lane; idx < points; idx += 64) {
            PairPoint p = patch_point(rule, count, patch, idx, swap ? fs : ft, swap ? ft : fs);
            if (swap)
                for (int i = 0; i < 3; ++i) {
                    const double v = p.test[i];
                    p.test[i] = p.trial[i];
                    p.trial[i] = v;
                }
            const double ltx = gt[G_E1] * p.test[1] + gt[G_E2] * p.test[2],
                         lty = gt[G_E1 + 1] * p.test[1] + gt[G_E2 + 1] * p.test[2],
                         ltz = gt[G_E1 + 2] * p.test[1] + gt[G_E2 + 2] * p.test[2];
            const double lsx = gs[G_E1] * p.trial[1] + gs[G_E2] * p.trial[2],
                         lsy = gs[G_E1 + 1] * p.trial[1] + gs[G_E2 + 1] * p.trial[2],
                         lsz = gs[G_E1 + 2] * p.trial[1] + gs[G_E2 + 2] * p.trial[2];
            const SampleT<R> s = sample(static_cast<R>(bx + (ltx - lsx)), static_cast<R>(by + (lty - lsy)),
                                        static_cast<R>(bz + (ltz - lsz)), static_cast<R>(p.weight * wscale), pc,
                                        false, bad);
            const R xt[3] = {static_cast<R>(p.test[0]), static_cast<R>(p.test[1]), static_cast<R>(p.test[2])};
            const R xs[3] = {static_cast<R>(p.trial[0]), static_cast<R>(p.trial[1]), static_cast<R>(p.trial[2])};
            add_sample_t(m, s, xt, xs);
        }
    }
    if (bad)
        atomicExch(args.error, ERR_DISTANCE);
    wave_reduce_t(m);
    if (lane == 0)
        sink(args, m, job, nj.test, nj.trial);
}

// Tensor near patches, one per thread (degree-8 items first, then degree 14).
template <class R, class Sink>
__global__ __launch_bounds__(256) void near_patches(AssemblyArgs args, const NearPatchItem *items, std::uint32_t count,
                                                    const PairPatch *patches, const double *rules, Sink sink) {
    __shared__ R srule[4 * (16 + 64)];
    for (int e = threadIdx.x; e < 4 * 16; e += blockDim.x)
        srule[e] = static_cast<R>(rules[4 * kRule8 + e]);
    for (int e = threadIdx.x; e < 4 * 64; e += blockDim.x)
        srule[4 * 16 + e] = static_cast<R>(rules[4 * kRule14 + e]);
    __syncthreads();
    const std::uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count)
        return;
    const NearPatchItem it = items[idx];
    const PairPatch &patch = patches[it.patch];
    const PanelView &pv = args.panels;
    double gt[G_COUNT], gs[G_COUNT];
    load_geo(pv, it.test, gt);
    load_geo(pv, it.trial, gs);
    const bool high = patch.degree == 14;
    MomentsT<R> m;
    zero_t(m);
    bool bad = false;
    const double *Tt = it.swap ? patch.trial : patch.test, *Ts = it.swap ? patch.test : patch.trial;
    contract<R>(m, sub_frame(frame_of(gt), Tt), sub_frame(frame_of(gs), Ts),
                4.0 * gt[G_AREA] * gs[G_AREA] * 0.07957747154594766788 * patch.area_fraction,
                srule + (high ? 4 * 16 : 0), high ? 64 : 16, args.k, bad);
    to_parent(m, Tt, Ts);
    if (bad)
        atomicExch(args.error, ERR_DISTANCE);
    sink(args, m, idx, it.test, it.trial);
}

template <class M> __global__ void mass_terms(AssemblyArgs args) {
    const std::uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const PanelView &pv = args.panels;
    if (t >= pv.nt)
        return;
    M *A = matrix_of<M>(args);
    const double area = g_at(pv.geo, G_AREA, pv.nt, t);
    const std::int32_t fc = args.flux_column[t];
    for (int i = 0; i < 3; ++i) {
        const std::uint32_t row = pv.nodes[static_cast<std::size_t>(i) * pv.nt + t];
        const double rs = args.row_scale[row];
        for (int j = 0; j < 3; ++j) {
            const std::uint32_t col = pv.nodes[static_cast<std::size_t>(j) * pv.nt + t];
            unsafeAtomicAdd(A + 2 * (static_cast<std::size_t>(col) * pv.nv + row),
                            static_cast<M>(rs * area * (i == j ? 2.0 : 1.0) / 24.0));
        }
        if (fc >= 0)
            unsafeAtomicAdd(args.BD + 2 * (static_cast<std::size_t>(fc) * pv.nv + row) + 1, area / (6.0 * args.k) * rs);
    }
}

// rhs[e] = BD * qD[e]: one thread per (row, excitation).
__global__ void flux_rhs(const double *BD, std::uint32_t nv, std::uint32_t nd, const double *q, std::uint32_t nrhs,
                         double *rhs) {
    const std::size_t idx = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<std::size_t>(nv) * nrhs)
        return;
    const std::uint32_t row = idx % nv, e = idx / nv;
    double re = 0, im = 0;
    for (std::uint32_t c = 0; c < nd; ++c) {
        const double br = BD[2 * (static_cast<std::size_t>(c) * nv + row)],
                     bi = BD[2 * (static_cast<std::size_t>(c) * nv + row) + 1];
        const double qr = q[2 * (static_cast<std::size_t>(e) * nd + c)], qi = q[2 * (static_cast<std::size_t>(e) * nd + c) + 1];
        re += br * qr - bi * qi;
        im += br * qi + bi * qr;
    }
    rhs[2 * idx] = re;
    rhs[2 * idx + 1] = im;
}

// Regular (disjoint, not near) pairs of a list, one thread per pair: the same
// rule selection as regular_tiles_sym (all-pairs degree byte, else the qualified
// degree at the solve wavenumber) and FP64 contraction.
template <class Sink>
__global__ __launch_bounds__(256) void regular_pairs(AssemblyArgs args, NearView near, const std::uint32_t *pairs,
                                                     std::uint32_t count, const double *rules, Sink sink) {
    __shared__ double srule[4 * kRuleTableCount];
    for (int e = threadIdx.x; e < 4 * kRuleTableCount; e += blockDim.x)
        srule[e] = rules[e];
    __syncthreads();
    const std::uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count)
        return;
    const std::uint32_t ti = pairs[2 * idx], si = pairs[2 * idx + 1];
    const PanelView &pv = args.panels;
    int offset = -1, n = 0;
    if (near.whole) {
        const unsigned d = near.whole[unordered_pair_index(ti, si)];
        if (d) {
            offset = d == 14 ? kRule14 : kRule8;
            n = d == 14 ? 64 : 16;
        }
    }
    if (offset < 0) {
        PanelBounds bt, bs;
        bt.center = {pv.bounds[static_cast<std::size_t>(B_C) * pv.nt + ti], pv.bounds[static_cast<std::size_t>(B_C + 1) * pv.nt + ti],
                     pv.bounds[static_cast<std::size_t>(B_C + 2) * pv.nt + ti]};
        bt.radius = pv.bounds[static_cast<std::size_t>(B_R) * pv.nt + ti];
        bt.max_edge = pv.bounds[static_cast<std::size_t>(B_H) * pv.nt + ti];
        bt.qualified_shape = pv.shape[ti];
        bs.center = {pv.bounds[static_cast<std::size_t>(B_C) * pv.nt + si], pv.bounds[static_cast<std::size_t>(B_C + 1) * pv.nt + si],
                     pv.bounds[static_cast<std::size_t>(B_C + 2) * pv.nt + si]};
        bs.radius = pv.bounds[static_cast<std::size_t>(B_R) * pv.nt + si];
        bs.max_edge = pv.bounds[static_cast<std::size_t>(B_H) * pv.nt + si];
        bs.qualified_shape = pv.shape[si];
        const unsigned d = qualified_regular_degree(bt, bs, args.k, args.tolerance, args.regular_order);
        if (!d) {
            atomicExch(args.error, ERR_ESCAPED);
            return;
        }
        offset = d == 4 ? kRule4 : d == 6 ? kRule6 : kRule8;
        n = d == 4 ? 6 : d == 6 ? 12 : 16;
    }
    double gt[G_COUNT], gs[G_COUNT];
    load_geo(pv, ti, gt);
    load_geo(pv, si, gs);
    MomentsT<double> m;
    zero_t(m);
    bool bad = false;
    contract<double>(m, frame_of(gt), frame_of(gs), 4.0 * gt[G_AREA] * gs[G_AREA] * 0.07957747154594766788,
                     srule + 4 * offset, n, args.k, bad);
    if (bad)
        atomicExch(args.error, ERR_DISTANCE);
    sink(args, m, idx, ti, si);
}

// Mass terms into the near CSR: 1/2 M_pp (every panel's 3 x 3 block is near)
// and -alpha 1/2 M_pq into the active face columns.
__global__ void mass_csr(AssemblyArgs args, NearCsrView v) {
    const std::uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const PanelView &pv = args.panels;
    if (t >= pv.nt)
        return;
    const double area = g_at(pv.geo, G_AREA, pv.nt, t);
    const std::uint32_t fc = v.face_column[t];
    for (int i = 0; i < 3; ++i) {
        const std::uint32_t row = pv.nodes[static_cast<std::size_t>(i) * pv.nt + t];
        const double rs = args.row_scale[row];
        for (int j = 0; j < 3; ++j) {
            const std::uint32_t col = pv.nodes[static_cast<std::size_t>(j) * pv.nt + t];
            unsafeAtomicAdd(v.values + 2 * csr_slot(v.row_offsets, v.columns, row, col),
                            rs * area * (i == j ? 2.0 : 1.0) / 24.0);
        }
        if (fc != UINT32_MAX)
            unsafeAtomicAdd(v.flux_values + 2 * csr_slot(v.flux_offsets, v.flux_columns, row, fc) + 1,
                            area / (6.0 * args.k) * rs);
    }
}

// y = A x on the pressure CSR: one wavefront per row, lanes over columns.
__global__ __launch_bounds__(256) void csr_apply(const std::uint32_t *offsets, const std::uint32_t *columns,
                                                 const double *values, std::uint32_t rows, const double *x, double *y) {
    const std::uint32_t row = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (row >= rows)
        return;
    double re = 0, im = 0;
    for (std::uint32_t e = offsets[row] + lane; e < offsets[row + 1]; e += 64) {
        const double ar = values[2 * e], ai = values[2 * e + 1];
        const std::uint32_t c = columns[e];
        const double xr = x[2 * c], xi = x[2 * c + 1];
        re += ar * xr - ai * xi;
        im += ar * xi + ai * xr;
    }
    re = wave_sum_t(re);
    im = wave_sum_t(im);
    if (lane == 0) {
        y[2 * row] = re;
        y[2 * row + 1] = im;
    }
}

} // namespace

void launch_near_operator(const NearOperatorLaunch &L, hipStream_t stream) {
    NearCsrView v = L.csr;
    auto with = [&](const std::uint32_t *masks) {
        v.masks = masks;
        return CsrSink{v};
    };
    if (L.regular_count)
        hipLaunchKernelGGL(regular_pairs<CsrSink>, dim3((L.regular_count + 255) / 256), dim3(256), 0, stream, L.args,
                           L.near, L.regular, L.regular_count, L.rules, with(L.regular_masks));
    if (L.touching_count)
        hipLaunchKernelGGL(touching_jobs<CsrSink>, dim3((L.touching_count + 3) / 4), dim3(256), 0, stream, L.args,
                           L.touching, L.touching_count, L.gauss, L.singular_order, with(L.touching_masks));
    if (L.patch_count)
        hipLaunchKernelGGL((near_patches<double, CsrSink>), dim3((L.patch_count + 255) / 256), dim3(256), 0, stream,
                           L.args, L.patch_items, L.patch_count, L.patches, L.rules, with(L.patch_masks));
    if (L.fan_count)
        hipLaunchKernelGGL((near_jobs<double, CsrSink>), dim3((L.fan_count + 3) / 4), dim3(256), 0, stream, L.args,
                           L.fan_jobs, L.fan_count, L.patches, L.rules, with(L.fan_masks));
    hipLaunchKernelGGL(mass_csr, dim3((L.args.panels.nt + 255) / 256), dim3(256), 0, stream, L.args, L.csr);
}

void launch_csr_apply(const std::uint32_t *offsets, const std::uint32_t *columns, const double *values,
                      std::uint32_t rows, const double *x, double *y, hipStream_t stream) {
    hipLaunchKernelGGL(csr_apply, dim3((rows + 3) / 4), dim3(256), 0, stream, offsets, columns, values, rows, x, y);
}

namespace {

// The degree 4, 6 and 8 rules in the kernel-argument form (same points and order
// as the device rule table).
template <class R> TileRules<R> make_tile_rules() {
    TileRules<R> t{};
    auto fill = [](R(*dst)[4], unsigned degree, std::size_t count) {
        const auto rule = triangle_rule(degree);
        if (rule.size() != count)
            throw std::logic_error("regular rule of degree " + std::to_string(degree) + " has " +
                                   std::to_string(rule.size()) + " points, the tile kernels expect " +
                                   std::to_string(count));
        for (std::size_t i = 0; i < count; ++i) {
            for (int c = 0; c < 3; ++c)
                dst[i][c] = static_cast<R>(rule[i].bary[c]);
            dst[i][3] = static_cast<R>(rule[i].weight);
        }
    };
    fill(t.d4, 4, 6);
    fill(t.d6, 6, 12);
    fill(t.d8, 8, 16);
    return t;
}

} // namespace

// The tile list holds unordered cluster pairs (a <= b): tiles[0, far_tiles) take the
// uniform far path at this wavenumber, the rest the general one. The mixed form forms
// both orientations of a pair in one workgroup; the FP64 form launches every listed
// tile and its mirror.
void launch_regular(const AssemblyLaunch &L, hipStream_t stream) {
    const std::uint32_t general = L.tile_count - L.far_tiles;
    auto shift_of = [](std::uint32_t block) {
        if (block != 4 && block != 8 && block != 16)
            throw std::invalid_argument("regular tiles: the trial block must be 4, 8 or 16 panels");
        return block == 4 ? 2u : block == 8 ? 3u : 4u;
    };
    const std::uint32_t shift = shift_of(L.trial_block), far_shift = shift_of(L.far_trial_block);
    const std::uint32_t *general_tiles = L.tiles + 2 * static_cast<std::size_t>(L.far_tiles);
    if (L.args.A32) {
        static const TileRules<float> fixed = make_tile_rules<float>();
        if (L.far_tiles)
            hipLaunchKernelGGL((regular_tiles_sym<float, true>), dim3(L.far_tiles), dim3(kTileThreads), 0, stream,
                               L.args, L.clusters, L.near, L.rules, L.tiles, fixed, far_shift);
        if (general)
            hipLaunchKernelGGL((regular_tiles_sym<float, false>), dim3(general), dim3(kTileThreads), 0, stream,
                               L.args, L.clusters, L.near, L.rules, general_tiles, fixed, shift);
    } else {
        static const TileRules<double> fixed = make_tile_rules<double>();
        if (L.far_tiles)
            hipLaunchKernelGGL((regular_tiles_sym<double, true>), dim3(L.far_tiles), dim3(kTileThreads), 0, stream,
                               L.args, L.clusters, L.near, L.rules, L.tiles, fixed, far_shift);
        if (general)
            hipLaunchKernelGGL((regular_tiles_sym<double, false>), dim3(general), dim3(kTileThreads), 0, stream,
                               L.args, L.clusters, L.near, L.rules, general_tiles, fixed, shift);
    }
}

void launch_touching(const AssemblyLaunch &L, hipStream_t stream) {
    if (!L.touching_count)
        return;
    const dim3 grid((L.touching_count + 3) / 4);
    if (L.args.A32)
        hipLaunchKernelGGL(touching_jobs<DenseSink<float>>, grid, dim3(256), 0, stream, L.args, L.touching,
                           L.touching_count, L.gauss, L.singular_order, DenseSink<float>{});
    else
        hipLaunchKernelGGL(touching_jobs<DenseSink<double>>, grid, dim3(256), 0, stream, L.args, L.touching,
                           L.touching_count, L.gauss, L.singular_order, DenseSink<double>{});
}

void launch_near(const AssemblyLaunch &L, hipStream_t stream) {
    if (L.near_item_count) {
        const dim3 grid((L.near_item_count + 255) / 256);
        if (L.args.A32)
            hipLaunchKernelGGL((near_patches<float, DenseSink<float>>), grid, dim3(256), 0, stream, L.args,
                               L.near_items, L.near_item_count, L.patches, L.rules, DenseSink<float>{});
        else
            hipLaunchKernelGGL((near_patches<double, DenseSink<double>>), grid, dim3(256), 0, stream, L.args,
                               L.near_items, L.near_item_count, L.patches, L.rules, DenseSink<double>{});
    }
    if (L.near_count) {
        const dim3 grid((L.near_count + 3) / 4);
        if (L.args.A32)
            hipLaunchKernelGGL((near_jobs<float, DenseSink<float>>), grid, dim3(256), 0, stream, L.args, L.near_jobs,
                               L.near_count, L.patches, L.rules, DenseSink<float>{});
        else
            hipLaunchKernelGGL((near_jobs<double, DenseSink<double>>), grid, dim3(256), 0, stream, L.args,
                               L.near_jobs, L.near_count, L.patches, L.rules, DenseSink<double>{});
    }
}

void launch_mass(const AssemblyLaunch &L, hipStream_t stream) {
    const dim3 grid((L.args.panels.nt + 255) / 256);
    if (L.args.A32)
        hipLaunchKernelGGL(mass_terms<float>, grid, dim3(256), 0, stream, L.args);
    else
        hipLaunchKernelGGL(mass_terms<double>, grid, dim3(256), 0, stream, L.args);
}

void launch_assembly(const AssemblyLaunch &L, hipStream_t stream) {
    launch_regular(L, stream);
    launch_touching(L, stream);
    launch_near(L, stream);
    launch_mass(L, stream);
}

void launch_flux_rhs(const double *BD, std::uint32_t nv, std::uint32_t nd, const double *q, std::uint32_t nrhs,
                     double *rhs, hipStream_t stream) {
    const std::size_t n = static_cast<std::size_t>(nv) * nrhs;
    hipLaunchKernelGGL(flux_rhs, dim3((n + 255) / 256), dim3(256), 0, stream, BD, nv, nd, q, nrhs, rhs);
}

} // namespace hsbem::dev
