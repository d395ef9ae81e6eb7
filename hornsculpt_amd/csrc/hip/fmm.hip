// Far field of the wideband FMM operator (low-frequency spherical form), FP64.
//
// Replaces the device side of SupportLeafOperator and the LF translations of
// SupportFmmOperator (modules/bem/cuda/wideband_leaf.cu:14-903,
// wideband_rotation.cu, wideband_axial.cu, support_fmm_operator.cu:1103-1296):
//   * leaf moments   per (leaf, support panel): the panel rule's samples of the
//                    regular waves R_lm(y - c) = Jhat_l(k r) (r/a)^l Y_lm and their
//                    gradients, reduced over a 16-lane group per panel, give the
//                    emission rows conj(n . grad R_lm) phi_j and the testing rows
//                    (-R_lm - alpha n . grad R_lm) phi_i / m_i of the leaf's P1 DOFs;
//   * emission       multipole of a leaf = E_leaf p (one block per leaf);
//   * translation    every M2M / M2L / L2L link is rotate-to-axis, coaxial m-block
//                    translation, rotate-back (O(p^3)): the rotations run per chunk
//                    of links sharing a direction with the Wigner blocks staged in
//                    LDS, the coaxial products per link, and one block per
//                    destination sums its links (no atomics);
//   * testing        y_i += T_leaf L_leaf (one wavefront per DOF).
#include "fmm_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>

namespace hsbem::dev {
namespace {

constexpr int kGroupSmall = 16; // lanes per (leaf, panel) item: rules of at most 16 points
constexpr int kGroupLarge = 64; // one wavefront per item: rules of at most 64 points (degrees 12, 14)
constexpr int kMillerMax = 256; // Miller recurrence start bound (z <= 96 at the leaves)

struct C2 {
    double re, im;
};
__device__ __forceinline__ C2 operator+(C2 a, C2 b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ C2 operator-(C2 a, C2 b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ C2 operator*(C2 a, C2 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ C2 operator*(double s, C2 a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ C2 conj(C2 a) { return {a.re, -a.im}; }

// FP32 complex (the FP32-bulk form's expansions and translation workspaces)
struct C2f {
    float re, im;
};
__device__ __forceinline__ C2f operator+(C2f a, C2f b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ C2f operator*(C2f a, C2f b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ C2f operator*(float s, C2f a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ C2f conj(C2f a) { return {a.re, -a.im}; }
// complex type of an expansion scalar E (double: C2, float: C2f)
template <class E> struct CplxOf {
    using type = C2;
};
template <> struct CplxOf<float> {
    using type = C2f;
};
template <class E> using Cx = typename CplxOf<E>::type;
template <class E> __device__ __forceinline__ Cx<E> to_cx(C2 v) { return {static_cast<E>(v.re), static_cast<E>(v.im)}; }

// value and gradient of one solid harmonic
struct Wave {
    C2 v, g[3];
};

__device__ __forceinline__ Wave product(const Wave &a, const Wave &b) {
    Wave r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        r.g[i] = a.g[i] * b.v + a.v * b.g[i];
    return r;
}

__device__ __forceinline__ Wave scaled(Wave a, double s) {
    a.v = s * a.v;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        a.g[i] = s * a.g[i];
    return a;
}

// Jhat_l(z) = (2l+1)!! j_l(z) / z^l for l = 0..top (top <= kMaxBigOrder + 1): the
// power series while its first correction z^2 / (2 (2l+3)) is at most one, else
// j_l by normalised Miller recursion (downward from past the turning point,
// normalised by the larger of j0 and j1) divided by z^l / (2l+1)!!.
__device__ bool normalized_regular(int top, double z, double *J) {
    int split = 0; // l < split use the j_l table
    if (z > 1)
        while (split <= top && z * z > 2 * (2.0 * split + 3))
            ++split;
    if (split) {
        const int start = max(split + 48, static_cast<int>(ceil(z + 48 + 8 * cbrt(z))));
        if (start + 2 > kMillerMax)
            return false;
        double mil[kMillerMax];
        mil[start + 1] = 0;
        mil[start] = 1;
        for (int l = start; l > 0; --l) {
            mil[l - 1] = ((2.0 * l + 1) / z) * mil[l] - mil[l + 1];
            if (fabs(mil[l - 1]) > 1e200)
                for (int j = l - 1; j <= start + 1; ++j)
                    mil[j] *= 1e-200;
        }
        double s, c;
        sincos(z, &s, &c);
        const double j0 = s / z, j1 = (j0 - c) / z;
        const double norm = fabs(j0) > fabs(j1) ? j0 / mil[0] : j1 / mil[1];
        double lead = 1;
        for (int l = 0; l < split; ++l) {
            if (l)
                lead *= z / (2.0 * l + 1);
            J[l] = norm * mil[l] / lead;
        }
    }
    for (int l = split; l <= top; ++l) {
        double term = 1, sum = 1;
        for (int n = 1; n < 128; ++n) {
            term *= -z * z / (2.0 * n * (2.0 * l + 2 * n + 1));
            sum += term;
            if (fabs(term) < 2.220446049250313e-16 * fabs(sum))
                break;
        }
        J[l] = sum;
    }
    return true;
}

template <int kGroup> __device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = kGroup / 2; off > 0; off >>= 1)
        v += __shfl_xor(v, off, kGroup);
    return v;
}

__device__ __forceinline__ void atomic_add_c(double *p, C2 v) {
    atomicAdd(p, v.re);
    atomicAdd(p + 1, v.im);
}

// One kGroup-lane group per (leaf, support panel) item, lane = rule point.
template <int kGroup, int kTop> __global__ __launch_bounds__(64) void leaf_moments(FmmLeafMomentsLaunch L) {
    const int lane = threadIdx.x % kGroup;
    const std::uint32_t item = blockIdx.x * (64 / kGroup) + threadIdx.x / kGroup;
    const bool live_item = item < L.item_count;
    const FmmLeafItem it = live_item ? L.items[item] : FmmLeafItem{};
    const FmmLeaf leaf = L.leaves[live_item ? it.leaf : 0];
    const std::uint32_t nt = L.nt, t = it.panel;
    const bool live = live_item && lane < L.rule_points;
    const double *rp = L.rule + 4 * (live ? lane : 0);
    const double b0 = rp[0], b1 = rp[1], b2 = rp[2];
    const double *g = L.geo;
    const double area = g[static_cast<std::size_t>(G_AREA) * nt + t];
    const double w = live ? rp[3] * 2 * area : 0.0; // rule weights sum to 1/2
    double y[3], n[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double v0 = g[static_cast<std::size_t>(G_V0 + i) * nt + t];
        y[i] = v0 + b1 * g[static_cast<std::size_t>(G_E1 + i) * nt + t] + b2 * g[static_cast<std::size_t>(G_E2 + i) * nt + t];
        n[i] = g[static_cast<std::size_t>(G_N + i) * nt + t];
    }
    const double a = leaf.radius, k = L.k;
    const double x[3] = {y[0] - leaf.center[0], y[1] - leaf.center[1], y[2] - leaf.center[2]};
    const double u[3] = {x[0] / a, x[1] / a, x[2] / a};
    const int p = static_cast<int>(leaf.order);
    double J[kTop + 2];
    const bool ok = normalized_regular(p + 1, k * sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]), J);
    if (!ok && live && L.error)
        atomicMax(L.error, 1);
    const double bc[3] = {b0, b1, b2};
    const C2 alpha{L.alpha_re, L.alpha_im};
    const std::size_t nc = static_cast<std::size_t>(p + 1) * (p + 1);
    // solid harmonics |u|^l Y_lm(u) as polynomials in x + iy, z, r^2 with gradients in u
    const Wave xy{{u[0], u[1]}, {{1, 0}, {0, 1}, {0, 0}}}, zz{{u[2], 0}, {{0, 0}, {0, 0}, {1, 0}}},
        r2{{u[0] * u[0] + u[1] * u[1] + u[2] * u[2], 0}, {{2 * u[0], 0}, {2 * u[1], 0}, {2 * u[2], 0}}};
    Wave diag{{0.28209479177387814, 0}, {{0, 0}, {0, 0}, {0, 0}}}; // 1 / sqrt(4 pi)
    for (int m = 0; m <= p; ++m) {
        if (m)
            diag = scaled(product(xy, diag), -sqrt((2.0 * m + 1) / (2.0 * m)));
        Wave prev{}, cur = diag;
        for (int l = m; l <= p; ++l) {
            if (l == m + 1) {
                prev = cur;
                cur = scaled(product(zz, diag), sqrt(2.0 * m + 3));
            } else if (l > m + 1) {
                const double den = static_cast<double>(l) * l - static_cast<double>(m) * m;
                const double ca = sqrt((4.0 * l * l - 1) / den),
                             cb = sqrt((2.0 * l + 1) * ((l - 1.0) * (l - 1.0) - static_cast<double>(m) * m) /
                                       ((2.0 * l - 3) * den));
                const Wave za = scaled(product(zz, cur), ca), rb = scaled(product(r2, prev), cb);
                Wave nx;
                nx.v = za.v - rb.v;
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    nx.g[i] = za.g[i] - rb.g[i];
                prev = cur;
                cur = nx;
            }
            // R_lm and n . grad R_lm
            const double jl = J[l], jl1 = J[l + 1];
            const C2 R = jl * cur.v;
            C2 dn{0, 0};
#pragma unroll
            for (int i = 0; i < 3; ++i)
                dn = dn + n[i] * ((jl / a) * cur.g[i] - (k * k * x[i] / (2.0 * l + 3) * jl1) * cur.v);
            // emission conj(dn) w, testing (-R - alpha dn) w; the -m partners follow from
            // R_{l,-m} = (-1)^m conj(R_lm) (real n and Jhat)
            const C2 e = w * conj(dn), tt = w * (C2{-R.re, -R.im} - alpha * dn);
            const C2 em = (m % 2 ? -w : w) * dn,
                     tm = (m % 2 ? -w : w) * (C2{-R.re, R.im} - alpha * conj(dn));
            const std::size_t jp = static_cast<std::size_t>(l) * l + l + m, jm = static_cast<std::size_t>(l) * l + l - m;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double s = bc[c];
                const C2 sums[4] = {{group_sum<kGroup>(s * e.re), group_sum<kGroup>(s * e.im)},
                                    {group_sum<kGroup>(s * tt.re), group_sum<kGroup>(s * tt.im)},
                                    {group_sum<kGroup>(s * em.re), group_sum<kGroup>(s * em.im)},
                                    {group_sum<kGroup>(s * tm.re), group_sum<kGroup>(s * tm.im)}};
                const std::int32_t j = it.corner[c];
                if (live_item && lane == 0 && j >= 0) {
                    const std::size_t row = static_cast<std::size_t>(j) * nc;
                    const double scale = L.row_scale[L.leaf_dofs[leaf.first_dof + j]];
                    atomic_add_c(L.emission + 2 * (leaf.emission + row + jp), sums[0]);
                    atomic_add_c(L.testing + 2 * (leaf.emission + row + jp), scale * sums[1]);
                    if (m) {
                        atomic_add_c(L.emission + 2 * (leaf.emission + row + jm), sums[2]);
                        atomic_add_c(L.testing + 2 * (leaf.emission + row + jm), scale * sums[3]);
                    }
                }
            }
        }
    }
}

// M_leaf = E_leaf p: one block per leaf, thread per coefficient; the expansions are
// stored as S (the sums run in FP64).
template <class S>
__global__ __launch_bounds__(256) void leaf_emit(const FmmLeaf *leaves, const std::uint32_t *leaf_dofs,
                                                 const double *emission, const double *p, S *expansions) {
    __shared__ double sp[2 * kMaxLeafDofs];
    const FmmLeaf leaf = leaves[blockIdx.x];
    for (std::uint32_t j = threadIdx.x; j < leaf.dof_count; j += blockDim.x) {
        const std::uint32_t d = leaf_dofs[leaf.first_dof + j];
        sp[2 * j] = p[2 * d];
        sp[2 * j + 1] = p[2 * d + 1];
    }
    __syncthreads();
    const std::size_t nc = static_cast<std::size_t>(leaf.order + 1) * (leaf.order + 1);
    const double2 *E = reinterpret_cast<const double2 *>(emission) + leaf.emission;
    for (std::size_t c = threadIdx.x; c < nc; c += blockDim.x) {
        double re = 0, im = 0;
        for (std::uint32_t j = 0; j < leaf.dof_count; ++j) {
            const double2 e = E[j * nc + c];
            re += e.x * sp[2 * j] - e.y * sp[2 * j + 1];
            im += e.x * sp[2 * j + 1] + e.y * sp[2 * j];
        }
        expansions[2 * (leaf.multipole + c)] = static_cast<S>(re);
        expansions[2 * (leaf.multipole + c) + 1] = static_cast<S>(im);
    }
}

// y_i += T_leaf[i] . L_leaf: one wavefront per DOF; the expansions are read as S.
template <class S>
__global__ __launch_bounds__(256) void leaf_test(const FmmLeaf *leaves, const std::uint32_t *leaf_dofs,
                                                 const double *testing, const S *expansions, double *y) {
    const FmmLeaf leaf = leaves[blockIdx.x];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const std::size_t nc = static_cast<std::size_t>(leaf.order + 1) * (leaf.order + 1);
    const double2 *T = reinterpret_cast<const double2 *>(testing) + leaf.emission;
    const Cx<S> *Lc = reinterpret_cast<const Cx<S> *>(expansions) + leaf.local;
    for (std::uint32_t j = wave; j < leaf.dof_count; j += blockDim.x / 64) {
        double re = 0, im = 0;
        for (std::size_t c = lane; c < nc; c += 64) {
            const double2 t = T[j * nc + c];
            const double lr = Lc[c].re, li = Lc[c].im;
            re += t.x * lr - t.y * li;
            im += t.x * li + t.y * lr;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            re += __shfl_xor(re, off, 64);
            im += __shfl_xor(im, off, 64);
        }
        if (lane == 0) {
            const std::uint32_t d = leaf_dofs[leaf.first_dof + j];
            y[2 * d] += re;
            y[2 * d + 1] += im;
        }
    }
}

__device__ __forceinline__ void split_index(int e, int &l, int &t) {
    l = static_cast<int>(sqrt(static_cast<double>(e)));
    while (l * l > e)
        --l;
    while ((l + 1) * (l + 1) <= e)
        ++l;
    t = e - l * l;
}

// Offset of the |m| block of a coaxial plan: sum_{j<m} (pt+1-j)(ps+1-j).
__device__ __forceinline__ std::size_t axial_block(int m, int ps, int pt) {
    std::size_t off = 0;
    for (int j = 0; j < m; ++j)
        off += static_cast<std::size_t>(pt + 1 - j) * (ps + 1 - j);
    return off;
}

// Position of coefficient (l, m) of an order-p expansion in |m|-major order: blocks
// |m| = 0..p, each holding l = |m|..p for +|m| and then (|m| > 0) for -|m|; block |m|
// starts at (2|m|-1)(p+1) - (|m|-1)|m| (0 for |m| = 0). The forward rotation writes
// the rotated source in this order so the coaxial products read contiguous columns.
__device__ __forceinline__ std::size_t m_major(int p, int l, int m) {
    const int a = m < 0 ? -m : m;
    const std::size_t block = a ? static_cast<std::size_t>(2 * a - 1) * (p + 1) - static_cast<std::size_t>(a - 1) * a : 0;
    return block + (m < 0 ? p - a + 1 : 0) + (l - a);
}

// Offset of degree l in a rotation's small-d table: sum_{j<l} (2j+1)^2 = l(2l-1)(2l+1)/3.
__device__ __forceinline__ std::size_t degree_block(int l) {
    return static_cast<std::size_t>(l) * (2 * l - 1) * (2 * l + 1) / 3;
}

// Translations run as four passes over the links of one level:
//   rotate    per chunk of links sharing a direction: the Wigner blocks d_l are
//             staged in LDS degree by degree and applied to every source of the
//             chunk (Y = d^T e^{i m phi} X, stored |m|-major);
//   coaxial   one wavefront per link: per-|m| block product Z = T Y;
//   back      per chunk again: W = e^{-i m phi} d Z;
//   gather    one block per destination: the sum of its links' W.
// Y and W share one workspace slot per link, Z uses a second.
// Two degrees per stage, l and top - l: their Wigner blocks d_l and the chunk's
// slices of both degrees are staged in LDS together, so every stage carries about
// the same work ((2l+1)^2 + (2(top-l)+1)^2 multiply-adds per link) and the number of
// barrier pairs halves; the next stage's values are loaded into registers while the
// current one is computed. Workgroup (chunk, y) takes the stages h = y, y + Y, ... of
// its chunk (Y = gridDim.y): the degrees are independent, so a level with few chunks
// spreads its stages over more workgroups instead of running them one after another.
// Instantiated for orders up to kP = 32, 40 and 48 (LDS and registers sized to kP, so
// the common orders 24-32 keep three workgroups per CU).
template <int kP, class E>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kP <= 32 ? 3 : 1))) void rotate_links(FmmTranslateLaunch L, bool back) {
    using CE = Cx<E>;
    constexpr int kStage = (2 * kP + 1) * (2 * kP + 1) + 1; // >= n_a^2 + n_b^2 for n_a + n_b <= 2 kP + 2
    constexpr int kStagePerThread = (kStage + 255) / 256;
    constexpr int kSliceWidth = 2 * kP + 2;                 // >= n_a + n_b
    constexpr int kSlicePerThread = (kFmmChunk * kSliceWidth + 255) / 256;
    __shared__ E dl[kStage];
    __shared__ CE xs[kFmmChunk][kSliceWidth]; // rows >= count: stale, their results are dropped
    __shared__ CE ph[2 * kP + 1];
    __shared__ std::uint64_t src_off[kFmmChunk], dst_off[kFmmChunk];
    __shared__ int ord[kFmmChunk];
    const FmmChunk ch = L.chunks[blockIdx.x];
    const int split = static_cast<int>(blockIdx.y), splits = static_cast<int>(gridDim.y);
    if (split > static_cast<int>(ch.max_order) / 2)
        return; // this chunk has fewer stages than the launch has splits
    const FmmRotation rot = L.rotations[ch.rotation];
    const C2 *az = reinterpret_cast<const C2 *>(L.azimuth) + rot.azimuth;
    for (int m = threadIdx.x; m <= 2 * static_cast<int>(rot.order); m += blockDim.x)
        ph[m] = to_cx<E>(az[m]);
    if (threadIdx.x < ch.count) {
        const FmmLink lk = L.links[L.order[ch.first + threadIdx.x]];
        src_off[threadIdx.x] = back ? lk.slot : lk.source;
        dst_off[threadIdx.x] = lk.slot;
        ord[threadIdx.x] = static_cast<int>(back ? lk.target_order : lk.source_order);
    }
    __syncthreads();
    const CE *in = reinterpret_cast<const CE *>(back ? L.work_z : L.source_expansions);
    CE *out = reinterpret_cast<CE *>(L.work_y);
    const int count = static_cast<int>(ch.count), P = static_cast<int>(rot.order), top = static_cast<int>(ch.max_order);
    const int stages = top / 2 + 1; // stage h: degrees h and top - h (one degree when they coincide)
    // degrees of stage h and their sizes; nb = 0 when the stage holds one degree
    auto degrees = [&](int h, int &la, int &lb, int &na, int &nb) {
        la = h;
        lb = top - h;
        na = 2 * la + 1;
        nb = lb > la ? 2 * lb + 1 : 0;
    };
    E pd[kStagePerThread];
    CE px[kSlicePerThread];
    auto fetch = [&](int h) {
        int la, lb, na, nb;
        degrees(h, la, lb, na, nb);
        const double *srca = L.small_d + rot.small_d + degree_block(la);
        const double *srcb = L.small_d + rot.small_d + degree_block(lb);
#pragma unroll
        for (int i = 0; i < kStagePerThread; ++i) {
            const int e = threadIdx.x + 256 * i;
            pd[i] = static_cast<E>(e < na * na ? srca[e] : e < na * na + nb * nb ? srcb[e - na * na] : 0.0);
        }
        const int w = na + nb;
#pragma unroll
        for (int i = 0; i < kSlicePerThread; ++i) {
            const int e = threadIdx.x + 256 * i;
            CE v{0, 0};
            if (e < count * w) {
                const int j = e / w, q = e % w;
                const int l = q < na ? la : lb, qq = q < na ? q : q - na;
                if (l <= ord[j])
                    v = in[src_off[j] + static_cast<std::size_t>(l) * l + qq];
            }
            px[i] = v;
        }
    };
    fetch(split);
    for (int h = split; h < stages; h += splits) {
        int la, lb, na, nb;
        degrees(h, la, lb, na, nb);
        const int w = na + nb;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kStagePerThread; ++i) {
            const int e = threadIdx.x + 256 * i;
            if (e < na * na + nb * nb)
                dl[e] = pd[i];
        }
#pragma unroll
        for (int i = 0; i < kSlicePerThread; ++i) {
            const int e = threadIdx.x + 256 * i;
            if (e < count * w) {
                const int j = e / w, q = e % w;
                const int l = q < na ? la : lb, qq = q < na ? q : q - na;
                xs[j][q] = back ? px[i] : ph[P + qq - l] * px[i];
            }
        }
        __syncthreads();
        if (h + splits < stages)
            fetch(h + splits);
        // thread = (output t of one of the two degrees, group of 4 links): one Wigner
        // load feeds 4 accumulators
        const int groups = (count + 3) / 4;
        for (int e = threadIdx.x; e < groups * w; e += blockDim.x) {
            const int g = e / w, tq = e % w, j0 = 4 * g;
            const bool first = tq < na;
            const int l = first ? la : lb, n = first ? na : nb, t = first ? tq : tq - na, x0 = first ? 0 : na;
            const E *dbl = dl + (first ? 0 : na * na);
            CE s0{0, 0}, s1{0, 0}, s2{0, 0}, s3{0, 0};
            const int stride = back ? 1 : n;
            const E *drow = back ? dbl + t * n : dbl + t;
            // three steps per pass: their 15 LDS reads are issued before the products
            // (three, not four, keeps the order-32 kernel within the 168 registers of
            // three waves per SIMD: measured 4.18 ms per horn apply against 4.57 ms)
            constexpr int kSteps = 3;
            int q = 0;
            for (; q + kSteps <= n; q += kSteps) {
                E d[kSteps];
                CE x[kSteps][4];
#pragma unroll
                for (int u = 0; u < kSteps; ++u) {
                    d[u] = drow[(q + u) * stride];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        x[u][i] = xs[j0 + i][x0 + q + u];
                }
#pragma unroll
                for (int u = 0; u < kSteps; ++u) {
                    s0 = s0 + d[u] * x[u][0];
                    s1 = s1 + d[u] * x[u][1];
                    s2 = s2 + d[u] * x[u][2];
                    s3 = s3 + d[u] * x[u][3];
                }
            }
            for (; q < n; ++q) {
                const E d = drow[q * stride];
                s0 = s0 + d * xs[j0][x0 + q];
                s1 = s1 + d * xs[j0 + 1][x0 + q];
                s2 = s2 + d * xs[j0 + 2][x0 + q];
                s3 = s3 + d * xs[j0 + 3][x0 + q];
            }
            const CE sv[4] = {s0, s1, s2, s3};
            const CE rot_back = back ? conj(ph[P + t - l]) : CE{1, 0};
            const std::size_t base = static_cast<std::size_t>(l) * l;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j = j0 + i;
                if (j < count && l <= ord[j])
                    out[dst_off[j] + (back ? base + t : m_major(ord[j], l, t - l))] = back ? rot_back * sv[i] : sv[i];
            }
        }
    }
}

// Z = T Y per |m| block for a chunk of links sharing one coaxial plan, two blocks per
// stage (|m| = h and mm - h, which balances the rows x columns of the two): the blocks
// T_m and the chunk's order-m columns of Y are staged in LDS, the next stage's values
// prefetched into registers during the current products. Workgroup (chunk, y) takes
// the stages y, y + Y, ... (Y = gridDim.y) of its chunk's stage sequence.
// T_m rows are stored with an odd stride (cols | 1) of 16-byte values, so the
// consecutive rows read by consecutive lanes fall in distinct bank groups.
template <int kP, class E>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kP <= 32 ? 4 : 1))) void coaxial_links(FmmTranslateLaunch L) {
    using CE = Cx<E>;
    constexpr int kAxialRows = kP + 1;
    // A stage pairs |m| = h with mm - h while their columns fit kColumns (always when
    // ps <= pt: cols_a + cols_b = ps + 2), else holds one block. Its blocks hold at most
    // (pt + 1)(ps + 2) <= kAxialRows^2 + kAxialRows values, stored with odd row strides
    // (at most rows_a + rows_b <= 2 kAxialRows more).
    constexpr int kBlockElems = kAxialRows * kAxialRows + kAxialRows;
    constexpr int kBlock = kAxialRows * (kAxialRows + 3);
    constexpr int kBlockPerThread = (kBlockElems + 255) / 256;
    constexpr int kColumns = kAxialRows + 1;
    constexpr int kColumnPerThread = (kFmmChunk * 2 * kColumns + 255) / 256;
    __shared__ CE tm[kBlock];
    __shared__ CE ys[kFmmChunk][2][kColumns];
    __shared__ std::uint64_t slot[kFmmChunk];
    const FmmChunk ch = L.axial_chunks[blockIdx.x];
    const int split = static_cast<int>(blockIdx.y), splits = static_cast<int>(gridDim.y);
    const FmmLink first = L.links[L.axial_order[ch.first]];
    const int ps = static_cast<int>(first.source_order), pt = static_cast<int>(first.target_order), mm = min(ps, pt);
    const C2 *T = reinterpret_cast<const C2 *>(L.axial) + L.axial_offsets[first.axial];
    const CE *y = reinterpret_cast<const CE *>(L.work_y);
    CE *z = reinterpret_cast<CE *>(L.work_z);
    const int count = static_cast<int>(ch.count);
    if (threadIdx.x < count)
        slot[threadIdx.x] = L.links[L.axial_order[ch.first + threadIdx.x]].slot;
    // degrees above min(ps, pt) in |m| have no source: those outputs are zero
    const int nct = (pt + 1) * (pt + 1);
    __syncthreads();
    if (split == 0)
        for (int e = threadIdx.x; e < count * nct; e += blockDim.x) {
            const int j = e / nct, c = e % nct;
            int l, t;
            split_index(c, l, t);
            if (abs(t - l) > mm)
                z[slot[j] + c] = {0, 0};
        }
    struct Pair {
        int ma, mb, ra, ca, rb, cb; // rb = cb = 0 when the stage holds one block
        int lo, hi;                 // the |m| still to do after this stage: lo .. hi
    };
    // the stage that starts with |m| in lo .. hi still to do
    auto blocks = [&](int lo, int hi) {
        Pair q;
        q.ma = lo;
        q.ra = pt - lo + 1;
        q.ca = ps - lo + 1;
        const bool two = hi > lo && q.ca + (ps - hi + 1) <= kColumns;
        q.mb = two ? hi : lo;
        q.rb = two ? pt - hi + 1 : 0;
        q.cb = two ? ps - hi + 1 : 0;
        q.lo = lo + 1;
        q.hi = two ? hi - 1 : hi;
        return q;
    };
    CE pt_[kBlockPerThread], py[kColumnPerThread];
    auto fetch = [&](const Pair &q) {
        const C2 *Ta = T + axial_block(q.ma, ps, pt), *Tb = T + axial_block(q.mb, ps, pt);
        const int sa = q.ra * q.ca, sb = q.rb * q.cb;
#pragma unroll
        for (int i = 0; i < kBlockPerThread; ++i) {
            const int e = threadIdx.x + 256 * i;
            pt_[i] = to_cx<E>(e < sa ? Ta[e] : e < sa + sb ? Tb[e - sa] : C2{0, 0});
        }
        const int w = q.ca + q.cb;
#pragma unroll
        for (int i = 0; i < kColumnPerThread; ++i) {
            const int e = threadIdx.x + 256 * i;
            const bool live = e < count * 2 * w;
            const int ee = live ? e : 0;
            const int j = ee / (2 * w), r = ee % (2 * w), sgn = r / w, c = r % w;
            const int m = c < q.ca ? q.ma : q.mb, n = m + (c < q.ca ? c : c - q.ca);
            py[i] = live ? y[slot[j] + m_major(ps, n, sgn ? -m : m)] : CE{0, 0};
        }
    };
    // the stage `steps` stages after p; false when the sequence ends first
    auto advance = [&](Pair &p, int steps) {
        for (int i = 0; i < steps; ++i) {
            if (p.lo > p.hi)
                return false;
            p = blocks(p.lo, p.hi);
        }
        return true;
    };
    Pair q = blocks(0, mm);
    if (!advance(q, split))
        return;
    fetch(q);
    for (;;) {
        const int sa = q.ra * q.ca, w = q.ca + q.cb;
        const int stride_a = q.ca | 1, stride_b = q.cb | 1, base_b = q.ra * stride_a;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kBlockPerThread; ++i) {
            const int e = threadIdx.x + 256 * i;
            if (e < sa)
                tm[(e / q.ca) * stride_a + e % q.ca] = pt_[i];
            else if (e < sa + q.rb * q.cb)
                tm[base_b + ((e - sa) / q.cb) * stride_b + (e - sa) % q.cb] = pt_[i];
        }
#pragma unroll
        for (int i = 0; i < kColumnPerThread; ++i) {
            const int e = threadIdx.x + 256 * i;
            if (e < count * 2 * w) {
                const int j = e / (2 * w), r = e % (2 * w), sgn = r / w;
                ys[j][sgn][r % w] = py[i];
            }
        }
        __syncthreads();
        Pair next = q;
        const bool more = advance(next, splits);
        if (more)
            fetch(next);
        // thread = (row of one of the two blocks, sign, group of 4 links): one T load
        // feeds 4 accumulators
        const int groups = (count + 3) / 4, rows2 = 2 * (q.ra + q.rb);
        for (int e = threadIdx.x; e < groups * rows2; e += blockDim.x) {
            const int g = e / rows2, r = e % rows2, j0 = 4 * g;
            const int half = r < 2 * q.ra ? 0 : 1;
            const int rr = half ? r - 2 * q.ra : r, rows = half ? q.rb : q.ra;
            const int sgn = rr / rows, m = half ? q.mb : q.ma, l = m + rr % rows;
            if (sgn && !m)
                continue;
            const int cols = half ? q.cb : q.ca, c0 = half ? q.ca : 0;
            const CE *trow = tm + (half ? base_b + (l - m) * stride_b : (l - m) * stride_a);
            CE s0{0, 0}, s1{0, 0}, s2{0, 0}, s3{0, 0};
            // two columns per pass: their 10 LDS reads are issued before the products (two,
            // not four, keeps the order-32 kernel within the 128 registers of four waves
            // per SIMD, which its LDS allows; both measured alike on the horn apply)
            constexpr int kSteps = 2;
            int c = 0;
            for (; c + kSteps <= cols; c += kSteps) {
                CE tv[kSteps], yv[kSteps][4];
#pragma unroll
                for (int u = 0; u < kSteps; ++u) {
                    tv[u] = trow[c + u];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        yv[u][i] = ys[j0 + i][sgn][c0 + c + u];
                }
#pragma unroll
                for (int u = 0; u < kSteps; ++u) {
                    s0 = s0 + tv[u] * yv[u][0];
                    s1 = s1 + tv[u] * yv[u][1];
                    s2 = s2 + tv[u] * yv[u][2];
                    s3 = s3 + tv[u] * yv[u][3];
                }
            }
            for (; c < cols; ++c) {
                const CE tv = trow[c];
                s0 = s0 + tv * ys[j0][sgn][c0 + c];
                s1 = s1 + tv * ys[j0 + 1][sgn][c0 + c];
                s2 = s2 + tv * ys[j0 + 2][sgn][c0 + c];
                s3 = s3 + tv * ys[j0 + 3][sgn][c0 + c];
            }
            const CE sv[4] = {s0, s1, s2, s3};
            const std::size_t idx = static_cast<std::size_t>(l) * l + l + (sgn ? -m : m);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (j0 + i < count)
                    z[slot[j0 + i] + idx] = sv[i];
        }
        if (!more)
            break;
        q = next;
    }
}

// Orders above kMaxFmmOrder (up to kMaxWideOrder): the same two passes with the
// Wigner block staged in column (forward) or row (back) tiles of kWideTile and no
// register prefetch; one workgroup per chunk holds ~100 KB of LDS.
constexpr int kWideN = 2 * kMaxWideOrder + 1;
constexpr int kWideTile = kMaxWideOrder + 1;

__global__ __launch_bounds__(256) void rotate_links_wide(FmmTranslateLaunch L, bool back) {
    __shared__ double dt[kWideN * kWideTile];
    __shared__ C2 xs[kFmmChunk][kWideN];
    __shared__ C2 ph[kWideN];
    __shared__ std::uint64_t src_off[kFmmChunk], dst_off[kFmmChunk];
    __shared__ int ord[kFmmChunk];
    const FmmChunk ch = L.chunks[blockIdx.x];
    const FmmRotation rot = L.rotations[ch.rotation];
    const C2 *az = reinterpret_cast<const C2 *>(L.azimuth) + rot.azimuth;
    for (int m = threadIdx.x; m <= 2 * static_cast<int>(rot.order); m += blockDim.x)
        ph[m] = az[m];
    if (threadIdx.x < ch.count) {
        const FmmLink lk = L.links[L.order[ch.first + threadIdx.x]];
        src_off[threadIdx.x] = back ? lk.slot : lk.source;
        dst_off[threadIdx.x] = lk.slot;
        ord[threadIdx.x] = static_cast<int>(back ? lk.target_order : lk.source_order);
    }
    __syncthreads();
    const C2 *in = reinterpret_cast<const C2 *>(back ? L.work_z : L.source_expansions);
    C2 *out = reinterpret_cast<C2 *>(L.work_y);
    const int count = static_cast<int>(ch.count), P = static_cast<int>(rot.order), top = static_cast<int>(ch.max_order);
    const int groups = (count + 3) / 4;
    for (int l = 0; l <= top; ++l) {
        const int n = 2 * l + 1;
        const std::size_t base = static_cast<std::size_t>(l) * l;
        const double *dl = L.small_d + rot.small_d + static_cast<std::size_t>(l) * (2 * l - 1) * (2 * l + 1) / 3;
        __syncthreads();
        for (int e = threadIdx.x; e < count * n; e += blockDim.x) {
            const int j = e / n, q = e % n;
            C2 v{0, 0};
            if (l <= ord[j]) {
                v = in[src_off[j] + base + q];
                if (!back)
                    v = ph[P + q - l] * v;
            }
            xs[j][q] = v;
        }
        for (int t0 = 0; t0 < n; t0 += kWideTile) {
            const int tw = min(kWideTile, n - t0);
            __syncthreads();
            // forward: dt[q][t - t0] = d[q][t]; back: dt[t - t0][q] = d[t][q]
            for (int e = threadIdx.x; e < n * tw; e += blockDim.x) {
                if (back) {
                    const int t = e / n, q = e % n;
                    dt[t * n + q] = dl[static_cast<std::size_t>(t0 + t) * n + q];
                } else {
                    const int q = e / tw, t = e % tw;
                    dt[q * tw + t] = dl[static_cast<std::size_t>(q) * n + t0 + t];
                }
            }
            __syncthreads();
            for (int e = threadIdx.x; e < groups * tw; e += blockDim.x) {
                const int g = e / tw, tt = e % tw, t = t0 + tt, j0 = 4 * g;
                C2 s0{0, 0}, s1{0, 0}, s2{0, 0}, s3{0, 0};
                const int stride = back ? 1 : tw;
                const double *drow = back ? dt + tt * n : dt + tt;
                for (int q = 0; q < n; ++q) {
                    const double d = drow[q * stride];
                    s0 = s0 + d * xs[j0][q];
                    s1 = s1 + d * xs[min(j0 + 1, kFmmChunk - 1)][q];
                    s2 = s2 + d * xs[min(j0 + 2, kFmmChunk - 1)][q];
                    s3 = s3 + d * xs[min(j0 + 3, kFmmChunk - 1)][q];
                }
                const C2 sv[4] = {s0, s1, s2, s3};
                const C2 rot_back = back ? conj(ph[P + t - l]) : C2{1, 0};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int j = j0 + i;
                    if (j < count && l <= ord[j])
                        out[dst_off[j] + (back ? base + t : m_major(ord[j], l, t - l))] = back ? rot_back * sv[i] : sv[i];
                }
            }
        }
    }
}

constexpr int kWideRows = kMaxWideOrder + 1;

__global__ __launch_bounds__(256) void coaxial_links_wide(FmmTranslateLaunch L) {
    __shared__ C2 tm[kWideRows * kWideRows];
    __shared__ C2 ys[kFmmChunk][2][kWideRows];
    __shared__ std::uint64_t slot[kFmmChunk];
    const FmmChunk ch = L.axial_chunks[blockIdx.x];
    const FmmLink first = L.links[L.axial_order[ch.first]];
    const int ps = static_cast<int>(first.source_order), pt = static_cast<int>(first.target_order), mm = min(ps, pt);
    const C2 *T = reinterpret_cast<const C2 *>(L.axial) + L.axial_offsets[first.axial];
    const C2 *y = reinterpret_cast<const C2 *>(L.work_y);
    C2 *z = reinterpret_cast<C2 *>(L.work_z);
    const int count = static_cast<int>(ch.count);
    if (threadIdx.x < count)
        slot[threadIdx.x] = L.links[L.axial_order[ch.first + threadIdx.x]].slot;
    const int nct = (pt + 1) * (pt + 1);
    __syncthreads();
    for (int e = threadIdx.x; e < count * nct; e += blockDim.x) {
        const int j = e / nct, c = e % nct;
        int l, t;
        split_index(c, l, t);
        if (abs(t - l) > mm)
            z[slot[j] + c] = {0, 0};
    }
    std::size_t off = 0;
    const int groups = (count + 3) / 4;
    for (int m = 0; m <= mm; ++m) {
        const int rows = pt - m + 1, cols = ps - m + 1;
        __syncthreads();
        const int stride = cols | 1;
        for (int e = threadIdx.x; e < rows * cols; e += blockDim.x)
            tm[(e / cols) * stride + e % cols] = T[off + e];
        for (int e = threadIdx.x; e < kFmmChunk * 2 * cols; e += blockDim.x) {
            const int j = e / (2 * cols), r = e % (2 * cols), sgn = r / cols, n = m + r % cols;
            ys[j][sgn][r % cols] = j < count ? y[slot[j] + m_major(ps, n, sgn ? -m : m)]
                                             : C2{0, 0};
        }
        __syncthreads();
        off += static_cast<std::size_t>(rows) * cols;
        for (int e = threadIdx.x; e < groups * 2 * rows; e += blockDim.x) {
            const int g = e / (2 * rows), r = e % (2 * rows), sgn = r / rows, l = m + r % rows, j0 = 4 * g;
            if (sgn && !m)
                continue;
            C2 s0{0, 0}, s1{0, 0}, s2{0, 0}, s3{0, 0};
            const C2 *trow = tm + (l - m) * stride;
            for (int c = 0; c < cols; ++c) {
                const C2 tv = trow[c];
                s0 = s0 + tv * ys[j0][sgn][c];
                s1 = s1 + tv * ys[j0 + 1][sgn][c];
                s2 = s2 + tv * ys[j0 + 2][sgn][c];
                s3 = s3 + tv * ys[j0 + 3][sgn][c];
            }
            const C2 sv[4] = {s0, s1, s2, s3};
            const std::size_t idx = static_cast<std::size_t>(l) * l + l + (sgn ? -m : m);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (j0 + i < count)
                    z[slot[j0 + i] + idx] = sv[i];
        }
    }
}

// Orders above kMaxWideOrder (up to kMaxBigOrder): no staging of a whole Wigner or
// coaxial block fits in LDS, so the reduction index is tiled. Rotations: per degree,
// all of the chunk's inputs (one half of the chunk, 8 links, at a time) stay in LDS
// while 32-row slabs of d_l pass through; every thread keeps the running sums of
// its (t, 4-link group) items in registers. Coaxial products: per |m| block the
// chunk's columns stay in LDS while 16-column slabs of T_m pass through.
constexpr int kBigN = 2 * kMaxBigOrder + 1;
constexpr int kBigHalf = 8; // links per pass (a chunk of kFmmChunk links takes kFmmChunk / 8 passes)
constexpr int kBigSlab = 32;
constexpr int kBigItems = (kBigN * (kBigHalf / 4) + 255) / 256; // (t, group) items per thread

__global__ __launch_bounds__(256) void rotate_links_big(FmmTranslateLaunch L, bool back) {
    __shared__ C2 xs[kBigHalf][kBigN];
    __shared__ double dt[kBigSlab * kBigN];
    __shared__ C2 ph[kBigN];
    __shared__ std::uint64_t src_off[kFmmChunk], dst_off[kFmmChunk];
    __shared__ int ord[kFmmChunk];
    const FmmChunk ch = L.chunks[blockIdx.x];
    const FmmRotation rot = L.rotations[ch.rotation];
    const C2 *az = reinterpret_cast<const C2 *>(L.azimuth) + rot.azimuth;
    for (int m = threadIdx.x; m <= 2 * static_cast<int>(rot.order); m += blockDim.x)
        ph[m] = az[m];
    if (threadIdx.x < ch.count) {
        const FmmLink lk = L.links[L.order[ch.first + threadIdx.x]];
        src_off[threadIdx.x] = back ? lk.slot : lk.source;
        dst_off[threadIdx.x] = lk.slot;
        ord[threadIdx.x] = static_cast<int>(back ? lk.target_order : lk.source_order);
    }
    __syncthreads();
    const C2 *in = reinterpret_cast<const C2 *>(back ? L.work_z : L.source_expansions);
    C2 *out = reinterpret_cast<C2 *>(L.work_y);
    const int count = static_cast<int>(ch.count), P = static_cast<int>(rot.order), top = static_cast<int>(ch.max_order);
    for (int h = 0; h * kBigHalf < count; ++h) {
        const int j_first = h * kBigHalf, hc = min(kBigHalf, count - j_first);
        for (int l = 0; l <= top; ++l) {
            const int n = 2 * l + 1;
            const std::size_t base = static_cast<std::size_t>(l) * l;
            const double *dl = L.small_d + rot.small_d + static_cast<std::size_t>(l) * (2 * l - 1) * (2 * l + 1) / 3;
            __syncthreads();
            for (int e = threadIdx.x; e < kBigHalf * n; e += blockDim.x) {
                const int jj = e / n, q = e % n, j = j_first + jj;
                C2 v{0, 0};
                if (jj < hc && l <= ord[j]) {
                    v = in[src_off[j] + base + q];
                    if (!back)
                        v = ph[P + q - l] * v;
                }
                xs[jj][q] = v;
            }
            C2 acc[kBigItems][4];
#pragma unroll
            for (int it = 0; it < kBigItems; ++it)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    acc[it][i] = {0, 0};
            const int items = n * (kBigHalf / 4);
            for (int q0 = 0; q0 < n; q0 += kBigSlab) {
                const int qw = min(kBigSlab, n - q0);
                __syncthreads();
                // dt[qq][t] = d[q][t] (forward) or d[t][q] (back), q = q0 + qq
                for (int e = threadIdx.x; e < qw * n; e += blockDim.x) {
                    if (back) {
                        const int t = e / qw, qq = e % qw;
                        dt[qq * n + t] = dl[static_cast<std::size_t>(t) * n + q0 + qq];
                    } else {
                        const int qq = e / n, t = e % n;
                        dt[qq * n + t] = dl[static_cast<std::size_t>(q0 + qq) * n + t];
                    }
                }
                __syncthreads();
#pragma unroll
                for (int it = 0; it < kBigItems; ++it) {
                    const int e = threadIdx.x + 256 * it;
                    if (e >= items)
                        continue;
                    const int g = e / n, t = e % n, j0 = 4 * g;
                    for (int qq = 0; qq < qw; ++qq) {
                        const double d = dt[qq * n + t];
                        const int q = q0 + qq;
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            acc[it][i] = acc[it][i] + d * xs[j0 + i][q];
                    }
                }
            }
#pragma unroll
            for (int it = 0; it < kBigItems; ++it) {
                const int e = threadIdx.x + 256 * it;
                if (e >= items)
                    continue;
                const int g = e / n, t = e % n;
                const C2 rot_back = back ? conj(ph[P + t - l]) : C2{1, 0};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int jj = 4 * g + i, j = j_first + jj;
                    if (jj < hc && l <= ord[j])
                        out[dst_off[j] + (back ? base + t : m_major(ord[j], l, t - l))] =
                            back ? rot_back * acc[it][i] : acc[it][i];
                }
            }
        }
    }
}

constexpr int kBigRows = kMaxBigOrder + 1;
constexpr int kBigCols = 16;
constexpr int kBigCoaxItems = (kBigRows * 2 * (kFmmChunk / 4) + 255) / 256; // (row, sign, group) per thread

__global__ __launch_bounds__(256) void coaxial_links_big(FmmTranslateLaunch L) {
    __shared__ C2 ys[kFmmChunk][2][kBigCols]; // the current column tile of every link's +m / -m sources
    __shared__ C2 tm[kBigRows * kBigCols];
    __shared__ std::uint64_t slot[kFmmChunk];
    const FmmChunk ch = L.axial_chunks[blockIdx.x];
    const FmmLink first = L.links[L.axial_order[ch.first]];
    const int ps = static_cast<int>(first.source_order), pt = static_cast<int>(first.target_order), mm = min(ps, pt);
    const C2 *T = reinterpret_cast<const C2 *>(L.axial) + L.axial_offsets[first.axial];
    const C2 *y = reinterpret_cast<const C2 *>(L.work_y);
    C2 *z = reinterpret_cast<C2 *>(L.work_z);
    const int count = static_cast<int>(ch.count);
    if (threadIdx.x < count)
        slot[threadIdx.x] = L.links[L.axial_order[ch.first + threadIdx.x]].slot;
    const int nct = (pt + 1) * (pt + 1);
    __syncthreads();
    for (int e = threadIdx.x; e < count * nct; e += blockDim.x) {
        const int j = e / nct, c = e % nct;
        int l, t;
        split_index(c, l, t);
        if (abs(t - l) > mm)
            z[slot[j] + c] = {0, 0};
    }
    std::size_t off = 0;
    for (int m = 0; m <= mm; ++m) {
        const int rows = pt - m + 1, cols = ps - m + 1;
        C2 acc[kBigCoaxItems][4];
#pragma unroll
        for (int it = 0; it < kBigCoaxItems; ++it)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                acc[it][i] = {0, 0};
        const int items = rows * 2 * (kFmmChunk / 4);
        for (int c0 = 0; c0 < cols; c0 += kBigCols) {
            const int cw = min(kBigCols, cols - c0);
            __syncthreads();
            for (int e = threadIdx.x; e < rows * cw; e += blockDim.x) {
                const int r = e / cw, cc = e % cw;
                tm[r * kBigCols + cc] = T[off + static_cast<std::size_t>(r) * cols + c0 + cc];
            }
            for (int e = threadIdx.x; e < kFmmChunk * 2 * cw; e += blockDim.x) {
                const int j = e / (2 * cw), r = e % (2 * cw), sgn = r / cw, cc = r % cw, n = m + c0 + cc;
                ys[j][sgn][cc] = j < count ? y[slot[j] + m_major(ps, n, sgn ? -m : m)] : C2{0, 0};
            }
            __syncthreads();
#pragma unroll
            for (int it = 0; it < kBigCoaxItems; ++it) {
                const int e = threadIdx.x + 256 * it;
                if (e >= items)
                    continue;
                const int g = e / (2 * rows), r = e % (2 * rows), sgn = r / rows, row = r % rows, j0 = 4 * g;
                if (sgn && !m)
                    continue;
                for (int cc = 0; cc < cw; ++cc) {
                    const C2 tv = tm[row * kBigCols + cc];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        acc[it][i] = acc[it][i] + tv * ys[j0 + i][sgn][cc];
                }
            }
        }
#pragma unroll
        for (int it = 0; it < kBigCoaxItems; ++it) {
            const int e = threadIdx.x + 256 * it;
            if (e >= items)
                continue;
            const int g = e / (2 * rows), r = e % (2 * rows), sgn = r / rows, l = m + r % rows, j0 = 4 * g;
            if (sgn && !m)
                continue;
            const std::size_t idx = static_cast<std::size_t>(l) * l + l + (sgn ? -m : m);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (j0 + i < count)
                    z[slot[j0 + i] + idx] = acc[it][i];
        }
        off += static_cast<std::size_t>(rows) * cols;
    }
}

// target = sum of the destination's link results.
template <class E> __global__ __launch_bounds__(256) void gather_links(FmmTranslateLaunch L) {
    using CE = Cx<E>;
    const FmmDestination dst = L.destinations[blockIdx.x];
    const int nct = static_cast<int>((dst.order + 1) * (dst.order + 1));
    const CE *w = reinterpret_cast<const CE *>(L.work_y);
    CE *out = reinterpret_cast<CE *>(L.target_expansions) + dst.target;
    const FmmLink *links = L.links + dst.first_link;
    const std::uint32_t n = dst.link_count;
    for (int e = threadIdx.x; e < nct; e += blockDim.x) {
        // the sum runs in link order; the loads of four links are issued together
        CE s{0, 0};
        std::uint32_t i = 0;
        for (; i + 4 <= n; i += 4) {
            const CE a = w[links[i].slot + e], b = w[links[i + 1].slot + e];
            const CE c = w[links[i + 2].slot + e], d = w[links[i + 3].slot + e];
            s = s + a;
            s = s + b;
            s = s + c;
            s = s + d;
        }
        for (; i < n; ++i)
            s = s + w[links[i].slot + e];
        out[e] = s;
    }
}

// Block-Jacobi blocks of the near operator: block b holds rows and columns
// order[b*W .. b*W+W) (tree order); entries absent from the CSR are zero.
__global__ void gather_blocks(const std::uint32_t *row_offsets, const std::uint32_t *columns, const double *values,
                              const std::uint32_t *order, const std::uint32_t *position, std::uint32_t n,
                              std::uint32_t width, double *blocks) {
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const std::uint32_t W = width, row = order[i], blk = i / W, r = i % W;
    double *B = blocks + static_cast<std::size_t>(blk) * W * W * 2;
    for (std::uint32_t e = row_offsets[row]; e < row_offsets[row + 1]; ++e) {
        const std::uint32_t pc = position[columns[e]];
        if (pc / W != blk)
            continue;
        const std::uint32_t c = pc % W;
        B[2 * (static_cast<std::size_t>(c) * W + r)] = values[2 * e];
        B[2 * (static_cast<std::size_t>(c) * W + r) + 1] = values[2 * e + 1];
    }
}

// Far part of B q for the FMM path, by direct pair quadrature (the flux panels
// are the few driven ones): for test panel s and flux panel t, every corner i of
// s whose pressure leaf is not a near leaf pair of t's flux leaf receives
//   (1/m_i) q_t sum_a sum_b w_a w_b phi_i(x_a) (-G - alpha dG/dn_x)(x_a, y_b).
__global__ __launch_bounds__(256) void far_flux(FmmFarFluxLaunch L) {
    const std::uint64_t pair = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (pair >= static_cast<std::uint64_t>(L.nt) * L.flux_count)
        return;
    const std::uint32_t s = static_cast<std::uint32_t>(pair / L.flux_count);
    const std::uint32_t t = L.flux_panels[pair % L.flux_count];
    const std::uint32_t fleaf = L.flux_leaf[t];
    bool far[3];
    std::uint32_t node[3];
    bool any = false;
    for (int c = 0; c < 3; ++c) {
        node[c] = L.nodes[static_cast<std::size_t>(c) * L.nt + s];
        const std::uint32_t pl = L.pressure_leaf[node[c]];
        bool near = false;
        for (std::uint32_t e = L.near_offsets[pl]; e < L.near_offsets[pl + 1]; ++e)
            near |= L.near_sources[e] == fleaf;
        far[c] = !near;
        any |= far[c];
    }
    if (!any)
        return;
    const double *g = L.geo;
    const std::size_t nt = L.nt;
    double vs[3], e1s[3], e2s[3], ns[3], vt[3], e1t[3], e2t[3];
    for (int i = 0; i < 3; ++i) {
        vs[i] = g[(G_V0 + i) * nt + s];
        e1s[i] = g[(G_E1 + i) * nt + s];
        e2s[i] = g[(G_E2 + i) * nt + s];
        ns[i] = g[(G_N + i) * nt + s];
        vt[i] = g[(G_V0 + i) * nt + t];
        e1t[i] = g[(G_E1 + i) * nt + t];
        e2t[i] = g[(G_E2 + i) * nt + t];
    }
    const double ws = 2 * g[G_AREA * nt + s], wt = 2 * g[G_AREA * nt + t], k = L.k;
    const C2 alpha{0, -1 / k};
    C2 acc[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (std::uint32_t a = 0; a < L.rule_points; ++a) {
        const double *ra = L.rule + 4 * a;
        double x[3];
        for (int i = 0; i < 3; ++i)
            x[i] = vs[i] + ra[1] * e1s[i] + ra[2] * e2s[i];
        C2 kern{0, 0};
        for (std::uint32_t b = 0; b < L.rule_points; ++b) {
            const double *rb = L.rule + 4 * b;
            double d[3], r2 = 0, nd = 0;
            for (int i = 0; i < 3; ++i) {
                d[i] = x[i] - (vt[i] + rb[1] * e1t[i] + rb[2] * e2t[i]);
                r2 += d[i] * d[i];
                nd += ns[i] * d[i];
            }
            const double r = sqrt(r2);
            double sn, cs;
            sincos(k * r, &sn, &cs);
            const C2 G = (1.0 / (4 * 3.141592653589793 * r)) * C2{cs, -sn};
            const C2 dG = (nd / r) * (C2{-1 / r, -k} * G);
            kern = kern + (rb[3] * wt) * (C2{-G.re, -G.im} - alpha * dG);
        }
        for (int c = 0; c < 3; ++c)
            acc[c] = acc[c] + (ra[3] * ws * ra[c]) * kern;
    }
    for (std::uint32_t e = 0; e < L.excitations; ++e) {
        const C2 q{L.q[2 * (static_cast<std::size_t>(e) * L.nt_all + t)], L.q[2 * (static_cast<std::size_t>(e) * L.nt_all + t) + 1]};
        if (q.re == 0 && q.im == 0)
            continue;
        for (int c = 0; c < 3; ++c)
            if (far[c]) {
                const C2 v = L.row_scale[node[c]] * (q * acc[c]);
                atomic_add_c(L.out + 2 * (static_cast<std::size_t>(e) * L.nv + node[c]), v);
            }
    }
}

// Direct far action of the Burton-Miller pressure operator. For a test panel s and a
// trial panel t the far corner pairs (ci, cj) contribute
//   (1/m_i) x_j int_s int_t phi_i(x) phi_j(y) [-dG/dn_y - alpha d2G/dn_x dn_y] dy dx,
// G = exp(-ikr)/(4 pi r), d = x - y, g' = -(ik + 1/r) G, g'' = G/r^2 + (ik + 1/r)^2 G:
//   dG/dn_y = -g' (d.n_y)/r,
//   d2G/dn_x dn_y = -(d.n_x)(d.n_y)(g''/r^2 - g'/r^3) - (g'/r)(n_x.n_y).
// Block per test panel, threads over trial panels; per-thread row sums reduced in
// the block, one atomic add per test corner.
__global__ __launch_bounds__(256) void direct_far(FmmDirectLaunch L) {
    __shared__ double red[2][3][4];
    const std::uint32_t s = blockIdx.x;
    const std::size_t nt = L.nt;
    const double *g = L.geo;
    double vs[3], e1s[3], e2s[3], ns[3];
    std::uint32_t node_s[3], leaf_s[3];
    for (int i = 0; i < 3; ++i) {
        vs[i] = g[(G_V0 + i) * nt + s];
        e1s[i] = g[(G_E1 + i) * nt + s];
        e2s[i] = g[(G_E2 + i) * nt + s];
        ns[i] = g[(G_N + i) * nt + s];
        node_s[i] = L.nodes[static_cast<std::size_t>(i) * nt + s];
        leaf_s[i] = L.vertex_leaf[node_s[i]];
    }
    const double ws = 2 * g[G_AREA * nt + s], k = L.k;
    const int Q = static_cast<int>(L.rule_points);
    const C2 alpha{0, -1 / k};
    C2 acc[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (std::uint32_t t = threadIdx.x; t < L.nt; t += blockDim.x) {
        std::uint32_t node_t[3];
        bool far[3][3], any = false;
        for (int c = 0; c < 3; ++c)
            node_t[c] = L.nodes[static_cast<std::size_t>(c) * nt + t];
        for (int ci = 0; ci < 3; ++ci)
            for (int cj = 0; cj < 3; ++cj) {
                far[ci][cj] = !L.near[static_cast<std::size_t>(leaf_s[ci]) * L.leaf_count + L.vertex_leaf[node_t[cj]]];
                any |= far[ci][cj];
            }
        if (!any)
            continue;
        double vt[3], e1t[3], e2t[3], nn[3];
        for (int i = 0; i < 3; ++i) {
            vt[i] = g[(G_V0 + i) * nt + t];
            e1t[i] = g[(G_E1 + i) * nt + t];
            e2t[i] = g[(G_E2 + i) * nt + t];
            nn[i] = g[(G_N + i) * nt + t];
        }
        const double wt = 2 * g[G_AREA * nt + t];
        const double nxny = ns[0] * nn[0] + ns[1] * nn[1] + ns[2] * nn[2];
        C2 M[3][3] = {};
        for (int a = 0; a < Q; ++a) {
            const double *ra = L.rule + 4 * a;
            double x[3];
            for (int i = 0; i < 3; ++i)
                x[i] = vs[i] + ra[1] * e1s[i] + ra[2] * e2s[i];
            C2 T[3] = {{0, 0}, {0, 0}, {0, 0}};
            for (int b = 0; b < Q; ++b) {
                const double *rb = L.rule + 4 * b;
                double d[3], r2 = 0, dnx = 0, dny = 0;
                for (int i = 0; i < 3; ++i) {
                    d[i] = x[i] - (vt[i] + rb[1] * e1t[i] + rb[2] * e2t[i]);
                    r2 += d[i] * d[i];
                    dnx += ns[i] * d[i];
                    dny += nn[i] * d[i];
                }
                const double r = sqrt(r2), ir = 1 / r;
                double sn, cs;
                sincos(k * r, &sn, &cs);
                const C2 G = (ir / (4 * 3.141592653589793)) * C2{cs, -sn};
                const C2 ikr{ir, k};                       // ik + 1/r
                const C2 g1 = C2{-1, 0} * (ikr * G);        // g'
                const C2 g2 = (ir * ir) * G + (ikr * ikr) * G; // g''
                const C2 dGy = (-dny * ir) * g1;
                const C2 H = (-dnx * dny) * ((ir * ir) * g2 - (ir * ir * ir) * g1) - (nxny * ir) * g1;
                const C2 K = (rb[3] * wt) * (C2{-dGy.re, -dGy.im} - alpha * H);
                for (int c = 0; c < 3; ++c)
                    T[c] = T[c] + rb[c] * K;
            }
            for (int ci = 0; ci < 3; ++ci)
                for (int cj = 0; cj < 3; ++cj)
                    M[ci][cj] = M[ci][cj] + (ra[3] * ws * ra[ci]) * T[cj];
        }
        for (int cj = 0; cj < 3; ++cj) {
            const C2 xj{L.x[2 * node_t[cj]], L.x[2 * node_t[cj] + 1]};
            for (int ci = 0; ci < 3; ++ci)
                if (far[ci][cj])
                    acc[ci] = acc[ci] + M[ci][cj] * xj;
        }
    }
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    for (int c = 0; c < 3; ++c) {
        double re = acc[c].re, im = acc[c].im;
        for (int off = 32; off > 0; off >>= 1) {
            re += __shfl_xor(re, off, 64);
            im += __shfl_xor(im, off, 64);
        }
        if (lane == 0) {
            red[0][c][wave] = re;
            red[1][c][wave] = im;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int c = threadIdx.x;
        double re = 0, im = 0;
        for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) {
            re += red[0][c][w];
            im += red[1][c][w];
        }
        const double sc = L.row_scale[node_s[c]];
        atomic_add_c(L.y + 2 * static_cast<std::size_t>(node_s[c]), C2{sc * re, sc * im});
    }
}

// out[:, e] += B_near q_e with B_near the P1 x face-column near CSR.
__global__ void flux_csr_apply(const std::uint32_t *offsets, const std::uint32_t *columns, const double *values,
                               std::uint32_t rows, const double *q, std::uint32_t faces, std::uint32_t excitations,
                               double *out) {
    const std::uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows)
        return;
    for (std::uint32_t e = 0; e < excitations; ++e) {
        C2 s{0, 0};
        for (std::uint32_t j = offsets[r]; j < offsets[r + 1]; ++j) {
            const C2 v{values[2 * j], values[2 * j + 1]};
            const std::size_t f = static_cast<std::size_t>(e) * faces + columns[j];
            s = s + v * C2{q[2 * f], q[2 * f + 1]};
        }
        out[2 * (static_cast<std::size_t>(e) * rows + r)] += s.re;
        out[2 * (static_cast<std::size_t>(e) * rows + r) + 1] += s.im;
    }
}

// out[map[i]] = in[i] (forward) or out[i] = in[map[i]] (backward), per vector.
__global__ void permute_vectors(const double *in, double *out, const std::uint32_t *map, std::uint32_t n,
                                std::uint32_t count, bool forward) {
    const std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<std::uint64_t>(n) * count)
        return;
    const std::uint64_t v = i / n, j = i % n, base = v * n;
    const std::uint64_t src = base + (forward ? j : map[j]), dst = base + (forward ? map[j] : j);
    out[2 * dst] = in[2 * src];
    out[2 * dst + 1] = in[2 * src + 1];
}

} // namespace

void launch_leaf_moments(const FmmLeafMomentsLaunch &launch, hipStream_t stream) {
    if (!launch.item_count)
        return;
    // register table of the radial factors up to the leaves' largest order
    const bool wide = launch.max_order > static_cast<std::uint32_t>(kMaxLeafOrder);
    if (launch.rule_points <= static_cast<std::uint32_t>(kGroupSmall)) {
        const std::uint32_t per_block = 64 / kGroupSmall;
        const dim3 grid((launch.item_count + per_block - 1) / per_block);
        if (wide)
            hipLaunchKernelGGL((leaf_moments<kGroupSmall, kMaxBigOrder>), grid, dim3(64), 0, stream, launch);
        else
            hipLaunchKernelGGL((leaf_moments<kGroupSmall, kMaxLeafOrder>), grid, dim3(64), 0, stream, launch);
    } else if (wide) {
        hipLaunchKernelGGL((leaf_moments<kGroupLarge, kMaxBigOrder>), dim3(launch.item_count), dim3(64), 0, stream,
                           launch);
    } else {
        hipLaunchKernelGGL((leaf_moments<kGroupLarge, kMaxLeafOrder>), dim3(launch.item_count), dim3(64), 0, stream,
                           launch);
    }
}

void launch_leaf_emit(const FmmLeaf *leaves, std::uint32_t leaf_count, const std::uint32_t *leaf_dofs,
                      const double *emission, const double *p, double *expansions, hipStream_t stream, bool fp32) {
    if (!leaf_count)
        return;
    if (fp32)
        hipLaunchKernelGGL(leaf_emit<float>, dim3(leaf_count), dim3(256), 0, stream, leaves, leaf_dofs, emission, p,
                           reinterpret_cast<float *>(expansions));
    else
        hipLaunchKernelGGL(leaf_emit<double>, dim3(leaf_count), dim3(256), 0, stream, leaves, leaf_dofs, emission, p,
                           expansions);
}

void launch_leaf_test(const FmmLeaf *leaves, std::uint32_t leaf_count, const std::uint32_t *leaf_dofs,
                      const double *testing, const double *expansions, double *y, hipStream_t stream, bool fp32) {
    if (!leaf_count)
        return;
    if (fp32)
        hipLaunchKernelGGL(leaf_test<float>, dim3(leaf_count), dim3(256), 0, stream, leaves, leaf_dofs, testing,
                           reinterpret_cast<const float *>(expansions), y);
    else
        hipLaunchKernelGGL(leaf_test<double>, dim3(leaf_count), dim3(256), 0, stream, leaves, leaf_dofs, testing,
                           expansions, y);
}

// The order-templated translations of one launch (orders up to kMidFmmOrder) in scalar E.
template <class E>
void launch_links(const FmmTranslateLaunch &launch, std::uint32_t max_order, dim3 rot_grid, dim3 axial_grid,
                  hipStream_t stream) {
    if (max_order > kMaxFmmOrder) {
        hipLaunchKernelGGL((rotate_links<kMidFmmOrder, E>), rot_grid, dim3(256), 0, stream, launch, false);
        hipLaunchKernelGGL((coaxial_links<kMidFmmOrder, E>), axial_grid, dim3(256), 0, stream, launch);
        hipLaunchKernelGGL((rotate_links<kMidFmmOrder, E>), rot_grid, dim3(256), 0, stream, launch, true);
    } else if (max_order > kSmallFmmOrder) {
        hipLaunchKernelGGL((rotate_links<kMaxFmmOrder, E>), rot_grid, dim3(256), 0, stream, launch, false);
        hipLaunchKernelGGL((coaxial_links<kMaxFmmOrder, E>), axial_grid, dim3(256), 0, stream, launch);
        hipLaunchKernelGGL((rotate_links<kMaxFmmOrder, E>), rot_grid, dim3(256), 0, stream, launch, true);
    } else {
        hipLaunchKernelGGL((rotate_links<kSmallFmmOrder, E>), rot_grid, dim3(256), 0, stream, launch, false);
        hipLaunchKernelGGL((coaxial_links<kSmallFmmOrder, E>), axial_grid, dim3(256), 0, stream, launch);
        hipLaunchKernelGGL((rotate_links<kSmallFmmOrder, E>), rot_grid, dim3(256), 0, stream, launch, true);
    }
}

void launch_translate(const FmmTranslateLaunch &launch, std::uint32_t destination_count, std::uint32_t link_count,
                      std::uint32_t chunk_count, std::uint32_t axial_chunk_count, std::uint32_t max_order,
                      std::uint32_t split_target, hipStream_t stream) {
    if (!destination_count)
        return;
    // Levels with few chunks spread each chunk's stages over up to `splits` workgroups,
    // so that a launch holds about split_target workgroups (a rotation chunk has
    // max_order / 2 + 1 stages, a coaxial chunk at most as many); 0 = no split.
    auto splits = [&](std::uint32_t chunks) {
        const std::uint32_t stages = max_order / 2 + 1;
        return chunks && split_target ? std::max(1u, std::min(stages, (split_target + chunks - 1) / chunks)) : 1u;
    };
    const dim3 rot_grid(chunk_count, splits(chunk_count)), axial_grid(axial_chunk_count, splits(axial_chunk_count));
    if (launch.fp32 && link_count && max_order > kMidFmmOrder)
        throw std::invalid_argument("FP32 translations take orders up to 48 only");
    if (link_count && max_order > kMaxWideOrder) {
        hipLaunchKernelGGL(rotate_links_big, dim3(chunk_count), dim3(256), 0, stream, launch, false);
        hipLaunchKernelGGL(coaxial_links_big, dim3(axial_chunk_count), dim3(256), 0, stream, launch);
        hipLaunchKernelGGL(rotate_links_big, dim3(chunk_count), dim3(256), 0, stream, launch, true);
    } else if (link_count && max_order > kMidFmmOrder) {
        hipLaunchKernelGGL(rotate_links_wide, dim3(chunk_count), dim3(256), 0, stream, launch, false);
        hipLaunchKernelGGL(coaxial_links_wide, dim3(axial_chunk_count), dim3(256), 0, stream, launch);
        hipLaunchKernelGGL(rotate_links_wide, dim3(chunk_count), dim3(256), 0, stream, launch, true);
    } else if (link_count && launch.fp32) {
        launch_links<float>(launch, max_order, rot_grid, axial_grid, stream);
    } else if (link_count) {
        launch_links<double>(launch, max_order, rot_grid, axial_grid, stream);
    }
    if (launch.fp32)
        hipLaunchKernelGGL(gather_links<float>, dim3(destination_count), dim3(256), 0, stream, launch);
    else
        hipLaunchKernelGGL(gather_links<double>, dim3(destination_count), dim3(256), 0, stream, launch);
}

void launch_gather_blocks(const std::uint32_t *row_offsets, const std::uint32_t *columns, const double *values,
                          const std::uint32_t *order, const std::uint32_t *position, std::uint32_t n,
                          std::uint32_t width, double *blocks, hipStream_t stream) {
    hipLaunchKernelGGL(gather_blocks, dim3((n + 255) / 256), dim3(256), 0, stream, row_offsets, columns, values, order,
                       position, n, width, blocks);
}

void launch_direct_far(const FmmDirectLaunch &launch, hipStream_t stream) {
    if (launch.nt)
        hipLaunchKernelGGL(direct_far, dim3(launch.nt), dim3(256), 0, stream, launch);
}

void launch_far_flux(const FmmFarFluxLaunch &launch, hipStream_t stream) {
    const std::uint64_t pairs = static_cast<std::uint64_t>(launch.nt) * launch.flux_count;
    if (pairs)
        hipLaunchKernelGGL(far_flux, dim3(static_cast<std::uint32_t>((pairs + 255) / 256)), dim3(256), 0, stream, launch);
}

void launch_flux_csr_apply(const std::uint32_t *offsets, const std::uint32_t *columns, const double *values,
                           std::uint32_t rows, const double *q, std::uint32_t faces, std::uint32_t excitations,
                           double *out, hipStream_t stream) {
    if (rows && faces)
        hipLaunchKernelGGL(flux_csr_apply, dim3((rows + 255) / 256), dim3(256), 0, stream, offsets, columns, values,
                           rows, q, faces, excitations, out);
}

void launch_permute(const double *in, double *out, const std::uint32_t *map, std::uint32_t n, std::uint32_t count,
                    bool forward, hipStream_t stream) {
    const std::uint64_t total = static_cast<std::uint64_t>(n) * count;
    if (total)
        hipLaunchKernelGGL(permute_vectors, dim3(static_cast<std::uint32_t>((total + 255) / 256)), dim3(256), 0,
                           stream, in, out, map, n, count, forward);
}

} // namespace hsbem::dev
