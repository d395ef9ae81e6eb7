// Device-resident restart cycle of GMRES (one right-hand side).
//
// The Arnoldi step index j, the Hessenberg column, the Givens rotations, the
// rotated right-hand side g and the stop decision live in device memory, so one
// Arnoldi step is a fixed sequence of launches (current = v_j, the operator, CGS2
// against v_0..v_j, the norm, v_{j+1}, the rotation) whose arguments never change:
// the host captures it once into a HIP graph and replays it once per iteration,
// reading the state back only every few steps. Each kernel reads j from the state
// and returns at once when the cycle has stopped (target met, breakdown or the
// step limit), so replays queued past the stop cost only their launches.
//
// The arithmetic is that of the host-driven cycle: the same projection, subtraction
// and norm reductions in the same order, and the rotation formulas of the host
// Givens loop evaluated in FP64 on one lane.
#include "device.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace hsbem::dev {
namespace {

constexpr int kRows = 512; // rows per projection workgroup (as the host-driven cycle)

enum : int { S_J = 0, S_DONE = 1, S_USED = 2, S_LIMIT = 3, S_ERROR = 4 };

struct Layout {
    std::uint32_t stride; // restart + 1
    __host__ __device__ std::size_t h1() const { return 0; }
    __host__ __device__ std::size_t h2() const { return 2 * static_cast<std::size_t>(stride); }
    __host__ __device__ std::size_t nrm() const { return 4 * static_cast<std::size_t>(stride); }
    __host__ __device__ std::size_t target() const { return nrm() + 1; }
    __host__ __device__ std::size_t resid() const { return nrm() + 2; }
    __host__ __device__ std::size_t g() const { return nrm() + 4; }
    __host__ __device__ std::size_t cs() const { return g() + 2 * static_cast<std::size_t>(stride); }
    __host__ __device__ std::size_t sn() const { return cs() + stride; }
    __host__ __device__ std::size_t H() const { return sn() + 2 * static_cast<std::size_t>(stride); }
    __host__ __device__ std::size_t total() const { return H() + 2 * static_cast<std::size_t>(stride) * stride; }
};

__device__ __forceinline__ bool stopped(const int *st) { return st[S_DONE] != 0; }

__global__ void cycle_begin(int *st, double *small, Layout L, double beta, double target, std::uint32_t limit) {
    const std::uint32_t t = threadIdx.x;
    for (std::size_t i = t; i < 2 * static_cast<std::size_t>(L.stride); i += blockDim.x)
        small[L.g() + i] = 0.0;
    __syncthreads();
    if (t == 0) {
        small[L.g()] = beta;
        small[L.target()] = target;
        small[L.resid()] = beta;
        st[S_J] = 0;
        st[S_DONE] = limit == 0 ? 1 : 0;
        st[S_USED] = 0;
        st[S_LIMIT] = static_cast<int>(limit);
        st[S_ERROR] = 0;
    }
}

__global__ void load_current(const double *basis, std::uint32_t n, const int *st, double *current) {
    if (stopped(st))
        return;
    const std::size_t j = static_cast<std::size_t>(st[S_J]);
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * n)
        current[i] = basis[2 * j * n + i];
}

// part[b][c] = sum over rows of block b of conj(V[:, c]) w, c <= j (one wavefront per column).
__global__ __launch_bounds__(256) void project_step(const double *V, std::uint32_t n, const int *st, const double *w,
                                                   double *part) {
    if (stopped(st))
        return;
    const std::uint32_t m = static_cast<std::uint32_t>(st[S_J]) + 1;
    const std::uint32_t r0 = blockIdx.x * kRows;
    const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    const std::uint32_t c = blockIdx.y * 4 + wave;
    if (c >= m)
        return;
    double re = 0, im = 0;
    for (std::uint32_t r = r0 + lane; r < min(r0 + kRows, n); r += 64) {
        const double vr = V[2 * (static_cast<std::size_t>(c) * n + r)], vi = V[2 * (static_cast<std::size_t>(c) * n + r) + 1];
        const double xr = w[2 * r], xi = w[2 * r + 1];
        re += vr * xr + vi * xi;
        im += vr * xi - vi * xr;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) {
        part[2 * (static_cast<std::size_t>(blockIdx.x) * m + c)] = re;
        part[2 * (static_cast<std::size_t>(blockIdx.x) * m + c) + 1] = im;
    }
}

__global__ void sum_step(const double *part, std::uint32_t blocks, const int *st, double *out) {
    if (stopped(st))
        return;
    const std::uint32_t m = static_cast<std::uint32_t>(st[S_J]) + 1;
    const std::uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m)
        return;
    double re = 0, im = 0;
    for (std::uint32_t b = 0; b < blocks; ++b) {
        re += part[2 * (static_cast<std::size_t>(b) * m + c)];
        im += part[2 * (static_cast<std::size_t>(b) * m + c) + 1];
    }
    out[2 * c] = re;
    out[2 * c + 1] = im;
}

__global__ void subtract_step(const double *V, std::uint32_t n, const int *st, const double *h, double *w) {
    if (stopped(st))
        return;
    const std::uint32_t m = static_cast<std::uint32_t>(st[S_J]) + 1;
    const std::uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    double re = w[2 * r], im = w[2 * r + 1];
    for (std::uint32_t c = 0; c < m; ++c) {
        const double vr = V[2 * (static_cast<std::size_t>(c) * n + r)], vi = V[2 * (static_cast<std::size_t>(c) * n + r) + 1];
        const double hr = h[2 * c], hi = h[2 * c + 1];
        re -= vr * hr - vi * hi;
        im -= vr * hi + vi * hr;
    }
    w[2 * r] = re;
    w[2 * r + 1] = im;
}

__global__ __launch_bounds__(256) void norm_step_partial(const double *x, std::uint32_t n, const int *st, double *part) {
    if (stopped(st))
        return;
    __shared__ double red[4];
    double s = 0;
    for (std::uint32_t r = blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256) {
        const double a = x[2 * r] * 1.0, b = x[2 * r + 1] * 1.0;
        s += a * a + b * b;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void norm_step_finish(const double *part, std::uint32_t blocks, const int *st, double *out) {
    if (stopped(st))
        return;
    double s = 0;
    for (std::uint32_t b = threadIdx.x; b < blocks; b += 64)
        s += part[b];
    s = wave_sum(s);
    if (threadIdx.x == 0)
        out[0] = s;
}

// v_{j+1} = w / |w| (0 when |w| = 0).
__global__ void store_next(const double *w, std::uint32_t n, const int *st, const double *norm2, double *basis) {
    if (stopped(st))
        return;
    const std::size_t j = static_cast<std::size_t>(st[S_J]);
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n)
        return;
    const double v = *norm2;
    basis[2 * (j + 1) * n + i] = v > 0 ? w[i] / sqrt(v) : 0.0;
}

struct Cx {
    double re, im;
};
__device__ __forceinline__ Cx cmul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cx cadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cx cscale(Cx a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ Cx cconj(Cx a) { return {a.re, -a.im}; }
__device__ __forceinline__ Cx cneg(Cx a) { return {-a.re, -a.im}; }
__device__ __forceinline__ double cabs(Cx a) { return hypot(a.re, a.im); }

// Column j of the Hessenberg matrix from the two CGS passes and |w|, the earlier
// rotations applied to it, the new rotation, g rotated, and the stop decision.
__global__ void givens_step(int *st, double *small, Layout L) {
    if (threadIdx.x != 0 || stopped(st))
        return;
    const std::uint32_t j = static_cast<std::uint32_t>(st[S_J]);
    Cx *col = reinterpret_cast<Cx *>(small + L.H()) + static_cast<std::size_t>(j) * L.stride;
    const Cx *h1 = reinterpret_cast<const Cx *>(small + L.h1()), *h2 = reinterpret_cast<const Cx *>(small + L.h2());
    Cx *g = reinterpret_cast<Cx *>(small + L.g()), *sn = reinterpret_cast<Cx *>(small + L.sn());
    double *cs = small + L.cs();
    for (std::uint32_t i = 0; i <= j; ++i)
        col[i] = cadd(h1[i], h2[i]);
    for (std::uint32_t i = j + 2; i < L.stride; ++i)
        col[i] = {0.0, 0.0};
    const double next = sqrt(small[L.nrm()]);
    col[j + 1] = {next, 0.0};
    const bool breakdown = !(next > 0);
    for (std::uint32_t i = 0; i < j; ++i) {
        const Cx t = cadd(cscale(col[i], cs[i]), cmul(sn[i], col[i + 1]));
        col[i + 1] = cadd(cmul(cneg(cconj(sn[i])), col[i]), cscale(col[i + 1], cs[i]));
        col[i] = t;
    }
    const Cx a = col[j], bb = col[j + 1];
    const double absa = cabs(a), len = hypot(absa, cabs(bb));
    if (!(len > 0) || !isfinite(len)) {
        st[S_ERROR] = 1;
        st[S_DONE] = 1;
        return;
    }
    const Cx phase = absa > 0 ? Cx{a.re / absa, a.im / absa} : Cx{1.0, 0.0};
    cs[j] = absa / len;
    const Cx s = cmul(phase, cconj(bb));
    sn[j] = {s.re / len, s.im / len};
    col[j] = cscale(phase, len);
    col[j + 1] = {0.0, 0.0};
    const Cx t = cadd(cscale(g[j], cs[j]), cmul(sn[j], g[j + 1]));
    g[j + 1] = cadd(cmul(cneg(cconj(sn[j])), g[j]), cscale(g[j + 1], cs[j]));
    g[j] = t;
    const double resid = cabs(g[j + 1]);
    small[L.resid()] = resid;
    st[S_USED] = static_cast<int>(j + 1);
    st[S_J] = static_cast<int>(j + 1);
    if (resid <= small[L.target()] || breakdown || static_cast<int>(j + 1) >= st[S_LIMIT])
        st[S_DONE] = 1;
}

} // namespace

std::size_t gmres_cycle_small_doubles(std::uint32_t restart) { return Layout{restart + 1}.total(); }
std::size_t gmres_cycle_scratch_doubles(std::uint32_t n, std::uint32_t restart) {
    const std::size_t blocks = (n + kRows - 1) / kRows;
    return std::max<std::size_t>(2 * blocks * (restart + 1) + 1024, 1024 + 2048);
}
std::size_t gmres_cycle_h_offset(std::uint32_t restart) { return Layout{restart + 1}.H(); }
std::size_t gmres_cycle_g_offset(std::uint32_t restart) { return Layout{restart + 1}.g(); }

void gmres_cycle_begin(const GmresDeviceCycle &c, double beta, double target, std::uint32_t limit, hipStream_t s) {
    hipLaunchKernelGGL(cycle_begin, dim3(1), dim3(256), 0, s, c.state, c.small, Layout{c.restart + 1}, beta, target,
                       limit);
}

void gmres_cycle_load(const GmresDeviceCycle &c, hipStream_t s) {
    hipLaunchKernelGGL(load_current, dim3((2 * c.n + 255) / 256), dim3(256), 0, s, c.basis, c.n, c.state, c.current);
}

void gmres_cycle_orthogonalize(const GmresDeviceCycle &c, hipStream_t s) {
    const Layout L{c.restart + 1};
    const std::uint32_t n = c.n, blocks = (n + kRows - 1) / kRows, groups = (c.restart + 1 + 3) / 4;
    for (const std::size_t h : {L.h1(), L.h2()}) {
        hipLaunchKernelGGL(project_step, dim3(blocks, groups), dim3(256), 0, s, c.basis, n, c.state, c.w, c.scratch);
        hipLaunchKernelGGL(sum_step, dim3((c.restart + 1 + 63) / 64), dim3(64), 0, s, c.scratch, blocks, c.state,
                           c.small + h);
        hipLaunchKernelGGL(subtract_step, dim3((n + 255) / 256), dim3(256), 0, s, c.basis, n, c.state, c.small + h, c.w);
    }
    const std::uint32_t nblocks = std::min<std::uint32_t>(1024, (n + 255) / 256);
    hipLaunchKernelGGL(norm_step_partial, dim3(nblocks), dim3(256), 0, s, c.w, n, c.state, c.scratch);
    hipLaunchKernelGGL(norm_step_finish, dim3(1), dim3(64), 0, s, c.scratch, nblocks, c.state, c.small + L.nrm());
    hipLaunchKernelGGL(store_next, dim3((2 * n + 255) / 256), dim3(256), 0, s, c.w, n, c.state, c.small + L.nrm(),
                       c.basis);
    hipLaunchKernelGGL(givens_step, dim3(1), dim3(64), 0, s, c.state, c.small, L);
}

} // namespace hsbem::dev
