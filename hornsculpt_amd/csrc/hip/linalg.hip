// Dense complex FP64 linear algebra for the Burton-Miller system, gfx950.
//
// Replaces the vendor calls of the reference dense path
// (cusolverDnCgetrf/Cgetrs, cuda_solver.cu:1485-1513; cublasZgemv/Zaxpy/Dznrm2
// in device_gmres.cu:115-157) with own kernels:
//   gemv            y = A X for up to 8 vectors (A FP64 or FP32, vectors FP64):
//                   64 rows x 512 columns per workgroup, A streamed once (HBM
//                   bound), deterministic split-K partials reduced in a second pass;
//                   for one vector the column chunk (512, 1024 or 2048) is chosen by
//                   timing the candidates once per (n, precision) on the live stream.
//   project         CGS2 projections V^H w, one wavefront per (basis column, row block).
//   lu_factor       right-looking blocked LU (panel 32, 4-column leaves): pivot
//                   search, row interchanges, unit-lower TRSM, and trailing
//                   updates on the FP64 matrix cores.
//   lu_solve        blocked forward/backward substitution.
#include "../autotune.hpp"
#include "device.hpp"

#include <algorithm>
#include <cstdlib>
#include <optional>
#include <string>
#include <utility>
#include <vector>

namespace hsbem::dev {
namespace {

constexpr int kGemvRows = 64;    // rows per workgroup (one per lane)
constexpr int kGemvWaves = 4;    // waves per workgroup, splitting the column chunk
constexpr int kGemvCols = 512;   // columns per workgroup (split-K chunk)
constexpr int kGemvUnroll = 8;   // column loads in flight per lane
constexpr int kMaxVec = 8;

// Split-K partial products: workgroup (row block, column chunk); lane = row, the
// four waves take interleaved columns of the chunk and are summed in LDS, so one
// partial per (chunk, row) reaches HBM.
template <int NV, class T, int COLS = kGemvCols>
__global__ __launch_bounds__(kGemvRows *kGemvWaves) void gemv_partial(const T *__restrict__ A, std::uint32_t n,
                                                                     const double *__restrict__ x,
                                                                     std::uint32_t chunks, double *__restrict__ part,
                                                                     const int *skip) {
    if (skip && *skip) // a device-resident solver loop has already finished
        return;
    __shared__ double sx[NV][2 * COLS];
    __shared__ double red[kGemvWaves - 1][NV][2][kGemvRows];
    const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    const std::uint32_t row = blockIdx.x * kGemvRows + lane;
    const std::uint32_t c0 = blockIdx.y * COLS;
    const std::uint32_t cn = min(static_cast<std::uint32_t>(COLS), n - c0);
    for (std::uint32_t e = threadIdx.x; e < NV * 2 * COLS; e += kGemvRows * kGemvWaves) {
        const std::uint32_t v = e / (2 * COLS), c = e % (2 * COLS);
        sx[v][c] = c / 2 < cn ? x[2 * (static_cast<std::size_t>(v) * n + c0) + c] : 0.0;
    }
    __syncthreads();
    double re[NV], im[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v)
        re[v] = im[v] = 0.0;
    if (row < n) {
        const T *col = A + static_cast<std::size_t>(c0) * n + row;
        std::uint32_t c = wave;
        for (; c + kGemvWaves * (kGemvUnroll - 1) < cn; c += kGemvWaves * kGemvUnroll) {
            T a[kGemvUnroll];
#pragma unroll
            for (int u = 0; u < kGemvUnroll; ++u)
                a[u] = col[static_cast<std::size_t>(c + kGemvWaves * u) * n];
#pragma unroll
            for (int u = 0; u < kGemvUnroll; ++u) {
                const double ar = a[u].x, ai = a[u].y;
                const std::uint32_t cc = c + kGemvWaves * u;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const double xr = sx[v][2 * cc], xi = sx[v][2 * cc + 1];
                    re[v] += ar * xr - ai * xi;
                    im[v] += ar * xi + ai * xr;
                }
            }
        }
        for (; c < cn; c += kGemvWaves) {
            const T a = col[static_cast<std::size_t>(c) * n];
            const double ar = a.x, ai = a.y;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double xr = sx[v][2 * c], xi = sx[v][2 * c + 1];
                re[v] += ar * xr - ai * xi;
                im[v] += ar * xi + ai * xr;
            }
        }
    }
    if (wave > 0)
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            red[wave - 1][v][0][lane] = re[v];
            red[wave - 1][v][1][lane] = im[v];
        }
    __syncthreads();
    if (wave == 0 && row < n)
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double sr = re[v], si = im[v];
#pragma unroll
            for (int w = 0; w < kGemvWaves - 1; ++w) {
                sr += red[w][v][0][lane];
                si += red[w][v][1][lane];
            }
            const std::size_t o = (static_cast<std::size_t>(v) * chunks + blockIdx.y) * n + row;
            part[2 * o] = sr;
            part[2 * o + 1] = si;
        }
}

// The same split-K partials for an FP32 matrix and one vector with ROWS (2 or 4)
// consecutive rows per lane: ROWS / 2 adjacent 16-byte loads per column, so a
// wavefront covers 64 ROWS rows with fewer, wider load instructions. Needs n a
// multiple of ROWS (every column then starts on a 16-byte boundary and no lane is
// partly past the last row).
template <int COLS, int ROWS>
__global__ __launch_bounds__(kGemvRows *kGemvWaves) void gemv_partial_rows(const float2 *__restrict__ A,
                                                                          std::uint32_t n,
                                                                          const double *__restrict__ x,
                                                                          std::uint32_t chunks,
                                                                          double *__restrict__ part, const int *skip) {
    static_assert(ROWS == 2 || ROWS == 4, "two or four rows per lane");
    constexpr int kQuads = ROWS / 2;                 // float4 loads per column per lane
    constexpr int kUnroll = kGemvUnroll * 2 / ROWS;  // the same bytes in flight for both widths
    if (skip && *skip)
        return;
    __shared__ double sx[2 * COLS];
    __shared__ double red[kGemvWaves - 1][2 * ROWS][kGemvRows];
    const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    const std::uint32_t row = ROWS * (blockIdx.x * kGemvRows + lane);
    const std::uint32_t c0 = blockIdx.y * COLS;
    const std::uint32_t cn = min(static_cast<std::uint32_t>(COLS), n - c0);
    for (std::uint32_t c = threadIdx.x; c < 2 * COLS; c += kGemvRows * kGemvWaves)
        sx[c] = c / 2 < cn ? x[2 * static_cast<std::size_t>(c0) + c] : 0.0;
    __syncthreads();
    double acc[2 * ROWS] = {}; // re, im per row
    if (row < n) {
        const std::size_t stride = n / 2; // one column in float4 units
        const float4 *col = reinterpret_cast<const float4 *>(A + static_cast<std::size_t>(c0) * n + row);
        auto mac = [&](const float4 a, int q, std::uint32_t cc) {
            const double xr = sx[2 * cc], xi = sx[2 * cc + 1];
            const double a0 = a.x, b0 = a.y, a1 = a.z, b1 = a.w;
            acc[4 * q] += a0 * xr - b0 * xi;
            acc[4 * q + 1] += a0 * xi + b0 * xr;
            acc[4 * q + 2] += a1 * xr - b1 * xi;
            acc[4 * q + 3] += a1 * xi + b1 * xr;
        };
        std::uint32_t c = wave;
        for (; c + kGemvWaves * (kUnroll - 1) < cn; c += kGemvWaves * kUnroll) {
            float4 a[kUnroll][kQuads];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u)
#pragma unroll
                for (int q = 0; q < kQuads; ++q)
                    a[u][q] = col[static_cast<std::size_t>(c + kGemvWaves * u) * stride + q];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u)
#pragma unroll
                for (int q = 0; q < kQuads; ++q)
                    mac(a[u][q], q, c + kGemvWaves * u);
        }
        for (; c < cn; c += kGemvWaves)
#pragma unroll
            for (int q = 0; q < kQuads; ++q)
                mac(col[static_cast<std::size_t>(c) * stride + q], q, c);
    }
    if (wave > 0)
#pragma unroll
        for (int q = 0; q < 2 * ROWS; ++q)
            red[wave - 1][q][lane] = acc[q];
    __syncthreads();
    if (wave == 0 && row < n) {
#pragma unroll
        for (int w = 0; w < kGemvWaves - 1; ++w)
#pragma unroll
            for (int q = 0; q < 2 * ROWS; ++q)
                acc[q] += red[w][q][lane];
        const std::size_t o = static_cast<std::size_t>(blockIdx.y) * n + row;
#pragma unroll
        for (int q = 0; q < 2 * ROWS; ++q)
            part[2 * o + q] = acc[q];
    }
}

__global__ void gemv_reduce(const double *part, std::uint32_t n, std::uint32_t chunks, std::uint32_t nvec, double *y,
                            const int *skip) {
    if (skip && *skip)
        return;
    const std::size_t idx = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<std::size_t>(n) * nvec)
        return;
    const std::uint32_t row = idx % n, v = idx / n;
    double re = 0, im = 0;
    for (std::uint32_t c = 0; c < chunks; ++c) {
        const std::size_t o = (static_cast<std::size_t>(v) * chunks + c) * n + row;
        re += part[2 * o];
        im += part[2 * o + 1];
    }
    y[2 * idx] = re;
    y[2 * idx + 1] = im;
}

// Projection partials: workgroup (b, g) reduces rows [b*kProjectRows, ...) for the
// columns j = 4g + wave, one wavefront per column, into part[b][j].
constexpr int kProjectRows = 512;
__global__ __launch_bounds__(256) void project_partial(const double *V, std::uint32_t n, std::uint32_t m,
                                                      const double *w, double *part) {
    constexpr int kRows = kProjectRows;
    const std::uint32_t r0 = blockIdx.x * kRows;
    const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    const std::uint32_t j = blockIdx.y * 4 + wave;
    if (j >= m)
        return;
    double re = 0, im = 0;
    for (std::uint32_t r = r0 + lane; r < min(r0 + kRows, n); r += 64) {
        const double vr = V[2 * (static_cast<std::size_t>(j) * n + r)], vi = V[2 * (static_cast<std::size_t>(j) * n + r) + 1];
        const double xr = w[2 * r], xi = w[2 * r + 1];
        re += vr * xr + vi * xi; // conj(v) * x
        im += vr * xi - vi * xr;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) {
        part[2 * (static_cast<std::size_t>(blockIdx.x) * m + j)] = re;
        part[2 * (static_cast<std::size_t>(blockIdx.x) * m + j) + 1] = im;
    }
}

__global__ void sum_columns(const double *part, std::uint32_t blocks, std::uint32_t m, double *out) {
    const std::uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m)
        return;
    double re = 0, im = 0;
    for (std::uint32_t b = 0; b < blocks; ++b) {
        re += part[2 * (static_cast<std::size_t>(b) * m + j)];
        im += part[2 * (static_cast<std::size_t>(b) * m + j) + 1];
    }
    out[2 * j] = re;
    out[2 * j + 1] = im;
}

__global__ void subtract_combination(const double *V, std::uint32_t n, std::uint32_t m, const double *h, double *w) {
    const std::uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    double re = w[2 * r], im = w[2 * r + 1];
    for (std::uint32_t j = 0; j < m; ++j) {
        const double vr = V[2 * (static_cast<std::size_t>(j) * n + r)], vi = V[2 * (static_cast<std::size_t>(j) * n + r) + 1];
        const double hr = h[2 * j], hi = h[2 * j + 1];
        re -= vr * hr - vi * hi;
        im -= vr * hi + vi * hr;
    }
    w[2 * r] = re;
    w[2 * r + 1] = im;
}

__global__ __launch_bounds__(256) void norm_partial(const double *x, std::uint32_t n, const double *weight,
                                                   double *part) {
    __shared__ double red[4];
    double s = 0;
    for (std::uint32_t r = blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256) {
        const double f = weight ? weight[r] : 1.0;
        const double a = x[2 * r] * f, b = x[2 * r + 1] * f;
        s += a * a + b * b;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void norm_finish(const double *part, std::uint32_t blocks, double *out) {
    double s = 0;
    for (std::uint32_t b = threadIdx.x; b < blocks; b += 64)
        s += part[b];
    s = wave_sum(s);
    if (threadIdx.x == 0)
        out[0] = s;
}

__global__ void axpy_kernel(std::uint32_t n, const double *alpha, const double *x, double *y) {
    const std::uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    const double ar = alpha[0], ai = alpha[1], xr = x[2 * r], xi = x[2 * r + 1];
    y[2 * r] += ar * xr - ai * xi;
    y[2 * r + 1] += ar * xi + ai * xr;
}

__global__ void add_kernel(std::uint32_t n, const double *x, double *y) {
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n)
        return;
    y[i] += x[i];
}

__global__ void subtract_from_kernel(std::uint32_t n, const double *b, double *r) {
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n)
        return;
    r[i] = b[i] - r[i];
}

// y = x / sqrt(*norm2) with the squared norm read on the device (0 when the norm is 0).
__global__ void normalize_kernel(std::uint32_t n, const double *norm2, const double *x, double *y) {
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n)
        return;
    const double v = *norm2;
    y[i] = v > 0 ? x[i] / sqrt(v) : 0.0;
}

__global__ void scale_kernel(std::uint32_t n, double s, const double *x, double *y) {
    const std::uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    y[2 * r] = s * x[2 * r];
    y[2 * r + 1] = s * x[2 * r + 1];
}

__global__ void subtract_vectors(const double *b, const double *ax, std::size_t count, double *r) {
    const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count)
        r[i] = b[i] - ax[i];
}

// ---------------- LU ----------------
constexpr int kPanel = 32;
constexpr int kPanelThreads = 1024;

// Factor the leaf columns [j0, j0 + w) of the panel [k0, pe) over rows [j0, n).
// Per column j, three barriers: (1) pivot search (largest squared modulus, ties to
// the lowest row) reduced per wavefront into LDS; (2) every thread finishes the
// reduction itself, a few threads stage the pivot row p's leaf entries in LDS and
// one records the pivot; (3) one pass over the rows r > j writes
// l = a[r][j] / a[p][j] and a[r][c] - l u[c] for the leaf columns c > j, reading
// row j in place of row p (the interchange folded in), while the thread of row p
// also writes the pivot row into row j, and the other panel columns (the leaf's
// factored columns before j included) swap rows j and p. One workgroup; pivot rows
// are recorded as absolute indices.
constexpr int kLeaf = 4;
__global__ __launch_bounds__(kPanelThreads) void lu_leaf(double *A, std::uint32_t n, std::uint32_t k0, std::uint32_t j0,
                                                        std::uint32_t w, std::uint32_t pe, std::int32_t *piv, int *info) {
    __shared__ double best[kPanelThreads / 64];
    __shared__ std::uint32_t where[kPanelThreads / 64];
    __shared__ double2 u[kLeaf];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid / 64;
    double2 *A2 = reinterpret_cast<double2 *>(A);
    auto at = [&](std::uint32_t c, std::uint32_t r) -> double2 & { return A2[static_cast<std::size_t>(c) * n + r]; };
    const std::uint32_t je = j0 + w;
    for (std::uint32_t j = j0; j < je; ++j) {
        double bv = -1;
        std::uint32_t bi = j;
        for (std::uint32_t r = j + tid; r < n; r += kPanelThreads) {
            const double2 v = at(j, r);
            const double m = v.x * v.x + v.y * v.y;
            if (m > bv) {
                bv = m;
                bi = r;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off, 64);
            const std::uint32_t oi = __shfl_xor(bi, off, 64);
            if (ov > bv || (ov == bv && oi < bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            best[wave] = bv;
            where[wave] = bi;
        }
        __syncthreads(); // (1)
        double v = best[0];
        std::uint32_t p = where[0];
        for (int q = 1; q < kPanelThreads / 64; ++q)
            if (best[q] > v || (best[q] == v && where[q] < p)) {
                v = best[q];
                p = where[q];
            }
        if (tid < static_cast<int>(je - j))
            u[tid] = at(j + tid, p);
        if (tid == 0) {
            piv[j] = static_cast<std::int32_t>(p);
            if (!(v > 0) && *info == 0)
                *info = static_cast<int>(j) + 1;
        }
        __syncthreads(); // (2): every thread has p and the pivot row's leaf entries
        const double2 d = u[0];
        const double dd = d.x * d.x + d.y * d.y;
        const double ir = dd > 0 ? d.x / dd : 0.0, ii = dd > 0 ? -d.y / dd : 0.0; // 1/a = conj(a)/|a|^2
        for (std::uint32_t r = j + 1 + tid; r < n; r += kPanelThreads) {
            const std::uint32_t src = r == p ? j : r; // row p receives the old row j
            const double2 a = at(j, src);
            const double lr = a.x * ir - a.y * ii, li = a.x * ii + a.y * ir;
            for (std::uint32_t c = j + 1; c < je; ++c) {
                const double2 b = at(c, src), uc = u[c - j];
                at(c, r) = double2{b.x - (lr * uc.x - li * uc.y), b.y - (lr * uc.y + li * uc.x)};
            }
            at(j, r) = double2{lr, li};
            if (r == p) // this thread read the old row j; the pivot row takes its place
                for (std::uint32_t c = j; c < je; ++c)
                    at(c, j) = u[c - j];
        }
        if (p != j) // the other panel columns (incl. this leaf's L columns before j): interchange rows j and p
            for (std::uint32_t c = k0 + tid; c < pe; c += kPanelThreads)
                if (c < j || c >= je) {
                    const double2 t = at(c, j);
                    at(c, j) = at(c, p);
                    at(c, p) = t;
                }
        __syncthreads(); // (3)
    }
}

// Apply the panel's row interchanges to the columns outside it.
__global__ void lu_swap(double *A, std::uint32_t n, std::uint32_t k0, std::uint32_t nb, const std::int32_t *piv) {
    const std::uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || (c >= k0 && c < k0 + nb))
        return;
    double *col = A + 2 * static_cast<std::size_t>(c) * n;
    for (std::uint32_t j = k0; j < k0 + nb; ++j) {
        const std::uint32_t p = static_cast<std::uint32_t>(piv[j]);
        if (p != j) {
            const double t0 = col[2 * j], t1 = col[2 * j + 1];
            col[2 * j] = col[2 * p];
            col[2 * j + 1] = col[2 * p + 1];
            col[2 * p] = t0;
            col[2 * p + 1] = t1;
        }
    }
}

// U12 = L11^{-1} A12 (unit lower nb x nb L11 at [k0, k0 + nb)) for the columns
// [k0 + nb, c_end): 64 threads per 64 columns. The block's nb x 64 slice of A12 is
// staged through LDS with consecutive threads on consecutive rows of a column
// (coalesced), each thread solves one column from LDS, and the slice is written back
// the same way. Rows of the staged slice are padded against LDS bank conflicts.
__global__ __launch_bounds__(64) void lu_trsm(double *A, std::uint32_t n, std::uint32_t k0, std::uint32_t nb,
                                              std::uint32_t c_end) {
    __shared__ double L[kPanel][kPanel][2];
    __shared__ double2 X[64][kPanel + 1];
    for (std::uint32_t e = threadIdx.x; e < nb * nb; e += 64) {
        const std::uint32_t r = e % nb, c = e / nb;
        L[r][c][0] = A[2 * ((static_cast<std::size_t>(k0) + c) * n + k0 + r)];
        L[r][c][1] = A[2 * ((static_cast<std::size_t>(k0) + c) * n + k0 + r) + 1];
    }
    const std::uint32_t c_first = k0 + nb + blockIdx.x * 64;
    const std::uint32_t cols = min(64u, c_end - c_first);
    double2 *A2 = reinterpret_cast<double2 *>(A);
    for (std::uint32_t e = threadIdx.x; e < cols * nb; e += 64) {
        const std::uint32_t q = e / nb, r = e % nb;
        X[q][r] = A2[static_cast<std::size_t>(c_first + q) * n + k0 + r];
    }
    __syncthreads();
    const std::uint32_t t = threadIdx.x;
    if (t < cols) {
        double xr[kPanel], xi[kPanel];
        for (std::uint32_t i = 0; i < nb; ++i) {
            double vr = X[t][i].x, vi = X[t][i].y;
            for (std::uint32_t j = 0; j < i; ++j) {
                vr -= L[i][j][0] * xr[j] - L[i][j][1] * xi[j];
                vi -= L[i][j][0] * xi[j] + L[i][j][1] * xr[j];
            }
            xr[i] = vr;
            xi[i] = vi;
            X[t][i] = double2{vr, vi};
        }
    }
    __syncthreads();
    for (std::uint32_t e = threadIdx.x; e < cols * nb; e += 64) {
        const std::uint32_t q = e / nb, r = e % nb;
        A2[static_cast<std::size_t>(c_first + q) * n + k0 + r] = X[q][r];
    }
}

// C -= L U over rows [r_begin, r_end) and columns [c_begin, c_end) with L = A[rows][k0 ..
// k0 + nb) and U = A[k0 .. k0 + nb)[columns], K = nb <= 2 kPanel (staged through LDS in chunks
// of kPanel, the accumulators carried across the chunks), on the FP64 matrix cores: 256
// threads per 64 x 64 complex output tile, each wavefront a 32 x 32 quarter as 2 x 2 blocks
// of 16 x 16. The product is formed transposed, D = U^T L^T (v_mfma_f64_16x16x4f64: A[i][k]
// from lane i + 16 k, B[k][j] from lane j + 16 k, D[i][j] in lane j + 16 (i mod 4), register
// i / 4), so D's index j runs over rows and each 16-lane group writes 16 consecutive rows of
// one column. A complex MAC is four real MFMAs: Re += Ur Lr - Ui Li, Im += Ui Lr + Ur Li.
// Each chunk's K is padded with zeros to a multiple of 4.
typedef double lu_v4d __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void lu_update_mfma(double *A, std::uint32_t n, std::uint32_t k0, std::uint32_t nb,
                                                      std::uint32_t r_begin, std::uint32_t r_end,
                                                      std::uint32_t c_begin, std::uint32_t c_end) {
    __shared__ double2 Ls[kPanel][65]; // Ls[k][q] = A[r0 + q][k0 + kc + k] (padded row: no bank conflicts across k)
    __shared__ double2 Us[kPanel][65]; // Us[k][q] = A[k0 + kc + k][c0 + q]
    const std::uint32_t r0 = r_begin + blockIdx.x * 64, c0 = c_begin + blockIdx.y * 64;
    const double2 *A2 = reinterpret_cast<const double2 *>(A);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, i = lane & 15, kq = lane >> 4;
    const int rh = wave & 1, ch = wave >> 1;
    // this lane's old C values, loaded first so their latency hides behind the staging and the MFMAs
    double2 cold[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const std::uint32_t r = r0 + rh * 32 + a * 16 + i, c = c0 + ch * 32 + b * 16 + 4 * v + kq;
                cold[a][b][v] = r < r_end && c < c_end ? A2[static_cast<std::size_t>(c) * n + r] : double2{0.0, 0.0};
            }
    lu_v4d accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
            accr[a][b] = acci[a][b] = lu_v4d{0.0, 0.0, 0.0, 0.0};
    for (std::uint32_t kc = 0; kc < nb; kc += kPanel) {
        const std::uint32_t kn = min(static_cast<std::uint32_t>(kPanel), nb - kc), kpad = (kn + 3) & ~3u;
        if (kc)
            __syncthreads(); // the previous chunk's reads are done
        for (std::uint32_t e = threadIdx.x; e < kpad * 64; e += 256) {
            const std::uint32_t kk = e / 64, q = e % 64, r = r0 + q, c = c0 + q;
            const std::size_t k = static_cast<std::size_t>(k0) + kc + kk;
            Ls[kk][q] = kk < kn && r < r_end ? A2[k * n + r] : double2{0.0, 0.0};
            Us[kk][q] = kk < kn && c < c_end ? A2[static_cast<std::size_t>(c) * n + k] : double2{0.0, 0.0};
        }
        __syncthreads();
        for (std::uint32_t ks = 0; ks < kpad / 4; ++ks) {
            const std::uint32_t kk = 4 * ks + kq;
            double lr[2], li[2], ur[2], ui[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const double2 l = Ls[kk][rh * 32 + a * 16 + i];
                lr[a] = l.x;
                li[a] = l.y;
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const double2 u = Us[kk][ch * 32 + b * 16 + i];
                ur[b] = u.x;
                ui[b] = u.y;
            }
            // four passes over the eight independent accumulators: no MFMA waits on the one before it
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[b], lr[a], accr[a][b], 0, 0, 0);
                    acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ui[b], lr[a], acci[a][b], 0, 0, 0);
                }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ui[b], li[a], accr[a][b], 0, 0, 0);
                    acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[b], li[a], acci[a][b], 0, 0, 0);
                }
        }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const std::uint32_t r = r0 + rh * 32 + a * 16 + i, c = c0 + ch * 32 + b * 16 + 4 * v + kq;
                if (r < r_end && c < c_end)
                    reinterpret_cast<double2 *>(A)[static_cast<std::size_t>(c) * n + r] =
                        double2{cold[a][b][v].x - accr[a][b][v], cold[a][b][v].y - acci[a][b][v]};
            }
}

// Row interchanges of every panel applied to the right-hand sides.
__global__ void apply_pivots(double *B, std::uint32_t n, const std::int32_t *piv, std::uint32_t nrhs) {
    const std::uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrhs)
        return;
    double *x = B + 2 * static_cast<std::size_t>(e) * n;
    for (std::uint32_t j = 0; j < n; ++j) {
        const std::uint32_t p = static_cast<std::uint32_t>(piv[j]);
        if (p != j) {
            const double t0 = x[2 * j], t1 = x[2 * j + 1];
            x[2 * j] = x[2 * p];
            x[2 * j + 1] = x[2 * p + 1];
            x[2 * p] = t0;
            x[2 * p + 1] = t1;
        }
    }
}

constexpr int kSolveBlock = 64;

// Solve the diagonal block [j0, j0+nb) (lower unit or upper) for every rhs; one
// workgroup per rhs. The block is staged in LDS first (column by column, rows over
// the threads: coalesced, independent loads), so the nb sequential steps read LDS
// only; the arithmetic and its order are those of a column-by-column substitution.
__global__ __launch_bounds__(64) void block_substitute(const double *LU, std::uint32_t n, double *B, std::uint32_t j0,
                                                      std::uint32_t nb, int upper) {
    __shared__ double2 blk[kSolveBlock][kSolveBlock + 1]; // blk[c][r] = LU[j0 + r][j0 + c]
    __shared__ double2 x[kSolveBlock];
    const double2 *LU2 = reinterpret_cast<const double2 *>(LU);
    double2 *b = reinterpret_cast<double2 *>(B) + static_cast<std::size_t>(blockIdx.x) * n + j0;
    const int t = threadIdx.x, m = static_cast<int>(nb);
    if (t < m) {
        for (int c = 0; c < m; ++c)
            blk[c][t] = LU2[(static_cast<std::size_t>(j0) + c) * n + j0 + t];
        x[t] = b[t];
    }
    __syncthreads();
    if (!upper) {
        for (int j = 0; j < m; ++j) {
            if (t > j && t < m) {
                const double2 l = blk[j][t], xj = x[j];
                x[t].x -= l.x * xj.x - l.y * xj.y;
                x[t].y -= l.x * xj.y + l.y * xj.x;
            }
            __syncthreads();
        }
    } else {
        for (int jj = m - 1; jj >= 0; --jj) {
            if (t == jj) {
                const double2 u = blk[jj][jj];
                const double d = u.x * u.x + u.y * u.y;
                const double vr = x[jj].x, vi = x[jj].y;
                x[jj] = double2{(vr * u.x + vi * u.y) / d, (vi * u.x - vr * u.y) / d};
            }
            __syncthreads();
            if (t < jj) {
                const double2 u = blk[jj][t], xj = x[jj];
                x[t].x -= u.x * xj.x - u.y * xj.y;
                x[t].y -= u.x * xj.y + u.y * xj.x;
            }
            __syncthreads();
        }
    }
    if (t < m)
        b[t] = x[t];
}

// B[rows] -= LU[rows, j0:j0+nb] * B[j0:j0+nb] for rows outside the block
// (below for the lower sweep, above for the upper sweep).
__global__ __launch_bounds__(256) void block_eliminate(const double *LU, std::uint32_t n, double *B, std::uint32_t j0,
                                                      std::uint32_t nb, std::uint32_t r_begin, std::uint32_t r_end) {
    const std::uint32_t r = r_begin + blockIdx.x * 256 + threadIdx.x;
    const double *b = B + 2 * static_cast<std::size_t>(blockIdx.y) * n;
    if (r >= r_end)
        return;
    double re = 0, im = 0;
    for (std::uint32_t j = 0; j < nb; ++j) {
        const double lr = LU[2 * ((static_cast<std::size_t>(j0) + j) * n + r)],
                     li = LU[2 * ((static_cast<std::size_t>(j0) + j) * n + r) + 1];
        const double xr = b[2 * (j0 + j)], xi = b[2 * (j0 + j) + 1];
        re += lr * xr - li * xi;
        im += lr * xi + li * xr;
    }
    double *out = B + 2 * (static_cast<std::size_t>(blockIdx.y) * n + r);
    out[0] -= re;
    out[1] -= im;
}

// ---- block-Jacobi preconditioner: explicit inverses of the kJacobi x kJacobi
// diagonal blocks (consecutive internal nodes, i.e. spatial clusters) of the
// row-scaled system, one workgroup per block: LU with partial pivoting in LDS,
// then the inverse column by column (one thread per column of the identity,
// forward and back substitution against the shared factors). FP64, column-major
// per block. A block whose pivot vanishes is replaced by the identity. The apply
// is then a block-diagonal matrix-vector product.
constexpr int kJacobi = 64;

// Diagonal block entry (row r, column c of block blk) of a dense column-major matrix.
template <class T> struct DenseBlocks {
    const T *A;
    std::uint32_t n;
    __device__ double2 operator()(std::uint32_t blk, int r, int c) const {
        const T v = A[static_cast<std::size_t>(blk * kJacobi + c) * n + blk * kJacobi + r];
        return {static_cast<double>(v.x), static_cast<double>(v.y)};
    }
};

// ... of gathered blocks (complex FP64, column-major per block).
struct GatheredBlocks {
    const double *B;
    __device__ double2 operator()(std::uint32_t blk, int r, int c) const {
        const double *e = B + 2 * ((static_cast<std::size_t>(blk) * kJacobi + c) * kJacobi + r);
        return {e[0], e[1]};
    }
};

template <class Load>
__global__ __launch_bounds__(256) void jacobi_factor(Load load, std::uint32_t n, double *inverses) {
    __shared__ double a[kJacobi][kJacobi][2]; // LU, [col][row]
    __shared__ double x[kJacobi][kJacobi][2]; // inverse, [row][col] (thread = column: conflict free)
    __shared__ std::int32_t p[kJacobi];
    __shared__ int pivot_row, singular;
    const std::uint32_t blk = blockIdx.x, r0 = blk * kJacobi;
    const int m = static_cast<int>(min(static_cast<std::uint32_t>(kJacobi), n - r0));
    const int tid = threadIdx.x;
    for (int e = tid; e < kJacobi * kJacobi; e += 256) {
        const int c = e / kJacobi, r = e % kJacobi;
        double re = 0, im = 0;
        if (c < m && r < m) {
            const double2 v = load(blk, r, c);
            re = v.x;
            im = v.y;
        } else if (c == r) {
            re = 1;
        }
        a[c][r][0] = re;
        a[c][r][1] = im;
    }
    if (tid < kJacobi)
        p[tid] = tid;
    if (tid == 0)
        singular = 0;
    __syncthreads();
    for (int j = 0; j < m; ++j) {
        if (tid < 64) { // pivot search over rows j..m-1 of column j (wavefront 0)
            const int r = tid;
            double mag = r >= j && r < m ? a[j][r][0] * a[j][r][0] + a[j][r][1] * a[j][r][1] : -1.0;
            int idx = r;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double om = __shfl_xor(mag, off, 64);
                const int oi = __shfl_xor(idx, off, 64);
                if (om > mag || (om == mag && oi < idx)) {
                    mag = om;
                    idx = oi;
                }
            }
            if (tid == 0) {
                pivot_row = idx;
                if (!(mag > 0))
                    singular = 1;
            }
        }
        __syncthreads();
        if (singular)
            break;
        const int pr = pivot_row;
        if (pr != j) {
            for (int c = tid; c < kJacobi; c += 256) {
                const double t0 = a[c][j][0], t1 = a[c][j][1];
                a[c][j][0] = a[c][pr][0];
                a[c][j][1] = a[c][pr][1];
                a[c][pr][0] = t0;
                a[c][pr][1] = t1;
            }
            if (tid == 0) {
                const int t = p[j];
                p[j] = p[pr];
                p[pr] = t;
            }
        }
        __syncthreads();
        const double dr = a[j][j][0], di = a[j][j][1], inv = 1.0 / (dr * dr + di * di);
        for (int r = j + 1 + tid; r < m; r += 256) {
            const double xr = a[j][r][0], xi = a[j][r][1];
            a[j][r][0] = (xr * dr + xi * di) * inv;
            a[j][r][1] = (xi * dr - xr * di) * inv;
        }
        __syncthreads();
        const int w = m - j - 1;
        for (int e = tid; e < w * w; e += 256) {
            const int c = j + 1 + e / w, r = j + 1 + e % w;
            const double lr = a[j][r][0], li = a[j][r][1], ur = a[c][j][0], ui = a[c][j][1];
            a[c][r][0] -= lr * ur - li * ui;
            a[c][r][1] -= lr * ui + li * ur;
        }
        __syncthreads();
    }
    // inverse: column c of B^{-1} solves L U x = P e_c; thread c < m owns column c
    if (!singular && tid < m) {
        const int c = tid;
        for (int i = 0; i < m; ++i) { // y = P e_c
            x[i][c][0] = p[i] == c ? 1.0 : 0.0;
            x[i][c][1] = 0.0;
        }
        for (int j = 0; j < m; ++j) { // unit lower
            const double yr = x[j][c][0], yi = x[j][c][1];
            for (int i = j + 1; i < m; ++i) {
                const double lr = a[j][i][0], li = a[j][i][1];
                x[i][c][0] -= lr * yr - li * yi;
                x[i][c][1] -= lr * yi + li * yr;
            }
        }
        for (int j = m - 1; j >= 0; --j) { // upper
            const double dr = a[j][j][0], di = a[j][j][1], inv = 1.0 / (dr * dr + di * di);
            const double xr = x[j][c][0], xi = x[j][c][1];
            const double vr = (xr * dr + xi * di) * inv, vi = (xi * dr - xr * di) * inv;
            x[j][c][0] = vr;
            x[j][c][1] = vi;
            for (int i = 0; i < j; ++i) {
                const double ur = a[j][i][0], ui = a[j][i][1];
                x[i][c][0] -= ur * vr - ui * vi;
                x[i][c][1] -= ur * vi + ui * vr;
            }
        }
    }
    __syncthreads();
    double *f = inverses + static_cast<std::size_t>(blk) * kJacobi * kJacobi * 2;
    for (int e = tid; e < kJacobi * kJacobi; e += 256) {
        const int c = e / kJacobi, r = e % kJacobi;
        const bool id = singular != 0 || c >= m || r >= m;
        f[2 * e] = id ? (c == r ? 1.0 : 0.0) : x[r][c][0];
        f[2 * e + 1] = id ? 0.0 : x[r][c][1];
    }
}

// out = M^{-1} v: one wavefront per block, lane = row, columns streamed. Block
// rows are order[r0 .. r0 + 64) when order is given, else consecutive.
__global__ __launch_bounds__(64) void jacobi_apply(const double *inverses, std::uint32_t n, const double *v,
                                                   double *out, const std::uint32_t *order) {
    __shared__ double sv[kJacobi][2];
    const std::uint32_t blk = blockIdx.x, r0 = blk * kJacobi;
    const int m = static_cast<int>(min(static_cast<std::uint32_t>(kJacobi), n - r0));
    const int i = threadIdx.x;
    const std::uint32_t row = i < m ? (order ? order[r0 + i] : r0 + i) : 0;
    sv[i][0] = i < m ? v[2 * row] : 0.0;
    sv[i][1] = i < m ? v[2 * row + 1] : 0.0;
    __syncthreads();
    const double2 *f = reinterpret_cast<const double2 *>(inverses) + static_cast<std::size_t>(blk) * kJacobi * kJacobi;
    double re = 0, im = 0;
#pragma unroll 8
    for (int c = 0; c < kJacobi; ++c) {
        const double2 b = f[c * kJacobi + i];
        re += b.x * sv[c][0] - b.y * sv[c][1];
        im += b.x * sv[c][1] + b.y * sv[c][0];
    }
    if (i < m) {
        out[2 * row] = re;
        out[2 * row + 1] = im;
    }
}

// Block-Jacobi of any block width W (the FMM preconditioner's preconditionerBlockWidth).
// One workgroup per gathered W x W block (complex FP64, column-major a[c * W + r]),
// factored in place in global memory by a right-looking LU with partial pivoting in
// panels of kWidePanel columns: the panel is factored column by column (pivot search,
// full-row interchange, scaling, update of the remaining panel columns), then the
// panel's rows of U are solved against the unit L11 and the trailing matrix takes
// one rank-kWidePanel update, so the trailing matrix is read and written once per panel
// instead of once per column. The inverse X = U^{-1} L^{-1} P is formed the same
// way: X starts as P, a blocked forward sweep with L and a blocked backward sweep
// with U, each a small per-column triangular solve plus a rank-kWidePanel update of the
// rows below (above). A vanishing pivot gives the identity block, as in the LDS kernel.
constexpr int kWidePanel = 16;

// A[r][c] -= sum_t L(r, t) U(t, c) for r in [rb, re), c in [cb, ce), t < kb, with
// L(r, t) = lp[t * W + r], U(t, c) = up[c * W + t]; all column-major with leading
// dimension W. A thread keeps its row of L in registers and streams columns; when
// there are fewer rows than threads, the threads split the columns into groups.
__device__ void rank_update(double2 *A, std::uint32_t W, int rb, int re, int cb, int ce, const double2 *lp,
                            const double2 *up, int kb) {
    const int rows = re - rb;
    if (rows <= 0 || ce <= cb)
        return;
    const int threads = static_cast<int>(blockDim.x);
    const int groups = rows >= threads ? 1 : threads / rows;
    for (int e = static_cast<int>(threadIdx.x); e < rows * groups; e += threads) {
        const int r = rb + e % rows, g = e / rows;
        double2 l[kWidePanel];
#pragma unroll
        for (int t = 0; t < kWidePanel; ++t)
            l[t] = t < kb ? lp[static_cast<std::size_t>(t) * W + r] : double2{0, 0};
        for (int c = cb + g; c < ce; c += groups) {
            double2 acc = A[static_cast<std::size_t>(c) * W + r];
            const double2 *u = up + static_cast<std::size_t>(c) * W;
#pragma unroll
            for (int t = 0; t < kWidePanel; ++t)
                if (t < kb) {
                    const double2 v = u[t];
                    acc.x -= l[t].x * v.x - l[t].y * v.y;
                    acc.y -= l[t].x * v.y + l[t].y * v.x;
                }
            A[static_cast<std::size_t>(c) * W + r] = acc;
        }
    }
}

__global__ __launch_bounds__(256) void jacobi_factor_wide(double *blocks, std::uint32_t n, std::uint32_t width,
                                                          double *inverses, std::int32_t *pivots) {
    __shared__ double red_mag[256];
    __shared__ int red_idx[256];
    __shared__ int singular;
    const std::uint32_t blk = blockIdx.x, W = width, r0 = blk * W;
    const int m = static_cast<int>(min(W, n - r0));
    double2 *a = reinterpret_cast<double2 *>(blocks) + static_cast<std::size_t>(blk) * W * W;
    std::int32_t *p = pivots + static_cast<std::size_t>(blk) * W;
    const int tid = threadIdx.x;
    auto at = [&](int c, int r) -> double2 & { return a[static_cast<std::size_t>(c) * W + r]; };
    for (int r = tid; r < static_cast<int>(W); r += 256)
        p[r] = r;
    if (tid == 0)
        singular = 0;
    __syncthreads();
    for (int j0 = 0; j0 < m; j0 += kWidePanel) {
        const int je = min(j0 + kWidePanel, m);
        for (int j = j0; j < je; ++j) {
            double best = -1.0;
            int bi = j;
            for (int r = j + tid; r < m; r += 256) {
                const double2 v = at(j, r);
                const double mag = v.x * v.x + v.y * v.y;
                if (mag > best) {
                    best = mag;
                    bi = r;
                }
            }
            red_mag[tid] = best;
            red_idx[tid] = bi;
            __syncthreads();
            for (int off = 128; off > 0; off >>= 1) {
                if (tid < off) {
                    const double om = red_mag[tid + off];
                    const int oi = red_idx[tid + off];
                    if (om > red_mag[tid] || (om == red_mag[tid] && oi < red_idx[tid])) {
                        red_mag[tid] = om;
                        red_idx[tid] = oi;
                    }
                }
                __syncthreads();
            }
            if (!(red_mag[0] > 0)) { // uniform: every thread reads the reduced value
                if (tid == 0)
                    singular = 1;
                break;
            }
            const int pr = red_idx[0];
            if (pr != j) {
                for (int c = tid; c < m; c += 256) {
                    const double2 t = at(c, j);
                    at(c, j) = at(c, pr);
                    at(c, pr) = t;
                }
                if (tid == 0) {
                    const int t = p[j];
                    p[j] = p[pr];
                    p[pr] = t;
                }
            }
            __syncthreads();
            const double2 d = at(j, j);
            const double inv = 1.0 / (d.x * d.x + d.y * d.y);
            for (int r = j + 1 + tid; r < m; r += 256) {
                const double2 x = at(j, r);
                at(j, r) = {(x.x * d.x + x.y * d.y) * inv, (x.y * d.x - x.x * d.y) * inv};
            }
            __syncthreads();
            const int rows = m - j - 1, cols = je - j - 1;
            for (int e = tid; e < rows * cols; e += 256) {
                const int c = j + 1 + e / rows, r = j + 1 + e % rows;
                const double2 l = at(j, r), u = at(c, j);
                double2 &t = at(c, r);
                t.x -= l.x * u.x - l.y * u.y;
                t.y -= l.x * u.y + l.y * u.x;
            }
            __syncthreads();
        }
        __syncthreads();
        if (singular)
            break;
        for (int c = je + tid; c < m; c += 256) // U12 = L11^{-1} A12
            for (int j = j0; j < je; ++j) {
                const double2 u = at(c, j);
                for (int i = j + 1; i < je; ++i) {
                    const double2 l = at(j, i);
                    at(c, i).x -= l.x * u.x - l.y * u.y;
                    at(c, i).y -= l.x * u.y + l.y * u.x;
                }
            }
        __syncthreads();
        rank_update(a, W, je, m, je, m, a + static_cast<std::size_t>(j0) * W, a + j0, je - j0);
        __syncthreads();
    }
    double2 *x = reinterpret_cast<double2 *>(inverses) + static_cast<std::size_t>(blk) * W * W; // x[c * W + r]
    const bool identity = singular != 0;
    for (std::size_t e = tid; e < static_cast<std::size_t>(W) * W; e += 256) {
        const int c = static_cast<int>(e / W), r = static_cast<int>(e % W);
        const bool inside = !identity && r < m && c < m;
        x[e] = {inside ? (p[r] == c ? 1.0 : 0.0) : (r == c ? 1.0 : 0.0), 0.0};
    }
    __syncthreads();
    if (identity)
        return;
    auto xt = [&](int c, int r) -> double2 & { return x[static_cast<std::size_t>(c) * W + r]; };
    for (int j0 = 0; j0 < m; j0 += kWidePanel) { // L X = P
        const int je = min(j0 + kWidePanel, m);
        for (int c = tid; c < m; c += 256)
            for (int j = j0; j < je; ++j) {
                const double2 y = xt(c, j);
                for (int i = j + 1; i < je; ++i) {
                    const double2 l = at(j, i);
                    xt(c, i).x -= l.x * y.x - l.y * y.y;
                    xt(c, i).y -= l.x * y.y + l.y * y.x;
                }
            }
        __syncthreads();
        rank_update(x, W, je, m, 0, m, a + static_cast<std::size_t>(j0) * W, x + j0, je - j0);
        __syncthreads();
    }
    const int last = ((m - 1) / kWidePanel) * kWidePanel;
    for (int j0 = last; j0 >= 0; j0 -= kWidePanel) { // U X = Y
        const int je = min(j0 + kWidePanel, m);
        for (int c = tid; c < m; c += 256)
            for (int j = je - 1; j >= j0; --j) {
                const double2 d = at(j, j), y = xt(c, j);
                const double inv = 1.0 / (d.x * d.x + d.y * d.y);
                const double2 v = {(y.x * d.x + y.y * d.y) * inv, (y.y * d.x - y.x * d.y) * inv};
                xt(c, j) = v;
                for (int i = j0; i < j; ++i) {
                    const double2 u = at(j, i);
                    xt(c, i).x -= u.x * v.x - u.y * v.y;
                    xt(c, i).y -= u.x * v.y + u.y * v.x;
                }
            }
        __syncthreads();
        rank_update(x, W, 0, j0, 0, m, a + static_cast<std::size_t>(j0) * W, x + j0, je - j0);
        __syncthreads();
    }
}

// out = M^{-1} v for width W: one workgroup of min(W, 256) threads per block, the
// block's v in LDS, threads stride over the rows (column-major inverse: coalesced
// per column).
__global__ __launch_bounds__(256) void jacobi_apply_wide(const double *inverses, std::uint32_t n, std::uint32_t width,
                                                         const double *v, double *out, const std::uint32_t *order) {
    __shared__ double2 sv[kMaxJacobiWidth];
    const std::uint32_t blk = blockIdx.x, W = width, r0 = blk * W;
    const int m = static_cast<int>(min(W, n - r0)), threads = static_cast<int>(blockDim.x);
    for (int i = threadIdx.x; i < m; i += threads) {
        const std::uint32_t row = order ? order[r0 + i] : r0 + i;
        sv[i] = double2{v[2 * row], v[2 * row + 1]};
    }
    __syncthreads();
    const double2 *f = reinterpret_cast<const double2 *>(inverses) + static_cast<std::size_t>(blk) * W * W;
    for (int r = threadIdx.x; r < m; r += threads) {
        double re = 0, im = 0;
        for (int c = 0; c < m; ++c) {
            const double2 b = f[static_cast<std::size_t>(c) * W + r];
            re += b.x * sv[c].x - b.y * sv[c].y;
            im += b.x * sv[c].y + b.y * sv[c].x;
        }
        const std::uint32_t row = order ? order[r0 + r] : r0 + r;
        out[2 * row] = re;
        out[2 * row + 1] = im;
    }
}

// One thread per entry of the gathered blocks: block b, column c, row r takes
// A[(W b + c) n + W b + r] (complex, widened to FP64).
template <class T>
__global__ __launch_bounds__(256) void jacobi_gather_dense_kernel(const T *A, std::uint32_t n, std::uint32_t width,
                                                                 double *blocks) {
    const std::size_t e = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const std::size_t W = width, blocks_count = (n + W - 1) / W;
    if (e >= blocks_count * W * W)
        return;
    const std::size_t b = e / (W * W), c = (e / W) % W, r = e % W;
    const std::size_t row = b * W + r, col = b * W + c;
    if (row >= n || col >= n)
        return;
    const T a = A[col * n + row];
    blocks[2 * e] = static_cast<double>(a.x);
    blocks[2 * e + 1] = static_cast<double>(a.y);
}

} // namespace

void jacobi_gather_dense(const void *A, bool fp32, std::uint32_t n, std::uint32_t width, double *blocks,
                         hipStream_t stream) {
    const std::size_t count = static_cast<std::size_t>((n + width - 1) / width) * width * width;
    const dim3 grid(static_cast<std::uint32_t>((count + 255) / 256));
    if (fp32)
        hipLaunchKernelGGL(jacobi_gather_dense_kernel<float2>, grid, dim3(256), 0, stream,
                           static_cast<const float2 *>(A), n, width, blocks);
    else
        hipLaunchKernelGGL(jacobi_gather_dense_kernel<double2>, grid, dim3(256), 0, stream,
                           static_cast<const double2 *>(A), n, width, blocks);
}

std::size_t jacobi_wide_doubles(std::uint32_t n, std::uint32_t width) {
    return static_cast<std::size_t>((n + width - 1) / width) * width * width * 2;
}

void jacobi_factor_wide(double *blocks, std::uint32_t n, std::uint32_t width, double *inverses, std::int32_t *pivots,
                        hipStream_t stream) {
    hipLaunchKernelGGL(jacobi_factor_wide, dim3((n + width - 1) / width), dim3(256), 0, stream, blocks, n, width,
                       inverses, pivots);
}

void jacobi_apply_wide(const double *inverses, std::uint32_t n, std::uint32_t width, const double *v, double *out,
                       hipStream_t stream, const std::uint32_t *order) {
    hipLaunchKernelGGL(jacobi_apply_wide, dim3((n + width - 1) / width), dim3(std::min(width, 256u)), 0, stream,
                       inverses, n, width, v, out, order);
}

std::size_t gemv_scratch_doubles(std::uint32_t n, std::uint32_t nvec) {
    const std::size_t chunks = (n + kGemvCols - 1) / kGemvCols;
    return 2 * chunks * n * std::max<std::uint32_t>(nvec, 1);
}

namespace {

template <int COLS>
void launch_gemv_one(const void *A, bool fp32, std::uint32_t n, const double *x, double *y, double *scratch,
                     hipStream_t stream, const int *skip, std::uint32_t rows) {
    const std::uint32_t chunks = (n + COLS - 1) / COLS;
    const dim3 grid((n + kGemvRows - 1) / kGemvRows, chunks);
    if (fp32 && rows == 4 && n % 4 == 0)
        hipLaunchKernelGGL((gemv_partial_rows<COLS, 4>), dim3((n + 4 * kGemvRows - 1) / (4 * kGemvRows), chunks),
                           dim3(kGemvRows * kGemvWaves), 0, stream, static_cast<const float2 *>(A), n, x, chunks,
                           scratch, skip);
    else if (fp32 && rows == 2 && n % 2 == 0)
        hipLaunchKernelGGL((gemv_partial_rows<COLS, 2>), dim3((n + 2 * kGemvRows - 1) / (2 * kGemvRows), chunks),
                           dim3(kGemvRows * kGemvWaves), 0, stream, static_cast<const float2 *>(A), n, x, chunks,
                           scratch, skip);
    else if (fp32)
        hipLaunchKernelGGL((gemv_partial<1, float2, COLS>), grid, dim3(kGemvRows * kGemvWaves), 0, stream,
                           static_cast<const float2 *>(A), n, x, chunks, scratch, skip);
    else
        hipLaunchKernelGGL((gemv_partial<1, double2, COLS>), grid, dim3(kGemvRows * kGemvWaves), 0, stream,
                           static_cast<const double2 *>(A), n, x, chunks, scratch, skip);
    hipLaunchKernelGGL(gemv_reduce, dim3((n + 255) / 256), dim3(256), 0, stream, scratch, n, chunks, 1u, y, skip);
}

// A product shape: the column chunk (512 / 1024 / 2048, low 16 bits) and, for an FP32
// matrix whose order they divide, the rows per lane (1, 2 or 4; bits 16 and up).
constexpr std::uint32_t gemv_shape(std::uint32_t cols, std::uint32_t rows) { return cols | rows << 16; }

void launch_gemv_cols(std::uint32_t shape, const void *A, bool fp32, std::uint32_t n, const double *x, double *y,
                      double *scratch, hipStream_t stream, const int *skip = nullptr) {
    const std::uint32_t cols = shape & 0xffffu, rows = std::max(shape >> 16, 1u);
    if (cols == 2048)
        launch_gemv_one<2048>(A, fp32, n, x, y, scratch, stream, skip, rows);
    else if (cols == 1024)
        launch_gemv_one<1024>(A, fp32, n, x, y, scratch, stream, skip, rows);
    else
        launch_gemv_one<kGemvCols>(A, fp32, n, x, y, scratch, stream, skip, rows);
}

// Shape of the one-vector product per (n, precision), chosen by the autotuner among
// the column chunks 512 / 1024 / 2048 and, for an FP32 matrix whose order they
// divide, one, two or four rows per lane: three timed launches (after one warm launch)
// of each on the caller's stream and data; the product it leaves in y is the last
// candidate's, and the caller's own launch follows. The default chunk while the
// stream is being captured; HSBEM_GEMV_COLS = 512 / 1024 / 2048 pins the chunk (one
// row per lane), "COLSx2" / "COLSx4" the chunk and rows per lane (FP32 matrix whose
// order they divide).
std::uint32_t gemv_columns(const void *A, bool fp32, std::uint32_t n, const double *x, double *y, double *scratch,
                           hipStream_t stream) {
    std::optional<std::uint32_t> pinned;
    if (const char *pin = std::getenv("HSBEM_GEMV_COLS")) { // "COLS" or "COLSxROWS"
        char *end = nullptr;
        const unsigned long v = std::strtoul(pin, &end, 10);
        unsigned long rows = 1;
        if (end && *end == 'x')
            rows = std::strtoul(end + 1, nullptr, 10);
        const bool rows_ok = rows == 1 || ((rows == 2 || rows == 4) && fp32 && n % rows == 0);
        if ((v == 512 || v == 1024 || v == 2048) && rows_ok)
            pinned = gemv_shape(static_cast<std::uint32_t>(v), static_cast<std::uint32_t>(rows));
    }
    std::vector<std::uint32_t> shapes;
    for (const std::uint32_t rows : {1u, 2u, 4u})
        if (rows == 1 || (fp32 && n % rows == 0))
            for (const std::uint32_t cols : {512u, 1024u, 2048u})
                shapes.push_back(gemv_shape(cols, rows));
    const std::string key = "n=" + std::to_string(n) + (fp32 ? " fp32" : " fp64");
    return autotune::choose(
        "gemv_shape", key, shapes, kGemvCols, stream,
        [&](std::uint32_t shape) { launch_gemv_cols(shape, A, fp32, n, x, y, scratch, stream); }, 3, pinned);
}

} // namespace

void launch_gemv(const void *A, bool fp32, std::uint32_t n, const double *x, double *y, std::uint32_t nvec,
                 double *scratch, hipStream_t stream, const int *skip) {
    if (nvec == 1) {
        launch_gemv_cols(gemv_columns(A, fp32, n, x, y, scratch, stream), A, fp32, n, x, y, scratch, stream, skip);
        return;
    }
    const std::uint32_t chunks = (n + kGemvCols - 1) / kGemvCols;
    for (std::uint32_t v0 = 0; v0 < nvec; v0 += kMaxVec) {
        const std::uint32_t nv = std::min<std::uint32_t>(kMaxVec, nvec - v0);
        const dim3 grid((n + kGemvRows - 1) / kGemvRows, chunks);
        const double *xs = x + 2 * static_cast<std::size_t>(v0) * n;
        switch (nv) {
#define HS_GEMV_CASE(N)                                                                                          \
    case N:                                                                                                      \
        if (fp32)                                                                                                \
            hipLaunchKernelGGL((gemv_partial<N, float2>), grid, dim3(kGemvRows * kGemvWaves), 0, stream,                      \
                               static_cast<const float2 *>(A), n, xs, chunks, scratch, nullptr);                 \
        else                                                                                                     \
            hipLaunchKernelGGL((gemv_partial<N, double2>), grid, dim3(kGemvRows * kGemvWaves), 0, stream,                     \
                               static_cast<const double2 *>(A), n, xs, chunks, scratch, nullptr);                \
        break;
            HS_GEMV_CASE(1)
            HS_GEMV_CASE(2)
            HS_GEMV_CASE(3)
            HS_GEMV_CASE(4)
            HS_GEMV_CASE(5)
            HS_GEMV_CASE(6)
            HS_GEMV_CASE(7)
            HS_GEMV_CASE(8)
#undef HS_GEMV_CASE
        }
        const std::size_t count = static_cast<std::size_t>(n) * nv;
        hipLaunchKernelGGL(gemv_reduce, dim3((count + 255) / 256), dim3(256), 0, stream, scratch, n, chunks, nv,
                           y + 2 * static_cast<std::size_t>(v0) * n, nullptr);
    }
}

std::size_t reduction_scratch_doubles(std::uint32_t n, std::uint32_t m) {
    const std::size_t blocks = (n + kProjectRows - 1) / kProjectRows;
    return 2 * blocks * std::max<std::uint32_t>(m, 1) + 1024;
}

void launch_project(const double *V, std::uint32_t n, std::uint32_t m, const double *w, double *out, double *scratch,
                    hipStream_t stream) {
    const std::uint32_t blocks = (n + kProjectRows - 1) / kProjectRows;
    hipLaunchKernelGGL(project_partial, dim3(blocks, (m + 3) / 4), dim3(256), 0, stream, V, n, m, w, scratch);
    hipLaunchKernelGGL(sum_columns, dim3((m + 63) / 64), dim3(64), 0, stream, scratch, blocks, m, out);
}

void launch_subtract_combination(const double *V, std::uint32_t n, std::uint32_t m, const double *h, double *w,
                                 hipStream_t stream) {
    hipLaunchKernelGGL(subtract_combination, dim3((n + 255) / 256), dim3(256), 0, stream, V, n, m, h, w);
}

void launch_norm2(const double *x, std::uint32_t n, const double *weight, double *out, double *scratch,
                  hipStream_t stream) {
    const std::uint32_t blocks = std::min<std::uint32_t>(1024, (n + 255) / 256);
    hipLaunchKernelGGL(norm_partial, dim3(blocks), dim3(256), 0, stream, x, n, weight, scratch);
    hipLaunchKernelGGL(norm_finish, dim3(1), dim3(64), 0, stream, scratch, blocks, out);
}

void launch_add(std::uint32_t n, const double *x, double *y, hipStream_t stream) {
    hipLaunchKernelGGL(add_kernel, dim3((2 * n + 255) / 256), dim3(256), 0, stream, n, x, y);
}

void launch_subtract_from(std::uint32_t n, const double *b, double *r, hipStream_t stream) {
    hipLaunchKernelGGL(subtract_from_kernel, dim3((2 * n + 255) / 256), dim3(256), 0, stream, n, b, r);
}

void launch_axpy(std::uint32_t n, const double *alpha, const double *x, double *y, hipStream_t stream) {
    hipLaunchKernelGGL(axpy_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, alpha, x, y);
}

void launch_normalize(std::uint32_t n, const double *norm2, const double *x, double *y, hipStream_t stream) {
    hipLaunchKernelGGL(normalize_kernel, dim3((2 * n + 255) / 256), dim3(256), 0, stream, n, norm2, x, y);
}

void launch_scale(std::uint32_t n, double s, const double *x, double *y, hipStream_t stream) {
    hipLaunchKernelGGL(scale_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, s, x, y);
}

void launch_residual(const void *A, bool fp32, std::uint32_t n, const double *b, const double *x, std::uint32_t nvec,
                     double *residual, double *scratch, hipStream_t stream) {
    launch_gemv(A, fp32, n, x, residual, nvec, scratch, stream);
    const std::size_t count = 2 * static_cast<std::size_t>(n) * nvec;
    hipLaunchKernelGGL(subtract_vectors, dim3((count + 255) / 256), dim3(256), 0, stream, b, residual, count, residual);
}

std::size_t lu_work_doubles(std::uint32_t) { return 0; }

namespace {

// The leaves of panel [k0, k0 + nb): pivoted 4-column leaves (row interchanges inside the
// panel), each leaf's U rows of the panel and the rank-4 update of the panel below it.
void factor_panel(double *A, std::uint32_t n, std::uint32_t k0, std::uint32_t nb, std::int32_t *piv, int *info,
                  hipStream_t stream) {
    const std::uint32_t pe = k0 + nb;
    for (std::uint32_t j0 = k0; j0 < pe; j0 += kLeaf) {
        const std::uint32_t w = std::min<std::uint32_t>(kLeaf, pe - j0);
        hipLaunchKernelGGL(lu_leaf, dim3(1), dim3(kPanelThreads), 0, stream, A, n, k0, j0, w, pe, piv, info);
        if (j0 + w >= pe)
            continue;
        hipLaunchKernelGGL(lu_trsm, dim3(1), dim3(64), 0, stream, A, n, j0, w, pe);
        const std::uint32_t below = n - j0 - w;
        hipLaunchKernelGGL(lu_update_mfma, dim3((below + 63) / 64, 1), dim3(256), 0, stream, A, n, j0, w, j0 + w, n,
                           j0 + w, pe);
    }
}

// C -= L U over rows [r0, r1) x columns [c0, c1) with K = A's columns / rows [k0, k0 + nb).
void update(double *A, std::uint32_t n, std::uint32_t k0, std::uint32_t nb, std::uint32_t r0, std::uint32_t r1,
            std::uint32_t c0, std::uint32_t c1, hipStream_t stream) {
    if (r1 > r0 && c1 > c0)
        hipLaunchKernelGGL(lu_update_mfma, dim3((r1 - r0 + 63) / 64, (c1 - c0 + 63) / 64), dim3(256), 0, stream, A, n,
                           k0, nb, r0, r1, c0, c1);
}

} // namespace

// Right-looking blocked LU with partial pivoting: panels of kPanel columns, each
// factored in leaves of kLeaf columns (lu_leaf on one workgroup, then the leaf's
// rows of U for the rest of the panel and a rank-kLeaf update of the panel below
// it on every CU), then the panel's interchanges outside it, U12 and the trailing
// update on the matrix cores, two panels to one rank-2 kPanel trailing update.
void lu_factor(double *A, std::uint32_t n, std::int32_t *piv, int *info, double *, hipStream_t stream) {
    // Pairs of kPanel-column panels P1 = [k0, e1), P2 = [e1, e2) share one rank-2 kPanel
    // trailing update (half the trailing matrix's HBM traffic per flop of rank-kPanel
    // updates): P1 is factored, its interchanges applied and its U rows solved over every
    // later column; P1's update reaches P2's columns before P2 is factored; after P2's
    // interchanges (applied to the still un-updated trailing rows and to P1's L rows alike)
    // P1's update reaches P2's U rows, P2's U rows are solved, and the rows below P2 take
    // both panels' updates at once.
    for (std::uint32_t k0 = 0; k0 < n; k0 += 2 * kPanel) {
        const std::uint32_t n1 = std::min<std::uint32_t>(kPanel, n - k0), e1 = k0 + n1;
        factor_panel(A, n, k0, n1, piv, info, stream);
        hipLaunchKernelGGL(lu_swap, dim3((n + 63) / 64), dim3(64), 0, stream, A, n, k0, n1, piv);
        if (e1 == n)
            break;
        hipLaunchKernelGGL(lu_trsm, dim3((n - e1 + 63) / 64), dim3(64), 0, stream, A, n, k0, n1, n);
        const std::uint32_t n2 = std::min<std::uint32_t>(kPanel, n - e1), e2 = e1 + n2;
        update(A, n, k0, n1, e1, n, e1, e2, stream);
        factor_panel(A, n, e1, n2, piv, info, stream);
        hipLaunchKernelGGL(lu_swap, dim3((n + 63) / 64), dim3(64), 0, stream, A, n, e1, n2, piv);
        if (e2 == n)
            break;
        update(A, n, k0, n1, e1, e2, e2, n, stream);
        hipLaunchKernelGGL(lu_trsm, dim3((n - e2 + 63) / 64), dim3(64), 0, stream, A, n, e1, n2, n);
        update(A, n, k0, n1 + n2, e2, n, e2, n, stream);
    }
}

void lu_solve(const double *LU, std::uint32_t n, const std::int32_t *piv, double *B, std::uint32_t nrhs,
              hipStream_t stream) {
    hipLaunchKernelGGL(apply_pivots, dim3((nrhs + 63) / 64), dim3(64), 0, stream, B, n, piv, nrhs);
    for (std::uint32_t j0 = 0; j0 < n; j0 += kSolveBlock) {
        const std::uint32_t nb = std::min<std::uint32_t>(kSolveBlock, n - j0);
        hipLaunchKernelGGL(block_substitute, dim3(nrhs), dim3(64), 0, stream, LU, n, B, j0, nb, 0);
        const std::uint32_t rows = n - j0 - nb;
        if (rows)
            hipLaunchKernelGGL(block_eliminate, dim3((rows + 255) / 256, nrhs), dim3(256), 0, stream, LU, n, B, j0, nb,
                               j0 + nb, n);
    }
    const std::uint32_t last = ((n - 1) / kSolveBlock) * kSolveBlock;
    for (std::int64_t j0 = last; j0 >= 0; j0 -= kSolveBlock) {
        const std::uint32_t nb = std::min<std::uint32_t>(kSolveBlock, n - static_cast<std::uint32_t>(j0));
        hipLaunchKernelGGL(block_substitute, dim3(nrhs), dim3(64), 0, stream, LU, n, B, static_cast<std::uint32_t>(j0),
                           nb, 1);
        if (j0 > 0)
            hipLaunchKernelGGL(block_eliminate, dim3((static_cast<std::uint32_t>(j0) + 255) / 256, nrhs), dim3(256), 0,
                               stream, LU, n, B, static_cast<std::uint32_t>(j0), nb, 0u, static_cast<std::uint32_t>(j0));
    }
}

std::size_t jacobi_factor_doubles(std::uint32_t n) {
    return static_cast<std::size_t>((n + kJacobi - 1) / kJacobi) * kJacobi * kJacobi * 2;
}

void jacobi_factor(const void *A, bool fp32, std::uint32_t n, double *inverses, hipStream_t stream) {
    const dim3 grid((n + kJacobi - 1) / kJacobi);
    if (fp32)
        hipLaunchKernelGGL(jacobi_factor<DenseBlocks<float2>>, grid, dim3(256), 0, stream,
                           DenseBlocks<float2>{static_cast<const float2 *>(A), n}, n,
                           inverses);
    else
        hipLaunchKernelGGL(jacobi_factor<DenseBlocks<double2>>, grid, dim3(256), 0, stream,
                           DenseBlocks<double2>{static_cast<const double2 *>(A), n}, n,
                           inverses);
}

void jacobi_factor_gathered(const double *blocks, std::uint32_t n, double *inverses, hipStream_t stream) {
    const dim3 grid((n + kJacobi - 1) / kJacobi);
    hipLaunchKernelGGL(jacobi_factor<GatheredBlocks>, grid, dim3(256), 0, stream, GatheredBlocks{blocks}, n, inverses);
}

void jacobi_apply(const double *inverses, std::uint32_t n, const double *v, double *out, hipStream_t stream,
                  const std::uint32_t *order) {
    hipLaunchKernelGGL(jacobi_apply, dim3((n + kJacobi - 1) / kJacobi), dim3(64), 0, stream, inverses, n, v, out,
                       order);
}

} // namespace hsbem::dev
