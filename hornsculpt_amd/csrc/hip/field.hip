// Exterior pressure from the solved boundary data, gfx950 (FP64).
//
// far_field      F_o(d) = 1/(4 pi) int exp(+ik d.(y-o)) [ik (d.n) p(y) - q(y)] dS_y,
//                times exp(-ikR)/R (panel_integrals.cuh:379-408, cuda_solver.cu:636-661).
// observer_field p(x) = D[p](x) - S[q](x) with the fixed panel rule, and the
//                adaptive four-child subdivision for panels whose centroid lies
//                within two edges of the observer (panel_integrals.cuh:298-443).
// Both reduce over panels in 256-panel tiles (one workgroup per output x tile);
// the tiles are summed in a second pass, so the result is deterministic.
#include "device.hpp"

namespace hsbem::dev {
namespace {

constexpr int kFieldTile = 256;

__device__ __forceinline__ double geo_at(const PanelView &pv, int f, std::uint32_t t) {
    return pv.geo[static_cast<std::size_t>(f) * pv.nt + t];
}

__device__ __forceinline__ Cplx nodal_pressure(const FieldArgs &a, std::uint32_t t, std::uint32_t e, const double *b) {
    Cplx p = {0, 0};
    for (int l = 0; l < 3; ++l) {
        const std::uint32_t node = a.panels.nodes[static_cast<std::size_t>(l) * a.panels.nt + t];
        const double *v = a.pressure + 2 * (static_cast<std::size_t>(e) * a.panels.nv + node);
        p.re += v[0] * b[l];
        p.im += v[1] * b[l];
    }
    return p;
}

__device__ void block_store(Cplx v, double *partials) {
    __shared__ double red[2][4];
    v.re = wave_sum(v.re);
    v.im = wave_sum(v.im);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x / 64] = v.re;
        red[1][threadIdx.x / 64] = v.im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        partials[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

__global__ __launch_bounds__(256) void far_field_tiles(FieldArgs a, std::uint32_t tiles, double *partials) {
    const std::uint32_t out = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    const std::uint32_t e = out / a.count, o = out % a.count;
    const PanelView &pv = a.panels;
    const double dx = a.points[3 * o], dy = a.points[3 * o + 1], dz = a.points[3 * o + 2];
    Cplx sum = {0, 0};
    const std::uint32_t end = min((tile + 1) * kFieldTile, pv.nt);
    for (std::uint32_t t = tile * kFieldTile + threadIdx.x; t < end; t += blockDim.x) {
        const double ox = geo_at(pv, G_V0, t) - a.origin[0], oy = geo_at(pv, G_V0 + 1, t) - a.origin[1],
                     oz = geo_at(pv, G_V0 + 2, t) - a.origin[2];
        const double nf = a.k * (dx * geo_at(pv, G_N, t) + dy * geo_at(pv, G_N + 1, t) + dz * geo_at(pv, G_N + 2, t));
        const double area = geo_at(pv, G_AREA, t);
        const double *q = a.flux + 2 * (static_cast<std::size_t>(e) * pv.nt + t);
        for (std::uint32_t i = 0; i < a.rule_count; ++i) {
            const double *pt = a.rule + 4 * i;
            const double px = ox + (geo_at(pv, G_E1, t) * pt[1] + geo_at(pv, G_E2, t) * pt[2]);
            const double py = oy + (geo_at(pv, G_E1 + 1, t) * pt[1] + geo_at(pv, G_E2 + 1, t) * pt[2]);
            const double pz = oz + (geo_at(pv, G_E1 + 2, t) * pt[1] + geo_at(pv, G_E2 + 2, t) * pt[2]);
            const double phase = a.k * (dx * px + dy * py + dz * pz);
            if (!isfinite(phase)) {
                atomicExch(a.error, ERR_DISTANCE);
                continue;
            }
            double sn, cs;
            sincos(phase, &sn, &cs);
            const double scale = 2 * area * pt[3] / 12.56637061435917295385;
            const Cplx p = nodal_pressure(a, t, e, pt);
            const Cplx integrand = {-nf * p.im - q[0], nf * p.re - q[1]};
            sum = cadd(sum, cmul({cs * scale, sn * scale}, integrand));
        }
    }
    block_store(sum, partials);
}

struct Patch {
    double b[3][3];
    double fraction;
    unsigned depth;
};

__device__ Patch child_patch(const Patch &p, int c) {
    double ab[3], ac[3], bc[3];
    for (int i = 0; i < 3; ++i) {
        ab[i] = (p.b[0][i] + p.b[1][i]) * .5;
        ac[i] = (p.b[0][i] + p.b[2][i]) * .5;
        bc[i] = (p.b[1][i] + p.b[2][i]) * .5;
    }
    Patch r;
    r.fraction = p.fraction * .25;
    r.depth = p.depth + 1;
    const double *src[4][3] = {{p.b[0], ab, ac}, {ab, p.b[1], bc}, {ac, bc, p.b[2]}, {ab, bc, ac}};
    for (int v = 0; v < 3; ++v)
        for (int i = 0; i < 3; ++i)
            r.b[v][i] = src[c][v][i];
    return r;
}

// D[p] - S[q] over one sub-triangle with the fixed rule; scale accumulates |dp|+|sq|.
__device__ Cplx patch_field(const FieldArgs &a, double x, double y, double z, std::uint32_t t, std::uint32_t e,
                            const Patch &pa, double *scale) {
    const PanelView &pv = a.panels;
    const double *q = a.flux + 2 * (static_cast<std::size_t>(e) * pv.nt + t);
    const double nx = geo_at(pv, G_N, t), ny = geo_at(pv, G_N + 1, t), nz = geo_at(pv, G_N + 2, t);
    const double area = geo_at(pv, G_AREA, t);
    Cplx sum = {0, 0};
    for (std::uint32_t i = 0; i < a.rule_count; ++i) {
        const double *pt = a.rule + 4 * i;
        double b[3];
        for (int l = 0; l < 3; ++l)
            b[l] = pa.b[0][l] * pt[0] + pa.b[1][l] * pt[1] + pa.b[2][l] * pt[2];
        const double dx = (x - geo_at(pv, G_V0, t)) - (geo_at(pv, G_E1, t) * b[1] + geo_at(pv, G_E2, t) * b[2]);
        const double dy = (y - geo_at(pv, G_V0 + 1, t)) - (geo_at(pv, G_E1 + 1, t) * b[1] + geo_at(pv, G_E2 + 1, t) * b[2]);
        const double dz = (z - geo_at(pv, G_V0 + 2, t)) - (geo_at(pv, G_E1 + 2, t) * b[1] + geo_at(pv, G_E2 + 2, t) * b[2]);
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (!(r > 0)) {
            atomicExch(a.error, ERR_DISTANCE);
            return {0, 0};
        }
        double sn, cs;
        sincos(a.k * r, &sn, &cs);
        const double s = 2 * area * pa.fraction * pt[3] / (12.56637061435917295385 * r);
        const Cplx g = {cs * s, -sn * s};
        const double nf = (dx * nx + dy * ny + dz * nz) / (r * r);
        const Cplx dual = {(g.re - g.im * a.k * r) * nf, (g.im + g.re * a.k * r) * nf};
        const Cplx p = nodal_pressure(a, t, e, b);
        const Cplx dp = cmul(dual, p), sq = cmul(g, {q[0], q[1]});
        sum = cadd(sum, {dp.re - sq.re, dp.im - sq.im});
        *scale += hypot(dp.re, dp.im) + hypot(sq.re, sq.im);
    }
    return sum;
}

__device__ Cplx adaptive_field(const FieldArgs &a, double x, double y, double z, std::uint32_t t, std::uint32_t e) {
    Patch stack[34];
    unsigned pending = 1;
    stack[0] = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, 1.0, 0};
    Cplx result = {0, 0};
    while (pending) {
        const Patch p = stack[--pending];
        double cscale = 0, fscale = 0;
        const Cplx coarse = patch_field(a, x, y, z, t, e, p, &cscale);
        Cplx fine = {0, 0};
        for (int c = 0; c < 4; ++c)
            fine = cadd(fine, patch_field(a, x, y, z, t, e, child_patch(p, c), &fscale));
        const double est = hypot(fine.re - coarse.re, fine.im - coarse.im);
        if (est <= a.near_tolerance * fmax(fscale, 1e-300)) {
            result = cadd(result, fine);
        } else {
            if (p.depth >= a.max_depth || pending + 4 > 34) {
                atomicExch(a.error, ERR_FIELD);
                return {0, 0};
            }
            for (int c = 0; c < 4; ++c)
                stack[pending++] = child_patch(p, c);
        }
    }
    return result;
}

__global__ __launch_bounds__(256) void observer_field_tiles(FieldArgs a, std::uint32_t tiles, double *partials) {
    const std::uint32_t out = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    const std::uint32_t e = out / a.count, o = out % a.count;
    const PanelView &pv = a.panels;
    const double x = a.points[3 * o], y = a.points[3 * o + 1], z = a.points[3 * o + 2];
    Cplx sum = {0, 0};
    const std::uint32_t end = min((tile + 1) * kFieldTile, pv.nt);
    const Patch whole = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, 1.0, 0};
    for (std::uint32_t t = tile * kFieldTile + threadIdx.x; t < end; t += blockDim.x) {
        const double cx = x - pv.bounds[t], cy = y - pv.bounds[pv.nt + t], cz = z - pv.bounds[2 * pv.nt + t];
        const double h = pv.bounds[4 * static_cast<std::size_t>(pv.nt) + t];
        if (cx * cx + cy * cy + cz * cz > 4 * h * h) {
            double s = 0;
            sum = cadd(sum, patch_field(a, x, y, z, t, e, whole, &s));
        } else {
            sum = cadd(sum, adaptive_field(a, x, y, z, t, e));
        }
    }
    block_store(sum, partials);
}

__global__ void finish_field(const double *partials, std::uint32_t outputs, std::uint32_t tiles, int far, double k,
                             double R, double *out, int *error) {
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= outputs)
        return;
    double re = 0, im = 0;
    for (std::uint32_t t = 0; t < tiles; ++t) {
        re += partials[2 * (static_cast<std::size_t>(i) * tiles + t)];
        im += partials[2 * (static_cast<std::size_t>(i) * tiles + t) + 1];
    }
    if (far) {
        double sn, cs;
        sincos(k * R, &sn, &cs);
        const Cplx v = cmul({re, im}, {cs, -sn});
        re = v.re / R;
        im = v.im / R;
    }
    if (!isfinite(re) || !isfinite(im))
        atomicExch(error, ERR_OVERFLOW);
    out[2 * i] = re;
    out[2 * i + 1] = im;
}

} // namespace

std::size_t field_partial_doubles(std::uint32_t nt, std::uint32_t count, std::uint32_t nrhs) {
    const std::size_t tiles = (nt + kFieldTile - 1) / kFieldTile;
    return 2 * tiles * std::max<std::size_t>(1, static_cast<std::size_t>(count) * nrhs);
}

void launch_far_field(const FieldArgs &a, double *partials, hipStream_t stream) {
    const std::uint32_t tiles = (a.panels.nt + kFieldTile - 1) / kFieldTile, outputs = a.count * a.nrhs;
    if (!outputs)
        return;
    hipLaunchKernelGGL(far_field_tiles, dim3(outputs * tiles), dim3(256), 0, stream, a, tiles, partials);
    hipLaunchKernelGGL(finish_field, dim3((outputs + 127) / 128), dim3(128), 0, stream, partials, outputs, tiles, 1, a.k,
                       a.reference_distance, a.out, a.error);
}

void launch_observer_field(const FieldArgs &a, double *partials, hipStream_t stream) {
    const std::uint32_t tiles = (a.panels.nt + kFieldTile - 1) / kFieldTile, outputs = a.count * a.nrhs;
    if (!outputs)
        return;
    hipLaunchKernelGGL(observer_field_tiles, dim3(outputs * tiles), dim3(256), 0, stream, a, tiles, partials);
    hipLaunchKernelGGL(finish_field, dim3((outputs + 127) / 128), dim3(128), 0, stream, partials, outputs, tiles, 0, a.k,
                       1.0, a.out, a.error);
}

} // namespace hsbem::dev
