// Burton-Miller Galerkin assembly kernels (FP64), gfx950.
//
// Replaces the reference kernels assembleRegular / assembleSingular /
// assembleNear / addMass (modules/bem/cuda/cuda_solver.cu:266-540) with:
//   regular_tiles   one workgroup per (test cluster, trial cluster) of up to 64
//                   spatially sorted panels; panel geometry and rules in LDS; every
//                   3x3 pair block accumulated in an LDS node tile (ds_add_f64) and
//                   flushed once per tile to HBM, so global atomics drop from 18 per
//                   pair to about one per tile entry.
//   touching_jobs   one wavefront per touching pair: Sauter-Schwab (u,v,w) tuples
//                   over lanes with the radial coordinate innermost (every sector map
//                   is linear in r), DPP/shuffle reduction, one lane scatters.
//   near_jobs       one wavefront per near job, patch points over lanes.
//   mass            1/2 M_pp into A, -alpha 1/2 M_pq into the flux columns.
#include "../host/plan.hpp"
#include "../host/qualification.hpp"
#include "../host/rules.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace hsbem::dev {
namespace {

constexpr int kTilePanels = 64;                // panels per cluster
constexpr int kTileNodes = 64;                 // distinct pressure nodes per cluster
constexpr int kTileStride = 2 * kTileNodes + 2; // doubles per trial-node row (padded)
constexpr int kTileThreads = 512;

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

// Two-stage tensor contraction of the pair integral over rule x rule:
// inner sums over trial points carry the trial basis, the outer sum the test basis.
__device__ __forceinline__ void integrate_tensor(Moments &m, const double *gt, const double *gs, const double *rule,
                                                 int n, double fourAA, double k, double inv_k, double nn, int *error) {
    const double knn = k * nn, wscale = fourAA * 0.07957747154594766788; // 4 At As / (4 pi)
    const double bx = gt[G_V0] - gs[G_V0], by = gt[G_V0 + 1] - gs[G_V0 + 1], bz = gt[G_V0 + 2] - gs[G_V0 + 2];
    for (int xi = 0; xi < n; ++xi) {
        const double *xp = rule + 4 * xi;
        const double x1 = xp[1], x2 = xp[2], xw = xp[3] * wscale;
        const double ltx = gt[G_E1] * x1 + gt[G_E2] * x2, lty = gt[G_E1 + 1] * x1 + gt[G_E2 + 1] * x2,
                     ltz = gt[G_E1 + 2] * x1 + gt[G_E2 + 2] * x2;
        Cplx ga[3] = {{0, 0}, {0, 0}, {0, 0}}, gg = {0, 0}, gb = {0, 0};
        for (int yi = 0; yi < n; ++yi) {
            const double *yp = rule + 4 * yi;
            const double y0 = yp[0], y1 = yp[1], y2 = yp[2];
            const double lsx = gs[G_E1] * y1 + gs[G_E2] * y2, lsy = gs[G_E1 + 1] * y1 + gs[G_E2 + 1] * y2,
                         lsz = gs[G_E1 + 2] * y1 + gs[G_E2 + 2] * y2;
            const Sample s = bm_sample(bx + (ltx - lsx), by + (lty - lsy), bz + (ltz - lsz), xw * yp[3], k, inv_k,
                                       knn, gs[G_N], gs[G_N + 1], gs[G_N + 2], gt[G_N], gt[G_N + 1], gt[G_N + 2],
                                       false, error);
            ga[0] = cadd(ga[0], cscale(s.a, y0));
            ga[1] = cadd(ga[1], cscale(s.a, y1));
            ga[2] = cadd(ga[2], cscale(s.a, y2));
            gg = cadd(gg, s.g);
            gb = cadd(gb, s.b);
        }
        m.g = cadd(m.g, gg);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double xb = xp[i];
            m.b[i] = cadd(m.b[i], cscale(gb, xb));
#pragma unroll
            for (int j = 0; j < 3; ++j)
                m.a[3 * i + j] = cadd(m.a[3 * i + j], cscale(ga[j], xb));
        }
    }
}

__global__ __launch_bounds__(kTileThreads) void regular_tiles(AssemblyArgs args, ClusterView cl, NearView near,
                                                              const double *rules) {
    __shared__ double tile[kTileNodes * kTileStride];
    __shared__ double sgt[G_COUNT][kTilePanels], sgs[G_COUNT][kTilePanels];
    __shared__ double sbt[B_COUNT][kTilePanels], sbs[B_COUNT][kTilePanels];
    __shared__ std::uint32_t sidt[kTilePanels], sids[kTilePanels];
    __shared__ std::uint32_t snt[3][kTilePanels], sns[3][kTilePanels];
    __shared__ std::uint8_t slt[3][kTilePanels], sls[3][kTilePanels], shpt[kTilePanels], shps[kTilePanels];
    __shared__ std::int32_t sfc[kTilePanels];
    __shared__ double srule[4 * kRuleTableCount];

    const std::uint32_t ct = blockIdx.y, cs = blockIdx.x;
    const std::uint32_t t0 = cl.panel_offsets[ct], nT = cl.panel_offsets[ct + 1] - t0;
    const std::uint32_t s0 = cl.panel_offsets[cs], nS = cl.panel_offsets[cs + 1] - s0;
    const std::uint32_t a0 = cl.node_offsets[ct], nA = cl.node_offsets[ct + 1] - a0;
    const std::uint32_t b0 = cl.node_offsets[cs], nB = cl.node_offsets[cs + 1] - b0;
    const PanelView &pv = args.panels;
    const int tid = threadIdx.x;

    for (int e = tid; e < kTileNodes * kTileStride; e += kTileThreads)
        tile[e] = 0.0;
    for (int e = tid; e < 4 * kRuleTableCount; e += kTileThreads)
        srule[e] = rules[e];
    for (int e = tid; e < static_cast<int>(G_COUNT * kTilePanels); e += kTileThreads) {
        const int f = e / kTilePanels, p = e % kTilePanels;
        if (p < static_cast<int>(nT))
            sgt[f][p] = g_at(pv.geo, f, pv.nt, cl.panels[t0 + p]);
        if (p < static_cast<int>(nS))
            sgs[f][p] = g_at(pv.geo, f, pv.nt, cl.panels[s0 + p]);
    }
    for (int e = tid; e < static_cast<int>(B_COUNT * kTilePanels); e += kTileThreads) {
        const int f = e / kTilePanels, p = e % kTilePanels;
        if (p < static_cast<int>(nT))
            sbt[f][p] = pv.bounds[static_cast<std::size_t>(f) * pv.nt + cl.panels[t0 + p]];
        if (p < static_cast<int>(nS))
            sbs[f][p] = pv.bounds[static_cast<std::size_t>(f) * pv.nt + cl.panels[s0 + p]];
    }
    for (int p = tid; p < kTilePanels; p += kTileThreads) {
        if (p < static_cast<int>(nT)) {
            const std::uint32_t id = cl.panels[t0 + p];
            sidt[p] = id;
            shpt[p] = pv.shape[id];
            for (int l = 0; l < 3; ++l) {
                snt[l][p] = pv.nodes[static_cast<std::size_t>(l) * pv.nt + id];
                slt[l][p] = cl.local[static_cast<std::size_t>(l) * pv.nt + id];
            }
        }
        if (p < static_cast<int>(nS)) {
            const std::uint32_t id = cl.panels[s0 + p];
            sids[p] = id;
            shps[p] = pv.shape[id];
            sfc[p] = args.flux_column[id];
            for (int l = 0; l < 3; ++l) {
                sns[l][p] = pv.nodes[static_cast<std::size_t>(l) * pv.nt + id];
                sls[l][p] = cl.local[static_cast<std::size_t>(l) * pv.nt + id];
            }
        }
    }
    __syncthreads();

    const double k = args.k, inv_k = 1.0 / args.k;
    const std::uint32_t pairs = nT * nS;
    unsigned long long counted = 0;
    for (std::uint32_t p = tid; p < pairs; p += kTileThreads) {
        const std::uint32_t lt = p % nT, ls = p / nT;
        bool touching = false;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                touching |= snt[i][lt] == sns[j][ls];
        if (touching)
            continue;
        const std::uint32_t ti = sidt[lt], si = sids[ls];
        const std::uint32_t nb = near.offsets[ti];
        if (contains_sorted(near.trials + nb, near.offsets[ti + 1] - nb, si))
            continue;
        int offset = -1, count = 0;
        if (near.whole) {
            const unsigned d = near.whole[unordered_pair_index(ti, si)];
            if (d) {
                offset = d == 14 ? kRule14 : kRule8;
                count = d == 14 ? 64 : 16;
            }
        }
        if (offset < 0) {
            PanelBounds bt, bs;
            bt.center = {sbt[B_C][lt], sbt[B_C + 1][lt], sbt[B_C + 2][lt]};
            bt.radius = sbt[B_R][lt];
            bt.max_edge = sbt[B_H][lt];
            bt.qualified_shape = shpt[lt];
            bs.center = {sbs[B_C][ls], sbs[B_C + 1][ls], sbs[B_C + 2][ls]};
            bs.radius = sbs[B_R][ls];
            bs.max_edge = sbs[B_H][ls];
            bs.qualified_shape = shps[ls];
            const unsigned d = qualified_regular_degree(bt, bs, k, args.tolerance, args.regular_order);
            if (!d) {
                atomicExch(args.error, ERR_ESCAPED);
                continue;
            }
            offset = d == 4 ? kRule4 : d == 6 ? kRule6 : kRule8;
            count = d == 4 ? 6 : d == 6 ? 12 : 16;
        }
        double gt[G_COUNT], gs[G_COUNT];
#pragma unroll
        for (int f = 0; f < G_COUNT; ++f) {
            gt[f] = sgt[f][lt];
            gs[f] = sgs[f][ls];
        }
        const double nn = gt[G_N] * gs[G_N] + gt[G_N + 1] * gs[G_N + 1] + gt[G_N + 2] * gs[G_N + 2];
        Moments m;
        zero(m);
        integrate_tensor(m, gt, gs, srule + 4 * offset, count, 4.0 * gt[G_AREA] * gs[G_AREA], k, inv_k, nn,
                         args.error);
        counted += static_cast<unsigned long long>(count) * count;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int a = slt[i][lt];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int b = sls[j][ls];
                const double cc = gt[G_C0 + 3 * i] * gs[G_C0 + 3 * j] + gt[G_C0 + 3 * i + 1] * gs[G_C0 + 3 * j + 1] +
                                  gt[G_C0 + 3 * i + 2] * gs[G_C0 + 3 * j + 2];
                const Cplx v = {m.a[3 * i + j].re + m.g.im * cc / k, m.a[3 * i + j].im - m.g.re * cc / k};
                double *slot = tile + b * kTileStride + 2 * a;
                unsafeAtomicAdd(slot, v.re);
                unsafeAtomicAdd(slot + 1, v.im);
            }
        }
        const std::int32_t fc = sfc[ls];
        if (fc >= 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const std::uint32_t row = snt[i][lt];
                atomic_add_c(args.BD + 2 * (static_cast<std::size_t>(fc) * pv.nv + row),
                             cscale(m.b[i], args.row_scale[row]));
            }
        }
    }
    if (args.samples) {
        unsigned long long c = counted;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            c += __shfl_xor(c, off, 64);
        if ((tid & 63) == 0 && c)
            atomicAdd(args.samples, c);
    }
    __syncthreads();
    for (std::uint32_t e = tid; e < nA * nB; e += kTileThreads) {
        const std::uint32_t a = e % nA, b = e / nA;
        const double *slot = tile + b * kTileStride + 2 * a;
        const std::uint32_t row = cl.nodes[a0 + a], col = cl.nodes[b0 + b];
        const double rs = args.row_scale[row];
        const double re = slot[0], im = slot[1];
        if (re != 0.0 || im != 0.0)
            atomic_add_c(args.A + 2 * (static_cast<std::size_t>(col) * pv.nv + row), {re * rs, im * rs});
    }
}

__device__ __forceinline__ void scatter_pair(const AssemblyArgs &args, const Moments &m, std::uint32_t ti,
                                             std::uint32_t si) {
    const PanelView &pv = args.panels;
    const double k = args.k;
    const std::int32_t fc = args.flux_column[si];
    for (int i = 0; i < 3; ++i) {
        const std::uint32_t row = pv.nodes[static_cast<std::size_t>(i) * pv.nt + ti];
        const double rs = args.row_scale[row];
        for (int j = 0; j < 3; ++j) {
            const std::uint32_t col = pv.nodes[static_cast<std::size_t>(j) * pv.nt + si];
            double cc = 0;
            for (int c = 0; c < 3; ++c)
                cc += g_at(pv.geo, G_C0 + 3 * i + c, pv.nt, ti) * g_at(pv.geo, G_C0 + 3 * j + c, pv.nt, si);
            const Cplx v = {m.a[3 * i + j].re + m.g.im * cc / k, m.a[3 * i + j].im - m.g.re * cc / k};
            atomic_add_c(args.A + 2 * (static_cast<std::size_t>(col) * pv.nv + row), cscale(v, rs));
        }
        if (fc >= 0)
            atomic_add_c(args.BD + 2 * (static_cast<std::size_t>(fc) * pv.nv + row), cscale(m.b[i], rs));
    }
}

__device__ __forceinline__ void load_geo(const PanelView &pv, std::uint32_t t, double *g) {
#pragma unroll
    for (int f = 0; f < G_COUNT; ++f)
        g[f] = g_at(pv.geo, f, pv.nt, t);
}

__global__ __launch_bounds__(256) void touching_jobs(AssemblyArgs args, const TouchingJob *jobs, std::uint32_t njobs,
                                                     const double *gauss, unsigned order) {
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
    const double nn = gt[G_N] * gs[G_N] + gt[G_N + 1] * gs[G_N + 1] + gt[G_N + 2] * gs[G_N + 2];
    const double wscale = 4.0 * gt[G_AREA] * gs[G_AREA] * 0.07957747154594766788;
    const double inv_k = 1.0 / args.k, knn = args.k * nn;
    const double bx = gt[G_V0] - gs[G_V0], by = gt[G_V0 + 1] - gs[G_V0 + 1], bz = gt[G_V0 + 2] - gs[G_V0 + 2];
    const unsigned per = order * order * order;
    const unsigned total = sector_count(topo) * per;
    const int at = tj.permutation.test[0], as = tj.permutation.trial[0];
    Moments m;
    zero(m);
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
            const Sample s = bm_sample(r * dx, r * dy, r * dz, c.weight * r * r * r * gauss[2 * ri + 1] * wscale, args.k,
                                       inv_k, knn, gs[G_N], gs[G_N + 1], gs[G_N + 2], gt[G_N], gt[G_N + 1],
                                       gt[G_N + 2], coincident, args.error);
            add_sample(m, s, xt, xs);
        }
    }
    wave_reduce(m);
    if (lane == 0)
        scatter_pair(args, m, tj.test, tj.trial);
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

__global__ __launch_bounds__(256) void near_jobs(AssemblyArgs args, const NearJob *jobs, std::uint32_t njobs,
                                                 const PairPatch *patches, const double *rules) {
    const std::uint32_t job = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (job >= njobs)
        return;
    const NearJob nj = jobs[job];
    const PanelView &pv = args.panels;
    double gt[G_COUNT], gs[G_COUNT];
    load_geo(pv, nj.test, gt);
    load_geo(pv, nj.trial, gs);
    const double nn = gt[G_N] * gs[G_N] + gt[G_N + 1] * gs[G_N + 1] + gt[G_N + 2] * gs[G_N + 2];
    const double wscale = 4.0 * gt[G_AREA] * gs[G_AREA] * 0.07957747154594766788;
    const double inv_k = 1.0 / args.k, knn = args.k * nn;
    const double bx = gt[G_V0] - gs[G_V0], by = gt[G_V0 + 1] - gs[G_V0 + 1], bz = gt[G_V0 + 2] - gs[G_V0 + 2];
    const bool swap = nj.swap != 0;
    const FanPanel ft = fan_of(gt), fs = fan_of(gs);
    const TrianglePoint *r8 = reinterpret_cast<const TrianglePoint *>(rules + 4 * kRule8);
    const TrianglePoint *r14 = reinterpret_cast<const TrianglePoint *>(rules + 4 * kRule14);
    Moments m;
    zero(m);
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
            const Sample s = bm_sample(bx + (ltx - lsx), by + (lty - lsy), bz + (ltz - lsz), p.weight * wscale,
                                       args.k, inv_k, knn, gs[G_N], gs[G_N + 1], gs[G_N + 2], gt[G_N], gt[G_N + 1],
                                       gt[G_N + 2], false, args.error);
            add_sample(m, s, p.test, p.trial);
        }
    }
    wave_reduce(m);
    if (lane == 0)
        scatter_pair(args, m, nj.test, nj.trial);
}

__global__ void mass_terms(AssemblyArgs args) {
    const std::uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const PanelView &pv = args.panels;
    if (t >= pv.nt)
        return;
    const double area = g_at(pv.geo, G_AREA, pv.nt, t);
    const std::int32_t fc = args.flux_column[t];
    for (int i = 0; i < 3; ++i) {
        const std::uint32_t row = pv.nodes[static_cast<std::size_t>(i) * pv.nt + t];
        const double rs = args.row_scale[row];
        for (int j = 0; j < 3; ++j) {
            const std::uint32_t col = pv.nodes[static_cast<std::size_t>(j) * pv.nt + t];
            unsafeAtomicAdd(args.A + 2 * (static_cast<std::size_t>(col) * pv.nv + row),
                            rs * area * (i == j ? 2.0 : 1.0) / 24.0);
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

} // namespace

void launch_regular(const AssemblyLaunch &L, hipStream_t stream) {
    if (L.clusters.count)
        hipLaunchKernelGGL(regular_tiles, dim3(L.clusters.count, L.clusters.count), dim3(kTileThreads), 0, stream,
                           L.args, L.clusters, L.near, L.rules);
}

void launch_touching(const AssemblyLaunch &L, hipStream_t stream) {
    if (L.touching_count)
        hipLaunchKernelGGL(touching_jobs, dim3((L.touching_count + 3) / 4), dim3(256), 0, stream, L.args, L.touching,
                           L.touching_count, L.gauss, L.singular_order);
}

void launch_near(const AssemblyLaunch &L, hipStream_t stream) {
    if (L.near_count)
        hipLaunchKernelGGL(near_jobs, dim3((L.near_count + 3) / 4), dim3(256), 0, stream, L.args, L.near_jobs,
                           L.near_count, L.patches, L.rules);
}

void launch_mass(const AssemblyLaunch &L, hipStream_t stream) {
    hipLaunchKernelGGL(mass_terms, dim3((L.args.panels.nt + 255) / 256), dim3(256), 0, stream, L.args);
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
