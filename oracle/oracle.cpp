// ORACLE - test infrastructure only. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker.
//
// An FP64 CPU restatement of the reference's dense Burton-Miller Galerkin path:
//   space ........................ modules/bem/galerkin/space.cpp:27-107
//   rules ........................ modules/bem/galerkin/pair_quadrature.cpp:685-769
//   Sauter-Schwab sectors ........ modules/bem/galerkin/pair_quadrature.hpp:112-207
//   projection fan maps .......... modules/bem/galerkin/near_projection.hpp:120-255
//   regular qualification ........ modules/bem/galerkin/regular_qualification.hpp:17-96
//   pair kernel (FP64 pairSample)  modules/bem/cuda/panel_integrals.cuh:166-198
//   curl term, row scale, scatter  modules/bem/cuda/panel_integrals.cuh:246-269
//   rule selection per pair ...... modules/bem/cuda/cuda_solver.cu:266-356, 382-518
//   mass terms ................... modules/bem/cuda/cuda_solver.cu:519-540
//   observer field ............... modules/bem/cuda/panel_integrals.cuh:298-443, cuda_solver.cu:609-634
//   far field .................... modules/bem/cuda/panel_integrals.cuh:379-408, cuda_solver.cu:636-661
//   surface outputs .............. modules/bem/cuda/cuda_solver.cu:1800-1846
// The interaction plan (touching / near jobs / patches / all-pairs bytes) is an
// input: tests take it from the reference planner built in oracle/_ref.
// Each piece is pinned in tests/ against oracle/_ref, the BEAT CSV fixtures,
// the analytic pulsating sphere and closed forms.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numbers>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {
using C = std::complex<double>;
struct V3 {
    double x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

thread_local std::string last_error;

struct Panel {
    std::uint32_t n[3];
    V3 v[3], centroid, normal, grad[3], curl[3];
    double area, edge;
};

struct Space {
    std::vector<V3> vertices;
    std::vector<Panel> panels;
    std::vector<double> lumped;
};

Space make_space(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri) {
    Space s;
    s.vertices.resize(nv);
    for (std::uint32_t i = 0; i < nv; ++i)
        s.vertices[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    s.lumped.assign(nv, 0);
    s.panels.resize(nt);
    for (std::uint32_t t = 0; t < nt; ++t) {
        Panel &p = s.panels[t];
        for (int l = 0; l < 3; ++l) {
            p.n[l] = tri[3 * t + l];
            if (p.n[l] >= nv)
                throw std::invalid_argument("bad vertex index");
            p.v[l] = s.vertices[p.n[l]];
        }
        const V3 ab = sub(p.v[1], p.v[0]), ac = sub(p.v[2], p.v[0]), bc = sub(p.v[2], p.v[1]);
        const V3 av = cross(ab, ac);
        const double twice = std::hypot(av.x, av.y, av.z);
        p.edge = std::max({std::hypot(ab.x, ab.y, ab.z), std::hypot(ac.x, ac.y, ac.z), std::hypot(bc.x, bc.y, bc.z)});
        if (!(twice > 128 * std::numeric_limits<double>::epsilon() * p.edge * p.edge))
            throw std::invalid_argument("degenerate panel");
        p.area = .5 * twice;
        p.normal = {av.x / twice, av.y / twice, av.z / twice};
        p.centroid = {p.v[0].x + (ab.x + ac.x) / 3, p.v[0].y + (ab.y + ac.y) / 3, p.v[0].z + (ab.z + ac.z) / 3};
        const V3 opp[3] = {bc, sub(p.v[0], p.v[2]), ab};
        for (int l = 0; l < 3; ++l) {
            const V3 g = cross(p.normal, opp[l]);
            p.grad[l] = {g.x / twice, g.y / twice, g.z / twice};
            p.curl[l] = cross(p.normal, p.grad[l]);
            s.lumped[p.n[l]] += p.area / 3;
        }
    }
    return s;
}

struct TriPt {
    double b[3], w;
};
struct GaussPt {
    double x, w;
};

// Gauss-Legendre on [0,1] by Newton iteration on the three-term recurrence.
std::vector<GaussPt> gauss01(unsigned n) {
    std::vector<GaussPt> g(n);
    for (unsigned i = 0; i < (n + 1) / 2; ++i) {
        double x = std::cos(std::numbers::pi * (i + .75) / (n + .5)), d = 0;
        for (int it = 0; it < 64; ++it) {
            double p0 = 1, p1 = x;
            for (unsigned m = 2; m <= n; ++m) {
                const double p = ((2.0 * m - 1) * x * p1 - (m - 1.0) * p0) / m;
                p0 = p1;
                p1 = p;
            }
            d = n * (x * p1 - p0) / (x * x - 1);
            const double dx = p1 / d;
            x -= dx;
            if (std::abs(dx) < 1e-16)
                break;
        }
        const double w = 1 / ((1 - x * x) * d * d);
        g[i] = {(1 - x) / 2, w};
        g[n - 1 - i] = {(1 + x) / 2, w};
    }
    return g;
}

// Dunavant degree 4/6/8 orbits (weights normalised to sum 1/2) and collapsed
// Gauss tensor rules of degree 12/14.
std::vector<TriPt> tri_rule(unsigned degree) {
    std::vector<TriPt> r;
    if (degree == 12 || degree == 14) {
        const auto g = gauss01((degree + 2) / 2);
        for (const auto &u : g)
            for (const auto &v : g)
                r.push_back({{(1 - u.x) * (1 - v.x), u.x, (1 - u.x) * v.x}, u.w * v.w * (1 - u.x)});
        return r;
    }
    auto o3 = [&](double a, double b, double w) {
        r.push_back({{b, a, a}, w});
        r.push_back({{a, b, a}, w});
        r.push_back({{a, a, b}, w});
    };
    auto o6 = [&](double a, double b, double c, double w) {
        const double p[3] = {a, b, c};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (i != j)
                    r.push_back({{p[i], p[j], p[3 - i - j]}, w});
    };
    if (degree == 2) {
        o3(1.0 / 6, 2.0 / 3, 1.0 / 6);
    } else if (degree == 4) {
        o3(.445948490915965, .108103018168070, .1116907948390055);
        o3(.091576213509771, .816847572980459, .054975871827661);
    } else if (degree == 6) {
        o3(.249286745170910, .501426509658179, .0583931378631895);
        o3(.063089014491502, .873821971016996, .0254224531851035);
        o6(.310352451033785, .636502499121399, .053145049844816, .041425537809187);
    } else if (degree == 8) {
        r.push_back({{1.0 / 3, 1.0 / 3, 1.0 / 3}, .0721578038388935});
        o3(.459292588292723, .081414823414554, .0475458171336425);
        o3(.170569307751760, .658861384496480, .051608685267359);
        o3(.050547228317031, .898905543365938, .016229248811599);
        o6(.263112829634638, .728492392955404, .008394777409958, .0136151570872175);
    } else {
        throw std::invalid_argument("triangle rule degree");
    }
    return r;
}

struct PairPt {
    double x[3], y[3], w;
};

// Sauter-Schwab sector maps on {0<=t<=s<=1}^2, returning reference-triangle
// barycentrics (1-s, s-t, t) for test and trial and the 4-D Jacobian.
PairPt ss_point(int topo, unsigned sector, double r, double u, double v, double w) {
    double xs, xt, ys, yt, J;
    if (topo == 1) { // shared vertex
        xs = r, xt = r * u, ys = r * v, yt = r * v * w, J = r * r * r * v;
        if (sector == 1)
            std::swap(xs, ys), std::swap(xt, yt);
    } else if (topo == 2) { // shared edge
        J = r * r * r * u * u;
        if (sector == 0)
            xs = r, xt = r * u * w, ys = r * (1 - u * v), yt = r * u * (1 - v);
        else if (sector == 1)
            xs = r, xt = r * u, ys = r * (1 - u * v * w), yt = r * u * v * (1 - w), J *= v;
        else if (sector == 2)
            xs = r * (1 - u * v), xt = r * u * (1 - v), ys = r, yt = r * u * v * w, J *= v;
        else if (sector == 3)
            xs = r * (1 - u * v * w), xt = r * u * v * (1 - w), ys = r, yt = r * u, J *= v;
        else
            xs = r * (1 - u * v * w), xt = r * u * (1 - v * w), ys = r, yt = r * u * v, J *= v;
    } else { // coincident
        J = r * r * r * u * u * v;
        const unsigned d = sector % 3;
        if (d == 0)
            xs = r, xt = r * (1 - u + u * v), ys = r * (1 - u * v * w), yt = r * (1 - u);
        else if (d == 1)
            xs = r, xt = r * u * (1 - v + v * w), ys = r * (1 - u * v), yt = r * u * (1 - v);
        else
            xs = r * (1 - u * v * w), xt = r * u * (1 - v * w), ys = r, yt = r * u * (1 - v);
        if (sector >= 3)
            std::swap(xs, ys), std::swap(xt, yt);
    }
    return {{1 - xs, xs - xt, xt}, {1 - ys, ys - yt, yt}, J};
}

// Shared vertices first (ascending global id), then the others in local order.
struct Perm {
    int topo;
    int t[3], s[3];
};
Perm classify(const Panel &a, const Panel &b) {
    std::vector<std::uint32_t> shared;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (a.n[i] == b.n[j])
                shared.push_back(a.n[i]);
    std::sort(shared.begin(), shared.end());
    Perm p{static_cast<int>(shared.size()), {}, {}};
    int nt = 0, ns = 0;
    for (auto id : shared)
        for (int i = 0; i < 3; ++i) {
            if (a.n[i] == id)
                p.t[nt++] = i;
            if (b.n[i] == id)
                p.s[ns++] = i;
        }
    for (int i = 0; i < 3; ++i) {
        if (std::find(shared.begin(), shared.end(), a.n[i]) == shared.end())
            p.t[nt++] = i;
        if (std::find(shared.begin(), shared.end(), b.n[i]) == shared.end())
            p.s[ns++] = i;
    }
    return p;
}

struct Bounds {
    V3 c;
    double radius, edge;
    bool shape;
};
Bounds bounds_of(const Panel &p) {
    Bounds b{p.centroid, 0, p.edge, true};
    double e[3];
    for (int i = 0; i < 3; ++i) {
        const V3 d = sub(p.v[i], p.v[(i + 1) % 3]), q = sub(p.v[i], p.centroid);
        e[i] = std::hypot(d.x, d.y, d.z);
        b.radius = std::max(b.radius, std::hypot(q.x, q.y, q.z));
    }
    b.radius = std::nextafter(b.radius * (1 + 8 * std::numeric_limits<double>::epsilon()), INFINITY);
    for (int i = 0; i < 3; ++i) {
        const double a = e[i], c = e[(i + 1) % 3], d = e[(i + 2) % 3];
        b.shape = b.shape && a > 0 && c > 0 && (a * a + c * c - d * d) <= 2 * a * c * .939692620785908384;
    }
    return b;
}
unsigned qualified(const Bounds &t, const Bounds &s, double k, double tol, unsigned minimum) {
    if (!t.shape || !s.shape)
        return 0;
    const double h = std::max(t.edge, s.edge), kh = k * h;
    const V3 d = sub(t.c, s.c);
    const double d2 = dot(d, d), rs = t.radius + s.radius, budget = .2 * tol;
    auto sep = [&](double g) { return d2 >= (rs + g * h) * (rs + g * h); };
    if (minimum <= 4 && ((kh <= .5 && sep(3) && 1e-6 + 4 * 2.3e-6 <= budget) ||
                         (kh <= .25 && sep(8) && 1e-6 + 4 * 1e-7 <= budget)))
        return 4;
    if (minimum <= 6 && ((kh <= 2 && sep(1) && 1e-6 + 4 * 6e-7 <= budget) ||
                         (kh <= 1 && sep(2) && 1e-6 + 4 * 2.4e-8 <= budget)))
        return 6;
    if (minimum <= 8 && kh <= 2 && ((sep(.5) && 1e-6 + 4 * 7.5e-7 <= budget) || (sep(1) && 1e-6 + 4 * 3.5e-8 <= budget)))
        return 8;
    return 0;
}

// Patch layout shared with the tests: test[9], trial[9], fraction, degree, projection, intervals.
struct Patch {
    double t[9], s[9], f;
    unsigned degree, projection, intervals;
};

// Fan projection of the trial panel about the foot of the test point.
struct FanOuter {
    double x[3], proj[3];
    V3 rel;
    double height, h, w;
};
FanOuter fan_outer(const TriPt &o, const Patch &pa, const Panel &t, const Panel &s) {
    FanOuter f{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            f.x[i] += o.b[j] * pa.t[3 * j + i];
    const V3 e1 = sub(s.v[1], s.v[0]), e2 = sub(s.v[2], s.v[0]);
    f.rel = add(sub(t.v[0], s.v[0]), add(mul(sub(t.v[1], t.v[0]), f.x[1]), mul(sub(t.v[2], t.v[0]), f.x[2])));
    f.proj[1] = dot(f.rel, cross(e2, s.normal)) / (2 * s.area);
    f.proj[2] = dot(f.rel, cross(s.normal, e1)) / (2 * s.area);
    f.proj[0] = 1 - f.proj[1] - f.proj[2];
    f.height = dot(f.rel, s.normal);
    f.h = std::abs(f.height);
    f.w = o.w * pa.f;
    return f;
}
PairPt fan_point(const FanOuter &o, unsigned edge, const Panel &s, GaussPt ang, GaussPt rad) {
    const unsigned nx = (edge + 1) % 3, op = (edge + 2) % 3;
    const V3 a = sub(s.v[edge], s.v[0]), dir = sub(s.v[nx], s.v[edge]);
    const V3 off = add(sub(a, o.rel), mul(s.normal, o.height));
    const double l2 = dot(dir, dir), ctr = -dot(off, dir) / l2;
    const V3 perp = add(off, mul(dir, ctr));
    const double width = std::sqrt((dot(perp, perp) + o.h * o.h) / l2);
    const double lo = std::asinh(-ctr / width), hi = std::asinh((1 - ctr) / width);
    const double par = lo + (hi - lo) * ang.x, u = ctr + width * std::sinh(par);
    const double du = width * (hi - lo) * std::cosh(par);
    const V3 ray = add(off, mul(dir, u));
    const double st = std::asinh(std::sqrt(dot(ray, ray)) / o.h), den = std::sinh(st);
    const double tt = st == 0 ? rad.x : std::sinh(st * rad.x) / den;
    const double dt = st == 0 ? 1 : st * std::cosh(st * rad.x) / den;
    double bnd[3] = {0, 0, 0};
    bnd[edge] = 1 - u;
    bnd[nx] = u;
    PairPt p{};
    for (int i = 0; i < 3; ++i) {
        p.x[i] = o.x[i];
        p.y[i] = o.proj[i] + tt * (bnd[i] - o.proj[i]);
    }
    p.w = o.w * o.proj[op] * tt * dt * du * ang.w * rad.w;
    return p;
}

// Accumulated pair integrals: 9 A moments, 3 B moments and the G sum.
struct Moments {
    C a[9]{}, b[3]{}, g{};
};

// FP64 Burton-Miller sample (alpha=-i/k, exp(+iwt), G=exp(-ikr)/(4 pi r)).
void pair_sample(Moments &m, const Panel &t, const Panel &s, const double *xt, const double *xs, double w, double k,
                 bool coincident) {
    const V3 lt = add(mul(sub(t.v[1], t.v[0]), xt[1]), mul(sub(t.v[2], t.v[0]), xt[2]));
    const V3 ls = add(mul(sub(s.v[1], s.v[0]), xs[1]), mul(sub(s.v[2], s.v[0]), xs[2]));
    const V3 d = add(sub(t.v[0], s.v[0]), sub(lt, ls));
    const double r = std::sqrt(dot(d, d));
    if (!(r > 0) || !std::isfinite(r))
        throw std::runtime_error("BEM quadrature encountered zero or invalid source distance.");
    const C g = std::polar(w * 4 * t.area * s.area / (4 * std::numbers::pi * r), -k * r);
    m.g += g;
    const C radial = g * C(1, k * r); // (1+ikr) G
    const double ir2 = 1 / (r * r);
    const C dn = coincident ? C(0) : radial * (dot(d, s.normal) * ir2);   // dG/dn_y
    const C adj = coincident ? C(0) : radial * (-dot(d, t.normal) * ir2); // -dG/dn_x... sign folded below
    const double nn = dot(t.normal, s.normal);
    const C a = -dn + C(0, 1) * k * nn * g;   // -D + alpha*(-k^2 nx.ny G) part
    const C b = -g + C(0, 1) * adj / k;       // -S - alpha K'
    for (int i = 0; i < 3; ++i) {
        m.b[i] += b * xt[i];
        for (int j = 0; j < 3; ++j)
            m.a[3 * i + j] += a * (xt[i] * xs[j]);
    }
}

struct Problem {
    Space space;
    std::vector<Bounds> bounds;
    std::vector<double> row_scale;
};

// Output rows: the whole matrix (map == nullptr, leading dimension np) or a
// compact set of rows (map[row] = position or -1, leading dimension ld).
struct RowSink {
    const std::int64_t *map;
    std::size_t ld;
    std::int64_t at(std::uint32_t row) const { return map ? map[row] : std::int64_t(row); }
};

void scatter(const Problem &pr, const Moments &m, std::uint32_t ti, std::uint32_t si, double k, C *A, C *B,
             const RowSink &sink) {
    const Space &sp = pr.space;
    const Panel &t = sp.panels[ti], &s = sp.panels[si];
    for (int i = 0; i < 3; ++i) {
        const std::uint32_t row = t.n[i];
        const std::int64_t at = sink.at(row);
        if (at < 0)
            continue;
        const double rs = pr.row_scale[row];
        for (int j = 0; j < 3; ++j) {
            const double cc = dot(t.curl[i], s.curl[j]);
            const C val = m.a[3 * i + j] + C(m.g.imag() * cc / k, -m.g.real() * cc / k); // -i G curl.curl / k
            A[at + std::size_t(s.n[j]) * sink.ld] += val * rs;
        }
        B[at + std::size_t(si) * sink.ld] += m.b[i] * rs;
    }
}

struct PlanView {
    const std::uint32_t *touching; // 9 per job: test, trial, topology, perm t[3], perm s[3]
    std::uint64_t n_touching;
    const std::uint32_t *near;     // 5 per job: test, trial, first patch, count, swap
    std::uint64_t n_near;
    const double *patches;         // 22 per patch
    const std::uint8_t *whole;     // all-pairs bytes or null
    const std::uint32_t *offsets, *trials;
};

Patch unpack(const double *p) {
    Patch q;
    std::memcpy(q.t, p, 9 * sizeof(double));
    std::memcpy(q.s, p + 9, 9 * sizeof(double));
    q.f = p[18];
    q.degree = unsigned(p[19]);
    q.projection = unsigned(p[20]);
    q.intervals = unsigned(p[21]);
    return q;
}

template <class F> void parallel_for(std::size_t n, F &&f, unsigned threads) {
    std::atomic<std::size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < std::max(1u, threads); ++w)
        pool.emplace_back([&] {
            for (std::size_t i; (i = next.fetch_add(1)) < n;)
                f(i);
        });
    for (auto &t : pool)
        t.join();
}

// Assemble the row-scaled A (np x np) and B (np x nt), column-major, for one k.
// Rows are owned by test panels -> per-thread accumulation is race free when
// each thread owns whole test panels only if rows are not shared; rows ARE
// shared between panels, so accumulate per test panel into private buffers.
void assemble(const Problem &pr, const PlanView &plan, double k, double tol, unsigned regular_order,
              unsigned singular_order, C *A, C *B, const std::uint32_t *test_subset, std::uint64_t n_subset,
              unsigned threads, const RowSink &sink) {
    const Space &sp = pr.space;
    const std::size_t np = sp.vertices.size(), nt = sp.panels.size();
    const auto gs = gauss01(singular_order);
    const auto rule8 = tri_rule(8), rule14 = tri_rule(14);
    std::vector<std::vector<TriPt>> regular(9);
    for (unsigned d : {4u, 6u, 8u})
        regular[d] = tri_rule(d);
    // touching jobs and near jobs indexed by test panel for the per-test loop
    std::vector<std::vector<std::uint64_t>> touch_of(nt), near_of(nt);
    for (std::uint64_t i = 0; i < plan.n_touching; ++i)
        touch_of[plan.touching[9 * i]].push_back(i);
    for (std::uint64_t i = 0; i < plan.n_near; ++i)
        near_of[plan.near[5 * i]].push_back(i);
    std::vector<std::uint32_t> tests;
    if (test_subset)
        tests.assign(test_subset, test_subset + n_subset);
    else
        for (std::uint32_t t = 0; t < nt; ++t)
            tests.push_back(t);
    // one lock per matrix column block would be heavy; use per-thread copies of
    // the 3 rows touched by each test panel: moments are computed in parallel,
    // scattered under a mutex-free ordering by collecting per test panel.
    struct Block {
        std::uint32_t s;
        Moments m;
    };
    std::vector<std::vector<Block>> results(tests.size());
    parallel_for(
        tests.size(),
        [&](std::size_t idx) {
            const std::uint32_t ti = tests[idx];
            const Panel &t = sp.panels[ti];
            auto &out = results[idx];
            out.reserve(nt);
            std::vector<char> special(nt, 0);
            for (auto j : touch_of[ti]) {
                const std::uint32_t *job = plan.touching + 9 * j;
                const std::uint32_t si = job[1];
                special[si] = 1;
                const Panel &s = sp.panels[si];
                const int topo = int(job[2]);
                Moments m;
                for (unsigned sec = 0; sec < (topo == 3 ? 6u : topo == 2 ? 5u : 2u); ++sec)
                    for (const auto &r : gs)
                        for (const auto &u : gs)
                            for (const auto &v : gs)
                                for (const auto &w : gs) {
                                    const PairPt c = ss_point(topo, sec, r.x, u.x, v.x, w.x);
                                    double x[3], y[3];
                                    for (int l = 0; l < 3; ++l) {
                                        x[job[3 + l]] = c.x[l];
                                        y[job[6 + l]] = c.y[l];
                                    }
                                    pair_sample(m, t, s, x, y, c.w * r.w * u.w * v.w * w.w, k, topo == 3);
                                }
                out.push_back({si, m});
            }
            for (auto j : near_of[ti]) {
                const std::uint32_t *job = plan.near + 5 * j;
                const std::uint32_t si = job[1];
                special[si] = 1;
                const Panel &s = sp.panels[si];
                const bool swap = job[4] != 0;
                const Panel &pt = swap ? s : t, &ps = swap ? t : s; // planning orientation
                Moments m;
                for (std::uint32_t q = 0; q < job[3]; ++q) {
                    const Patch pa = unpack(plan.patches + 22 * std::size_t(job[2] + q));
                    const auto &rule = pa.degree == 14 ? rule14 : rule8;
                    if (!pa.projection) {
                        for (const auto &x : rule)
                            for (const auto &y : rule) {
                                PairPt p{};
                                p.w = x.w * y.w * pa.f;
                                for (int i = 0; i < 3; ++i)
                                    for (int l = 0; l < 3; ++l) {
                                        p.x[i] += x.b[l] * pa.t[3 * l + i];
                                        p.y[i] += y.b[l] * pa.s[3 * l + i];
                                    }
                                if (swap)
                                    pair_sample(m, t, s, p.y, p.x, p.w, k, false);
                                else
                                    pair_sample(m, t, s, p.x, p.y, p.w, k, false);
                            }
                    } else {
                        const unsigned order = 20 * pa.intervals;
                        const auto g20 = gauss01(20);
                        for (const auto &x : rule) {
                            const FanOuter o = fan_outer(x, pa, pt, ps);
                            for (unsigned e = 0; e < 3; ++e)
                                for (unsigned ia = 0; ia < order; ++ia)
                                    for (unsigned ir = 0; ir < order; ++ir) {
                                        GaussPt ga = g20[ia % 20], gr = g20[ir % 20];
                                        ga = {(ia / 20 + ga.x) / pa.intervals, ga.w / pa.intervals};
                                        gr = {(ir / 20 + gr.x) / pa.intervals, gr.w / pa.intervals};
                                        const PairPt p = fan_point(o, e, ps, ga, gr);
                                        if (swap)
                                            pair_sample(m, t, s, p.y, p.x, p.w, k, false);
                                        else
                                            pair_sample(m, t, s, p.x, p.y, p.w, k, false);
                                    }
                        }
                    }
                }
                out.push_back({si, m});
            }
            for (std::uint32_t si = 0; si < nt; ++si) {
                if (special[si])
                    continue;
                const Panel &s = sp.panels[si];
                bool touching = false;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        touching |= t.n[i] == s.n[j];
                if (touching)
                    continue; // covered by the touching list
                const std::vector<TriPt> *rule = nullptr;
                std::vector<TriPt> whole_rule;
                if (plan.whole) {
                    const std::size_t lo = std::min(ti, si), hi = std::max(ti, si);
                    const unsigned d = plan.whole[hi * (hi - 1) / 2 + lo];
                    if (d)
                        rule = d == 14 ? &rule14 : &rule8;
                }
                if (!rule) {
                    const unsigned d = qualified(pr.bounds[ti], pr.bounds[si], k, tol, regular_order);
                    if (!d)
                        throw std::runtime_error("BEM dense regular pair escaped its qualified adaptive plan.");
                    rule = &regular[d];
                }
                Moments m;
                for (const auto &x : *rule)
                    for (const auto &y : *rule)
                        pair_sample(m, t, s, x.b, y.b, x.w * y.w, k, false);
                out.push_back({si, m});
            }
        },
        threads);
    for (std::size_t idx = 0; idx < tests.size(); ++idx)
        for (const auto &blk : results[idx])
            scatter(pr, blk.m, tests[idx], blk.s, k, A, B, sink);
    // Galerkin mass: 1/2 M_pp in A and -alpha 1/2 M_pq = i A/(6k) in B.
    for (std::uint32_t ti : tests) {
        const Panel &t = sp.panels[ti];
        for (int i = 0; i < 3; ++i) {
            const std::uint32_t row = t.n[i];
            const std::int64_t at = sink.at(row);
            if (at < 0)
                continue;
            const double rs = pr.row_scale[row];
            for (int j = 0; j < 3; ++j)
                A[at + std::size_t(t.n[j]) * sink.ld] += rs * t.area * (i == j ? 2.0 : 1.0) / 24.0;
            B[at + std::size_t(ti) * sink.ld] += C(0, t.area / (6 * k) * rs);
        }
    }
}

// Blocked right-looking complex LU with partial pivoting (column-major, lda=n),
// trailing updates split over threads by column. Solves A X = R in place (R n x m).
void lu_solve(std::size_t n, C *a, std::size_t m, C *rhs, unsigned threads) {
    std::vector<std::size_t> piv(n);
    const std::size_t nb = 48;
    for (std::size_t k0 = 0; k0 < n; k0 += nb) {
        const std::size_t kb = std::min(nb, n - k0);
        // panel factorisation (unblocked) on columns k0..k0+kb
        for (std::size_t j = k0; j < k0 + kb; ++j) {
            std::size_t p = j;
            double best = std::abs(a[j + j * n]);
            for (std::size_t i = j + 1; i < n; ++i)
                if (std::abs(a[i + j * n]) > best)
                    best = std::abs(a[i + j * n]), p = i;
            if (!(best > 0))
                throw std::runtime_error("singular matrix");
            piv[j] = p;
            if (p != j)
                for (std::size_t c = 0; c < n; ++c)
                    std::swap(a[j + c * n], a[p + c * n]);
            const C inv = 1.0 / a[j + j * n];
            for (std::size_t i = j + 1; i < n; ++i)
                a[i + j * n] *= inv;
            for (std::size_t c = j + 1; c < k0 + kb; ++c) {
                const C f = a[j + c * n];
                for (std::size_t i = j + 1; i < n; ++i)
                    a[i + c * n] -= a[i + j * n] * f;
            }
        }
        const std::size_t rest = n - (k0 + kb);
        if (!rest)
            continue;
        // U12 = L11^{-1} A12, then A22 -= L21 U12, columns in parallel
        parallel_for(
            rest,
            [&](std::size_t cc) {
                const std::size_t c = k0 + kb + cc;
                for (std::size_t j = k0; j < k0 + kb; ++j) {
                    const C f = a[j + c * n];
                    for (std::size_t i = j + 1; i < k0 + kb; ++i)
                        a[i + c * n] -= a[i + j * n] * f;
                }
                for (std::size_t j = k0; j < k0 + kb; ++j) {
                    const C f = a[j + c * n];
                    if (f == C(0))
                        continue;
                    const C *l = a + j * n;
                    C *col = a + c * n;
                    for (std::size_t i = k0 + kb; i < n; ++i)
                        col[i] -= l[i] * f;
                }
            },
            threads);
    }
    for (std::size_t e = 0; e < m; ++e) {
        C *x = rhs + e * n;
        for (std::size_t j = 0; j < n; ++j)
            if (piv[j] != j)
                std::swap(x[j], x[piv[j]]);
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t i = j + 1; i < n; ++i)
                x[i] -= a[i + j * n] * x[j];
        for (std::size_t j = n; j-- > 0;) {
            x[j] /= a[j + j * n];
            for (std::size_t i = 0; i < j; ++i)
                x[i] -= a[i + j * n] * x[j];
        }
    }
}

// Observer field p = D[p] - S[q] on one (sub)triangle with a fixed rule.
struct FieldValue {
    C v;
    double scale;
};
FieldValue field_patch(V3 obs, const Panel &p, const double (*bary)[3], double fraction,
                       const std::vector<TriPt> &rule, const C *nodal, C q, double k) {
    FieldValue out{0, 0};
    for (const auto &x : rule) {
        double b[3];
        for (int i = 0; i < 3; ++i)
            b[i] = bary[0][i] * x.b[0] + bary[1][i] * x.b[1] + bary[2][i] * x.b[2];
        const V3 d = sub(sub(obs, p.v[0]), add(mul(sub(p.v[1], p.v[0]), b[1]), mul(sub(p.v[2], p.v[0]), b[2])));
        const double r = std::sqrt(dot(d, d));
        if (!(r > 0))
            throw std::runtime_error("BEM quadrature encountered zero or invalid source distance.");
        const C g = std::polar(2 * p.area * fraction * x.w / (4 * std::numbers::pi * r), -k * r);
        const C dual = g * C(1, k * r) * (dot(d, p.normal) / (r * r));
        const C pv = nodal[p.n[0]] * b[0] + nodal[p.n[1]] * b[1] + nodal[p.n[2]] * b[2];
        const C dp = dual * pv, sq = g * q;
        out.v += dp - sq;
        out.scale += std::abs(dp) + std::abs(sq);
    }
    return out;
}

C near_field(V3 obs, const Panel &p, const std::vector<TriPt> &rule, const C *nodal, C q, double k, double tol,
             unsigned max_depth) {
    struct Node {
        double b[3][3];
        double f;
        unsigned depth;
    };
    std::vector<Node> stack{{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, 1, 0}};
    C sum = 0;
    auto child = [](const Node &n, int c) {
        Node r;
        double ab[3], ac[3], bc[3];
        for (int i = 0; i < 3; ++i) {
            ab[i] = (n.b[0][i] + n.b[1][i]) * .5;
            ac[i] = (n.b[0][i] + n.b[2][i]) * .5;
            bc[i] = (n.b[1][i] + n.b[2][i]) * .5;
        }
        const double *v[4][3] = {{n.b[0], ab, ac}, {ab, n.b[1], bc}, {ac, bc, n.b[2]}, {ab, bc, ac}};
        for (int a = 0; a < 3; ++a)
            for (int i = 0; i < 3; ++i)
                r.b[a][i] = v[c][a][i];
        r.f = n.f * .25;
        r.depth = n.depth + 1;
        return r;
    };
    while (!stack.empty()) {
        const Node n = stack.back();
        stack.pop_back();
        const FieldValue coarse = field_patch(obs, p, n.b, n.f, rule, nodal, q, k);
        FieldValue fine{0, 0};
        Node kids[4];
        for (int c = 0; c < 4; ++c) {
            kids[c] = child(n, c);
            const FieldValue v = field_patch(obs, p, kids[c].b, kids[c].f, rule, nodal, q, k);
            fine.v += v.v;
            fine.scale += v.scale;
        }
        if (std::abs(fine.v - coarse.v) <= tol * std::max(fine.scale, 1e-300)) {
            sum += fine.v;
        } else {
            if (n.depth >= max_depth || stack.size() + 4 > 34)
                throw std::runtime_error("BEM field integration did not reach its near-field tolerance.");
            for (int c = 0; c < 4; ++c)
                stack.push_back(kids[c]);
        }
    }
    return sum;
}

template <class F> int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        last_error = e.what();
        return 1;
    }
}

Problem make_problem(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri) {
    Problem pr;
    pr.space = make_space(nv, xyz, nt, tri);
    for (const auto &p : pr.space.panels)
        pr.bounds.push_back(bounds_of(p));
    for (double m : pr.space.lumped)
        pr.row_scale.push_back(1 / m);
    return pr;
}
} // namespace

extern "C" {

const char *oracle_last_error() { return last_error.c_str(); }

int oracle_triangle_rule(unsigned degree, double *out, unsigned *count) {
    return guarded([&] {
        const auto r = tri_rule(degree);
        *count = unsigned(r.size());
        if (out)
            for (std::size_t i = 0; i < r.size(); ++i) {
                std::memcpy(out + 4 * i, r[i].b, 3 * sizeof(double));
                out[4 * i + 3] = r[i].w;
            }
    });
}

int oracle_gauss(unsigned order, double *out) {
    return guarded([&] {
        const auto g = gauss01(order);
        for (std::size_t i = 0; i < g.size(); ++i)
            out[2 * i] = g[i].x, out[2 * i + 1] = g[i].w;
    });
}

// Sauter-Schwab points of one touching pair in original local vertex order
// (7 doubles per point: test[3], trial[3], weight), sector-major then r,u,v,w.
int oracle_touching_points(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri,
                           std::uint32_t a, std::uint32_t b, unsigned order, double *out, unsigned *count) {
    return guarded([&] {
        const Space s = make_space(nv, xyz, nt, tri);
        const Perm pm = classify(s.panels[a], s.panels[b]);
        if (pm.topo == 0)
            throw std::invalid_argument("pair is not touching");
        const auto g = gauss01(order);
        const unsigned sectors = pm.topo == 3 ? 6 : pm.topo == 2 ? 5 : 2;
        unsigned n = 0;
        for (unsigned sec = 0; sec < sectors; ++sec)
            for (const auto &r : g)
                for (const auto &u : g)
                    for (const auto &v : g)
                        for (const auto &w : g) {
                            const PairPt c = ss_point(pm.topo, sec, r.x, u.x, v.x, w.x);
                            if (out) {
                                double *o = out + 7 * n;
                                for (int l = 0; l < 3; ++l) {
                                    o[pm.t[l]] = c.x[l];
                                    o[3 + pm.s[l]] = c.y[l];
                                }
                                o[6] = c.w * r.w * u.w * v.w * w.w;
                            }
                            ++n;
                        }
        *count = n;
    });
}

// The 36 G / grad G P1 moments of one disjoint patch in the planner's moment
// convention (pair_quadrature.cpp:163-197), built from this oracle's own patch
// point maps. Lets tests pin the point maps against estimatePairPatchCpu.
int oracle_patch_moments(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri,
                         std::uint32_t a, std::uint32_t b, const double *patch, double k, double *out) {
    return guarded([&] {
        const Space s = make_space(nv, xyz, nt, tri);
        const Panel &t = s.panels[a], &q = s.panels[b];
        const Patch pa = unpack(patch);
        const double scale = std::max(t.edge, q.edge);
        C m[36]{};
        auto acc = [&](const PairPt &p) {
            const V3 lt = add(mul(sub(t.v[1], t.v[0]), p.x[1]), mul(sub(t.v[2], t.v[0]), p.x[2]));
            const V3 ls = add(mul(sub(q.v[1], q.v[0]), p.y[1]), mul(sub(q.v[2], q.v[0]), p.y[2]));
            const V3 d = add(sub(t.v[0], q.v[0]), sub(lt, ls));
            const double r = std::sqrt(dot(d, d));
            const C g = std::polar(1 / r, -k * r), dg = -g * C(1, k * r) * (scale / (r * r));
            const C kv[4] = {g, dg * d.x, dg * d.y, dg * d.z};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int c = 0; c < 4; ++c)
                        m[9 * c + 3 * i + j] += kv[c] * (p.w * p.x[i] * p.y[j]);
        };
        const auto rule = tri_rule(pa.degree == 14 ? 14 : 8);
        if (!pa.projection) {
            for (const auto &x : rule)
                for (const auto &y : rule) {
                    PairPt p{};
                    p.w = x.w * y.w * pa.f;
                    for (int i = 0; i < 3; ++i)
                        for (int l = 0; l < 3; ++l) {
                            p.x[i] += x.b[l] * pa.t[3 * l + i];
                            p.y[i] += y.b[l] * pa.s[3 * l + i];
                        }
                    acc(p);
                }
        } else {
            const unsigned order = 20 * pa.intervals;
            const auto g20 = gauss01(20);
            for (const auto &x : rule) {
                const FanOuter o = fan_outer(x, pa, t, q);
                for (unsigned e = 0; e < 3; ++e)
                    for (unsigned ia = 0; ia < order; ++ia)
                        for (unsigned ir = 0; ir < order; ++ir) {
                            GaussPt ga = g20[ia % 20], gr = g20[ir % 20];
                            ga = {(ia / 20 + ga.x) / pa.intervals, ga.w / pa.intervals};
                            gr = {(ir / 20 + gr.x) / pa.intervals, gr.w / pa.intervals};
                            acc(fan_point(o, e, q, ga, gr));
                        }
            }
        }
        for (int i = 0; i < 36; ++i)
            out[2 * i] = m[i].real(), out[2 * i + 1] = m[i].imag();
    });
}

// Unscaled 3x3 A block (curl term included) and 3 B moments of one ordered
// disjoint pair integrated over one tensor patch with the FP64 pair kernel.
int oracle_pair_block(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri,
                      std::uint32_t a, std::uint32_t b, const double *patch, double k, double *block, double *rhs) {
    return guarded([&] {
        const Space s = make_space(nv, xyz, nt, tri);
        const Panel &t = s.panels[a], &q = s.panels[b];
        const Patch pa = unpack(patch);
        const auto rule = tri_rule(pa.degree == 14 ? 14 : 8);
        Moments m;
        for (const auto &x : rule)
            for (const auto &y : rule) {
                PairPt p{};
                p.w = x.w * y.w * pa.f;
                for (int i = 0; i < 3; ++i)
                    for (int l = 0; l < 3; ++l) {
                        p.x[i] += x.b[l] * pa.t[3 * l + i];
                        p.y[i] += y.b[l] * pa.s[3 * l + i];
                    }
                pair_sample(m, t, q, p.x, p.y, p.w, k, false);
            }
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
                const double cc = dot(t.curl[i], q.curl[j]);
                const C v = m.a[3 * i + j] + C(m.g.imag() * cc / k, -m.g.real() * cc / k);
                block[2 * (3 * i + j)] = v.real();
                block[2 * (3 * i + j) + 1] = v.imag();
            }
            rhs[2 * i] = m.b[i].real();
            rhs[2 * i + 1] = m.b[i].imag();
        }
    });
}

// Row-scaled A (np x np) and B (np x nt), complex interleaved, column-major.
// test_subset (optional) restricts assembly to the listed test panels.
int oracle_assemble(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri, double k,
                    double tolerance, unsigned regular_order, unsigned singular_order, const std::uint32_t *touching,
                    std::uint64_t n_touching, const std::uint32_t *near, std::uint64_t n_near, const double *patches,
                    const std::uint8_t *whole, const std::uint32_t *test_subset, std::uint64_t n_subset,
                    unsigned threads, double *A, double *B) {
    return guarded([&] {
        const Problem pr = make_problem(nv, xyz, nt, tri);
        const PlanView plan{touching, n_touching, near, n_near, patches, whole, nullptr, nullptr};
        assemble(pr, plan, k, tolerance, regular_order, singular_order, reinterpret_cast<C *>(A),
                 reinterpret_cast<C *>(B), test_subset, n_subset, threads, RowSink{nullptr, nv});
    });
}

// Rows `rows[0..n_rows)` of the same A and B, compact (column-major with leading
// dimension n_rows): only the test panels that touch one of the rows are
// integrated, and nothing of the full matrices is formed, so rows of meshes far
// beyond a dense host matrix stay cheap.
int oracle_assemble_rows(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri, double k,
                         double tolerance, unsigned regular_order, unsigned singular_order,
                         const std::uint32_t *touching, std::uint64_t n_touching, const std::uint32_t *near,
                         std::uint64_t n_near, const double *patches, const std::uint8_t *whole,
                         const std::uint32_t *rows, std::uint32_t n_rows, unsigned threads, double *A, double *B) {
    return guarded([&] {
        std::vector<std::int64_t> map(nv, -1);
        for (std::uint32_t r = 0; r < n_rows; ++r) {
            if (rows[r] >= nv || map[rows[r]] >= 0)
                throw std::invalid_argument("rows must be distinct vertex indices");
            map[rows[r]] = r;
        }
        std::vector<std::uint32_t> tests;
        for (std::uint32_t t = 0; t < nt; ++t)
            if (map[tri[3 * t]] >= 0 || map[tri[3 * t + 1]] >= 0 || map[tri[3 * t + 2]] >= 0)
                tests.push_back(t);
        const Problem pr = make_problem(nv, xyz, nt, tri);
        const PlanView plan{touching, n_touching, near, n_near, patches, whole, nullptr, nullptr};
        assemble(pr, plan, k, tolerance, regular_order, singular_order, reinterpret_cast<C *>(A),
                 reinterpret_cast<C *>(B), tests.data(), tests.size(), threads, RowSink{map.data(), n_rows});
    });
}

// In-place dense solve of A X = R (complex, column-major, A n x n, R n x m).
int oracle_lu_solve(std::uint64_t n, double *A, std::uint64_t m, double *R, unsigned threads) {
    return guarded([&] { lu_solve(n, reinterpret_cast<C *>(A), m, reinterpret_cast<C *>(R), threads); });
}

// Observer pressures (nobs per excitation) for nodal pressures p (np x nrhs) and
// DP0 fluxes q (nt x nrhs), at one rule degree; far != 0 evaluates the
// asymptotic far field about origin normalised to reference distance R.
int oracle_field(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri, double k,
                 unsigned degree, const double *pressure, const double *flux, std::uint32_t nrhs,
                 const double *observers, std::uint32_t nobs, int far, const double *origin, double R,
                 double near_tolerance, unsigned max_depth, double *out) {
    return guarded([&] {
        const Space s = make_space(nv, xyz, nt, tri);
        const auto rule = tri_rule(degree);
        const C *p = reinterpret_cast<const C *>(pressure), *q = reinterpret_cast<const C *>(flux);
        C *o = reinterpret_cast<C *>(out);
        for (std::uint32_t e = 0; e < nrhs; ++e)
            for (std::uint32_t j = 0; j < nobs; ++j) {
                const V3 x{observers[3 * j], observers[3 * j + 1], observers[3 * j + 2]};
                C sum = 0;
                if (far) {
                    const V3 org{origin[0], origin[1], origin[2]};
                    const V3 rel = sub(x, org);
                    const double len = std::hypot(rel.x, rel.y, rel.z);
                    const V3 dir = mul(rel, 1 / len);
                    for (std::uint32_t t = 0; t < nt; ++t) {
                        const Panel &pn = s.panels[t];
                        const double nf = k * dot(dir, pn.normal);
                        for (const auto &pt : rule) {
                            const V3 pos = add(sub(pn.v[0], org), add(mul(sub(pn.v[1], pn.v[0]), pt.b[1]),
                                                                      mul(sub(pn.v[2], pn.v[0]), pt.b[2])));
                            const C pv = p[e * nv + pn.n[0]] * pt.b[0] + p[e * nv + pn.n[1]] * pt.b[1] +
                                         p[e * nv + pn.n[2]] * pt.b[2];
                            const C integrand = C(0, nf) * pv - q[e * nt + t];
                            sum += std::polar(2 * pn.area * pt.w / (4 * std::numbers::pi), k * dot(dir, pos)) * integrand;
                        }
                    }
                    sum *= std::polar(1 / R, -k * R);
                } else {
                    for (std::uint32_t t = 0; t < nt; ++t) {
                        const Panel &pn = s.panels[t];
                        const V3 off = sub(x, pn.centroid);
                        const double whole[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
                        if (dot(off, off) > 4 * pn.edge * pn.edge)
                            sum += field_patch(x, pn, whole, 1, rule, p + e * nv, q[e * nt + t], k).v;
                        else
                            sum += near_field(x, pn, rule, p + e * nv, q[e * nt + t], k, near_tolerance, max_depth);
                    }
                }
                o[e * nobs + j] = sum;
            }
    });
}
}
