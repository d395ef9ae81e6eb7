// C ABI (include/hsbem.h) over MeshSolver. Exceptions map to status codes the
// way the reference throws them (std::invalid_argument / std::runtime_error).
#include "../../include/hsbem.h"

#include "autotune.hpp"
#include "fmm_operator.hpp"
#include "host/acoustic_analysis.hpp"
#include "host/frequency_schedule.hpp"
#include "host/fmm_levels.hpp"
#include "host/fmm_rotation.hpp"
#include "host/fmm_scaled.hpp"
#include "host/fmm_transfer.hpp"
#include "host/support_near_plan.hpp"
#include "host/support_tree.hpp"
#include "server.hpp"
#include "solver.hpp"
#include "host/mesh_check.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <optional>
#include <string>

struct hsbem_context {
    hsbem::Context ctx;
};
struct hsbem_mesh {
    hsbem_context *context;
    std::shared_ptr<hsbem::MeshSolver> solver; // shared with serving tickets in flight
    std::uint64_t hash;
};
struct hsbem_near_operator {
    hsbem::NearOperator op;
    std::uint32_t nv;
};
struct hsbem_fmm_operator {
    std::shared_ptr<hsbem::MeshSolver> solver; // keeps the mesh's device data alive
    hsbem::Context *ctx;
    bool graph{};
    std::unique_ptr<hsbem::FmmOperator> op;
    hsbem::GmresWorkspace gmres;
    hsbem::DeviceBuffer<double> x, y;
};
struct hsbem_plan_cache {
    hsbem::TranslationPlanCache cache;
    explicit hsbem_plan_cache(std::size_t bytes) : cache(bytes) {}
};
struct hsbem_server {
    std::unique_ptr<hsbem::Server> server;
};
struct hsbem_result {
    hsbem::Result result;
};
struct hsbem_plan {
    hsbem::InteractionPlan plan;
};

namespace {
thread_local std::string last_error;

template <class F> int guarded(F &&f) {
    try {
        f();
        return HSBEM_OK;
    } catch (const hsbem::Cancelled &e) {
        last_error = e.what();
        return HSBEM_CANCELLED;
    } catch (const hsbem::DeviceError &e) {
        last_error = e.what();
        return HSBEM_DEVICE_ERROR;
    } catch (const std::invalid_argument &e) {
        last_error = e.what();
        return HSBEM_INVALID_ARGUMENT;
    } catch (const std::exception &e) {
        last_error = e.what();
        return HSBEM_RUNTIME_ERROR;
    } catch (...) {
        last_error = "unknown error";
        return HSBEM_RUNTIME_ERROR;
    }
}

void require(bool ok, const char *what) {
    if (!ok)
        throw std::invalid_argument(what);
}

hsbem::Numerics to_numerics(const hsbem_numerics &n) {
    hsbem::Numerics o;
    o.quadrature_order = n.quadrature_order;
    o.singular_order = n.singular_order;
    o.near_ratio = n.near_distance_ratio;
    o.near_tolerance = n.near_relative_tolerance;
    o.near_depth = n.near_maximum_depth;
    o.residual_tolerance = n.relative_residual_tolerance;
    o.refinement_steps = n.maximum_refinement_steps;
    o.solver = n.solver;
    o.gmres_restart = n.gmres_restart;
    o.gmres_max_iterations = n.gmres_maximum_iterations;
    o.precision = n.precision;
    o.plan_threads = n.plan_threads;
    o.preconditioner = n.preconditioner;
    o.acceleration = n.acceleration;
    o.fmm_leaf_capacity = n.fmm_leaf_capacity;
    o.fmm_separation = n.fmm_separation_ratio;
    o.fmm_maximum_iterations = n.fmm_maximum_iterations;
    o.fmm_restart = n.fmm_restart_length;
    o.fmm_block_width = n.fmm_block_width;
    return o;
}

hsbem::Request to_request(const hsbem_request &rr) {
    const hsbem_request *r = &rr;
    require(r->frequency_count == 0 || r->frequencies, "null frequencies");
    require(r->excitation_count == 0 || r->excitation_offsets, "null excitation offsets");
    require(r->observer_count == 0 || r->observers, "null observers");
    hsbem::Request q;
    q.density = r->density;
    q.sound_speed = r->sound_speed;
    q.frequencies.assign(r->frequencies, r->frequencies + r->frequency_count);
    if (r->planning_frequency_count) {
        require(r->planning_frequencies, "null planning frequencies");
        q.planning_frequencies.assign(r->planning_frequencies, r->planning_frequencies + r->planning_frequency_count);
    }
    for (uint32_t e = 0; e < r->excitation_count; ++e) {
        hsbem::Excitation ex;
        for (uint32_t j = r->excitation_offsets[e]; j < r->excitation_offsets[e + 1]; ++j)
            ex.faces.emplace_back(r->face_ids[j], std::complex<double>(r->velocities[2 * j], r->velocities[2 * j + 1]));
        q.excitations.push_back(std::move(ex));
    }
    for (uint32_t i = 0; i < r->observer_count; ++i)
        q.observers.push_back({r->observers[3 * i], r->observers[3 * i + 1], r->observers[3 * i + 2]});
    q.far_field = r->far_field != 0;
    q.far_origin = {r->far_origin[0], r->far_origin[1], r->far_origin[2]};
    q.far_reference_distance = r->far_reference_distance;
    q.numerics = to_numerics(r->numerics);
    return q;
}

const hsbem::ExcitationResult &excitation(const hsbem_result *r, std::uint32_t f, std::uint32_t e) {
    require(r && f < r->result.frequencies.size() && e < r->result.frequencies[f].excitations.size(),
            "result index out of range");
    return r->result.frequencies[f].excitations[e];
}

void copy_complex(const std::vector<std::complex<double>> &v, double *out) {
    for (std::size_t i = 0; i < v.size(); ++i) {
        out[2 * i] = v[i].real();
        out[2 * i + 1] = v[i].imag();
    }
}

std::uint64_t fnv1a(std::uint64_t h, const void *data, std::size_t bytes) {
    const auto *p = static_cast<const unsigned char *>(data);
    for (std::size_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 1099511628211ULL;
    }
    return h;
}
std::vector<double> serialize_tree(const hsbem::SupportTree &t, std::uint32_t nv, std::uint32_t nt) {
    std::vector<double> b;
    auto put = [&](double v) { b.push_back(v); };
    auto box = [&](const hsbem::SupportBox &x) {
        for (double v : {x.lower.x, x.lower.y, x.lower.z, x.upper.x, x.upper.y, x.upper.z, x.max_edge})
            put(v);
    };
    const auto &pp = t.pressure_pressure, &pf = t.pressure_flux;
    for (double v : {double(t.nodes.size()), double(t.levels.size()), double(pp.near.size()), double(pp.far.size()),
                     double(pf.near.size()), double(pf.far.size()), double(pp.far_levels.size()),
                     double(pf.far_levels.size()), double(nv), double(nt), double(pp.represented_near),
                     double(pp.represented_far), double(pf.represented_near), double(pf.represented_far)})
        put(v);
    for (const auto &n : t.nodes) {
        put(n.center.x), put(n.center.y), put(n.center.z), put(n.half_width);
        for (auto c : n.coords)
            put(c);
        put(n.level), put(n.parent);
        for (auto c : n.children)
            put(c);
        put(n.first_pressure), put(n.pressure_count), put(n.first_flux), put(n.flux_count);
        box(n.pressure_box);
        box(n.flux_box);
        put(n.pressure_radius), put(n.flux_radius), put(n.active);
    }
    for (const auto *v : {&t.pressure_order, &t.flux_order, &t.pressure_leaf, &t.flux_leaf})
        for (auto x : *v)
            put(x);
    for (const auto *l : {&pp.near, &pp.far, &pf.near, &pf.far})
        for (const auto &p : *l)
            put(p.target), put(p.source);
    for (const auto *l : {&pp.far_levels, &pf.far_levels})
        for (const auto &e : *l)
            put(e.target_level), put(e.source_level), put(double(e.interactions)), put(double(e.represented_pairs));
    for (const auto &l : t.levels) {
        put(double(l.nodes.size()));
        for (auto id : l.nodes)
            put(id);
        put(l.max_radius), put(l.max_active_radius);
        for (auto v : {l.pressure_outgoing, l.flux_outgoing, l.pressure_incoming, l.pp_far, l.pf_far, l.pp_offsets,
                       l.pf_offsets, l.union_offsets, l.pp_distance_classes, l.pf_distance_classes,
                       l.pp_polar_classes, l.pf_polar_classes, l.union_polar_classes})
            put(double(v));
        put(double(l.squared_offsets.size()));
        for (auto v : l.squared_offsets)
            put(double(v));
    }
    return b;
}
} // namespace

namespace {

// Host vector in, device operator, host vector out (the context stream).
template <class F> void fmm_host_call(hsbem_fmm_operator *o, const double *in, double *out, F &&f) {
    require(o && in && out, "null FMM operator argument");
    hsbem::hip_check(hipSetDevice(o->ctx->device), "set device");
    const std::size_t n = 2 * static_cast<std::size_t>(o->op->size());
    hipStream_t s = o->ctx->stream;
    o->x.upload(std::vector<double>(in, in + n), s);
    o->y.resize(n);
    f(o->x.data(), o->y.data(), s);
    hsbem::hip_check(hipMemcpyAsync(out, o->y.data(), n * sizeof(double), hipMemcpyDeviceToHost, s), "FMM download");
    hsbem::hip_check(hipStreamSynchronize(s), "FMM sync");
}

} // namespace

extern "C" {

void hsbem_numerics_default(hsbem_numerics *n) {
    if (!n)
        return;
    n->quadrature_order = 4;
    n->singular_order = 4;
    n->near_distance_ratio = 1.0;
    n->near_relative_tolerance = 1e-4;
    n->near_maximum_depth = 6;
    n->relative_residual_tolerance = 5e-5;
    n->maximum_refinement_steps = 3;
    n->solver = HSBEM_SOLVER_LU;
    n->gmres_restart = 50;
    n->gmres_maximum_iterations = 1000;
    n->precision = HSBEM_PRECISION_FP64;
    n->plan_threads = 0;
    n->preconditioner = HSBEM_PRECONDITIONER_BLOCK_JACOBI;
    n->acceleration = HSBEM_ACCELERATION_DENSE;
    n->fmm_leaf_capacity = 16;
    n->fmm_separation_ratio = 1.5;
    n->fmm_maximum_iterations = 300;
    n->fmm_restart_length = 40;
    n->fmm_block_width = 32;
}

const char *hsbem_last_error(void) { return last_error.c_str(); }

int hsbem_device_description(char *buffer, size_t length) {
    return guarded([&] {
        require(buffer && length, "null buffer");
        int device = 0;
        hsbem::hip_check(hipGetDevice(&device), "device query");
        hipDeviceProp_t prop{};
        hsbem::hip_check(hipGetDeviceProperties(&prop, device), "device properties");
        const std::string s = std::string(prop.name) + " (" + prop.gcnArchName + ", " +
                              std::to_string(prop.multiProcessorCount) + " CUs, " +
                              std::to_string(prop.totalGlobalMem >> 20) + " MiB)";
        std::strncpy(buffer, s.c_str(), length - 1);
        buffer[length - 1] = 0;
    });
}

int hsbem_tuning_report(char *buffer, size_t length, size_t *needed) {
    return guarded([&] {
        require(buffer && length, "null buffer");
        const std::string s = hsbem::autotune::records_json();
        if (needed)
            *needed = s.size() + 1;
        std::strncpy(buffer, s.c_str(), length - 1);
        buffer[length - 1] = 0;
    });
}

int hsbem_context_create(int32_t device, hsbem_context **out) {
    return guarded([&] {
        require(out, "null context output");
        auto c = std::make_unique<hsbem_context>();
        c->ctx.device = device;
        hsbem::hip_check(hipSetDevice(device), "set device");
        hsbem::hip_check(hipStreamCreateWithFlags(&c->ctx.stream, hipStreamNonBlocking), "stream create");
        *out = c.release();
    });
}

int hsbem_context_destroy(hsbem_context *c) {
    return guarded([&] {
        if (!c)
            return;
        (void)hipStreamSynchronize(c->ctx.stream);
        (void)hipStreamDestroy(c->ctx.stream);
        delete c;
    });
}

int hsbem_mesh_create(hsbem_context *c, const double *vertices, uint32_t nv, const uint32_t *triangles,
                      const uint32_t *face_ids, const uint8_t *boundary, uint32_t nt, hsbem_mesh **out) {
    return guarded([&] {
        require(c && out && vertices && triangles && nv && nt, "invalid mesh arguments");
        hsbem::SurfaceMesh m;
        m.vertices.resize(nv);
        for (uint32_t i = 0; i < nv; ++i)
            m.vertices[i] = {vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]};
        m.triangles.resize(nt);
        for (uint32_t t = 0; t < nt; ++t) {
            m.triangles[t].vertices = {triangles[3 * t], triangles[3 * t + 1], triangles[3 * t + 2]};
            m.triangles[t].face_id = face_ids ? face_ids[t] : 0;
            m.triangles[t].boundary =
                boundary && boundary[t] == HSBEM_BOUNDARY_DRIVEN ? hsbem::Boundary::driven : hsbem::Boundary::wall;
        }
        std::uint64_t h = 14695981039346656037ULL;
        h = fnv1a(h, vertices, sizeof(double) * 3 * nv);
        h = fnv1a(h, triangles, sizeof(uint32_t) * 3 * nt);
        if (face_ids)
            h = fnv1a(h, face_ids, sizeof(uint32_t) * nt);
        if (boundary)
            h = fnv1a(h, boundary, nt);
        hsbem::hip_check(hipSetDevice(c->ctx.device), "set device");
        auto mesh = std::make_unique<hsbem_mesh>();
        mesh->context = c;
        mesh->hash = h;
        mesh->solver = std::make_shared<hsbem::MeshSolver>(&c->ctx, std::move(m));
        *out = mesh.release();
    });
}

int hsbem_mesh_validate(const double *vertices, uint32_t nv, const uint32_t *triangles, uint32_t nt) {
    return guarded([&] {
        require(vertices && triangles && nv && nt, "invalid mesh arguments");
        hsbem::SurfaceMesh m;
        m.vertices.resize(nv);
        const hsbem::Vec3 origin{vertices[0], vertices[1], vertices[2]};
        for (uint32_t i = 0; i < nv; ++i) {
            m.vertices[i] = hsbem::vsub({vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]}, origin);
            if (!std::isfinite(m.vertices[i].x) || !std::isfinite(m.vertices[i].y) || !std::isfinite(m.vertices[i].z))
                throw std::invalid_argument("mesh vertex coordinates must be finite");
        }
        m.triangles.resize(nt);
        for (uint32_t t = 0; t < nt; ++t)
            m.triangles[t].vertices = {triangles[3 * t], triangles[3 * t + 1], triangles[3 * t + 2]};
        hsbem::validate_closed_exterior(m);
    });
}

int hsbem_mesh_destroy(hsbem_mesh *m) {
    return guarded([&] { delete m; });
}

int hsbem_mesh_info(const hsbem_mesh *m, uint64_t *hash, uint32_t *nv, uint32_t *nt, double *area) {
    return guarded([&] {
        require(m, "null mesh");
        if (hash)
            *hash = m->hash;
        if (nv)
            *nv = m->solver->nv();
        if (nt)
            *nt = m->solver->nt();
        if (area)
            *area = m->solver->space().surface_area;
    });
}

int hsbem_mesh_cache_info(hsbem_mesh *m, uint64_t *cached_plans, uint64_t *resident_bytes) {
    return guarded([&] {
        require(m, "null mesh");
        if (cached_plans)
            *cached_plans = m->solver->cached_plans();
        if (resident_bytes)
            *resident_bytes = m->solver->resident_bytes();
    });
}

int hsbem_solve(hsbem_mesh *m, const hsbem_request *r, hsbem_progress_fn progress, void *user,
                const volatile int32_t *cancel, hsbem_result **out) {
    return guarded([&] {
        require(m && r && out, "invalid solve arguments");
        const hsbem::Request q = to_request(*r);
        hsbem::hip_check(hipSetDevice(m->context->ctx.device), "set device");
        auto res = std::make_unique<hsbem_result>();
        res->result = m->solver->solve(q, progress, user, cancel);
        *out = res.release();
    });
}

int hsbem_server_create(hsbem_context *c, const hsbem_server_options *options, hsbem_server **out) {
    return guarded([&] {
        require(c && out, "invalid server arguments");
        hsbem::ServerOptions o;
        if (options) {
            o.slots = options->slots;
            o.memory_budget = options->memory_budget_bytes;
        }
        auto srv = std::make_unique<hsbem_server>();
        srv->server = std::make_unique<hsbem::Server>(&c->ctx, o);
        *out = srv.release();
    });
}

int hsbem_server_destroy(hsbem_server *s) {
    return guarded([&] { delete s; });
}

int hsbem_submit(hsbem_server *s, hsbem_mesh *m, const hsbem_request *r, uint64_t *ticket) {
    return guarded([&] {
        require(s && m && r && ticket, "invalid submit arguments");
        *ticket = s->server->submit(m->solver, to_request(*r));
    });
}

int hsbem_poll(hsbem_server *s, uint64_t ticket, int32_t wait_ms, hsbem_result **out) {
    int pending = 0;
    const int status = guarded([&] {
        require(s && out, "invalid poll arguments");
        *out = nullptr;
        auto res = std::make_unique<hsbem_result>();
        if (s->server->poll(ticket, wait_ms, &res->result))
            *out = res.release();
        else
            pending = 1;
    });
    return status == HSBEM_OK && pending ? HSBEM_PENDING : status;
}

int hsbem_cancel(hsbem_server *s, uint64_t ticket) {
    return guarded([&] {
        require(s, "null server");
        s->server->cancel(ticket);
    });
}

int hsbem_server_stats(const hsbem_server *s, hsbem_server_statistics *out) {
    return guarded([&] {
        require(s && out, "invalid stats arguments");
        const hsbem::ServerStats st = s->server->stats();
        out->submitted = st.submitted;
        out->completed = st.completed;
        out->failed = st.failed;
        out->points = st.points;
        out->running = st.running;
        out->queued = st.queued;
        out->max_running = st.max_running;
        out->budget_bytes = st.budget;
        out->held_bytes = st.held_bytes;
        out->peak_held_bytes = st.peak_held_bytes;
        out->resident_bytes = st.resident_bytes;
        out->peak_total_bytes = st.peak_total_bytes;
    });
}

int hsbem_result_served_latency(const hsbem_result *r, uint32_t f, double *milliseconds) {
    return guarded([&] {
        require(r && milliseconds && f < r->result.frequencies.size(), "result index out of range");
        *milliseconds = r->result.frequencies[f].served_ms;
    });
}

int hsbem_result_destroy(hsbem_result *r) {
    return guarded([&] { delete r; });
}

int hsbem_result_sizes(const hsbem_result *r, uint32_t *nf, uint32_t *ne, uint32_t *nt, uint32_t *nv, uint32_t *no) {
    return guarded([&] {
        require(r, "null result");
        if (nf)
            *nf = static_cast<uint32_t>(r->result.frequencies.size());
        if (ne)
            *ne = r->result.nexc;
        if (nt)
            *nt = r->result.nt;
        if (nv)
            *nv = r->result.nv;
        if (no)
            *no = r->result.nobs;
    });
}

int hsbem_result_surface_pressure(const hsbem_result *r, uint32_t f, uint32_t e, double *out) {
    return guarded([&] { copy_complex(excitation(r, f, e).surface_pressure, out); });
}
int hsbem_result_pressure_trace(const hsbem_result *r, uint32_t f, uint32_t e, double *out) {
    return guarded([&] { copy_complex(excitation(r, f, e).pressure_trace, out); });
}
int hsbem_result_normal_derivative(const hsbem_result *r, uint32_t f, uint32_t e, double *out) {
    return guarded([&] { copy_complex(excitation(r, f, e).normal_derivative, out); });
}
int hsbem_result_observer_pressure(const hsbem_result *r, uint32_t f, uint32_t e, double *out) {
    return guarded([&] { copy_complex(excitation(r, f, e).observer_pressure, out); });
}

int hsbem_result_impedances(const hsbem_result *r, uint32_t f, uint32_t e, uint32_t *count, uint32_t *face_ids,
                            double *area, double *mean_pressure, double *volume_velocity, double *impedance) {
    return guarded([&] {
        const auto &x = excitation(r, f, e);
        require(count, "null count");
        *count = static_cast<uint32_t>(x.impedances.size());
        for (std::size_t i = 0; i < x.impedances.size(); ++i) {
            const auto &z = x.impedances[i];
            if (face_ids)
                face_ids[i] = z.face_id;
            if (area)
                area[i] = z.area;
            if (mean_pressure)
                mean_pressure[2 * i] = z.mean_pressure.real(), mean_pressure[2 * i + 1] = z.mean_pressure.imag();
            if (volume_velocity)
                volume_velocity[2 * i] = z.volume_velocity.real(), volume_velocity[2 * i + 1] = z.volume_velocity.imag();
            if (impedance)
                impedance[2 * i] = z.impedance.real(), impedance[2 * i + 1] = z.impedance.imag();
        }
    });
}

int hsbem_result_residual(const hsbem_result *r, uint32_t f, uint32_t e, double *residual, uint32_t *iterations,
                          uint32_t *refinements) {
    return guarded([&] {
        const auto &x = excitation(r, f, e);
        if (residual)
            *residual = x.residual;
        if (iterations)
            *iterations = x.iterations;
        if (refinements)
            *refinements = x.refinements;
    });
}

int hsbem_result_timings(const hsbem_result *r, uint32_t f, double *ms) {
    return guarded([&] {
        require(r && f < r->result.frequencies.size() && ms, "result index out of range");
        const auto &fr = r->result.frequencies[f];
        for (int i = 0; i < 6; ++i)
            ms[i] = fr.ms[i];
        ms[6] = fr.regular_kernel_ms;
        ms[7] = fr.touching_kernel_ms;
        ms[8] = fr.near_kernel_ms;
        ms[9] = fr.regular_samples;
        ms[10] = fr.touching_samples;
        ms[11] = fr.near_samples;
    });
}

int hsbem_result_fmm(const hsbem_result *r, uint32_t f, double *out) {
    return guarded([&] {
        require(r && f < r->result.frequencies.size() && out, "result index out of range");
        const auto &fr = r->result.frequencies[f];
        out[0] = fr.fmm_audit;
        out[1] = fr.fmm_apply_ms;
        out[2] = fr.fmm_corrected ? 1 : 0;
        out[3] = static_cast<double>(fr.fmm_plane_wave_levels);
        out[4] = static_cast<double>(fr.fmm_m2l_links);
        out[5] = static_cast<double>(fr.fmm_near_nnz);
        out[6] = static_cast<double>(fr.fmm_device_bytes);
        out[7] = fr.fmm_direct_audit ? 1 : 0;
        out[8] = fr.fmm_quadrature_discrepancy;
        out[9] = fr.fmm_endpoint_degree;
        out[10] = fr.fmm_promotions;
        out[11] = fr.fmm_verification_seed;
        out[12] = fr.fmm_fp32 ? 1 : 0;
    });
}

int hsbem_result_peak_device_bytes(const hsbem_result *r, uint64_t *bytes) {
    return guarded([&] {
        require(r && bytes, "null result");
        *bytes = r->result.peak_bytes;
    });
}

int hsbem_plan_create(hsbem_mesh *m, const hsbem_numerics *n, double k_max, hsbem_plan **out) {
    return guarded([&] {
        require(m && n && out, "invalid plan arguments");
        auto p = std::make_unique<hsbem_plan>();
        // HSBEM_PLAN_ESTIMATOR=device: the device patch estimates (the fast multipole
        // form's near operator); else host threads (the dense form's plan)
        const char *estimator = std::getenv("HSBEM_PLAN_ESTIMATOR");
        p->plan = m->solver->plan(to_numerics(*n), k_max, nullptr, estimator && std::string(estimator) == "device");
        *out = p.release();
    });
}

int hsbem_plan_build(const double *vertices, uint32_t nv, const uint32_t *triangles, uint32_t nt,
                     const hsbem_numerics *n, double k_max, hsbem_plan **out) {
    return guarded([&] {
        require(vertices && triangles && nv && nt && n && out, "invalid plan arguments");
        hsbem::SurfaceMesh m;
        m.vertices.resize(nv);
        const hsbem::Vec3 origin{vertices[0], vertices[1], vertices[2]};
        for (uint32_t i = 0; i < nv; ++i)
            m.vertices[i] = hsbem::vsub({vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]}, origin);
        m.triangles.resize(nt);
        for (uint32_t t = 0; t < nt; ++t)
            m.triangles[t].vertices = {triangles[3 * t], triangles[3 * t + 1], triangles[3 * t + 2]};
        const auto space = hsbem::build_space(m);
        const auto num = to_numerics(*n);
        hsbem::QuadratureSettings q;
        q.regular_order = num.quadrature_order;
        q.singular_order = num.singular_order;
        q.near_distance_ratio = num.near_ratio;
        q.near_tolerance = num.near_tolerance;
        q.near_max_depth = num.near_depth;
        auto p = std::make_unique<hsbem_plan>();
        // HSBEM_PLAN_BATCHED=1: the batched refinement loop on the host estimator (the same
        // plan; a check of the loop the device estimator feeds)
        const char *batched = std::getenv("HSBEM_PLAN_BATCHED");
        if (batched && std::string(batched) == "1") {
            const auto host = hsbem::host_patch_estimator(space, k_max, num.plan_threads);
            p->plan = hsbem::make_interaction_plan(space, q, k_max, num.plan_threads, &host);
        } else {
            p->plan = hsbem::make_interaction_plan(space, q, k_max, num.plan_threads);
        }
        *out = p.release();
    });
}

int hsbem_support_tree_build(const double *vertices, uint32_t nv, const uint32_t *triangles, uint32_t nt,
                             uint32_t leaf_capacity, uint32_t maximum_depth, double separation, double near_ratio,
                             double *blob, uint64_t *blob_size) {
    return guarded([&] {
        require(vertices && triangles && nv && nt && blob_size, "invalid support tree arguments");
        hsbem::SurfaceMesh m;
        m.vertices.resize(nv);
        for (uint32_t i = 0; i < nv; ++i)
            m.vertices[i] = {vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]};
        m.triangles.resize(nt);
        for (uint32_t t = 0; t < nt; ++t)
            m.triangles[t].vertices = {triangles[3 * t], triangles[3 * t + 1], triangles[3 * t + 2]};
        hsbem::SupportTreeSettings settings;
        settings.leaf_capacity = leaf_capacity;
        settings.maximum_depth = maximum_depth;
        settings.separation = separation;
        settings.near_ratio = near_ratio;
        const auto blob_v = serialize_tree(hsbem::make_support_tree(hsbem::build_space(m), settings), nv, nt);
        if (blob) {
            require(*blob_size >= blob_v.size(), "support tree buffer too small");
            std::copy(blob_v.begin(), blob_v.end(), blob);
        }
        *blob_size = blob_v.size();
    });
}

int hsbem_support_near_plan_build(const double *vertices, uint32_t nv, const uint32_t *triangles,
                                  const uint32_t *face_ids, const uint8_t *boundary, uint32_t nt,
                                  const uint32_t *active_faces, uint32_t nfaces, uint32_t leaf_capacity,
                                  uint32_t maximum_depth, double separation, double near_ratio, double *blob,
                                  uint64_t *blob_size) {
    return guarded([&] {
        require(vertices && triangles && face_ids && boundary && nv && nt && blob_size && (active_faces || !nfaces),
                "invalid near plan arguments");
        hsbem::SurfaceMesh m;
        m.vertices.resize(nv);
        for (uint32_t i = 0; i < nv; ++i)
            m.vertices[i] = {vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]};
        m.triangles.resize(nt);
        for (uint32_t t = 0; t < nt; ++t) {
            m.triangles[t].vertices = {triangles[3 * t], triangles[3 * t + 1], triangles[3 * t + 2]};
            m.triangles[t].face_id = face_ids[t];
            m.triangles[t].boundary = boundary[t] == HSBEM_BOUNDARY_DRIVEN ? hsbem::Boundary::driven : hsbem::Boundary::wall;
        }
        const auto space = hsbem::build_space(m);
        hsbem::SupportTreeSettings settings;
        settings.leaf_capacity = leaf_capacity;
        settings.maximum_depth = maximum_depth;
        settings.separation = separation;
        settings.near_ratio = near_ratio;
        const auto tree = hsbem::make_support_tree(space, settings);
        const auto plan =
            hsbem::make_support_near_plan(space, tree, std::vector<std::uint32_t>(active_faces, active_faces + nfaces));
        const auto jobs = hsbem::support_near_jobs(space, tree, plan);
        std::vector<double> b;
        for (double v : {double(plan.pressure.row_offsets.size()), double(plan.pressure.columns.size()),
                         double(plan.flux.columns.size()), double(plan.node_panel_offsets.size()),
                         double(plan.support_panels.size()), double(plan.flux_block_indices.size()), double(jobs.size())})
            b.push_back(v);
        for (const auto *v : {&plan.pressure.row_offsets, &plan.pressure.columns, &plan.flux.row_offsets,
                              &plan.flux.columns, &plan.panel_face_column, &plan.node_panel_offsets,
                              &plan.support_panels, &plan.flux_block_offsets, &plan.flux_block_indices})
            b.insert(b.end(), v->begin(), v->end());
        for (const auto &j : jobs)
            for (double v : {double(j.test), double(j.source), double(j.pressure_mask), double(j.flux_mask)})
                b.push_back(v);
        if (blob) {
            require(*blob_size >= b.size(), "near plan buffer too small");
            std::copy(b.begin(), b.end(), blob);
        }
        *blob_size = b.size();
    });
}

int hsbem_support_expansion_plan_build(const double *vertices, uint32_t nv, const uint32_t *triangles,
                                       const uint32_t *face_ids, const uint8_t *boundary, uint32_t nt,
                                       const uint32_t *active_faces, uint32_t nfaces, uint32_t leaf_capacity,
                                       uint32_t maximum_depth, double separation, double near_ratio, double k,
                                       double seed, double *blob, uint64_t *blob_size) {
    return guarded([&] {
        require(vertices && triangles && face_ids && boundary && nv && nt && blob_size && (active_faces || !nfaces),
                "invalid expansion plan arguments");
        hsbem::SurfaceMesh m;
        m.vertices.resize(nv);
        for (uint32_t i = 0; i < nv; ++i)
            m.vertices[i] = {vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]};
        m.triangles.resize(nt);
        for (uint32_t t = 0; t < nt; ++t)
            m.triangles[t].vertices = {triangles[3 * t], triangles[3 * t + 1], triangles[3 * t + 2]};
        const auto space = hsbem::build_space(m);
        hsbem::SupportTreeSettings settings;
        settings.leaf_capacity = leaf_capacity;
        settings.maximum_depth = maximum_depth;
        settings.separation = separation;
        settings.near_ratio = near_ratio;
        const auto tree = hsbem::make_support_tree(space, settings);
        std::vector<std::uint32_t> column(nt, UINT32_MAX);
        for (uint32_t t = 0; t < nt; ++t)
            for (uint32_t f = 0; f < nfaces; ++f)
                if (face_ids[t] == active_faces[f] && boundary[t] == HSBEM_BOUNDARY_DRIVEN) {
                    column[t] = f;
                    break;
                }
        const auto plan = hsbem::make_support_expansion_plan(tree, column, k, seed);
        std::vector<double> b{double(plan.levels.size()), double(plan.active.size())};
        for (const auto &l : plan.levels) {
            const auto &e = l.expansion;
            for (double v : {double(static_cast<std::uint32_t>(e.representation)), e.active_radius, e.electrical_radius,
                             e.accuracy_seed, double(e.signature_order), double(e.translation_order),
                             double(e.quadrature_order), double(e.spherical_coefficients), double(e.angular_samples),
                             l.source_separation, l.target_separation, l.policy_separation, l.max_active_radius,
                             l.minimum_far_distance, l.diagonal_condition})
                b.push_back(v);
        }
        for (auto a : plan.active)
            b.push_back(a);
        if (blob) {
            require(*blob_size >= b.size(), "expansion plan buffer too small");
            std::copy(b.begin(), b.end(), blob);
        }
        *blob_size = b.size();
    });
}

int hsbem_near_operator_create(hsbem_mesh *m, const hsbem_numerics *n, double k_max, double k, uint32_t leaf_capacity,
                               uint32_t maximum_depth, double separation, double near_ratio, const uint32_t *faces,
                               uint32_t nfaces, const double *x, hsbem_near_operator **out) {
    return guarded([&] {
        require(m && n && out && (faces || !nfaces), "invalid near operator arguments");
        require(std::isfinite(k) && k > 0 && std::isfinite(k_max) && k_max > 0, "invalid wavenumber");
        hsbem::SupportTreeSettings ts;
        ts.leaf_capacity = leaf_capacity;
        ts.maximum_depth = maximum_depth;
        ts.separation = separation;
        ts.near_ratio = near_ratio;
        hsbem::hip_check(hipSetDevice(m->context->ctx.device), "set device");
        auto op = std::make_unique<hsbem_near_operator>();
        op->nv = m->solver->nv();
        op->op = m->solver->near_operator(to_numerics(*n), k_max, k, ts, std::vector<std::uint32_t>(faces, faces + nfaces), x);
        *out = op.release();
    });
}

void hsbem_fmm_settings_default(hsbem_fmm_settings *s) {
    if (!s)
        return;
    const hsbem::FmmSettings d;
    s->leaf_capacity = d.tree.leaf_capacity;
    s->maximum_depth = d.tree.maximum_depth;
    s->separation = d.tree.separation;
    s->near_ratio = d.tree.near_ratio;
    s->accuracy_seed = d.accuracy_seed;
    s->endpoint_degree = d.endpoint_degree;
}

int hsbem_fmm_operator_create(hsbem_mesh *m, const hsbem_numerics *n, const hsbem_fmm_settings *settings, double k_max,
                              double k, const uint32_t *faces, uint32_t nfaces, hsbem_fmm_operator **out) {
    return guarded([&] {
        require(m && n && settings && out && (faces || !nfaces), "invalid FMM operator arguments");
        require(std::isfinite(k) && k > 0 && std::isfinite(k_max) && k_max > 0, "invalid wavenumber");
        hsbem::FmmSettings fs;
        fs.tree.leaf_capacity = settings->leaf_capacity;
        fs.tree.maximum_depth = settings->maximum_depth;
        fs.tree.separation = settings->separation;
        fs.tree.near_ratio = settings->near_ratio;
        fs.accuracy_seed = settings->accuracy_seed;
        fs.endpoint_degree = settings->endpoint_degree;
        require(fs.endpoint_degree == 0 || fs.endpoint_degree == 4 || fs.endpoint_degree == 6 ||
                    fs.endpoint_degree == 8 || fs.endpoint_degree == 12 || fs.endpoint_degree == 14,
                "FMM endpoint degree must be 0, 4, 6, 8, 12 or 14");
        fs.block_width = n->fmm_block_width; // the preconditioner's blocks follow the numerics
        hsbem::hip_check(hipSetDevice(m->context->ctx.device), "set device");
        auto op = std::make_unique<hsbem_fmm_operator>();
        op->solver = m->solver;
        op->ctx = &m->context->ctx;
        op->op = m->solver->fmm_operator(to_numerics(*n), k_max, k, fs, std::vector<std::uint32_t>(faces, faces + nfaces));
        *out = op.release();
    });
}


int hsbem_fmm_operator_apply(hsbem_fmm_operator *o, const double *x, double *y) {
    return guarded([&] {
        fmm_host_call(o, x, y, [&](const double *dx, double *dy, hipStream_t s) {
            if (o->graph)
                o->op->apply_graph(dx, dy, s);
            else
                o->op->apply(dx, dy, s);
        });
    });
}

int hsbem_fmm_operator_apply_fp32(hsbem_fmm_operator *o, const double *x, double *y) {
    return guarded([&] {
        require(o && o->op->fp32_available(), "the FP32-bulk apply needs spherical levels of order at most 48");
        fmm_host_call(o, x, y, [&](const double *dx, double *dy, hipStream_t s) { o->op->apply_fp32(dx, dy, s); });
    });
}

int hsbem_fmm_operator_apply_direct(hsbem_fmm_operator *o, const double *x, double *y) {
    return guarded([&] {
        fmm_host_call(o, x, y, [&](const double *dx, double *dy, hipStream_t s) { o->op->apply_direct(dx, dy, s); });
    });
}

int hsbem_fmm_operator_precondition(hsbem_fmm_operator *o, const double *v, double *z) {
    return guarded([&] {
        fmm_host_call(o, v, z, [&](const double *dv, double *dz, hipStream_t s) { o->op->precondition(dv, dz, s); });
    });
}

int hsbem_fmm_operator_solve(hsbem_fmm_operator *o, const double *b, double *x, uint32_t restart,
                             uint32_t maximum_iterations, double tolerance, int32_t precondition, double *achieved,
                             uint32_t *iterations) {
    return guarded([&] {
        require(achieved && iterations && restart >= 2 && maximum_iterations >= 1 && tolerance > 0,
                "invalid GMRES arguments");
        fmm_host_call(o, b, x, [&](const double *db, double *dx, hipStream_t s) {
            hsbem::KrylovOperator kop;
            const hsbem::FmmOperator *op = o->op.get();
            if (o->graph)
                kop.apply = [op](const double *v, double *w, hipStream_t st) { op->apply_graph(v, w, st); };
            else
                kop.apply = [op](const double *v, double *w, hipStream_t st) { op->apply(v, w, st); };
            if (precondition)
                kop.precondition = [op](const double *v, double *z, hipStream_t st) { op->precondition(v, z, st); };
            hsbem::gmres(s, kop, op->size(), db, dx, restart, maximum_iterations, tolerance, achieved, iterations,
                         nullptr, o->gmres);
        });
    });
}

int hsbem_fmm_operator_time_apply_fp32(hsbem_fmm_operator *o, uint32_t repetitions, double *milliseconds) {
    return guarded([&] {
        require(o && milliseconds && repetitions, "invalid FMM timing arguments");
        require(o->op->fp32_available(), "the FP32-bulk apply needs spherical levels of order at most 48");
        hsbem::hip_check(hipSetDevice(o->ctx->device), "set device");
        const std::size_t n = 2 * static_cast<std::size_t>(o->op->size());
        hipStream_t s = o->ctx->stream;
        o->x.upload(std::vector<double>(n, 1.0), s);
        o->y.resize(n);
        o->op->apply_fp32(o->x.data(), o->y.data(), s); // warm (and the FP32 buffers)
        hipEvent_t e0, e1;
        hsbem::hip_check(hipEventCreate(&e0), "event");
        hsbem::hip_check(hipEventCreate(&e1), "event");
        hsbem::hip_check(hipEventRecord(e0, s), "event");
        for (uint32_t r = 0; r < repetitions; ++r)
            o->op->apply_fp32(o->x.data(), o->y.data(), s);
        hsbem::hip_check(hipEventRecord(e1, s), "event");
        hsbem::hip_check(hipEventSynchronize(e1), "event sync");
        float ms = 0;
        hsbem::hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *milliseconds = ms / repetitions;
    });
}

int hsbem_fmm_operator_time_apply(hsbem_fmm_operator *o, uint32_t repetitions, double *milliseconds) {
    return guarded([&] {
        require(o && milliseconds && repetitions, "invalid FMM timing arguments");
        hsbem::hip_check(hipSetDevice(o->ctx->device), "set device");
        const std::size_t n = 2 * static_cast<std::size_t>(o->op->size());
        hipStream_t s = o->ctx->stream;
        o->x.upload(std::vector<double>(n, 1.0), s);
        o->y.resize(n);
        auto run = [&] {
            if (o->graph)
                o->op->apply_graph(o->x.data(), o->y.data(), s);
            else
                o->op->apply(o->x.data(), o->y.data(), s);
        };
        run(); // warm (and capture)
        hipEvent_t e0, e1;
        hsbem::hip_check(hipEventCreate(&e0), "event");
        hsbem::hip_check(hipEventCreate(&e1), "event");
        hsbem::hip_check(hipEventRecord(e0, s), "event");
        for (uint32_t r = 0; r < repetitions; ++r)
            run();
        hsbem::hip_check(hipEventRecord(e1, s), "event");
        hsbem::hip_check(hipEventSynchronize(e1), "event sync");
        float ms = 0;
        hsbem::hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *milliseconds = ms / repetitions;
    });
}

int hsbem_fmm_operator_statistics(const hsbem_fmm_operator *o, uint64_t *counts, double *times, uint32_t *orders,
                                  uint32_t *level_count) {
    return guarded([&] {
        require(o && counts && times && level_count, "null FMM statistics argument");
        const auto &st = o->op->statistics();
        const uint64_t v[16] = {st.nodes, st.levels, st.leaves, st.top_level, st.rule_points, st.near_nnz,
                                st.far_pairs, st.m2m_links, st.m2l_links, st.l2l_links, st.rotations, st.axial_plans,
                                st.expansion_values, st.moment_values, st.device_bytes, o->op->size()};
        std::copy(v, v + 16, counts);
        times[0] = st.tree_ms;
        times[1] = st.near_ms;
        times[2] = st.plans_ms;
        times[3] = st.moments_ms;
        times[4] = st.near_kernel_ms;
        if (orders)
            std::copy(st.level_orders.begin(), st.level_orders.end(), orders);
        *level_count = static_cast<uint32_t>(st.level_orders.size());
    });
}

int hsbem_fmm_operator_levels(const hsbem_fmm_operator *o, uint64_t *counts, uint32_t *representation,
                              uint32_t *rule_orders, uint32_t *level_count) {
    return guarded([&] {
        require(o && counts && level_count, "null FMM level argument");
        const auto &st = o->op->statistics();
        counts[0] = st.plane_wave_levels;
        counts[1] = st.angular_links;
        counts[2] = st.diagonal_classes;
        counts[3] = st.plane_wave_leaves;
        counts[4] = st.plan_cache_hits;
        counts[5] = st.plans_built;
        if (representation)
            std::copy(st.level_representation.begin(), st.level_representation.end(), representation);
        if (rule_orders)
            std::copy(st.level_rule_orders.begin(), st.level_rule_orders.end(), rule_orders);
        *level_count = static_cast<uint32_t>(st.level_representation.size());
    });
}

int hsbem_fmm_operator_set_graph(hsbem_fmm_operator *o, int32_t enable) {
    return guarded([&] {
        require(o, "null FMM operator");
        o->graph = enable != 0;
    });
}

int hsbem_fmm_operator_destroy(hsbem_fmm_operator *o) {
    return guarded([&] {
        if (o)
            hsbem::hip_check(hipSetDevice(o->ctx->device), "set device");
        delete o;
    });
}

int hsbem_near_operator_sizes(const hsbem_near_operator *o, uint64_t *s, double *ms) {
    return guarded([&] {
        require(o && s, "null near operator");
        const auto &op = o->op;
        const uint64_t v[8] = {op.row_offsets.size(), op.columns.size(), op.flux_columns.size(), op.jobs,
                               op.regular, op.touching, op.patches, op.fan_jobs};
        std::copy(v, v + 8, s);
        if (ms)
            *ms = op.kernel_ms;
    });
}

int hsbem_near_operator_copy(const hsbem_near_operator *o, uint32_t *ro, uint32_t *co, double *va, uint32_t *fro,
                             uint32_t *fco, double *fva, double *applied) {
    return guarded([&] {
        require(o, "null near operator");
        const auto &op = o->op;
        if (ro)
            std::copy(op.row_offsets.begin(), op.row_offsets.end(), ro);
        if (co)
            std::copy(op.columns.begin(), op.columns.end(), co);
        if (va)
            std::copy(op.values.begin(), op.values.end(), va);
        if (fro)
            std::copy(op.flux_offsets.begin(), op.flux_offsets.end(), fro);
        if (fco)
            std::copy(op.flux_columns.begin(), op.flux_columns.end(), fco);
        if (fva)
            std::copy(op.flux_values.begin(), op.flux_values.end(), fva);
        if (applied) {
            require(!op.applied.empty(), "near operator was created without a vector");
            std::copy(op.applied.begin(), op.applied.end(), applied);
        }
    });
}

int hsbem_near_operator_destroy(hsbem_near_operator *o) {
    return guarded([&] { delete o; });
}

int hsbem_rotation_plan(uint32_t order, const double *axis, double *small_d, double *azimuth, const double *coefficients,
                        int32_t direction, double *rotated) {
    return guarded([&] {
        require(axis, "null rotation axis");
        const auto p = hsbem::make_rotation_plan(order, {axis[0], axis[1], axis[2]});
        if (small_d)
            std::copy(p.small_d.begin(), p.small_d.end(), small_d);
        if (azimuth)
            for (std::size_t i = 0; i < p.azimuth.size(); ++i) {
                azimuth[2 * i] = p.azimuth[i].real();
                azimuth[2 * i + 1] = p.azimuth[i].imag();
            }
        if (coefficients) {
            require(rotated && direction != 0, "rotation needs an output and a direction");
            const std::size_t nc = static_cast<std::size_t>(order + 1) * (order + 1);
            std::vector<std::complex<double>> c(nc);
            for (std::size_t i = 0; i < nc; ++i)
                c[i] = {coefficients[2 * i], coefficients[2 * i + 1]};
            const auto r = direction > 0 ? hsbem::rotate_to_axis(p, c) : hsbem::rotate_from_axis(p, c);
            for (std::size_t i = 0; i < nc; ++i) {
                rotated[2 * i] = r[i].real();
                rotated[2 * i + 1] = r[i].imag();
            }
        }
    });
}

namespace {

std::vector<std::complex<double>> complex_in(const double *v, std::size_t n) {
    std::vector<std::complex<double>> out(n);
    for (std::size_t i = 0; i < n; ++i)
        out[i] = {v[2 * i], v[2 * i + 1]};
    return out;
}

void complex_out(const std::vector<std::complex<double>> &v, double *out) {
    for (std::size_t i = 0; i < v.size(); ++i) {
        out[2 * i] = v[i].real();
        out[2 * i + 1] = v[i].imag();
    }
}

std::size_t rule_size(uint32_t order) { return 2 * static_cast<std::size_t>(order + 1) * (order + 1); }

} // namespace

int hsbem_sphere_rule(uint32_t order, double *directions, double *weights) {
    return guarded([&] {
        require(directions && weights, "null sphere rule output");
        const auto r = hsbem::make_sphere_rule(order);
        for (std::size_t q = 0; q < r.samples.size(); ++q) {
            directions[3 * q] = r.samples[q].direction.x;
            directions[3 * q + 1] = r.samples[q].direction.y;
            directions[3 * q + 2] = r.samples[q].direction.z;
            weights[q] = r.samples[q].weight;
        }
    });
}

int hsbem_spherical_harmonics(uint32_t order, const double *direction, double *values) {
    return guarded([&] {
        require(direction && values, "null harmonic argument");
        complex_out(hsbem::spherical_harmonics(order, {direction[0], direction[1], direction[2]}), values);
    });
}

int hsbem_regular_radial_table(uint32_t order, double x, double *values, double *derivatives) {
    return guarded([&] {
        require(values && derivatives, "null radial output");
        const auto t = hsbem::regular_radial_table(order, x);
        std::copy(t.values.begin(), t.values.end(), values);
        std::copy(t.derivatives.begin(), t.derivatives.end(), derivatives);
    });
}

int hsbem_angular_transfer(uint32_t source_order, uint32_t target_order, uint32_t bandlimit, int32_t mode, double k,
                           const double *shift, const double *values, double *out) {
    return guarded([&] {
        require(values && out, "null spectrum");
        require(mode == HSBEM_TRANSFER_RESAMPLE || shift, "null shift");
        const auto source = hsbem::make_sphere_rule(source_order), target = hsbem::make_sphere_rule(target_order);
        const auto v = complex_in(values, rule_size(source_order));
        const hsbem::Vec3 d = shift ? hsbem::Vec3{shift[0], shift[1], shift[2]} : hsbem::Vec3{};
        switch (mode) {
        case HSBEM_TRANSFER_RESAMPLE:
            complex_out(hsbem::resample_spectrum(v, source, target, bandlimit), out);
            break;
        case HSBEM_TRANSFER_OUTGOING_TO_PARENT:
            complex_out(hsbem::outgoing_to_parent(v, source, target, bandlimit, k, d), out);
            break;
        case HSBEM_TRANSFER_INCOMING_TO_CHILD:
            complex_out(hsbem::incoming_to_child(v, source, target, bandlimit, k, d), out);
            break;
        default:
            throw std::invalid_argument("unknown angular transfer mode");
        }
    });
}

int hsbem_angular_transfer_device(hsbem_context *c, int32_t kind, uint32_t source_order, uint32_t target_order,
                                  uint32_t bandlimit, double k, double radius, const double *shift, const double *in,
                                  double *out) {
    return guarded([&] {
        require(c && shift && in && out, "null device angular transfer argument");
        require(kind >= HSBEM_DEVICE_SCALED_TO_PARENT && kind <= HSBEM_DEVICE_INCOMING_TO_CHILD,
                "unknown device angular transfer kind");
        hsbem::hip_check(hipSetDevice(c->ctx.device), "set device");
        using K = hsbem::dev::HfTransferKind;
        const K kinds[] = {K::up_handoff, K::up_resample, K::down_handoff, K::down_resample};
        const K kd = kinds[kind];
        const std::size_t n_in = kd == K::up_handoff ? static_cast<std::size_t>(source_order + 1) * (source_order + 1)
                                                     : rule_size(source_order);
        const auto v = complex_in(in, n_in);
        complex_out(hsbem::device_angular_transfer(kd, source_order, target_order, bandlimit, k, radius,
                                                   {shift[0], shift[1], shift[2]}, v, c->ctx.stream),
                    out);
    });
}

int hsbem_plan_cache_create(uint64_t maximum_bytes, hsbem_plan_cache **out) {
    return guarded([&] {
        require(out && maximum_bytes, "invalid plan cache arguments");
        *out = new hsbem_plan_cache(static_cast<std::size_t>(maximum_bytes));
    });
}

int hsbem_plan_cache_axial(hsbem_plan_cache *c, int32_t kind, uint32_t ps, uint32_t pt, double k, double as, double at,
                           double distance, uint32_t polar_count, double projection_radius, uint64_t *count,
                           double *coefficients) {
    return guarded([&] {
        require(c && count, "null plan cache argument");
        require(kind >= HSBEM_AXIAL_MULTIPOLE_TO_PARENT && kind <= HSBEM_AXIAL_LOCAL_TO_CHILD, "unknown axial kind");
        const auto plan = c->cache.axial({static_cast<hsbem::AxialKind>(kind), ps, pt, polar_count, k, as, at,
                                          distance, projection_radius});
        *count = plan->coefficients.size();
        if (coefficients)
            complex_out(plan->coefficients, coefficients);
    });
}

int hsbem_plan_cache_statistics(const hsbem_plan_cache *c, uint64_t *out, double *builder_ms) {
    return guarded([&] {
        require(c && out, "null plan cache argument");
        const auto st = c->cache.statistics();
        out[0] = st.hits;
        out[1] = st.plans_built;
        out[2] = st.evictions;
        out[3] = st.live_bytes;
        out[4] = st.peak_bytes;
        if (builder_ms)
            *builder_ms = st.builder_ms;
    });
}

int hsbem_plan_cache_destroy(hsbem_plan_cache *c) {
    return guarded([&] { delete c; });
}

int hsbem_harmonic_spectrum(uint32_t order, uint32_t rule_order, int32_t direction, const double *in, double *out) {
    return guarded([&] {
        require(in && out && direction != 0, "null spectrum or direction");
        const auto rule = hsbem::make_sphere_rule(rule_order);
        if (direction > 0)
            complex_out(hsbem::multipole_to_spectrum(complex_in(in, static_cast<std::size_t>(order + 1) * (order + 1)),
                                                     order, rule),
                        out);
        else
            complex_out(hsbem::spectrum_to_local(complex_in(in, rule_size(rule_order)), rule, order), out);
    });
}

int hsbem_diagonal_translation(uint32_t rule_order, uint32_t expansion_order, double k,
                               const double *target_minus_source, double *out) {
    return guarded([&] {
        require(target_minus_source && out, "null translation argument");
        const auto rule = hsbem::make_sphere_rule(rule_order);
        complex_out(hsbem::diagonal_translation(
                        rule, expansion_order, k,
                        {target_minus_source[0], target_minus_source[1], target_minus_source[2]}),
                    out);
    });
}

namespace {

void wave_out(const hsbem::WaveValue &w, double *out) {
    out[0] = w.value.real();
    out[1] = w.value.imag();
    for (int i = 0; i < 3; ++i) {
        out[2 + 2 * i] = w.gradient[i].real();
        out[3 + 2 * i] = w.gradient[i].imag();
    }
}

} // namespace

int hsbem_scaled_waves(uint32_t order, double k, double radius, const double *point, int32_t outgoing,
                       double *waves) {
    return guarded([&] {
        require(point && waves, "null wave argument");
        const hsbem::Vec3 p{point[0], point[1], point[2]};
        const auto w = outgoing ? hsbem::scaled_outgoing_waves(order, k, radius, p)
                                : hsbem::scaled_regular_waves(order, k, radius, p);
        for (std::size_t j = 0; j < w.size(); ++j)
            wave_out(w[j], waves + 8 * j);
    });
}

int hsbem_scaled_expansion_evaluate(uint32_t order, double radius, double k, const double *coefficients,
                                    int32_t local, const double *point, double *wave) {
    return guarded([&] {
        require(coefficients && point && wave, "null expansion argument");
        const hsbem::ScaledExpansion e{order, radius,
                                       complex_in(coefficients, static_cast<std::size_t>(order + 1) * (order + 1))};
        const hsbem::Vec3 p{point[0], point[1], point[2]};
        wave_out(local ? hsbem::evaluate_scaled_local(e, k, p) : hsbem::evaluate_scaled_multipole(e, k, p), wave);
    });
}

int hsbem_scaled_spectrum(uint32_t order, double radius, double k, uint32_t rule_order, int32_t direction,
                          const double *in, double *out) {
    return guarded([&] {
        require(in && out && direction != 0, "null spectrum or direction");
        const auto rule = hsbem::make_sphere_rule(rule_order);
        if (direction > 0)
            complex_out(hsbem::scaled_multipole_to_spectrum(
                            {order, radius, complex_in(in, static_cast<std::size_t>(order + 1) * (order + 1))}, k, rule),
                        out);
        else
            complex_out(hsbem::spectrum_to_scaled_local(complex_in(in, rule_size(rule_order)), rule, order, k, radius)
                            .coefficients,
                        out);
    });
}

int hsbem_axial_plan(int32_t kind, uint32_t source_order, uint32_t target_order, double k, double source_radius,
                     double target_radius, double distance, uint32_t polar_count, double projection_radius,
                     uint64_t *coefficient_count, double *coefficients, double *probe, const double *source,
                     double *applied) {
    return guarded([&] {
        require(coefficient_count, "null coefficient count");
        const auto plan = hsbem::make_axial_plan(static_cast<hsbem::AxialKind>(kind), source_order, target_order, k,
                                                 source_radius, target_radius, distance, polar_count,
                                                 projection_radius);
        *coefficient_count = plan.coefficients.size();
        if (coefficients)
            complex_out(plan.coefficients, coefficients);
        if (probe)
            *probe = plan.projection_radius;
        if (source) {
            require(applied, "null applied output");
            complex_out(hsbem::apply_axial_plan(
                            plan, complex_in(source, static_cast<std::size_t>(source_order + 1) * (source_order + 1))),
                        applied);
        }
    });
}

namespace {

std::vector<hsbem::AngularSample> samples_in(const double *angles, const double *levels, uint32_t n) {
    require(n == 0 || (angles && levels), "null angular samples");
    std::vector<hsbem::AngularSample> s(n);
    for (uint32_t i = 0; i < n; ++i)
        s[i] = {angles[i], levels[i]};
    return s;
}

hsbem::AcousticFrame frame_in(const double *origin, const double *forward, const double *horizontal) {
    require(origin && forward, "null acoustic frame");
    const hsbem::Vec3 o{origin[0], origin[1], origin[2]}, f{forward[0], forward[1], forward[2]};
    if (!horizontal)
        return hsbem::default_acoustic_frame(o, f);
    return hsbem::normalized_acoustic_frame({o, f, {horizontal[0], horizontal[1], horizontal[2]}});
}

} // namespace

int hsbem_spl_db(double re, double im, double reference, double *out) {
    return guarded([&] {
        require(out, "null SPL output");
        *out = hsbem::spl_db({re, im}, reference);
    });
}

int hsbem_normalized_volume_impedance(double re, double im, double area, double density, double sound_speed,
                                      double *out) {
    return guarded([&] {
        require(out, "null impedance output");
        const auto z = hsbem::normalized_volume_impedance({re, im}, area, density, sound_speed);
        out[0] = z.real();
        out[1] = z.imag();
    });
}

int hsbem_angular_level_db(const double *angles, const double *levels, uint32_t count, double angle, double *level,
                           int32_t *present) {
    return guarded([&] {
        require(level && present, "null angular level output");
        const auto v = hsbem::angular_level_db(samples_in(angles, levels, count), angle);
        *present = v.has_value();
        *level = v.value_or(0.0);
    });
}

int hsbem_beamwidth(const double *angles, const double *levels, uint32_t count, double attenuation, double *out,
                    int32_t *present) {
    return guarded([&] {
        require(out && present, "null beamwidth output");
        const auto b = hsbem::beamwidth(samples_in(angles, levels, count), attenuation);
        *present = b.has_value();
        out[0] = b ? b->left : 0.0;
        out[1] = b ? b->right : 0.0;
        out[2] = b ? b->width : 0.0;
    });
}

int hsbem_cta2034_cuts(const double *h_angles, const double *h_levels, uint32_t h_count, const double *v_angles,
                       const double *v_levels, uint32_t v_count, double *out, int32_t *present) {
    return guarded([&] {
        require(out && present, "null CTA output");
        const auto c = hsbem::cta2034_cuts(samples_in(h_angles, h_levels, h_count),
                                           samples_in(v_angles, v_levels, v_count));
        const std::optional<double> v[7] = {c.on_axis, c.listening_window, c.early_reflections,
                                            c.early_reflections_di, c.sound_power, c.sound_power_di,
                                            c.predicted_in_room};
        for (int i = 0; i < 7; ++i) {
            present[i] = v[i].has_value();
            out[i] = v[i].value_or(0.0);
        }
    });
}

int hsbem_angular_frequency_response(const double *frequencies, const uint32_t *row_offsets, const double *angles,
                                     const double *levels, uint32_t rows, double angle, int32_t relative,
                                     double reference_angle, double *levels_out, int32_t *present) {
    return guarded([&] {
        require(rows == 0 || (frequencies && row_offsets && levels_out && present), "null response arguments");
        std::vector<hsbem::AngularRow> r(rows);
        for (uint32_t i = 0; i < rows; ++i) {
            require(row_offsets[i + 1] >= row_offsets[i], "invalid row offsets");
            r[i].frequency = frequencies[i];
            r[i].samples = samples_in(angles ? angles + row_offsets[i] : nullptr,
                                      levels ? levels + row_offsets[i] : nullptr, row_offsets[i + 1] - row_offsets[i]);
        }
        const auto v = hsbem::angular_frequency_response(r, angle, relative != 0, reference_angle);
        for (uint32_t i = 0; i < rows; ++i) {
            present[i] = v[i].has_value();
            levels_out[i] = v[i].value_or(0.0);
        }
    });
}

int hsbem_acoustic_frame(const double *origin, const double *forward, const double *horizontal, double *out) {
    return guarded([&] {
        require(out, "null frame output");
        const auto f = frame_in(origin, forward, horizontal);
        const hsbem::Vec3 v = hsbem::vertical_direction(f);
        const double w[12] = {f.origin.x,     f.origin.y,     f.origin.z,     f.forward.x,
                              f.forward.y,    f.forward.z,    f.horizontal.x, f.horizontal.y,
                              f.horizontal.z, v.x,            v.y,            v.z};
        std::copy(w, w + 12, out);
    });
}

int hsbem_polar_directions(const double *origin, const double *forward, const double *horizontal, int32_t vertical,
                           const double *degrees, uint32_t count, double *directions) {
    return guarded([&] {
        require((degrees && directions) || count == 0, "null polar arguments");
        const auto d = hsbem::polar_directions(frame_in(origin, forward, horizontal), vertical != 0,
                                               std::vector<double>(degrees, degrees + count));
        for (uint32_t i = 0; i < count; ++i) {
            directions[3 * i] = d[i].x;
            directions[3 * i + 1] = d[i].y;
            directions[3 * i + 2] = d[i].z;
        }
    });
}

int hsbem_frequency_schedule(uint64_t count, uint64_t *order) {
    return guarded([&] {
        require(order || count == 0, "null schedule output");
        const auto s = hsbem::frequency_schedule(static_cast<std::size_t>(count));
        std::copy(s.begin(), s.end(), order);
    });
}

int hsbem_rcr_queue_policy(uint32_t source_order, uint32_t target_order, uint32_t reserved, uint64_t *out) {
    return guarded([&] {
        require(out, "null output");
        const auto p = hsbem::make_rcr_queue_policy(source_order, target_order, reserved);
        out[0] = p.batch_capacity;
        out[1] = p.maximum_chunk_batches;
        out[2] = p.work_per_link;
    });
}

int hsbem_plan_sizes(const hsbem_plan *p, uint64_t *s) {
    return guarded([&] {
        require(p && s, "null plan");
        s[0] = p->plan.touching.size();
        s[1] = p->plan.near.size();
        s[2] = p->plan.patches.size();
        s[3] = p->plan.whole_panel_degrees.size();
        s[4] = p->plan.candidate_pairs;
        s[5] = p->plan.maximum_depth;
    });
}

int hsbem_plan_copy(const hsbem_plan *p, uint32_t *touching, uint32_t *near, double *patches, uint8_t *whole,
                    uint32_t *offsets, uint32_t *trials, double *max_error) {
    return guarded([&] {
        require(p, "null plan");
        const auto &pl = p->plan;
        if (touching)
            for (std::size_t i = 0; i < pl.touching.size(); ++i) {
                const auto &j = pl.touching[i];
                uint32_t *o = touching + 9 * i;
                o[0] = j.test;
                o[1] = j.trial;
                o[2] = static_cast<uint32_t>(j.permutation.topology);
                for (int l = 0; l < 3; ++l) {
                    o[3 + l] = j.permutation.test[l];
                    o[6 + l] = j.permutation.trial[l];
                }
            }
        if (near)
            for (std::size_t i = 0; i < pl.near.size(); ++i) {
                const auto &j = pl.near[i];
                uint32_t *o = near + 5 * i;
                o[0] = j.test, o[1] = j.trial, o[2] = j.first_patch, o[3] = j.patch_count, o[4] = j.swap;
            }
        if (patches)
            for (std::size_t i = 0; i < pl.patches.size(); ++i) {
                const auto &q = pl.patches[i];
                double *o = patches + 22 * i;
                std::memcpy(o, q.test, 9 * sizeof(double));
                std::memcpy(o + 9, q.trial, 9 * sizeof(double));
                o[18] = q.area_fraction;
                o[19] = q.degree;
                o[20] = q.projection;
                o[21] = q.projection_intervals;
            }
        if (whole && !pl.whole_panel_degrees.empty())
            std::memcpy(whole, pl.whole_panel_degrees.data(), pl.whole_panel_degrees.size());
        if (offsets)
            std::memcpy(offsets, pl.near_offsets.data(), sizeof(uint32_t) * pl.near_offsets.size());
        if (trials && !pl.near_trials.empty())
            std::memcpy(trials, pl.near_trials.data(), sizeof(uint32_t) * pl.near_trials.size());
        if (max_error)
            *max_error = pl.maximum_estimated_relative_error;
    });
}

int hsbem_plan_destroy(hsbem_plan *p) {
    return guarded([&] { delete p; });
}

int hsbem_assemble(hsbem_mesh *m, const hsbem_numerics *n, double k_max, double k, double *A, double *B) {
    return guarded([&] {
        require(m && n && A && B && k > 0 && k_max >= k, "invalid assembly arguments");
        hsbem::hip_check(hipSetDevice(m->context->ctx.device), "set device");
        m->solver->assemble_dense(to_numerics(*n), k_max, k, A, B);
    });
}

int hsbem_dense_gmres(hsbem_context *c, uint32_t n, const double *A, const double *b, double *x, uint32_t restart,
                      uint32_t maximum_iterations, double tolerance, double *relative_residual, uint32_t *iterations) {
    return guarded([&] {
        require(c && A && b && x && n && restart >= 1 && maximum_iterations >= 1 && tolerance > 0,
                "invalid GMRES arguments");
        hsbem::hip_check(hipSetDevice(c->ctx.device), "set device");
        hipStream_t s = c->ctx.stream;
        const std::size_t nn = static_cast<std::size_t>(n) * n;
        hsbem::DeviceBuffer<double> dA(2 * nn), db(2 * static_cast<std::size_t>(n)), dx(2 * static_cast<std::size_t>(n));
        hsbem::hip_check(hipMemcpyAsync(dA.data(), A, dA.bytes(), hipMemcpyHostToDevice, s), "A upload");
        hsbem::hip_check(hipMemcpyAsync(db.data(), b, db.bytes(), hipMemcpyHostToDevice, s), "b upload");
        hsbem::GmresWorkspace ws;
        double achieved = 0;
        std::uint32_t its = 0;
        hsbem::dense_gmres(s, dA.data(), false, n, db.data(), dx.data(), restart, maximum_iterations, tolerance,
                           &achieved, &its, nullptr, ws, hsbem::BlockJacobi{});
        hsbem::hip_check(hipMemcpyAsync(x, dx.data(), dx.bytes(), hipMemcpyDeviceToHost, s), "x download");
        hsbem::hip_check(hipStreamSynchronize(s), "GMRES sync");
        if (relative_residual)
            *relative_residual = achieved;
        if (iterations)
            *iterations = its;
    });
}

int hsbem_dense_lu_solve(hsbem_context *c, uint32_t n, const double *A, uint32_t nrhs, double *B) {
    return guarded([&] {
        require(c && A && B && n && nrhs, "invalid LU arguments");
        hsbem::hip_check(hipSetDevice(c->ctx.device), "set device");
        hipStream_t s = c->ctx.stream;
        const std::size_t nn = static_cast<std::size_t>(n) * n;
        hsbem::DeviceBuffer<double> dA(2 * nn), dB(2 * static_cast<std::size_t>(n) * nrhs);
        hsbem::DeviceBuffer<std::int32_t> piv(n);
        hsbem::DeviceBuffer<int> info(1);
        hsbem::hip_check(hipMemcpyAsync(dA.data(), A, dA.bytes(), hipMemcpyHostToDevice, s), "A upload");
        hsbem::hip_check(hipMemcpyAsync(dB.data(), B, dB.bytes(), hipMemcpyHostToDevice, s), "B upload");
        hsbem::hip_check(hipMemsetAsync(info.data(), 0, sizeof(int), s), "info clear");
        hsbem::dev::lu_factor(dA.data(), n, piv.data(), info.data(), nullptr, s);
        int code = 0;
        hsbem::hip_check(hipMemcpyAsync(&code, info.data(), sizeof(int), hipMemcpyDeviceToHost, s), "info download");
        hsbem::hip_check(hipStreamSynchronize(s), "LU sync");
        if (code)
            throw std::runtime_error("LU factorization failed at pivot " + std::to_string(code));
        hsbem::dev::lu_solve(dA.data(), n, piv.data(), dB.data(), nrhs, s);
        hsbem::hip_check(hipMemcpyAsync(B, dB.data(), dB.bytes(), hipMemcpyDeviceToHost, s), "B download");
        hsbem::hip_check(hipStreamSynchronize(s), "LU solve sync");
    });
}

int hsbem_dense_lu_solve_noalloc(hsbem_context *c, uint32_t n, double *A, uint32_t nrhs, double *B) {
    return guarded([&] {
        require(c != nullptr, "invalid LU arguments");
        hsbem::dev::dense_lu_solve_noalloc(c->ctx.device, c->ctx.stream, n, A, nrhs, B);
    });
}

} // extern "C"
