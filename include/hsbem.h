/*
 * hsbem.h - C ABI of the MI355X Burton-Miller Galerkin BEM solver.
 *
 * The boundary replaces hornsculpt::CudaBemSolver
 * (reference: modules/bem/include/hornsculpt/bem.hpp:367-388):
 *   CudaBemSolver(SurfaceMesh)            -> hsbem_mesh_create      (bem.hpp:369)
 *   CudaBemSolver::solve(BemRequest, ...) -> hsbem_solve            (bem.hpp:384-385)
 *   BemResult / BemFrequencyResult / BemExcitationResult accessors
 *                                          -> hsbem_result_*        (bem.hpp:193-332)
 *   static deviceDescription()            -> hsbem_device_description (bem.hpp:386)
 * and exposes the internal seams the reference tests drive directly:
 *   galerkin::makeInteractionPlan         -> hsbem_plan_*  (galerkin/interaction_plan.hpp:68-71)
 *   dense Galerkin assembly kernels       -> hsbem_assemble (cuda/cuda_solver.cu:266-540)
 *   cuda_detail::DeviceGmresSolver        -> hsbem_dense_gmres (cuda/device_gmres.hpp:16-34)
 *   cusolverDnCgetrf/Cgetrs dense solve   -> hsbem_dense_lu_solve (cuda/cuda_solver.cu:1485-1513)
 *
 * Conventions (bem.hpp:356): RMS phasors with exp(+i omega t), outgoing
 * exp(-ikr), q = dp/dn = -i rho omega v, Burton-Miller alpha = -i/k.
 * Complex numbers are (re, im) double pairs. Every matrix is column-major.
 * Every call returns an hsbem_status; on failure hsbem_last_error() gives the
 * message of the calling thread (std::invalid_argument -> HSBEM_INVALID_ARGUMENT,
 * std::runtime_error -> HSBEM_RUNTIME_ERROR, as the reference throws).
 */
#ifndef HSBEM_H
#define HSBEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSBEM_API __attribute__((visibility("default")))

typedef enum {
    HSBEM_OK = 0,
    HSBEM_INVALID_ARGUMENT = 1, /* reference: std::invalid_argument */
    HSBEM_RUNTIME_ERROR = 2,    /* reference: std::runtime_error (numerics, budget, convergence) */
    HSBEM_DEVICE_ERROR = 3,     /* HIP runtime failure */
    HSBEM_CANCELLED = 4,        /* reference: "BEM calculation cancelled." */
    HSBEM_PENDING = 5           /* hsbem_poll: the request is still queued or running */
} hsbem_status;

enum { HSBEM_SOLVER_LU = 0, HSBEM_SOLVER_GMRES = 1 };
enum { HSBEM_PRECISION_FP64 = 0, HSBEM_PRECISION_MIXED = 1 };
/* Right preconditioner of the dense GMRES: LU of the 64 x 64 diagonal blocks of
 * the row-scaled system in the spatial (cluster) node order. */
enum { HSBEM_PRECONDITIONER_NONE = 0, HSBEM_PRECONDITIONER_BLOCK_JACOBI = 1 };
enum { HSBEM_BOUNDARY_WALL = 0, HSBEM_BOUNDARY_DRIVEN = 1 };

/* BemSolvePhase (bem.hpp:334-342). */
enum {
    HSBEM_PHASE_PREPARING = 0,
    HSBEM_PHASE_NEAR_CONSTRUCTION = 1,
    HSBEM_PHASE_FAR_PREPARATION = 2,
    HSBEM_PHASE_KRYLOV_SOLVE = 3,
    HSBEM_PHASE_RESIDUAL_AUDIT = 4,
    HSBEM_PHASE_FIELD_EVALUATION = 5,
    HSBEM_PHASE_FREQUENCY_COMPLETE = 6
};

#define HSBEM_ACCELERATION_DENSE 0
#define HSBEM_ACCELERATION_FAST_MULTIPOLE 1

typedef struct hsbem_context hsbem_context;
typedef struct hsbem_mesh hsbem_mesh;
typedef struct hsbem_result hsbem_result;
typedef struct hsbem_plan hsbem_plan;
typedef struct hsbem_server hsbem_server;
typedef struct hsbem_near_operator hsbem_near_operator;

/* BemNumerics (bem.hpp:113-123) plus the MI355X solver choices. */
typedef struct {
    uint32_t quadrature_order;           /* 2 or 4 (default 4) */
    uint32_t singular_order;             /* 2..20 (default 4) */
    double near_distance_ratio;          /* 0..10 (default 1) */
    double near_relative_tolerance;      /* 1e-10..1e-2 (default 1e-4) */
    uint32_t near_maximum_depth;         /* <= 10 (default 6) */
    double relative_residual_tolerance;  /* 1e-12..1e-2 (default 5e-5) */
    uint32_t maximum_refinement_steps;   /* <= 10 (default 3) */
    uint32_t solver;                     /* HSBEM_SOLVER_LU (default) or HSBEM_SOLVER_GMRES */
    uint32_t gmres_restart;              /* default 50 */
    uint32_t gmres_maximum_iterations;   /* default 1000 */
    uint32_t precision;                  /* HSBEM_PRECISION_FP64 (default) or HSBEM_PRECISION_MIXED */
    uint32_t plan_threads;               /* host planner threads, 0 = up to 16 */
    uint32_t preconditioner;             /* GMRES: HSBEM_PRECONDITIONER_BLOCK_JACOBI (default) or _NONE */
    /* BemNumerics::acceleration and FastMultipoleSettings (bem.hpp:60-71, 121): */
    uint32_t acceleration;               /* HSBEM_ACCELERATION_DENSE (default) or _FAST_MULTIPOLE (FP64) */
    uint32_t fmm_leaf_capacity;          /* leafSupportCapacity, 4..256 (default 16) */
    double fmm_separation_ratio;         /* separationRatio, 1.5..10 (default 1.5) */
    uint32_t fmm_maximum_iterations;     /* maximumIterations (default 300) */
    uint32_t fmm_restart_length;         /* restartLength (default 40) */
    uint32_t fmm_block_width;            /* preconditionerBlockWidth (bem.hpp:70): 32 (default), 64, ..., 2048 */
} hsbem_numerics;

/* BemRequest (bem.hpp:163-191) for a closed exterior mesh. Excitation e assigns
 * velocities[j] (RMS m/s, complex) to faces face_ids[j] for
 * j in [excitation_offsets[e], excitation_offsets[e+1]). */
typedef struct {
    double density;                      /* kg/m^3 */
    double sound_speed;                  /* m/s */
    const double *frequencies;           /* Hz */
    uint32_t frequency_count;
    uint32_t excitation_count;           /* 1..1024 */
    const uint32_t *excitation_offsets;  /* excitation_count + 1 entries */
    const uint32_t *face_ids;
    const double *velocities;            /* (re, im) per face entry */
    const double *observers;             /* xyz, world frame */
    uint32_t observer_count;             /* <= 100000 */
    int32_t far_field;                   /* nonzero: asymptotic far field (FarFieldObservation) */
    double far_origin[3];
    double far_reference_distance;       /* m, default 1 */
    const double *planning_frequencies;  /* optional raster (interactionPlanFrequencies) */
    uint32_t planning_frequency_count;
    hsbem_numerics numerics;
} hsbem_request;

typedef void (*hsbem_progress_fn)(void *user, uint32_t completed_frequencies, uint32_t total_frequencies,
                                  double frequency, int32_t phase);

HSBEM_API void hsbem_numerics_default(hsbem_numerics *numerics);
HSBEM_API const char *hsbem_last_error(void);
HSBEM_API int hsbem_device_description(char *buffer, size_t length);
/* The launch-shape autotuner's decisions so far in this process, as a JSON array of
 * {"tunable", "key", "chosen", "timings": [[candidate, ms], ...]} (tunables:
 * "tile_trial_block", "fmm_split_target", "gemv_shape" = column chunk | rows per lane << 16);
 * values pinned by environment variables are not listed.
 * Writes at most length - 1 characters and a terminating zero; *needed (optional)
 * receives the full length + 1. No counterpart in the reference (it has no autotuner). */
HSBEM_API int hsbem_tuning_report(char *buffer, size_t length, size_t *needed);

HSBEM_API int hsbem_context_create(int32_t device, hsbem_context **context);
HSBEM_API int hsbem_context_destroy(hsbem_context *context);

/* Copies, recentres (vertex 0 at the origin) and validates the mesh (closed,
 * consistently oriented, positive volume per component, no self-intersection,
 * no nested component), builds the Galerkin space and uploads device-resident
 * panel geometry. Replaces CudaBemSolver(SurfaceMesh) (bem.hpp:369) with the
 * checks of validatedSpace (cuda_solver.cu:782-789). */
HSBEM_API int hsbem_mesh_create(hsbem_context *context, const double *vertices, uint32_t vertex_count,
                                const uint32_t *triangles, const uint32_t *face_ids, const uint8_t *boundary,
                                uint32_t triangle_count, hsbem_mesh **mesh);
HSBEM_API int hsbem_mesh_destroy(hsbem_mesh *mesh);
/* The validation of hsbem_mesh_create on the host only (no device, no context):
 * HSBEM_OK, or HSBEM_INVALID_ARGUMENT with the reason in hsbem_last_error(). */
HSBEM_API int hsbem_mesh_validate(const double *vertices, uint32_t vertex_count, const uint32_t *triangles,
                                  uint32_t triangle_count);
/* 64-bit FNV-1a hash of the mesh arrays (state key for the serving loop). */
HSBEM_API int hsbem_mesh_info(const hsbem_mesh *mesh, uint64_t *hash, uint32_t *vertex_count,
                              uint32_t *triangle_count, double *surface_area);

/* Solves every frequency of the request. cancel (optional) is polled between
 * launches; a nonzero value ends the call with HSBEM_CANCELLED. */
/* Device state the mesh keeps between solves: the number of cached interaction
 * plans (at most 4 not held by a running solve; least recently used evicted) and
 * the resident device bytes (geometry, cached plans, FMM near structure). */
HSBEM_API int hsbem_mesh_cache_info(hsbem_mesh *mesh, uint64_t *cached_plans, uint64_t *resident_bytes);
HSBEM_API int hsbem_solve(hsbem_mesh *mesh, const hsbem_request *request, hsbem_progress_fn progress, void *user,
                          const volatile int32_t *cancel, hsbem_result **result);
HSBEM_API int hsbem_result_destroy(hsbem_result *result);

HSBEM_API int hsbem_result_sizes(const hsbem_result *result, uint32_t *frequency_count, uint32_t *excitation_count,
                                 uint32_t *triangle_count, uint32_t *vertex_count, uint32_t *observer_count);
/* DP0 triangle means of the P1 pressure (nt complex), BemBoundaryData::surfacePressure. */
HSBEM_API int hsbem_result_surface_pressure(const hsbem_result *result, uint32_t frequency, uint32_t excitation,
                                            double *out);
/* P1 nodal pressure (nv complex), BemBoundaryData::pressureTrace. */
HSBEM_API int hsbem_result_pressure_trace(const hsbem_result *result, uint32_t frequency, uint32_t excitation,
                                          double *out);
/* DP0 q = -i rho omega v (nt complex), BemBoundaryData::normalDerivative. */
HSBEM_API int hsbem_result_normal_derivative(const hsbem_result *result, uint32_t frequency, uint32_t excitation,
                                             double *out);
HSBEM_API int hsbem_result_observer_pressure(const hsbem_result *result, uint32_t frequency, uint32_t excitation,
                                             double *out);
/* FaceImpedance list (bem.hpp:193-199): one entry per excited face with nonzero velocity.
 * Pass null arrays to query *count. mean_pressure, volume_velocity, impedance are complex. */
HSBEM_API int hsbem_result_impedances(const hsbem_result *result, uint32_t frequency, uint32_t excitation,
                                      uint32_t *count, uint32_t *face_ids, double *area, double *mean_pressure,
                                      double *volume_velocity, double *impedance);
/* Relative L2 residual of the unscaled equation, Krylov iterations, refinement steps. */
HSBEM_API int hsbem_result_residual(const hsbem_result *result, uint32_t frequency, uint32_t excitation,
                                    double *residual, uint32_t *iterations, uint32_t *refinements);
/* ms[12]: milliseconds of assembly, solve, factorization, triangular solve, field, interaction plan;
 * device-event milliseconds of the regular, touching and near assembly kernels; and the
 * algorithmic sample counts of those three kernels (FP64 pair-kernel evaluations). */
HSBEM_API int hsbem_result_timings(const hsbem_result *result, uint32_t frequency, double *milliseconds);
HSBEM_API int hsbem_result_peak_device_bytes(const hsbem_result *result, uint64_t *bytes);
/* Fast multipole evidence of a frequency point (BemFrequencyResult's FMM statistics,
 * bem.hpp:193-332; zeros for the dense form): out[13] = seed-audit discrepancy
 * (relative; 0 under the direct audit), apply
 * milliseconds, defect correction applied (0/1), plane-wave levels, M2L links, near
 * CSR entries, operator device bytes, residual evaluated by the direct Galerkin
 * audit (0/1), endpoint-quadrature discrepancy against the next panel rule (A and B
 * actions, relative), endpoint rule degree, endpoint-rule promotions, and (out[11])
 * the accuracy seed of the verification operator (0 under the direct audit), and
 * (out[12]) whether the Krylov iterations applied FP32 translations (0/1: precision
 * mixed, the FP32-bulk form, on operators with spherical levels of order <= 48 only).
 * The seed audit discrepancy is that of the verification operator against the audit one. */
HSBEM_API int hsbem_result_fmm(const hsbem_result *result, uint32_t frequency, double *out);

/* ---- serving loop (no counterpart class in the reference: its callers run
 * CudaBemSolver::solve on worker threads, apps/desktop/simulation_dock.cpp:2425-2600,
 * apps/desktop/parameter_sweep.cpp:215-250) ----
 * Requests are split into frequency points. Each slot (own stream and workspace)
 * takes queued points in arrival order once the point's device footprint fits
 * the memory budget; points of requests submitted while others run start as
 * soon as a slot frees, and their kernels overlap on the device. Meshes stay
 * resident (geometry, clusters, interaction plans) and may be destroyed by the
 * caller while their requests are in flight. */
typedef struct {
    uint32_t slots;               /* concurrent frequency points, 1..16 (default 3) */
    uint64_t memory_budget_bytes; /* device bytes for slot workspaces; 0: 90 % of free memory */
} hsbem_server_options;

typedef struct {
    uint64_t submitted, completed, failed, points;
    uint32_t running, queued, max_running;
    uint64_t budget_bytes, held_bytes, peak_held_bytes;
    /* device bytes the served meshes keep between points (geometry, cached plans,
     * FMM near structures) and the largest held + resident total at an admission;
     * both count against budget_bytes */
    uint64_t resident_bytes, peak_total_bytes;
} hsbem_server_statistics;

HSBEM_API int hsbem_server_create(hsbem_context *context, const hsbem_server_options *options, hsbem_server **server);
/* Cancels queued and running requests and joins the slots. */
HSBEM_API int hsbem_server_destroy(hsbem_server *server);
/* Validates like hsbem_solve (HSBEM_INVALID_ARGUMENT also when one frequency point
 * exceeds the memory budget) and queues the request; returns at once. */
HSBEM_API int hsbem_submit(hsbem_server *server, hsbem_mesh *mesh, const hsbem_request *request, uint64_t *ticket);
/* Waits up to wait_ms (0: no wait; negative: until finished). HSBEM_OK with *result
 * when every point is solved, HSBEM_PENDING otherwise, or the request's error
 * status (HSBEM_CANCELLED after hsbem_cancel). A finished ticket is released. */
HSBEM_API int hsbem_poll(hsbem_server *server, uint64_t ticket, int32_t wait_ms, hsbem_result **result);
HSBEM_API int hsbem_cancel(hsbem_server *server, uint64_t ticket);
HSBEM_API int hsbem_server_stats(const hsbem_server *server, hsbem_server_statistics *statistics);
/* Milliseconds from hsbem_submit to the completion of this frequency point. */
HSBEM_API int hsbem_result_served_latency(const hsbem_result *result, uint32_t frequency, double *milliseconds);

/* ---- internal seams (component tests, benchmarks) ---- */

/* Interaction plan of the mesh for numerics at planning wavenumber k_max.
 * sizes: touching, near jobs, patches, all-pairs bytes, candidate pairs, max depth.
 * Copy layouts: touching 9 x uint32 (test, trial, topology, perm test[3], perm trial[3]);
 * near 5 x uint32 (test, trial, first patch, patch count, swap); patch 22 doubles
 * (test[9], trial[9], area fraction, degree, projection, intervals). */
HSBEM_API int hsbem_plan_create(hsbem_mesh *mesh, const hsbem_numerics *numerics, double k_max, hsbem_plan **plan);
/* Host-only variant on raw arrays (recentred on vertex 0 like hsbem_mesh_create); no device use. */
HSBEM_API int hsbem_plan_build(const double *vertices, uint32_t vertex_count, const uint32_t *triangles,
                               uint32_t triangle_count, const hsbem_numerics *numerics, double k_max,
                               hsbem_plan **plan);
HSBEM_API int hsbem_plan_sizes(const hsbem_plan *plan, uint64_t *sizes);
HSBEM_API int hsbem_plan_copy(const hsbem_plan *plan, uint32_t *touching, uint32_t *near, double *patches,
                              uint8_t *whole_panel_degrees, uint32_t *near_offsets, uint32_t *near_trials,
                              double *maximum_estimated_error);
HSBEM_API int hsbem_plan_destroy(hsbem_plan *plan);

/* Wideband FMM support octree and near/far partition (host-only; replaces
 * makeSupportTree, modules/bem/fmm/support_tree.hpp:96-104, support_tree.cpp:395-414)
 * on the given coordinates (no recentring). Serialised into doubles (every
 * integer exact): pass blob = NULL to query *blob_size. Layout: header[14]
 * (nodes, levels, pp near, pp far, pf near, pf far, pp far-level pairs, pf
 * far-level pairs, nv, nt, represented pp near, pp far, pf near, pf far);
 * per node 38 values (center[3], half width, coords[3], level, parent,
 * children[8], first/count pressure, first/count flux, pressure box[7], flux
 * box[7], pressure radius, flux radius, active channels); pressure order,
 * flux order, pressure leaf, flux leaf; the four interaction lists (target,
 * source); the two far-level lists (target level, source level,
 * interactions, represented pairs); per level: node count, node ids, max
 * radius, max active radius, 13 counters, squared-offset count and values. */
HSBEM_API int hsbem_support_tree_build(const double *vertices, uint32_t vertex_count, const uint32_t *triangles,
                                       uint32_t triangle_count, uint32_t leaf_capacity, uint32_t maximum_depth,
                                       double separation_ratio, double near_distance_ratio, double *blob,
                                       uint64_t *blob_size);

/* Near-field DOF plan of the wideband FMM path over the support tree of the
 * same settings (host-only; replaces makeSupportNearPlan and
 * SupportNearJobIterator, modules/bem/fmm/support_near_plan.hpp:13-72) for the
 * given active face columns. Serialised into doubles: header[7] (nv + 1,
 * pressure nnz, flux nnz, nodes + 1, support panels, flux blocks, jobs), then
 * pressure row offsets and columns, flux row offsets and columns, panel face
 * column (UINT32_MAX when inactive), node panel offsets, support panels, flux
 * block offsets and indices, and per job (test panel, source panel, P1/P1
 * mask, P1/DP0 mask). blob = NULL queries *blob_size. */
HSBEM_API int hsbem_support_near_plan_build(const double *vertices, uint32_t vertex_count, const uint32_t *triangles,
                                            const uint32_t *face_ids, const uint8_t *boundary, uint32_t triangle_count,
                                            const uint32_t *active_faces, uint32_t active_face_count,
                                            uint32_t leaf_capacity, uint32_t maximum_depth, double separation_ratio,
                                            double near_distance_ratio, double *blob, uint64_t *blob_size);

/* Per-level expansion plan of the wideband FMM path at wavenumber k and
 * accuracy seed (host-only; replaces makeSupportExpansionPlan and
 * makeWidebandLevelPlan, modules/bem/fmm/support_level_plan.hpp:8-44,
 * wideband_transfer.hpp:57-72) over the support tree of the given settings;
 * panels of the active faces carry flux columns. Serialised into doubles:
 * header[2] (levels, nodes); per level 15 values (representation 0 inactive /
 * 1 spherical / 2 plane wave, active radius, electrical radius, accuracy seed,
 * signature, translation and quadrature orders, spherical coefficients,
 * angular samples, source, target and policy separations, max active radius,
 * minimum far distance, diagonal condition indicator); then per node its
 * active channels. blob = NULL queries *blob_size. */
HSBEM_API int hsbem_support_expansion_plan_build(const double *vertices, uint32_t vertex_count,
                                                 const uint32_t *triangles, const uint32_t *face_ids,
                                                 const uint8_t *boundary, uint32_t triangle_count,
                                                 const uint32_t *active_faces, uint32_t active_face_count,
                                                 uint32_t leaf_capacity, uint32_t maximum_depth,
                                                 double separation_ratio, double near_distance_ratio,
                                                 double wavenumber, double accuracy_seed, double *blob,
                                                 uint64_t *blob_size);

/* Near part of the FMM operator split on the device (replaces
 * SupportNearOperator::assemble/applyPressureDevice,
 * modules/bem/cuda/support_near_operator.hpp:13-95): the exact near entries of
 * the row-scaled A (P1 x P1 CSR) and B (P1 x active-face CSR, per unit face
 * flux) at wavenumber k, for the support tree of the given settings, in mesh
 * vertex order (the mesh is recentred like hsbem_mesh_create). x (optional,
 * nv complex) is multiplied by A_near on the device.
 * sizes[8]: nv + 1, nnz, flux nnz, near jobs, regular, touching, tensor
 * patches, projection-fan jobs. */
HSBEM_API int hsbem_near_operator_create(hsbem_mesh *mesh, const hsbem_numerics *numerics, double k_max, double k,
                                         uint32_t leaf_capacity, uint32_t maximum_depth, double separation_ratio,
                                         double near_distance_ratio, const uint32_t *active_faces,
                                         uint32_t active_face_count, const double *x, hsbem_near_operator **out);
HSBEM_API int hsbem_near_operator_sizes(const hsbem_near_operator *op, uint64_t *sizes, double *kernel_milliseconds);
HSBEM_API int hsbem_near_operator_copy(const hsbem_near_operator *op, uint32_t *row_offsets, uint32_t *columns,
                                       double *values, uint32_t *flux_offsets, uint32_t *flux_columns,
                                       double *flux_values, double *applied);
HSBEM_API int hsbem_near_operator_destroy(hsbem_near_operator *op);

/* Matrix-free FMM pressure operator (replaces SupportFmmOperator,
 * modules/bem/cuda/support_fmm_operator.hpp:66-136, with its near part
 * SupportNearOperator and its preconditioner SupportBlockJacobi,
 * support_preconditioner.hpp:9-25): y = A x for the row-scaled Burton-Miller
 * pressure operator of the mesh at wavenumber k (near entries planned for
 * k_max), A x = exact near CSR x + far field through the support octree in the
 * low-frequency spherical basis (orders from the level plan at accuracy_seed).
 * Vectors are nv complex (interleaved), global P1 numbering. */
typedef struct hsbem_fmm_settings {
    uint32_t leaf_capacity, maximum_depth;
    double separation, near_ratio;
    double accuracy_seed;
    uint32_t endpoint_degree; /* leaf panel rule degree (4, 6, 8, 12, 14); 0: minimumEndpointDegree */
} hsbem_fmm_settings;
typedef struct hsbem_fmm_operator hsbem_fmm_operator;
HSBEM_API void hsbem_fmm_settings_default(hsbem_fmm_settings *settings);
HSBEM_API int hsbem_fmm_operator_create(hsbem_mesh *mesh, const hsbem_numerics *numerics,
                                        const hsbem_fmm_settings *settings, double k_max, double k,
                                        const uint32_t *active_faces, uint32_t active_face_count,
                                        hsbem_fmm_operator **out);
/* applyPressureDevice (support_fmm_operator.hpp:86-88): y = A x. */
HSBEM_API int hsbem_fmm_operator_apply(hsbem_fmm_operator *op, const double *x, double *y);
/* preconditionDevice (support_fmm_operator.hpp:91): z = M^{-1} v, 64 x 64 near
 * diagonal blocks in tree order. */
/* y = A x with the far part integrated directly over every far DOF pair
 * (SupportDirectAudit::applyPressureDevice, support_direct_audit.hpp:11-22); meshes
 * of at most 20000 panels (HSBEM_RUNTIME_ERROR above). */
HSBEM_API int hsbem_fmm_operator_apply_direct(hsbem_fmm_operator *op, const double *x, double *y);
/* y = A x in the FP32-bulk form: expansions, rotations, coaxial products and gathers in
 * FP32 (the near part and the emission and testing sums in FP64); operators with
 * spherical levels of order <= 48 only (else HSBEM_INVALID_ARGUMENT). No counterpart
 * entry point in the reference (its FP32-bulk form is internal to the request solve);
 * for callers checking that form's action. */
HSBEM_API int hsbem_fmm_operator_apply_fp32(hsbem_fmm_operator *op, const double *x, double *y);
HSBEM_API int hsbem_fmm_operator_precondition(hsbem_fmm_operator *op, const double *v, double *z);
/* Restarted GMRES on A x = b with the FMM operator (DeviceGmresSolver::solve,
 * modules/bem/cuda/device_gmres.hpp:11-35), right block-Jacobi when precondition != 0. */
HSBEM_API int hsbem_fmm_operator_solve(hsbem_fmm_operator *op, const double *b, double *x, uint32_t restart,
                                       uint32_t maximum_iterations, double tolerance, int32_t precondition,
                                       double *achieved, uint32_t *iterations);
/* Average device time of one apply over repetitions (HIP events, context stream). */
HSBEM_API int hsbem_fmm_operator_time_apply(hsbem_fmm_operator *op, uint32_t repetitions, double *milliseconds);
/* The same timing of hsbem_fmm_operator_apply_fp32 (no counterpart in the reference). */
HSBEM_API int hsbem_fmm_operator_time_apply_fp32(hsbem_fmm_operator *op, uint32_t repetitions, double *milliseconds);
/* counts[16] = nodes, levels, leaves, top level, rule points, near nnz, far pairs,
 * M2M links, M2L links, L2L links, rotations, coaxial plans, expansion values,
 * moment values, device bytes, nv; times[5] = tree, near, plans, moments, near
 * kernels (ms);
 * orders[level_count] = expansion order per level (0: inactive). */
HSBEM_API int hsbem_fmm_operator_statistics(const hsbem_fmm_operator *op, uint64_t *counts, double *times,
                                            uint32_t *orders, uint32_t *level_count);
/* Plane-wave levels of an FMM operator: counts[6] = plane-wave levels, angular
 * transfer links, diagonal translation classes, leaves on plane-wave levels,
 * translation plans taken from the mesh's plan cache, plans built; per level
 * representation[level_count] (0 inactive, 1 scaled spherical, 2 plane wave) and
 * rule_orders[level_count] (sphere rule order of a plane-wave level, else 0). */
HSBEM_API int hsbem_fmm_operator_levels(const hsbem_fmm_operator *op, uint64_t *counts, uint32_t *representation,
                                        uint32_t *rule_orders, uint32_t *level_count);
/* Route later applies and solves through one captured graph of the whole apply
 * (replaces the M2L chunk graph replay, support_m2l_graph.hpp:13-91, and
 * support_fmm_operator.cu:865-1027): captured on first use, replayed after. */
HSBEM_API int hsbem_fmm_operator_set_graph(hsbem_fmm_operator *op, int32_t enable);
HSBEM_API int hsbem_fmm_operator_destroy(hsbem_fmm_operator *op);

/* Result post-processing (modules/analysis/include/hornsculpt/acoustic_analysis.hpp:11-78,
 * acoustic_frame.hpp:8-26). Angular samples: angles (degrees, strictly increasing)
 * and levels (dB) arrays; optional results set *present (0: no value). */
/* soundPressureLevelDecibels: 20 log10(|p| / reference); a zero pressure gives -inf. */
HSBEM_API int hsbem_spl_db(double re, double im, double reference, double *out);
/* normalizedVolumeImpedance: Z S / (rho c). out[2]. */
HSBEM_API int hsbem_normalized_volume_impedance(double re, double im, double area, double density,
                                                double sound_speed, double *out);
/* angularLevelDecibels: linear dB interpolation between finite samples. */
HSBEM_API int hsbem_angular_level_db(const double *angles, const double *levels, uint32_t count, double angle,
                                     double *level, int32_t *present);
/* beamwidthDegrees: out = left, right, width (degrees). */
HSBEM_API int hsbem_beamwidth(const double *angles, const double *levels, uint32_t count, double attenuation,
                              double *out, int32_t *present);
/* cta2034Cuts: out[7] = on axis, listening window, early reflections, early
 * reflections DI, sound power, sound power DI, predicted in-room (dB). */
HSBEM_API int hsbem_cta2034_cuts(const double *h_angles, const double *h_levels, uint32_t h_count,
                                 const double *v_angles, const double *v_levels, uint32_t v_count, double *out,
                                 int32_t *present);
/* angularFrequencyResponse: rows i = frequencies[i] with samples
 * [row_offsets[i], row_offsets[i+1]); relative != 0 subtracts the level at
 * reference_angle per row. */
HSBEM_API int hsbem_angular_frequency_response(const double *frequencies, const uint32_t *row_offsets,
                                               const double *angles, const double *levels, uint32_t rows,
                                               double angle, int32_t relative, double reference_angle,
                                               double *levels_out, int32_t *present);
/* normalizedAcousticFrame (horizontal given) or defaultAcousticFrame (horizontal
 * null) and verticalDirection: out[12] = origin, forward, horizontal, vertical. */
HSBEM_API int hsbem_acoustic_frame(const double *origin, const double *forward, const double *horizontal,
                                   double *out);
/* Far-field directions of an H (vertical == 0) or V polar sweep in that frame:
 * cos(a) forward + sin(a) axis; directions[3 count]. No reference counterpart:
 * the observer directions a directivity request hands to hsbem_solve. */
HSBEM_API int hsbem_polar_directions(const double *origin, const double *forward, const double *horizontal,
                                     int32_t vertical, const double *degrees, uint32_t count, double *directions);

/* Progressive solve order of a raster of count frequencies for realtime sweeps
 * (replaces frequencySchedule, apps/desktop/frequency_schedule.hpp:11-32): the
 * highest index, the lowest, then the midpoint of the widest remaining index gap,
 * repeatedly, equal gaps higher in the raster first; order[count] is a permutation
 * of 0..count-1. */
HSBEM_API int hsbem_frequency_schedule(uint64_t count, uint64_t *order);

/* Spherical-harmonic rotation plan of the LF FMM translations (host-only;
 * replaces makeSphericalRotationPlan / rotateToAxis / rotateFromAxis,
 * modules/bem/fmm/spherical_rotation.hpp:8-27): Wigner small-d blocks of
 * degrees 0..order (small_d: sum (2l+1)^2 doubles, row-major per degree) and
 * azimuth phases exp(+i m phi) (2 order + 1 complex). When coefficients
 * (order+1)^2 complex, indexed l*l+l+m) are given, they are rotated to the axis
 * (direction > 0) or back (direction < 0) into rotated. */
HSBEM_API int hsbem_rotation_plan(uint32_t order, const double *axis, double *small_d, double *azimuth,
                                  const double *coefficients, int32_t direction, double *rotated);
/* Angular and radial numerics of the wideband translations (host-only;
 * modules/bem/fmm/wideband_transfer.hpp:14-104, wideband_transfer.cpp:1-498).
 * Complex arrays are interleaved (re, im). Sphere rules are the canonical ones
 * of makeSphereRule: order L has 2 (L+1)^2 samples, polar-major, L+1 rings of
 * 2 (L+1) azimuths. */
/* makeSphereRule (wideband_transfer.hpp:28-30): directions 3 per sample, weights. */
HSBEM_API int hsbem_sphere_rule(uint32_t order, double *directions, double *weights);
/* sphericalHarmonics (wideband_transfer.hpp:32-34): (order+1)^2 complex values,
 * index l*l + l + m, for a finite unit direction. */
HSBEM_API int hsbem_spherical_harmonics(uint32_t order, const double *direction, double *values);
/* regularRadialTable (wideband_transfer.hpp:36-42): j_l(x) and j_l'(x), l = 0..order. */
HSBEM_API int hsbem_regular_radial_table(uint32_t order, double x, double *values, double *derivatives);
#define HSBEM_TRANSFER_RESAMPLE 0
#define HSBEM_TRANSFER_OUTGOING_TO_PARENT 1
#define HSBEM_TRANSFER_INCOMING_TO_CHILD 2
/* resampleSpectrum / outgoingToParent / incomingToChild (wideband_transfer.hpp:82-92)
 * from the rule of source_order to the rule of target_order at the given
 * bandlimit; shift = child centre - parent centre (unused for a resample). */
HSBEM_API int hsbem_angular_transfer(uint32_t source_order, uint32_t target_order, uint32_t bandlimit, int32_t mode,
                                     double k, const double *shift, const double *values, double *out);
/* multipoleToSpectrum (direction > 0: (order+1)^2 multipoles in, rule samples out)
 * or spectrumToLocal (direction < 0: rule samples in, (order+1)^2 locals out),
 * wideband_transfer.hpp:72-78. */
HSBEM_API int hsbem_harmonic_spectrum(uint32_t order, uint32_t rule_order, int32_t direction, const double *in,
                                      double *out);
/* diagonalTranslation (wideband_transfer.hpp:94-101): one complex factor per rule sample. */
HSBEM_API int hsbem_diagonal_translation(uint32_t rule_order, uint32_t expansion_order, double k,
                                         const double *target_minus_source, double *out);
/* Scaled LF spherical waves and coaxial translation plans (host-only;
 * modules/bem/fmm/scaled_translation.hpp:8-69, scaled_translation.cpp:1-399).
 * A wave is 8 doubles: value (re, im) then the gradient's x, y, z (re, im). */
/* scaledRegularWaves (outgoing == 0) / scaledOutgoingWaves (outgoing != 0),
 * scaled_translation.hpp:19-20: (order+1)^2 waves at point - centre. */
HSBEM_API int hsbem_scaled_waves(uint32_t order, double k, double radius, const double *point, int32_t outgoing,
                                 double *waves);
/* evaluateScaledMultipole (local == 0) / evaluateScaledLocal (local != 0),
 * scaled_translation.hpp:35-36: one wave. */
HSBEM_API int hsbem_scaled_expansion_evaluate(uint32_t order, double radius, double k, const double *coefficients,
                                              int32_t local, const double *point, double *wave);
/* scaledMultipoleToSpectrum (direction > 0) / spectrumToScaledLocal (direction < 0),
 * scaled_translation.hpp:37-40, on the canonical rule of rule_order. */
HSBEM_API int hsbem_scaled_spectrum(uint32_t order, double radius, double k, uint32_t rule_order, int32_t direction,
                                    const double *in, double *out);
#define HSBEM_DEVICE_SCALED_TO_PARENT 0
#define HSBEM_DEVICE_OUTGOING_TO_PARENT 1
#define HSBEM_DEVICE_INCOMING_TO_SCALED_CHILD 2
#define HSBEM_DEVICE_INCOMING_TO_CHILD 3
/* One plane-wave transfer on the device, through the FMM operator's angular kernels
 * (AngularTransferOperator::scaledOutgoingToParentDevice / outgoingToParentDevice /
 * incomingToScaledChildDevice / incomingToChildDevice, wideband_operator.hpp:39-51,
 * wideband_angular.cu:479-505). source_order / target_order: rule orders of the
 * signatures, or the spherical order (= bandlimit) of a hand-off's scaled
 * expansion of basis radius `radius`; shift = child centre - parent centre.
 * in: rule samples or (order+1)^2 coefficients; out likewise. */
HSBEM_API int hsbem_angular_transfer_device(hsbem_context *ctx, int32_t kind, uint32_t source_order,
                                            uint32_t target_order, uint32_t bandlimit, double k, double radius,
                                            const double *shift, const double *in, double *out);
/* TranslationPlanCache (modules/bem/fmm/translation_plan_cache.hpp:23-42): a bounded
 * LRU of coaxial plans keyed by the exact request. `axial` returns the plan of
 * hsbem_axial_plan's arguments (coefficients when given, their count); statistics
 * out[5] = hits, plans built, evictions, live bytes, peak bytes; *builder_ms the
 * summed build time. */
typedef struct hsbem_plan_cache hsbem_plan_cache;
HSBEM_API int hsbem_plan_cache_create(uint64_t maximum_bytes, hsbem_plan_cache **out);
HSBEM_API int hsbem_plan_cache_axial(hsbem_plan_cache *cache, int32_t kind, uint32_t source_order,
                                     uint32_t target_order, double k, double source_radius, double target_radius,
                                     double distance, uint32_t polar_count, double projection_radius,
                                     uint64_t *coefficient_count, double *coefficients);
HSBEM_API int hsbem_plan_cache_statistics(const hsbem_plan_cache *cache, uint64_t *out, double *builder_ms);
HSBEM_API int hsbem_plan_cache_destroy(hsbem_plan_cache *cache);
#define HSBEM_AXIAL_MULTIPOLE_TO_PARENT 0
#define HSBEM_AXIAL_MULTIPOLE_TO_LOCAL 1
#define HSBEM_AXIAL_LOCAL_TO_CHILD 2
/* makeScaledAxialPlan / applyScaledAxialPlan (scaled_translation.hpp:42-67): the
 * per-|m| blocks (rows l = m..target_order, columns n = m..source_order,
 * row-major, blocks in m order) into coefficients when given, their count, the
 * probe radius used, and with source ((source_order+1)^2 complex) the translated
 * coefficients into applied. distance: child - parent (M2M, L2L) or target -
 * source (M2L) along +z; projection_radius 0 selects the default probe. */
HSBEM_API int hsbem_axial_plan(int32_t kind, uint32_t source_order, uint32_t target_order, double k,
                               double source_radius, double target_radius, double distance, uint32_t polar_count,
                               double projection_radius, uint64_t *coefficient_count, double *coefficients,
                               double *probe, const double *source, double *applied);
/* RCR work model of the LF queue (makeRcrQueuePolicy, modules/bem/fmm/rcr_work.hpp:23-34):
 * out = (batch capacity, chunk batches, work per link). */
HSBEM_API int hsbem_rcr_queue_policy(uint32_t source_order, uint32_t target_order, uint32_t reserved_batches,
                                     uint64_t *out);

/* Device assembly of the row-scaled Galerkin system at wavenumber k with the
 * plan for k_max: A (nv x nv) and B (nv x nt, every panel a flux column), both
 * complex column-major, copied to host. */
HSBEM_API int hsbem_assemble(hsbem_mesh *mesh, const hsbem_numerics *numerics, double k_max, double k, double *A,
                             double *B);

/* Restarted GMRES (CGS2 Arnoldi) on a dense complex n x n system on the device. */
HSBEM_API int hsbem_dense_gmres(hsbem_context *context, uint32_t n, const double *A, const double *b, double *x,
                                uint32_t restart, uint32_t maximum_iterations, double tolerance,
                                double *relative_residual, uint32_t *iterations);
/* Partial-pivot complex LU of a dense n x n matrix and solve of nrhs right-hand sides in place. */
HSBEM_API int hsbem_dense_lu_solve(hsbem_context *context, uint32_t n, const double *A, uint32_t nrhs, double *B);
/* The same LU solve on caller-owned DEVICE buffers, allocating no device memory: the pivots and the
 * singular-pivot status live in a static device array of the library (256 KiB + 256 B per GPU, held
 * until the process ends). No counterpart in the reference: it is for callers that cannot allocate
 * (code under stream capture, tight serving loops, harnesses that forbid hipMalloc). A (column-major
 * complex n x n, interleaved re/im) is overwritten by its LU factors and B (n x nrhs) by the
 * solutions; n <= 65536 (larger orders are refused with HSBEM_INVALID_ARGUMENT). The work runs on the
 * context's stream, which is synchronised before the call returns; the caller orders its own writes
 * of A and B before the call. Calls that overlap on one GPU share its array and race. */
HSBEM_API int hsbem_dense_lu_solve_noalloc(hsbem_context *context, uint32_t n, double *A, uint32_t nrhs, double *B);

#ifdef __cplusplus
}
#endif

#endif /* HSBEM_H */
