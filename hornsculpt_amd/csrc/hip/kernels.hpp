// Host-callable launchers of the device kernels (plain C++ declarations).
#pragma once

#include "../host/plan.hpp"
#include "../host/rules.hpp"

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace hsbem::dev {

// Panel SoA fields (geo[field][nt]) and bound fields (bounds[field][nt]).
enum : int { G_V0 = 0, G_E1 = 3, G_E2 = 6, G_N = 9, G_C0 = 12, G_AREA = 21, G_COUNT = 22 };
enum : int { B_C = 0, B_R = 3, B_H = 4, B_COUNT = 5 };

// Error codes of the reference kernels (cuda_solver.cu:705-715).
enum : int { ERR_DISTANCE = 1, ERR_OVERFLOW = 2, ERR_FIELD = 3, ERR_ESCAPED = 4, ERR_RANGE = 5 };

// Rule table layout (TrianglePoint = bary[3], weight): degree 4, 6, 8, 14.
enum : int { kRule4 = 0, kRule6 = 6, kRule8 = 18, kRule14 = 34, kRuleTableCount = 98 };

struct PanelView {
    const double *geo;
    const std::uint32_t *nodes;
    const double *bounds;
    const std::uint8_t *shape;
    std::uint32_t nt, nv;
};

struct AssemblyArgs {
    PanelView panels;
    const double *row_scale;        // nv, 1/lumped mass, internal node order
    const std::int32_t *flux_column; // nt, column of BD or -1
    double *A;                      // nv x nv complex (plain form) or null
    float *A32;                     // nv x nv complex FP32 (mixed form) or null
    double *BD;                     // nv x nd complex
    std::uint32_t nd;
    double k, tolerance;
    unsigned regular_order;
    int *error;
    unsigned long long *samples; // optional: regular pair samples of the launch (algorithmic count)
};

// Doubles per cluster in ClusterView::far: centre (3), reach, largest panel area,
// largest |curl phi_i|, largest number of the cluster's panels sharing a node, pad.
constexpr int kClusterFar = 8;

struct ClusterView {
    const std::uint32_t *panel_offsets, *panels; // cluster -> panel ids
    const std::uint32_t *node_offsets, *nodes;   // cluster -> internal node ids
    const std::uint8_t *local;                   // [3][nt] slot of each panel vertex in its cluster node list
    const double *far;                           // kClusterFar per cluster (far-tile fixed-point bounds)
    std::uint32_t count;
};

struct NearView {
    const std::uint32_t *offsets, *trials; // near CSR by test panel
    const std::uint8_t *whole;             // all-pairs degree bytes or null
};

// One tensor-product near patch of a near job (no projection fan): the pair
// (test, trial) of the job, the patch index and the job's swap flag.
struct NearPatchItem {
    std::uint32_t test, trial, patch, swap;
};

// Near operator of the FMM path: exact near entries into CSR (original node
// numbering, rows scaled by 1/lumped mass like the dense matrix).
struct NearCsrView {
    const std::uint32_t *masks;                       // per job: P1/P1 bits 0-8 (3 i + j), P1/DP0 bits 9-11
    const std::uint32_t *row_offsets, *columns;        // pressure CSR
    double *values;                                    // complex, nnz
    const std::uint32_t *flux_offsets, *flux_columns; // P1 x active-face CSR
    double *flux_values;                              // complex
    const std::uint32_t *face_column;                 // per panel, UINT32_MAX when inactive
};

struct NearOperatorLaunch {
    AssemblyArgs args;                     // panels (original numbering), row_scale, k, tolerance, error
    NearCsrView csr;
    NearView near;                         // interaction plan near CSR / all-pairs degrees (qualification)
    const double *rules;
    const std::uint32_t *regular;          // (test, trial) pairs
    std::uint32_t regular_count;
    const TouchingJob *touching;
    std::uint32_t touching_count;
    const double *gauss;
    unsigned singular_order;
    const NearJob *fan_jobs;
    std::uint32_t fan_count;
    const NearPatchItem *patch_items;
    std::uint32_t patch_count;
    const PairPatch *patches;
    // per-list mask arrays (indexed like each list)
    const std::uint32_t *regular_masks, *touching_masks, *fan_masks, *patch_masks;
};

void launch_near_operator(const NearOperatorLaunch &launch, hipStream_t stream);
// y[r] = sum_c values[r][c] x[c] over the pressure CSR (complex FP64).
void launch_csr_apply(const std::uint32_t *row_offsets, const std::uint32_t *columns, const double *values,
                      std::uint32_t rows, const double *x, double *y, hipStream_t stream);

struct AssemblyLaunch {
    AssemblyArgs args;
    ClusterView clusters;
    // (test cluster, trial cluster) tiles: the first far_tiles are far tiles whose
    // pairs all qualify for degree 4 at this wavenumber (uniform path).
    const std::uint32_t *tiles;
    std::uint32_t tile_count, far_tiles;
    NearView near;
    const double *rules;                   // kRuleTableCount x 4 doubles
    const TouchingJob *touching;
    std::uint32_t touching_count;
    const double *gauss;                   // singular Gauss rule (position, weight)
    unsigned singular_order;
    const NearJob *near_jobs;              // jobs with a projection-fan patch (per-job kernel)
    std::uint32_t near_count;
    const NearPatchItem *near_items;       // tensor patches of the other jobs (per-patch kernel)
    std::uint32_t near_item_count;
    const PairPatch *patches;
    // regular tiles: trial panels per wavefront block (4, 8 or 16) of the general and
    // of the far tiles
    std::uint32_t trial_block = 8, far_trial_block = 8;
};

// ---- adaptive near-pair patch estimates (estimator.hip) ----

struct EstimatorPanel {
    FanPanel fan;    // vertices (recentred), unit normal, area
    double max_edge;
};
struct EstimatorTask {
    std::uint32_t test, trial;
    PairPatch patch;
};
struct EstimatorResult {
    double integral[72]; // the higher rule's 36 moments (re, im), moment m = c * 9 + i * 3 + j
    double outer_error, inner_error;
    int status;          // 0 ok, 1 coincident sample points (overlapping panels), 2 non-finite moment
};
struct EstimatorLaunch {
    const EstimatorPanel *panels;
    const EstimatorTask *tasks;
    std::uint32_t count;
    const TrianglePoint *rules; // triangle rules of degree 6 (12), 8 (16), 12 (49), 14 (64) in order
    double k;
    EstimatorResult *out;
};
// projection: every task of the launch is a projection patch, else none is.
void launch_estimate_patches(const EstimatorLaunch &launch, bool projection, hipStream_t stream);

void launch_assembly(const AssemblyLaunch &launch, hipStream_t stream);
void launch_regular(const AssemblyLaunch &launch, hipStream_t stream);
void launch_touching(const AssemblyLaunch &launch, hipStream_t stream);
void launch_near(const AssemblyLaunch &launch, hipStream_t stream);
void launch_mass(const AssemblyLaunch &launch, hipStream_t stream);
void launch_flux_rhs(const double *BD, std::uint32_t nv, std::uint32_t nd, const double *q, std::uint32_t nrhs,
                     double *rhs, hipStream_t stream);

// ---- dense linear algebra (linalg.hip) ----

// y[:, e] = A x[:, e] for e < nvec (A n x n complex, column-major, FP64 or FP32
// when fp32 != 0; vectors and accumulation FP64); y may not alias x.
// skip (device int, single-vector products only): when nonzero at run time the
// kernels return at once (a captured solver step after its loop has finished).
void launch_gemv(const void *A, bool fp32, std::uint32_t n, const double *x, double *y, std::uint32_t nvec,
                 double *scratch, hipStream_t stream, const int *skip = nullptr);
// Scratch doubles launch_gemv needs for n and nvec.
std::size_t gemv_scratch_doubles(std::uint32_t n, std::uint32_t nvec);

// out[j] = V[:, j]^H w for j < m (V n x m column-major); out complex m.
void launch_project(const double *V, std::uint32_t n, std::uint32_t m, const double *w, double *out,
                    double *scratch, hipStream_t stream);
// w -= V h (h complex m).
void launch_subtract_combination(const double *V, std::uint32_t n, std::uint32_t m, const double *h, double *w,
                                 hipStream_t stream);
// out[0] = ||x||_2^2 (one complex vector of length n, weighted by mass if given).
void launch_norm2(const double *x, std::uint32_t n, const double *weight, double *out, double *scratch,
                  hipStream_t stream);
std::size_t reduction_scratch_doubles(std::uint32_t n, std::uint32_t m);

// Complex LU with partial pivoting in place (A n x n), pivots n.
void lu_factor(double *A, std::uint32_t n, std::int32_t *pivots, int *info, double *work, hipStream_t stream);
std::size_t lu_work_doubles(std::uint32_t n);
// Solve with the factors for nrhs right-hand sides in place (B n x nrhs).
void lu_solve(const double *LU, std::uint32_t n, const std::int32_t *pivots, double *B, std::uint32_t nrhs,
              hipStream_t stream);

// Device-resident GMRES restart cycle (one right-hand side): the step index, the
// Hessenberg columns, the Givens rotations, g and the stop flag live in `small` /
// `state`, so one Arnoldi step (load, operator, orthogonalize) is a fixed launch
// sequence a HIP graph can replay; every kernel of a step returns at once after the
// cycle has stopped. state: [j, done, used, limit, error]; small layout: h1 at 0
// ((restart+1) complex, reused for the cycle's y), g and H at the offsets below
// (H column-major, stride restart+1).
struct GmresDeviceCycle {
    double *basis;    // (restart + 1) x n complex
    double *current;  // n complex: v_j
    double *w;        // n complex: the operator's output being orthogonalized
    double *scratch;  // >= gmres_cycle_scratch_doubles(n, restart)
    double *small;    // gmres_cycle_small_doubles(restart)
    int *state;       // 5 ints
    std::uint32_t n, restart;
};
std::size_t gmres_cycle_small_doubles(std::uint32_t restart);
std::size_t gmres_cycle_scratch_doubles(std::uint32_t n, std::uint32_t restart);
std::size_t gmres_cycle_h_offset(std::uint32_t restart);
std::size_t gmres_cycle_g_offset(std::uint32_t restart);
// Starts a cycle: g = (beta, 0, ...), the absolute target, at most `limit` steps.
void gmres_cycle_begin(const GmresDeviceCycle &c, double beta, double target, std::uint32_t limit, hipStream_t s);
// current = v_j.
void gmres_cycle_load(const GmresDeviceCycle &c, hipStream_t s);
// CGS2 of w against v_0..v_j, v_{j+1} = w / |w|, column j rotated, j += 1.
void gmres_cycle_orthogonalize(const GmresDeviceCycle &c, hipStream_t s);

// LU factorization and solve (as lu_factor + lu_solve) whose pivots and status
// word live in a static device array instead of allocated memory: A (n x n,
// column-major complex) is overwritten by its factors, B (n x nrhs) by the
// solutions; n <= noalloc_lu_max_order(); synchronises the stream; calls that
// overlap on one device race on the shared array.
void dense_lu_solve_noalloc(int device, hipStream_t stream, std::uint32_t n, double *A, std::uint32_t nrhs,
                            double *B);
std::uint32_t noalloc_lu_max_order();

// r = b - A x, scaled by weight (lumped mass) into weighted; r unscaled into residual.
// Block-Jacobi of any width W <= kMaxJacobiWidth (a power of two) over gathered
// blocks (nb x W x W complex, column-major per block, factored in place by a panel
// LU; pivots nb x W); the inverses are column-major W x W per block and
// out = M^{-1} v with block rows order[W b .. W b + W) when order is given.
inline constexpr std::uint32_t kMaxJacobiWidth = 2048;
std::size_t jacobi_wide_doubles(std::uint32_t n, std::uint32_t width);
void jacobi_factor_wide(double *blocks, std::uint32_t n, std::uint32_t width, double *inverses, std::int32_t *pivots,
                        hipStream_t stream);
void jacobi_apply_wide(const double *inverses, std::uint32_t n, std::uint32_t width, const double *v, double *out,
                       hipStream_t stream, const std::uint32_t *order = nullptr);
// The diagonal W x W blocks of a dense column-major complex matrix (FP32 or FP64) in
// the gathered layout of jacobi_factor_wide (FP64); the tail block's padding is left.
void jacobi_gather_dense(const void *A, bool fp32, std::uint32_t n, std::uint32_t width, double *blocks,
                         hipStream_t stream);
// Block-Jacobi preconditioner of 64 x 64 diagonal blocks (consecutive unknowns):
// explicit FP64 inverses per block (LU with partial pivoting), and out = M^{-1} v.
std::size_t jacobi_factor_doubles(std::uint32_t n);
void jacobi_factor(const void *A, bool fp32, std::uint32_t n, double *inverses, hipStream_t stream);
// Inverses of gathered blocks (complex FP64, column-major per 64 x 64 block).
void jacobi_factor_gathered(const double *blocks, std::uint32_t n, double *inverses, hipStream_t stream);
// out = M^{-1} v; block rows are order[64 b .. 64 b + 64) when order is given.
void jacobi_apply(const double *inverses, std::uint32_t n, const double *v, double *out, hipStream_t stream,
                  const std::uint32_t *order = nullptr);
void launch_residual(const void *A, bool fp32, std::uint32_t n, const double *b, const double *x, std::uint32_t nvec,
                     double *residual, double *scratch, hipStream_t stream);

// Elementwise helpers.
void launch_axpy(std::uint32_t n, const double *alpha, const double *x, double *y, hipStream_t stream);
// y = s * x (complex vector of length n, real s).
// y += x (n complex)
void launch_add(std::uint32_t n, const double *x, double *y, hipStream_t stream);
// r = b - r (complex vectors of length n).
void launch_subtract_from(std::uint32_t n, const double *b, double *r, hipStream_t stream);
// y = x / sqrt(*norm2) (device-resident squared norm; zero vector when it is 0).
void launch_normalize(std::uint32_t n, const double *norm2, const double *x, double *y, hipStream_t stream);
void launch_scale(std::uint32_t n, double s, const double *x, double *y, hipStream_t stream);

// ---- fields (field.hip) ----

struct FieldArgs {
    PanelView panels;
    const double *pressure; // nv x nrhs complex, internal node order
    const double *flux;     // nt x nrhs complex (DP0 q per panel)
    std::uint32_t nrhs;
    const double *points;   // observers (local frame) or unit directions
    std::uint32_t count;
    const double *rule;     // triangle rule (bary[3], w)
    std::uint32_t rule_count;
    double k;
    double near_tolerance;
    unsigned max_depth;
    double origin[3];       // far field: phase origin (local frame)
    double reference_distance;
    double *out;            // count x nrhs complex
    int *error;
};

void launch_far_field(const FieldArgs &args, double *partials, hipStream_t stream);
void launch_observer_field(const FieldArgs &args, double *partials, hipStream_t stream);
std::size_t field_partial_doubles(std::uint32_t nt, std::uint32_t count, std::uint32_t nrhs);

} // namespace hsbem::dev
