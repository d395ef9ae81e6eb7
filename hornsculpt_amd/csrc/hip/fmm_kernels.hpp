// Host-callable launchers of the FMM far-field kernels (fmm.hip).
#pragma once

#include "kernels.hpp"

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace hsbem::dev {

inline constexpr int kMaxLeafOrder = 64;   // leaf orders with the smaller radial register table (up to kMaxBigOrder)
inline constexpr int kMaxLeafDofs = 512;   // P1 DOFs of one leaf
inline constexpr int kSmallFmmOrder = 32;  // smaller instantiation of the staged translations
inline constexpr int kMaxFmmOrder = 40;    // translations with the whole d_l block staged: (2p+1)^2 doubles
inline constexpr int kMidFmmOrder = 48;    // largest whole-block instantiation (one workgroup per CU)
inline constexpr int kMaxWideOrder = 64;   // tiled d_l stage above kMidFmmOrder
inline constexpr int kMaxBigOrder = 128;   // any translated expansion (reduction-tiled kernels above kMaxWideOrder)
inline constexpr int kFmmChunk = 16;       // links per rotation chunk
inline constexpr int kFmmJacobi = 64;      // block width served by the LDS factorisation (others: the wide path)

// One leaf of the support tree with P1 DOFs. Offsets count complex values.
struct FmmLeaf {
    double center[3];
    double radius;               // basis normalisation scale of the leaf's level
    std::uint32_t order;         // expansion order of the leaf's level
    std::uint32_t first_dof, dof_count; // into the leaf DOF list (tree order)
    std::uint32_t pad;
    std::uint64_t emission;      // E and T rows of this leaf: dof_count x (order+1)^2
    std::uint64_t multipole, local; // into the expansion pool
};

// (leaf, support panel): local DOF index of each panel corner in the leaf, or -1.
struct FmmLeafItem {
    std::uint32_t leaf, panel;
    std::int32_t corner[3];
    std::uint32_t pad;
};

struct FmmLeafMomentsLaunch {
    const double *geo;            // panel SoA (original numbering)
    std::uint32_t nt;
    const double *rule;           // rule_points x (b0, b1, b2, w), weights sum to 1/2
    std::uint32_t rule_points;    // <= 64
    double k, alpha_re, alpha_im; // Burton-Miller alpha = -i/k
    const double *row_scale;      // 1 / lumped mass, original numbering
    const FmmLeaf *leaves;
    const std::uint32_t *leaf_dofs;
    const FmmLeafItem *items;
    std::uint32_t item_count;
    std::uint32_t max_order;      // largest leaf order
    double *emission, *testing;   // complex, zeroed before the launch
    int *error;                   // set when a radial table leaves its bounded domain
};

// A direction of translation: Wigner small-d blocks of degrees 0..order and the
// azimuth phases exp(i m phi), m = -order..order.
struct FmmRotation {
    std::uint64_t small_d; // doubles
    std::uint64_t azimuth; // complex values
    std::uint32_t order, pad;
};

// One translation into a destination: source expansion (pool offset), the
// orders, the rotation of the link axis, the coaxial plan and the link's
// workspace slot (complex offset, room for the larger of both orders).
struct FmmLink {
    std::uint64_t source, slot;
    std::uint32_t source_order, target_order, rotation, axial;
};

// Consecutive links (in the direction-sorted order) sharing one rotation.
struct FmmChunk {
    std::uint32_t rotation, first, count, max_order;
};

struct FmmDestination {
    std::uint64_t target;     // pool offset of the destination expansion (overwritten)
    std::uint32_t order;      // destination order
    std::uint32_t first_link, link_count, pad;
};

struct FmmTranslateLaunch {
    const FmmDestination *destinations; // of this launch
    const FmmLink *links;               // of this launch (destination order)
    const std::uint32_t *order;         // link indices sorted by rotation
    const FmmChunk *chunks;             // rotation chunks
    const std::uint32_t *axial_order;   // link indices sorted by coaxial plan
    const FmmChunk *axial_chunks;       // chunks sharing a coaxial plan (rotation field = plan)
    double *work_y, *work_z;            // link workspaces (complex)
    const FmmRotation *rotations;
    const double *small_d;
    const double *azimuth;          // complex
    const double *axial;            // complex m-block coefficients of every plan
    const std::uint64_t *axial_offsets; // complex offset of each plan
    const double *source_expansions;
    double *target_expansions;
    // FP32-bulk form: work_y, work_z and the expansions hold float pairs (orders up to
    // 48; the tables stay FP64 and are rounded when staged)
    bool fp32{};
};

// Far part of B q by direct pair quadrature over the flux panels (original numbering).
struct FmmFarFluxLaunch {
    const double *geo;
    const std::uint32_t *nodes;          // 3 x nt, mesh vertex ids
    std::uint32_t nt, nv, nt_all;        // nt_all: panels per excitation block of q (= nt)
    const std::uint32_t *flux_panels;    // active driven panels
    std::uint32_t flux_count;
    const std::uint32_t *pressure_leaf;  // per vertex
    const std::uint32_t *flux_leaf;      // per panel
    const std::uint32_t *near_offsets, *near_sources; // per pressure leaf: near flux leaves
    const double *rule;                  // (b0, b1, b2, w), weights sum to 1/2
    std::uint32_t rule_points;
    double k;
    const double *row_scale;             // 1 / lumped mass
    const double *q;                     // per excitation, per panel (complex)
    std::uint32_t excitations;
    double *out;                         // nv x excitations (complex), accumulated
};

// ---- plane-wave levels (fmm_hf.hip)

inline constexpr int kMaxHfAzimuths = 2048; // azimuths of a plane-wave rule: 2 (order + 1), order <= 1023

// Sphere rule of a plane-wave level: rings x azimuths samples, polar-major.
struct HfRule {
    std::uint64_t legendre;          // doubles: ring r, degree l, order m >= 0 at r * T + l(l+1)/2 + m,
                                     // T = (bmax+1)(bmax+2)/2: orthonormal P_lm(cos theta) (Y_lm at phi = 0)
    std::uint64_t polar;             // doubles: per ring (cos theta, sin theta, sample weight)
    std::uint32_t rings, azimuths, bmax, pad;
};

// One angular transfer into a destination. Offsets count complex values.
struct HfLink {
    std::uint64_t source;            // pool offset of the source signature or scaled multipole
    std::uint64_t target;            // pool offset of the scaled local (down hand-off), else unused
    std::uint64_t fourier;           // workspace: max(in, out rings) x (2 order + 1)
    std::uint64_t harmonic;          // workspace: (order + 1)^2
    double shift[3];                 // +k (c_child - c_parent) upward, -k (c_child - c_parent) downward
    double ka;                       // k a of the scaled spherical level of a hand-off
    std::uint32_t in_rule, out_rule; // rules of the source and destination signatures
    std::uint32_t order;             // bandlimit (the spherical order of a hand-off)
    std::uint32_t pad;
};

enum class HfTransferKind : std::uint32_t {
    up_handoff = 0,    // scaled multipole of a spherical child -> plane-wave parent signature
    up_resample = 1,   // plane-wave child -> plane-wave parent signature
    down_handoff = 2,  // plane-wave parent local -> scaled local of a spherical child (added)
    down_resample = 3, // plane-wave parent local -> plane-wave child local (added)
};

struct HfTransferLaunch {
    const HfRule *rules;
    const double *legendre, *polar;
    const HfLink *links;               // of this launch, in destination order
    const FmmDestination *destinations; // target = pool offset of the destination signature
    double *fourier_ws, *harmonic_ws;
    const double *source_expansions;
    double *target_expansions;
};

struct HfDiagonalLink {
    std::uint64_t source;       // pool offset of the source signature
    std::uint64_t coefficients; // offset of the direction class's diagonal translation
};

struct HfDiagonalLaunch {
    const FmmDestination *destinations; // target: pool offset of the destination local signature (overwritten)
    const HfDiagonalLink *links;
    const double *coefficients;         // complex
    std::uint64_t samples;              // samples of the level's rule
    const double *source_expansions;
    double *target_expansions;
};

// A leaf on a plane-wave level: its signature is emitted and tested directly on the
// level's sphere rule (no moment tables).
struct HfLeaf {
    double center[3];
    std::uint64_t multipole, local;  // pool offsets of the outgoing and incoming signatures
    std::uint32_t rule;              // HfRule of the level
    std::uint32_t first_item, item_count; // support panels of the leaf in the HF item list
    std::uint32_t first_dof, dof_count;   // into the leaf DOF list
    std::uint32_t pad;
};

// A support panel of a plane-wave leaf: geometry, corner DOFs (local index or -1).
struct HfLeafItem {
    double v0[3], e1[3], e2[3], normal[3];
    double area2;                    // twice the panel area (rule weights sum to 1/2)
    std::int32_t corner[3];
    std::uint32_t leaf;
};

struct HfLeafLaunch {
    const HfLeaf *leaves;
    std::uint32_t leaf_count;
    const HfLeafItem *items;
    std::uint32_t item_count;
    const std::uint32_t *leaf_dofs;
    const double *row_scale;         // per leaf DOF entry (1 / lumped mass)
    const HfRule *rules;
    const double *polar;
    const double *rule;              // panel rule: rule_points x (b0, b1, b2, w)
    std::uint32_t rule_points;       // <= 64
    std::uint32_t max_samples;       // largest rule sample count of the leaves
    double k;
    double *field;                   // workspace: item_count x rule_points complex
    double *expansions;              // pool
};

// Emission: F(d) = i k sum_panels sum_q 2A w_q (n.d) p(y_q) exp(i k d.(y_q - c)).
void launch_hf_leaf_emit(const HfLeafLaunch &launch, const double *p, hipStream_t stream);
// Testing: y_i += (1/m_i) sum_panels sum_q 2A w_q phi_i(y_q) sum_d w_d (n.d - 1) exp(-i k d.(y_q - c)) L(d).
void launch_hf_leaf_test(const HfLeafLaunch &launch, double *y, hipStream_t stream);

// in_rings / out_rings: the largest ring counts of the links' input / output rules;
// out_samples: samples of the destinations' rule.
void launch_hf_transfer(const HfTransferLaunch &launch, HfTransferKind kind, std::uint32_t link_count,
                        std::uint32_t destination_count, std::uint32_t in_rings, std::uint32_t out_rings,
                        std::uint32_t out_samples, hipStream_t stream);
void launch_hf_diagonal(const HfDiagonalLaunch &launch, std::uint32_t destination_count, hipStream_t stream);

// Direct far part of A x for the direct Galerkin audit: every P1 DOF pair whose
// pressure leaves are not near, integrated with the panel rule (original numbering).
struct FmmDirectLaunch {
    const double *geo;                 // panel SoA
    const std::uint32_t *nodes;        // 3 x nt mesh vertex ids
    std::uint32_t nt, nv;
    const std::uint32_t *vertex_leaf;  // dense pressure-leaf index per vertex
    const std::uint8_t *near;          // leaf_count x leaf_count: near[target * leaf_count + source]
    std::uint32_t leaf_count;
    const double *rule;                // rule_points x (b0, b1, b2, w), weights sum to 1/2
    std::uint32_t rule_points;         // <= 64
    double k;
    const double *row_scale;           // 1 / lumped mass
    const double *x;                   // complex, nv
    double *y;                         // complex, nv, accumulated
};
// y += A_far x (block per test panel, atomics per test DOF).
void launch_direct_far(const FmmDirectLaunch &launch, hipStream_t stream);

void launch_leaf_moments(const FmmLeafMomentsLaunch &launch, hipStream_t stream);
void launch_far_flux(const FmmFarFluxLaunch &launch, hipStream_t stream);
// out[:, e] += B_near q_e (q: faces x excitations complex).
void launch_flux_csr_apply(const std::uint32_t *offsets, const std::uint32_t *columns, const double *values,
                           std::uint32_t rows, const double *q, std::uint32_t faces, std::uint32_t excitations,
                           double *out, hipStream_t stream);
// Per vector of length n: out[map[i]] = in[i] (forward) or out[i] = in[map[i]].
void launch_permute(const double *in, double *out, const std::uint32_t *map, std::uint32_t n, std::uint32_t count,
                    bool forward, hipStream_t stream);
// fp32: the expansions hold float pairs (the FP32-bulk form); the sums run in FP64.
void launch_leaf_emit(const FmmLeaf *leaves, std::uint32_t leaf_count, const std::uint32_t *leaf_dofs,
                      const double *emission, const double *p, double *expansions, hipStream_t stream,
                      bool fp32 = false);
void launch_leaf_test(const FmmLeaf *leaves, std::uint32_t leaf_count, const std::uint32_t *leaf_dofs,
                      const double *testing, const double *expansions, double *y, hipStream_t stream,
                      bool fp32 = false);
// max_order: largest source or target order of the launch's links and rotation plans
// (selects the tiled kernels above kMaxFmmOrder and the reduction-tiled ones above
// kMaxWideOrder; at most kMaxBigOrder).
// split_target: a level with few chunks spreads each chunk's degree / |m| stages over
// more workgroups, up to about split_target workgroups per launch (0: one per chunk).
void launch_translate(const FmmTranslateLaunch &launch, std::uint32_t destination_count, std::uint32_t link_count,
                      std::uint32_t chunk_count, std::uint32_t axial_chunk_count, std::uint32_t max_order,
                      std::uint32_t split_target, hipStream_t stream);
// blocks (zeroed, n rounded up to width) from the near CSR in tree order;
// position[dof] = index of dof in order.
void launch_gather_blocks(const std::uint32_t *row_offsets, const std::uint32_t *columns, const double *values,
                          const std::uint32_t *order, const std::uint32_t *position, std::uint32_t n,
                          std::uint32_t width, double *blocks, hipStream_t stream);

} // namespace hsbem::dev
