// Launch wrappers of the XPBD kernels (host side).  Every wrapper enqueues on
// `stream` and returns the hipError_t of the launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tc_device.h"

namespace tc {

struct SubstepArgs {
    float dt;          // substep length (dt / substeps)
    float gx, gy, gz;  // gravity
    float damping;     // air damping
    float tension_damping, compression_damping, shear_damping, bending_damping;
    // Projection kernels return at once when run_flag is set and *run_flag == 0
    // (the fallback iterations of the warm-started solve); nullptr: always run.
    const uint32_t *run_flag = nullptr;
};

// --- core substep (tc_xpbd.hip) ---
hipError_t launch_predict(hipStream_t s, float4 *P, const float4 *X, float4 *V, uint32_t n, const SubstepArgs &a,
                          uint32_t block);
// One colour of distance constraints. first: lambda treated as zero (it is
// cleared at every substep start, dispatch.rs:117-119).
hipError_t launch_distance(hipStream_t s, float4 *P, const float4 *X, const uint4 *rec, const float *cc, float *lam,
                           uint32_t off, uint32_t cnt, const SubstepArgs &a, bool first, bool split_compliance,
                           uint32_t block);
// solver_mode 1 (distance_atomic.wgsl:78-182): all distance constraints at once into fixed-point
// sums `acc` (zero on entry, left zero), then the averaged, over-relaxed corrections on P.
hipError_t launch_distance_jacobi(hipStream_t s, float4 *P, const float4 *X, const uint4 *rec, const float *cc,
                                  float *lam, uint32_t nd, int4 *acc, uint32_t nv, const SubstepArgs &a, bool first,
                                  bool split);
hipError_t launch_bending(hipStream_t s, float4 *P, const float4 *X, const uint4 *rec, float *lam, uint32_t off,
                          uint32_t cnt, const SubstepArgs &a, bool first, uint32_t block, const uint4 *gate = nullptr);
// sewing (tc_sewing.hip, sewing.wgsl / sewing_shrink.wgsl): rec {v0, v1, current rest bits, target rest bits}
struct SewArgs {
    float dt;              // substep length
    float compliance;      // 0 for stiffness >= 5000, else 1 / max(stiffness, 1)
    float lock;            // > 0.5: lock pairs closer than max(thickness sum, lock_dist)
    float lock_dist;
    float speed;           // shrink speed of the current rest lengths (m/s)
};
hipError_t launch_sew_shrink(hipStream_t s, uint4 *rec, float *lam, uint32_t n, const SewArgs &a);
hipError_t launch_sewing(hipStream_t s, float4 *P, const float4 *X, uint4 *rec, float *lam, uint32_t off, uint32_t cnt,
                         const SewArgs &a, uint32_t block);
// SDF colliders (set_bone_sdf_colliders, crates/cloth_core/src/simulation/mod.rs:1630-1727): an
// Rg16Float volume (pack2x16float(distance, weight) per texel, Z/Y/X order), 20 floats of
// GpuBoneInfo per collider and its world / inverse-world matrices (16 + 16 floats, each given
// as matrix[row][column] the way the reference uploads them).
struct SdfArgs {
    const uint32_t *tex = nullptr;
    uint32_t w = 0, h = 0, d = 0;
    const float *info = nullptr;
    const float *xf = nullptr;
    uint32_t nb = 0;
};
// order: the vertices in the order the lanes take them (a permutation of [0, n), or null for 0..n-1)
hipError_t launch_collision(hipStream_t s, float4 *P, float4 *X, uint32_t n, const uint32_t *order,
                            const Collider *cols, uint32_t ncols, const MeshTri *tris, const float4 *bounds,
                            uint32_t ntris, const float4 *cbox, int single_sided_recovery, float dt, uint32_t block,
                            const SdfArgs &sdf);
hipError_t launch_update_vel(hipStream_t s, const float4 *P, float4 *X, float4 *V, uint32_t n, float dt,
                             uint32_t block);
hipError_t launch_extract_positions(hipStream_t s, const float4 *X, float *out, uint32_t n);
// Compliant pins (pin.wgsl:55-92) on P: pin_t = (target xyz, weight), one distinct vertex per pin.
hipError_t launch_pins(hipStream_t s, float4 *P, const uint32_t *pin_v, const float4 *pin_t, uint32_t npins,
                       const SubstepArgs &a);
hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t value, size_t n);
// Residual of the XPBD system after the last sweep (double sums, deterministic order):
// out[0] = sum h^2, out[1] = sum L0^2, out[2] = max |h|/L0, out[3] = sqrt(out[0] / out[1]).
// part holds 3 * 1024 doubles.  With gate != nullptr the kernel also writes gate[0] = (out[3] >
// threshold) and counts checks in gate[1], fallbacks in gate[2], fallback chunks in gate[3] and, for
// the final check after the last chunk (final_check), substeps that end above the threshold in gate[4].
// X (the committed positions) is read only with constraint damping in `a`, for the damping term.
hipError_t launch_residual(hipStream_t s, const float4 *P, const float4 *X, const uint4 *drec, const float *dcc,
                           const float *dlam, uint32_t nd, const uint4 *brec, const float *blam, uint32_t nb,
                           const SubstepArgs &a, bool split, double *part, double *out, uint32_t *gate,
                           float threshold, const uint32_t *run_flag = nullptr, uint32_t *gate_host = nullptr,
                           bool final_check = false);

// --- one iteration per launch over cone tiles (tc_fused.hip, plan from tc_tiles.cpp) ---
struct FusedArgs {
    uint32_t ntiles, nphases, nd_phases, max_ext;
    uint32_t ns_phases;                          // sewing colours after the distance colours
    float sew_comp, sew_lock, sew_lock_dist;     // as SewArgs::compliance, lock, lock_dist
    // 1: fast projection (hardware reciprocal square root and reciprocal, one Newton-Raphson step
    // each, instead of the IEEE square root and divisions); 0: the exact operation sequence
    uint32_t fast;
    float dtc[16], dcc[16], bcomp[16];           // compliance classes
    float dalpha_t[16], dalpha_c[16], balpha[16];  // compliance / (dt * dt) per class for this launch
};
struct FusedPlanDev {
    const uint32_t *ext_off, *ext_vid, *n_own, *ph_off, *items;
    uint4 *srec;          // sewing records (current rest length written by the lock), or null
    const uint4 *brec;    // bending records: the rest length of seam-gated items
    const uint4 *bgate;   // seam gate table (hinge vertices, gap threshold), or null
};
// Sewing lambdas (slam_in / slam_out) are double-buffered like the others.
hipError_t launch_tile_iteration(hipStream_t s, const float4 *Pin, float4 *Pout, const float4 *X,
                                 const FusedPlanDev &p, const float *dlam_in, float *dlam_out, const float *blam_in,
                                 float *blam_out, const float *slam_in, float *slam_out, const FusedArgs &fa,
                                 const SubstepArgs &a, bool first, uint32_t block);
hipError_t set_tile_lds_limit(size_t bytes);
// fills the alpha~ tables of `fa` for substep length dt_sub (same float operations as the kernels)
void fused_alphas(FusedArgs &fa, float dt_sub);

// --- spatial hash + self collision (tc_selfcol.hip) ---
struct HashBufs {
    uint32_t *keys;     // n
    uint32_t *counts;   // table + table/1024: bucket counts, then scatter cursors; the tail holds the scan's row totals
    uint32_t *starts;   // table + 1
    uint32_t *sorted;   // n
};
// Vertex hash over P.xyz (spatial_grid_sort.wgsl) and edge hash over the edge midpoints
// (spatial_grid_sort_edges.wgsl), built together, table buckets each; within a cell indices
// are ascending.
hipError_t launch_hashes(hipStream_t s, const float4 *P, uint32_t nv, const uint2 *edges, uint32_t ne, float cell,
                         uint32_t table, const HashBufs &vh, const HashBufs &eh);

struct SelfColArgs {
    float cell;
    uint32_t table;
    uint32_t nv, ne;
    float relief_factor, max_disp_ratio;
    int enable_relief, exclude_neighbors;
    uint32_t max_search_iterations;
    int untangle;  // normal untangling in effect: enabled and the cloth has more than one layer
};
struct SelfColTopo {
    const float *local_len;
    const uint32_t *th_off, *th_idx;
    const float *th_rest;
    const uint32_t *star_off;
    const uint2 *star;
    const uint2 *edges;
    const uint32_t *edge_bbits;
    // multi-layer cloth (read only when SelfColArgs::untangle is set): layer ids, the area-weighted
    // vertex normals of the substep's start positions, the 1-ring adjacency
    const uint32_t *layer;
    const float4 *normals;
    const uint32_t *adj_off, *adj_idx;
};
// compute_normals.wgsl: per vertex the normalised sum of its star triangles' cross products at X
hipError_t launch_compute_normals(hipStream_t s, const float4 *X, const SelfColTopo &t, float4 *normals, uint32_t nv);
// self_collision_untangle.wgsl: an outer-layer vertex recorded behind an inner triangle is pushed out
hipError_t launch_sc_untangle(hipStream_t s, const float4 *P, const float4 *X, const SelfColTopo &t, Accum *acc,
                              const SelfColArgs &a);
hipError_t launch_sc_solve(hipStream_t s, const float4 *P, const float4 *X, const SelfColTopo &t,
                           const HashBufs &vh, Accum *acc, const SelfColArgs &a);
hipError_t launch_sc_solve_ee(hipStream_t s, const float4 *P, const float4 *X, const SelfColTopo &t,
                              const HashBufs &eh, Accum *acc, const SelfColArgs &a);
hipError_t launch_sc_apply(hipStream_t s, float4 *P, const float *local_len, Accum *acc, const SelfColArgs &a,
                           uint32_t block);

// Wrinkle field (WrinkleFieldParams, simulation/types.rs:110-139)
struct WrinkleArgs {
    float origin[3] = {0.0f, 0.0f, 0.0f}, influence_radius = 0.0f;
    float axis[3] = {0.0f, 0.0f, 0.0f}, bone_radius = 0.0f;
    float normal[3] = {0.0f, 0.0f, 0.0f}, stiffness = 0.0f;
    float binormal[3] = {0.0f, 0.0f, 0.0f}, blend = 0.0f;
    float z_min = 0.0f, z_max = 0.0f, r_min = 0.0f, r_max = 0.0f;
    float valley_window = 0.015f, crest_window = 0.020f;
};

// Collider-edge collision (edge_collision.wgsl:153-276): pushes into the accumulators' pin channel.
struct EdgeColArgs {
    float cell;
    uint32_t table;
    float margin_scale, margin_offset;
};
hipError_t launch_edge_collision(hipStream_t s, const float4 *P, const float4 *X, const uint2 *edges, uint32_t ne,
                                 const ColliderEdge *ce, uint32_t nce, const uint32_t *starts, const uint32_t *sorted,
                                 const EdgeColArgs &a, Accum *acc);

// Pair cache (self_collision_algorithm 1).  Slots are vertex-major: vertex i owns
// [i * quota, (i + 1) * quota) of each list; an unused slot has 0xFFFFFFFF in its first word.
// V-T slot: (vertex, triangle c0, c1, c2 with c0 the smallest index); E-E slot: (vertex, its
// neighbour, min(j, vj), max(j, vj)).  counters: [V-T evaluated, E-E evaluated, V-T dropped,
// E-E dropped] (the caller clears them before a collect).
struct PairArgs {
    float cell;
    uint32_t table, nv;
    float safety_margin;
    int exclude_neighbors;
    float mode_f;  // 1 with the velocity horizon (margin mode 1), 0 without
    float dt_frame, horizon_scale, max_horizon;
    uint32_t quota_vt, quota_ee;  // 1..8
};
struct PairTopo {
    const uint32_t *adj_off, *adj_idx, *island;
};
hipError_t launch_pair_collect(hipStream_t s, const float4 *P, const float4 *X, const float4 *V, const SelfColTopo &T,
                               const PairTopo &PT, const HashBufs &vh, uint4 *vt, uint4 *ee, uint32_t *counters,
                               const PairArgs &a);
hipError_t launch_pair_solve_vt(hipStream_t s, const float4 *P, const float4 *X, const uint4 *vt, uint32_t slots,
                                Accum *acc, uint32_t nv, const SelfColTopo &T, int untangle);
hipError_t launch_pair_solve_ee(hipStream_t s, const float4 *P, const float4 *X, const uint4 *ee, uint32_t slots,
                                Accum *acc, uint32_t nv, const SelfColTopo &T, int untangle);

}  // namespace tc
