/*
 * tc_sim.h -- C ABI of the MI355X XPBD cloth substep library (libtcloth_hip.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, no framework types.
 * Every entry point returns an int status (0 = TC_OK) unless stated; on error
 * a thread-local message is available from tc_last_error().
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference repository root).  The PyO3 class
 * `taremin_cloth_core.ClothSimulator` (crates/cloth_py/src/lib.rs:1549-3258)
 * wraps GpuClothSimulator (crates/cloth_core/src/simulation/mod.rs:57-3223);
 * this header is the boundary at that level.
 *
 * Threading: a handle is used by one thread at a time.  Each simulator owns
 * its device memory and one HIP stream.  Host inputs are copied at the call
 * (the reference copies numpy inputs into owned Vecs, lib.rs:1584-1712);
 * outputs go to caller-owned buffers.
 */
#ifndef TC_SIM_H
#define TC_SIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TC_OK = 0,
    TC_ERR_VALUE = 1,    /* bad shape / index / argument  (PyValueError in the reference) */
    TC_ERR_RUNTIME = 2,  /* device or allocation failure  (PyRuntimeError in the reference) */
    TC_ERR_UNSUPPORTED = 3,
    TC_ERR_IO = 4        /* file could not be written       (PyOSError in the reference) */
};

typedef struct tc_sim tc_sim;

/* Constructor keyword arguments of ClothSimulator::new (lib.rs:1556-1582).
 * compression_stiffness / shear_stiffness < 0 mean "not given"
 * (defaults: stiffness and 0.5 * stiffness, lib.rs:1720-1722). */
typedef struct {
    float thickness;              /* default 0.005 */
    uint32_t layer_id;            /* default 0 */
    float stiffness;              /* default 1000 */
    float bending_stiffness;      /* default 10 */
    float compression_stiffness;  /* < 0: = stiffness */
    float shear_stiffness;        /* < 0: = 0.5 * stiffness */
    uint32_t workgroup_size;      /* 0: autotuned; otherwise 64/128/256 */
    uint32_t solver_mode;         /* 0 = graph colouring (Gauss-Seidel per colour), 1 = fixed-point Jacobi */
} tc_sim_params;

/* Fills the reference defaults. */
void tc_sim_params_default(tc_sim_params *p);

/* Sewing keyword arguments of ClothSimulator::new (lib.rs:1557-1582): sewing_shrink_speed (m/s,
 * default 1.0), sewing_stiffness (< 0: not given -> 10000), enable_sewing_lock (< 0: not given ->
 * on), sewing_lock_distance (< 0: not given -> 0.02 m).  As in the reference, the stiffness and lock
 * given here are stored in the sewing records only (mesh.rs:911-941); the solver uses the simulator's
 * values, which start at 10000 and on (mod.rs:637-640) and follow tc_sim_set_sewing_stiffness /
 * tc_sim_set_enable_sewing_lock.  The lock distance is applied (lib.rs:1749-1751). */
typedef struct {
    float shrink_speed;
    float stiffness;
    int32_t lock;
    float lock_distance;
} tc_sewing_params;
void tc_sewing_params_default(tc_sewing_params *p);

/* ClothSimulator::new, lib.rs:1557-1766 -> ClothMesh::from_raw_with_opts (mesh.rs:407)
 * + GpuClothSimulator::with_options (simulation/mod.rs:363).
 * faces / inv_masses / thicknesses / layer_ids may be NULL. */
int tc_sim_create(const float *positions, uint32_t num_vertices,
                  const uint32_t *edges, uint32_t num_edges,
                  const uint32_t *faces, uint32_t num_faces,
                  const float *inv_masses, const float *thicknesses, const uint32_t *layer_ids,
                  const tc_sim_params *params, tc_sim **out);
void tc_sim_destroy(tc_sim *sim);
/* tc_sim_create with sewing springs (lib.rs:1643-1665 -> mesh.rs:706-1191): `sewing` holds
 * num_sewing vertex pairs; an index >= num_vertices is TC_ERR_VALUE.  Sewn vertices join one node of
 * the self-collision topology, seams get gated bending constraints, and the sewing springs shrink
 * and project every iteration (sewing.wgsl, sewing_shrink.wgsl). */
int tc_sim_create_sewn(const float *positions, uint32_t num_vertices, const uint32_t *edges, uint32_t num_edges,
                       const uint32_t *faces, uint32_t num_faces, const float *inv_masses, const float *thicknesses,
                       const uint32_t *layer_ids, const uint32_t *sewing, uint32_t num_sewing,
                       const tc_sewing_params *sewing_params, const tc_sim_params *params, tc_sim **out);
/* ClothSimulator.set_sewing_stiffness (lib.rs:2800, clamped to >= 1), set_sewing_shrink_speed
 * (lib.rs:2806, non-finite -> 1, >= 0; the current rest lengths are kept), set_enable_sewing_lock
 * (lib.rs:2817), set_sewing_lock_distance (lib.rs:2823, >= 0). */
int tc_sim_set_sewing_stiffness(tc_sim *sim, float stiffness);
int tc_sim_set_sewing_shrink_speed(tc_sim *sim, float speed);
int tc_sim_set_enable_sewing_lock(tc_sim *sim, int enable);
int tc_sim_set_sewing_lock_distance(tc_sim *sim, float distance);
/* out[5]: stiffness, shrink speed, lock (0|1), lock distance, number of sewing constraints
 * (get_sewing_shrink_speed, get_sewing_lock_distance, num_sewing_constraints: lib.rs:2811-2947). */
int tc_sim_get_sewing_state(const tc_sim *sim, float *out);
/* get_sewing_current_rest_lengths / set_sewing_current_rest_lengths (lib.rs:2901-2908 ->
 * dispatch.rs:912-950), in the constraints' colour order.  get: n must equal the constraint count;
 * set: a list of another length is ignored, as in the reference. */
int tc_sim_get_sewing_rest_lengths(tc_sim *sim, float *out, uint32_t n);
int tc_sim_set_sewing_rest_lengths(tc_sim *sim, const float *values, uint32_t n);

/* Sewing priority (mod.rs:1262-1420, lib.rs:2842-2923): a one-way latch advanced on the host that
 * scales gravity and the wrinkle field's blend (dispatch.rs:13-21, 53-60, 151-165) by 0 until a share
 * `threshold` of the sewing pairs is closer than `merge_dist` (or `max_frames` updates have passed,
 * 0 = never), then by smoothstep over `ramp_frames` updates.  Off, or without sewing pairs, the scale
 * is 1.
 *   tc_sewing_closure_ratio: the share of closed pairs among those whose indices fall inside the
 *     positions (num_vertices x 3 floats); 1 when none count (mod.rs:1381-1420).  Needs no GPU.
 *   set_options: clamps threshold to [0, 1], merge_dist to [0.0005, 0.1], ramp_frames to 120 and
 *     max_frames to 100000; a change resets the latch, an unchanged setting keeps it.
 *   set_state: restores (latched, frame, ramp_t, ratio), ramp_t and ratio clamped to [0, 1].
 *   get_state out[12]: enabled, threshold, merge_dist, ramp_frames, max_frames, latched, frame,
 *     ramp_t, ratio (cached), scale, closure ratio (1 when off), active (latch not yet reached).
 *   update: measures the closure on the given positions (no GPU read-back), advances the latch and
 *     ramp; out[3] = ratio, scale, latched (1, 1, 1 when off or without sewing pairs).
 *   tc_sim_reset clears the latch (dispatch.rs:1414-1417). */
int tc_sewing_closure_ratio(const float *positions, uint32_t num_vertices, const uint32_t *pairs, uint32_t num_pairs,
                            float merge_dist, float *ratio);
int tc_sim_set_sewing_priority_options(tc_sim *sim, int enabled, float threshold, float merge_dist,
                                       uint32_t ramp_frames, uint32_t max_frames);
int tc_sim_set_sewing_priority_state(tc_sim *sim, int latched, uint32_t frame, float ramp_t, float ratio);
int tc_sim_get_sewing_priority_state(const tc_sim *sim, double *out);
int tc_sim_update_sewing_priority(tc_sim *sim, const float *positions, uint32_t num_vertices, float *out);

/* One cloth instance of a batch: the arguments of tc_sim_create for that cloth (vertex indices are
 * local to the instance; inv_masses / thicknesses / faces may be NULL). */
typedef struct {
    const float *positions;
    uint32_t num_vertices;
    const uint32_t *edges;
    uint32_t num_edges;
    const uint32_t *faces;
    uint32_t num_faces;
    const float *inv_masses;
    const float *thicknesses;
} tc_instance;

/* Batch of independent cloth instances in one simulator: no reference counterpart (the reference
 * runs one GpuClothSimulator per cloth, simulation/mod.rs:363).  Each instance is built and
 * coloured on its own; the simulator holds their disjoint union, colour c of the union being
 * colour c of every instance, so one launch per colour (or per iteration) advances all of them
 * with each instance's own Gauss-Seidel order.  Tiles of the per-iteration path never mix two
 * instances.  Self-collision starts disabled (instances that overlap in space would collide);
 * colliders, gravity and damping are shared.  Every other tc_sim_* call works on the union;
 * tc_sim_instance_range gives each instance's vertex range in it. */
int tc_sim_create_batch(const tc_instance *instances, uint32_t num_instances, const tc_sim_params *params,
                        tc_sim **out);
int tc_sim_instance_count(const tc_sim *sim, uint32_t *num_instances);
int tc_sim_instance_range(const tc_sim *sim, uint32_t instance, uint32_t *first_vertex, uint32_t *num_vertices);

/* ClothSimulator.step(dt, substeps, solver_iterations), lib.rs:1768-1786 -> dispatch.rs:610-671.
 * substeps 0 -> 20 (base_substeps), or the adaptive policy's choice when it is on; iterations 0 -> keep current.
 * Asynchronous on the simulator's stream; *effective_substeps may be NULL. */
int tc_sim_step(tc_sim *sim, float dt, uint32_t substeps, uint32_t iterations, uint32_t *effective_substeps);
/* Adaptive substeps and iterations: set_adaptive_substep_options(enabled, min, max), lib.rs:1797-1800
 * -> simulation/mod.rs:723-744; set_auto_compensate_iterations, lib.rs:1950 -> mod.rs:760;
 * set_strain_adaptive_options(enabled, tolerance), lib.rs:1961 -> mod.rs:775 (tolerance clamped to
 * [1e-4, 0.1]); note_external_displacement(disp), lib.rs:1803 -> mod.rs:787.  With the policy on,
 * tc_sim_step / _step_async / _step_buffered pick the frame's substeps (CFL on the largest per-axis
 * displacement of the last frame read back or noted, a free-fall floor, a strain floor, hysteresis of
 * +-2 per frame with an emergency jump, dispatch.rs:476-570) and its iterations (a propagation budget
 * spread over the substeps, dispatch.rs:573-607); the effective substeps come back through
 * *effective_substeps.  The feedback is read from the last frame's positions, copied asynchronously
 * after the frame to pinned host memory and consumed at the next step once it has landed. */
int tc_sim_set_adaptive_substep_options(tc_sim *sim, int enabled, uint32_t min_substeps, uint32_t max_substeps);
int tc_sim_set_auto_compensate_iterations(tc_sim *sim, int enabled);
int tc_sim_set_strain_adaptive_options(tc_sim *sim, int enabled, float tolerance);
/* set_enable_compact_readback(enabled): whether step copies the positions for the policy's feedback. */
int tc_sim_set_enable_compact_readback(tc_sim *sim, int enabled);
int tc_sim_note_external_displacement(tc_sim *sim, float displacement);
/* get_effective_substeps / get_effective_solver_iterations / get_last_step_max_strain, lib.rs:1787-1795,
 * 1955: out[8] = [effective substeps of the last step, its iterations, P95 strain of the last feedback,
 * largest displacement of the last feedback, displacement the last decision used, CFL characteristic
 * length, mesh diameter in edge hops, mean edge length]. */
int tc_sim_frame_policy_state(const tc_sim *sim, double *out);
/* ClothSimulator.step_async(dt, substeps, solver_iterations), lib.rs:1823-1831 -> dispatch.rs:738-807:
 * the frame is enqueued and its positions copied to pinned host memory without waiting; a step_async
 * still uncollected is collected first.  fetch_positions(out), lib.rs:1833-1845 -> dispatch.rs:809-863:
 * waits for that copy, writes min(n_floats, 3N) floats and sets *ok = 1; *ok = 0 when nothing is pending. */
int tc_sim_step_async(tc_sim *sim, float dt, uint32_t substeps, uint32_t iterations);
int tc_sim_fetch_positions(tc_sim *sim, float *out, size_t n_floats, int *ok);
/* step_buffered(dt, substeps, solver_iterations) -> frames held, lib.rs:1847-1855 -> dispatch.rs:1293-1339:
 * each frame's positions go to a device ring of 8 frames (the last slot is overwritten once full);
 * fetch_buffered_positions(out) -> frames, lib.rs:1857-1870 -> dispatch.rs:1345-1398: one copy of all
 * held frames (frame-major, 3N floats each), then the ring is empty; get_buffered_frame_count,
 * clear_frame_buffer, lib.rs:1872-1880. */
int tc_sim_step_buffered(tc_sim *sim, float dt, uint32_t substeps, uint32_t iterations, uint32_t *count);
int tc_sim_fetch_buffered_positions(tc_sim *sim, float *out, size_t n_floats, uint32_t *count);
int tc_sim_get_buffered_frame_count(const tc_sim *sim, uint32_t *count);
int tc_sim_clear_frame_buffer(tc_sim *sim);

/* ClothSimulator.step_single_substep(dt_sub), lib.rs:1808-1813 */
int tc_sim_step_single_substep(tc_sim *sim, float dt_sub);
/* ClothSimulator.get_positions(out), lib.rs:2110-2121 (blocking, n_floats >= 3N) */
int tc_sim_get_positions(tc_sim *sim, float *out, size_t n_floats);
/* ClothSimulator.set_positions_and_velocities(pos, vel=None), lib.rs:2124-2153 */
int tc_sim_set_positions_velocities(tc_sim *sim, const float *positions, const float *velocities, uint32_t n);
/* ClothSimulator.reset(), dispatch.rs:1411-1427 */
int tc_sim_reset(tc_sim *sim);

/* Dynamic pins, lib.rs:2616-2675 -> simulation/mod.rs:2768-2905.  set_pins: set_pin (n = 1) and
 * set_pins_batch; targets hold 3 floats per pin; an index >= num_vertices is ignored; a vertex's
 * inverse mass becomes its base inverse mass * (1 - clamp(weight, 0, 1)) (a non-finite weight
 * counts as 0).  Each iteration, and once after the collider pass, a pin pulls its vertex towards
 * the target with compliance 1e-3 (1 - w) / w, or snaps it there when its inverse mass is 0
 * (pin.wgsl:55-92).  release_pins: release_pin (n = 1) and release_pins_batch, restoring the base
 * inverse masses.  clear_pins: clear_pins. */
int tc_sim_set_pins(tc_sim *sim, const uint32_t *indices, const float *targets, const float *weights, uint32_t n);
int tc_sim_release_pins(tc_sim *sim, const uint32_t *indices, uint32_t n);
int tc_sim_clear_pins(tc_sim *sim);
int tc_sim_pin_count(const tc_sim *sim, uint32_t *n);
/* The pin table in ascending vertex order: up to cap entries, *count = number of pins (each output
 * may be NULL). */
int tc_sim_get_pins(const tc_sim *sim, uint32_t *indices, float *targets, float *weights, uint32_t cap,
                    uint32_t *count);
/* set_vertex_inv_masses(indices, inv_masses) / get_vertex_inv_masses(), lib.rs:2660-2676: sets the
 * base inverse masses (the device value keeps a pin's scaling); get returns the base values
 * (cap >= num_vertices). */
int tc_sim_set_vertex_inv_masses(tc_sim *sim, const uint32_t *indices, const float *inv_masses, uint32_t n);
int tc_sim_get_vertex_inv_masses(const tc_sim *sim, float *out, size_t cap);

/* Colliders: add_sphere_collider / add_capsule_collider / add_plane_collider /
 * clear_colliders, lib.rs:2157-2172 -> simulation/mod.rs:1571-1629 */
int tc_sim_add_sphere(tc_sim *sim, const float center[3], float radius, float friction, float restitution);
int tc_sim_add_capsule(tc_sim *sim, const float a[3], const float b[3], float radius, float friction,
                       float restitution);
int tc_sim_add_plane(tc_sim *sim, const float point[3], const float normal[3], float friction, float restitution);
int tc_sim_clear_colliders(tc_sim *sim);
/* ClothSimulator.set_bone_sdf_colliders(width, height, depth, texture_bytes, bone_infos),
 * lib.rs:2245-2284 -> simulation/mod.rs:1630-1727: an Rg16Float volume (4 bytes per texel,
 * pack2x16float(distance, weight), Z/Y/X order) and nb rows of 20 floats
 * [aabb_min(3), centre.x, aabb_max(3), centre.y | width, uvw_scale(3), centre.z, uvw_offset(3),
 *  radius, friction, thickness, restitution, weight threshold].  Transforms start as identity.
 * Sampled trilinearly in the collider resolve after the primitives and the mesh triangles. */
int tc_sim_set_sdf_colliders(tc_sim *sim, uint32_t width, uint32_t height, uint32_t depth, const void *texels,
                             size_t nbytes, const float *infos, uint32_t nb);
/* ClothSimulator.update_bone_transforms(world, inv_world), lib.rs:2287-2330: n 4x4 matrices each,
 * row-major as given (the reference uploads them unchanged); n must equal the collider count. */
int tc_sim_update_bone_transforms(tc_sim *sim, const float *world, const float *inv_world, uint32_t n);
/* ClothSimulator.clear_bone_sdf_colliders, lib.rs:2342-2345 -> simulation/mod.rs:1941-1948: removes the
 * SDF colliders and the dynamic bone SDF. */
int tc_sim_clear_bone_sdf_colliders(tc_sim *sim);
/* ClothSimulator.setup_dynamic_bone_sdf, lib.rs:2362-2588 -> simulation/mod.rs:1998-2312.  A skinned
 * collider mesh whose bone-SDF atlas (width x height x depth texels, res^3 tiles, bone b's tile at
 * column b % (width/res), row (b / (width/res)) % (height/res), layer b / ((width/res)(height/res)))
 * is re-baked on the device at the start of every update_interval-th step / step_async /
 * step_buffered frame.  infos: num_bones x 20 floats (the collider infos; the bake box is the
 * local min/max of slots 0-2 and 4-6); rest_positions: num_verts x 3; bone_indices / bone_weights:
 * num_verts x 4 (the four skin influences); tri_sources: num_tris x 4 (i0, i1, i2, bone), each
 * bone's triangles contiguous; tri_weights: num_tris x 3 (the corners' weights); bind_inv: num_bind
 * row-major 4x4 bind matrices (world -> bone), num_bind >= num_bones.  TC_ERR_VALUE for a zero
 * size, num_bind < num_bones, a weighted bone index, triangle vertex or triangle bone out of range,
 * or a tile outside the atlas (the reference reads or writes out of range there).  An empty mesh,
 * triangle list or bone list changes nothing.  The atlas stays zero until the first update, which
 * bakes every bone. */
int tc_sim_setup_dynamic_bone_sdf(tc_sim *sim, uint32_t width, uint32_t height, uint32_t depth, uint32_t res,
                                  const float *infos, uint32_t num_bones, const float *rest_positions,
                                  uint32_t num_verts, const uint32_t *bone_indices, const float *bone_weights,
                                  const uint32_t *tri_sources, const float *tri_weights, uint32_t num_tris,
                                  const float *bind_inv, uint32_t num_bind, uint32_t update_interval);
/* set_dynamic_bone_sdf_dirty_bones(dirty_bone_indices=None), lib.rs:2332-2340 -> simulation/mod.rs:
 * 2314-2317: has_list 0 re-bakes every bone at each update; otherwise only the listed bones (an empty
 * list skips the bake).  update_bone_transforms in Python sets it as well. */
int tc_sim_set_dynamic_bone_sdf_dirty_bones(tc_sim *sim, const uint32_t *bones, uint32_t n, int has_list);
/* set_dynamic_bone_sdf_enabled, lib.rs:2591-2594 (on only after a setup) and
 * set_dynamic_bone_sdf_update_interval, lib.rs:2596-2599 (at least 1). */
int tc_sim_set_dynamic_bone_sdf_enabled(tc_sim *sim, int enabled);
int tc_sim_set_dynamic_bone_sdf_update_interval(tc_sim *sim, uint32_t interval);
/* Diagnostics (no reference counterpart): out[0] enabled, [1] baked at least once, [2] bone tiles
 * baked so far, [3] frame counter, [4] updates skipped by an empty dirty list. */
int tc_sim_dynamic_bone_sdf_stats(const tc_sim *sim, uint32_t *out);

/* set_mesh_collider_triangles(tris f32[T,3,3], friction, thickness, restitution, single_sided, attributes=None),
 * lib.rs:2175-2233 -> upload_colliders (simulation/mod.rs:2582-2765). attrs: NULL or T*4 floats
 * (friction, thickness, restitution, flags) overriding the scalars per triangle. */
int tc_sim_set_mesh_triangles(tc_sim *sim, const float *tris, uint32_t num_tris, const float *attrs,
                              float friction, float thickness, float restitution, int single_sided);

/* Parameter setters, lib.rs:1830-2000 -> simulation/mod.rs:1029-1447 */
int tc_sim_set_gravity(tc_sim *sim, float gx, float gy, float gz);
int tc_sim_set_damping(tc_sim *sim, float damping);
int tc_sim_set_damping_all(tc_sim *sim, float tension, float compression, float shear, float bending);
int tc_sim_set_solver_iterations(tc_sim *sim, uint32_t iterations);
int tc_sim_set_enable_self_collision(tc_sim *sim, int enable);
/* set_self_collision_options, simulation/mod.rs:920-953. Normal untangling acts only on multi-layer
 * meshes (several islands or layer ids; simulation/mod.rs:885-890): it adds the vertex-normal pass
 * and the untangle pass (compute_normals.wgsl, self_collision_untangle.wgsl) to each substep's
 * self-collision and exempts outer-behind-inner contacts from the V-V, V-T and E-E responses. */
int tc_sim_set_self_collision_options(tc_sim *sim, float relief_factor, float max_displacement_ratio,
                                      int exclude_neighbors, int enable_normal_untangling, uint32_t max_iterations);
/* set_stiffness_all, simulation/mod.rs:1106-1168 */
int tc_sim_set_stiffness_all(tc_sim *sim, float tension, float compression, float shear, float bending);

/* Extensions with no reference counterpart:
 * raw XPBD compliances per constraint kind (the reference maps k >= 5000 N/m to 0),
 * and an explicit self-collision hash cell size (the reference clamps cell >= 1 cm). */
int tc_sim_set_compliances_raw(tc_sim *sim, float stretch, float shear, float bend);
int tc_sim_set_hash_cell_size(tc_sim *sim, float cell_size);

/* Solver schedule (no reference counterpart; both give identical results):
 * 0 auto, 1 one launch per colour (the reference's dispatch pattern, dispatch.rs:1449-1455),
 * 2 one launch per solver iteration over dependency-cone tiles held in LDS. */
int tc_sim_set_solver_path(tc_sim *sim, int path);
/* set_solver_mode / get_solver_mode, lib.rs:1887-1895 -> simulation/mod.rs:2975-2984: 1 solves the
 * distance constraints of an iteration in one pass with fixed-point Jacobi sums, averaged and
 * over-relaxed by 1.25 (distance_atomic.wgsl:78-182; lambdas accumulate a quarter of each step);
 * any other value colours.  Mode 1 always takes the one-launch-per-colour path. */
int tc_sim_set_solver_mode(tc_sim *sim, uint32_t mode);
int tc_sim_get_solver_mode(const tc_sim *sim, uint32_t *mode);
/* Pair-cache self-collision (self_collision_algorithm 1), lib.rs:2029-2107 -> simulation/mod.rs:703-870,
 * dispatch.rs:1554-1645.  On the first substep of a frame a broadphase keeps, per vertex, the
 * quota = clamp(budget / num_vertices, 1, 8) closest V-T triples and E-E quadruples
 * (self_collision_collect_pairs.wgsl); every substep then resolves only those
 * (self_collision_solve_vt.wgsl, self_collision_solve_ee.wgsl) and applies the sums; with the final
 * fallback (default on) the last substep takes the direct path instead.  Algorithm 2 (virtual
 * vertices) returns TC_ERR_UNSUPPORTED.  set_pair_cache_options clamps as the reference: budgets to
 * [64, 262144], safety margin to [0, 0.05], horizon scale to [0, 4], max horizon to [0, 0.1];
 * margin_mode 0 drops the velocity horizon.  pair_cache_state out[9] = [max V-T, max E-E, quota V-T,
 * quota E-E, margin mode, safety margin, horizon scale, max horizon, final fallback]. */
int tc_sim_set_enable_pair_cache(tc_sim *sim, int enable);
int tc_sim_get_enable_pair_cache(const tc_sim *sim, int *enable);
int tc_sim_set_self_collision_algorithm(tc_sim *sim, uint32_t algo);
int tc_sim_get_self_collision_algorithm(const tc_sim *sim, uint32_t *algo);
int tc_sim_set_pair_cache_options(tc_sim *sim, uint32_t max_vt_pairs, uint32_t max_ee_pairs, uint32_t margin_mode,
                                  float safety_margin, float horizon_scale, float max_horizon);
int tc_sim_set_enable_pair_cache_final_fallback(tc_sim *sim, int enable);
int tc_sim_pair_cache_state(const tc_sim *sim, double *out);
/* get_active_pairs, dispatch.rs:1255-1291: counts[4] = [V-T evaluated, E-E evaluated, V-T dropped,
 * E-E dropped] of the last collect; the held pairs in slot order (vertex-major), 4 indices each,
 * up to *_cap pairs written, *n_* = pairs held (outputs may be NULL to count). */
int tc_sim_get_active_pairs(tc_sim *sim, uint32_t *counts, uint32_t *vt, size_t vt_cap, size_t *n_vt, uint32_t *ee,
                            size_t ee_cap, size_t *n_ee);
/* Collider-edge collision, lib.rs:2735-2749 -> simulation/mod.rs:1007-1025: cloth edges against the
 * unique edges of the mesh collider (edge_collision.wgsl), after the collider pass and the final pin,
 * pushing into the self-collision accumulators' hard channel and applied at once.  set_edge_margin:
 * target distance = (cloth + collider thickness) * scale + offset, scale >= 1, offset >= 0
 * (set_edge_margin_scale / set_edge_margin_offset).  collider_edge_count: unique collider edges. */
int tc_sim_set_enable_edge_collision(tc_sim *sim, int enable);
int tc_sim_set_edge_margin(tc_sim *sim, float scale, float offset);
int tc_sim_collider_edge_count(const tc_sim *sim, uint32_t *n);
/* The collider edges, 12 words each: p0 xyz, thickness, p1 xyz, friction, normal xyz, flags (u32 bits);
 * cap = edges the buffer holds (no reference counterpart: exposed for parity checks). */
int tc_sim_get_collider_edges(tc_sim *sim, float *out, uint32_t cap);
/* Wrinkle field (drape guide), lib.rs:3049-3137 -> simulation/mod.rs:2997-3130, wrinkle_field.wgsl:
 * after predict, a radial acceleration around a bone read from an RGBA8 texture over the bone's
 * cylinder (R valley strength, G crest strength, B / A valley / crest radius in [r_min, r_max]).
 * set_wrinkle_field_params: frame[16] = origin xyz + influence radius, axis xyz + bone radius,
 * normal xyz + stiffness, binormal xyz + blend weight; the z and r ranges and the two windows.
 * get_wrinkle_field_params: out[22] in that order, then z_min, z_max, r_min, r_max, windows.
 * set_wrinkle_field_texture_2d: width * height * 4 bytes (rows of RGBA, v = row); a zero size drops
 * the texture (the pass then does nothing); also sets the z and r ranges. */
int tc_sim_set_enable_wrinkle_field(tc_sim *sim, int enable);
int tc_sim_wrinkle_field_enabled(const tc_sim *sim, int *enabled);
int tc_sim_set_wrinkle_field_params(tc_sim *sim, const float *frame, float z_min, float z_max, float r_min,
                                    float r_max, float valley_window, float crest_window);
int tc_sim_get_wrinkle_field_params(const tc_sim *sim, float *out);
int tc_sim_set_wrinkle_field_texture_2d(tc_sim *sim, uint32_t width, uint32_t height, const uint8_t *rgba,
                                        size_t nbytes, float z_min, float z_max, float r_min, float r_max);
/* Fast projection (no reference counterpart; opt-in, off by default): the per-iteration tile path
 * computes each projection's length, direction and lambda step with the hardware reciprocal square
 * root and reciprocal, each refined by one Newton-Raphson step, in place of the IEEE square root and
 * divisions of distance.wgsl:60-124 / bending.wgsl:59-126.  Results are no longer bit-identical to
 * the exact path (within 1e-5 m of it over 100 substeps on the tested drapes); the per-colour and
 * Jacobi paths are unaffected. */
int tc_sim_set_fast_projection(tc_sim *sim, int enable);
int tc_sim_get_fast_projection(const tc_sim *sim, int *enabled);
/* out[10]: [active path (1|2), tiles, largest cone (vertices), cone vertices total,
 *           items per iteration, constraints per iteration, plan available (0|1), colour phases,
 *           tile size (vertices), block size of the per-iteration kernel] */
int tc_sim_solver_info(tc_sim *sim, double *out);

/* Autotuner (no reference counterpart; the reference exposes a fixed workgroup_size knob,
 * pipeline_cache.rs:25-48): times 16 solver iterations of the per-colour schedule and of the
 * per-iteration schedule for tile sizes {256, 512, 1024, 2048} vertices x block sizes {256, 512, 1024}
 * with HIP events on the simulator's stream, keeps the fastest, and writes
 * out[4] = [tile vertices (0 = per-colour kept), block size, ms per iteration, per-colour ms per iteration].
 * Leaves the cloth state unchanged (the trial iterations start from a fresh prediction). */
int tc_sim_autotune(tc_sim *sim, float dt_sub, double *out);

/* num_vertices / num_distance_constraints / num_bending_constraints getters, lib.rs:2928-2950 */
uint32_t tc_sim_num_vertices(const tc_sim *sim);
uint32_t tc_sim_num_distance_constraints(const tc_sim *sim);
uint32_t tc_sim_num_bending_constraints(const tc_sim *sim);

/* Introspection for parity checks (no reference counterpart; the reference
 * exposes the same data only to its own Rust tests). `which` selects:
 *  0 dist v0 (u32, colour order)   1 dist v1          2 dist colour offsets   3 dist colour counts
 *  4 bend v2                        5 bend v3          6 bend colour offsets   7 bend colour counts
 *  8 dist rest (f32)                9 bend rest (f32)  10 vertex hash keys of the last build (u32)
 * 11 vertex hash cell starts (T+1)  12 vertex hash sorted indices            13 predicted positions p (f32 xyz)
 * 14 velocities (f32 xyz)          15 distance lambdas (f32)                16 bending lambdas (f32)
 * 17 scalars [cell_size(f32 bits), table_size, n_virtual, num_edges]       18 Morton-sorted edges (u32 pairs)
 * 19 edge hash keys                 20 edge hash cell starts                 21 edge hash sorted indices
 * 22 inverse masses the solver reads (f32; the base values scaled by the pins)
 * 23 SDF texels (u32, width x height x depth)  24 SDF collider infos (f32, 20 per collider)
 * 25 dynamic bone SDF skinned positions (f32 xyz)  26 dynamic bone SDF bake triangles (16 f32 each:
 *    p0, w0, p1, w1, p2, w2, normal, open-edge mask)
 * *count receives the element count; out may be NULL to query it. */
int tc_sim_get_array(tc_sim *sim, int which, void *out, size_t capacity, size_t *count);
/* Rebuild the spatial hashes from the current predicted positions (parity checks of the hash). */
int tc_sim_build_hashes(tc_sim *sim);
/* Relative XPBD residual ||C + alpha~ lambda|| / ||L0|| after the last projection sweep of the
 * last substep (project definition, no reference counterpart), and max |h_j| / L0_j. */
int tc_sim_residual(tc_sim *sim, float dt_sub, double *rel, double *max_rel);
/* Run the parts of one substep up to the velocity update (residual / warm-start studies). */
int tc_sim_substep_parts(tc_sim *sim, float dt_sub, int collide, int self_collide, int update_velocity);
/* Times `reps` substeps launched kernel by kernel (no graph) with HIP events recorded on the
 * simulator's stream, using the active solver schedule.  Launches of one kind that follow each other
 * (the iterations of the per-iteration kernel; one iteration's distance colours; its bending colours)
 * sit under one event pair, so no event packet separates them.  Writes per-substep averages (milliseconds)
 * to out[12]: [0] predict  [1] all distance launches  [2] distance launches per substep
 * [3] all bending launches  [4] bending launches per substep  [5] collider resolve
 * [6] self-collision (hash + solve + apply)  [7] update_vel  [8] wall time of the substep between its
 * first and last event  [9] all one-launch-per-iteration launches  [10] those launches per substep
 * [11] reserved (0).  No reference counterpart (the reference's wgpu-profiler scopes, simulation/profile.rs). */
int tc_sim_time_kernels(tc_sim *sim, float dt_sub, uint32_t reps, double *out);
/* ClothSimulator.set_profiling_enabled(enable) -> bool, lib.rs:1975-1978 -> GpuProfiler::set_enabled
 * (simulation/profile.rs:76-83).  While on, frames launch kernel by kernel and HIP events on the
 * simulator's stream bracket each pass under the reference's pass names: predict, solver_iter (one per
 * iteration), collider_decoupled, self_collision_Outer (a scope holding hash_build, sc_solve,
 * sc_solve_ee, sc_apply), update_vel, plus warm_start, residual_gate and warm_start_update for the
 * learned warm start.  *enabled (may be NULL) receives the resulting state; HIP events are always
 * available, so it equals the request.  Turning it on or off clears what was recorded. */
int tc_sim_set_profiling_enabled(tc_sim *sim, int enable, int *enabled);
/* ClothSimulator.take_profile() -> [(pass, ms)], lib.rs:1980-1983 -> take_ms (profile.rs:206-230):
 * waits for the simulator's stream and keeps the passes of the last frame stepped; *count receives
 * their number (0 when profiling is off or nothing ran).  A later call without a new frame returns the
 * same passes, as the reference returns its previous values.  Entries, in recording order (a scope
 * before its passes), come from tc_sim_profile_entry: name (owned by the simulator, valid until the
 * next take), milliseconds, start relative to the frame's first pass, nesting depth. */
int tc_sim_take_profile(tc_sim *sim, size_t *count);
int tc_sim_profile_entry(const tc_sim *sim, size_t index, const char **name, float *ms, float *start_ms,
                         uint32_t *depth);
/* ClothSimulator.save_profile_trace(path), lib.rs:1985-1990 -> save_chrometrace (profile.rs:232-237):
 * the passes of the last take as Chrome trace JSON (complete events in microseconds).  TC_ERR_IO when
 * the file cannot be written (the reference raises OSError). */
int tc_sim_save_profile_trace(tc_sim *sim, const char *path);

/* Block until all work on the simulator's stream has finished. */
int tc_sim_synchronize(tc_sim *sim);

/* ---- Learned warm start (no reference counterpart; it runs between predict and the
 * projection loop of the substep, crates/cloth_core/src/simulation/dispatch.rs:137 -> 169).
 * A message-passing network over the cloth's stretch-edge graph (hidden width 128, `rounds`
 * rounds, two-layer MLPs, BF16 MFMA with FP32 accumulation) predicts each free vertex's position
 * correction for the coming substep from local features of the predicted state and the previous
 * substep's correction; plain XPBD iterations always follow, and a residual gate runs further
 * plain iterations when the relative residual after them exceeds a threshold. */
typedef struct tc_ws_net tc_ws_net;

/* New network with zero parameters; rounds 1..4.  The network must outlive every simulator
 * it is attached to. */
int tc_ws_net_create(uint32_t rounds, tc_ws_net **out);
void tc_ws_net_destroy(tc_ws_net *net);
uint32_t tc_ws_net_num_params(const tc_ws_net *net);
uint32_t tc_ws_net_rounds(const tc_ws_net *net);
/* Flat fp32 parameter vector: per layer W (K x N, row-major) then b (N), layers in the order
 * encoder 1 (32 x 128), encoder 2 (128 x 128), per round (256 x 128, 128 x 128),
 * decoder 1 (128 x 128), decoder 2 (128 x 16).  Setting parameters resets the Adam state. */
int tc_ws_net_set_params(tc_ws_net *net, const float *params, size_t count);
int tc_ws_net_get_params(tc_ws_net *net, float *params, size_t count);
/* Seeded He-uniform initialisation; the last layer starts at zero. */
int tc_ws_net_init(tc_ws_net *net, uint64_t seed);
/* One forward pass (and, with targets, the loss and its gradient) on host arrays:
 * features rows x 32 (rounded to BF16), a symmetric neighbour graph in CSR form,
 * targets rows x 4 (target correction xyz in units of |g| dt^2, weight 0|1) or NULL.
 * Outputs: y rows x 16, grad (num_params), loss; each may be NULL.  The gradient stays on the
 * device for tc_ws_net_adam_step.  Without targets only the forward pass runs (no activations are
 * kept, no gradient is formed). */
int tc_ws_net_forward_backward(tc_ws_net *net, const float *features, uint32_t rows, const uint32_t *graph_off,
                               const uint32_t *graph_nbr, const float *targets, float *y, float *grad,
                               double *loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with the last computed gradient. */
int tc_ws_net_adam_step(tc_ws_net *net, float lr);
/* Allocation-free forward pass.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same pass as tc_ws_net_forward_backward with targets NULL (features rounded to
 * BF16 the same way, the same kernels, bit-identical y), but every device buffer it needs -- the
 * features, the two hidden-state buffers, the outputs and the graph -- lives in one static device
 * array of TC_WS_NOALLOC_SCRATCH_BYTES held by each GPU until the process ends; it calls no
 * hipMalloc or hipFree.  Limits: rows <= TC_WS_NOALLOC_MAX_ROWS (a 128 x 128 cloth) and
 * graph_off[rows] <= TC_WS_NOALLOC_MAX_NBR; larger inputs return TC_ERR_VALUE.  The graph need not
 * be symmetric (there is no backward pass).  Inputs and y are host arrays; the call blocks until y
 * is written.  Calls on one GPU share the array (it keeps the last call's contents, none of which a
 * call reads before writing), so calls that overlap in time on one GPU race. */
#define TC_WS_NOALLOC_MAX_ROWS 16384u
#define TC_WS_NOALLOC_MAX_NBR 131072u
#define TC_WS_NOALLOC_SCRATCH_BYTES 13172992u
int tc_ws_net_forward_noalloc(tc_ws_net *net, const float *features, uint32_t rows, const uint32_t *graph_off,
                              const uint32_t *graph_nbr, float *y);
/* Trains on the samples the simulators recorded (tc_sim_ws_record): `steps` Adam steps, each on
 * one recorded substep drawn with a seeded generator; losses[steps] may be NULL. */
int tc_ws_train(tc_ws_net *net, tc_sim *const *sims, uint32_t nsims, uint32_t steps, float lr, uint64_t seed,
                float *losses);

/* Mean loss of the network over every sample the simulators recorded (no parameter update). */
int tc_ws_eval(tc_ws_net *net, tc_sim *const *sims, uint32_t nsims, double *mean_loss);

/* mode 0: off (plain XPBD with the simulator's iteration count).
 * mode 1: start the projection from prediction + w * previous substep's correction.
 * mode 2: start it from prediction + w * previous correction + network output.
 * w = history_weight in [0, 1]; the previous correction also reaches the substep through the
 * velocity, so w near 1 amplifies error modes the iterations correct slowly (0.5 or less keeps
 * them decaying).  A network must be used with the weight it was trained for.
 * Modes 1 and 2 run `iterations` plain iterations, then a residual check; while the relative
 * residual exceeds `residual_threshold`, fallback iterations run in chunks of `iterations`
 * (rounded up to even), each followed by another check, up to `fallback_iterations` (rounded up
 * to even) in all. */
int tc_sim_set_warm_start(tc_sim *sim, tc_ws_net *net, int mode, uint32_t iterations,
                          uint32_t fallback_iterations, float residual_threshold, float history_weight);
/* out[8]: [substeps checked, substeps that fell back, relative residual at the last check,
 *          max |h|/L0 at the last check, mode, warm-start iterations, fallback chunks run,
 *          substeps that spent the whole fallback budget and still ended above the threshold].
 * A substep whose last fallback chunk runs gets one more check after it, so "the last check" is the
 * residual the substep ended with. */
int tc_sim_warm_start_stats(tc_sim *sim, double *out);
int tc_sim_ws_reset_stats(tc_sim *sim);
/* Records training samples from the coming substeps: every `stride`-th substep stores its
 * features, target correction, prediction and previous correction, up to `capacity` samples.
 * The target is (correction - history_weight * previous correction) / s.
 * capacity 0 stops recording and frees the samples.  Recording launches kernel by kernel. */
int tc_sim_ws_record(tc_sim *sim, uint32_t capacity, uint32_t stride, float history_weight);
uint32_t tc_sim_ws_num_records(const tc_sim *sim);
/* Copies sample `idx`: features N x 32, target N x 4, prediction N x 4 (xyz, inverse mass),
 * previous correction N x 4; each may be NULL. */
int tc_sim_ws_get_record(tc_sim *sim, uint32_t idx, float *features, float *target, float *prediction,
                         float *previous);
/* Network output of the last warm-started substep (N x 16). */
int tc_sim_ws_last_output(tc_sim *sim, float *y, size_t count);
/* Reads and/or replaces the previous-substep correction (N x 3 each; either may be NULL). */
int tc_sim_ws_history(tc_sim *sim, float *get_xyz, const float *set_xyz, uint32_t n);
/* Runs `reps` warm-started substeps kernel by kernel with HIP events on the simulator's stream
 * and writes per-substep averages (ms) to out[16]: [0] predict [1] features [2] encoder
 * [3] message-passing rounds (all rounds) [4] decoder [5] apply [6] warm-start iterations
 * [7] residual gate [8] fallback launches [9] correction update [10] collider resolve +
 * self-collision [11] update_vel [12] whole substep [13] rounds [14] warm-start iterations
 * ([15] unused).  The simulation advances by `reps` substeps. */
int tc_sim_ws_time(tc_sim *sim, float dt_sub, uint32_t reps, double *out);

/* ---- Mesh SDF bake (module functions bake_mesh_sdf_gpu / bake_mesh_sdf_hierarchical_gpu,
 * lib.rs:1350-1485 -> sdf_baker.rs:1114-1323, 1625-1882).  method 0: dense bake, 1: hierarchical
 * (coarse top-8 pass + fine pass), 2: the coarse pass alone (bake_mesh_sdf_coarse_ext,
 * sdf_baker.rs:1350-1600: cell width 4 x voxel_size, 8 nearest triangle indices per cell).
 * The result holds width x height x depth texels (pack2x16float(signed distance, 1.0), Z/Y/X
 * order), the 20-float collider info and, for method 2, the nearest-triangle indices; method 3 is the
 * dense bake with an fp32 brick atlas as well (tc_sdf_result_atlas). */
typedef struct tc_sdf_result tc_sdf_result;
int tc_bake_mesh_sdf(const float *verts, uint32_t num_verts, const int32_t *tris, uint32_t num_tris, float voxel_size,
                     float margin, float friction, float thickness, float restitution, int close_open_boundaries,
                     int method, tc_sdf_result **out);
/* dims[3], bone_info[20], coarse cell width (method 2); each may be NULL. */
int tc_sdf_result_info(const tc_sdf_result *res, uint32_t *dims, float *bone_info, float *coarse_voxel);
/* Device time (ms, HIP events on the launching stream) of the bake's kernels: kernel_ms[0] the dense
 * traversal or the coarse pass, kernel_ms[1] the fine pass with its fallback (0 unless hierarchical);
 * fallback_voxels: voxels the fine pass left to its exhaustive fallback.  Either may be NULL.
 * Diagnostics only; no reference counterpart (bake_mesh_sdf_gpu_ext, sdf_baker.rs:1114-1323, reports
 * neither). */
int tc_sdf_result_timing(const tc_sdf_result *res, float *kernel_ms, uint32_t *fallback_voxels);
const uint32_t *tc_sdf_result_texels(const tc_sdf_result *res);
/* method 3 (no reference counterpart; the reference stores only pack2x16float texels,
 * bake_mesh_sdf.wgsl:120): the dense bake's texels plus an fp32 brick atlas of the same voxels,
 * ceil(w/8) x ceil(h/8) x ceil(d/8) bricks of 8^3 floats (signed distance to the nearest triangle,
 * sign from its face normal), bricks and voxels within a brick in Z/Y/X order, voxels past the
 * domain's edge holding the nearest voxel's value.  Returns NULL for other methods; brick_dims
 * (may be NULL) receives the brick counts. */
const float *tc_sdf_result_atlas(const tc_sdf_result *res, uint32_t *brick_dims);
const uint32_t *tc_sdf_result_nearest(const tc_sdf_result *res);
void tc_sdf_result_free(tc_sdf_result *res);
/* ---- Bone SDF atlas bake (module function bake_bone_sdf_gpu, lib.rs:1190-1348 ->
 * bake_bone_sdf_gpu_ext, sdf_baker.rs:494-930, kernel bake_sdf.wgsl).  weights: num_bones x
 * num_verts floats, bone-major; binds: num_bones row-major 4x4 bind (world -> bone) matrices.
 * Bones whose largest weight exceeds weight_threshold each get a res^3 tile of one 3D atlas
 * (compute_3d_atlas_layout, sdf_baker.rs:137-163, at most 2048 texels per axis); each texel holds
 * pack2x16float(signed distance to the bone's skinned triangles in bone space, bone weight at the
 * nearest point) in Z/Y/X order.  TC_ERR_VALUE for resolution 0 or null arrays; TC_ERR_RUNTIME when
 * no bone passes the threshold or the atlas does not fit. */
typedef struct tc_bone_sdf_result tc_bone_sdf_result;
int tc_bake_bone_sdf(const float *verts, uint32_t num_verts, const int32_t *tris, uint32_t num_tris,
                     const float *weights, const float *binds, uint32_t num_bones, uint32_t resolution, float margin,
                     float weight_threshold, float blend_k, float friction, float thickness, float restitution,
                     int close_open_boundaries, tc_bone_sdf_result **out);
/* dims[3] = width, height, depth; n_active = active bones; kernel_ms = device time of the bake
 * kernel (HIP events; no reference counterpart).  Each may be NULL. */
int tc_bone_sdf_result_info(const tc_bone_sdf_result *res, uint32_t *dims, uint32_t *n_active, float *kernel_ms);
/* texels (w*h*d u32), infos (20 floats per active bone: local min, tile index, local max, res,
 * atlas uvw scale + blend_k, atlas uvw offset + 0, friction, thickness, restitution, weight
 * threshold; sdf_baker.rs:603-615), active (input index of each active bone), binds (16 floats per
 * active bone).  Each may be NULL. */
int tc_bone_sdf_result_arrays(const tc_bone_sdf_result *res, uint32_t *texels, float *infos, uint32_t *active,
                              float *binds);
void tc_bone_sdf_result_free(tc_bone_sdf_result *res);
/* close_open_boundary_loops_mesh (sdf_baker.rs:455-462): caps the closed boundary loops of a solid
 * mesh.  With NULL outputs only the counts are written. */
int tc_close_open_boundary_loops(const float *verts, uint32_t num_verts, const int32_t *tris, uint32_t num_tris,
                                 float *out_verts, size_t cap_verts, int32_t *out_tris, size_t cap_tris,
                                 uint32_t *num_verts_out, uint32_t *num_tris_out, uint32_t *loops_closed);

/* Host-only mesh build (no device work; usable without a GPU): builds the constraints, colouring
 * and self-collision topology exactly as tc_sim_create does and returns array `which`:
 * 0-9, 17, 18 as in tc_sim_get_array, and
 * 22 two-hop offsets   23 two-hop indices   24 two-hop rest lengths (f32)  25 local edge lengths (f32)
 * 26 star offsets      27 star pairs (u32 pairs)  28 open-boundary edge bits  29 island ids
 * 30 input edge -> distance constraint (colour order, UINT32_MAX if invalid)   31 distance type (0 stretch, 1 shear)
 * 32 tension compliance (f32)   33 compression compliance (f32)   34 bending compliance (f32)
 * 35 adjacency offsets          36 adjacency indices               37 layer ids */
int tc_mesh_build_arrays(const float *positions, uint32_t num_vertices, const uint32_t *edges, uint32_t num_edges,
                         const uint32_t *faces, uint32_t num_faces, const float *inv_masses,
                         const float *thicknesses, const tc_sim_params *params, int which, void *out,
                         size_t capacity, size_t *count);
/* The same with sewing springs; further array ids:
 * 38 bending v0   39 bending v1   40 bending _pad0 (f32)   41 bending _pad1 (f32)
 * 42 sewing v0    43 sewing v1    44 sewing rest lengths (f32)   45 sewing colour offsets   46 sewing colour counts */
int tc_mesh_build_arrays_sewn(const float *positions, uint32_t num_vertices, const uint32_t *edges,
                              uint32_t num_edges, const uint32_t *faces, uint32_t num_faces,
                              const float *inv_masses, const float *thicknesses, const uint32_t *sewing,
                              uint32_t num_sewing, const tc_sewing_params *sewing_params,
                              const tc_sim_params *params, int which, void *out, size_t capacity, size_t *count);

/* Module functions */
/* color_edge_pairs(n, edges) -> colours, lib.rs:636-648 -> coloring.rs:180 */
int tc_color_edge_pairs(uint32_t num_vertices, const uint32_t *pairs, uint32_t num_pairs, uint32_t *colors_out);
/* compute_combined_inv_masses (faces, pins; no strands), lib.rs:552 -> mesh.rs:237-312 */
int tc_compute_inv_masses(const float *positions, uint32_t num_vertices, const uint32_t *faces, uint32_t num_faces,
                          const float *pin_weights, float areal_density, float *out);

/* Last error message of the calling thread ("" if none). */
const char *tc_last_error(void);
/* Library version string. */
const char *tc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* TC_SIM_H */
