// Rust declarations of the C ABI in include/tc_sim.h, for the reference's PyO3 crate
// (crates/cloth_py): a maintainer would add this file as crates/cloth_py/src/tc_ffi.rs, link
// libtcloth_hip.so from build.rs (`println!("cargo:rustc-link-lib=dylib=tcloth_hip")`) and route
// ClothSimulator's methods through these functions instead of GpuClothSimulator.
//
// Unverified by a Rust compiler (cargo is not available where this library is built).  The test
// suite checks that every function of tc_sim.h appears here once, in the same order, with the same
// number of parameters and the same pointer / scalar shape for each, and that the three #[repr(C)]
// structures list the fields of tc_sim_params, tc_instance and tc_sewing_params in order.
#![allow(non_camel_case_types, dead_code)]

use std::os::raw::{c_char, c_void};

pub const TC_OK: i32 = 0;
pub const TC_ERR_VALUE: i32 = 1;       // PyValueError
pub const TC_ERR_RUNTIME: i32 = 2;     // PyRuntimeError
pub const TC_ERR_UNSUPPORTED: i32 = 3; // PyNotImplementedError
pub const TC_ERR_IO: i32 = 4;          // PyOSError
pub const TC_WS_NOALLOC_MAX_ROWS: u32 = 16384;
pub const TC_WS_NOALLOC_MAX_NBR: u32 = 131072;
pub const TC_WS_NOALLOC_SCRATCH_BYTES: u32 = 13172992;

#[repr(C)] pub struct TcSim { _private: [u8; 0] }
#[repr(C)] pub struct TcWsNet { _private: [u8; 0] }
#[repr(C)] pub struct TcSdfResult { _private: [u8; 0] }
#[repr(C)] pub struct TcBoneSdfResult { _private: [u8; 0] }

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct TcParams {
    pub thickness: f32,
    pub layer_id: u32,
    pub stiffness: f32,
    pub bending_stiffness: f32,
    pub compression_stiffness: f32,
    pub shear_stiffness: f32,
    pub workgroup_size: u32,
    pub solver_mode: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct TcInstance {
    pub positions: *const f32,
    pub num_vertices: u32,
    pub edges: *const u32,
    pub num_edges: u32,
    pub faces: *const u32,
    pub num_faces: u32,
    pub inv_masses: *const f32,
    pub thicknesses: *const f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct TcSewingParams {
    pub shrink_speed: f32,
    pub stiffness: f32,
    pub lock: i32,
    pub lock_distance: f32,
}

#[link(name = "tcloth_hip")]
extern "C" {
    pub fn tc_sim_params_default(p: *mut TcParams);
    pub fn tc_sewing_params_default(p: *mut TcSewingParams);
    pub fn tc_sim_create(positions: *const f32, num_vertices: u32, edges: *const u32, num_edges: u32, faces: *const u32, num_faces: u32, inv_masses: *const f32, thicknesses: *const f32, layer_ids: *const u32, params: *const TcParams, out: *mut *mut TcSim) -> i32;
    pub fn tc_sim_destroy(sim: *mut TcSim);
    pub fn tc_sim_create_sewn(positions: *const f32, num_vertices: u32, edges: *const u32, num_edges: u32, faces: *const u32, num_faces: u32, inv_masses: *const f32, thicknesses: *const f32, layer_ids: *const u32, sewing: *const u32, num_sewing: u32, sewing_params: *const TcSewingParams, params: *const TcParams, out: *mut *mut TcSim) -> i32;
    pub fn tc_sim_set_sewing_stiffness(sim: *mut TcSim, stiffness: f32) -> i32;
    pub fn tc_sim_set_sewing_shrink_speed(sim: *mut TcSim, speed: f32) -> i32;
    pub fn tc_sim_set_enable_sewing_lock(sim: *mut TcSim, enable: i32) -> i32;
    pub fn tc_sim_set_sewing_lock_distance(sim: *mut TcSim, distance: f32) -> i32;
    pub fn tc_sim_get_sewing_state(sim: *const TcSim, out: *mut f32) -> i32;
    pub fn tc_sim_get_sewing_rest_lengths(sim: *mut TcSim, out: *mut f32, n: u32) -> i32;
    pub fn tc_sim_set_sewing_rest_lengths(sim: *mut TcSim, values: *const f32, n: u32) -> i32;
    pub fn tc_sewing_closure_ratio(positions: *const f32, num_vertices: u32, pairs: *const u32, num_pairs: u32, merge_dist: f32, ratio: *mut f32) -> i32;
    pub fn tc_sim_set_sewing_priority_options(sim: *mut TcSim, enabled: i32, threshold: f32, merge_dist: f32, ramp_frames: u32, max_frames: u32) -> i32;
    pub fn tc_sim_set_sewing_priority_state(sim: *mut TcSim, latched: i32, frame: u32, ramp_t: f32, ratio: f32) -> i32;
    pub fn tc_sim_get_sewing_priority_state(sim: *const TcSim, out: *mut f64) -> i32;
    pub fn tc_sim_update_sewing_priority(sim: *mut TcSim, positions: *const f32, num_vertices: u32, out: *mut f32) -> i32;
    pub fn tc_sim_create_batch(instances: *const TcInstance, num_instances: u32, params: *const TcParams, out: *mut *mut TcSim) -> i32;
    pub fn tc_sim_instance_count(sim: *const TcSim, num_instances: *mut u32) -> i32;
    pub fn tc_sim_instance_range(sim: *const TcSim, instance: u32, first_vertex: *mut u32, num_vertices: *mut u32) -> i32;
    pub fn tc_sim_step(sim: *mut TcSim, dt: f32, substeps: u32, iterations: u32, effective_substeps: *mut u32) -> i32;
    pub fn tc_sim_set_adaptive_substep_options(sim: *mut TcSim, enabled: i32, min_substeps: u32, max_substeps: u32) -> i32;
    pub fn tc_sim_set_auto_compensate_iterations(sim: *mut TcSim, enabled: i32) -> i32;
    pub fn tc_sim_set_strain_adaptive_options(sim: *mut TcSim, enabled: i32, tolerance: f32) -> i32;
    pub fn tc_sim_set_enable_compact_readback(sim: *mut TcSim, enabled: i32) -> i32;
    pub fn tc_sim_note_external_displacement(sim: *mut TcSim, displacement: f32) -> i32;
    pub fn tc_sim_frame_policy_state(sim: *const TcSim, out: *mut f64) -> i32;
    pub fn tc_sim_step_async(sim: *mut TcSim, dt: f32, substeps: u32, iterations: u32) -> i32;
    pub fn tc_sim_fetch_positions(sim: *mut TcSim, out: *mut f32, n_floats: usize, ok: *mut i32) -> i32;
    pub fn tc_sim_step_buffered(sim: *mut TcSim, dt: f32, substeps: u32, iterations: u32, count: *mut u32) -> i32;
    pub fn tc_sim_fetch_buffered_positions(sim: *mut TcSim, out: *mut f32, n_floats: usize, count: *mut u32) -> i32;
    pub fn tc_sim_get_buffered_frame_count(sim: *const TcSim, count: *mut u32) -> i32;
    pub fn tc_sim_clear_frame_buffer(sim: *mut TcSim) -> i32;
    pub fn tc_sim_step_single_substep(sim: *mut TcSim, dt_sub: f32) -> i32;
    pub fn tc_sim_get_positions(sim: *mut TcSim, out: *mut f32, n_floats: usize) -> i32;
    pub fn tc_sim_set_positions_velocities(sim: *mut TcSim, positions: *const f32, velocities: *const f32, n: u32) -> i32;
    pub fn tc_sim_reset(sim: *mut TcSim) -> i32;
    pub fn tc_sim_set_pins(sim: *mut TcSim, indices: *const u32, targets: *const f32, weights: *const f32, n: u32) -> i32;
    pub fn tc_sim_release_pins(sim: *mut TcSim, indices: *const u32, n: u32) -> i32;
    pub fn tc_sim_clear_pins(sim: *mut TcSim) -> i32;
    pub fn tc_sim_pin_count(sim: *const TcSim, n: *mut u32) -> i32;
    pub fn tc_sim_get_pins(sim: *const TcSim, indices: *mut u32, targets: *mut f32, weights: *mut f32, cap: u32, count: *mut u32) -> i32;
    pub fn tc_sim_set_vertex_inv_masses(sim: *mut TcSim, indices: *const u32, inv_masses: *const f32, n: u32) -> i32;
    pub fn tc_sim_get_vertex_inv_masses(sim: *const TcSim, out: *mut f32, cap: usize) -> i32;
    pub fn tc_sim_add_sphere(sim: *mut TcSim, center: *const f32, radius: f32, friction: f32, restitution: f32) -> i32;
    pub fn tc_sim_add_capsule(sim: *mut TcSim, a: *const f32, b: *const f32, radius: f32, friction: f32, restitution: f32) -> i32;
    pub fn tc_sim_add_plane(sim: *mut TcSim, point: *const f32, normal: *const f32, friction: f32, restitution: f32) -> i32;
    pub fn tc_sim_clear_colliders(sim: *mut TcSim) -> i32;
    pub fn tc_sim_set_sdf_colliders(sim: *mut TcSim, width: u32, height: u32, depth: u32, texels: *const c_void, nbytes: usize, infos: *const f32, nb: u32) -> i32;
    pub fn tc_sim_update_bone_transforms(sim: *mut TcSim, world: *const f32, inv_world: *const f32, n: u32) -> i32;
    pub fn tc_sim_clear_bone_sdf_colliders(sim: *mut TcSim) -> i32;
    pub fn tc_sim_setup_dynamic_bone_sdf(sim: *mut TcSim, width: u32, height: u32, depth: u32, res: u32, infos: *const f32, num_bones: u32, rest_positions: *const f32, num_verts: u32, bone_indices: *const u32, bone_weights: *const f32, tri_sources: *const u32, tri_weights: *const f32, num_tris: u32, bind_inv: *const f32, num_bind: u32, update_interval: u32) -> i32;
    pub fn tc_sim_set_dynamic_bone_sdf_dirty_bones(sim: *mut TcSim, bones: *const u32, n: u32, has_list: i32) -> i32;
    pub fn tc_sim_set_dynamic_bone_sdf_enabled(sim: *mut TcSim, enabled: i32) -> i32;
    pub fn tc_sim_set_dynamic_bone_sdf_update_interval(sim: *mut TcSim, interval: u32) -> i32;
    pub fn tc_sim_dynamic_bone_sdf_stats(sim: *const TcSim, out: *mut u32) -> i32;
    pub fn tc_sim_set_mesh_triangles(sim: *mut TcSim, tris: *const f32, num_tris: u32, attrs: *const f32, friction: f32, thickness: f32, restitution: f32, single_sided: i32) -> i32;
    pub fn tc_sim_set_gravity(sim: *mut TcSim, gx: f32, gy: f32, gz: f32) -> i32;
    pub fn tc_sim_set_damping(sim: *mut TcSim, damping: f32) -> i32;
    pub fn tc_sim_set_damping_all(sim: *mut TcSim, tension: f32, compression: f32, shear: f32, bending: f32) -> i32;
    pub fn tc_sim_set_solver_iterations(sim: *mut TcSim, iterations: u32) -> i32;
    pub fn tc_sim_set_enable_self_collision(sim: *mut TcSim, enable: i32) -> i32;
    pub fn tc_sim_set_self_collision_options(sim: *mut TcSim, relief_factor: f32, max_displacement_ratio: f32, exclude_neighbors: i32, enable_normal_untangling: i32, max_iterations: u32) -> i32;
    pub fn tc_sim_set_stiffness_all(sim: *mut TcSim, tension: f32, compression: f32, shear: f32, bending: f32) -> i32;
    pub fn tc_sim_set_compliances_raw(sim: *mut TcSim, stretch: f32, shear: f32, bend: f32) -> i32;
    pub fn tc_sim_set_hash_cell_size(sim: *mut TcSim, cell_size: f32) -> i32;
    pub fn tc_sim_set_solver_path(sim: *mut TcSim, path: i32) -> i32;
    pub fn tc_sim_set_solver_mode(sim: *mut TcSim, mode: u32) -> i32;
    pub fn tc_sim_get_solver_mode(sim: *const TcSim, mode: *mut u32) -> i32;
    pub fn tc_sim_set_enable_pair_cache(sim: *mut TcSim, enable: i32) -> i32;
    pub fn tc_sim_get_enable_pair_cache(sim: *const TcSim, enable: *mut i32) -> i32;
    pub fn tc_sim_set_self_collision_algorithm(sim: *mut TcSim, algo: u32) -> i32;
    pub fn tc_sim_get_self_collision_algorithm(sim: *const TcSim, algo: *mut u32) -> i32;
    pub fn tc_sim_set_pair_cache_options(sim: *mut TcSim, max_vt_pairs: u32, max_ee_pairs: u32, margin_mode: u32, safety_margin: f32, horizon_scale: f32, max_horizon: f32) -> i32;
    pub fn tc_sim_set_enable_pair_cache_final_fallback(sim: *mut TcSim, enable: i32) -> i32;
    pub fn tc_sim_pair_cache_state(sim: *const TcSim, out: *mut f64) -> i32;
    pub fn tc_sim_get_active_pairs(sim: *mut TcSim, counts: *mut u32, vt: *mut u32, vt_cap: usize, n_vt: *mut usize, ee: *mut u32, ee_cap: usize, n_ee: *mut usize) -> i32;
    pub fn tc_sim_set_enable_edge_collision(sim: *mut TcSim, enable: i32) -> i32;
    pub fn tc_sim_set_edge_margin(sim: *mut TcSim, scale: f32, offset: f32) -> i32;
    pub fn tc_sim_collider_edge_count(sim: *const TcSim, n: *mut u32) -> i32;
    pub fn tc_sim_get_collider_edges(sim: *mut TcSim, out: *mut f32, cap: u32) -> i32;
    pub fn tc_sim_set_enable_wrinkle_field(sim: *mut TcSim, enable: i32) -> i32;
    pub fn tc_sim_wrinkle_field_enabled(sim: *const TcSim, enabled: *mut i32) -> i32;
    pub fn tc_sim_set_wrinkle_field_params(sim: *mut TcSim, frame: *const f32, z_min: f32, z_max: f32, r_min: f32, r_max: f32, valley_window: f32, crest_window: f32) -> i32;
    pub fn tc_sim_get_wrinkle_field_params(sim: *const TcSim, out: *mut f32) -> i32;
    pub fn tc_sim_set_wrinkle_field_texture_2d(sim: *mut TcSim, width: u32, height: u32, rgba: *const u8, nbytes: usize, z_min: f32, z_max: f32, r_min: f32, r_max: f32) -> i32;
    pub fn tc_sim_set_fast_projection(sim: *mut TcSim, enable: i32) -> i32;
    pub fn tc_sim_get_fast_projection(sim: *const TcSim, enabled: *mut i32) -> i32;
    pub fn tc_sim_solver_info(sim: *mut TcSim, out: *mut f64) -> i32;
    pub fn tc_sim_autotune(sim: *mut TcSim, dt_sub: f32, out: *mut f64) -> i32;
    pub fn tc_sim_num_vertices(sim: *const TcSim) -> u32;
    pub fn tc_sim_num_distance_constraints(sim: *const TcSim) -> u32;
    pub fn tc_sim_num_bending_constraints(sim: *const TcSim) -> u32;
    pub fn tc_sim_get_array(sim: *mut TcSim, which: i32, out: *mut c_void, capacity: usize, count: *mut usize) -> i32;
    pub fn tc_sim_build_hashes(sim: *mut TcSim) -> i32;
    pub fn tc_sim_residual(sim: *mut TcSim, dt_sub: f32, rel: *mut f64, max_rel: *mut f64) -> i32;
    pub fn tc_sim_substep_parts(sim: *mut TcSim, dt_sub: f32, collide: i32, self_collide: i32, update_velocity: i32) -> i32;
    pub fn tc_sim_time_kernels(sim: *mut TcSim, dt_sub: f32, reps: u32, out: *mut f64) -> i32;
    pub fn tc_sim_set_profiling_enabled(sim: *mut TcSim, enable: i32, enabled: *mut i32) -> i32;
    pub fn tc_sim_take_profile(sim: *mut TcSim, count: *mut usize) -> i32;
    pub fn tc_sim_profile_entry(sim: *const TcSim, index: usize, name: *mut *const c_char, ms: *mut f32, start_ms: *mut f32, depth: *mut u32) -> i32;
    pub fn tc_sim_save_profile_trace(sim: *mut TcSim, path: *const c_char) -> i32;
    pub fn tc_sim_synchronize(sim: *mut TcSim) -> i32;
    pub fn tc_ws_net_create(rounds: u32, out: *mut *mut TcWsNet) -> i32;
    pub fn tc_ws_net_destroy(net: *mut TcWsNet);
    pub fn tc_ws_net_num_params(net: *const TcWsNet) -> u32;
    pub fn tc_ws_net_rounds(net: *const TcWsNet) -> u32;
    pub fn tc_ws_net_set_params(net: *mut TcWsNet, params: *const f32, count: usize) -> i32;
    pub fn tc_ws_net_get_params(net: *mut TcWsNet, params: *mut f32, count: usize) -> i32;
    pub fn tc_ws_net_init(net: *mut TcWsNet, seed: u64) -> i32;
    pub fn tc_ws_net_forward_backward(net: *mut TcWsNet, features: *const f32, rows: u32, graph_off: *const u32, graph_nbr: *const u32, targets: *const f32, y: *mut f32, grad: *mut f32, loss: *mut f64) -> i32;
    pub fn tc_ws_net_adam_step(net: *mut TcWsNet, lr: f32) -> i32;
    pub fn tc_ws_net_forward_noalloc(net: *mut TcWsNet, features: *const f32, rows: u32, graph_off: *const u32, graph_nbr: *const u32, y: *mut f32) -> i32;
    pub fn tc_ws_train(net: *mut TcWsNet, sims: *const *mut TcSim, nsims: u32, steps: u32, lr: f32, seed: u64, losses: *mut f32) -> i32;
    pub fn tc_ws_eval(net: *mut TcWsNet, sims: *const *mut TcSim, nsims: u32, mean_loss: *mut f64) -> i32;
    pub fn tc_sim_set_warm_start(sim: *mut TcSim, net: *mut TcWsNet, mode: i32, iterations: u32, fallback_iterations: u32, residual_threshold: f32, history_weight: f32) -> i32;
    pub fn tc_sim_warm_start_stats(sim: *mut TcSim, out: *mut f64) -> i32;
    pub fn tc_sim_ws_reset_stats(sim: *mut TcSim) -> i32;
    pub fn tc_sim_ws_record(sim: *mut TcSim, capacity: u32, stride: u32, history_weight: f32) -> i32;
    pub fn tc_sim_ws_num_records(sim: *const TcSim) -> u32;
    pub fn tc_sim_ws_get_record(sim: *mut TcSim, idx: u32, features: *mut f32, target: *mut f32, prediction: *mut f32, previous: *mut f32) -> i32;
    pub fn tc_sim_ws_last_output(sim: *mut TcSim, y: *mut f32, count: usize) -> i32;
    pub fn tc_sim_ws_history(sim: *mut TcSim, get_xyz: *mut f32, set_xyz: *const f32, n: u32) -> i32;
    pub fn tc_sim_ws_time(sim: *mut TcSim, dt_sub: f32, reps: u32, out: *mut f64) -> i32;
    pub fn tc_bake_mesh_sdf(verts: *const f32, num_verts: u32, tris: *const i32, num_tris: u32, voxel_size: f32, margin: f32, friction: f32, thickness: f32, restitution: f32, close_open_boundaries: i32, method: i32, out: *mut *mut TcSdfResult) -> i32;
    pub fn tc_sdf_result_info(res: *const TcSdfResult, dims: *mut u32, bone_info: *mut f32, coarse_voxel: *mut f32) -> i32;
    pub fn tc_sdf_result_timing(res: *const TcSdfResult, kernel_ms: *mut f32, fallback_voxels: *mut u32) -> i32;
    pub fn tc_sdf_result_texels(res: *const TcSdfResult) -> *const u32;
    pub fn tc_sdf_result_atlas(res: *const TcSdfResult, brick_dims: *mut u32) -> *const f32;
    pub fn tc_sdf_result_nearest(res: *const TcSdfResult) -> *const u32;
    pub fn tc_sdf_result_free(res: *mut TcSdfResult);
    pub fn tc_bake_bone_sdf(verts: *const f32, num_verts: u32, tris: *const i32, num_tris: u32, weights: *const f32, binds: *const f32, num_bones: u32, resolution: u32, margin: f32, weight_threshold: f32, blend_k: f32, friction: f32, thickness: f32, restitution: f32, close_open_boundaries: i32, out: *mut *mut TcBoneSdfResult) -> i32;
    pub fn tc_bone_sdf_result_info(res: *const TcBoneSdfResult, dims: *mut u32, n_active: *mut u32, kernel_ms: *mut f32) -> i32;
    pub fn tc_bone_sdf_result_arrays(res: *const TcBoneSdfResult, texels: *mut u32, infos: *mut f32, active: *mut u32, binds: *mut f32) -> i32;
    pub fn tc_bone_sdf_result_free(res: *mut TcBoneSdfResult);
    pub fn tc_close_open_boundary_loops(verts: *const f32, num_verts: u32, tris: *const i32, num_tris: u32, out_verts: *mut f32, cap_verts: usize, out_tris: *mut i32, cap_tris: usize, num_verts_out: *mut u32, num_tris_out: *mut u32, loops_closed: *mut u32) -> i32;
    pub fn tc_mesh_build_arrays(positions: *const f32, num_vertices: u32, edges: *const u32, num_edges: u32, faces: *const u32, num_faces: u32, inv_masses: *const f32, thicknesses: *const f32, params: *const TcParams, which: i32, out: *mut c_void, capacity: usize, count: *mut usize) -> i32;
    pub fn tc_mesh_build_arrays_sewn(positions: *const f32, num_vertices: u32, edges: *const u32, num_edges: u32, faces: *const u32, num_faces: u32, inv_masses: *const f32, thicknesses: *const f32, sewing: *const u32, num_sewing: u32, sewing_params: *const TcSewingParams, params: *const TcParams, which: i32, out: *mut c_void, capacity: usize, count: *mut usize) -> i32;
    pub fn tc_color_edge_pairs(num_vertices: u32, pairs: *const u32, num_pairs: u32, colors_out: *mut u32) -> i32;
    pub fn tc_compute_inv_masses(positions: *const f32, num_vertices: u32, faces: *const u32, num_faces: u32, pin_weights: *const f32, areal_density: f32, out: *mut f32) -> i32;
    pub fn tc_last_error() -> *const c_char;
    pub fn tc_version() -> *const c_char;
}
