"""ctypes binding of libtcloth_hip.so (the C ABI declared in include/tc_sim.h).

Loading fails loudly: there is no CPU fallback.  If the library has not been
built, ``load()`` raises with the build command to run.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libtcloth_hip.so")

# Every symbol include/tc_sim.h declares (checked by the test suite).
EXPORTS = [
    "tc_sim_params_default", "tc_sim_create", "tc_sim_create_batch", "tc_sim_instance_count",
    "tc_sim_instance_range", "tc_sim_destroy", "tc_sim_step", "tc_sim_step_single_substep",
    "tc_sim_get_positions", "tc_sim_set_positions_velocities",
    "tc_sim_set_adaptive_substep_options", "tc_sim_set_auto_compensate_iterations", "tc_sim_set_strain_adaptive_options",
    "tc_sim_set_enable_compact_readback", "tc_sim_note_external_displacement", "tc_sim_frame_policy_state",
    "tc_sim_step_async", "tc_sim_fetch_positions", "tc_sim_step_buffered", "tc_sim_fetch_buffered_positions",
    "tc_sim_get_buffered_frame_count", "tc_sim_clear_frame_buffer", "tc_sim_reset", "tc_sim_set_pins", "tc_sim_release_pins",
    "tc_sim_clear_pins", "tc_sim_pin_count", "tc_sim_get_pins", "tc_sim_set_vertex_inv_masses",
    "tc_sim_get_vertex_inv_masses", "tc_sim_set_solver_mode", "tc_sim_get_solver_mode",
    "tc_sim_set_enable_pair_cache", "tc_sim_get_enable_pair_cache", "tc_sim_set_self_collision_algorithm",
    "tc_sim_get_self_collision_algorithm", "tc_sim_set_pair_cache_options", "tc_sim_set_enable_pair_cache_final_fallback",
    "tc_sim_pair_cache_state", "tc_sim_get_active_pairs", "tc_sim_set_enable_edge_collision", "tc_sim_set_edge_margin",
    "tc_sim_collider_edge_count", "tc_sim_get_collider_edges", "tc_sim_set_enable_wrinkle_field",
    "tc_sim_wrinkle_field_enabled", "tc_sim_set_wrinkle_field_params", "tc_sim_get_wrinkle_field_params",
    "tc_sim_set_wrinkle_field_texture_2d", "tc_sim_add_sphere",
    "tc_sim_add_capsule", "tc_sim_add_plane", "tc_sim_clear_colliders", "tc_sim_set_mesh_triangles",
    "tc_sim_set_gravity", "tc_sim_set_damping", "tc_sim_set_damping_all", "tc_sim_set_solver_iterations",
    "tc_sim_set_enable_self_collision", "tc_sim_set_self_collision_options", "tc_sim_set_stiffness_all",
    "tc_sim_set_compliances_raw", "tc_sim_set_hash_cell_size", "tc_sim_set_solver_path", "tc_sim_solver_info", "tc_sim_set_fast_projection", "tc_sim_get_fast_projection", "tc_sim_autotune",
    "tc_sim_num_vertices",
    "tc_sim_num_distance_constraints", "tc_sim_num_bending_constraints", "tc_sim_get_array",
    "tc_sim_build_hashes", "tc_sim_residual", "tc_sim_substep_parts", "tc_sim_time_kernels", "tc_sim_synchronize",
    "tc_sim_set_profiling_enabled", "tc_sim_take_profile", "tc_sim_profile_entry", "tc_sim_save_profile_trace",
    "tc_mesh_build_arrays", "tc_color_edge_pairs", "tc_compute_inv_masses", "tc_last_error", "tc_version",
    "tc_ws_net_create", "tc_ws_net_destroy", "tc_ws_net_num_params", "tc_ws_net_rounds", "tc_ws_net_set_params",
    "tc_ws_net_get_params", "tc_ws_net_init", "tc_ws_net_forward_backward", "tc_ws_net_adam_step", "tc_ws_train",
    "tc_sim_set_warm_start", "tc_sim_warm_start_stats", "tc_sim_ws_reset_stats", "tc_sim_ws_record",
    "tc_sim_ws_num_records", "tc_sim_ws_get_record", "tc_sim_ws_last_output", "tc_sim_ws_history",
    "tc_ws_eval", "tc_sim_ws_time", "tc_ws_net_forward_noalloc",
    "tc_sim_set_sdf_colliders", "tc_sim_update_bone_transforms", "tc_sim_clear_bone_sdf_colliders",
    "tc_sim_setup_dynamic_bone_sdf", "tc_sim_set_dynamic_bone_sdf_dirty_bones", "tc_sim_set_dynamic_bone_sdf_enabled",
    "tc_sim_set_dynamic_bone_sdf_update_interval", "tc_sim_dynamic_bone_sdf_stats", "tc_bake_mesh_sdf", "tc_sdf_result_info",
    "tc_sdf_result_timing", "tc_sdf_result_texels", "tc_sdf_result_atlas", "tc_sdf_result_nearest", "tc_sdf_result_free", "tc_close_open_boundary_loops",
    "tc_bake_bone_sdf", "tc_bone_sdf_result_info", "tc_bone_sdf_result_arrays", "tc_bone_sdf_result_free",
    "tc_sewing_params_default", "tc_sim_create_sewn", "tc_sim_set_sewing_stiffness", "tc_sim_set_sewing_shrink_speed",
    "tc_sim_set_enable_sewing_lock", "tc_sim_set_sewing_lock_distance", "tc_sim_get_sewing_state",
    "tc_sim_get_sewing_rest_lengths", "tc_sim_set_sewing_rest_lengths", "tc_mesh_build_arrays_sewn",
    "tc_sewing_closure_ratio", "tc_sim_set_sewing_priority_options", "tc_sim_set_sewing_priority_state",
    "tc_sim_get_sewing_priority_state", "tc_sim_update_sewing_priority",
]

TC_OK, TC_ERR_VALUE, TC_ERR_RUNTIME, TC_ERR_UNSUPPORTED, TC_ERR_IO = 0, 1, 2, 3, 4


class TcInstance(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("num_vertices", C.c_uint32), ("edges", C.c_void_p),
                ("num_edges", C.c_uint32), ("faces", C.c_void_p), ("num_faces", C.c_uint32),
                ("inv_masses", C.c_void_p), ("thicknesses", C.c_void_p)]


class TcParams(C.Structure):
    _fields_ = [
        ("thickness", C.c_float),
        ("layer_id", C.c_uint32),
        ("stiffness", C.c_float),
        ("bending_stiffness", C.c_float),
        ("compression_stiffness", C.c_float),
        ("shear_stiffness", C.c_float),
        ("workgroup_size", C.c_uint32),
        ("solver_mode", C.c_uint32),
    ]


class TcSewingParams(C.Structure):
    _fields_ = [
        ("shrink_speed", C.c_float),
        ("stiffness", C.c_float),
        ("lock", C.c_int32),
        ("lock_distance", C.c_float),
    ]


_LIB: C.CDLL | None = None
_vp, _u32, _f32, _i = C.c_void_p, C.c_uint32, C.c_float, C.c_int


def load() -> C.CDLL:
    """Load libtcloth_hip.so and declare every entry point's signature."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m taremincloth_amd.build` "
            "(there is no CPU fallback for the XPBD substep)")
    L = C.CDLL(LIB_PATH)
    sig = {
        "tc_sim_params_default": (None, [C.POINTER(TcParams)]),
        "tc_sim_create": (_i, [_vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp, _vp, C.POINTER(TcParams), C.POINTER(_vp)]),
        "tc_sim_create_batch": (_i, [_vp, _u32, C.POINTER(TcParams), C.POINTER(_vp)]),
        "tc_sim_instance_count": (_i, [_vp, C.POINTER(_u32)]),
        "tc_sim_instance_range": (_i, [_vp, _u32, C.POINTER(_u32), C.POINTER(_u32)]),
        "tc_sim_destroy": (None, [_vp]),
        "tc_sim_step": (_i, [_vp, _f32, _u32, _u32, C.POINTER(_u32)]),
        "tc_sim_step_single_substep": (_i, [_vp, _f32]),
        "tc_sim_get_positions": (_i, [_vp, _vp, C.c_size_t]),
        "tc_sim_set_adaptive_substep_options": (_i, [_vp, _i, _u32, _u32]),
        "tc_sim_set_auto_compensate_iterations": (_i, [_vp, _i]),
        "tc_sim_set_strain_adaptive_options": (_i, [_vp, _i, _f32]),
        "tc_sim_set_enable_compact_readback": (_i, [_vp, _i]),
        "tc_sim_note_external_displacement": (_i, [_vp, _f32]),
        "tc_sim_frame_policy_state": (_i, [_vp, C.POINTER(C.c_double)]),
        "tc_sim_step_async": (_i, [_vp, _f32, _u32, _u32]),
        "tc_sim_fetch_positions": (_i, [_vp, _vp, C.c_size_t, C.POINTER(_i)]),
        "tc_sim_step_buffered": (_i, [_vp, _f32, _u32, _u32, C.POINTER(_u32)]),
        "tc_sim_fetch_buffered_positions": (_i, [_vp, _vp, C.c_size_t, C.POINTER(_u32)]),
        "tc_sim_get_buffered_frame_count": (_i, [_vp, C.POINTER(_u32)]),
        "tc_sim_clear_frame_buffer": (_i, [_vp]),
        "tc_sim_set_positions_velocities": (_i, [_vp, _vp, _vp, _u32]),
        "tc_sim_reset": (_i, [_vp]),
        "tc_sim_set_pins": (_i, [_vp, _vp, _vp, _vp, _u32]),
        "tc_sim_release_pins": (_i, [_vp, _vp, _u32]),
        "tc_sim_clear_pins": (_i, [_vp]),
        "tc_sim_pin_count": (_i, [_vp, _vp]),
        "tc_sim_get_pins": (_i, [_vp, _vp, _vp, _vp, _u32, _vp]),
        "tc_sim_set_vertex_inv_masses": (_i, [_vp, _vp, _vp, _u32]),
        "tc_sim_get_vertex_inv_masses": (_i, [_vp, _vp, C.c_size_t]),
        "tc_sim_set_solver_mode": (_i, [_vp, _u32]),
        "tc_sim_get_solver_mode": (_i, [_vp, C.POINTER(_u32)]),
        "tc_sim_set_enable_pair_cache": (_i, [_vp, _i]),
        "tc_sim_get_enable_pair_cache": (_i, [_vp, C.POINTER(_i)]),
        "tc_sim_set_self_collision_algorithm": (_i, [_vp, _u32]),
        "tc_sim_get_self_collision_algorithm": (_i, [_vp, C.POINTER(_u32)]),
        "tc_sim_set_pair_cache_options": (_i, [_vp, _u32, _u32, _u32, _f32, _f32, _f32]),
        "tc_sim_set_enable_pair_cache_final_fallback": (_i, [_vp, _i]),
        "tc_sim_pair_cache_state": (_i, [_vp, C.POINTER(C.c_double)]),
        "tc_sim_set_enable_edge_collision": (_i, [_vp, _i]),
        "tc_sim_set_edge_margin": (_i, [_vp, _f32, _f32]),
        "tc_sim_collider_edge_count": (_i, [_vp, C.POINTER(_u32)]),
        "tc_sim_get_collider_edges": (_i, [_vp, _vp, _u32]),
        "tc_sim_set_enable_wrinkle_field": (_i, [_vp, _i]),
        "tc_sim_wrinkle_field_enabled": (_i, [_vp, C.POINTER(_i)]),
        "tc_sim_set_wrinkle_field_params": (_i, [_vp, _vp, _f32, _f32, _f32, _f32, _f32, _f32]),
        "tc_sim_get_wrinkle_field_params": (_i, [_vp, _vp]),
        "tc_sim_set_wrinkle_field_texture_2d": (_i, [_vp, _u32, _u32, _vp, C.c_size_t, _f32, _f32, _f32, _f32]),
        "tc_sim_get_active_pairs": (_i, [_vp, _vp, _vp, C.c_size_t, C.POINTER(C.c_size_t), _vp, C.c_size_t,
                                         C.POINTER(C.c_size_t)]),
        "tc_sim_add_sphere": (_i, [_vp, _vp, _f32, _f32, _f32]),
        "tc_sim_add_capsule": (_i, [_vp, _vp, _vp, _f32, _f32, _f32]),
        "tc_sim_add_plane": (_i, [_vp, _vp, _vp, _f32, _f32]),
        "tc_sim_clear_colliders": (_i, [_vp]),
        "tc_sim_set_mesh_triangles": (_i, [_vp, _vp, _u32, _vp, _f32, _f32, _f32, _i]),
        "tc_sim_set_gravity": (_i, [_vp, _f32, _f32, _f32]),
        "tc_sim_set_damping": (_i, [_vp, _f32]),
        "tc_sim_set_damping_all": (_i, [_vp, _f32, _f32, _f32, _f32]),
        "tc_sim_set_solver_iterations": (_i, [_vp, _u32]),
        "tc_sim_set_enable_self_collision": (_i, [_vp, _i]),
        "tc_sim_set_self_collision_options": (_i, [_vp, _f32, _f32, _i, _i, _u32]),
        "tc_sim_set_stiffness_all": (_i, [_vp, _f32, _f32, _f32, _f32]),
        "tc_sim_set_compliances_raw": (_i, [_vp, _f32, _f32, _f32]),
        "tc_sim_set_hash_cell_size": (_i, [_vp, _f32]),
        "tc_sim_set_solver_path": (_i, [_vp, _i]),
        "tc_sim_set_fast_projection": (_i, [_vp, _i]),
        "tc_sim_get_fast_projection": (_i, [_vp, C.POINTER(_i)]),
        "tc_sim_solver_info": (_i, [_vp, C.POINTER(C.c_double)]),
        "tc_sim_autotune": (_i, [_vp, _f32, C.POINTER(C.c_double)]),
        "tc_sim_num_vertices": (_u32, [_vp]),
        "tc_sim_num_distance_constraints": (_u32, [_vp]),
        "tc_sim_num_bending_constraints": (_u32, [_vp]),
        "tc_sim_get_array": (_i, [_vp, _i, _vp, C.c_size_t, C.POINTER(C.c_size_t)]),
        "tc_sim_build_hashes": (_i, [_vp]),
        "tc_sim_residual": (_i, [_vp, _f32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "tc_sim_substep_parts": (_i, [_vp, _f32, _i, _i, _i]),
        "tc_sim_time_kernels": (_i, [_vp, _f32, _u32, C.POINTER(C.c_double)]),
        "tc_sim_synchronize": (_i, [_vp]),
        "tc_sim_set_profiling_enabled": (_i, [_vp, _i, C.POINTER(_i)]),
        "tc_sim_take_profile": (_i, [_vp, C.POINTER(C.c_size_t)]),
        "tc_sim_profile_entry": (_i, [_vp, C.c_size_t, C.POINTER(C.c_char_p), C.POINTER(_f32), C.POINTER(_f32),
                                      C.POINTER(_u32)]),
        "tc_sim_save_profile_trace": (_i, [_vp, C.c_char_p]),
        "tc_mesh_build_arrays": (_i, [_vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp, C.POINTER(TcParams), _i, _vp,
                                      C.c_size_t, C.POINTER(C.c_size_t)]),
        "tc_color_edge_pairs": (_i, [_u32, _vp, _u32, _vp]),
        "tc_compute_inv_masses": (_i, [_vp, _u32, _vp, _u32, _vp, _f32, _vp]),
        "tc_ws_net_create": (_i, [_u32, C.POINTER(_vp)]),
        "tc_ws_net_destroy": (None, [_vp]),
        "tc_ws_net_num_params": (_u32, [_vp]),
        "tc_ws_net_rounds": (_u32, [_vp]),
        "tc_ws_net_set_params": (_i, [_vp, _vp, C.c_size_t]),
        "tc_ws_net_get_params": (_i, [_vp, _vp, C.c_size_t]),
        "tc_ws_net_init": (_i, [_vp, C.c_uint64]),
        "tc_ws_net_forward_backward": (_i, [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_double)]),
        "tc_ws_net_adam_step": (_i, [_vp, _f32]),
        "tc_ws_net_forward_noalloc": (_i, [_vp, _vp, _u32, _vp, _vp, _vp]),
        "tc_ws_train": (_i, [_vp, _vp, _u32, _u32, _f32, C.c_uint64, _vp]),
        "tc_sim_set_warm_start": (_i, [_vp, _vp, _i, _u32, _u32, _f32, _f32]),
        "tc_sim_warm_start_stats": (_i, [_vp, C.POINTER(C.c_double)]),
        "tc_sim_ws_reset_stats": (_i, [_vp]),
        "tc_sim_ws_record": (_i, [_vp, _u32, _u32, _f32]),
        "tc_sim_ws_num_records": (_u32, [_vp]),
        "tc_sim_ws_get_record": (_i, [_vp, _u32, _vp, _vp, _vp, _vp]),
        "tc_sim_ws_last_output": (_i, [_vp, _vp, C.c_size_t]),
        "tc_sim_ws_history": (_i, [_vp, _vp, _vp, _u32]),
        "tc_ws_eval": (_i, [_vp, _vp, _u32, C.POINTER(C.c_double)]),
        "tc_sim_ws_time": (_i, [_vp, _f32, _u32, C.POINTER(C.c_double)]),
        "tc_sim_set_sdf_colliders": (_i, [_vp, _u32, _u32, _u32, _vp, C.c_size_t, _vp, _u32]),
        "tc_sim_update_bone_transforms": (_i, [_vp, _vp, _vp, _u32]),
        "tc_sim_clear_bone_sdf_colliders": (_i, [_vp]),
        "tc_sim_setup_dynamic_bone_sdf": (_i, [_vp, _u32, _u32, _u32, _u32, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp,
                                               _u32, _vp, _u32, _u32]),
        "tc_sim_set_dynamic_bone_sdf_dirty_bones": (_i, [_vp, _vp, _u32, _i]),
        "tc_sim_set_dynamic_bone_sdf_enabled": (_i, [_vp, _i]),
        "tc_sim_set_dynamic_bone_sdf_update_interval": (_i, [_vp, _u32]),
        "tc_sim_dynamic_bone_sdf_stats": (_i, [_vp, _vp]),
        "tc_bake_mesh_sdf": (_i, [_vp, _u32, _vp, _u32, _f32, _f32, _f32, _f32, _f32, _i, _i, C.POINTER(_vp)]),
        "tc_sdf_result_info": (_i, [_vp, _vp, _vp, C.POINTER(_f32)]),
        "tc_sdf_result_timing": (_i, [_vp, _vp, _vp]),
        "tc_sdf_result_texels": (_vp, [_vp]),
        "tc_sdf_result_atlas": (_vp, [_vp, _vp]),
        "tc_sdf_result_nearest": (_vp, [_vp]),
        "tc_sdf_result_free": (None, [_vp]),
        "tc_bake_bone_sdf": (_i, [_vp, _u32, _vp, _u32, _vp, _vp, _u32, _u32, _f32, _f32, _f32, _f32, _f32, _f32, _i,
                                 C.POINTER(_vp)]),
        "tc_bone_sdf_result_info": (_i, [_vp, _vp, _vp, _vp]),
        "tc_bone_sdf_result_arrays": (_i, [_vp, _vp, _vp, _vp, _vp]),
        "tc_bone_sdf_result_free": (None, [_vp]),
        "tc_sewing_params_default": (None, [_vp]),
        "tc_sim_create_sewn": (_i, [_vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp, _u32, _vp, _vp,
                                   C.POINTER(_vp)]),
        "tc_sim_set_sewing_stiffness": (_i, [_vp, _f32]),
        "tc_sim_set_sewing_shrink_speed": (_i, [_vp, _f32]),
        "tc_sim_set_enable_sewing_lock": (_i, [_vp, _i]),
        "tc_sim_set_sewing_lock_distance": (_i, [_vp, _f32]),
        "tc_sim_get_sewing_state": (_i, [_vp, _vp]),
        "tc_sim_get_sewing_rest_lengths": (_i, [_vp, _vp, _u32]),
        "tc_sim_set_sewing_rest_lengths": (_i, [_vp, _vp, _u32]),
        "tc_sewing_closure_ratio": (_i, [_vp, _u32, _vp, _u32, C.c_float, _vp]),
        "tc_sim_set_sewing_priority_options": (_i, [_vp, _i, C.c_float, C.c_float, _u32, _u32]),
        "tc_sim_set_sewing_priority_state": (_i, [_vp, _i, _u32, C.c_float, C.c_float]),
        "tc_sim_get_sewing_priority_state": (_i, [_vp, _vp]),
        "tc_sim_update_sewing_priority": (_i, [_vp, _vp, _u32, _vp]),
        "tc_mesh_build_arrays_sewn": (_i, [_vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _u32, _vp, _vp, _i, _vp,
                                          C.c_size_t, C.POINTER(C.c_size_t)]),
        "tc_close_open_boundary_loops": (_i, [_vp, _u32, _vp, _u32, _vp, C.c_size_t, _vp, C.c_size_t,
                                              C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
        "tc_last_error": (C.c_char_p, []),
        "tc_version": (C.c_char_p, []),
    }
    for name in EXPORTS:
        fn = getattr(L, name)  # AttributeError if the symbol is missing
        fn.restype, fn.argtypes = sig[name]
    _LIB = L
    return L


def check(status: int) -> None:
    """Raise the reference's exception type for a non-zero status."""
    if status == TC_OK:
        return
    msg = (load().tc_last_error() or b"").decode(errors="replace")
    if status == TC_ERR_VALUE:
        raise ValueError(msg)
    if status == TC_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if status == TC_ERR_IO:
        raise OSError(msg)
    raise RuntimeError(msg)
