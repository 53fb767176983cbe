// Layout of the C ABI structures (include/tc_sim.h), checked at compile time by both compilers
// that build the library (g++ for the host mesh build, hipcc for the HIP sources).  Bindings in
// other languages (ctypes, Rust #[repr(C)]) must use the same sizes and offsets; the CPU test
// suite reads the numbers below and compares them with the ctypes structures.  This stands in
// for the reference's host/shader layout test (crates/cloth_core/tests/test_shader_alignment.rs:225),
// which checks its Rust structs against the WGSL structs they are uploaded into.
#pragma once
#include <cstddef>

#include "../../include/tc_sim.h"

static_assert(sizeof(tc_sim_params) == 32, "tc_sim_params size");
static_assert(offsetof(tc_sim_params, thickness) == 0, "tc_sim_params.thickness");
static_assert(offsetof(tc_sim_params, layer_id) == 4, "tc_sim_params.layer_id");
static_assert(offsetof(tc_sim_params, stiffness) == 8, "tc_sim_params.stiffness");
static_assert(offsetof(tc_sim_params, bending_stiffness) == 12, "tc_sim_params.bending_stiffness");
static_assert(offsetof(tc_sim_params, compression_stiffness) == 16, "tc_sim_params.compression_stiffness");
static_assert(offsetof(tc_sim_params, shear_stiffness) == 20, "tc_sim_params.shear_stiffness");
static_assert(offsetof(tc_sim_params, workgroup_size) == 24, "tc_sim_params.workgroup_size");
static_assert(offsetof(tc_sim_params, solver_mode) == 28, "tc_sim_params.solver_mode");

static_assert(sizeof(tc_instance) == 64, "tc_instance size");
static_assert(offsetof(tc_instance, positions) == 0, "tc_instance.positions");
static_assert(offsetof(tc_instance, num_vertices) == 8, "tc_instance.num_vertices");
static_assert(offsetof(tc_instance, edges) == 16, "tc_instance.edges");
static_assert(offsetof(tc_instance, num_edges) == 24, "tc_instance.num_edges");
static_assert(offsetof(tc_instance, faces) == 32, "tc_instance.faces");
static_assert(offsetof(tc_instance, num_faces) == 40, "tc_instance.num_faces");
static_assert(offsetof(tc_instance, inv_masses) == 48, "tc_instance.inv_masses");
static_assert(offsetof(tc_instance, thicknesses) == 56, "tc_instance.thicknesses");

static_assert(sizeof(tc_sewing_params) == 16, "tc_sewing_params size");
static_assert(offsetof(tc_sewing_params, shrink_speed) == 0, "tc_sewing_params.shrink_speed");
static_assert(offsetof(tc_sewing_params, stiffness) == 4, "tc_sewing_params.stiffness");
static_assert(offsetof(tc_sewing_params, lock) == 8, "tc_sewing_params.lock");
static_assert(offsetof(tc_sewing_params, lock_distance) == 12, "tc_sewing_params.lock_distance");
