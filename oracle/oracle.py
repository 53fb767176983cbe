"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module.  The product (``taremincloth_amd``) never
does.  ``OracleSim`` mirrors the reference ``ClothSimulator`` surface
(crates/cloth_py/src/lib.rs:1556-1786, 2110-2175) so parity tests read like the
reference's own tests.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "build")
_LIBS: dict[str, C.CDLL] = {}

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_vp = C.c_void_p

# orc_mesh_get field ids (xpbd_oracle.c)
FIELDS = [
    "dist_v0", "dist_v1", "dist_type", "dist_rest", "dist_tc", "dist_cc", "dist_col_off", "dist_col_cnt",
    "bend_v0", "bend_v1", "bend_v2", "bend_v3", "bend_rest", "bend_comp", "bend_pad0", "bend_col_off",
    "bend_col_cnt", "adj_off", "adj_idx", "th_off", "th_idx", "th_rest", "local_len", "star_off",
    "star_v0", "star_v1", "edge_v0", "edge_v1", "edge_bbits", "island", "scalars", "edge_to_constraint",
    "layer", "bend_pad1", "sew_v0", "sew_v1", "sew_rest", "sew_col_off", "sew_col_cnt",
]
_FLOAT_FIELDS = {"dist_rest", "dist_tc", "dist_cc", "bend_rest", "bend_comp", "bend_pad0", "bend_pad1", "th_rest",
                 "local_len", "sew_rest"}


def build() -> None:
    """Compile the oracle libraries with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _opt(arr, dtype):
    if arr is None:
        return None
    return np.ascontiguousarray(arr, dtype=dtype)


def lib(precision: str = "f32") -> C.CDLL:
    if precision in _LIBS:
        return _LIBS[precision]
    path = os.path.join(_BUILD, f"liborc_{precision}.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    L.orc_mesh_build.restype = _vp
    L.orc_mesh_build.argtypes = [_vp, C.c_uint32, _vp, C.c_uint32, _vp, C.c_uint32, _vp, _vp, _vp, C.c_uint32,
                                 C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]
    L.orc_mesh_build_sewn.restype = _vp
    L.orc_mesh_build_sewn.argtypes = [_vp, C.c_uint32, _vp, C.c_uint32, _vp, C.c_uint32, _vp, _vp, _vp, C.c_uint32,
                                      C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, _vp, C.c_uint32,
                                      C.c_float]
    L.orc_sim_create_sewn.restype = _vp
    L.orc_sim_create_sewn.argtypes = [_vp, C.c_uint32, _vp, C.c_uint32, _vp, C.c_uint32, _vp, _vp, _vp, C.c_uint32,
                                      C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, _vp, C.c_uint32,
                                      C.c_float, C.c_float]
    L.orc_sim_set_sewing.argtypes = [_vp, C.c_float, C.c_int, C.c_float, C.c_float]
    L.orc_sim_sewing_rest_lengths.argtypes = [_vp, _vp, _vp]
    L.orc_mesh_get.restype = C.c_uint32
    L.orc_mesh_get.argtypes = [_vp, C.c_int, _vp]
    L.orc_mesh_free.argtypes = [_vp]
    L.orc_sim_create.restype = _vp
    L.orc_sim_create.argtypes = [_vp, C.c_uint32, _vp, C.c_uint32, _vp, C.c_uint32, _vp, _vp, _vp, C.c_uint32,
                                 C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]
    L.orc_sim_free.argtypes = [_vp]
    L.orc_sim_mesh.restype = _vp
    L.orc_sim_mesh.argtypes = [_vp]
    L.orc_sim_set_gravity.argtypes = [_vp, C.c_float, C.c_float, C.c_float]
    L.orc_sim_set_damping.argtypes = [_vp, C.c_float]
    L.orc_sim_set_damping_all.argtypes = [_vp, C.c_float, C.c_float, C.c_float, C.c_float]
    L.orc_sim_set_iterations.argtypes = [_vp, C.c_uint32]
    L.orc_sim_set_self_collision.argtypes = [_vp, C.c_int]
    L.orc_sim_set_hash_cell.argtypes = [_vp, C.c_float]
    L.orc_sim_set_self_collision_options.argtypes = [_vp, C.c_float, C.c_float, C.c_int, C.c_int, C.c_uint32]
    L.orc_sim_set_compliances_raw.argtypes = [_vp, C.c_float, C.c_float, C.c_float]
    L.orc_sim_add_sphere.argtypes = [_vp, _f32p, C.c_float, C.c_float, C.c_float]
    L.orc_sim_add_capsule.argtypes = [_vp, _f32p, _f32p, C.c_float, C.c_float, C.c_float]
    L.orc_sim_add_plane.argtypes = [_vp, _f32p, _f32p, C.c_float, C.c_float]
    L.orc_sim_clear_colliders.argtypes = [_vp]
    L.orc_sim_set_mesh_triangles.argtypes = [_vp, _f32p, C.c_uint32, _vp, C.c_float, C.c_float, C.c_float, C.c_int]
    L.orc_sim_reset.argtypes = [_vp]
    L.orc_sim_set_state.argtypes = [_vp, _vp, _vp, _vp]
    L.orc_sim_get.argtypes = [_vp, C.c_int, _f64p]
    L.orc_sim_get_lambdas.argtypes = [_vp, _f64p, _f64p]
    L.orc_sim_set_lambdas.argtypes = [_vp, _vp, _vp]
    L.orc_sim_step.restype = C.c_uint32
    L.orc_sim_step.argtypes = [_vp, C.c_float, C.c_uint32]
    L.orc_sim_step_single_substep.argtypes = [_vp, C.c_float]
    L.orc_sim_substep_parts.argtypes = [_vp, C.c_float, C.c_int, C.c_int, C.c_int]
    L.orc_sim_residual.argtypes = [_vp, C.c_float, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_areal_inv_masses.argtypes = [_f32p, C.c_uint32, _u32p, C.c_uint32, C.c_float, _vp, _f32p]
    L.orc_color_edge_pairs.argtypes = [C.c_uint32, _u32p, C.c_uint32, _u32p]
    L.orc_hash_points.argtypes = [_f32p, C.c_uint32, C.c_float, C.c_uint32, _u32p, _u32p, _u32p]
    L.orc_real_bytes.restype = C.c_int
    L.orc_num_threads.restype = C.c_int
    L.orc_sim_stats.argtypes = [_vp, np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS"), C.c_int]
    # SDF bake and SDF colliders (sdf_oracle.c)
    L.orc_sdf_bake.restype = _vp
    L.orc_sdf_bake.argtypes = [_f32p, C.c_uint32, _vp, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float,
                               C.c_float, C.c_int, C.c_int]
    L.orc_sdf_info.argtypes = [_vp, _u32p, _f32p, C.POINTER(C.c_float)]
    L.orc_sdf_texels.argtypes = [_vp, _u32p]
    L.orc_sdf_nearest.argtypes = [_vp, _u32p]
    L.orc_sdf_free.argtypes = [_vp]
    L.orc_f32_to_f16_n.argtypes = [_f32p, _u16p, C.c_size_t]
    L.orc_f16_to_f32_n.argtypes = [_u16p, _f32p, C.c_size_t]
    L.orc_close_loops_copy.restype = C.c_uint32
    L.orc_close_loops_copy.argtypes = [_f32p, C.c_uint32, _vp, C.c_uint32, _vp, C.c_size_t, _vp, C.c_size_t,
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.orc_sim_set_sdf.argtypes = [_vp, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, _f32p, C.c_uint32]
    L.orc_sim_set_bone_transforms.argtypes = [_vp, _f32p, _f32p, C.c_uint32]
    L.orc_sim_clear_bone_sdf.argtypes = [_vp]
    L.orc_sim_normals.argtypes = [_vp, _vp]
    L.orc_sim_setup_dynamic_bone_sdf.restype = C.c_int
    L.orc_sim_setup_dynamic_bone_sdf.argtypes = [_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _vp,
                                                 C.c_uint32, _vp, C.c_uint32, _vp, _vp, _vp, _vp, C.c_uint32, _vp,
                                                 C.c_uint32, C.c_uint32]
    L.orc_sim_set_dynamic_dirty.argtypes = [_vp, _vp, C.c_uint32, C.c_int]
    L.orc_sim_set_dynamic_enabled.argtypes = [_vp, C.c_int]
    L.orc_sim_set_dynamic_interval.argtypes = [_vp, C.c_uint32]
    L.orc_sim_dynamic_arrays.argtypes = [_vp, _vp, _vp]
    L.orc_sim_sdf_arrays.argtypes = [_vp, _vp, _vp]
    L.orc_bone_bake_tiles.argtypes = [_vp, C.c_uint32, _vp, _vp]
    L.orc_sim_set_pin_table.argtypes = [_vp, _vp, _vp, _vp, C.c_uint32]
    L.orc_sim_set_inv_masses.argtypes = [_vp, _vp, _vp, C.c_uint32]
    L.orc_sim_get_inv_masses.argtypes = [_vp, C.c_int, _f32p]
    L.orc_sim_set_solver_mode.argtypes = [_vp, C.c_uint32]
    L.orc_sim_set_pair_cache.argtypes = [_vp, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                                         C.c_float, C.c_float, C.c_int]
    L.orc_sim_active_pairs.restype = C.c_uint32
    L.orc_sim_active_pairs.argtypes = [_vp, C.c_int, _vp, _vp]
    L.orc_sim_set_edge_collision.argtypes = [_vp, C.c_int, C.c_float, C.c_float]
    L.orc_atlas_layout.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, _u32p]
    L.orc_bone_sdf_bake.restype = _vp
    L.orc_bone_sdf_bake.argtypes = [_f32p, C.c_uint32, _vp, C.c_uint32, _f32p, _f32p, C.c_uint32, C.c_uint32,
                                    C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]
    L.orc_bone_sdf_info.argtypes = [_vp, _u32p, C.POINTER(C.c_uint32)]
    L.orc_bone_sdf_arrays.argtypes = [_vp, _u32p, _f32p, _u32p, _f32p]
    L.orc_bone_sdf_free.argtypes = [_vp]
    L.orc_sim_set_wrinkle.argtypes = [_vp, C.c_int, _f32p, _vp, C.c_uint32, C.c_uint32]
    L.orc_sim_collider_edges.restype = C.c_uint32
    L.orc_sim_collider_edges.argtypes = [_vp, _vp, _vp]
    _LIBS[precision] = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def mesh_fields(L: C.CDLL, mh) -> dict[str, np.ndarray]:
    out = {}
    for fid, name in enumerate(FIELDS):
        n = L.orc_mesh_get(mh, fid, None)
        arr = np.empty(n, dtype=np.float32 if name in _FLOAT_FIELDS else np.uint32)
        if n:
            L.orc_mesh_get(mh, fid, arr.ctypes.data_as(C.c_void_p))
        out[name] = arr
    s = out["scalars"]
    out["nv"], out["nd"], out["nb"], out["nd_col"], out["nb_col"], out["ne"], out["ntri"] = (int(x) for x in s[:7])
    out["n_virtual"], out["table_size"], out["has_multiple_layers"] = int(s[7]), int(s[8]), int(s[9])
    out["cell_size"], out["mean_edge"], out["max_edge"] = (float(np.array(x, np.uint32).view(np.float32)) for x in s[10:13])
    return out


def build_mesh(positions, edges, faces=None, inv_masses=None, thicknesses=None, layer_ids=None, layer_id=0,
               thickness=0.005, stiffness=1000.0, bending_stiffness=10.0, compression_stiffness=None,
               shear_stiffness=None, sewing_springs=None, sewing_shrink_speed=1.0,
               precision="f32") -> dict[str, np.ndarray]:
    """ClothMesh::from_raw_with_opts restated (mesh.rs:407-1377), returned as arrays."""
    L = lib(precision)
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    e = np.ascontiguousarray(edges, np.uint32).reshape(-1, 2)
    f = _opt(faces, np.uint32)
    comp = stiffness if compression_stiffness is None else compression_stiffness
    shear = stiffness * 0.5 if shear_stiffness is None else shear_stiffness
    im, th, ly = _opt(inv_masses, np.float32), _opt(thicknesses, np.float32), _opt(layer_ids, np.uint32)
    sw = None if sewing_springs is None else np.ascontiguousarray(sewing_springs, np.uint32).reshape(-1, 2)
    mh = L.orc_mesh_build_sewn(_ptr(pos), len(pos), _ptr(e), len(e), _ptr(f),
                               0 if f is None else len(f.reshape(-1, 3)), _ptr(im), _ptr(th), _ptr(ly), layer_id,
                               thickness, stiffness, comp, shear, bending_stiffness, _ptr(sw),
                               0 if sw is None else len(sw), sewing_shrink_speed)
    try:
        return mesh_fields(L, mh)
    finally:
        L.orc_mesh_free(mh)


def areal_inv_masses(positions, faces, density, pin_weights=None) -> np.ndarray:
    """compute_combined_inv_masses restated (mesh.rs:237-312), faces only."""
    L = lib("f32")
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
    out = np.empty(len(pos), np.float32)
    pw = _opt(pin_weights, np.float32)
    L.orc_areal_inv_masses(pos, len(pos), f, len(f), density, _ptr(pw), out)
    return out


def color_edge_pairs(n, pairs) -> np.ndarray:
    L = lib("f32")
    p = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    out = np.empty(len(p), np.uint32)
    L.orc_color_edge_pairs(n, p, len(p), out)
    return out


def hash_points(points, cell, table):
    """Spatial-hash keys, cell starts (table+1) and stably sorted indices."""
    L = lib("f32")
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    keys = np.empty(len(p), np.uint32)
    starts = np.empty(table + 1, np.uint32)
    srt = np.empty(len(p), np.uint32)
    L.orc_hash_points(p, len(p), cell, table, keys, starts, srt)
    return keys, starts, srt


class OracleSim:
    """CPU restatement of ClothSimulator (decoupled colliders, self-collision DIRECT)."""

    def __init__(self, positions, edges, faces=None, inv_masses=None, thicknesses=None, layer_ids=None,
                 layer_id=0, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0,
                 compression_stiffness=None, shear_stiffness=None, sewing_springs=None, sewing_shrink_speed=1.0,
                 sewing_lock_distance=None, precision="f32"):
        self._L = lib(precision)
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        e = np.ascontiguousarray(edges, np.uint32).reshape(-1, 2)
        f = _opt(faces, np.uint32)
        if f is not None and f.size == 0:
            f = None
        comp = stiffness if compression_stiffness is None else compression_stiffness
        shear = stiffness * 0.5 if shear_stiffness is None else shear_stiffness
        im, th, ly = _opt(inv_masses, np.float32), _opt(thicknesses, np.float32), _opt(layer_ids, np.uint32)
        sw = None if sewing_springs is None or np.asarray(sewing_springs).size == 0 else \
            np.ascontiguousarray(sewing_springs, np.uint32).reshape(-1, 2)
        self._h = self._L.orc_sim_create_sewn(_ptr(pos), len(pos), _ptr(e), len(e), _ptr(f),
                                              0 if f is None else len(f.reshape(-1, 3)), _ptr(im), _ptr(th), _ptr(ly),
                                              layer_id, thickness, stiffness, bending_stiffness, comp, shear, _ptr(sw),
                                              0 if sw is None else len(sw), sewing_shrink_speed,
                                              -1.0 if sewing_lock_distance is None else sewing_lock_distance)
        self._sew = [10000.0, 1, 0.02 if sewing_lock_distance is None else max(sewing_lock_distance, 0.0),
                     max(sewing_shrink_speed, 0.0)]
        self.n = len(pos)
        self._pins = {}  # vertex -> (target float32[3], weight float32)
        self._inv_orig = self.inv_masses(1)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.orc_sim_free(h)
            self._h = None

    def mesh(self) -> dict[str, np.ndarray]:
        return mesh_fields(self._L, self._L.orc_sim_mesh(self._h))

    # sewing: stiffness, lock, lock distance, shrink speed (the simulator's values, mod.rs:1211-1255)
    def _apply_sewing(self):
        st, lk, ld, sp = self._sew
        self._L.orc_sim_set_sewing(self._h, st, int(lk), ld, sp)

    def set_sewing_stiffness(self, v):
        self._sew[0] = max(float(v), 1.0)
        self._apply_sewing()

    def set_enable_sewing_lock(self, on):
        self._sew[1] = 1 if on else 0
        self._apply_sewing()

    def set_sewing_lock_distance(self, d):
        self._sew[2] = max(float(d), 0.0)
        self._apply_sewing()

    def set_sewing_shrink_speed(self, v):
        self._sew[3] = max(float(v), 0.0) if np.isfinite(v) else 1.0
        self._apply_sewing()

    def sewing_rest_lengths(self, values=None) -> np.ndarray:
        n = len(self.mesh()["sew_v0"])
        out = np.empty(n, np.float64)
        vin = None if values is None else np.ascontiguousarray(values, np.float64)
        self._L.orc_sim_sewing_rest_lengths(self._h, out.ctypes.data_as(C.c_void_p), _ptr(vin))
        return out

    def set_gravity(self, gx, gy, gz):
        self._L.orc_sim_set_gravity(self._h, gx, gy, gz)

    def set_damping(self, d):
        self._L.orc_sim_set_damping(self._h, d)

    def set_damping_all(self, t, c, s, b):
        self._L.orc_sim_set_damping_all(self._h, t, c, s, b)

    def set_solver_iterations(self, it):
        self._L.orc_sim_set_iterations(self._h, it)

    def set_pair_cache(self, on=True, final_fallback=True, quota_vt=None, quota_ee=None, margin_mode=1,
                       safety_margin=0.005, horizon_scale=1.3, max_horizon=0.02, brute=False):
        """Pair-cache self-collision (self_collision_algorithm 1); quotas default to the reference's
        clamp(clamp(8 N, 8192, 262144) / N, 1, 8).  brute: candidates from a scan of every vertex
        instead of the hash tables (the same lists)."""
        budget = min(max(8 * self.n, 8192), 262144)
        q = min(max(budget // max(self.n, 1), 1), 8)
        self._L.orc_sim_set_pair_cache(self._h, int(on), int(final_fallback), q if quota_vt is None else quota_vt,
                                       q if quota_ee is None else quota_ee, margin_mode, safety_margin,
                                       horizon_scale, max_horizon, int(brute))

    def active_pairs(self):
        """(counters [V-T evaluated, E-E evaluated, V-T dropped, E-E dropped], V-T (n, 4), E-E (n, 4))."""
        cnt = np.zeros(4, np.uint32)
        nvt = self._L.orc_sim_active_pairs(self._h, 0, None, None)
        nee = self._L.orc_sim_active_pairs(self._h, 1, None, None)
        vt = np.zeros((nvt, 4), np.uint32)
        ee = np.zeros((nee, 4), np.uint32)
        self._L.orc_sim_active_pairs(self._h, 0, _ptr(vt), _ptr(cnt))
        self._L.orc_sim_active_pairs(self._h, 1, _ptr(ee), None)
        return cnt, vt, ee

    def set_wrinkle(self, on, params22, texture_bytes=None, width=0, height=0):
        """The wrinkle field: on, its 22 parameters (origin + influence radius, axis + bone radius,
        normal + stiffness, binormal + blend, z_min, z_max, r_min, r_max, valley / crest window), the
        RGBA8 texture."""
        tex = None if not texture_bytes else np.frombuffer(bytes(texture_bytes), np.uint8).copy()
        self._wtex = tex
        self._L.orc_sim_set_wrinkle(self._h, int(on), np.ascontiguousarray(params22, np.float32), _ptr(tex),
                                    width, height)

    def set_edge_collision(self, on=True, margin_scale=1.0, margin_offset=0.0):
        self._L.orc_sim_set_edge_collision(self._h, int(on), margin_scale, margin_offset)

    def collider_edges(self):
        """The unique collider edges: (n, 11) [p0, thickness, p1, friction, normal] and flags."""
        n = self._L.orc_sim_collider_edges(self._h, None, None)
        e = np.zeros((n, 11), np.float32)
        fl = np.zeros(n, np.uint32)
        self._L.orc_sim_collider_edges(self._h, _ptr(e), _ptr(fl))
        return e, fl

    def set_solver_mode(self, mode):
        """0: colours; 1: fixed-point Jacobi distance pass (distance_atomic.wgsl)."""
        self._L.orc_sim_set_solver_mode(self._h, int(mode))

    def set_enable_self_collision(self, on):
        self._L.orc_sim_set_self_collision(self._h, 1 if on else 0)

    def set_self_collision_options(self, relief_factor, max_displacement_ratio, exclude_neighbors=True,
                                   enable_normal_untangling=True, max_iterations=256):
        self._L.orc_sim_set_self_collision_options(self._h, relief_factor, max_displacement_ratio,
                                                   1 if exclude_neighbors else 0, 1 if enable_normal_untangling else 0,
                                                   max_iterations)

    def set_hash_cell_size(self, cell):
        self._L.orc_sim_set_hash_cell(self._h, cell)

    def set_compliances_raw(self, stretch, shear, bend):
        self._L.orc_sim_set_compliances_raw(self._h, stretch, shear, bend)

    def add_sphere_collider(self, center, radius, friction=0.0, restitution=0.0):
        self._L.orc_sim_add_sphere(self._h, np.asarray(center, np.float32), radius, friction, restitution)

    def add_capsule_collider(self, a, b, radius, friction=0.0, restitution=0.0):
        self._L.orc_sim_add_capsule(self._h, np.asarray(a, np.float32), np.asarray(b, np.float32), radius,
                                    friction, restitution)

    def add_plane_collider(self, point, normal, friction=0.0, restitution=0.0):
        self._L.orc_sim_add_plane(self._h, np.asarray(point, np.float32), np.asarray(normal, np.float32),
                                  friction, restitution)

    def clear_colliders(self):
        self._L.orc_sim_clear_colliders(self._h)

    def set_mesh_collider_triangles(self, tris, friction=0.0, thickness=0.0, restitution=0.0, single_sided=False,
                                    attributes=None):
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        at = _opt(attributes, np.float32)
        self._L.orc_sim_set_mesh_triangles(self._h, t, len(t), _ptr(at), friction, thickness, restitution,
                                           1 if single_sided else 0)

    def set_bone_sdf_colliders(self, width, height, depth, texture_bytes, bone_infos):
        tex = np.frombuffer(bytes(texture_bytes), dtype=np.uint32).copy()
        infos = np.ascontiguousarray(np.asarray(bone_infos, np.float32)[:, :20])
        self._L.orc_sim_set_sdf(self._h, width, height, depth, tex, infos, len(infos))

    def update_bone_transforms(self, world_matrices, inv_world_matrices, dirty_bone_indices=None):
        """lib.rs:2287-2330: transforms, the bounding spheres in the infos, and the dirty bones."""
        w = np.ascontiguousarray(world_matrices, np.float32)
        iw = np.ascontiguousarray(inv_world_matrices, np.float32)
        self._L.orc_sim_set_bone_transforms(self._h, w, iw, len(w))
        self.set_dynamic_bone_sdf_dirty_bones(dirty_bone_indices)

    def clear_bone_sdf_colliders(self):
        self._L.orc_sim_clear_bone_sdf(self._h)

    def vertex_normals(self):
        """compute_normals.wgsl at the current positions: (nv, 3)."""
        out = np.zeros((len(self.state(0).reshape(-1, 3)), 3), np.float32)
        self._L.orc_sim_normals(self._h, _ptr(out))
        return out

    def setup_dynamic_bone_sdf(self, width, height, depth, res, bone_infos, rest_verts, bone_indices, bone_weights,
                               tri_sources, tri_weights, bone_bind_inv_matrices, update_interval=1):
        """setup_dynamic_bone_sdf (lib.rs:2362-2588); raises ValueError where the build rejects the input."""
        infos = np.ascontiguousarray(np.asarray(bone_infos, np.float32)[:, :20])
        rv = np.ascontiguousarray(np.asarray(rest_verts, np.float32)[:, :3])
        bi = np.ascontiguousarray(np.asarray(bone_indices, np.uint32)[:, :4])
        bw = np.ascontiguousarray(np.asarray(bone_weights, np.float32)[:, :4])
        ts = np.ascontiguousarray(np.asarray(tri_sources, np.uint32).reshape(-1, 4)[:, :4])
        tw = np.ascontiguousarray(np.asarray(tri_weights, np.float32).reshape(-1, 3)[:, :3])
        inv = np.ascontiguousarray(bone_bind_inv_matrices, np.float32)
        r = self._L.orc_sim_setup_dynamic_bone_sdf(self._h, width, height, depth, res, _ptr(infos), len(infos),
                                                   _ptr(rv), len(rv), _ptr(bi), _ptr(bw), _ptr(ts), _ptr(tw), len(ts),
                                                   _ptr(inv), len(inv), update_interval)
        if r:
            raise ValueError("setup_dynamic_bone_sdf: input rejected")

    def set_dynamic_bone_sdf_dirty_bones(self, dirty_bone_indices=None):
        if dirty_bone_indices is None:
            self._L.orc_sim_set_dynamic_dirty(self._h, None, 0, 0)
        else:
            d = np.ascontiguousarray(np.asarray(dirty_bone_indices, np.uint32).reshape(-1))
            self._L.orc_sim_set_dynamic_dirty(self._h, _ptr(d), len(d), 1)

    def set_dynamic_bone_sdf_enabled(self, enabled):
        self._L.orc_sim_set_dynamic_enabled(self._h, 1 if enabled else 0)

    def set_dynamic_bone_sdf_update_interval(self, interval):
        self._L.orc_sim_set_dynamic_interval(self._h, int(interval))

    def dynamic_arrays(self, nv, nt):
        """(skinned positions (nv, 3), bake triangles (nt, 16)) of the last dynamic update."""
        sk = np.zeros((nv, 3), np.float32)
        tr = np.zeros((nt, 16), np.float32)
        self._L.orc_sim_dynamic_arrays(self._h, _ptr(sk), _ptr(tr))
        return sk, tr

    def sdf_arrays(self, nvox, nb):
        """(texels (nvox,) u32, infos (nb, 20)) of the SDF colliders."""
        tex = np.zeros(nvox, np.uint32)
        info = np.zeros((nb, 20), np.float32)
        self._L.orc_sim_sdf_arrays(self._h, _ptr(tex), _ptr(info))
        return tex, info

    def reset(self):
        """dispatch.rs:1411-1427: initial state and inverse masses, no pins; the base inverse masses
        of set_vertex_inv_masses stay."""
        self._L.orc_sim_reset(self._h)
        self._pins.clear()
        self._upload_pins()

    # --- dynamic pins: the host policy of simulation/mod.rs:2768-2905 -----------------------------
    @staticmethod
    def _clamped(w):
        w = np.float32(w)
        return np.float32(min(max(w, np.float32(0)), np.float32(1))) if np.isfinite(w) else np.float32(0)

    def inv_masses(self, which=0) -> np.ndarray:
        """which 0: the working inverse masses (what the solver reads), 1: the mesh's at creation."""
        out = np.empty(self.n, np.float32)
        self._L.orc_sim_get_inv_masses(self._h, which, out)
        return out

    def _upload_pins(self):
        keys = sorted(self._pins)
        v = np.array(keys, np.uint32)
        t = np.array([self._pins[k][0] for k in keys], np.float32).reshape(-1, 3)
        w = np.array([self._pins[k][1] for k in keys], np.float32)
        self._L.orc_sim_set_pin_table(self._h, _ptr(v), _ptr(t), _ptr(w), len(keys))

    def _write_inv(self, idx, vals):
        i = np.asarray(idx, np.uint32)
        v = np.asarray(vals, np.float32)
        self._L.orc_sim_set_inv_masses(self._h, _ptr(i), _ptr(v), len(i))

    def _effective(self, i):
        if i in self._pins:
            return np.float32(self._inv_orig[i] * (np.float32(1) - self._clamped(self._pins[i][1])))
        return self._inv_orig[i]

    def set_pin(self, vertex_idx, target_pos, weight=1.0):
        self.set_pins_batch([vertex_idx], [target_pos], [weight])

    def set_pins_batch(self, indices, targets, weights):
        n = min(len(indices), len(targets), len(weights))
        idx, vals = [], []
        for k in range(n):
            i = int(indices[k])
            if i >= self.n:
                continue
            self._pins[i] = (np.asarray(targets[k], np.float32).reshape(3), np.float32(weights[k]))
            idx.append(i)
            vals.append(np.float32(self._inv_orig[i] * (np.float32(1) - self._clamped(weights[k]))))
        self._upload_pins()
        self._write_inv(idx, vals)

    def release_pin(self, vertex_idx):
        self.release_pins_batch([vertex_idx])

    def release_pins_batch(self, indices):
        idx = [int(i) for i in indices if self._pins.pop(int(i), None) is not None]
        if idx:
            self._upload_pins()
            self._write_inv(idx, [self._inv_orig[i] for i in idx])

    def clear_pins(self):
        idx = list(self._pins)
        self._pins.clear()
        self._upload_pins()
        self._write_inv(idx, [self._inv_orig[i] for i in idx])

    def set_vertex_inv_masses(self, indices, inv_masses):
        idx, vals = [], []
        for i, m in zip(indices, inv_masses):
            i = int(i)
            if i >= self.n:
                continue
            self._inv_orig[i] = np.float32(m)
            idx.append(i)
            vals.append(self._effective(i))
        self._write_inv(idx, vals)

    def get_vertex_inv_masses(self) -> np.ndarray:
        return self._inv_orig.copy()

    def set_positions_and_velocities(self, positions, velocities=None):
        """dispatch.rs:952-986: x = p = positions, v = velocities (or zero), inverse masses from the
        base values and the pins."""
        pos = np.asarray(positions, np.float64).reshape(-1, 3)
        if len(pos) != self.n:
            return
        vel = np.zeros_like(pos) if velocities is None else np.asarray(velocities, np.float64).reshape(-1, 3)
        self.set_state(x=pos, p=pos, v=vel)
        self._write_inv(range(self.n), [self._effective(i) for i in range(self.n)])

    def step(self, dt=1.0 / 60.0, substeps=None, solver_iterations=None):
        if solver_iterations is not None:
            self.set_solver_iterations(solver_iterations)
        return self._L.orc_sim_step(self._h, dt, 0 if substeps is None else substeps)

    def step_single_substep(self, dt_sub):
        self._L.orc_sim_step_single_substep(self._h, dt_sub)

    def substep_parts(self, dt_sub, collide=True, self_collide=True, update=True):
        self._L.orc_sim_substep_parts(self._h, dt_sub, int(collide), int(self_collide), int(update))

    def residual(self, dt_sub):
        a, b = C.c_double(), C.c_double()
        self._L.orc_sim_residual(self._h, dt_sub, C.byref(a), C.byref(b))
        return a.value, b.value

    def state(self, which=0) -> np.ndarray:
        out = np.empty(self.n * 3, np.float64)
        self._L.orc_sim_get(self._h, which, out)
        return out.reshape(-1, 3)

    def set_state(self, x=None, p=None, v=None):
        cv = [None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1) for a in (x, p, v)]
        self._L.orc_sim_set_state(self._h, *[_ptr(a) for a in cv])

    def lambdas(self):
        m = self.mesh()
        ld, lb = np.empty(m["nd"], np.float64), np.empty(m["nb"], np.float64)
        self._L.orc_sim_get_lambdas(self._h, ld, lb)
        return ld, lb

    def set_lambdas(self, ld=None, lb=None):
        """Overwrite the lambdas (distance in colour order, bending); a test hook."""
        a = None if ld is None else np.ascontiguousarray(ld, np.float64)
        b = None if lb is None else np.ascontiguousarray(lb, np.float64)
        self._L.orc_sim_set_lambdas(self._h, _ptr(a), _ptr(b))

    def contact_stats(self, clear=True) -> dict[str, int]:
        """Contacts resolved since the last call: vv, vt_ccd, vt_prox, ee_ccd, ee_prox."""
        out = np.zeros(6, np.uint64)
        self._L.orc_sim_stats(self._h, out, 1 if clear else 0)
        return dict(zip(["vv", "vt_ccd", "vt_prox", "ee_ccd", "ee_prox"], (int(x) for x in out[:5])))

    def get_positions(self, out=None):
        x = self.state(0).astype(np.float32).reshape(-1)
        if out is not None:
            out[: x.size] = x
            return out
        return x


def f32_to_f16_bits(x, precision: str = "f32") -> np.ndarray:
    """binary16 bits of float32 values, round to nearest even (sdf_oracle.c)."""
    a = np.ascontiguousarray(x, np.float32).reshape(-1)
    out = np.empty(a.size, np.uint16)
    lib(precision).orc_f32_to_f16_n(a, out, a.size)
    return out


def f16_bits_to_f32(h, precision: str = "f32") -> np.ndarray:
    a = np.ascontiguousarray(h, np.uint16).reshape(-1)
    out = np.empty(a.size, np.float32)
    lib(precision).orc_f16_to_f32_n(a, out, a.size)
    return out


def close_open_boundary_loops(verts, tris, precision: str = "f32"):
    """sdf_baker.rs:244-462 restated: (verts, tris, loops closed)."""
    L = lib(precision)
    v = np.ascontiguousarray(np.asarray(verts, np.float32).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, np.int32).reshape(-1, 3))
    nv, nt = C.c_uint32(), C.c_uint32()
    L.orc_close_loops_copy(v.reshape(-1), len(v), _ptr(t), len(t), None, 0, None, 0, C.byref(nv), C.byref(nt))
    ov = np.empty((nv.value, 3), np.float32)
    ot = np.empty((nt.value, 3), np.int32)
    loops = L.orc_close_loops_copy(v.reshape(-1), len(v), _ptr(t), len(t), _ptr(ov), ov.size, _ptr(ot), ot.size,
                                   C.byref(nv), C.byref(nt))
    return ov, ot, int(loops)


def atlas_layout(n_bones, resolution, max_dim=2048):
    """compute_3d_atlas_layout (sdf_baker.rs:137-163): (cols, rows, layers, res, width, height, depth)."""
    o = np.zeros(7, np.uint32)
    lib("f32").orc_atlas_layout(int(n_bones), int(resolution), int(max_dim), o)
    return tuple(int(x) for x in o)


def bake_bone_sdf(verts, tris, weights, binds, resolution=64, margin=0.2, weight_threshold=0.02, blend_k=0.05,
                  friction=0.5, thickness=0.005, restitution=0.0, close_open_boundaries=True,
                  precision: str = "f32"):
    """Bone SDF atlas bake restated (sdf_baker.rs:494-930, bake_sdf.wgsl).  weights (n_bones, n_verts),
    binds (n_bones, 4, 4).  Returns texels, width, height, depth, bone_infos, active (input indices),
    bind_matrices, or None where the reference bake fails (no bone above the threshold)."""
    L = lib(precision)
    v = np.ascontiguousarray(np.asarray(verts, np.float32).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, np.int32).reshape(-1, 3))
    w = np.ascontiguousarray(np.asarray(weights, np.float32).reshape(-1))
    b = np.ascontiguousarray(np.asarray(binds, np.float32).reshape(-1))
    nb = b.size // 16
    h = L.orc_bone_sdf_bake(v.reshape(-1), len(v), _ptr(t), len(t), w if w.size else np.zeros(1, np.float32),
                            b if b.size else np.zeros(1, np.float32), nb, int(resolution), margin, weight_threshold,
                            blend_k, friction, thickness, restitution, 1 if close_open_boundaries else 0)
    if not h:
        return None
    try:
        dims = np.zeros(3, np.uint32)
        na = C.c_uint32()
        L.orc_bone_sdf_info(h, dims, C.byref(na))
        n = int(na.value)
        tex = np.empty(int(dims[0]) * int(dims[1]) * int(dims[2]), np.uint32)
        infos = np.empty((n, 20), np.float32)
        act = np.empty(n, np.uint32)
        mats = np.empty((n, 4, 4), np.float32)
        L.orc_bone_sdf_arrays(h, tex, infos.reshape(-1), act, mats.reshape(-1))
        return {"texels": tex, "width": int(dims[0]), "height": int(dims[1]), "depth": int(dims[2]),
                "bone_infos": infos, "active": act, "bind_matrices": mats}
    finally:
        L.orc_bone_sdf_free(h)


def bake_mesh_sdf(verts, tris, voxel_size=0.004, margin=0.02, friction=0.5, thickness=0.005, restitution=0.0,
                  close_open_boundaries=True, method=0, precision: str = "f32"):
    """Brute-force restatement of the reference bakes (method 0 dense, 1 hierarchical, 2 coarse pass);
    returns the same dict as the product's bake functions, or None where the reference bake fails."""
    L = lib(precision)
    v = np.ascontiguousarray(np.asarray(verts, np.float32).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, np.int32).reshape(-1, 3))
    h = L.orc_sdf_bake(v.reshape(-1), len(v), _ptr(t), len(t), voxel_size, margin, friction, thickness,
                       restitution, 1 if close_open_boundaries else 0, method)
    if not h:
        return None
    try:
        dims = np.zeros(3, np.uint32)
        info = np.zeros(20, np.float32)
        cv = C.c_float()
        L.orc_sdf_info(h, dims, info, C.byref(cv))
        n = int(dims[0]) * int(dims[1]) * int(dims[2])
        tex = np.empty(n, np.uint32)
        L.orc_sdf_texels(h, tex)
        out = {"texture_bytes": tex.tobytes(), "width": int(dims[0]), "height": int(dims[1]),
               "depth": int(dims[2]), "bone_info": info, "texels": tex}
        if method == 2:
            near = np.empty(n * 8, np.uint32)
            L.orc_sdf_nearest(h, near)
            out["nearest_tri"] = near
            out["coarse_voxel"] = float(cv.value)
        return out
    finally:
        L.orc_sdf_free(h)
