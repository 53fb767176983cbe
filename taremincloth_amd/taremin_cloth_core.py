"""Drop-in mirror of the reference's PyO3 module ``taremin_cloth_core`` for the substep path.

Same class and method names, argument meaning and error behaviour as
crates/cloth_py/src/lib.rs (ClothSimulator at :1549-3258, module functions at
:636-648, :552 and :3539-3580); the work runs in libtcloth_hip.so on the GPU.
Shape and index errors raise ``ValueError`` before any GPU work, device
failures raise ``RuntimeError`` (lib.rs:1590-1718).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from . import sim_config as _sc

__all__ = ["ClothSimulator", "color_edge_pairs", "compute_combined_inv_masses", "bake_mesh_sdf_gpu",
           "bake_mesh_sdf_hierarchical_gpu"]


def _f32(a, shape_tail=None, name="array"):
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape_tail is not None:
        if arr.ndim != 1 + len(shape_tail) or tuple(arr.shape[1:]) != tuple(shape_tail):
            raise ValueError(f"{name} must have shape (N, {', '.join(map(str, shape_tail))})")
    return arr


def _u32(a, cols, name):
    arr = np.asarray(a)
    if arr.ndim != 2 or arr.shape[1] != cols:
        raise ValueError(f"{name} must have shape (N, {cols})")
    if arr.size and np.issubdtype(arr.dtype, np.signedinteger) and (arr < 0).any():
        raise ValueError(f"{name} contains negative indices")
    return np.ascontiguousarray(arr, dtype=np.uint32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _u32_arg(v):
    """An unsigned 32-bit argument, rejected with OverflowError outside that range as PyO3 does."""
    i = int(v)
    if not 0 <= i < 2 ** 32:
        raise OverflowError("argument must fit an unsigned 32-bit integer")
    return i


def _vec3(v, name):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1))
    if a.size != 3:
        raise ValueError(f"{name} must have 3 components")
    return a


class ClothSimulator:
    """GPU XPBD cloth simulator (graph-coloured distance/shear/bending, colliders, self-collision)."""

    def __init__(self, positions, edges, faces=None, inv_masses=None, sewing_springs=None, layer_ids=None,
                 thicknesses=None, layer_id=0, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0,
                 sewing_shrink_speed=1.0, sewing_stiffness=None, enable_sewing_lock=None, sewing_lock_distance=None,
                 compression_stiffness=None, shear_stiffness=None, workgroup_size=32, solver_mode=0,
                 enable_compact_readback=None, enable_pair_cache=None, areal_density=None,
                 enable_coarse_constraints=None, self_collision_algorithm=None):
        L = _lib.load()
        self._L = L
        self._h = None
        pos = _f32(positions, (3,), "positions")
        e = _u32(edges, 2, "edges")
        n = len(pos)
        if e.size and int(e.max()) >= n:
            bad = e[(e >= n).any(axis=1)][0]
            raise ValueError(f"edges contain an index >= number of vertices ({n}): [{bad[0]}, {bad[1]}]")
        f = None
        if faces is not None:
            f = _u32(faces, 3, "faces")
            if f.size and int(f.max()) >= n:
                raise ValueError(f"faces contain an index >= number of vertices ({n})")
            if f.size == 0:
                f = None
        sew = None
        if sewing_springs is not None:  # lib.rs:1643-1665
            sa = np.asarray(sewing_springs)
            if sa.size:
                if sa.ndim != 2 or sa.shape[1] != 2:
                    raise ValueError("each pair of sewing_springs must hold two vertices (v0, v1)")
                sew = np.ascontiguousarray(sa.astype(np.uint32))
                bad = (sew >= n).any(axis=1)
                if bad.any():
                    b = sew[bad][0]
                    raise ValueError(f"sewing_springs contain an index >= number of vertices ({n}): [{b[0]}, {b[1]}]")
        if enable_coarse_constraints:
            raise NotImplementedError("coarse (two-hop) constraints are not part of this build")
        if self_collision_algorithm == 2:
            raise NotImplementedError("self_collision_algorithm=2 (virtual-vertex V-V) is not part of this build")
        solver_mode = int(solver_mode)
        if not 0 <= solver_mode < 2 ** 32:
            raise OverflowError("solver_mode must fit an unsigned 32-bit integer")

        def vec(a, dtype, name):
            if a is None:
                return None
            v = np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1))
            if len(v) != n:
                raise ValueError(f"{name} length ({len(v)}) does not match the number of vertices ({n})")
            return v

        im = vec(inv_masses, np.float32, "inv_masses")
        ly = vec(layer_ids, np.uint32, "layer_ids")
        th = vec(thicknesses, np.float32, "thicknesses")
        p = _lib.TcParams()
        L.tc_sim_params_default(C.byref(p))
        p.thickness = float(thickness)
        p.layer_id = int(layer_id)
        p.stiffness = float(stiffness)
        p.bending_stiffness = float(bending_stiffness)
        p.compression_stiffness = -1.0 if compression_stiffness is None else float(compression_stiffness)
        p.shear_stiffness = -1.0 if shear_stiffness is None else float(shear_stiffness)
        # The reference's workgroup_size selects 32/64-wide WGSL groups; wave64 hardware
        # uses its own tuned block size, so the value is accepted and not forwarded.
        p.workgroup_size = 0
        p.solver_mode = solver_mode
        sp = _lib.TcSewingParams()
        L.tc_sewing_params_default(C.byref(sp))
        sp.shrink_speed = float(sewing_shrink_speed)
        sp.stiffness = -1.0 if sewing_stiffness is None else float(sewing_stiffness)
        sp.lock = -1 if enable_sewing_lock is None else (1 if enable_sewing_lock else 0)
        sp.lock_distance = -1.0 if sewing_lock_distance is None else float(sewing_lock_distance)
        h = C.c_void_p()
        _lib.check(L.tc_sim_create_sewn(_ptr(pos), n, _ptr(e), len(e), _ptr(f), 0 if f is None else len(f), _ptr(im),
                                        _ptr(th), _ptr(ly), _ptr(sew), 0 if sew is None else len(sew), C.byref(sp),
                                        C.byref(p), C.byref(h)))
        self._h = h
        self._n = n
        self._keep = (pos, e, f)
        # set_areal_density only records the value (simulation/mod.rs:1241-1245); the substep
        # uses the inverse masses given at construction.
        self.areal_density = 0.15 if areal_density is None else float(min(max(areal_density, 0.0), 100.0))
        self._init_config(stiffness, compression_stiffness, shear_stiffness, bending_stiffness, workgroup_size)
        self._cfg["solver_mode"] = solver_mode
        self._sync_pair_cfg()
        if enable_pair_cache is not None:  # lib.rs:1755-1760
            self.set_enable_pair_cache(enable_pair_cache)
        if self_collision_algorithm is not None:
            self.set_self_collision_algorithm(self_collision_algorithm)
        if sewing_lock_distance is not None:
            self._cfg["sewing_lock_distance"] = float(max(sewing_lock_distance, 0.0))
        self._cfg["sewing_shrink_speed"] = float(max(sewing_shrink_speed, 0.0))
        self._ctor = dict(positions=pos, edges=e, faces=f, inv_masses=im, thicknesses=th, layer_ids=ly,
                          sewing_springs=sew, sewing_shrink_speed=float(sewing_shrink_speed),
                          sewing_stiffness=sewing_stiffness, enable_sewing_lock=enable_sewing_lock,
                          sewing_lock_distance=sewing_lock_distance,
                          layer_id=int(layer_id), thickness=float(thickness), stiffness=float(stiffness),
                          bending_stiffness=float(bending_stiffness), compression_stiffness=compression_stiffness,
                          shear_stiffness=shear_stiffness)
        self._colliders = []
        self._recorder = None

    def _init_config(self, stiffness, compression_stiffness, shear_stiffness, bending_stiffness, workgroup_size):
        """The simulator's parameters as SimConfig fields, with the runtime defaults of
        simulation/mod.rs:607-691 (not SimConfig's own defaults, config.rs:156-213)."""
        cfg = _sc.defaults()
        cfg.update(damping=1.0, tension_damping=0.0, compression_damping=0.0, shear_damping=0.0,
                   bending_damping=0.0, tension_stiffness=float(stiffness),
                   compression_stiffness=float(stiffness if compression_stiffness is None else compression_stiffness),
                   shear_stiffness=float(stiffness * 0.5 if shear_stiffness is None else shear_stiffness),
                   bending_stiffness=float(bending_stiffness), solver_iterations=2, solver_mode=0,
                   workgroup_size=int(workgroup_size), enable_self_collision=True, relief_factor=0.2,
                   max_displacement_ratio=0.2, post_relaxation_iters=1, areal_density=self.areal_density)
        self._cfg = cfg

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.tc_sim_destroy(h)
            self._h = None

    # --- stepping ----------------------------------------------------------------------------
    def step(self, dt=0.016666667, substeps=None, solver_iterations=None) -> int:
        """One frame; returns the substeps it ran (the adaptive policy's choice when it is on)."""
        eff = C.c_uint32(0)
        if solver_iterations is not None:
            self.set_solver_iterations(solver_iterations)
        _lib.check(self._L.tc_sim_step(self._h, float(dt), 0 if substeps is None else int(substeps), 0,
                                       C.byref(eff)))
        if self._recorder is not None:
            self._recorder.record(self, float(dt), int(eff.value))
        return int(eff.value)

    def step_async(self, dt=0.016666667, substeps=None, solver_iterations=None):
        """Enqueues a frame and the copy of its positions without waiting (lib.rs:1823-1831)."""
        if solver_iterations is not None:
            self.set_solver_iterations(solver_iterations)
        _lib.check(self._L.tc_sim_step_async(self._h, float(dt), 0 if substeps is None else int(substeps), 0))

    def fetch_positions(self, out_array) -> bool:
        """Positions of the last step_async into out_array; False when none is pending (lib.rs:1833-1845)."""
        out = self._writable_f32(out_array)
        ok = C.c_int(0)
        _lib.check(self._L.tc_sim_fetch_positions(self._h, out.ctypes.data_as(C.c_void_p), out.size, C.byref(ok)))
        return bool(ok.value)

    def step_buffered(self, dt=0.016666667, substeps=None, solver_iterations=None) -> int:
        """A frame whose positions stay in a device ring of 8; returns the frames held (lib.rs:1847-1855)."""
        if solver_iterations is not None:
            self.set_solver_iterations(solver_iterations)
        cnt = C.c_uint32(0)
        _lib.check(self._L.tc_sim_step_buffered(self._h, float(dt), 0 if substeps is None else int(substeps), 0,
                                                C.byref(cnt)))
        return int(cnt.value)

    def fetch_buffered_positions(self, out_array) -> int:
        """All held frames (frame-major) in one copy; returns their number and empties the ring."""
        out = self._writable_f32(out_array)
        cnt = C.c_uint32(0)
        _lib.check(self._L.tc_sim_fetch_buffered_positions(self._h, out.ctypes.data_as(C.c_void_p), out.size,
                                                           C.byref(cnt)))
        return int(cnt.value)

    def get_buffered_frame_count(self) -> int:
        cnt = C.c_uint32(0)
        _lib.check(self._L.tc_sim_get_buffered_frame_count(self._h, C.byref(cnt)))
        return int(cnt.value)

    def clear_frame_buffer(self):
        _lib.check(self._L.tc_sim_clear_frame_buffer(self._h))

    @staticmethod
    def _writable_f32(a):
        if not isinstance(a, np.ndarray) or a.dtype != np.float32 or not a.flags.c_contiguous or not a.flags.writeable:
            raise ValueError("out_array must be a writable C-contiguous float32 numpy array")
        return a

    # --- adaptive substeps (lib.rs:1787-1805, 1950-1962) ----------------------------------------
    def set_adaptive_substep_options(self, enabled, min_substeps, max_substeps):
        lo = max(int(min_substeps), 1)
        hi = max(int(max_substeps), lo)
        _lib.check(self._L.tc_sim_set_adaptive_substep_options(self._h, 1 if enabled else 0, lo, hi))
        self._cfg.update(enable_adaptive_substep=bool(enabled), min_substeps=lo, max_substeps=hi)

    def set_auto_compensate_iterations(self, enabled):
        _lib.check(self._L.tc_sim_set_auto_compensate_iterations(self._h, 1 if enabled else 0))
        self._cfg["auto_compensate_iterations"] = bool(enabled)

    def set_strain_adaptive_options(self, enabled, tolerance):
        tol = float(np.float32(min(max(float(tolerance), 0.0001), 0.1)))
        _lib.check(self._L.tc_sim_set_strain_adaptive_options(self._h, 1 if enabled else 0, tol))
        self._cfg.update(enable_strain_adaptive=bool(enabled), strain_tolerance=tol)

    def set_enable_compact_readback(self, enabled):
        _lib.check(self._L.tc_sim_set_enable_compact_readback(self._h, 1 if enabled else 0))

    def note_external_displacement(self, disp):
        _lib.check(self._L.tc_sim_note_external_displacement(self._h, float(disp)))

    def _policy(self):
        out = (C.c_double * 8)()
        _lib.check(self._L.tc_sim_frame_policy_state(self._h, out))
        return list(out)

    def get_effective_substeps(self) -> int:
        return int(self._policy()[0])

    def get_effective_solver_iterations(self) -> int:
        return int(self._policy()[1])

    def get_last_step_max_strain(self) -> float:
        return float(self._policy()[2])

    # --- SimConfig (lib.rs:3028-3048, simulation/config_impl.rs) ---------------------------------
    def get_config_json(self) -> str:
        return _sc.to_json(self._cfg)

    def get_config_hash(self) -> int:
        return _sc.fnv1a64(self.get_config_json())

    def apply_config_json(self, json_str):
        """Applies every field through its setter (config_impl.rs:91-166).  Options whose solver is not
        part of this build raise NotImplementedError unless they keep their inert value."""
        c = _sc.parse(json_str)
        inert = {"coupled_mode": 0, "coupled_collider": False, "substep_interval": 1, "ee_substep_interval": 1}
        if c["self_collision_algorithm"] == 2:
            raise NotImplementedError("SimConfig self_collision_algorithm=2: that solver option is not part of this build")
        for k, v in inert.items():
            if c[k] != v:
                raise NotImplementedError(f"SimConfig {k}={c[k]!r}: that solver option is not part of this build")
        self.set_gravity(*c["gravity"])
        self.set_damping(c["damping"])
        self.set_damping_all(c["tension_damping"], c["compression_damping"], c["shear_damping"], c["bending_damping"])
        self.set_stiffness_all(c["tension_stiffness"], c["compression_stiffness"], c["shear_stiffness"],
                               c["bending_stiffness"])
        self.set_solver_iterations(c["solver_iterations"])
        self.set_solver_mode(c["solver_mode"])
        self.set_workgroup_size(c["workgroup_size"])
        self.set_enable_self_collision(c["enable_self_collision"])
        self.set_self_collision_options(c["relief_factor"], c["max_displacement_ratio"], c["exclude_neighbors"],
                                        c["enable_normal_untangling"], c["self_collision_max_iterations"])
        self.set_self_collision_algorithm(c["self_collision_algorithm"])
        self.set_enable_pair_cache(c["enable_pair_cache"])
        self.set_pair_cache_options(c["pair_max_pairs"], c["pair_max_pairs"], c["pair_margin_mode"],
                                    c["pair_safety_margin"], c["pair_horizon_scale"], c["pair_max_horizon"])
        self.set_enable_pair_cache_final_fallback(c["enable_pair_final_fallback"])
        self.set_enable_edge_collision(c["enable_edge_collision"])
        self.set_edge_margin_scale(c["edge_margin_scale"])
        self.set_edge_margin_offset(c["edge_margin_offset"])
        self.set_adaptive_substep_options(c["enable_adaptive_substep"], c["min_substeps"], c["max_substeps"])
        self.set_auto_compensate_iterations(c["auto_compensate_iterations"])
        self.set_strain_adaptive_options(c["enable_strain_adaptive"], c["strain_tolerance"])
        self.areal_density = float(min(max(c["areal_density"], 0.0), 100.0))
        self.set_sewing_stiffness(c["sewing_stiffness"])
        self.set_sewing_shrink_speed(c["sewing_shrink_speed"])
        self.set_enable_sewing_lock(c["enable_sewing_lock"])
        self.set_sewing_lock_distance(c["sewing_lock_distance"])
        self.set_sewing_priority_options(c["sewing_priority_enabled"], c["sewing_priority_threshold"],
                                         c["sewing_priority_merge_dist"], c["sewing_priority_ramp_frames"],
                                         c["sewing_priority_max_frames"])
        # recorded as given: options of solvers this build does not have, kept at their inert values
        for k in ("post_relaxation_iters", "auto_coupled_on_low_substeps", "areal_density", "version"):
            self._cfg[k] = c[k]

    # --- debug recording (lib.rs:2956-3012, debug_recorder.rs) ----------------------------------
    def start_debug_recording(self, object_name, max_frames=None):
        from .debug_recorder import DebugRecorder
        self._recorder = DebugRecorder(self, str(object_name), 3600 if max_frames is None else int(max_frames))

    def is_debug_recording(self) -> bool:
        return self._recorder is not None

    def stop_debug_recording(self):
        self._recording_kept = self._recorder
        self._recorder = None

    def save_debug_recording(self, file_path) -> str:
        rec = self._recorder if self._recorder is not None else getattr(self, "_recording_kept", None)
        if rec is None:
            raise RuntimeError("no debug recording to save")
        try:
            return rec.save(str(file_path))
        except OSError as e:
            raise OSError(str(e)) from None

    def step_single_substep(self, dt_sub=0.0016666667):
        _lib.check(self._L.tc_sim_step_single_substep(self._h, float(dt_sub)))

    def get_positions(self, out):
        if not isinstance(out, np.ndarray) or out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float32 numpy array")
        if out.size < 3 * self._n:
            raise ValueError(f"out must hold at least {3 * self._n} floats")
        _lib.check(self._L.tc_sim_get_positions(self._h, out.ctypes.data_as(C.c_void_p), out.size))

    def set_positions_and_velocities(self, positions, velocities=None):
        pos = _f32(positions).reshape(-1)
        vel = None if velocities is None else _f32(velocities).reshape(-1)
        if pos.size % 3 or (vel is not None and vel.size != pos.size):
            raise ValueError("positions / velocities must be (N, 3)")
        _lib.check(self._L.tc_sim_set_positions_velocities(self._h, _ptr(pos), _ptr(vel), pos.size // 3))

    def reset(self):
        _lib.check(self._L.tc_sim_reset(self._h))

    # --- wrinkle field (lib.rs:3049-3137) ----------------------------------------------------------
    def set_enable_wrinkle_field(self, enable):
        _lib.check(self._L.tc_sim_set_enable_wrinkle_field(self._h, 1 if enable else 0))

    def wrinkle_field_enabled(self) -> bool:
        e = C.c_int(0)
        _lib.check(self._L.tc_sim_wrinkle_field_enabled(self._h, C.byref(e)))
        return bool(e.value)

    def _wrinkle_params(self):
        out = np.zeros(22, np.float32)
        _lib.check(self._L.tc_sim_get_wrinkle_field_params(self._h, _ptr(out)))
        return out

    def set_wrinkle_field_params(self, origin, axis, normal, binormal, influence_radius=0.05, bone_radius=0.05,
                                 stiffness=1.0, blend_weight=1.0, z_min=None, z_max=None, r_min=None, r_max=None,
                                 valley_window=0.015, crest_window=0.020):
        cur = self._wrinkle_params()
        frame = np.concatenate([_vec3(origin, "origin"), [influence_radius], _vec3(axis, "axis"), [bone_radius],
                                _vec3(normal, "normal"), [stiffness], _vec3(binormal, "binormal"),
                                [blend_weight]]).astype(np.float32)
        zr = (cur[16] if z_min is None else z_min, cur[17] if z_max is None else z_max)
        rr = (cur[18] if r_min is None else r_min, cur[19] if r_max is None else r_max)
        vw = valley_window if valley_window > 0.0 else 0.015  # lib.rs:3093-3094
        cw = crest_window if crest_window > 0.0 else 0.020
        _lib.check(self._L.tc_sim_set_wrinkle_field_params(self._h, _ptr(frame), float(zr[0]), float(zr[1]),
                                                           float(rr[0]), float(rr[1]), float(vw), float(cw)))

    def set_wrinkle_field_texture_2d(self, width, height, texture_bytes, z_min, z_max, r_min, r_max):
        data = bytes(texture_bytes)
        if len(data) != int(width) * int(height) * 4:
            raise ValueError(f"texture_bytes has the wrong size: actual={len(data)}, "
                             f"expected={int(width) * int(height) * 4} (width={width}, height={height})")
        buf = np.frombuffer(data, np.uint8) if data else np.zeros(0, np.uint8)
        _lib.check(self._L.tc_sim_set_wrinkle_field_texture_2d(self._h, int(width), int(height), _ptr(buf),
                                                               len(data), float(z_min), float(z_max),
                                                               float(r_min), float(r_max)))

    # --- collider-edge collision (lib.rs:2735-2749) ------------------------------------------------
    def set_enable_edge_collision(self, enable):
        _lib.check(self._L.tc_sim_set_enable_edge_collision(self._h, 1 if enable else 0))
        self._cfg["enable_edge_collision"] = bool(enable)

    def _edge_margin(self, scale, offset):
        sc, off = float(np.float32(max(scale, 1.0))), float(np.float32(max(offset, 0.0)))
        _lib.check(self._L.tc_sim_set_edge_margin(self._h, sc, off))
        self._cfg["edge_margin_scale"], self._cfg["edge_margin_offset"] = sc, off

    def set_edge_margin_scale(self, scale):
        self._edge_margin(float(scale), self._cfg["edge_margin_offset"])

    def set_edge_margin_offset(self, offset):
        self._edge_margin(self._cfg["edge_margin_scale"], float(offset))

    def collider_edge_count(self) -> int:
        n = C.c_uint32(0)
        _lib.check(self._L.tc_sim_collider_edge_count(self._h, C.byref(n)))
        return int(n.value)

    def collider_edges(self):
        """The unique collider edges: (n, 11) [p0, thickness, p1, friction, normal] and their flags."""
        n = self.collider_edge_count()
        buf = np.zeros((n, 12), np.float32)
        if n:
            _lib.check(self._L.tc_sim_get_collider_edges(self._h, _ptr(buf), n))
        return buf[:, :11].copy(), buf[:, 11].view(np.uint32).copy()

    # --- pair-cache self-collision (lib.rs:2029-2107) ----------------------------------------------
    def _pair_state(self):
        out = (C.c_double * 9)()
        _lib.check(self._L.tc_sim_pair_cache_state(self._h, out))
        return list(out)

    def _sync_pair_cfg(self):
        st = self._pair_state()
        e = C.c_int(0)
        _lib.check(self._L.tc_sim_get_enable_pair_cache(self._h, C.byref(e)))
        a = C.c_uint32(0)
        _lib.check(self._L.tc_sim_get_self_collision_algorithm(self._h, C.byref(a)))
        self._cfg.update(enable_pair_cache=bool(e.value), self_collision_algorithm=int(a.value),
                         pair_max_pairs=int(max(st[0], st[1])), pair_margin_mode=int(st[4]),
                         pair_safety_margin=float(np.float32(st[5])), pair_horizon_scale=float(np.float32(st[6])),
                         pair_max_horizon=float(np.float32(st[7])), enable_pair_final_fallback=bool(st[8]))

    def set_enable_pair_cache(self, enable):
        _lib.check(self._L.tc_sim_set_enable_pair_cache(self._h, 1 if enable else 0))
        self._sync_pair_cfg()

    def get_enable_pair_cache(self) -> bool:
        e = C.c_int(0)
        _lib.check(self._L.tc_sim_get_enable_pair_cache(self._h, C.byref(e)))
        return bool(e.value)

    def set_self_collision_algorithm(self, algo):
        """0: DIRECT, 1: PAIR_CACHE; 2 (virtual-vertex V-V) raises NotImplementedError."""
        algo = int(algo)
        if not 0 <= algo < 2 ** 32:
            raise OverflowError("algo must fit an unsigned 32-bit integer")
        if algo == 2:
            raise NotImplementedError("self_collision_algorithm=2 (virtual-vertex V-V) is not part of this build")
        _lib.check(self._L.tc_sim_set_self_collision_algorithm(self._h, algo))
        self._sync_pair_cfg()

    def get_self_collision_algorithm(self) -> int:
        a = C.c_uint32(0)
        _lib.check(self._L.tc_sim_get_self_collision_algorithm(self._h, C.byref(a)))
        return int(a.value)

    def set_pair_cache_options(self, max_vt_pairs=65536, max_ee_pairs=65536, margin_mode=1, safety_margin=0.005,
                               horizon_scale=1.3, max_horizon=0.02):
        _lib.check(self._L.tc_sim_set_pair_cache_options(self._h, int(max_vt_pairs), int(max_ee_pairs),
                                                         int(margin_mode), float(safety_margin), float(horizon_scale),
                                                         float(max_horizon)))
        self._sync_pair_cfg()

    def set_enable_pair_cache_final_fallback(self, enable):
        _lib.check(self._L.tc_sim_set_enable_pair_cache_final_fallback(self._h, 1 if enable else 0))
        self._sync_pair_cfg()

    def get_enable_pair_cache_final_fallback(self) -> bool:
        return bool(self._pair_state()[8])

    def get_pair_cache_stats(self):
        """(V-T evaluated, E-E evaluated, V-T budget, E-E budget), dispatch.rs:1111-1119."""
        cnt = np.zeros(4, np.uint32)
        a, b = C.c_size_t(0), C.c_size_t(0)
        _lib.check(self._L.tc_sim_get_active_pairs(self._h, _ptr(cnt), None, 0, C.byref(a), None, 0, C.byref(b)))
        st = self._pair_state()
        return int(cnt[0]), int(cnt[1]), int(st[0]), int(st[1])

    def get_active_pairs(self):
        """(vt_count, ee_count, V-T pairs, E-E pairs) as lists of 4 indices, slot order
        (dispatch.rs:1255-1291)."""
        cnt = np.zeros(4, np.uint32)
        a, b = C.c_size_t(0), C.c_size_t(0)
        _lib.check(self._L.tc_sim_get_active_pairs(self._h, _ptr(cnt), None, 0, C.byref(a), None, 0, C.byref(b)))
        vt = np.zeros((a.value, 4), np.uint32)
        ee = np.zeros((b.value, 4), np.uint32)
        _lib.check(self._L.tc_sim_get_active_pairs(self._h, _ptr(cnt), _ptr(vt), a.value, C.byref(a), _ptr(ee),
                                                   b.value, C.byref(b)))
        return int(cnt[0]), int(cnt[1]), vt.tolist(), ee.tolist()

    def active_pairs_arrays(self):
        """(counters [V-T evaluated, E-E evaluated, V-T dropped, E-E dropped], V-T (n, 4), E-E (n, 4))."""
        cnt = np.zeros(4, np.uint32)
        a, b = C.c_size_t(0), C.c_size_t(0)
        _lib.check(self._L.tc_sim_get_active_pairs(self._h, _ptr(cnt), None, 0, C.byref(a), None, 0, C.byref(b)))
        vt = np.zeros((a.value, 4), np.uint32)
        ee = np.zeros((b.value, 4), np.uint32)
        _lib.check(self._L.tc_sim_get_active_pairs(self._h, _ptr(cnt), _ptr(vt), a.value, C.byref(a), _ptr(ee),
                                                   b.value, C.byref(b)))
        return cnt, vt, ee

    def set_solver_mode(self, mode):
        """0: graph colouring; 1: fixed-point Jacobi for the distance constraints (lib.rs:1887-1890)."""
        mode = int(mode)
        if not 0 <= mode < 2 ** 32:
            raise OverflowError("mode must fit an unsigned 32-bit integer")
        _lib.check(self._L.tc_sim_set_solver_mode(self._h, mode))
        self._cfg["solver_mode"] = mode

    def get_solver_mode(self) -> int:
        m = C.c_uint32(0)
        _lib.check(self._L.tc_sim_get_solver_mode(self._h, C.byref(m)))
        return int(m.value)

    # --- dynamic pins and inverse masses (lib.rs:2616-2676) ----------------------------------------
    def set_pin(self, vertex_idx, target_pos, weight=1.0):
        idx = np.array([int(vertex_idx)], np.uint32) if 0 <= int(vertex_idx) < 2 ** 32 else None
        if idx is None:
            raise OverflowError("vertex_idx must fit an unsigned 32-bit integer")
        t = _vec3(target_pos, "target_pos")
        w = np.array([weight], np.float32)
        _lib.check(self._L.tc_sim_set_pins(self._h, _ptr(idx), _ptr(t), _ptr(w), 1))

    def set_pins_batch(self, indices, targets, weights):
        """lib.rs:2637-2657: the shortest of the three lengths counts; targets is (N, 3)."""
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.int64).reshape(-1))
        if idx.size and (idx.min() < 0 or idx.max() >= 2 ** 32):
            raise OverflowError("indices must fit unsigned 32-bit integers")
        t = np.asarray(targets, dtype=np.float32)
        if t.ndim != 2:
            raise ValueError("targets must be a 2-D array")
        if t.shape[0] and t.shape[1] != 3:
            raise ValueError("each target must have 3 components")
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float32).reshape(-1))
        n = min(idx.size, t.shape[0], w.size)
        idx32 = np.ascontiguousarray(idx[:n].astype(np.uint32))
        t = np.ascontiguousarray(t[:n])
        _lib.check(self._L.tc_sim_set_pins(self._h, _ptr(idx32), _ptr(t), _ptr(w[:n].copy()), n))

    def release_pin(self, vertex_idx):
        self.release_pins_batch([vertex_idx])

    def release_pins_batch(self, indices):
        idx = np.asarray(indices, dtype=np.int64).reshape(-1)
        if idx.size and (idx.min() < 0 or idx.max() >= 2 ** 32):
            raise OverflowError("indices must fit unsigned 32-bit integers")
        idx32 = np.ascontiguousarray(idx.astype(np.uint32))
        _lib.check(self._L.tc_sim_release_pins(self._h, _ptr(idx32), idx32.size))

    def clear_pins(self):
        _lib.check(self._L.tc_sim_clear_pins(self._h))

    def get_pins(self):
        """The pins as (indices uint32, targets (N, 3) float32, weights float32), ascending index."""
        n = C.c_uint32(0)
        _lib.check(self._L.tc_sim_get_pins(self._h, None, None, None, 0, C.byref(n)))
        idx = np.empty(n.value, np.uint32)
        t = np.empty((n.value, 3), np.float32)
        w = np.empty(n.value, np.float32)
        _lib.check(self._L.tc_sim_get_pins(self._h, _ptr(idx), _ptr(t), _ptr(w), n.value, C.byref(n)))
        return idx, t, w

    def set_vertex_inv_masses(self, indices, inv_masses):
        idx = np.asarray(indices)
        inv = np.asarray(inv_masses)
        if idx.dtype != np.uint32 or inv.dtype != np.float32 or idx.ndim != 1 or inv.ndim != 1:
            raise TypeError("indices must be a 1-D uint32 array and inv_masses a 1-D float32 array")
        if not idx.flags.c_contiguous or not inv.flags.c_contiguous:
            raise TypeError("indices and inv_masses must be contiguous")
        n = min(idx.size, inv.size)
        _lib.check(self._L.tc_sim_set_vertex_inv_masses(self._h, _ptr(idx), _ptr(inv), n))

    def get_vertex_inv_masses(self) -> np.ndarray:
        out = np.empty(self._n, np.float32)
        _lib.check(self._L.tc_sim_get_vertex_inv_masses(self._h, _ptr(out), out.size))
        return out

    # --- colliders -----------------------------------------------------------------------------
    def add_sphere_collider(self, center, radius, friction=0.0, restitution=0.0):
        c = _vec3(center, "center")
        _lib.check(self._L.tc_sim_add_sphere(self._h, _ptr(c), float(radius), float(friction), float(restitution)))
        self._colliders.append({"type": "sphere", "center": c.tolist(), "radius": float(radius),
                                "friction": float(friction), "restitution": float(restitution)})

    def add_capsule_collider(self, a, b, radius, friction=0.0, restitution=0.0):
        pa, pb = _vec3(a, "a"), _vec3(b, "b")
        _lib.check(self._L.tc_sim_add_capsule(self._h, _ptr(pa), _ptr(pb), float(radius), float(friction),
                                              float(restitution)))
        self._colliders.append({"type": "capsule", "point_a": pa.tolist(), "point_b": pb.tolist(),
                                "radius": float(radius), "friction": float(friction), "restitution": float(restitution)})

    def add_plane_collider(self, point, normal, friction=0.0, restitution=0.0):
        pt, nm = _vec3(point, "point"), _vec3(normal, "normal")
        _lib.check(self._L.tc_sim_add_plane(self._h, _ptr(pt), _ptr(nm), float(friction), float(restitution)))
        self._colliders.append({"type": "plane", "point": pt.tolist(), "normal": nm.tolist(),
                                "friction": float(friction), "restitution": float(restitution)})

    def set_mesh_collider_triangles(self, tris, friction=0.0, thickness=0.0, restitution=0.0, single_sided=False,
                                    attributes=None):
        t = _f32(tris)
        if t.ndim != 3 or t.shape[1:] != (3, 3):
            raise ValueError("tris must have shape (T, 3, 3)")
        at = None
        if attributes is not None:
            at = _f32(attributes)
            if at.shape != (len(t), 4):
                raise ValueError("attributes must have shape (T, 4)")
        _lib.check(self._L.tc_sim_set_mesh_triangles(self._h, _ptr(t), len(t), _ptr(at), float(friction),
                                                     float(thickness), float(restitution), 1 if single_sided else 0))
        self._colliders = [c for c in self._colliders if c["type"] != "mesh"]
        if len(t):
            self._colliders.append({"type": "mesh", "triangles": t.reshape(-1, 9).tolist(), "friction": float(friction),
                                    "thickness": float(thickness), "restitution": float(restitution),
                                    "single_sided": bool(single_sided)})

    def clear_colliders(self):
        _lib.check(self._L.tc_sim_clear_colliders(self._h))
        self._colliders = []

    # --- parameters ----------------------------------------------------------------------------
    def set_gravity(self, gx, gy, gz):
        _lib.check(self._L.tc_sim_set_gravity(self._h, float(gx), float(gy), float(gz)))
        self._cfg["gravity"] = [float(np.float32(gx)), float(np.float32(gy)), float(np.float32(gz))]

    def set_damping(self, damping):
        _lib.check(self._L.tc_sim_set_damping(self._h, float(damping)))
        self._cfg["damping"] = float(np.float32(max(float(damping), 0.0)))

    def set_damping_all(self, tension, compression, shear, bending):
        _lib.check(self._L.tc_sim_set_damping_all(self._h, float(tension), float(compression), float(shear),
                                                  float(bending)))
        for k, v in zip(("tension_damping", "compression_damping", "shear_damping", "bending_damping"),
                        (tension, compression, shear, bending)):
            self._cfg[k] = float(np.float32(max(float(v), 0.0)))

    def set_solver_iterations(self, iterations):
        _lib.check(self._L.tc_sim_set_solver_iterations(self._h, int(iterations)))
        self._cfg["solver_iterations"] = int(iterations)

    def set_enable_self_collision(self, enable):
        _lib.check(self._L.tc_sim_set_enable_self_collision(self._h, 1 if enable else 0))
        self._cfg["enable_self_collision"] = bool(enable)

    def set_self_collision_options(self, relief_factor, max_displacement_ratio, exclude_neighbors=True,
                                   enable_normal_untangling=True, max_iterations=256):
        _lib.check(self._L.tc_sim_set_self_collision_options(self._h, float(relief_factor),
                                                             float(max_displacement_ratio),
                                                             1 if exclude_neighbors else 0,
                                                             1 if enable_normal_untangling else 0,
                                                             int(max_iterations)))
        self._cfg.update(relief_factor=float(np.float32(relief_factor)),
                         max_displacement_ratio=float(np.float32(max_displacement_ratio)),
                         exclude_neighbors=bool(exclude_neighbors),
                         enable_normal_untangling=bool(enable_normal_untangling),
                         self_collision_max_iterations=int(max_iterations))

    def set_stiffness_all(self, tension, compression, shear, bending):
        _lib.check(self._L.tc_sim_set_stiffness_all(self._h, float(tension), float(compression), float(shear),
                                                    float(bending)))
        for k, v in zip(("tension_stiffness", "compression_stiffness", "shear_stiffness", "bending_stiffness"),
                        (tension, compression, shear, bending)):
            self._cfg[k] = float(np.float32(v))

    def set_workgroup_size(self, size):
        """Accepted for API compatibility; the block size is tuned per kernel on gfx950."""
        self._cfg["workgroup_size"] = int(size)

    def set_areal_density(self, density):
        """Recorded only (simulation/mod.rs:1241-1245); the masses are those given at construction."""
        self.areal_density = float(min(max(float(density), 0.0), 100.0))
        self._cfg["areal_density"] = float(np.float32(self.areal_density))

    # extensions (no reference counterpart)
    def set_compliances_raw(self, stretch, shear, bend):
        _lib.check(self._L.tc_sim_set_compliances_raw(self._h, float(stretch), float(shear), float(bend)))

    def set_fast_projection(self, enable: bool):
        """Opt-in fast projection on the per-iteration tile path: refined hardware reciprocals in place
        of the IEEE square root and divisions (not bit-identical; see tc_sim_set_fast_projection)."""
        _lib.check(self._L.tc_sim_set_fast_projection(self._h, 1 if enable else 0))

    def get_fast_projection(self) -> bool:
        e = C.c_int(0)
        _lib.check(self._L.tc_sim_get_fast_projection(self._h, C.byref(e)))
        return bool(e.value)

    def set_solver_path(self, path):
        """0 auto, 1 one launch per colour, 2 one launch per iteration (identical results)."""
        _lib.check(self._L.tc_sim_set_solver_path(self._h, int(path)))

    def autotune(self, dt_sub=1.0 / 600.0) -> dict:
        """Pick the fastest solver schedule / tile size / block size for this mesh (identical results)."""
        out = (C.c_double * 4)()
        _lib.check(self._L.tc_sim_autotune(self._h, float(dt_sub), out))
        return {"tile_vertices": int(out[0]), "block": int(out[1]), "ms_per_iteration": out[2],
                "per_colour_ms_per_iteration": out[3]}

    def solver_info(self) -> dict:
        out = (C.c_double * 10)()
        _lib.check(self._L.tc_sim_solver_info(self._h, out))
        keys = ["path", "tiles", "max_cone", "cone_vertices", "items_per_iteration", "constraints_per_iteration",
                "plan_available", "phases", "tile_vertices", "block"]
        info = dict(zip(keys, list(out)))
        info["plan_note"] = (self._L.tc_last_error() or b"").decode(errors="replace")
        return info

    def set_hash_cell_size(self, cell_size):
        _lib.check(self._L.tc_sim_set_hash_cell_size(self._h, float(cell_size)))

    # --- getters (PyO3 #[getter] properties, lib.rs:2927-2950) ---------------------------------
    @property
    def num_vertices(self) -> int:
        return int(self._L.tc_sim_num_vertices(self._h))

    @property
    def num_distance_constraints(self) -> int:
        return int(self._L.tc_sim_num_distance_constraints(self._h))

    @property
    def num_bending_constraints(self) -> int:
        return int(self._L.tc_sim_num_bending_constraints(self._h))

    @property
    def num_sewing_constraints(self) -> int:
        return int(self._sewing_state()[4])

    # --- sewing (lib.rs:2800-2947, sewing.wgsl, sewing_shrink.wgsl) ------------------------------
    def _sewing_state(self):
        out = np.zeros(5, np.float32)
        _lib.check(self._L.tc_sim_get_sewing_state(self._h, _ptr(out)))
        return out

    def set_sewing_stiffness(self, stiffness):
        _lib.check(self._L.tc_sim_set_sewing_stiffness(self._h, float(stiffness)))
        self._cfg["sewing_stiffness"] = float(self._sewing_state()[0])

    def set_sewing_shrink_speed(self, speed):
        _lib.check(self._L.tc_sim_set_sewing_shrink_speed(self._h, float(speed)))
        self._cfg["sewing_shrink_speed"] = float(self._sewing_state()[1])

    def get_sewing_shrink_speed(self) -> float:
        return float(self._sewing_state()[1])

    def set_enable_sewing_lock(self, enabled):
        _lib.check(self._L.tc_sim_set_enable_sewing_lock(self._h, 1 if enabled else 0))
        self._cfg["enable_sewing_lock"] = bool(enabled)

    def set_sewing_lock_distance(self, distance):
        _lib.check(self._L.tc_sim_set_sewing_lock_distance(self._h, float(distance)))
        self._cfg["sewing_lock_distance"] = float(self._sewing_state()[3])

    def get_sewing_lock_distance(self) -> float:
        return float(self._sewing_state()[3])

    def get_sewing_current_rest_lengths(self) -> list:
        """Current (shrinking) rest lengths, in the sewing constraints' colour order."""
        n = self.num_sewing_constraints
        out = np.empty(n, np.float32)
        _lib.check(self._L.tc_sim_get_sewing_rest_lengths(self._h, _ptr(out), n))
        return [float(x) for x in out]

    def set_sewing_current_rest_lengths(self, values):
        v = np.ascontiguousarray(np.asarray(values, np.float32).reshape(-1))
        _lib.check(self._L.tc_sim_set_sewing_rest_lengths(self._h, _ptr(v), len(v)))

    # --- sewing priority (lib.rs:2842-2923, mod.rs:1262-1420) ------------------------------------
    def _priority(self):
        out = np.empty(12, np.float64)
        _lib.check(self._L.tc_sim_get_sewing_priority_state(self._h, _ptr(out)))
        return out

    def set_sewing_priority_options(self, enabled, threshold=0.9, merge_dist=0.005, ramp_frames=3, max_frames=600):
        """Holds gravity and the wrinkle field at zero until a share `threshold` of the sewing pairs is
        closer than `merge_dist`, then ramps them back over `ramp_frames` updates."""
        _lib.check(self._L.tc_sim_set_sewing_priority_options(self._h, 1 if enabled else 0, float(threshold),
                                                              float(merge_dist), _u32_arg(ramp_frames),
                                                              _u32_arg(max_frames)))
        st = self._priority()
        self._cfg["sewing_priority_enabled"] = bool(st[0])
        self._cfg["sewing_priority_threshold"] = float(np.float32(st[1]))
        self._cfg["sewing_priority_merge_dist"] = float(np.float32(st[2]))
        self._cfg["sewing_priority_ramp_frames"] = int(st[3])
        self._cfg["sewing_priority_max_frames"] = int(st[4])

    def set_sewing_priority_state(self, latched, frame, ramp_t, ratio):
        _lib.check(self._L.tc_sim_set_sewing_priority_state(self._h, 1 if latched else 0, _u32_arg(frame),
                                                            float(ramp_t), float(ratio)))

    def get_sewing_priority_state(self) -> tuple:
        """(latched, frame, ramp_t, ratio): the latch's internal state (mod.rs:1296-1303)."""
        st = self._priority()
        return bool(st[5]), int(st[6]), float(st[7]), float(st[8])

    def update_sewing_priority(self, positions) -> tuple:
        """Measures the closure on the given positions (N x 3, no read-back) and advances the latch;
        returns (closure ratio, gravity scale, latched)."""
        p = np.ascontiguousarray(np.asarray(positions, np.float32))
        if p.ndim != 2 or p.shape[1] < 3:
            raise ValueError("positions must be an N x 3 array")
        p = np.ascontiguousarray(p[:, :3])
        out = np.empty(3, np.float32)
        _lib.check(self._L.tc_sim_update_sewing_priority(self._h, _ptr(p), len(p), _ptr(out)))
        return float(out[0]), float(out[1]), bool(out[2])

    def get_sewing_priority_scale(self) -> float:
        return float(np.float32(self._priority()[9]))

    def get_sewing_closure_ratio(self) -> float:
        return float(np.float32(self._priority()[10]))

    def is_sewing_priority_active(self) -> bool:
        return bool(self._priority()[11])

    # --- diagnostics: GPU pass timing (lib.rs:1975-1990, simulation/profile.rs) -----------------
    def set_profiling_enabled(self, enable) -> bool:
        """Times each pass of the coming frames with HIP events; returns whether it is on."""
        on = C.c_int(0)
        _lib.check(self._L.tc_sim_set_profiling_enabled(self._h, 1 if enable else 0, C.byref(on)))
        return bool(on.value)

    def take_profile(self) -> list:
        """(pass name, milliseconds) of the last frame stepped, scopes before their passes."""
        n = C.c_size_t(0)
        _lib.check(self._L.tc_sim_take_profile(self._h, C.byref(n)))
        out = []
        name, ms = C.c_char_p(), C.c_float()
        for i in range(n.value):
            _lib.check(self._L.tc_sim_profile_entry(self._h, i, C.byref(name), C.byref(ms), None, None))
            out.append((name.value.decode(), float(ms.value)))
        return out

    def save_profile_trace(self, path):
        """Chrome trace JSON of the last frame taken; OSError when the file cannot be written."""
        _lib.check(self._L.tc_sim_save_profile_trace(self._h, str(path).encode()))

    # --- introspection used by parity checks ---------------------------------------------------
    _FLOAT_ARRAYS = {8, 9, 13, 14, 15, 16, 22, 24, 25, 26}

    def sdf_collider_arrays(self) -> dict[str, np.ndarray]:
        """The SDF colliders as the resolve reads them: texels (u32) and infos (N, 20), and the dynamic
        bone SDF's last skinned positions (V, 3) and bone-space bake triangles (T, 16)."""
        return {"texels": self._array(23), "infos": self._array(24).reshape(-1, 20),
                "skinned": self._array(25).reshape(-1, 3), "bake_triangles": self._array(26).reshape(-1, 16)}

    def device_inv_masses(self) -> np.ndarray:
        """The inverse masses the solver reads (base values scaled by the pins)."""
        return self._array(22)

    def _array(self, which: int) -> np.ndarray:
        cnt = C.c_size_t(0)
        _lib.check(self._L.tc_sim_get_array(self._h, which, None, 0, C.byref(cnt)))
        out = np.empty(cnt.value, np.float32 if which in self._FLOAT_ARRAYS else np.uint32)
        if cnt.value:
            _lib.check(self._L.tc_sim_get_array(self._h, which, out.ctypes.data_as(C.c_void_p), out.size,
                                                C.byref(cnt)))
        return out

    def colouring(self) -> dict[str, np.ndarray]:
        return {"dist_v0": self._array(0), "dist_v1": self._array(1), "dist_col_off": self._array(2),
                "dist_col_cnt": self._array(3), "bend_v2": self._array(4), "bend_v3": self._array(5),
                "bend_col_off": self._array(6), "bend_col_cnt": self._array(7), "dist_rest": self._array(8),
                "bend_rest": self._array(9)}

    def hash_tables(self, rebuild=True) -> dict[str, np.ndarray]:
        if rebuild:
            _lib.check(self._L.tc_sim_build_hashes(self._h))
        sc = self._array(17)
        return {"keys": self._array(10), "starts": self._array(11), "sorted": self._array(12),
                "edge_keys": self._array(19), "edge_starts": self._array(20), "edge_sorted": self._array(21),
                "edges": self._array(18).reshape(-1, 2), "cell_size": float(sc[:1].view(np.float32)[0]),
                "table_size": int(sc[1]), "n_virtual": int(sc[2])}

    def predicted_positions(self) -> np.ndarray:
        return self._array(13).reshape(-1, 3)

    def velocities(self) -> np.ndarray:
        return self._array(14).reshape(-1, 3)

    def lambdas(self):
        return self._array(15), self._array(16)

    def substep_parts(self, dt_sub, collide=True, self_collide=True, update=True):
        _lib.check(self._L.tc_sim_substep_parts(self._h, float(dt_sub), int(collide), int(self_collide), int(update)))

    def residual(self, dt_sub):
        a, b = C.c_double(), C.c_double()
        _lib.check(self._L.tc_sim_residual(self._h, float(dt_sub), C.byref(a), C.byref(b)))
        return a.value, b.value

    def time_kernels(self, dt_sub, reps=5) -> dict:
        """Per-substep kernel times (ms) from HIP events on the simulator's stream."""
        out = (C.c_double * 12)()
        _lib.check(self._L.tc_sim_time_kernels(self._h, float(dt_sub), int(reps), out))
        keys = ["predict", "distance", "distance_launches", "bending", "bending_launches", "collision",
                "self_collision", "update_vel", "substep_wall", "fused", "fused_launches"]
        return dict(zip(keys, list(out)[:11]))

    def synchronize(self):
        _lib.check(self._L.tc_sim_synchronize(self._h))

    def set_bone_sdf_colliders(self, width, height, depth, texture_bytes, bone_infos):
        """SDF colliders from a baked Rg16Float volume and rows of 20 floats (lib.rs:2245-2284)."""
        infos = np.ascontiguousarray(np.asarray(bone_infos, dtype=np.float32))
        if infos.ndim != 2 or (infos.size and infos.shape[1] < 20):
            raise ValueError("each bone_infos row needs at least 20 elements")
        infos = np.ascontiguousarray(infos[:, :20])
        tex = np.frombuffer(bytes(texture_bytes), dtype=np.uint8)
        _lib.check(self._L.tc_sim_set_sdf_colliders(self._h, int(width), int(height), int(depth),
                                                    tex.ctypes.data_as(C.c_void_p), tex.size,
                                                    infos.ctypes.data_as(C.c_void_p), len(infos)))
        self._sdf_texture = tex

    def update_bone_transforms(self, world_matrices, inv_world_matrices, dirty_bone_indices=None):
        """Per-collider world and inverse-world 4x4 matrices (lib.rs:2287-2330); also sets the dynamic
        bone SDF's dirty bones (None: every bone), as the reference does."""
        w = np.ascontiguousarray(world_matrices, np.float32)
        iw = np.ascontiguousarray(inv_world_matrices, np.float32)
        if w.ndim != 3 or w.shape[1:] != (4, 4) or iw.shape != w.shape:
            raise ValueError("world_matrices and inv_world_matrices must both have shape (N, 4, 4)")
        _lib.check(self._L.tc_sim_update_bone_transforms(self._h, w.ctypes.data_as(C.c_void_p),
                                                         iw.ctypes.data_as(C.c_void_p), len(w)))
        self.set_dynamic_bone_sdf_dirty_bones(dirty_bone_indices)

    def clear_bone_sdf_colliders(self):
        """Removes the SDF colliders and the dynamic bone SDF (lib.rs:2342-2345)."""
        _lib.check(self._L.tc_sim_clear_bone_sdf_colliders(self._h))
        self._sdf_texture = None

    def setup_dynamic_bone_sdf(self, width, height, depth, res, bone_infos, rest_verts, bone_indices, bone_weights,
                               tri_sources, tri_weights, bone_bind_inv_matrices, update_interval=1):
        """Skinned collider mesh whose bone-SDF atlas is re-baked on the device every update_interval
        frames (lib.rs:2347-2588).  bone_infos (N_bones, >=20); rest_verts (V, >=3) (columns 3-5, the
        rest normals, are not used by any pass); bone_indices / bone_weights (V, >=4); tri_sources
        (T, >=4) = (i0, i1, i2, bone), each bone's triangles contiguous; tri_weights (T, >=3);
        bone_bind_inv_matrices (>=N_bones, 4, 4) row-major bind matrices (world -> bone)."""
        if int(res) == 0 or int(width) == 0 or int(height) == 0 or int(depth) == 0:
            raise ValueError("width, height, depth and res must all be 1 or more")
        infos = np.asarray(bone_infos, dtype=np.float32)
        if infos.ndim != 2 or (infos.shape[0] and infos.shape[1] < 20):
            raise ValueError("each bone_infos row needs at least 20 elements")
        rv = np.asarray(rest_verts, dtype=np.float32)
        bi = np.asarray(bone_indices, dtype=np.uint32)
        bw = np.asarray(bone_weights, dtype=np.float32)
        if rv.ndim != 2 or bi.ndim != 2 or bw.ndim != 2 or rv.shape[1] < 3 or bi.shape[1] < 4 or bw.shape[1] < 4:
            raise ValueError("rest_verts needs >= 3 columns, bone_indices and bone_weights >= 4")
        if bi.shape[0] != rv.shape[0] or bw.shape[0] != rv.shape[0]:
            raise ValueError("rest_verts, bone_indices and bone_weights must have the same number of rows")
        ts = np.asarray(tri_sources, dtype=np.uint32)
        tw = np.asarray(tri_weights, dtype=np.float32)
        if ts.ndim != 2 or tw.ndim != 2 or ts.shape[0] != tw.shape[0]:
            raise ValueError("tri_sources and tri_weights must have the same number of rows")
        if ts.shape[0] and (ts.shape[1] < 4 or tw.shape[1] < 3):
            raise ValueError("tri_sources needs 4 columns (i0, i1, i2, bone) and tri_weights 3 (w0, w1, w2)")
        inv = np.asarray(bone_bind_inv_matrices, dtype=np.float32)
        n_bones = infos.shape[0]
        if inv.ndim != 3 or inv.shape[0] < n_bones or inv.shape[1:] != (4, 4):
            raise ValueError("bone_bind_inv_matrices must have shape (N_bones, 4, 4)")
        infos = np.ascontiguousarray(infos[:, :20]) if n_bones else np.zeros((0, 20), np.float32)
        rv = np.ascontiguousarray(rv[:, :3])
        bi = np.ascontiguousarray(bi[:, :4])
        bw = np.ascontiguousarray(bw[:, :4])
        ts = np.ascontiguousarray(ts[:, :4]) if len(ts) else np.zeros((0, 4), np.uint32)
        tw = np.ascontiguousarray(tw[:, :3]) if len(tw) else np.zeros((0, 3), np.float32)
        inv = np.ascontiguousarray(inv)
        vp = C.c_void_p
        _lib.check(self._L.tc_sim_setup_dynamic_bone_sdf(
            self._h, int(width), int(height), int(depth), int(res), infos.ctypes.data_as(vp), n_bones,
            rv.ctypes.data_as(vp), len(rv), bi.ctypes.data_as(vp), bw.ctypes.data_as(vp), ts.ctypes.data_as(vp),
            tw.ctypes.data_as(vp), len(ts), inv.ctypes.data_as(vp), inv.shape[0], int(update_interval)))

    def set_dynamic_bone_sdf_dirty_bones(self, dirty_bone_indices=None):
        """Bones whose tiles the next updates re-bake (None: every bone; empty: none) (lib.rs:2332-2340)."""
        if dirty_bone_indices is None:
            _lib.check(self._L.tc_sim_set_dynamic_bone_sdf_dirty_bones(self._h, None, 0, 0))
            return
        d = np.ascontiguousarray(np.asarray(dirty_bone_indices, dtype=np.uint32).reshape(-1))
        _lib.check(self._L.tc_sim_set_dynamic_bone_sdf_dirty_bones(self._h, d.ctypes.data_as(C.c_void_p), len(d), 1))

    def set_dynamic_bone_sdf_enabled(self, enabled):
        _lib.check(self._L.tc_sim_set_dynamic_bone_sdf_enabled(self._h, 1 if enabled else 0))

    def set_dynamic_bone_sdf_update_interval(self, interval):
        _lib.check(self._L.tc_sim_set_dynamic_bone_sdf_update_interval(self._h, int(interval)))

    def dynamic_bone_sdf_stats(self) -> dict:
        """Diagnostics (no reference counterpart): enabled, baked, tiles baked, frame counter, skipped."""
        out = (C.c_uint32 * 5)()
        _lib.check(self._L.tc_sim_dynamic_bone_sdf_stats(self._h, out))
        return dict(zip(("enabled", "baked", "tiles_baked", "frame_counter", "skipped"), list(out)))

    # --- learned warm start (no reference counterpart; see taremincloth_amd.warmstart) -------
    def set_warm_start(self, net=None, mode=0, iterations=8, fallback_iterations=56, residual_threshold=1e-4,
                       history_weight=None):
        """mode 0 off, 1 previous correction, 2 network (`net`).  history_weight defaults to the
        network's training value (0.5 without a network)."""
        if history_weight is None:
            history_weight = getattr(net, "history_weight", 0.5) if net is not None else 0.5
        self._ws_net = net  # keep the network alive while it is attached
        _lib.check(self._L.tc_sim_set_warm_start(self._h, None if net is None else net.handle, int(mode),
                                                 int(iterations), int(fallback_iterations),
                                                 float(residual_threshold), float(history_weight)))

    def warm_start_stats(self) -> dict:
        out = (C.c_double * 8)()
        _lib.check(self._L.tc_sim_warm_start_stats(self._h, out))
        keys = ["checked", "fallbacks", "last_relative_residual", "last_max_relative", "mode", "iterations",
                "fallback_chunks", "ended_above_threshold"]
        return dict(zip(keys, list(out)))

    def ws_reset_stats(self):
        _lib.check(self._L.tc_sim_ws_reset_stats(self._h))

    def ws_record(self, capacity, stride=1, history_weight=0.5):
        _lib.check(self._L.tc_sim_ws_record(self._h, int(capacity), int(stride), float(history_weight)))

    def ws_num_records(self) -> int:
        return int(self._L.tc_sim_ws_num_records(self._h))

    def ws_get_record(self, idx) -> dict:
        n = self._n
        f = np.empty((n, 32), np.float32)
        t, p, d = (np.empty((n, 4), np.float32) for _ in range(3))
        vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        _lib.check(self._L.tc_sim_ws_get_record(self._h, int(idx), vp(f), vp(t), vp(p), vp(d)))
        return {"features": f, "target": t, "prediction": p, "previous": d[:, :3].copy()}

    def ws_last_output(self) -> np.ndarray:
        y = np.empty((self._n, 16), np.float32)
        _lib.check(self._L.tc_sim_ws_last_output(self._h, y.ctypes.data_as(C.c_void_p), y.size))
        return y

    def ws_time(self, dt_sub, reps=3) -> dict:
        """Per-part times (ms) of a warm-started substep from HIP events; advances the simulation."""
        out = (C.c_double * 16)()
        _lib.check(self._L.tc_sim_ws_time(self._h, float(dt_sub), int(reps), out))
        keys = ["predict", "features", "encoder", "round_layers", "decoder", "apply",
                "iterations", "gate", "fallback", "correction_update", "collide", "update_vel", "substep",
                "rounds", "warm_iterations"]
        return dict(zip(keys, list(out)))

    def ws_history(self, new=None) -> np.ndarray:
        """Previous-substep correction (N x 3); with `new`, replaces it and returns the old one."""
        out = np.empty((self._n, 3), np.float32)
        src = None if new is None else np.ascontiguousarray(new, np.float32).reshape(self._n, 3)
        _lib.check(self._L.tc_sim_ws_history(self._h, out.ctypes.data_as(C.c_void_p),
                                             None if src is None else src.ctypes.data_as(C.c_void_p), self._n))
        return out


def color_edge_pairs(n, edges) -> list:
    """Welsh-Powell colour per vertex pair (lib.rs:636-648 -> coloring.rs:180)."""
    L = _lib.load()
    e = _u32(edges, 2, "edges")
    out = np.empty(len(e), np.uint32)
    _lib.check(L.tc_color_edge_pairs(int(n), _ptr(e), len(e), _ptr(out)))
    return out.tolist()


def compute_combined_inv_masses(positions, faces=None, strand_edges=None, pin_weights=None, areal_density=0.15,
                                linear_density=0.0) -> np.ndarray:
    """Area/3 mass split with pin weights (mesh.rs:237-312); strands are not part of this build."""
    if strand_edges is not None and np.asarray(strand_edges).size and linear_density > 0:
        raise NotImplementedError("strand masses are not part of this build")
    L = _lib.load()
    pos = _f32(positions, (3,), "positions")
    f = None if faces is None else _u32(faces, 3, "faces")
    pw = None if pin_weights is None else _f32(pin_weights).reshape(-1)
    if pw is not None and len(pw) != len(pos):
        raise ValueError("pin_weights length does not match the number of vertices")
    out = np.empty(len(pos), np.float32)
    _lib.check(L.tc_compute_inv_masses(_ptr(pos), len(pos), _ptr(f), 0 if f is None else len(f), _ptr(pw),
                                       float(areal_density), _ptr(out)))
    return out


class ClothBatch(ClothSimulator):
    """Many independent cloth instances advanced together (no reference counterpart: the reference
    runs one simulator per cloth).  ``instances`` is a sequence of dicts with the keys
    ``positions``, ``edges`` and optionally ``faces``, ``inv_masses`` and ``thicknesses`` (the
    ClothSimulator arguments of that cloth, indices local to it).  Each instance keeps its own
    colouring; one launch per colour (or per iteration) advances all of them.  The shared
    keyword arguments are those of ClothSimulator.  Every ClothSimulator method works on the
    concatenation of the instances' vertices; ``instance_range`` gives each one's slice.
    Self-collision starts disabled."""

    def __init__(self, instances, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0,
                 compression_stiffness=None, shear_stiffness=None, areal_density=None):
        L = _lib.load()
        self._L = L
        self._h = None
        descs = (_lib.TcInstance * max(len(instances), 1))()
        keep = []
        total = 0
        for k, inst in enumerate(instances):
            pos = _f32(inst["positions"], (3,), f"instances[{k}].positions")
            e = _u32(inst["edges"], 2, f"instances[{k}].edges")
            n = len(pos)
            if e.size and int(e.max()) >= n:
                raise ValueError(f"instances[{k}]: edges contain an index >= number of vertices ({n})")
            f = inst.get("faces")
            if f is not None:
                f = _u32(f, 3, f"instances[{k}].faces")
                if f.size and int(f.max()) >= n:
                    raise ValueError(f"instances[{k}]: faces contain an index >= number of vertices ({n})")
                if f.size == 0:
                    f = None
            arrs = []
            for name in ("inv_masses", "thicknesses"):
                a = inst.get(name)
                if a is not None:
                    a = np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1))
                    if len(a) != n:
                        raise ValueError(f"instances[{k}].{name} length ({len(a)}) does not match the number of "
                                         f"vertices ({n})")
                arrs.append(a)
            d = descs[k]
            d.positions, d.num_vertices = _ptr(pos), n
            d.edges, d.num_edges = _ptr(e), len(e)
            d.faces, d.num_faces = _ptr(f), 0 if f is None else len(f)
            d.inv_masses, d.thicknesses = _ptr(arrs[0]), _ptr(arrs[1])
            keep.append((pos, e, f, arrs))
            total += n
        p = _lib.TcParams()
        L.tc_sim_params_default(C.byref(p))
        p.thickness = float(thickness)
        p.stiffness = float(stiffness)
        p.bending_stiffness = float(bending_stiffness)
        p.compression_stiffness = -1.0 if compression_stiffness is None else float(compression_stiffness)
        p.shear_stiffness = -1.0 if shear_stiffness is None else float(shear_stiffness)
        h = C.c_void_p()
        _lib.check(L.tc_sim_create_batch(descs, len(instances), C.byref(p), C.byref(h)))
        self._h = h
        self._n = total
        self._keep = keep
        self.areal_density = 0.15 if areal_density is None else float(min(max(areal_density, 0.0), 100.0))
        self._init_config(stiffness, compression_stiffness, shear_stiffness, bending_stiffness, 32)
        self._sync_pair_cfg()
        self._ctor = None
        self._colliders = []
        self._recorder = None

    def num_instances(self) -> int:
        n = C.c_uint32(0)
        _lib.check(self._L.tc_sim_instance_count(self._h, C.byref(n)))
        return int(n.value)

    def instance_range(self, i):
        """(first vertex, vertex count) of instance i in the batch's vertex arrays."""
        a, n = C.c_uint32(0), C.c_uint32(0)
        _lib.check(self._L.tc_sim_instance_range(self._h, int(i), C.byref(a), C.byref(n)))
        return int(a.value), int(n.value)

    def get_instance_positions(self, i) -> np.ndarray:
        a, n = self.instance_range(i)
        out = np.empty(3 * self._n, np.float32)
        self.get_positions(out)
        return out.reshape(-1, 3)[a:a + n].copy()


def _bake(mesh_verts, mesh_tris, voxel_size, margin, friction, thickness, restitution, close_open_boundaries,
          method):
    L = _lib.load()
    v = np.ascontiguousarray(np.asarray(mesh_verts, dtype=np.float32))
    t = np.ascontiguousarray(np.asarray(mesh_tris, dtype=np.int32))
    if v.ndim != 2 or v.shape[1] < 3 or t.ndim != 2 or (t.size and t.shape[1] < 3):
        raise ValueError("mesh_verts must be (N, 3) and mesh_tris (T, 3)")
    v = np.ascontiguousarray(v[:, :3])
    t = np.ascontiguousarray(t[:, :3]) if t.size else np.zeros((0, 3), np.int32)
    res = C.c_void_p()
    status = L.tc_bake_mesh_sdf(_ptr(v), len(v), _ptr(t), len(t), float(voxel_size), float(margin), float(friction),
                                float(thickness), float(restitution), 1 if close_open_boundaries else 0, int(method),
                                C.byref(res))
    if status == _lib.TC_ERR_RUNTIME:  # the reference raises RuntimeError for every bake failure (lib.rs:1410)
        raise RuntimeError("GPU mesh SDF bake failed: " + (L.tc_last_error() or b"").decode(errors="replace"))
    _lib.check(status)
    try:
        dims = (C.c_uint32 * 3)()
        info = (C.c_float * 20)()
        cv = C.c_float(0.0)
        _lib.check(L.tc_sdf_result_info(res, dims, info, C.byref(cv)))
        kms = (C.c_float * 2)()
        nfb = C.c_uint32(0)
        _lib.check(L.tc_sdf_result_timing(res, kms, C.byref(nfb)))
        n = dims[0] * dims[1] * dims[2]
        tex = np.ctypeslib.as_array(C.cast(L.tc_sdf_result_texels(res), C.POINTER(C.c_uint32)), (n,)).copy()
        out = {"texture_bytes": tex.tobytes(), "width": int(dims[0]), "height": int(dims[1]), "depth": int(dims[2]),
               "bone_info": np.array(info, dtype=np.float32), "kernel_ms": (float(kms[0]), float(kms[1])),
               "fallback_voxels": int(nfb.value)}
        if method == 2:
            near = np.ctypeslib.as_array(C.cast(L.tc_sdf_result_nearest(res), C.POINTER(C.c_uint32)), (n * 8,))
            out["nearest_tri"] = near.copy()
            out["dist_packed"] = tex
            out["coarse_voxel"] = float(cv.value)
        if method == 3:
            bd = (C.c_uint32 * 3)()
            ptr = L.tc_sdf_result_atlas(res, bd)
            if not ptr:
                raise RuntimeError("the bake returned no fp32 atlas")
            nb = bd[0] * bd[1] * bd[2]
            out["atlas"] = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float)), (nb * 512,)).copy()
            out["brick_dims"] = (int(bd[0]), int(bd[1]), int(bd[2]))
        return out
    finally:
        L.tc_sdf_result_free(res)


def bake_mesh_sdf_gpu(mesh_verts, mesh_tris, voxel_size=0.004, margin=0.02, friction=0.5, thickness=0.005,
                      restitution=0.0, close_open_boundaries=True) -> dict:
    """Dense mesh SDF bake (lib.rs:1350-1417 -> sdf_baker.rs:1114-1323)."""
    return _bake(mesh_verts, mesh_tris, voxel_size, margin, friction, thickness, restitution, close_open_boundaries, 0)


def bake_mesh_sdf_hierarchical_gpu(mesh_verts, mesh_tris, voxel_size=0.004, margin=0.02, friction=0.5,
                                   thickness=0.005, restitution=0.0, close_open_boundaries=True) -> dict:
    """Hierarchical mesh SDF bake: coarse top-8 pass + fine pass (lib.rs:1419-1485 -> sdf_baker.rs:1625-1882)."""
    return _bake(mesh_verts, mesh_tris, voxel_size, margin, friction, thickness, restitution, close_open_boundaries, 1)


def bake_mesh_sdf_coarse(mesh_verts, mesh_tris, voxel_size=0.004, margin=0.02, friction=0.5, thickness=0.005,
                         restitution=0.0, close_open_boundaries=True) -> dict:
    """The coarse pass alone (sdf_baker.rs:1328-1600): adds nearest_tri (8 per cell) and coarse_voxel."""
    return _bake(mesh_verts, mesh_tris, voxel_size, margin, friction, thickness, restitution, close_open_boundaries, 2)


def bake_mesh_sdf_atlas_fp32(mesh_verts, mesh_tris, voxel_size=0.004, margin=0.02, friction=0.5, thickness=0.005,
                             restitution=0.0, close_open_boundaries=True) -> dict:
    """Dense mesh SDF bake that also returns the signed distances as an fp32 brick atlas ("atlas":
    bx*by*bz bricks of 8^3 floats, "brick_dims": (bx, by, bz)); brick b = (bz*by_n + by)*bx_n + bx holds
    voxel (8bx + lx, 8by + ly, 8bz + lz) at offset (lz*8 + ly)*8 + lx, clamped to the domain's edge.
    The texels are those of bake_mesh_sdf_gpu; the reference keeps only those (bake_mesh_sdf.wgsl:120)."""
    return _bake(mesh_verts, mesh_tris, voxel_size, margin, friction, thickness, restitution, close_open_boundaries, 3)


def atlas_to_dense(atlas, brick_dims, dims):
    """The fp32 brick atlas as a dense (depth, height, width) array."""
    bx, by, bz = brick_dims
    w, h, d = dims
    a = np.asarray(atlas, np.float32).reshape(bz, by, bx, 8, 8, 8)
    return a.transpose(0, 3, 1, 4, 2, 5).reshape(bz * 8, by * 8, bx * 8)[:d, :h, :w]


def bake_bone_sdf_gpu(mesh_verts, mesh_tris, bone_names, bone_weights, bone_bind_matrices, resolution=64, margin=0.2,
                      weight_threshold=0.02, blend_k=0.05, friction=0.5, thickness=0.005, restitution=0.0,
                      close_open_boundaries=True) -> dict:
    """Bone SDF atlas bake (lib.rs:1190-1348 -> sdf_baker.rs:494-930): one res^3 tile per bone whose
    largest weight exceeds weight_threshold, each texel pack2x16float(signed distance in bone space,
    interpolated bone weight).  Returns texture_bytes, width, height, depth, bone_infos (N, 20),
    active_bones (names) and bind_matrices (N, 4, 4); "kernel_ms" is the bake kernel's device time."""
    if int(resolution) <= 0:
        raise ValueError("resolution must be 1 or more")
    v = np.asarray(mesh_verts, dtype=np.float32)
    if v.ndim != 2 or v.shape[1] < 3:
        raise ValueError("mesh_verts must have shape (N, 3)")
    t = np.asarray(mesh_tris, dtype=np.int32)
    if t.ndim != 2 or t.shape[1] < 3:
        raise ValueError("mesh_tris must have shape (M, 3)")
    v = np.ascontiguousarray(v[:, :3])
    t = np.ascontiguousarray(t[:, :3])
    names = [str(n) for n in bone_names]
    nb = len(names)
    if len(bone_weights) != nb:
        raise ValueError("bone_names and bone_weights differ in length")
    bm = np.asarray(bone_bind_matrices, dtype=np.float32)
    if bm.ndim != 3 or bm.shape[0] < nb or bm.shape[1:] != (4, 4):
        raise ValueError("bone_bind_matrices must have shape (N_bones, 4, 4)")
    w = np.zeros((nb, len(v)), np.float32)
    for i, bw in enumerate(bone_weights):
        a = np.asarray(bw, dtype=np.float32).reshape(-1)
        if a.size != len(v):
            raise ValueError(f"bone_weights[{i}] has length {a.size}, the mesh has {len(v)} vertices")
        w[i] = a
    binds = np.ascontiguousarray(bm[:nb].reshape(nb, 16))
    L = _lib.load()
    res = C.c_void_p()
    try:
        _lib.check(L.tc_bake_bone_sdf(_ptr(v), len(v), _ptr(t), len(t), _ptr(w), _ptr(binds), nb, int(resolution),
                                      float(margin), float(weight_threshold), float(blend_k), float(friction),
                                      float(thickness), float(restitution), 1 if close_open_boundaries else 0,
                                      C.byref(res)))
    except RuntimeError as e:
        raise RuntimeError(f"GPU SDF bake failed: {e}") from None
    try:
        dims = np.zeros(3, np.uint32)
        na, ms = C.c_uint32(), C.c_float()
        _lib.check(L.tc_bone_sdf_result_info(res, _ptr(dims), C.byref(na), C.byref(ms)))
        n = int(na.value)
        tex = np.empty(int(dims[0]) * int(dims[1]) * int(dims[2]), np.uint32)
        infos = np.empty((n, 20), np.float32)
        act = np.empty(n, np.uint32)
        mats = np.empty((n, 4, 4), np.float32)
        _lib.check(L.tc_bone_sdf_result_arrays(res, _ptr(tex), _ptr(infos), _ptr(act), _ptr(mats)))
    finally:
        L.tc_bone_sdf_result_free(res)
    return {"texture_bytes": tex.tobytes(), "width": int(dims[0]), "height": int(dims[1]), "depth": int(dims[2]),
            "bone_infos": infos, "active_bones": [names[int(i)] for i in act], "bind_matrices": mats,
            "kernel_ms": float(ms.value)}


def close_open_boundary_loops(mesh_verts, mesh_tris):
    """Caps closed boundary loops of a solid mesh (sdf_baker.rs:244-462): (verts, tris, loops_closed)."""
    L = _lib.load()
    v = np.ascontiguousarray(np.asarray(mesh_verts, dtype=np.float32).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(mesh_tris, dtype=np.int32).reshape(-1, 3))
    nv, nt, loops = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _lib.check(L.tc_close_open_boundary_loops(_ptr(v), len(v), _ptr(t), len(t), None, 0, None, 0, C.byref(nv),
                                              C.byref(nt), C.byref(loops)))
    ov = np.empty((nv.value, 3), np.float32)
    ot = np.empty((nt.value, 3), np.int32)
    _lib.check(L.tc_close_open_boundary_loops(_ptr(v), len(v), _ptr(t), len(t), _ptr(ov), ov.size, _ptr(ot), ot.size,
                                              C.byref(nv), C.byref(nt), C.byref(loops)))
    return ov, ot, int(loops.value)


_MESH_FLOAT_ARRAYS = {8, 9, 24, 25, 32, 33, 34, 40, 41, 44}
MESH_ARRAYS = {
    "dist_v0": 0, "dist_v1": 1, "dist_col_off": 2, "dist_col_cnt": 3, "bend_v2": 4, "bend_v3": 5,
    "bend_col_off": 6, "bend_col_cnt": 7, "dist_rest": 8, "bend_rest": 9, "scalars": 17, "edges": 18,
    "th_off": 22, "th_idx": 23, "th_rest": 24, "local_len": 25, "star_off": 26, "star": 27, "edge_bbits": 28,
    "island": 29, "edge_to_constraint": 30, "dist_type": 31, "dist_tc": 32, "dist_cc": 33, "bend_comp": 34,
    "adj_off": 35, "adj_idx": 36, "layer": 37, "bend_v0": 38, "bend_v1": 39, "bend_pad0": 40, "bend_pad1": 41,
    "sew_v0": 42, "sew_v1": 43, "sew_rest": 44, "sew_col_off": 45, "sew_col_cnt": 46,
}


def sewing_closure_ratio(positions, pairs, merge_dist=0.005) -> float:
    """Share of sewing pairs closer than merge_dist (mod.rs:1381-1420); pairs indexing outside the
    positions are skipped, and with none counted the share is 1.  Host only, no GPU."""
    p = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1, 3))
    q = np.ascontiguousarray(np.asarray(pairs, np.int64).reshape(-1, 2))
    keep = ((q >= 0) & (q < 2 ** 32)).all(axis=1)
    q = np.ascontiguousarray(q[keep].astype(np.uint32))
    out = C.c_float(0.0)
    _lib.check(_lib.load().tc_sewing_closure_ratio(_ptr(p), len(p), _ptr(q), len(q), float(merge_dist),
                                                   C.byref(out)))
    return float(out.value)


def build_mesh_arrays(positions, edges, faces=None, inv_masses=None, thicknesses=None, thickness=0.005,
                      stiffness=1000.0, bending_stiffness=10.0, compression_stiffness=None,
                      shear_stiffness=None, layer_id=0, sewing_springs=None, sewing_shrink_speed=1.0,
                      sewing_stiffness=None, enable_sewing_lock=None) -> dict:
    """Host-side mesh build of the library (no GPU needed): constraints, colours, topology."""
    L = _lib.load()
    pos = _f32(positions, (3,), "positions")
    e = _u32(edges, 2, "edges")
    f = None if faces is None else _u32(faces, 3, "faces")
    im = None if inv_masses is None else _f32(inv_masses).reshape(-1)
    th = None if thicknesses is None else _f32(thicknesses).reshape(-1)
    p = _lib.TcParams()
    L.tc_sim_params_default(C.byref(p))
    p.thickness, p.layer_id, p.stiffness, p.bending_stiffness = thickness, layer_id, stiffness, bending_stiffness
    p.compression_stiffness = -1.0 if compression_stiffness is None else compression_stiffness
    p.shear_stiffness = -1.0 if shear_stiffness is None else shear_stiffness
    sew = None if sewing_springs is None or np.asarray(sewing_springs).size == 0 else _u32(sewing_springs, 2,
                                                                                          "sewing_springs")
    sp = _lib.TcSewingParams()
    L.tc_sewing_params_default(C.byref(sp))
    sp.shrink_speed = float(sewing_shrink_speed)
    sp.stiffness = -1.0 if sewing_stiffness is None else float(sewing_stiffness)
    sp.lock = -1 if enable_sewing_lock is None else (1 if enable_sewing_lock else 0)
    out = {}
    for name, which in MESH_ARRAYS.items():
        cnt = C.c_size_t(0)
        args = (_ptr(pos), len(pos), _ptr(e), len(e), _ptr(f), 0 if f is None else len(f), _ptr(im), _ptr(th),
                _ptr(sew), 0 if sew is None else len(sew), C.byref(sp), C.byref(p), which)
        _lib.check(L.tc_mesh_build_arrays_sewn(*args, None, 0, C.byref(cnt)))
        arr = np.empty(cnt.value, np.float32 if which in _MESH_FLOAT_ARRAYS else np.uint32)
        if cnt.value:
            _lib.check(L.tc_mesh_build_arrays_sewn(*args, arr.ctypes.data_as(C.c_void_p), arr.size, C.byref(cnt)))
        out[name] = arr
    sc = out["scalars"]
    out["cell_size"] = float(sc[:1].view(np.float32)[0])
    out["table_size"], out["n_virtual"], out["ne"] = int(sc[1]), int(sc[2]), int(sc[3])
    return out
