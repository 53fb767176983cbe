"""Debug recording of a simulation as gzip-compressed JSON lines, and its replay.

Follows the record layout of crates/cloth_core/src/debug_recorder.rs (JsonlRecord tagged by
"record_type": one "metadata" line, then one "frame" line per recorded step; FrameRecord,
FrameStats, ColliderRecord tagged by "type", SimulationMetadata) and the capture of
simulation/recording.rs:110-230: after each step the positions and velocities, the colliders, the
frame statistics (largest speed and displacement, NaN/Inf, bounding box, largest stretch strain,
config hash), the dynamic pins (PinRecord: vertex_idx, target_pos, weight), on the frames where the
SimConfig changed the whole SimConfig as param_deltas, and on the frames where base inverse masses
changed their new values as inv_mass_deltas (debug_recorder.rs:478-505: a value counts as changed
when its bits differ, +0 and -0 being the same).  What this build does not simulate (bone
SDFs, pair counters) is written empty.

`load(path)` reads such a file; `replay(path)` builds a simulator from the metadata and steps it
through the recorded frames (dt, substeps, iterations, configuration changes), for replay-driven
parity: the replayed positions equal the recorded ones bit for bit on the same device.
"""
from __future__ import annotations

import datetime
import gzip
import json

import numpy as np

from . import sim_config as _sc


def _stats(pos, vel, prev, edges, rest, cfg_hash):
    finite = bool(np.isfinite(pos).all() and np.isfinite(vel).all())
    disp = float(np.abs(pos - prev).max()) if prev is not None and len(pos) else 0.0
    strain = 0.0
    if len(edges):
        d = np.linalg.norm(pos[edges[:, 0]] - pos[edges[:, 1]], axis=1)
        ok = rest > 1e-9
        if ok.any():
            strain = float(max(((d[ok] - rest[ok]) / rest[ok]).max(), 0.0))
    return {"max_velocity": float(np.linalg.norm(vel, axis=1).max()) if len(vel) else 0.0,
            "max_displacement": disp, "has_nan_or_inf": not finite,
            "aabb_min": pos.min(0).tolist() if len(pos) else [0.0, 0.0, 0.0],
            "aabb_max": pos.max(0).tolist() if len(pos) else [0.0, 0.0, 0.0],
            "vt_count": 0, "ee_count": 0, "vt_saturated": False, "max_strain": strain, "config_hash": cfg_hash}


def _priority_record(sim):
    """[latched01, frame, ramp_t, ratio] of the sewing priority latch, while it is on (recording.rs:226-232)."""
    if not sim._cfg.get("sewing_priority_enabled") or not sim.num_sewing_constraints:
        return None
    latched, frame, ramp_t, ratio = sim.get_sewing_priority_state()
    return [1.0 if latched else 0.0, float(np.float32(frame)), ramp_t, ratio]


class DebugRecorder:
    """Frames of one simulator, kept in memory up to max_frames (the oldest dropped first)."""

    def __init__(self, sim, object_name: str, max_frames: int):
        if sim._ctor is None:
            raise NotImplementedError("debug recording of a batch of instances is not part of this build")
        c = sim._ctor
        self.max_frames = max(int(max_frames), 1)
        self.frames = []
        self.seq = 0
        pos = np.asarray(c["positions"], np.float32)
        edges = np.asarray(c["edges"], np.uint32)
        self._edges = edges.astype(np.int64)
        self._rest = np.linalg.norm(pos[self._edges[:, 0]] - pos[self._edges[:, 1]], axis=1) if len(edges) else \
            np.zeros(0)
        inv = sim.get_vertex_inv_masses()
        self._prev_inv = inv.copy()
        self.metadata = {
            "object_name": object_name, "num_vertices": int(len(pos)), "num_edges": int(len(edges)),
            "num_faces": 0 if c["faces"] is None else int(len(c["faces"])),
            "edges": edges.tolist(), "faces": [] if c["faces"] is None else np.asarray(c["faces"]).tolist(),
            "inv_masses": np.asarray(inv, np.float32).tolist(), "initial_rest_lengths": self._rest.tolist(),
            "initial_positions": pos.tolist(),
            "thicknesses": None if c["thicknesses"] is None else np.asarray(c["thicknesses"]).tolist(),
            "layer_ids": None if c["layer_ids"] is None else np.asarray(c["layer_ids"]).tolist(),
            "sewing_springs": None if c.get("sewing_springs") is None else np.asarray(c["sewing_springs"]).tolist(),
            "stiffness": c["stiffness"], "compression_stiffness": c["compression_stiffness"],
            "shear_stiffness": c["shear_stiffness"], "bending_stiffness": c["bending_stiffness"],
            "thickness": c["thickness"], "layer_id": c["layer_id"],
            "gravity": list(sim._cfg["gravity"]), "damping": sim._cfg["damping"], "solver_mode": sim.get_solver_mode(),
            "config": json.loads(sim.get_config_json()),
        }
        self._prev_pos = None
        self._prev_hash = sim.get_config_hash()

    def record(self, sim, dt: float, substeps: int):
        n = sim.num_vertices
        flat = np.empty(3 * n, np.float32)
        sim.get_positions(flat)
        pos = flat.reshape(-1, 3)
        vel = sim.velocities()
        h = sim.get_config_hash()
        pidx, ptgt, pw = sim.get_pins()
        pins = [{"vertex_idx": int(i), "target_pos": [float(x) for x in t], "weight": float(w)}
                for i, t, w in zip(pidx, ptgt, pw)]
        inv = sim.get_vertex_inv_masses()
        changed = (inv.view(np.uint32) != self._prev_inv.view(np.uint32)) & ~((inv == 0) & (self._prev_inv == 0))
        deltas = [{"vertex_idx": int(i), "inv_mass": float(inv[i])} for i in np.flatnonzero(changed)]
        self._prev_inv = inv.copy()
        rec = {"frame_index": self.seq, "dt": dt, "substeps": substeps,
               "solver_iterations": sim.get_effective_solver_iterations(),
               "positions": pos.tolist(), "velocities": vel.tolist(),
               "pinned_indices": [p["vertex_idx"] for p in pins], "pins": pins,
               "colliders": [dict(c) for c in sim._colliders],
               "stats": _stats(pos, vel, self._prev_pos, self._edges, self._rest, h),
               "param_deltas": json.loads(sim.get_config_json()) if h != self._prev_hash else None,
               "elastic_scales": None, "bone_transforms": None,
               "sewing_rest_lengths": sim.get_sewing_current_rest_lengths() if sim.num_sewing_constraints else None,
               "sewing_priority": _priority_record(sim),
               "inv_mass_deltas": deltas or None}
        self._prev_pos = pos.copy()
        self._prev_hash = h
        self.seq += 1
        self.frames.append(rec)
        if len(self.frames) > self.max_frames:
            self.frames.pop(0)

    def save(self, path: str) -> str:
        with gzip.open(path, "wt", encoding="utf-8") as f:
            f.write(json.dumps({"record_type": "metadata", "version": 1,
                                "created_at": datetime.datetime.now(datetime.timezone.utc).isoformat(),
                                "metadata": self.metadata}) + "\n")
            for r in self.frames:
                f.write(json.dumps(dict({"record_type": "frame"}, **r)) + "\n")
        return path


def load(path: str):
    """(metadata, frames) of a recording written by save()."""
    meta, frames = None, []
    with gzip.open(path, "rt", encoding="utf-8") as f:
        for line in f:
            r = json.loads(line)
            if r.get("record_type") == "metadata":
                meta = r["metadata"]
            elif r.get("record_type") == "frame":
                frames.append(r)
    if meta is None:
        raise ValueError(f"{path}: no metadata record")
    return meta, frames


def replay(path: str):
    """Rebuilds the recorded simulator and steps it through the recorded frames; yields
    (frame record, replayed positions N x 3) per frame."""
    from .taremin_cloth_core import ClothSimulator
    meta, frames = load(path)
    sim = ClothSimulator(np.asarray(meta["initial_positions"], np.float32), np.asarray(meta["edges"], np.uint32),
                         np.asarray(meta["faces"], np.uint32).reshape(-1, 3) if meta["faces"] else None,
                         np.asarray(meta["inv_masses"], np.float32),
                         sewing_springs=None if meta.get("sewing_springs") is None else
                         np.asarray(meta["sewing_springs"], np.uint32).reshape(-1, 2),
                         thicknesses=None if meta["thicknesses"] is None else np.asarray(meta["thicknesses"], np.float32),
                         layer_ids=None if meta["layer_ids"] is None else np.asarray(meta["layer_ids"], np.uint32),
                         layer_id=meta.get("layer_id", 0), thickness=meta["thickness"], stiffness=meta["stiffness"],
                         bending_stiffness=meta["bending_stiffness"],
                         compression_stiffness=meta["compression_stiffness"], shear_stiffness=meta["shear_stiffness"])
    sim.apply_config_json(json.dumps(meta["config"]))
    colliders = None
    for fr in frames:
        if fr["param_deltas"] is not None:
            sim.apply_config_json(json.dumps(fr["param_deltas"]))
        if fr["colliders"] != colliders:
            sim.clear_colliders()
            for c in fr["colliders"]:
                if c["type"] == "sphere":
                    sim.add_sphere_collider(c["center"], c["radius"], c["friction"], c["restitution"])
                elif c["type"] == "capsule":
                    sim.add_capsule_collider(c["point_a"], c["point_b"], c["radius"], c["friction"], c["restitution"])
                elif c["type"] == "plane":
                    sim.add_plane_collider(c["point"], c["normal"], c["friction"], c["restitution"])
                elif c["type"] == "mesh":
                    sim.set_mesh_collider_triangles(np.asarray(c["triangles"], np.float32).reshape(-1, 3, 3),
                                                    c["friction"], c["thickness"], c["restitution"], c["single_sided"])
            colliders = fr["colliders"]
        if fr.get("inv_mass_deltas"):
            sim.set_vertex_inv_masses(np.array([d["vertex_idx"] for d in fr["inv_mass_deltas"]], np.uint32),
                                      np.array([d["inv_mass"] for d in fr["inv_mass_deltas"]], np.float32))
        pins = fr.get("pins") or []
        gidx, gt, gw = sim.get_pins()
        want = ([p["vertex_idx"] for p in pins], [p["target_pos"] for p in pins], [p["weight"] for p in pins])
        if (list(gidx) != want[0] or not np.array_equal(gt, np.asarray(want[1], np.float32).reshape(-1, 3))
                or not np.array_equal(gw.view(np.uint32), np.asarray(want[2], np.float32).view(np.uint32))):
            gone = sorted(set(int(i) for i in gidx) - set(want[0]))
            if gone:
                sim.release_pins_batch(gone)
            if pins:
                sim.set_pins_batch(want[0], np.asarray(want[1], np.float32).reshape(-1, 3), want[2])
        # the latch state recorded with the frame is the state the frame stepped with (replayer.py:537-558)
        st = fr.get("sewing_priority")
        if isinstance(st, list) and len(st) >= 4:
            sim.set_sewing_priority_state(float(st[0]) > 0.5, int(st[1]), float(st[2]), float(st[3]))
        # the recorded substeps and iterations are what ran; the adaptive policy is not re-run
        sim.set_adaptive_substep_options(False, sim._cfg["min_substeps"], sim._cfg["max_substeps"])
        sim.set_solver_iterations(fr["solver_iterations"])
        sim.step(fr["dt"], fr["substeps"])
        out = np.empty(3 * sim.num_vertices, np.float32)
        sim.get_positions(out)
        yield fr, out.reshape(-1, 3)
