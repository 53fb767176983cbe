"""SimConfig: the simulator's adjustable parameters as one JSON object.

Mirrors crates/cloth_core/src/config.rs:45-153 (fields, order and defaults; a missing field takes its
default, config.rs:238-246) and the PyO3 methods get_config_json / apply_config_json /
get_config_hash (crates/cloth_py/src/lib.rs:3028-3048, simulation/config_impl.rs:16-166).

The JSON text follows serde_json's compact form for the struct: fields in declaration order, no
spaces, f32 values in their shortest round-trip decimal with a trailing ".0" for integral values,
booleans as true/false.  The hash is 64-bit FNV-1a over that text (config.rs:215-224).
"""
from __future__ import annotations

import json
from collections import OrderedDict

import numpy as np

# (name, kind, default) in config.rs declaration order; kind: f = f32, u = u32, b = bool, v3 = [f32; 3]
FIELDS = [
    ("version", "u", 1), ("gravity", "v3", [0.0, 0.0, -9.81]), ("damping", "f", 1.0),
    ("tension_damping", "f", 5.0), ("compression_damping", "f", 5.0), ("shear_damping", "f", 5.0),
    ("bending_damping", "f", 0.5), ("tension_stiffness", "f", 1500.0), ("compression_stiffness", "f", 150.0),
    ("shear_stiffness", "f", 100.0), ("bending_stiffness", "f", 15.0), ("solver_iterations", "u", 2),
    ("solver_mode", "u", 0), ("workgroup_size", "u", 32), ("enable_self_collision", "b", False),
    ("relief_factor", "f", 1.0), ("max_displacement_ratio", "f", 0.5), ("exclude_neighbors", "b", True),
    ("enable_normal_untangling", "b", True), ("self_collision_max_iterations", "u", 256),
    ("coupled_mode", "u", 0), ("coupled_collider", "b", False), ("post_relaxation_iters", "u", 2),
    ("substep_interval", "u", 1), ("ee_substep_interval", "u", 1), ("self_collision_algorithm", "u", 0),
    ("enable_pair_cache", "b", False), ("pair_margin_mode", "u", 1), ("pair_safety_margin", "f", 0.005),
    ("pair_horizon_scale", "f", 1.3), ("pair_max_horizon", "f", 0.02), ("pair_max_pairs", "u", 65536),
    ("enable_pair_final_fallback", "b", True), ("enable_edge_collision", "b", False),
    ("edge_margin_scale", "f", 1.0), ("edge_margin_offset", "f", 0.0), ("sewing_stiffness", "f", 10000.0),
    ("sewing_shrink_speed", "f", 1.0), ("enable_sewing_lock", "b", True), ("sewing_lock_distance", "f", 0.02),
    ("sewing_priority_enabled", "b", False), ("sewing_priority_threshold", "f", 0.9),
    ("sewing_priority_merge_dist", "f", 0.005), ("sewing_priority_ramp_frames", "u", 3),
    ("sewing_priority_max_frames", "u", 600), ("areal_density", "f", 0.15),
    ("enable_adaptive_substep", "b", False), ("min_substeps", "u", 4), ("max_substeps", "u", 64),
    ("auto_coupled_on_low_substeps", "b", True), ("auto_compensate_iterations", "b", True),
    ("enable_strain_adaptive", "b", True), ("strain_tolerance", "f", 0.008),
]
KINDS = {n: k for n, k, _ in FIELDS}


def _as_f32(x) -> float:
    return float(np.float32(x))


def defaults() -> "OrderedDict[str, object]":
    """The defaults of config.rs, f32 fields held as the f32 values they are."""
    def val(k, d):
        if k == "v3":
            return [_as_f32(x) for x in d]
        return _as_f32(d) if k == "f" else d
    return OrderedDict((n, val(k, d)) for n, k, d in FIELDS)


def _f32_text(x) -> str:
    """Shortest decimal that reads back as the same f32, in the notation of the ryu formatter that
    serde_json uses: plain decimals (with ".0" when integral) for exponents -5..16, else d.ddde+x."""
    f = np.float32(x)
    if not np.isfinite(f):
        return "null"
    sci = np.format_float_scientific(f, unique=True, trim="-", exp_digits=1)  # e.g. '9.81e+0'
    mant, exp = sci.split("e")
    e = int(exp)
    neg = mant.startswith("-")
    digits = mant.lstrip("-").replace(".", "")
    if f == 0:
        return "-0.0" if neg else "0.0"
    k = len(digits)
    kk = e + 1  # position of the decimal point relative to the digit string
    if 0 <= kk - k and kk <= 16:
        s = digits + "0" * (kk - k) + ".0"
    elif 0 < kk <= 16:
        s = digits[:kk] + "." + digits[kk:]
    elif -5 < kk <= 0:
        s = "0." + "0" * (-kk) + digits
    else:
        s = (digits[0] + ("." + digits[1:] if k > 1 else "")) + "e" + str(kk - 1)
    return ("-" if neg else "") + s


def _value_text(kind, v) -> str:
    if kind == "b":
        return "true" if v else "false"
    if kind == "u":
        return str(int(v))
    if kind == "v3":
        return "[" + ",".join(_f32_text(x) for x in v) + "]"
    return _f32_text(v)


def to_json(cfg) -> str:
    return "{" + ",".join(f'"{n}":{_value_text(KINDS[n], cfg[n])}' for n, _, _ in FIELDS) + "}"


def fnv1a64(text: str) -> int:
    h = 0xCBF29CE484222325
    for b in text.encode():
        h ^= b
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def parse(json_str: str) -> "OrderedDict[str, object]":
    """Reads a SimConfig JSON object: missing fields take their defaults, unknown fields are ignored
    (serde's default behaviour); a malformed text or a value of the wrong type raises ValueError."""
    try:
        raw = json.loads(json_str)
    except json.JSONDecodeError as e:
        raise ValueError(f"invalid SimConfig JSON: {e}") from None
    if not isinstance(raw, dict):
        raise ValueError("SimConfig JSON must be an object")
    cfg = defaults()
    for n, k, _ in FIELDS:
        if n not in raw:
            continue
        v = raw[n]
        if k == "b":
            if not isinstance(v, bool):
                raise ValueError(f"SimConfig field {n} must be a boolean")
        elif k == "u":
            if isinstance(v, bool) or not isinstance(v, int) or v < 0 or v > 0xFFFFFFFF:
                raise ValueError(f"SimConfig field {n} must be an unsigned 32-bit integer")
        elif k == "v3":
            if not isinstance(v, list) or len(v) != 3 or not all(isinstance(x, (int, float)) for x in v):
                raise ValueError(f"SimConfig field {n} must be an array of 3 numbers")
            v = [float(np.float32(x)) for x in v]
        else:
            if isinstance(v, bool) or not isinstance(v, (int, float)):
                raise ValueError(f"SimConfig field {n} must be a number")
            v = float(np.float32(v))
        cfg[n] = v
    return cfg
