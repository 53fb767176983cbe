"""Wrinkle field (wrinkle_field.wgsl:58-154; parameters lib.rs:3049-3137, simulation/types.rs:110-139;
after predict, dispatch.rs:151-165).

CPU tests pin the oracle's restatement:
  * atan2_det against numpy's arctan2 on every octant (within 3e-7 rad);
  * a uniform texture: one substep of a free vertex at radius r moves it by dir * a dt^2 with the
    closed-form valley pull a = R blend stiffness clamp((r - max(r_valley, bone + 2 mm)) / window)
    and crest push a = G blend stiffness clamp((r_crest - r) / window), both at once, and the
    velocity by dir * a dt;
  * a 4-column texture: a vertex at a column centre's angle gets exactly that column's strength
    (bilinear weight 0), halfway between two columns the mean of both; u repeats across 2 pi;
  * nothing moves outside the z range, beyond r_max + influence radius, on the axis, when pinned,
    when the field is off or has no texture.
GPU tests: the product bit-identical to the oracle on a cloth tube around a bone with a random
texture, over frames, and the API's argument handling.
"""
import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O


def _T():
    from taremincloth_amd import taremin_cloth_core as T
    return T


def _pos(sim):
    if isinstance(sim, O.OracleSim):
        return sim.get_positions().reshape(-1, 3)
    out = np.empty(3 * sim.num_vertices, np.float32)
    sim.get_positions(out)
    return out.reshape(-1, 3)


def _params(influence=0.05, bone_r=0.05, stiffness=2.0, blend=0.8, z=(-0.2, 0.2), r=(0.02, 0.2), vw=0.015,
            cw=0.02):
    return np.array([0, 0, 0, influence, 0, 0, 1, bone_r, 1, 0, 0, stiffness, 0, 1, 0, blend,
                     z[0], z[1], r[0], r[1], vw, cw], np.float32)


def _free(points, precision="f64"):
    """Isolated vertices (pairs joined by inactive constraints), no gravity, damping or iterations."""
    pts = np.asarray(points, np.float32)
    n = len(pts)
    e = np.array([[2 * k, 2 * k + 1] for k in range(n // 2)], np.uint32)
    o = O.OracleSim(pts, e, None, np.ones(n, np.float32), precision=precision)
    o.set_compliances_raw(1e10, 1e10, 1e10)
    o.set_gravity(0.0, 0.0, 0.0)
    o.set_damping(0.0)
    o.set_solver_iterations(0)
    o.set_enable_self_collision(False)
    return o


def test_atan2_det_matches_arctan2():
    L = O.lib("f32")
    import ctypes as C
    fn = getattr(L, "orc_atan2_det", None)
    assert fn is not None
    fn.restype = C.c_float
    fn.argtypes = [C.c_float, C.c_float]
    rng = np.random.default_rng(0)
    pts = np.concatenate([rng.normal(size=(4000, 2)), [[0, 1], [1, 0], [0, -1], [-1, 0], [1, 1], [-1, -1],
                                                        [1e-30, 1.0], [3.0, -1e-20]]])
    for y, x in pts.astype(np.float32):
        got = fn(float(y), float(x))
        want = np.arctan2(np.float64(y), np.float64(x))
        assert abs(got - want) < 3e-7, (y, x, got, want)


def _step(o, dt=1.0 / 600.0):
    o.step_single_substep(dt)


@pytest.mark.parametrize("r", [0.03, 0.065, 0.075, 0.09, 0.12])
def test_oracle_uniform_texture_closed_form(r):
    R, G, B, A = 200, 100, 51, 153  # valley 0.784, crest 0.392, radii r_min + (B, A)/255 * (r_max - r_min)
    tex = bytes([R, G, B, A]) * (8 * 4)
    ang = 0.7
    p = np.array([[r * np.cos(ang), r * np.sin(ang), 0.05], [0.5, 0.5, 0.5]], np.float32)
    o = _free(p)
    prm = _params()
    o.set_wrinkle(True, prm, tex, 8, 4)
    dt = np.float64(np.float32(1.0 / 600.0))
    _step(o, float(dt))
    f32 = lambda x: np.float64(np.float32(x))
    rr = np.hypot(np.float64(p[0, 0]), np.float64(p[0, 1]))
    rl = f32(0.2) - f32(0.02)
    vr = max(f32(0.02) + (B / 255.0) * rl, f32(0.05) + f32(0.002))
    cr = f32(0.02) + (A / 255.0) * rl
    pull = (R / 255.0) * f32(0.8) * f32(2.0) * np.clip(max(0.0, rr - vr) / f32(0.015), 0, 1)
    push = (G / 255.0) * f32(0.8) * f32(2.0) * np.clip(max(0.0, cr - rr) / f32(0.02), 0, 1)
    a = push - pull
    d = np.array([p[0, 0], p[0, 1], 0.0]) / rr
    want = p[0].astype(np.float64) + d * a * dt * dt
    np.testing.assert_allclose(o.state(0)[0], want, atol=1e-12)
    np.testing.assert_allclose(o.state(2)[0], d * a * dt, atol=1e-9)
    assert np.array_equal(o.state(0)[1], p[1].astype(np.float64))  # beyond r_max + influence


def test_oracle_columns_and_repeat():
    vals = [40, 120, 200, 255]
    tex = b"".join(bytes([v, 0, 255, 0]) for v in vals) * 2  # 4 x 2, valley only, target radius r_max
    r = 0.25  # beyond r_max: pull saturates once r - 0.2 >= window, here 0.05 / 0.015 > 1
    cols = [(k + 0.5) / 4 * 2 * np.pi for k in range(4)]
    mids = [(k + 1.0) / 4 * 2 * np.pi for k in range(4)]
    angs = cols + mids
    pts = [[r * np.cos(a), r * np.sin(a), 0.0] for a in angs]
    o = _free(pts)
    o.set_wrinkle(True, _params(influence=0.1, r=(0.02, 0.2)), tex, 4, 2)
    dt = np.float64(np.float32(1.0 / 600.0))
    _step(o, float(dt))
    p0 = np.asarray(pts, np.float32).astype(np.float64)
    moved = np.linalg.norm(o.state(0)[:, :2], axis=1) - np.linalg.norm(p0[:, :2], axis=1)
    a = -moved / (dt * dt) / (np.float64(np.float32(0.8)) * np.float64(np.float32(2.0)))  # = valley strength
    np.testing.assert_allclose(a[:4], np.array(vals) / 255.0, atol=2e-3)
    want_mid = [(vals[k] + vals[(k + 1) % 4]) / 2 / 255.0 for k in range(4)]
    np.testing.assert_allclose(a[4:], want_mid, atol=2e-3)


def test_oracle_inert_cases():
    pts = np.array([[0.1, 0.0, 0.5], [0.1, 0.0, 0.0], [0.0, 0.0, 0.0], [0.4, 0.0, 0.0],
                    [0.1, 0.0, 0.0], [0.0, 0.1, 0.0]], np.float32)
    tex = bytes([255, 255, 128, 128]) * 4
    o = _free(pts)
    o.set_wrinkle(True, _params(influence=0.05, r=(0.02, 0.2)), tex, 2, 2)
    inv = np.ones(6, np.float32)
    o.set_vertex_inv_masses([4], [0.0])
    _step(o)
    got = o.state(0)
    # z out of range, on the axis, beyond r_max + influence, pinned: unmoved
    for k in (0, 2, 3, 4):
        assert np.array_equal(got[k], pts[k].astype(np.float64)), k
    assert not np.array_equal(got[1], pts[1].astype(np.float64))
    for on, t in ((False, tex), (True, None)):
        q = _free(pts)
        q.set_wrinkle(on, _params(), t, 2 if t else 0, 2 if t else 0)
        _step(q)
        assert np.array_equal(q.state(0), pts.astype(np.float64))


# ------------------------------------------------------------------ GPU: the product against the oracle
def _tube(n_around=24, n_along=16, radius=0.09, length=0.3):
    th = np.arange(n_around) * 2 * np.pi / n_around
    zs = np.linspace(-length / 2, length / 2, n_along)
    pos = np.array([[radius * np.cos(t), radius * np.sin(t), z] for z in zs for t in th], np.float32)
    idx = lambda i, j: j * n_around + (i % n_around)
    edges, faces = set(), []
    for j in range(n_along - 1):
        for i in range(n_around):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i, j + 1), idx(i + 1, j + 1)
            faces += [[a, b, d], [a, d, c]]
            for u, v in ((a, b), (a, c), (a, d), (b, d), (c, d)):
                edges.add((min(u, v), max(u, v)))
    return pos, np.array(sorted(edges), np.uint32), np.array(faces, np.uint32)


@pytest.mark.gpu
def test_gpu_wrinkle_tube_bit_exact(require_gpu):
    T = _T()
    pos, e, f = _tube()
    inv = cases.areal_inv_masses_np(pos, f, 0.15)
    rng = np.random.default_rng(5)
    tw, th = 16, 8
    tex = rng.integers(0, 256, size=(th, tw, 4), dtype=np.uint8)
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.003)
    o = O.OracleSim(pos, e, f, inv, thickness=0.003)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_compliances_raw(1e-7, 1e-6, 1e-5)
        s.set_enable_self_collision(False)
    g.set_enable_wrinkle_field(True)
    g.set_wrinkle_field_params([0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                               influence_radius=0.05, bone_radius=0.05, stiffness=3.0, blend_weight=0.9,
                               valley_window=0.01, crest_window=0.02)
    g.set_wrinkle_field_texture_2d(tw, th, tex.tobytes(), -0.2, 0.2, 0.02, 0.16)
    assert g.wrinkle_field_enabled()
    o.set_wrinkle(True, g._wrinkle_params(), tex.tobytes(), tw, th)
    free = T.ClothSimulator(pos, e, f, inv, thickness=0.003)
    free.set_gravity(0.0, 0.0, -9.81)
    free.set_compliances_raw(1e-7, 1e-6, 1e-5)
    free.set_enable_self_collision(False)
    for fr in range(6):
        for s in (g, o, free):
            s.step(1.0 / 60.0, 5, 5)
        diff = float(np.abs(_pos(g) - _pos(o)).max())
        print(f"[wrinkle] frame {fr}: max |hip - oracle| = {diff:.3e}", flush=True)
        assert diff == 0.0
    assert float(np.abs(_pos(g) - _pos(free)).max()) > 1e-4  # the field acts
    with pytest.raises(ValueError):
        g.set_wrinkle_field_texture_2d(4, 4, b"\0" * 10, 0, 1, 0, 1)
    g.set_wrinkle_field_texture_2d(0, 0, b"", 0, 1, 0, 1)  # drops the texture: the field stops
    a = _pos(g)
    g2 = T.ClothSimulator(pos, e, f, inv, thickness=0.003)
    assert not g2.wrinkle_field_enabled()
