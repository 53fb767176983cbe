"""Pair-cache self-collision (self_collision_algorithm 1): per-vertex top-K candidate lists
(self_collision_collect_pairs.wgsl), narrowphase on the cached pairs (self_collision_solve_vt.wgsl,
self_collision_solve_ee.wgsl), the frame orchestration (dispatch.rs:1554-1645) and
get_active_pairs (dispatch.rs:1255-1291).

CPU tests pin the oracle's restatement:
  * the lists from the hash tables equal the lists from a scan of every vertex (same cells, same
    order), counters included;
  * closed form: a vertex hovering over the interior of a separate triangle holds exactly that
    triangle (smallest index first, orientation kept) while its height is under the summed
    thickness plus the safety margin, and nothing above it;
  * every held pair lies within the acceptance distance, the lists are sorted by distance and
    no vertex holds more than its quota;
  * the reference's own tests (tests/core/test_pair_caching.py): a two-layer stack is pushed
    apart with no pass-through, fixed margins collect nothing across a 30 mm gap while the
    velocity horizon catches the approaching layers, collection is deterministic, and a one-slot
    budget still covers at least half of the vertices.
GPU tests: the lists (slot order) and counters bit-identical to the oracle's, positions over
frames bit-identical with the final fallback on and off, and the reference's scenes on the
product.
"""
import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O


def _say(msg):
    print(f"[pair-cache] {msg}", flush=True)
    sys.stdout.flush()


def _T():
    from taremincloth_amd import taremin_cloth_core as T
    return T


def _pos(sim):
    if isinstance(sim, O.OracleSim):
        return sim.get_positions().reshape(-1, 3)
    out = np.empty(3 * sim.num_vertices, np.float32)
    sim.get_positions(out)
    return out.reshape(-1, 3)


def grid_cloth(nx=4, ny=4, dx=0.05, z=0.0, pin_first_row=True):
    """create_grid_cloth of tests/core/test_pair_caching.py:33-64."""
    pos, inv, edges, faces = [], [], [], []
    for y in range(ny):
        for x in range(nx):
            pos.append([x * dx, y * dx, z])
            inv.append(0.0 if pin_first_row and y == 0 else 1.0)
    for y in range(ny):
        for x in range(nx):
            i = y * nx + x
            if x + 1 < nx:
                edges.append([i, i + 1])
            if y + 1 < ny:
                edges.append([i, i + nx])
            if x + 1 < nx and y + 1 < ny:
                faces.append([i, i + 1, i + nx + 1])
                faces.append([i, i + nx + 1, i + nx])
    return (np.array(pos, np.float32), np.array(edges, np.uint32), np.array(faces, np.uint32),
            np.array(inv, np.float32))


def two_layers(nx=4, ny=4, dx=0.05, gap=0.003):
    p1, e1, f1, _ = grid_cloth(nx, ny, dx, 0.0, False)
    p2, e2, f2, _ = grid_cloth(nx, ny, dx, gap, False)
    n1 = len(p1)
    return (np.vstack([p1, p2]), np.vstack([e1, e2 + n1]), np.vstack([f1, f2 + n1]),
            np.ones(2 * n1, np.float32))


def _oracle(pos, e, f, inv, thickness=0.005, fallback=False, brute=False, **kw):
    o = O.OracleSim(pos, e, f, inv, thickness=thickness, **kw)
    o.set_enable_self_collision(True)
    o.set_pair_cache(True, final_fallback=fallback, brute=brute)
    return o


def _lists(sim):
    if isinstance(sim, O.OracleSim):
        return sim.active_pairs()
    return sim.active_pairs_arrays()


# ------------------------------------------------------------------ CPU: the oracle pinned
@pytest.mark.parametrize("scene", ["two_layers", "folded"])
def test_oracle_hash_and_scan_collect_the_same_lists(scene):
    if scene == "two_layers":
        pos, e, f, inv = two_layers(5, 5, 0.05, 0.003)
        vel = np.zeros_like(pos)
        vel[len(pos) // 2:, 2] = -0.4
    else:
        pos, e, f, inv = cases.folded_cloth(10, 0.02, 0.004)
        vel = np.zeros_like(pos)
        vel[:, 2] = np.where(pos[:, 2] > 0.002, -0.3, 0.1)
    out = []
    for brute in (False, True):
        o = _oracle(pos, e, f, inv, brute=brute)
        o.set_gravity(0.0, 0.0, -9.81)
        o.set_positions_and_velocities(pos, vel)
        o.step(1.0 / 60.0, 3)
        out.append(o.active_pairs())
    (c0, vt0, ee0), (c1, vt1, ee1) = out
    assert np.array_equal(c0, c1) and np.array_equal(vt0, vt1) and np.array_equal(ee0, ee1)
    assert len(vt0) > 0 and len(ee0) > 0


def _hover(h):
    """Triangle B (0, 1, 2) at z = 0 and a separate triangle A (3, 4, 5) whose vertex 3 hangs at
    height h over B's interior."""
    pos = np.array([[0.0, 0.0, 0.0], [0.04, 0.0, 0.0], [0.0, 0.04, 0.0],
                    [0.01, 0.01, h], [0.05, 0.01, h], [0.01, 0.05, h]], np.float32)
    e = np.array([[0, 1], [1, 2], [2, 0], [3, 4], [4, 5], [5, 3]], np.uint32)
    f = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    return pos, e, f, np.ones(6, np.float32)


@pytest.mark.parametrize("h,held", [(0.012, True), (0.0149, True), (0.0151, False), (0.02, False)])
def test_oracle_vertex_over_triangle_closed_form(h, held):
    """Thickness 5 mm each and a 5 mm safety margin, no velocity: vertex 3 holds triangle B iff
    h < 0.015; vertices 4 and 5 lie outside B's footprint and B's vertices outside A's."""
    pos, e, f, inv = _hover(h)
    o = _oracle(pos, e, f, inv, thickness=0.005, precision="f64")
    o.set_gravity(0.0, 0.0, 0.0)
    o.set_damping(0.0)
    o.set_solver_iterations(0)
    o.step(1.0 / 60.0, 2)
    _, vt, _ = o.active_pairs()
    assert [tuple(r) for r in vt if r[0] == 3] == ([(3, 0, 1, 2)] if held else [])
    assert all(r[0] == 3 for r in vt)


def _check_lists(sim_pos, vt, ee, quota_vt, quota_ee, thick, margin):
    """Every held pair within the acceptance distance (no velocity: horizons are zero), at most
    `quota` pairs per vertex, each vertex's pairs in ascending distance."""
    p = sim_pos.astype(np.float64)
    check = 2 * thick + margin
    for lst, q, kind in ((vt, quota_vt, "vt"), (ee, quota_ee, "ee")):
        owners = lst[:, 0] if len(lst) else np.zeros(0, np.uint32)
        if len(owners):
            assert np.bincount(owners.astype(np.int64)).max() <= q
        prev_owner, prev_d = None, -1.0
        for r in lst:
            i = int(r[0])
            if kind == "vt":
                a, b, c = p[r[1]], p[r[2]], p[r[3]]
                # closest point by projection onto the plane when inside, else the nearest edge
                n = np.cross(b - a, c - a)
                n /= np.linalg.norm(n)
                q_ = p[i] - np.dot(p[i] - a, n) * n
                def inside(u, v, w, x):
                    return np.dot(np.cross(v - u, x - u), n) >= 0
                if inside(a, b, c, q_) and inside(b, c, a, q_) and inside(c, a, b, q_):
                    d = abs(np.dot(p[i] - a, n))
                else:
                    d = min(_seg_point(p[i], u, v) for u, v in ((a, b), (b, c), (c, a)))
            else:
                d = _seg_seg(p[i], p[r[1]], p[r[2]], p[r[3]])
            assert d < check + 1e-6, (kind, r, d)
            if i == prev_owner:
                assert d >= prev_d - 1e-6, (kind, r)
            prev_owner, prev_d = i, d


def _seg_point(x, a, b):
    t = np.clip(np.dot(x - a, b - a) / np.dot(b - a, b - a), 0.0, 1.0)
    return np.linalg.norm(x - (a + t * (b - a)))


def _seg_seg(p1, q1, p2, q2):
    """Exact distance between two segments: the interior stationary point if it lies in the unit
    square, else the best of the four edge cases (an endpoint against the other segment)."""
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, b = np.dot(d1, d1), np.dot(d2, d2), np.dot(d1, d2)
    c, f = np.dot(d1, r), np.dot(d2, r)
    cands = [_seg_point(p1, p2, q2), _seg_point(q1, p2, q2), _seg_point(p2, p1, q1), _seg_point(q2, p1, q1)]
    den = a * e - b * b
    if den > 1e-18:
        s_ = (b * f - c * e) / den
        t_ = (a * f - b * c) / den
        if 0.0 <= s_ <= 1.0 and 0.0 <= t_ <= 1.0:
            cands.append(np.linalg.norm((p1 + s_ * d1) - (p2 + t_ * d2)))
    return min(cands)


def test_oracle_lists_respect_distance_quota_and_order():
    pos, e, f, inv = cases.folded_cloth(8, 0.02, 0.004)
    o = _oracle(pos, e, f, inv, thickness=0.002, precision="f64")
    o.set_gravity(0.0, 0.0, 0.0)
    o.set_damping(0.0)
    o.set_solver_iterations(0)
    o.step(1.0 / 60.0, 2)
    # the lists of the collect at the frame's first substep: positions after predict of substep 1
    # equal the start positions (no gravity, no velocity)
    c, vt, ee = o.active_pairs()
    budget = min(max(8 * len(pos), 8192), 262144)
    q = min(max(budget // len(pos), 1), 8)
    assert len(vt) > 0 and len(ee) > 0
    _check_lists(pos, vt, ee, q, q, 0.002, 0.005)
    assert c[0] >= len(vt) and c[1] >= len(ee)


def test_oracle_reference_double_layer_repulsion():
    """test_pair_caching.py:108-141: a pinned floor layer and a free layer 3 mm above it."""
    p1, e1, f1, i1 = grid_cloth(3, 3, 0.05, 0.0, False)
    i1[:] = 0.0
    p2, e2, f2, i2 = grid_cloth(3, 3, 0.05, 0.003, False)
    n1 = len(p1)
    pos, e = np.vstack([p1, p2]), np.vstack([e1, e2 + n1])
    f, inv = np.vstack([f1, f2 + n1]), np.concatenate([i1, i2])
    o = _oracle(pos, e, f, inv, thickness=0.005)
    o.set_gravity(0.0, 0.0, 0.0)
    o.set_self_collision_options(1.0, 0.5)
    o.step(1.0 / 60.0, 4)
    up = _pos(o)[n1:]
    assert int((up[:, 2] < 0.0).sum()) == 0
    assert float(up[:, 2].mean()) > 0.006


def _swept(factory, margin_mode):
    """test_pair_caching.py:143-181: layers 30 mm apart closing at 2 m/s."""
    p1, e1, f1, _ = grid_cloth(3, 3, 0.05, 0.0, False)
    p2, e2, f2, _ = grid_cloth(3, 3, 0.05, 0.030, False)
    n1 = len(p1)
    pos, e, f = np.vstack([p1, p2]), np.vstack([e1, e2 + n1]), np.vstack([f1, f2 + n1])
    vel = np.zeros_like(pos)
    vel[:n1, 2] = 1.0
    vel[n1:, 2] = -1.0
    s = factory(pos, e, f, np.ones(len(pos), np.float32), thickness=0.005)
    s.set_enable_self_collision(True)
    s.set_self_collision_options(0.2, 0.2, True, False, 256)
    s.set_positions_and_velocities(pos, vel)
    if isinstance(s, O.OracleSim):
        s.set_pair_cache(True, final_fallback=False, quota_vt=min(max(32768 // len(pos), 1), 8),
                         quota_ee=min(max(32768 // len(pos), 1), 8), margin_mode=margin_mode)
    else:
        s.set_enable_pair_cache(True)
        s.set_pair_cache_options(32768, 32768, margin_mode, 0.005, 1.3, 0.02)
        s.set_enable_pair_cache_final_fallback(False)
    s.step(1.0 / 60.0, 10)
    c, vt, ee = _lists(s)
    return c


def test_oracle_reference_swept_collection():
    cf = _swept(O.OracleSim, 0)
    assert cf[0] == 0 and cf[1] == 0
    ca = _swept(O.OracleSim, 1)
    assert ca[0] > 0 and ca[1] > 0


def test_oracle_reference_quota_fairness():
    """test_pair_caching.py:236-259: a budget of 64 pairs (one slot per vertex)."""
    pos, e, f, inv = two_layers(5, 5, 0.05, 0.003)
    o = O.OracleSim(pos, e, f, inv, thickness=0.005)
    o.set_enable_self_collision(True)
    o.set_pair_cache(True, final_fallback=False, quota_vt=1, quota_ee=1)
    o.step(1.0 / 60.0, 1)
    _, vt, ee = o.active_pairs()
    owners = set(int(r[0]) for r in vt) | set(int(r[0]) for r in ee)
    assert len(owners) >= len(pos) // 2


# ------------------------------------------------------------------ GPU: the product against the oracle
def _pair_scene(scene):
    if scene == "two_layers":
        pos, e, f, inv = two_layers(6, 6, 0.04, 0.003)
        vel = np.zeros_like(pos)
        vel[len(pos) // 2:, 2] = -0.5
        return pos, e, f, inv, vel, 0.005
    pos, e, f, inv = cases.folded_cloth(16, 0.015, 0.004)
    vel = np.zeros_like(pos)
    vel[:, 2] = np.where(pos[:, 2] > 0.002, -0.3, 0.1)
    return pos, e, f, inv, vel, 0.003


@pytest.mark.gpu
@pytest.mark.parametrize("scene", ["two_layers", "folded"])
@pytest.mark.parametrize("margin_mode", [0, 1])
def test_gpu_pair_lists_bit_exact(require_gpu, scene, margin_mode):
    T = _T()
    pos, e, f, inv, vel, th = _pair_scene(scene)
    g = T.ClothSimulator(pos, e, f, inv, thickness=th, enable_pair_cache=True)
    g.set_enable_self_collision(True)
    g.set_enable_pair_cache_final_fallback(False)
    g.set_pair_cache_options(65536, 65536, margin_mode, 0.005, 1.3, 0.02)
    st = g._pair_state()
    o = O.OracleSim(pos, e, f, inv, thickness=th)
    o.set_enable_self_collision(True)
    o.set_pair_cache(True, final_fallback=False, quota_vt=int(st[2]), quota_ee=int(st[3]), margin_mode=margin_mode)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_positions_and_velocities(pos, vel)
        # two separate sheets count as layers; these lists are compared without the untangling branches
        s.set_self_collision_options(0.2, 0.2, True, False, 256)
    _say(f"lists {scene} margin_mode={margin_mode}: {len(pos)} vertices, quota {int(st[2])}")
    held = [0, 0]
    for fr in range(3):
        g.step(1.0 / 60.0, 3)
        o.step(1.0 / 60.0, 3)
        cg, vtg, eeg = _lists(g)
        held = [held[0] + len(vtg), held[1] + len(eeg)]
        co, vto, eeo = _lists(o)
        _say(f"  frame {fr}: {len(vtg)} V-T, {len(eeg)} E-E held; counters {cg.tolist()} / oracle {co.tolist()}; "
             f"max |hip - oracle| = {float(np.abs(_pos(g) - _pos(o)).max()):.2e}")
        assert np.array_equal(vtg, vto) and np.array_equal(eeg, eeo)
        assert np.array_equal(cg, co)
        assert float(np.abs(_pos(g) - _pos(o)).max()) == 0.0
    assert held[0] > 0 and held[1] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("fallback", [True, False])
def test_gpu_pair_cache_frames_bit_exact(require_gpu, fallback):
    T = _T()
    pos, e, f, inv = cases.folded_cloth(20, 0.015, 0.004)
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.003, self_collision_algorithm=1)
    g.set_enable_pair_cache_final_fallback(fallback)
    st = g._pair_state()
    o = O.OracleSim(pos, e, f, inv, thickness=0.003)
    o.set_pair_cache(True, final_fallback=fallback, quota_vt=int(st[2]), quota_ee=int(st[3]))
    for s in (g, o):
        s.set_enable_self_collision(True)
        s.set_gravity(0.0, 0.0, -9.81)
        s.add_sphere_collider([0.15, 0.15, -0.2], 0.19, 0.3, 0.0)
    assert g.get_self_collision_algorithm() == 1 and g.get_enable_pair_cache()
    _say(f"frames fallback={fallback}: {len(pos)} vertices")
    for fr in range(5):
        g.step(1.0 / 60.0, 4, 4)
        o.step(1.0 / 60.0, 4, 4)
        diff = float(np.abs(_pos(g) - _pos(o)).max())
        _say(f"  frame {fr}: max |hip - oracle| = {diff:.3e}")
        assert diff == 0.0


@pytest.mark.gpu
def test_gpu_reference_pair_caching_scenes(require_gpu):
    T = _T()
    pos, e, f, inv = grid_cloth()
    s = T.ClothSimulator(pos, e, f, inv)
    assert not s.get_enable_pair_cache()
    s.set_enable_pair_cache(True)
    assert s.get_enable_pair_cache()
    s.set_enable_pair_cache(False)
    assert not s.get_enable_pair_cache()
    s2 = T.ClothSimulator(pos, e, f, inv, enable_pair_cache=True)
    assert s2.get_enable_pair_cache() and s2.get_enable_pair_cache_final_fallback()
    s2.set_enable_self_collision(True)
    s2.step(1.0 / 60.0, 2)
    out = _pos(s2)
    assert np.isfinite(out).all() and (out[4:, 2] < 0.0).all()
    # double layer (test_pair_caching.py:108-141)
    p1, e1, f1, i1 = grid_cloth(3, 3, 0.05, 0.0, False)
    i1[:] = 0.0
    p2, e2, f2, i2 = grid_cloth(3, 3, 0.05, 0.003, False)
    n1 = len(p1)
    d = T.ClothSimulator(np.vstack([p1, p2]), np.vstack([e1, e2 + n1]), np.vstack([f1, f2 + n1]),
                         np.concatenate([i1, i2]), thickness=0.005, enable_pair_cache=True)
    d.set_gravity(0.0, 0.0, 0.0)
    d.set_enable_self_collision(True)
    d.set_enable_pair_cache_final_fallback(False)
    d.set_self_collision_options(relief_factor=1.0, max_displacement_ratio=0.5, enable_normal_untangling=False)
    d.step(1.0 / 60.0, 4)
    up = _pos(d)[n1:]
    assert int((up[:, 2] < 0.0).sum()) == 0 and float(up[:, 2].mean()) > 0.006
    # swept collection
    cf = _swept(T.ClothSimulator, 0)
    ca = _swept(T.ClothSimulator, 1)
    _say(f"swept: fixed {cf.tolist()}, auto {ca.tolist()}")
    assert cf[0] == 0 and cf[1] == 0 and ca[0] > 0 and ca[1] > 0
    # determinism and fairness
    pos, e, f, inv = two_layers(5, 5, 0.05, 0.003)

    def once(budget):
        t = T.ClothSimulator(pos, e, f, inv, thickness=0.005, enable_pair_cache=True)
        t.set_enable_self_collision(True)
        t.set_self_collision_options(0.2, 0.2, True, False, 256)
        if budget:
            t.set_pair_cache_options(budget, budget, 1, 0.005, 1.3, 0.02)
        t.set_enable_pair_cache_final_fallback(False)
        t.step(1.0 / 60.0, 2 if not budget else 1)
        return t.get_active_pairs()
    a, b = once(None), once(None)
    assert a[0] == b[0] and a[1] == b[1] and sorted(map(tuple, a[2])) == sorted(map(tuple, b[2]))
    assert len(a[2]) + len(a[3]) > 0
    # with one substep the last substep takes the direct path: the pair lists stay empty unless the
    # fallback is off, which it is here
    _, _, vt, ee = once(64)
    owners = set(r[0] for r in vt) | set(r[0] for r in ee)
    assert len(owners) >= len(pos) // 2
