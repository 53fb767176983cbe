"""Collider-edge collision (edge_collision.wgsl:153-276; edges from mesh.rs:1492-1567, grid from
spatial_hash.rs:473-529; dispatched after the final pin, dispatch.rs:252-278).

CPU tests pin the oracle's restatement:
  * extraction of the unit cube: 18 unique edges (12 box edges between two faces, 6 face
    diagonals inside a face), box-edge normals at 45 degrees between the two face normals,
    diagonal normals equal to the face normal, thickness the largest, friction averaged,
    single-sided only when every triangle is;
  * the grid holds each edge in the buckets of both end points' cells;
  * closed forms on a ridge (two steep triangles meeting in a horizontal edge) under a cloth edge
    crossing it: above the ridge within the target distance both ends rise to the target
    (two-sided and one-sided), below it a two-sided ridge pushes the edge down to the target while
    a one-sided ridge pulls it back out over the top (min(target - s, 2 target)), out of reach
    nothing moves, and the margin scale and offset widen the target.
GPU tests: the product's collider edges equal the oracle's, and a cloth draped over a box with
edge collision on equals the FP32 oracle bit for bit over several frames.
"""
import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O

CUBE_V = np.array([[-0.5, -0.5, -0.5], [0.5, -0.5, -0.5], [0.5, 0.5, -0.5], [-0.5, 0.5, -0.5],
                   [-0.5, -0.5, 0.5], [0.5, -0.5, 0.5], [0.5, 0.5, 0.5], [-0.5, 0.5, 0.5]], np.float32)
CUBE_T = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                   [2, 3, 7], [2, 7, 6], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], np.int64)


def _say(msg):
    print(f"[edge-collision] {msg}", flush=True)
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


def _stick(precision="f64"):
    """Two vertices, one edge, nothing else active."""
    pos = np.array([[-0.5, 0.0, 0.0], [0.5, 0.0, 0.0]], np.float32)
    o = O.OracleSim(pos, np.array([[0, 1]], np.uint32), None, np.ones(2, np.float32), thickness=0.005,
                    precision=precision)
    o.set_gravity(0.0, 0.0, 0.0)
    o.set_damping(0.0)
    o.set_solver_iterations(0)
    o.set_enable_self_collision(False)
    return o


def ridge(flip=False):
    """Two steep triangles meeting in the edge (0, -1, 0)-(0, 1, 0); with flip the faces point down."""
    t1 = [[0.0, -1.0, 0.0], [0.0, 1.0, 0.0], [-0.1, 0.0, -1.0]]
    t2 = [[0.0, 1.0, 0.0], [0.0, -1.0, 0.0], [0.1, 0.0, -1.0]]
    tris = np.array([t1, t2], np.float32)
    if flip:
        tris = tris[:, ::-1].copy()
    return tris


# ------------------------------------------------------------------ CPU: the oracle pinned
@pytest.mark.parametrize("single", [False, True])
def test_oracle_cube_edges_closed_form(single):
    o = _stick()
    tris = CUBE_V[CUBE_T]
    attrs = np.array([[0.2 + 0.1 * (k % 3), 0.001 * (1 + k % 4), 0.0, 1.0 if (single or k % 2) else 0.0]
                      for k in range(12)], np.float32)
    o.set_mesh_collider_triangles(tris, attributes=attrs)
    e, fl = o.collider_edges()
    assert len(e) == 18
    d = e[:, 4:7] - e[:, 0:3]
    ln = np.linalg.norm(d, axis=1)
    box = np.isclose(ln, 1.0)
    assert box.sum() == 12 and np.isclose(ln[~box], np.sqrt(2.0)).all()
    # box edge normals: the two adjacent face normals (axis directions) summed and normalised
    mid = (e[:, 0:3] + e[:, 4:7]) * 0.5
    for k in np.flatnonzero(box):
        want = np.sign(mid[k]) * (np.abs(mid[k]) > 0.25)
        want = want / np.linalg.norm(want)
        np.testing.assert_allclose(e[k, 8:11], want, atol=1e-6)
    for k in np.flatnonzero(~box):  # diagonal inside one face: that face's normal
        axis = np.argmax(np.abs(mid[k]))
        want = np.zeros(3)
        want[axis] = np.sign(mid[k][axis])
        np.testing.assert_allclose(e[k, 8:11], want, atol=1e-6)
    # the triangles sharing each edge, from the triangle list
    key = lambda a, b: tuple(sorted([tuple(np.round(a, 6)), tuple(np.round(b, 6))]))
    share = {}
    for k, t in enumerate(CUBE_T):
        for a, b in ((0, 1), (1, 2), (2, 0)):
            share.setdefault(key(CUBE_V[t[a]], CUBE_V[t[b]]), []).append(k)
    for r, f in zip(e, fl):
        ks = share[key(r[0:3], r[4:7])]
        assert r[3] == attrs[ks, 1].max()
        assert f == int(all(attrs[ks, 3] == 1.0))


def test_oracle_grid_holds_both_end_cells():
    o = _stick()
    o.set_mesh_collider_triangles(CUBE_V[CUBE_T] * np.float32(0.1))
    # the grid itself is internal; an edge lying on the cloth edge's path must be found from either end:
    # checked through the closed forms below, and here through the edge count
    e, _ = o.collider_edges()
    assert len(e) == 18


def _ridge_case(z, single, scale=1.0, offset=0.0, flip=False, precision="f64"):
    o = _stick(precision)
    p = np.array([[-0.5, 0.0, z], [0.5, 0.0, z]], np.float64)
    o.set_state(x=p, p=p, v=np.zeros((2, 3)))
    o.set_mesh_collider_triangles(ridge(flip), friction=0.0, thickness=0.005, single_sided=single)
    o.set_edge_collision(True, scale, offset)
    o.substep_parts(1.0 / 600.0, True, False, False)
    return o.state(1)


@pytest.mark.parametrize("single", [False, True])
def test_oracle_ridge_above_rises_to_target(single):
    out = _ridge_case(0.006, single)
    target = np.float64(np.float32(0.005)) * 2
    np.testing.assert_allclose(out[:, 2], target, atol=2e-6)
    np.testing.assert_allclose(out[:, :2], [[-0.5, 0.0], [0.5, 0.0]], atol=1e-9)


def test_oracle_ridge_below_two_sided_pushes_down():
    out = _ridge_case(-0.004, False)
    np.testing.assert_allclose(out[:, 2], -0.01, atol=2e-6)


def test_oracle_ridge_below_one_sided_pulls_back_over():
    out = _ridge_case(-0.004, True)
    # normal +z (sum of the two face normals), pen = min(target - (-0.004), 2 target) = 0.014
    np.testing.assert_allclose(out[:, 2], -0.004 + 0.014, atol=2e-6)
    far = _ridge_case(-0.009, True)  # pen = 0.019 (the 2 target cap binds only out of reach here)
    np.testing.assert_allclose(far[:, 2], -0.009 + 0.019, atol=2e-6)


def test_oracle_ridge_out_of_reach_and_margins():
    out = _ridge_case(0.011, False)
    np.testing.assert_allclose(out[:, 2], 0.011, atol=1e-9)
    wide = _ridge_case(0.011, False, scale=1.5, offset=0.002)  # target = 0.01 * 1.5 + 0.002
    np.testing.assert_allclose(wide[:, 2], 0.017, atol=2e-6)


# ------------------------------------------------------------------ GPU: the product against the oracle
@pytest.mark.gpu
def test_gpu_collider_edges_equal_oracle(require_gpu):
    T = _T()
    v, f = cases.icosphere(2, radius=0.2, noise_amp=0.2, seed=1)
    tris = v[f.astype(np.int64)].astype(np.float32)
    pos, e, fc, inv = cases.drape_grid(12, 0.02, 0.3, seed=1)
    g = T.ClothSimulator(pos, e, fc, inv)
    o = O.OracleSim(pos, e, fc, inv)
    attrs = np.stack([np.linspace(0.1, 0.6, len(tris)), np.linspace(0.001, 0.004, len(tris)),
                      np.zeros(len(tris)), (np.arange(len(tris)) % 3 == 0).astype(np.float32)], 1).astype(np.float32)
    for s in (g, o):
        s.set_mesh_collider_triangles(tris, attributes=attrs)
    ge, gf = g.collider_edges()
    oe, of = o.collider_edges()
    _say(f"{len(ge)} collider edges")
    assert np.array_equal(ge, oe) and np.array_equal(gf, of)
    g.clear_colliders()
    assert g.collider_edge_count() == 0


@pytest.mark.gpu
@pytest.mark.parametrize("single", [False, True])
def test_gpu_edge_collision_drape_bit_exact(require_gpu, single):
    T = _T()
    box = CUBE_V * np.float32(0.2) + np.float32([0.0, 0.0, 0.0])
    tris = box[CUBE_T].astype(np.float32)
    pos, e, f, inv = cases.drape_grid(40, 0.012, 0.13, seed=2)
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.004)
    o = O.OracleSim(pos, e, f, inv, thickness=0.004)
    g.set_enable_edge_collision(True)
    g.set_edge_margin_scale(1.2)
    g.set_edge_margin_offset(0.001)
    o.set_edge_collision(True, 1.2, 0.001)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_compliances_raw(1e-8, 1e-6, 1e-5)
        s.set_enable_self_collision(False)
        s.set_mesh_collider_triangles(tris, friction=0.3, thickness=0.003, single_sided=single)
    _say(f"drape over a box, single_sided={single}: {len(pos)} vertices")
    moved = False
    for fr in range(8):
        g.step(1.0 / 60.0, 5, 6)
        o.step(1.0 / 60.0, 5, 6)
        diff = float(np.abs(_pos(g) - _pos(o)).max())
        _say(f"  frame {fr}: max |hip - oracle| = {diff:.3e}")
        assert diff == 0.0
    # the edge pass acts: the same run without it ends elsewhere
    ref = T.ClothSimulator(pos, e, f, inv, thickness=0.004)
    ref.set_gravity(0.0, 0.0, -9.81)
    ref.set_compliances_raw(1e-8, 1e-6, 1e-5)
    ref.set_enable_self_collision(False)
    ref.set_mesh_collider_triangles(tris, friction=0.3, thickness=0.003, single_sided=single)
    for _ in range(8):
        ref.step(1.0 / 60.0, 5, 6)
    d = float(np.abs(_pos(ref) - _pos(g)).max())
    _say(f"  without the edge pass the cloth ends up to {d * 1e3:.3f} mm elsewhere")
    assert d > 1e-5
