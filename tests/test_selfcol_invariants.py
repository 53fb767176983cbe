"""Self-collision invariants from the reference's own tests, on the oracle (CPU) and on the HIP path.

The reference checks action/reaction of the DIRECT self-collision solve with two facing 3x3 sheets
0.02 m apart and a combined thickness of 0.03 m (tests/core/test_self_collision_symmetry.py:8-60):
after one substep without gravity the two sheets move apart, with equal and opposite summed
displacements for equal masses (relative error < 5 %, :64-118) and equal and opposite momentum for a
2:1 mass ratio, the lighter sheet moving further (relative error < 10 %, :120-177).  The golden
frames hold no self-contacts, so these invariants are what pins the oracle's V-V / V-T solve and the
apply pass (crates/cloth_core/src/shaders/self_collision.wgsl:317-681,
self_collision_apply.wgsl:57-177); the HIP path is then checked against the oracle on the same case.

The edge-edge solve (self_collision_ee_direct.wgsl:266-776) has no reference fixture of its own.
Two isolated edges without faces exercise it alone: their endpoints stay farther apart than the
contact distance (no vertex-vertex contact) and without faces there is no vertex-triangle test.
Crossing at right angles 2 cm apart with a contact distance of 3 cm, they must separate along
the line between their closest points, with equal and opposite momentum for any mass ratio,
split between an edge's endpoints by the lever rule (3:1 for a crossing at a quarter of its
length). Farther apart than the contact distance, nothing moves. An edge swept through the other
within one substep must end on its own side (the continuous test).
"""
import numpy as np
import pytest

from oracle import oracle as O


def two_planes(nx=3, ny=3, z_top=0.01, z_bottom=-0.01, inv_top=1.0, inv_bottom=1.0, dx=0.05):
    pos, inv, edges, faces = [], [], [], []
    for z, w in ((z_bottom, inv_bottom), (z_top, inv_top)):
        base = len(pos)
        for y in range(ny):
            for x in range(nx):
                pos.append([x * dx, y * dx, z])
                inv.append(w)
        for y in range(ny):
            for x in range(nx):
                i = base + y * nx + x
                if x + 1 < nx:
                    edges.append([i, i + 1])
                if y + 1 < ny:
                    edges.append([i, i + nx])
                if x + 1 < nx and y + 1 < ny:
                    faces.append([i, i + 1, i + nx + 1])
                    faces.append([i, i + nx + 1, i + nx])
    return (np.array(pos, np.float32), np.array(edges, np.uint32), np.array(faces, np.uint32),
            np.array(inv, np.float32), nx * ny)


def one_substep(make_sim, inv_top, inv_bottom):
    pos, edges, faces, inv, top = two_planes(inv_top=inv_top, inv_bottom=inv_bottom)
    sim = make_sim(pos, edges, faces, inv)
    sim.set_gravity(0.0, 0.0, 0.0)
    sim.set_enable_self_collision(True)
    sim.set_self_collision_options(1.0, 1.0, True, False, 128)
    sim.step_single_substep(0.005)
    out = np.zeros(pos.size, np.float32)
    sim.get_positions(out)
    disp = out.reshape(-1, 3) - pos
    return out.reshape(-1, 3), disp[:top, 2].sum(dtype=np.float64), disp[top:, 2].sum(dtype=np.float64)


def oracle_sim(pos, edges, faces, inv):
    return O.OracleSim(pos, edges, faces, inv, thickness=0.015, stiffness=10000.0)


def check_equal_mass(bottom, top):
    assert bottom < 0.0 and top > 0.0
    assert abs(bottom + top) / max(abs(bottom), abs(top), 1e-6) < 0.05


def check_unequal_mass(bottom, top):
    pb, pt = bottom / 1.0, top / 0.5          # momentum = displacement / inverse mass
    assert abs(bottom) > abs(top)
    assert abs(pb + pt) / max(abs(pb), abs(pt), 1e-6) < 0.10


def test_oracle_equal_mass_symmetry():
    _, b, t = one_substep(oracle_sim, 1.0, 1.0)
    check_equal_mass(b, t)
    assert t > 1e-4                            # a real push, not a no-op


def test_oracle_unequal_mass_momentum():
    _, b, t = one_substep(oracle_sim, 0.5, 1.0)
    check_unequal_mass(b, t)


def test_oracle_push_vanishes_without_overlap():
    """Sheets 0.04 m apart with a combined thickness of 0.03 m: nothing moves."""
    pos, edges, faces, inv, top = two_planes(z_top=0.02, z_bottom=-0.02)
    sim = oracle_sim(pos, edges, faces, inv)
    sim.set_gravity(0.0, 0.0, 0.0)
    sim.set_self_collision_options(1.0, 1.0, True, False, 128)
    sim.step_single_substep(0.005)
    out = np.zeros(pos.size, np.float32)
    sim.get_positions(out)
    assert np.array_equal(out.reshape(-1, 3), pos)


@pytest.mark.gpu
@pytest.mark.parametrize("inv_top", [1.0, 0.5])
def test_gpu_two_sheet_push_matches_oracle(require_gpu, inv_top):
    from taremincloth_amd import taremin_cloth_core as T

    def gpu_sim(pos, edges, faces, inv):
        return T.ClothSimulator(pos, edges, faces, inv, thickness=0.015, stiffness=10000.0)

    g, gb, gt = one_substep(gpu_sim, inv_top, 1.0)
    o, ob, ot = one_substep(oracle_sim, inv_top, 1.0)
    assert np.max(np.abs(g - o)) <= 1e-6
    (check_equal_mass if inv_top == 1.0 else check_unequal_mass)(gb, gt)


def two_edges(d=0.02, length=0.1, frac=0.5, inv_a=1.0, inv_b=1.0):
    """Edge A = (0, 1) along x at z = d/2, crossed at fraction `frac` of its length by edge
    B = (2, 3) along y at z = -d/2."""
    pos = np.array([[-frac * length, 0.0, d / 2], [(1.0 - frac) * length, 0.0, d / 2],
                    [0.0, -length / 2, -d / 2], [0.0, length / 2, -d / 2]], np.float32)
    return (pos, np.array([[0, 1], [2, 3]], np.uint32), np.zeros((0, 3), np.uint32),
            np.array([inv_a, inv_a, inv_b, inv_b], np.float32))


def ee_substep(make_sim, d=0.02, frac=0.5, inv_a=1.0, inv_b=1.0, vz_a=0.0, self_collision=True):
    pos, edges, faces, inv = two_edges(d, frac=frac, inv_a=inv_a, inv_b=inv_b)
    vel = np.zeros_like(pos)
    vel[:2, 2] = vz_a
    sim = make_sim(pos, edges, faces, inv)
    sim.set_gravity(0.0, 0.0, 0.0)
    sim.set_enable_self_collision(self_collision)
    sim.set_self_collision_options(1.0, 1.0, True, False, 128)
    if vz_a != 0.0:
        if isinstance(sim, O.OracleSim):
            sim.set_state(v=vel)
        else:
            sim.set_positions_and_velocities(pos, vel)
    sim.step_single_substep(0.005)
    out = np.zeros(pos.size, np.float32)
    sim.get_positions(out)
    return pos, out.reshape(-1, 3)


def test_oracle_ee_push_along_the_separation_with_equal_and_opposite_momentum():
    pos, out = ee_substep(oracle_sim)
    disp = (out - pos).astype(np.float64)
    assert np.all(disp[:, :2] == 0.0)                  # along the line between the closest points
    assert disp[0, 2] > 1e-4 and disp[2, 2] < -1e-4    # apart: A up, B down
    assert disp[0, 2] == disp[1, 2] and disp[2, 2] == disp[3, 2]
    assert abs(disp[:2, 2].sum() + disp[2:, 2].sum()) <= 1e-6 * abs(disp[:2, 2].sum())


def test_oracle_ee_lever_rule():
    """B crosses A at a quarter of its length: A's near end takes three times the far end's push."""
    pos, out = ee_substep(oracle_sim, frac=0.25)
    disp = (out - pos).astype(np.float64)
    assert abs(disp[0, 2] / disp[1, 2] - 3.0) < 2e-3
    assert disp[2, 2] == disp[3, 2] < 0.0
    assert abs(disp[:2, 2].sum() + disp[2:, 2].sum()) <= 1e-3 * abs(disp[:2, 2].sum())


def test_oracle_ee_unequal_mass_momentum():
    """A twice as heavy as B: equal and opposite momentum, B moves further."""
    pos, out = ee_substep(oracle_sim, inv_a=0.5, inv_b=1.0)
    disp = (out - pos).astype(np.float64)
    pa, pb = disp[:2, 2].sum() / 0.5, disp[2:, 2].sum() / 1.0
    assert abs(disp[2, 2]) > abs(disp[0, 2]) > 0.0
    assert abs(pa + pb) <= 1e-6 * abs(pa)


def test_oracle_ee_no_push_without_overlap():
    pos, out = ee_substep(oracle_sim, d=0.04)
    assert np.array_equal(out, pos)


def test_oracle_ee_swept_edge_stays_on_its_side():
    """A moves down at 12 m/s: in 5 ms it would pass 2 cm through B. Without self-collision it
    ends below B; with it, above B, with equal and opposite corrections."""
    pos, free = ee_substep(oracle_sim, d=0.04, vz_a=-12.0, self_collision=False)
    assert free[:2, 2].mean() < free[2:, 2].mean()
    _, out = ee_substep(oracle_sim, d=0.04, vz_a=-12.0)
    assert out[:2, 2].mean() > out[2:, 2].mean()
    corr = (out - free).astype(np.float64)
    assert abs(corr[:2, 2].sum() + corr[2:, 2].sum()) <= 1e-4 * abs(corr[:2, 2].sum())


@pytest.mark.gpu
@pytest.mark.parametrize("case", [dict(), dict(frac=0.25), dict(inv_a=0.5), dict(d=0.04),
                                  dict(d=0.04, vz_a=-12.0)])
def test_gpu_two_edges_match_oracle(require_gpu, case):
    from taremincloth_amd import taremin_cloth_core as T

    def gpu_sim(pos, edges, faces, inv):
        return T.ClothSimulator(pos, edges, faces, inv, thickness=0.015, stiffness=10000.0)

    _, g = ee_substep(gpu_sim, **case)
    _, o = ee_substep(oracle_sim, **case)
    assert np.array_equal(g, o), np.abs(g - o).max()
