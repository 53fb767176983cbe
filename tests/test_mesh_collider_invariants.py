"""Closed forms that pin the triangle-mesh collider (crates/cloth_core/src/shaders/collision.wgsl:
463-542 for the per-triangle test, 83-108 for the contact response).

The reference's golden frames hold primitive colliders only, so the mesh path has no fixture of its
own. A large flat square of two triangles under a small cloth reduces it to cases whose answers
follow from the shader's own definitions:

- resting contact: a vertex within the contact distance of a face ends exactly at that distance
  (cloth thickness + triangle thickness) above it, the answer the plane collider gives;
- sliding with friction mu: the tangential step shrinks by (1 - mu), as on the plane collider;
- a vertex swept through the face within one substep (the segment test) ends on the front side at
  the contact distance, without lateral drift;
- single-sided faces recover a vertex slightly behind them to the front side; double-sided faces
  push it to the near side instead;
- beyond the contact distance nothing moves.

The cloth is offset so that no vertex sits over the square's diagonal, where the closest feature
is an edge (no recovery there, by the reference's face-only rule). The HIP path is checked against
the oracle on every case.
"""
import numpy as np
import pytest

import cases
from oracle import oracle as O

TH = 0.005          # cloth thickness
DT = 0.005          # substep
DAMPING = 1.0       # the simulator's default air damping (velocity scale 1 - damping * dt)


def _cloth():
    pos, e, f = cases.grid(8, 8, 0.05, "xy")
    pos = (pos - pos.mean(0) + np.array([0.013, -0.007, 0.0])).astype(np.float32)
    return pos, e, f


QUAD = np.array([[[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [1.0, 1.0, 0.0]],
                 [[-1.0, -1.0, 0.0], [1.0, 1.0, 0.0], [-1.0, 1.0, 0.0]]], np.float32)


def run(make_sim, z, vel=(0.0, 0.0, 0.0), single=False, mesh=True, plane=False, friction=0.0, tri_thickness=0.0):
    pos, e, f = _cloth()
    pos[:, 2] = z
    sim = make_sim(pos, e, f)
    sim.set_gravity(0.0, 0.0, 0.0)
    sim.set_enable_self_collision(False)
    if mesh:
        sim.set_mesh_collider_triangles(QUAD, friction=friction, thickness=tri_thickness, restitution=0.0,
                                        single_sided=single)
    if plane:
        sim.add_plane_collider([0.0, 0.0, 0.0], [0.0, 0.0, 1.0], friction, 0.0)
    v = np.tile(np.asarray(vel, np.float32), (len(pos), 1))
    if isinstance(sim, O.OracleSim):
        sim.set_state(v=v)
    else:
        sim.set_positions_and_velocities(pos, v)
    sim.step_single_substep(DT)
    out = np.zeros(pos.size, np.float32)
    sim.get_positions(out)
    return pos, out.reshape(-1, 3)


def oracle_sim(pos, e, f):
    return O.OracleSim(pos, e, f, None, thickness=TH)


def test_oracle_resting_contact_sits_at_the_contact_distance():
    _, out = run(oracle_sim, 0.004)
    assert np.abs(out[:, 2] - TH).max() <= 1e-7
    _, ref = run(oracle_sim, 0.004, mesh=False, plane=True)
    assert np.abs(out - ref).max() <= 1e-6
    _, thick = run(oracle_sim, 0.006, tri_thickness=0.002)     # target = cloth + triangle thickness
    assert np.abs(thick[:, 2] - (TH + 0.002)).max() <= 1e-7


def test_oracle_friction_scales_the_tangential_step():
    pos, out = run(oracle_sim, 0.004, vel=(1.0, 0.0, 0.0), friction=0.4)
    step = 1.0 * (1.0 - DAMPING * DT) * DT
    assert np.abs((out[:, 0] - pos[:, 0]) - (1.0 - 0.4) * step).max() <= 1e-6
    assert np.abs(out[:, 2] - TH).max() <= 1e-6
    _, ref = run(oracle_sim, 0.004, vel=(1.0, 0.0, 0.0), mesh=False, plane=True, friction=0.4)
    assert np.abs(out - ref).max() <= 1e-6


def test_oracle_swept_vertex_ends_in_front():
    """10 m/s downwards from 1 cm: the prediction is 4 cm below the face."""
    pos, free = run(oracle_sim, 0.01, vel=(0.0, 0.0, -10.0), mesh=False)
    assert free[:, 2].max() < -0.03
    _, out = run(oracle_sim, 0.01, vel=(0.0, 0.0, -10.0))
    assert np.abs(out[:, 2] - TH).max() <= 1e-7
    assert np.array_equal(out[:, :2], pos[:, :2])


def test_oracle_single_sided_recovers_and_double_sided_pushes_to_the_near_side():
    _, single = run(oracle_sim, -0.002, single=True)
    assert np.abs(single[:, 2] - TH).max() <= 1e-6
    _, double = run(oracle_sim, -0.002, single=False)
    assert np.abs(double[:, 2] + TH).max() <= 1e-6


def test_oracle_no_contact_beyond_the_contact_distance():
    pos, out = run(oracle_sim, 0.02)
    assert np.array_equal(out, pos)


CASES = [dict(z=0.004), dict(z=0.006, tri_thickness=0.002), dict(z=0.004, vel=(1.0, 0.0, 0.0), friction=0.4),
         dict(z=0.01, vel=(0.0, 0.0, -10.0)), dict(z=-0.002, single=True), dict(z=-0.002), dict(z=0.02)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_flat_square_matches_oracle(require_gpu, case):
    from taremincloth_amd import taremin_cloth_core as T

    def gpu_sim(pos, e, f):
        return T.ClothSimulator(pos, e, f, None, thickness=TH)

    _, g = run(gpu_sim, **case)
    _, o = run(oracle_sim, **case)
    assert np.array_equal(g, o), np.abs(g - o).max()
