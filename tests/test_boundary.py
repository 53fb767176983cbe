"""The drop-in boundary as the reference's own Python tests use it (tests/core/test_cloth_core.py:49-131,
tests/core/test_ffi_boundary.py:21-80), and the profiling trio (lib.rs:1975-1990)."""
import json
import os

import numpy as np
import pytest

import cases
from taremincloth_amd import taremin_cloth_core as T


def test_non_contiguous_pin_weights_reach_the_masses():
    """compute_combined_inv_masses with a strided pin_weights view (test_ffi_boundary.py:40-51)."""
    pw_raw = np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.0], np.float32)
    pw = pw_raw[::2]
    assert not pw.flags["C_CONTIGUOUS"]
    pos3 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], np.float32)
    inv = T.compute_combined_inv_masses(pos3, np.array([[0, 1, 2]], np.uint32), None, pw, 0.15, 0.0)
    assert float(inv[0]) == 0.0 and float(inv[1]) > 0.0 and float(inv[2]) == 0.0


@pytest.mark.gpu
def test_grid_drop_as_the_reference_test(require_gpu):
    """4x4 grid, two pinned top corners, stiffness 1e4, 60 frames of 30 substeps: finite, pins
    hold to 1e-5, the free corner falls, edge strain under 5%, reset restores to 1e-6."""
    nx = ny = 4
    dx = 0.1
    pos = np.array([[x * dx, y * dx, 0.0] for y in range(ny) for x in range(nx)], np.float32)
    inv = np.array([0.0 if (y == 0 and x in (0, nx - 1)) else 1.0 for y in range(ny) for x in range(nx)],
                   np.float32)
    edges = []
    for y in range(ny):
        for x in range(nx):
            i = y * nx + x
            if x + 1 < nx:
                edges.append([i, i + 1])
            if y + 1 < ny:
                edges.append([i, i + nx])
    edges = np.array(edges, np.uint32)
    sim = T.ClothSimulator(positions=pos, edges=edges, inv_masses=inv, layer_id=0, thickness=0.005,
                           stiffness=10000.0)
    assert sim.num_vertices == nx * ny
    assert sim.num_distance_constraints == len(edges)
    assert sim.num_sewing_constraints == 0
    for _ in range(60):
        sim.step(dt=1.0 / 60.0, substeps=30)
    out = np.zeros(nx * ny * 3, np.float32)
    sim.get_positions(out)
    p = out.reshape(-1, 3)
    assert np.all(np.isfinite(p))
    np.testing.assert_allclose(p[0], pos[0], atol=1e-5)
    np.testing.assert_allclose(p[nx - 1], pos[nx - 1], atol=1e-5)
    assert p[-1, 2] < -0.01
    strain = max(abs(np.linalg.norm(p[a] - p[b]) - np.linalg.norm(pos[a] - pos[b])) / np.linalg.norm(pos[a] - pos[b])
                 for a, b in edges)
    assert strain < 0.05, strain
    sim.reset()
    sim.get_positions(out)
    np.testing.assert_allclose(out.reshape(-1, 3), pos, atol=1e-6)


@pytest.mark.gpu
def test_non_contiguous_positions_are_copied(require_gpu):
    raw = np.zeros((10, 3), np.float32)
    raw[:, 0] = np.arange(10, dtype=np.float32) * 0.1
    sliced = raw[::2]
    assert not sliced.flags["C_CONTIGUOUS"]
    sim = T.ClothSimulator(positions=sliced, edges=np.array([[0, 1], [1, 2], [2, 3], [3, 4]], np.uint32))
    out = np.zeros(15, np.float32)
    sim.get_positions(out)
    np.testing.assert_allclose(out.reshape(-1, 3), sliced, atol=1e-6)
    sim.step(1.0 / 60.0, 5)
    sim.get_positions(out)
    assert np.all(np.isfinite(out))


@pytest.mark.gpu
def test_extreme_substeps_dt_and_negative_damping_stay_finite(require_gpu):
    pos = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0]], np.float32)
    sim = T.ClothSimulator(positions=pos, edges=np.array([[0, 1], [1, 2], [2, 0]], np.uint32),
                           faces=np.array([[0, 1, 2]], np.uint32))
    sim.set_damping(-10.0)
    sim.set_damping_all(-1.0, -1.0, -1.0, -1.0)
    sim.set_stiffness_all(0.0, 0.0, 0.0, 0.0)
    sim.step(1e-5, 1, 1)
    out = np.zeros(9, np.float32)
    sim.get_positions(out)
    assert np.all(np.isfinite(out))


@pytest.mark.gpu
def test_profiling_trio(require_gpu, tmp_path):
    """Pass times of the last frame under the reference's pass names, and its Chrome trace."""
    pos, e, f = cases.grid(12, 12, 0.02)
    sim = T.ClothSimulator(pos, e, f)
    sim.add_sphere_collider([0.1, 0.1, -0.2], 0.15, 0.3, 0.0)
    assert sim.take_profile() == []                 # off: nothing recorded
    sim.step(1.0 / 60.0, 4, 3)
    assert sim.take_profile() == []
    assert sim.set_profiling_enabled(True) is True
    sim.step(1.0 / 60.0, 4, 3)
    prof = sim.take_profile()
    names = [n for n, _ in prof]
    assert names.count("predict") == 4 and names.count("update_vel") == 4
    assert names.count("solver_iter") == 12 and names.count("collider_decoupled") == 4
    assert names.count("self_collision_Outer") == 4
    for child in ("hash_build", "sc_solve", "sc_solve_ee", "sc_apply"):
        assert names.count(child) == 4, child
    assert all(ms >= 0.0 and np.isfinite(ms) for _, ms in prof)
    assert sum(ms for n, ms in prof if n == "solver_iter") > 0.0
    # a scope covers its passes
    k = names.index("self_collision_Outer")
    assert prof[k][1] >= max(ms for n, ms in prof[k + 1:k + 5]) - 1e-3
    assert sim.take_profile() == prof              # no new frame: the same passes again
    path = tmp_path / "trace.json"
    sim.save_profile_trace(str(path))
    ev = json.load(open(path))["traceEvents"]
    assert [x["name"] for x in ev] == names and all(x["ph"] == "X" and x["dur"] >= 0 for x in ev)
    with pytest.raises(OSError):
        sim.save_profile_trace(os.path.join(str(tmp_path), "no_such_dir", "t.json"))
    # profiled frames run the same kernels in the same order as graph replay
    ref = T.ClothSimulator(pos, e, f)
    ref.add_sphere_collider([0.1, 0.1, -0.2], 0.15, 0.3, 0.0)
    ref.step(1.0 / 60.0, 4, 3)
    ref.step(1.0 / 60.0, 4, 3)
    a = np.zeros(pos.size, np.float32)
    b = np.zeros(pos.size, np.float32)
    sim.get_positions(a)
    ref.get_positions(b)
    assert np.array_equal(a, b)
    assert sim.set_profiling_enabled(False) is False
    assert sim.take_profile() == []
