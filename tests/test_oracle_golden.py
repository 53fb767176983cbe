"""The CPU oracle against the reference's own golden fixtures (tests/fixtures/golden_master).

Pins predict + colouring order + distance/shear/bending + decoupled
sphere/capsule/plane resolve + (inactive) self-collision + update_vel of the
oracle: one frame (20 substeps x 2 iterations) from the fixture input must match
the reference output within 1e-5 m max L2 (test_single_step_exact.py:45-130).
"""
import numpy as np
import pytest

import cases
from oracle import oracle as O


@pytest.mark.parametrize("case", ["case1", "case2", "case3"])
def test_single_step_exact(case):
    c = cases.GOLDEN_CASES[case]
    step = cases.load_npz(c["step"])
    pos_in = step.get("pos_in")
    sim, _, _ = cases.make_golden_sim(O.OracleSim, case, positions=pos_in)
    sim.step(1.0 / 60.0, 20, 2)
    out = sim.get_positions().reshape(-1, 3)
    diff = float(np.linalg.norm(out - step["pos_out"], axis=1).max())
    assert diff < 1e-5, f"{case}: {diff * 1e6:.3f} um"


def test_single_step_case3_bit_exact():
    """Case 3 reproduces the reference's fp32 output bit for bit."""
    sim, _, _ = cases.make_golden_sim(O.OracleSim, "case3")
    sim.step(1.0 / 60.0, 20, 2)
    out = sim.get_positions().reshape(-1, 3)
    assert np.array_equal(out, cases.load_npz("step_case3_f1.npz")["pos_out"])


def _edge_strain(pos, edges, rest):
    d = pos[edges[:, 0].astype(np.int64)] - pos[edges[:, 1].astype(np.int64)]
    return float((np.linalg.norm(d, axis=1) / np.maximum(rest, 1e-6)).max())


@pytest.mark.parametrize("case,short_frames,short_tol,com_tol,strain", [
    ("case1", 10, 1e-3, 0.010, 1.15),
    ("case2", 10, 1e-3, 0.015, 1.15),
    ("case3", 10, 1e-3, 0.025, 1.20),
    ("case4", 5, 5e-3, 0.020, 1.25),
])
def test_golden_trajectory(case, short_frames, short_tol, com_tol, strain):
    """25-frame trajectories with the reference's bounds (test_golden_regression.py:66-304):
    short-term max vertex error, centre of mass, edge strain, pins, non-penetration."""
    sim, d, pos0 = cases.make_golden_sim(O.OracleSim, case)
    gp = d["positions"]
    edges = d["edges"]
    rest = np.linalg.norm(pos0[edges[:, 0].astype(np.int64)] - pos0[edges[:, 1].astype(np.int64)], axis=1)
    _, _, inv = cases.golden_inputs(case)
    pinned = inv == 0.0
    for fi in range(1, len(gp)):
        sim.step(1.0 / 60.0, 20, 2)
        a = sim.get_positions().reshape(-1, 3)
        assert np.isfinite(a).all()
        diff = float(np.linalg.norm(a - gp[fi], axis=1).max())
        if fi <= short_frames:
            assert diff < short_tol, f"{case} frame {fi}: {diff * 1e3:.4f} mm"
        com = float(np.linalg.norm(a.mean(0) - gp[fi].mean(0)))
        assert com < com_tol, f"{case} frame {fi}: centre of mass off by {com * 1e3:.3f} mm"
        assert _edge_strain(a, edges, rest) < strain
        if pinned.any():
            assert float(np.linalg.norm(a[pinned] - pos0[pinned], axis=1).max()) < 1e-5
        if case == "case2":
            assert np.linalg.norm(a - np.array([0, 0, 0.16], np.float32), axis=1).min() >= 0.14 - 0.002
        if case == "case3":
            assert np.linalg.norm(a - np.array([0, 0, 0.3], np.float32), axis=1).min() >= 0.12 - 0.003
            assert a[:, 2].min() >= -0.002
        if case == "case4":
            assert np.linalg.norm(a - np.array([0, 0, 0.2], np.float32), axis=1).min() >= 0.20 - 0.003


def test_case1_whole_trajectory_within_microns():
    """Case 1 (pinned curtain, no contacts) stays within 2 um of the reference for all 25 frames."""
    sim, d, _ = cases.make_golden_sim(O.OracleSim, "case1")
    worst = 0.0
    for fi in range(1, 26):
        sim.step(1.0 / 60.0, 20, 2)
        worst = max(worst, float(np.linalg.norm(sim.get_positions().reshape(-1, 3) - d["positions"][fi], axis=1).max()))
    assert worst < 2e-6


def test_contact_trajectories_are_chaotic():
    """Why late frames of cases 2-4 are bounded by the reference's loose tolerances only:
    a 1e-7 m perturbation of the input grows by more than four orders of magnitude by
    frame 25 on case 2 (sphere contact)."""
    d, pos, _ = cases.golden_inputs("case2")
    a, _, _ = cases.make_golden_sim(O.OracleSim, "case2")
    jit = pos + np.float32(1e-7) * np.random.default_rng(0).standard_normal(pos.shape).astype(np.float32)
    b, _, _ = cases.make_golden_sim(O.OracleSim, "case2", positions=jit)
    for _ in range(25):
        a.step(1.0 / 60.0, 20, 2)
        b.step(1.0 / 60.0, 20, 2)
    growth = np.abs(a.get_positions() - b.get_positions()).max()
    assert growth > 1e-3


def test_fp32_vs_fp64_oracle_one_frame():
    """State the fp32 error against the FP64 twin on case 1 (one frame): below 1 um."""
    s32, _, _ = cases.make_golden_sim(lambda *a, **k: O.OracleSim(*a, precision="f32", **k), "case1")
    s64, _, _ = cases.make_golden_sim(lambda *a, **k: O.OracleSim(*a, precision="f64", **k), "case1")
    s32.step(1.0 / 60.0, 20, 2)
    s64.step(1.0 / 60.0, 20, 2)
    err = np.abs(s32.state(0) - s64.state(0)).max()
    assert err < 1e-6


def test_openmp_oracle_equals_the_serial_oracle():
    """The multithreaded build of the fp32 restatement (the bench's CPU baseline) splits each
    colour and each per-vertex pass over threads; the result is the serial one bit for bit."""
    from oracle import oracle as O
    pos, edges, faces = cases.grid(24, 24, 0.02)
    pos = pos + np.array([-0.23, -0.23, 0.2], np.float32)
    sims = []
    for prec in ("f32", "f32_omp"):
        s = O.OracleSim(pos, edges, faces, precision=prec)
        s.set_enable_self_collision(False)
        s.set_compliances_raw(1e-8, 1e-6, 1e-5)
        s.add_sphere_collider([0.0, 0.0, 0.0], 0.15, 0.3, 0.0)
        s.add_capsule_collider([-0.2, 0.1, 0.05], [0.2, 0.1, 0.05], 0.05, 0.3, 0.0)
        s.set_solver_iterations(8)
        for _ in range(12):
            s.step(1.0 / 60.0, 10)
        sims.append(s)
    assert np.array_equal(sims[0].state(0), sims[1].state(0))
    assert np.array_equal(sims[0].state(2), sims[1].state(2))
