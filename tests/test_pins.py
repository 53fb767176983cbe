"""Dynamic pins (pin.wgsl:55-92 at the start of every iteration and after the collider,
dispatch.rs:176-181, 238-250) and the inverse-mass policy around them (simulation/mod.rs:2768-2905).

CPU tests pin the oracle's restatement:
  * closed form: a free vertex with no active constraint, pulled by a pin of weight w, ends one
    substep at t + (x0 - t) r^(N+1) with r = alpha / (inv_m + alpha), inv_m = (1 - w) inv_m0,
    alpha = 1e-3 (1 - w) / w / dt^2 (N iterations plus the final pin), in FP64 and FP32;
  * a weight-1 pin snaps the vertex onto the target and leaves it at rest; weights 0 and NaN
    change nothing;
  * the reference's chain tests (simulation/tests.rs:556-625: full hold within 1 cm of the target,
    half weight strictly between rest and target) and its inverse-mass tests (tests.rs:2009-2031,
    1070-1122: set_positions keeps base inverse masses and pins; large tables and batch release).
GPU tests: the product bit-identical to the FP32 oracle with pins of every kind on a draped
cloth (per-colour path and one-launch-per-iteration path, graph replay, pins moved, released
and re-set between frames, with a collider and with self-collision), the reference scenes, and
the inverse masses the solver reads.
"""
import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O


def _say(msg):
    print(f"[pins] {msg}", flush=True)
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


def chain(factory, **kw):
    """The reference's 3-vertex chain 0-1-2, ends held (inverse mass 0), no gravity
    (simulation/tests.rs:556-584: tension 1000, compression 1000, shear 5000, bending 0)."""
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [2.0, 0.0, 0.0]], np.float32)
    e = np.array([[0, 1], [1, 2]], np.uint32)
    s = factory(pos, e, None, np.array([0.0, 1.0, 0.0], np.float32), thickness=0.005, stiffness=1000.0,
                compression_stiffness=1000.0, shear_stiffness=5000.0, bending_stiffness=0.0, **kw)
    s.set_gravity(0.0, 0.0, 0.0)
    return s


def free_pair(precision):
    """Two vertices whose only constraints are inactive (compliance >= 1e9), no gravity or damping."""
    pos = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 0.1]], np.float32)
    s = O.OracleSim(pos, np.array([[0, 1]], np.uint32), None, np.array([1.0, 1.0], np.float32),
                    precision=precision)
    s.set_compliances_raw(1e10, 1e10, 1e10)
    s.set_gravity(0.0, 0.0, 0.0)
    s.set_damping(0.0)
    s.set_enable_self_collision(False)
    return s, pos


# ------------------------------------------------------------------ CPU: the oracle pinned
@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 2e-6)])
@pytest.mark.parametrize("w", [0.2, 0.5, 0.9, 0.999])
@pytest.mark.parametrize("iters", [1, 3, 8])
def test_oracle_soft_pin_closed_form(precision, tol, w, iters):
    s, pos = free_pair(precision)
    s.set_solver_iterations(iters)
    t = np.array([0.1, 0.4, -0.25], np.float32)
    s.set_pin(1, t, w)
    dt = 1.0 / 600.0
    s.step_single_substep(dt)
    inv_m = np.float64(np.float32(1.0) * (np.float32(1.0) - np.float32(w)))
    c_ref = np.float64(np.float32(1e-3))  # the kernel's constant is the f32 value nearest 1e-3
    alpha = c_ref * (1.0 - np.float64(np.float32(w))) / np.float64(np.float32(w)) / (np.float64(np.float32(dt)) ** 2)
    r = alpha / (inv_m + alpha)
    want = t.astype(np.float64) + (pos[1].astype(np.float64) - t) * r ** (iters + 1)
    got = s.state(0)[1]
    assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max()), (got, want)
    np.testing.assert_allclose(s.state(2)[1], (want - pos[1]) / np.float64(np.float32(dt)), rtol=1e-4, atol=1e-6)
    assert np.array_equal(s.state(0)[0], pos[0].astype(np.float64))  # the unpinned vertex stays
    np.testing.assert_array_equal(s.inv_masses(), np.array([1.0, np.float32(inv_m)], np.float32))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_oracle_hard_pin_snaps_and_rests(precision):
    s, pos = free_pair(precision)
    s.set_solver_iterations(2)
    t = np.array([0.5, 0.5, 0.5], np.float32)
    s.set_pin(1, t, 1.0)
    assert s.inv_masses()[1] == 0.0
    for _ in range(3):
        s.step_single_substep(1.0 / 600.0)
        assert np.array_equal(s.state(0)[1], t.astype(np.float64))
    assert np.array_equal(s.state(2)[1], np.zeros(3))
    s.release_pin(1)
    assert s.inv_masses()[1] == 1.0


@pytest.mark.parametrize("w", [0.0, float("nan"), -2.0])
def test_oracle_inert_weights(w):
    a, _ = free_pair("f32")
    b, _ = free_pair("f32")
    for s in (a, b):
        s.set_gravity(0.0, 0.0, -9.81)
    a.set_pin(1, [5.0, 5.0, 5.0], w)
    assert a.inv_masses()[1] == 1.0
    for _ in range(4):
        a.step_single_substep(1.0 / 600.0)
        b.step_single_substep(1.0 / 600.0)
    assert np.array_equal(a.state(0), b.state(0))


def test_oracle_reference_chain_full_hold():
    s = chain(O.OracleSim)
    s.set_pin(1, [1.0, 0.5, 0.0], 1.0)
    for _ in range(30):
        s.step(1.0 / 60.0, 20)
    y1 = float(_pos(s)[1, 1])
    assert abs(y1 - 0.5) < 1e-2, y1


def test_oracle_reference_chain_partial_blend():
    s = chain(O.OracleSim)
    s.set_pin(1, [1.0, 0.5, 0.0], 0.5)
    for _ in range(150):
        s.step(1.0 / 60.0, 20)
    y1 = float(_pos(s)[1, 1])
    assert 1e-3 < y1 < 0.5 - 1e-3, y1


def quad(factory, **kw):
    """make_quad_mesh of simulation/tests.rs: a unit quad of two triangles at z = 0.5."""
    pos = np.array([[0, 0, 0.5], [1, 0, 0.5], [1, 1, 0.5], [0, 1, 0.5]], np.float32)
    e = np.array([[0, 1], [1, 2], [2, 3], [3, 0], [0, 2]], np.uint32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return factory(pos, e, f, None, **kw)


def test_oracle_set_positions_keeps_inv_masses_and_pins():
    """simulation/tests.rs:2009-2031."""
    s = quad(O.OracleSim)
    s.set_vertex_inv_masses([0], [0.5])
    s.set_pin(1, [1.0, 0.0, 0.5], 0.75)
    s.set_positions_and_velocities(np.array([[0, 0, 0.6], [1, 0, 0.6], [1, 1, 0.6], [0, 1, 0.6]], np.float32))
    inv = s.inv_masses()
    assert abs(inv[0] - 0.5) < 1e-6 and abs(inv[1] - 0.25) < 1e-6 and abs(inv[2] - 1.0) < 1e-6
    assert abs(s.state(0)[0, 2] - 0.6) < 1e-6
    np.testing.assert_array_equal(s.get_vertex_inv_masses(), np.array([0.5, 1.0, 1.0, 1.0], np.float32))
    s.reset()  # the pins go, the mesh's inverse masses come back; the base values stay
    np.testing.assert_array_equal(s.inv_masses(), np.ones(4, np.float32))
    np.testing.assert_array_equal(s.get_vertex_inv_masses(), np.array([0.5, 1.0, 1.0, 1.0], np.float32))


# ------------------------------------------------------------------ GPU: the product against the oracle
def _pinned_drape(factory, n=48, **kw):
    pos, e, f, inv = cases.drape_grid(n, 0.01, 0.3, seed=5)
    s = factory(pos, e, f, inv, thickness=0.004, stiffness=1000.0, bending_stiffness=10.0, **kw)
    s.set_compliances_raw(1e-8, 1e-6, 1e-5)
    s.set_gravity(0.0, 0.0, -9.81)
    return s, pos


def _pin_program(n, pos):
    """Frame -> list of ("set", idx, targets, weights) / ("release", idx) operations."""
    c = [0, n - 1, n * (n - 1), n * n - 1]
    mid = n * (n // 2) + n // 2
    prog = {0: [("set", c, pos[c] + np.float32([0.0, 0.0, 0.02]), [1.0, 1.0, 1.0, 1.0]),
                ("set", [mid, mid + 1], pos[[mid, mid + 1]] + np.float32([0.01, 0.0, 0.05]), [0.5, 0.9]),
                ("set", [n + 1, n + 2, n + 3], pos[[n + 1, n + 2, n + 3]], [0.0, float("nan"), 0.3])],
            2: [("set", c[:2], pos[c[:2]] + np.float32([-0.02, 0.0, 0.04]), [1.0, 0.7])],
            3: [("release", [mid, n + 3])],
            5: [("set", [mid], pos[[mid]] - np.float32([0.0, 0.0, 0.03]), [0.25])],
            6: [("clear",)],
            7: [("set", c[2:], pos[c[2:]], [1.0, 1.0])]}
    return prog


def _apply(s, ops):
    for op in ops:
        if op[0] == "set":
            s.set_pins_batch(op[1], np.asarray(op[2], np.float32), op[3])
        elif op[0] == "release":
            s.release_pins_batch(op[1])
        else:
            s.clear_pins()


@pytest.mark.gpu
@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("scene", ["sphere", "self"])
def test_gpu_pins_bit_exact_to_oracle(require_gpu, path, scene):
    T = _T()
    n = 48
    g, pos = _pinned_drape(T.ClothSimulator, n)
    o, _ = _pinned_drape(O.OracleSim, n)
    for s in (g, o):
        if scene == "sphere":
            s.set_enable_self_collision(False)
            s.add_sphere_collider([0.0, 0.0, 0.12], 0.15, 0.3, 0.0)
        else:
            s.set_enable_self_collision(True)
    g.set_solver_path(path)
    _say(f"pins path={path} scene={scene}: {n}x{n}")
    prog = _pin_program(n, pos)
    for fr in range(9):
        for s in (g, o):
            _apply(s, prog.get(fr, []))
        g.step(1.0 / 60.0, 4, 5)
        o.step(1.0 / 60.0, 4, 5)
        diff = float(np.abs(_pos(g) - _pos(o)).max())
        gi, oi = g.get_pins(), sorted(o._pins)
        _say(f"  frame {fr}: {len(oi)} pins, max |hip - oracle| = {diff:.3e} m")
        assert diff == 0.0, diff
        assert list(gi[0]) == oi
        np.testing.assert_array_equal(g.device_inv_masses(), o.inv_masses())
        np.testing.assert_array_equal(g.get_vertex_inv_masses(), o.get_vertex_inv_masses())
    assert g.solver_info()["path"] == path


@pytest.mark.gpu
def test_gpu_reference_chain_scenes(require_gpu):
    T = _T()
    for w, frames in ((1.0, 30), (0.5, 150)):
        g = chain(T.ClothSimulator)
        o = chain(O.OracleSim)
        for s in (g, o):
            s.set_pin(1, [1.0, 0.5, 0.0], w)
        for _ in range(frames):
            g.step(1.0 / 60.0, 20)
            o.step(1.0 / 60.0, 20)
        y1 = float(_pos(g)[1, 1])
        _say(f"chain w={w}: y1 = {y1:.6f}, oracle {float(_pos(o)[1, 1]):.6f}")
        assert np.array_equal(_pos(g), _pos(o))
        if w == 1.0:
            assert abs(y1 - 0.5) < 1e-2
        else:
            assert 1e-3 < y1 < 0.5 - 1e-3


@pytest.mark.gpu
def test_gpu_inv_mass_policy(require_gpu):
    """simulation/tests.rs:2009-2031 and 1070-1122 on the product."""
    T = _T()
    s = quad(T.ClothSimulator)
    s.set_vertex_inv_masses(np.array([0], np.uint32), np.array([0.5], np.float32))
    s.set_pin(1, [1.0, 0.0, 0.5], 0.75)
    s.set_positions_and_velocities(np.array([[0, 0, 0.6], [1, 0, 0.6], [1, 1, 0.6], [0, 1, 0.6]], np.float32))
    inv = s.device_inv_masses()
    assert abs(inv[0] - 0.5) < 1e-6 and abs(inv[1] - 0.25) < 1e-6 and abs(inv[2] - 1.0) < 1e-6
    assert abs(_pos(s)[0, 2] - 0.6) < 1e-6
    np.testing.assert_array_equal(s.get_vertex_inv_masses(), np.array([0.5, 1.0, 1.0, 1.0], np.float32))
    s.reset()
    np.testing.assert_array_equal(s.device_inv_masses(), np.ones(4, np.float32))
    assert s.get_pins()[0].size == 0
    np.testing.assert_array_equal(s.get_vertex_inv_masses(), np.array([0.5, 1.0, 1.0, 1.0], np.float32))
    # a 100 x 100 sheet: 9,000 pins set and released in one batch each, then more than the
    # first table holds, then cleared; the sheet still steps
    pos, e, f = cases.grid(100, 100, 0.01)
    big = T.ClothSimulator(pos, e, f, None, thickness=0.005)
    big.set_enable_self_collision(False)
    idx = np.arange(9000, dtype=np.uint32)
    big.set_pins_batch(idx, np.tile(np.float32([1.0, 2.0, 3.0]), (9000, 1)), [1.0] * 9000)
    assert big.get_pins()[0].size == 9000
    big.step(1.0 / 60.0, 2, 2)
    assert np.array_equal(_pos(big)[:9000], np.tile(np.float32([1.0, 2.0, 3.0]), (9000, 1)))
    big.release_pins_batch(idx)
    assert big.get_pins()[0].size == 0
    np.testing.assert_array_equal(big.device_inv_masses(), np.ones(10000, np.float32))
    big.set_pins_batch(np.arange(10000, dtype=np.uint32), pos, [0.5] * 10000)
    assert big.get_pins()[0].size == 10000
    big.step(1.0 / 60.0, 2, 2)
    big.clear_pins()
    assert big.get_pins()[0].size == 0
    big.step(1.0 / 60.0, 2, 2)
    assert np.isfinite(_pos(big)).all()
