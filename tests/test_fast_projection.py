"""Opt-in fast projection on the per-iteration tile path: refined hardware reciprocals in place of
the IEEE square root and divisions (no reference counterpart).

It is not bit-identical to the reference arithmetic, so it is held to the positional tolerance of
the parity tests instead: within 1e-5 m of the fp32 oracle after 1 and 100 substeps of the 32x32
poked grid (10 iterations per substep), and of the exact path on damped and undamped drapes over a
sphere; its distance from the fp64 oracle is printed beside the exact path's.  Switching it off
restores the exact path bit for bit (captured graphs are rebuilt)."""
import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(msg):
    print(f"[fast-projection] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _gpu(require_gpu):
    yield


def _T():
    from taremincloth_amd import taremin_cloth_core as T
    return T


def _pos(sim):
    out = np.empty(3 * sim.num_vertices, np.float32)
    sim.get_positions(out)
    return out.reshape(-1, 3)


def _config1(factory, **kw):
    pos, e, f, inv = cases.config1()
    s = factory(pos, e, f, inv, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0, **kw)
    s.set_compliances_raw(1e-8, 1e-6, 1e-5)
    s.set_gravity(0.0, 0.0, -9.81)
    s.set_enable_self_collision(False)
    s.set_solver_iterations(10)
    return s


def _product(fast):
    g = _config1(_T().ClothSimulator)
    g.set_solver_path(2)
    g.set_fast_projection(fast)
    assert g.get_fast_projection() == fast
    return g


def test_config1_one_and_hundred_substeps_within_tolerance():
    dt_sub = (1.0 / 60.0) / 10.0
    errs = {}
    for n_sub in (1, 100):
        gf, ge = _product(True), _product(False)
        o32 = _config1(O.OracleSim)
        o64 = _config1(lambda *a, **k: O.OracleSim(*a, precision="f64", **k))
        if n_sub == 1:
            for s in (gf, ge, o32, o64):
                s.step_single_substep(dt_sub)
        else:
            for _ in range(10):
                for s in (gf, ge, o32, o64):
                    s.step(1.0 / 60.0, 10, 10)
        p32 = o32.get_positions().reshape(-1, 3)
        p64 = o64.state(0)
        errs[n_sub] = (float(np.abs(_pos(gf) - p32).max()), float(np.abs(_pos(gf).astype(np.float64) - p64).max()),
                       float(np.abs(_pos(ge) - p32).max()), float(np.abs(_pos(ge).astype(np.float64) - p64).max()))
        _say(f"{n_sub} substeps: fast vs fp32 oracle {errs[n_sub][0]:.2e}, vs fp64 {errs[n_sub][1]:.2e}; "
             f"exact vs fp32 {errs[n_sub][2]:.2e}, vs fp64 {errs[n_sub][3]:.2e}")
        assert errs[n_sub][0] <= TOL, errs[n_sub]
        assert errs[n_sub][2] < TOL
        # the fast path is no farther from the FP64 solution than the exact fp32 path plus the tolerance
        assert errs[n_sub][1] <= errs[n_sub][3] + TOL
    assert errs[1][0] > 0.0 or errs[100][0] > 0.0  # it is a different arithmetic


@pytest.mark.parametrize("damp", [False, True])
def test_drape_on_sphere_within_tolerance_of_exact(damp):
    T = _T()
    pos, e, f, inv = cases.drape_grid(96, 0.004, 0.14, seed=3)
    out = {}
    for fast in (False, True):
        g = T.ClothSimulator(pos, e, f, inv, thickness=0.004)
        g.set_compliances_raw(1e-8, 1e-6, 1e-5)
        g.set_enable_self_collision(False)
        if damp:
            g.set_damping_all(2.0, 1.0, 3.0, 4.0)
        g.add_sphere_collider([0.0, 0.0, 0.0], 0.12, 0.3, 0.0)
        g.set_solver_path(2)
        g.set_fast_projection(fast)
        for _ in range(4):
            g.step(1.0 / 60.0, 5, 7)
        out[fast] = _pos(g)
        assert g.solver_info()["path"] == 2
    d = float(np.abs(out[True] - out[False]).max())
    _say(f"96x96 drape, damp={damp}: max |fast - exact| = {d:.2e} m after 20 substeps")
    assert 0.0 < d <= TOL


def test_switching_off_restores_the_exact_path():
    """A simulator that ran fast frames, reset and switched back, replays the exact trajectory bit for bit."""
    T = _T()
    pos, e, f, inv = cases.drape_grid(64, 0.005, 0.1, seed=8)
    ref = T.ClothSimulator(pos, e, f, inv, thickness=0.004)
    sw = T.ClothSimulator(pos, e, f, inv, thickness=0.004)
    for g in (ref, sw):
        g.set_compliances_raw(1e-8, 1e-6, 1e-5)
        g.add_sphere_collider([0.0, 0.0, 0.0], 0.1, 0.3, 0.0)
        g.set_solver_path(2)
    sw.set_fast_projection(True)
    for _ in range(2):
        sw.step(1.0 / 60.0, 4, 6)
    sw.reset()
    sw.set_fast_projection(False)
    for _ in range(3):
        ref.step(1.0 / 60.0, 4, 6)
        sw.step(1.0 / 60.0, 4, 6)
    assert np.array_equal(_pos(ref), _pos(sw))
