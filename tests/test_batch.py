"""Batches of independent cloth instances in one simulator (no reference counterpart).

Each instance is built and coloured on its own; the batch holds their disjoint union and
advances all of them with one launch per colour or per iteration.  Constraints of one colour
share no vertex, so every instance must follow exactly the Gauss-Seidel order it has alone:
the tests require each instance's slice of the batch to equal, bit for bit, the same cloth
simulated alone on the GPU, and to stay within 1e-5 m of the fp32 oracle run of that cloth.
"""
import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[batch] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _gpu(require_gpu):
    yield


def _T():
    from taremincloth_amd import taremin_cloth_core as T
    return T


def _configure(sim, iters=20):
    sim.set_compliances_raw(1e-8, 1e-6, 1e-5)
    sim.set_solver_iterations(iters)


def _run(sim, frames, substeps=10):
    for _ in range(frames):
        sim.step(1.0 / 60.0, substeps)


def _alone(factory, inst, colliders=False):
    s = factory(inst["positions"], inst["edges"], inst["faces"], inst["inv_masses"])
    _configure(s)
    s.set_enable_self_collision(False)
    if colliders:
        s.add_sphere_collider([0.3, 0.05, -0.15], 0.12, 0.3, 0.0)
    return s


def test_instance_ranges_and_errors():
    T = _T()
    inst = cases.batch_instances(4, 5, 12, seed=3)
    b = T.ClothBatch(inst)
    assert b.num_instances() == 4
    first = 0
    for k, d in enumerate(inst):
        a, n = b.instance_range(k)
        assert (a, n) == (first, len(d["positions"]))
        first += n
    assert b.num_vertices == first
    with pytest.raises(ValueError):
        b.instance_range(4)
    bad = [dict(inst[0]), dict(inst[1])]
    bad[1]["edges"] = np.array([[0, 10_000]], np.uint32)
    with pytest.raises(ValueError):
        T.ClothBatch(bad)


@pytest.mark.parametrize("path", [0, 1])
def test_batch_equals_instances_alone(path):
    """Per-iteration (path 0) and per-colour (path 1) schedules: each slice of the batch equals the
    instance simulated alone, bit for bit, after 3 frames of 10 substeps x 20 iterations."""
    T = _T()
    inst = cases.batch_instances(6, 4, 40, seed=11)
    _say(f"batch of {len(inst)} path {path}")
    b = T.ClothBatch(inst)
    _configure(b)
    b.set_solver_path(path)
    b.add_sphere_collider([0.3, 0.05, -0.15], 0.12, 0.3, 0.0)
    _run(b, 3)
    for k, d in enumerate(inst):
        s = _alone(T.ClothSimulator, d, colliders=True)
        s.set_solver_path(path)
        _run(s, 3)
        alone = np.empty(3 * len(d["positions"]), np.float32)
        s.get_positions(alone)
        got = b.get_instance_positions(k)
        assert np.array_equal(got, alone.reshape(-1, 3)), f"instance {k}: {np.abs(got - alone.reshape(-1, 3)).max()}"


def test_batch_matches_oracle_instances():
    T = _T()
    inst = cases.batch_instances(5, 4, 24, seed=5)
    _say("batch vs oracle")
    b = T.ClothBatch(inst)
    _configure(b)
    b.add_sphere_collider([0.3, 0.05, -0.15], 0.12, 0.3, 0.0)
    _run(b, 2)
    for k, d in enumerate(inst):
        o = _alone(O.OracleSim, d, colliders=True)
        _run(o, 2)
        diff = np.abs(b.get_instance_positions(k) - o.get_positions().reshape(-1, 3)).max()
        assert diff < 1e-5, f"instance {k}: {diff:.3e}"


def test_batch_autotune_and_reset():
    T = _T()
    inst = cases.batch_instances(8, 8, 64, seed=2)
    _say("batch autotune")
    b = T.ClothBatch(inst)
    _configure(b)
    _run(b, 1)
    before = np.empty(3 * b.num_vertices, np.float32)
    b.get_positions(before)
    b.reset()
    tune = b.autotune(1.0 / 600.0)
    assert tune is not None
    _run(b, 1)
    after = np.empty_like(before)
    b.get_positions(after)
    assert np.array_equal(before, after)
