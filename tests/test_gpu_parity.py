"""Parity of the HIP substep (through the C ABI) with the oracle and the reference's golden fixtures.

Tolerances: fp32 positions within 1e-5 m max-abs of the fp32 oracle after 1 and
100 substeps; spatial-hash keys, cell starts and sorted indices bit-exact;
golden single-step fixtures within 1e-5 m max L2 (the reference's own bound).
"""
import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O
from taremincloth_amd import taremin_cloth_core as T

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(msg):
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _gpu(require_gpu):
    yield


def _pos(sim):
    out = np.empty(3 * sim.num_vertices, np.float32)
    sim.get_positions(out)
    return out.reshape(-1, 3)


@pytest.mark.parametrize("case", ["case1", "case2", "case3"])
def test_golden_single_step(case):
    _say(f"golden single step {case}")
    c = cases.GOLDEN_CASES[case]
    step = cases.load_npz(c["step"])
    sim, _, _ = cases.make_golden_sim(T.ClothSimulator, case, positions=step.get("pos_in"))
    assert sim.step(1.0 / 60.0, 20, 2) == 20
    diff = float(np.linalg.norm(_pos(sim) - step["pos_out"], axis=1).max())
    assert diff < TOL, f"{case}: {diff * 1e6:.3f} um"


@pytest.mark.parametrize("case", ["case1", "case4"])
def test_golden_trajectory_matches_oracle(case):
    _say(f"golden trajectory vs oracle {case}")
    g, d, _ = cases.make_golden_sim(T.ClothSimulator, case)
    o, _, _ = cases.make_golden_sim(O.OracleSim, case)
    for fi in range(1, 9):
        g.step(1.0 / 60.0, 20, 2)
        o.step(1.0 / 60.0, 20, 2)
        diff = float(np.abs(_pos(g) - o.get_positions().reshape(-1, 3)).max())
        assert diff < TOL, f"{case} frame {fi}: {diff:.3e}"
        assert float(np.linalg.norm(_pos(g) - d["positions"][fi], axis=1).max()) < 1e-3


def _config1_pair(iters=10):
    pos, e, f, inv = cases.config1()
    sims = []
    for factory in (T.ClothSimulator, O.OracleSim, lambda *a, **k: O.OracleSim(*a, precision="f64", **k)):
        s = factory(pos, e, f, inv, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0)
        s.set_compliances_raw(1e-8, 1e-6, 1e-5)
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_enable_self_collision(False)
        s.set_solver_iterations(iters)
        sims.append(s)
    return sims


def test_config1_one_and_hundred_substeps():
    """32x32 poked grid, raw compliances 1e-8 / 1e-6 / 1e-5, 10 iterations per substep:
    after 1 substep and after 100 substeps (10 frames x 10 substeps)."""
    _say("config1 1 substep")
    g, o32, o64 = _config1_pair()
    dt_sub = (1.0 / 60.0) / 10.0
    for s in (g, o32, o64):
        s.step_single_substep(dt_sub)
    e1 = float(np.abs(_pos(g) - o32.get_positions().reshape(-1, 3)).max())
    e1_64 = float(np.abs(_pos(g).astype(np.float64) - o64.state(0)).max())
    assert e1 < TOL, e1
    _say("config1 100 substeps")
    g, o32, o64 = _config1_pair()
    for _ in range(10):
        for s in (g, o32, o64):
            s.step(1.0 / 60.0, 10, 10)
    e100 = float(np.abs(_pos(g) - o32.get_positions().reshape(-1, 3)).max())
    e100_64 = float(np.abs(_pos(g).astype(np.float64) - o64.state(0)).max())
    _say(f"config1 max-abs vs fp32 oracle: 1 substep {e1:.2e}, 100 substeps {e100:.2e}; "
         f"vs fp64 oracle: {e1_64:.2e}, {e100_64:.2e}")
    assert e100 < TOL, e100
    assert np.abs(g.velocities() - o32.state(2)).max() < 1e-2


def test_lambdas_and_residual_match_oracle():
    _say("lambdas and residual")
    g, o32, _ = _config1_pair(iters=4)
    dt_sub = (1.0 / 60.0) / 10.0
    for _ in range(3):
        g.step_single_substep(dt_sub)
        o32.step_single_substep(dt_sub)
    g.substep_parts(dt_sub, True, True, False)
    o32.substep_parts(dt_sub, True, True, False)
    ld_g, lb_g = g.lambdas()
    ld_o, lb_o = o32.lambdas()
    scale = max(np.abs(ld_o).max(), 1e-12)
    assert np.abs(ld_g - ld_o).max() / scale < 1e-3
    rg, mg = g.residual(dt_sub)
    ro, mo = o32.residual(dt_sub)
    assert abs(rg - ro) <= 1e-3 * ro + 1e-12, (rg, ro)


def test_damped_residual_matches_oracle():
    """With constraint damping (5, 5, 5, 0.5: the add-on's defaults) the residual carries the
    damping term gamma grad C . (p - x): GPU equals the oracle (whose damped residual is pinned by
    closed forms and by a single projection driving it to zero, tests/test_residual_pins.py)."""
    _say("damped residual")
    for iters in (4, 32):
        g, o32, _ = _config1_pair(iters=iters)
        dt_sub = (1.0 / 60.0) / 10.0
        for s in (g, o32):
            s.set_damping_all(5.0, 5.0, 5.0, 0.5)
            for _ in range(3):
                s.step_single_substep(dt_sub)
            s.substep_parts(dt_sub, True, True, False)
        rg, mg = g.residual(dt_sub)
        ro, mo = o32.residual(dt_sub)
        assert abs(rg - ro) <= 1e-3 * ro + 1e-12, (iters, rg, ro)
        assert abs(mg - mo) <= 1e-3 * mo + 1e-12, (iters, mg, mo)


def _rand_points(n, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)) * scale - scale / 2).astype(np.float32)


@pytest.mark.parametrize("n,seed", [(500, 1), (4096, 2), (40000, 3)])
def test_hash_tables_bit_exact(n, seed):
    """Vertex and edge hash keys, cell starts and sorted indices are bit-identical to the oracle."""
    _say(f"hash tables n={n}")
    pos, e, f = cases.grid(int(np.sqrt(n)), int(np.sqrt(n)), 0.01)
    sim = T.ClothSimulator(pos, e, f)
    pts = _rand_points(len(pos), seed)
    sim.set_positions_and_velocities(pts)
    ht = sim.hash_tables()
    keys, starts, srt = O.hash_points(pts, ht["cell_size"], ht["table_size"])
    assert np.array_equal(ht["keys"], keys)
    assert np.array_equal(ht["starts"], starts)
    assert np.array_equal(ht["sorted"], srt)
    ed = ht["edges"].astype(np.int64)
    mid = (pts[ed[:, 0]] + pts[ed[:, 1]]) * np.float32(0.5)
    k2, s2, o2 = O.hash_points(mid, ht["cell_size"], ht["table_size"])
    assert np.array_equal(ht["edge_keys"], k2)
    assert np.array_equal(ht["edge_starts"], s2)
    assert np.array_equal(ht["edge_sorted"], o2)


def test_hash_neighbour_pairs_match_brute_force():
    """All vertex pairs within one cell size are found among the 27 neighbour cells
    (canonical i<j pair lists), the same set brute force O(V^2) gives."""
    _say("hash neighbour pairs")
    pos, e, f = cases.grid(24, 24, 0.01)
    sim = T.ClothSimulator(pos, e, f)
    pts = _rand_points(len(pos), 7, scale=0.12)
    sim.set_positions_and_velocities(pts)
    ht = sim.hash_tables()
    cell, table = ht["cell_size"], ht["table_size"]
    keys, starts, srt = ht["keys"], ht["starts"], ht["sorted"]
    cells = np.floor(pts / np.float32(cell)).astype(np.int64)
    found = set()
    for i in range(len(pts)):
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    c = cells[i] + (dx, dy, dz)
                    h = ((((int(c[0]) & 0xFFFFFFFF) * 73856093) & 0xFFFFFFFF) ^
                         (((int(c[1]) & 0xFFFFFFFF) * 19349663) & 0xFFFFFFFF) ^
                         (((int(c[2]) & 0xFFFFFFFF) * 83492791) & 0xFFFFFFFF)) % table
                    for j in srt[starts[h]:starts[h + 1]]:
                        if j != i and (cells[j] == c).all() and np.linalg.norm(pts[i] - pts[j]) < cell:
                            found.add((min(i, int(j)), max(i, int(j))))
    d = np.linalg.norm(pts[:, None, :] - pts[None, :, :], axis=2)
    ii, jj = np.nonzero(np.triu(d < cell, 1))
    assert found == set(zip(ii.tolist(), jj.tolist()))


def test_self_collision_folded_cloth_matches_oracle():
    """A cloth folded onto itself: V-V, V-T and E-E contacts in the first substeps."""
    _say("self collision folded cloth")
    pos, e, f, inv = cases.folded_cloth()
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.003, stiffness=1000.0, bending_stiffness=10.0)
    o = O.OracleSim(pos, e, f, inv, thickness=0.003, stiffness=1000.0, bending_stiffness=10.0)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_self_collision_options(1.0, 0.5, True, False, 256)
    for step in range(3):
        g.step(1.0 / 60.0, 5, 2)
        o.step(1.0 / 60.0, 5, 2)
        diff = float(np.abs(_pos(g) - o.get_positions().reshape(-1, 3)).max())
        assert diff < TOL, f"frame {step}: {diff:.3e}"
    st = o.contact_stats()
    assert st["vv"] + st["vt_prox"] + st["vt_ccd"] > 0 and st["ee_prox"] + st["ee_ccd"] > 0, st


@pytest.mark.parametrize("cap,cell", [(2, 0.02), (3, 0.03), (256, 0.05)])
def test_self_collision_search_cap_matches_oracle(cap, cell):
    """The E-E search stops an edge after `cap` box-test passes, counted through the 27 cells in
    order. Small caps cut most edges short; the large cells fill the buckets, so that wavefronts
    queue more than 64 candidate pairs at once."""
    _say(f"self collision search cap {cap}, cell {cell}")
    pos, e, f, inv = cases.folded_cloth()
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.003, stiffness=1000.0, bending_stiffness=10.0)
    o = O.OracleSim(pos, e, f, inv, thickness=0.003, stiffness=1000.0, bending_stiffness=10.0)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_self_collision_options(1.0, 0.5, True, False, cap)
        s.set_hash_cell_size(cell)
    for step in range(3):
        g.step(1.0 / 60.0, 5, 2)
        o.step(1.0 / 60.0, 5, 2)
        diff = float(np.abs(_pos(g) - o.get_positions().reshape(-1, 3)).max())
        assert diff < TOL, f"frame {step}: {diff:.3e}"
    st = o.contact_stats()
    assert st["ee_prox"] + st["ee_ccd"] > 0, st


def test_drape_on_primitives_with_self_collision_matches_oracle():
    """64x64 cloth at 4 mm dropping onto a sphere and a capsule with friction 0.3,
    self-collision on (a reduced version of the 256^2 configuration)."""
    _say("drape on primitives")
    pos, e, f, inv = cases.drape_grid(64, 0.004, 0.12, seed=5)
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.002)
    o = O.OracleSim(pos, e, f, inv, thickness=0.002)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.add_sphere_collider([0.0, 0.0, 0.0], 0.1, 0.3, 0.0)
        s.add_capsule_collider([-0.1, 0.05, 0.05], [0.1, 0.05, 0.05], 0.03, 0.3, 0.0)
        s.set_solver_iterations(20)
    for fr in range(6):
        g.step(1.0 / 60.0, 10, 20)
        o.step(1.0 / 60.0, 10, 20)
        diff = float(np.abs(_pos(g) - o.get_positions().reshape(-1, 3)).max())
        assert diff < TOL, f"frame {fr}: {diff:.3e}"


def test_reset_and_set_positions():
    _say("reset / set positions")
    pos, e, f, inv = cases.config1()
    g = T.ClothSimulator(pos, e, f, inv)
    g.set_enable_self_collision(False)  # the poked centres have no edges: each is its own island
    g.step(1.0 / 60.0, 10, 4)
    g.reset()
    assert np.abs(_pos(g) - pos).max() <= 1e-6
    g.set_positions_and_velocities(pos + 0.01)
    assert np.abs(_pos(g) - (pos + 0.01)).max() <= 1e-6


def test_graph_replay_is_stable():
    """The frame graph replayed many times equals the oracle after each replay."""
    _say("graph replay")
    pos, e, f, inv = cases.config1()
    g = T.ClothSimulator(pos, e, f, inv)
    o = O.OracleSim(pos, e, f, inv)
    for s in (g, o):
        s.set_enable_self_collision(False)
    for k in range(5):
        g.step(1.0 / 60.0, 4, 3)
        o.step(1.0 / 60.0, 4, 3)
        assert np.abs(_pos(g) - o.get_positions().reshape(-1, 3)).max() < TOL


def test_empty_mesh_is_a_no_op():
    _say("empty mesh")
    g = T.ClothSimulator(np.zeros((0, 3), np.float32), np.zeros((0, 2), np.uint32))
    assert g.step(1.0 / 60.0) == 20
    g.get_positions(np.empty(0, np.float32))


@pytest.mark.parametrize("single_sided", [False, True])
def test_mesh_collider_matches_oracle(single_sided):
    """Triangle-mesh collider (Morton-ordered triangles, CCD + closest point + single-sided
    recovery) with friction 0.4: cloth dropped on a noisy icosphere (level 3)."""
    _say(f"mesh collider single_sided={single_sided}")
    sp, sf = cases.icosphere(3, radius=0.15, center=(0.0, 0.0, 0.0))
    tris = sp[sf.astype(np.int64)]
    pos, e, f, inv = cases.drape_grid(32, 0.01, 0.17, seed=9)
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.003)
    o = O.OracleSim(pos, e, f, inv, thickness=0.003)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_mesh_collider_triangles(tris, friction=0.4, thickness=0.002, restitution=0.0, single_sided=single_sided)
        s.set_solver_iterations(10)
    for fr in range(8):
        g.step(1.0 / 60.0, 10, 10)
        o.step(1.0 / 60.0, 10, 10)
        diff = float(np.abs(_pos(g) - o.get_positions().reshape(-1, 3)).max())
        assert diff < TOL, f"frame {fr}: {diff:.3e}"
    r = np.linalg.norm(_pos(g), axis=1)
    assert r.min() > 0.13  # resting on the surface, not through it


def test_mesh_collider_culling_large_mesh():
    """5,120 triangles (160 clusters in 5 super-clusters): the culled traversal against the
    oracle's scan of every triangle, bit for bit, with the cloth sliding over the mesh."""
    _say("mesh collider culling, 5120 triangles")
    sp, sf = cases.icosphere(4, radius=0.15, center=(0.0, 0.0, 0.0), seed=4)
    tris = sp[sf.astype(np.int64)]
    pos, e, f, inv = cases.drape_grid(40, 0.01, 0.17, seed=2)
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.003)
    o = O.OracleSim(pos, e, f, inv, thickness=0.003)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_enable_self_collision(False)
        s.set_mesh_collider_triangles(tris, friction=0.4, thickness=0.002, restitution=0.0)
        s.set_solver_iterations(10)
    for fr in range(10):
        g.step(1.0 / 60.0, 10, 10)
        o.step(1.0 / 60.0, 10, 10)
        diff = float(np.abs(_pos(g) - o.get_positions().reshape(-1, 3)).max())
        _say(f"  frame {fr}: {diff:.3e}")
        assert diff == 0.0, f"frame {fr}: {diff:.3e}"
    assert np.linalg.norm(_pos(g), axis=1).min() > 0.13


def test_mesh_collider_65k_vertex_cloth():
    """A 256 x 256 cloth (65,536 vertices) starting 1 mm above a 5,120-triangle sphere: the
    collider packs 8 vertices per wavefront at this size (4 on the smaller test meshes). Two
    substeps, bit for bit against the oracle's scan of every triangle."""
    _say("mesh collider, 65,536-vertex cloth")
    sp, sf = cases.icosphere(4, radius=0.3, center=(0.0, 0.0, 0.0), seed=4)
    tris = sp[sf.astype(np.int64)]
    pos, e, f, inv = cases.drape_grid(256, 0.004, 0.301, seed=6)
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.003)
    o = O.OracleSim(pos, e, f, inv, thickness=0.003)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_enable_self_collision(False)
        s.set_mesh_collider_triangles(tris, friction=0.4, thickness=0.002, restitution=0.0)
        s.set_solver_iterations(4)
    g.step(1.0 / 60.0, 2, 4)
    o.step(1.0 / 60.0, 2, 4)
    po = o.get_positions().reshape(-1, 3)
    assert float(np.abs(_pos(g) - po).max()) == 0.0
    assert int((po[:, 2] > pos[:, 2] + 1e-4).sum()) > 100  # the sphere pushed the centre up


def _run_path(path, pos, e, f, inv, frames, damp=False, iters=7, colliders=True):
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.004)
    g.set_compliances_raw(1e-8, 1e-6, 1e-5)
    g.set_enable_self_collision(False)
    if damp:
        g.set_damping_all(2.0, 1.0, 3.0, 4.0)
    if colliders:
        g.add_sphere_collider([0.0, 0.0, 0.0], 0.12, 0.3, 0.0)
    g.set_solver_path(path)
    for _ in range(frames):
        g.step(1.0 / 60.0, 5, iters)
    out = _pos(g)
    lam = g.lambdas()
    return g, out, lam


@pytest.mark.parametrize("damp", [False, True])
def test_per_iteration_path_bit_exact_to_per_colour_path(damp):
    """One launch per iteration over dependency-cone tiles gives bit-identical positions and
    lambdas to one launch per colour (odd iteration count exercises the buffer ping-pong)."""
    _say(f"fused vs per-colour damp={damp}")
    pos, e, f, inv = cases.drape_grid(96, 0.004, 0.14, seed=3)
    g1, p1, l1 = _run_path(1, pos, e, f, inv, 3, damp)
    g2, p2, l2 = _run_path(2, pos, e, f, inv, 3, damp)
    info = g2.solver_info()
    _say(f"plan: {info}")
    assert info["path"] == 2 and g1.solver_info()["path"] == 1
    assert np.array_equal(p1, p2)
    assert np.array_equal(l1[0], l2[0]) and np.array_equal(l1[1], l2[1])
    r1, r2 = g1.residual(1 / 300.0), g2.residual(1 / 300.0)
    assert abs(r1[0] - r2[0]) <= 1e-12 * r1[0] and r1[1] == r2[1]  # sums use double atomics


def test_per_iteration_path_matches_oracle_config1():
    _say("fused config1 vs oracle")
    g, o32, _ = _config1_pair()
    g.set_solver_path(2)
    for _ in range(10):
        g.step(1.0 / 60.0, 10, 10)
        o32.step(1.0 / 60.0, 10, 10)
    assert float(np.abs(_pos(g) - o32.get_positions().reshape(-1, 3)).max()) < TOL


def test_per_iteration_path_poked_and_pinned_bit_exact():
    _say("fused poked grid")
    pos, e, f, inv = cases.config1()
    _, p1, l1 = _run_path(1, pos, e, f, inv, 4, iters=5, colliders=False)
    _, p2, l2 = _run_path(2, pos, e, f, inv, 4, iters=5, colliders=False)
    assert np.array_equal(p1, p2) and np.array_equal(l1[0], l2[0])


@pytest.mark.parametrize("tile,block", [("256", "256"), ("2048", "1024")])
def test_per_iteration_path_tile_and_block_sizes(tile, block, monkeypatch):
    """Small tiles with small blocks (several items per thread and phase) and large tiles."""
    _say(f"fused tile={tile} block={block}")
    monkeypatch.setenv("TCLOTH_TILE_VERTICES", tile)
    monkeypatch.setenv("TCLOTH_TILE_BLOCK", block)
    pos, e, f, inv = cases.drape_grid(80, 0.004, 0.14, seed=4)
    _, p1, l1 = _run_path(1, pos, e, f, inv, 2, iters=6)
    g2, p2, l2 = _run_path(2, pos, e, f, inv, 2, iters=6)
    assert g2.solver_info()["path"] == 2
    assert np.array_equal(p1, p2) and np.array_equal(l1[0], l2[0]) and np.array_equal(l1[1], l2[1])


def test_autotune_keeps_results_identical():
    _say("autotune")
    pos, e, f, inv = cases.drape_grid(96, 0.004, 0.14, seed=6)
    ref = T.ClothSimulator(pos, e, f, inv, thickness=0.004)
    tun = T.ClothSimulator(pos, e, f, inv, thickness=0.004)
    for g in (ref, tun):
        g.set_compliances_raw(1e-8, 1e-6, 1e-5)
        g.set_enable_self_collision(False)
    ref.set_solver_path(1)
    tune = tun.autotune(1.0 / 300.0)
    _say(f"autotune: {tune}")
    for g in (ref, tun):
        for _ in range(3):
            g.step(1.0 / 60.0, 5, 9)
    assert np.array_equal(_pos(ref), _pos(tun))
