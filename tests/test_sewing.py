"""Sewing springs: mesh topology (mesh.rs:407-1377 with sewing: the sewing constraints and their
colours mesh.rs:911-941, seam-gated bending across sewn panels mesh.rs:~700-800, the contracted
self-collision adjacency and two-hop lists mesh.rs:991-1191 and 1382-1490, boundary bits
mesh.rs:1291-1344) and the solver (sewing_shrink.wgsl once per substep before predict,
sewing.wgsl after distance in every iteration and once more before the velocity update,
dispatch.rs).

CPU tests pin the oracle's restatement, and the product's host mesh build against it:
  * the reference's own mesh tests (mesh.rs:1671, 1859, 1995, 2052, 2141): zero-hop sharing of
    neighbours, two-hop exclusions and unfolded seam rest lengths (2.0 for panels 10 m apart),
    seam-gated bending rest 2.0 with gate 1.0, open boundary bits, point-seam slit rest 2.0;
  * invariants on random sewn meshes: symmetric adjacency holding every sewn pair and every
    original edge, symmetric two-hop rest lengths, the product's arrays equal to the oracle's;
  * closed forms of the solver on two free particles joined by one sewing spring: at zero
    compliance the distance ends each substep on the shrinking rest length (1 - speed * t) with the
    midpoint fixed; at compliance alpha one substep from rest leaves c * alpha~ / (2 + alpha~) of
    the violation whatever the iteration count (the XPBD fixed point, which pins the accumulation
    of lambda inside a substep), and several substeps follow the one-dimensional XPBD recursion with
    lambda cleared per substep; the lock closes a pair within max(thickness sum, lock distance) to
    zero length and latches rest length 0, and not outside it.
GPU tests: the product equals the FP32 oracle bit for bit on those cases and on two sewn panels
draped with self-collision, and reproduces the reference's simulator tests (tests.rs:1377, 1453,
2998): flap closure without crossing, a flat seam closing below 1 mm, a shrink speed changed
mid-run keeping the rest length already reached.
"""
import json
import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O
from taremincloth_amd import taremin_cloth_core as T

REF_KW = dict(thickness=0.005, stiffness=1000.0, compression_stiffness=1000.0, shear_stiffness=1000.0,
              bending_stiffness=10.0)


def _say(msg):
    print(f"[sewing] {msg}", flush=True)
    sys.stdout.flush()


def _two_hop(m, a, b):
    lo, hi = int(m["th_off"][a]), int(m["th_off"][a + 1])
    for k in range(lo, hi):
        if int(m["th_idx"][k]) == b:
            return float(m["th_rest"][k])
    return None


def _adj(m, v):
    return set(int(x) for x in m["adj_idx"][m["adj_off"][v]:m["adj_off"][v + 1]])


def _builders():
    return [("oracle", lambda *a, **k: O.build_mesh(*a, **k)), ("product", lambda *a, **k: T.build_mesh_arrays(*a, **k))]


def _edges_of(m, name):
    if name == "oracle":
        return np.stack([m["edge_v0"], m["edge_v1"]], 1)
    return m["edges"].reshape(-1, 2)


# ------------------------------------------------------------------ CPU: reference mesh tests
@pytest.mark.parametrize("name,build", _builders())
def test_sewing_topology_zero_hop(name, build):
    """mesh.rs:1671-1705: sewing (1, 2) makes 1 and 2 one node; each shares the other's neighbours."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0]], np.float32)
    m = build(pos, np.array([[0, 1], [2, 3]], np.uint32), None, sewing_springs=[[1, 2]], **REF_KW)
    assert {0, 2, 3} <= _adj(m, 1)
    assert {0, 1, 3} <= _adj(m, 2)


@pytest.mark.parametrize("name,build", _builders())
def test_two_hop_rest_lengths_sewing(name, build):
    """mesh.rs:1859-1905: 0-1 is an edge (rest 0), 2 is a contracted one-hop neighbour of 0 (rest 0)."""
    pos = np.array([[0, 0, 0], [0.1, 0, 0], [1.0, 0, 0], [1.1, 0, 0]], np.float32)
    m = build(pos, np.array([[0, 1], [2, 3]], np.uint32), None, sewing_springs=[[1, 2]], **REF_KW)
    assert _two_hop(m, 0, 1) == 0.0
    r = _two_hop(m, 0, 2)
    assert r is None or r == 0.0
    r = _two_hop(m, 0, 3)
    assert r is None or r == 0.0   # vertices that are not face vertices carry no rest length


@pytest.mark.parametrize("name,build", _builders())
def test_seam_gated_bending_across_sewn_panels(name, build):
    """mesh.rs:1995-2049: one bending constraint across the seam, rest h_a + h_b = 2.0, gate 1.0."""
    pos = np.array([[0, 0, 0], [0, 1, 0], [-1, 0.5, 0], [5, 0, 0], [5, 1, 0], [6, 0.5, 0]], np.float32)
    e = np.array([[0, 1], [1, 2], [2, 0], [3, 4], [4, 5], [5, 3]], np.uint32)
    f = np.array([[0, 1, 2], [4, 3, 5]], np.uint32)
    m = build(pos, e, f, sewing_springs=[[0, 3], [1, 4]], **REF_KW)
    assert len(m["bend_v2"]) == 1
    v2, v3 = int(m["bend_v2"][0]), int(m["bend_v3"][0])
    assert (min(v2, v3), max(v2, v3)) == (2, 5)
    assert abs(float(m["bend_rest"][0]) - 2.0) < 1e-5
    assert float(m["bend_pad0"][0]) > 0.5
    assert abs(float(m["bend_pad1"][0]) - 1.0) < 1e-5
    # the two sewing constraints, 5 m apart
    assert sorted(zip(m["sew_v0"].tolist(), m["sew_v1"].tolist())) == [(0, 3), (1, 4)]
    assert np.allclose(m["sew_rest"], 5.0)


def _quads_10m():
    pos = np.array([[0, 1, 0], [0, 0, 0], [-1, 1, 0], [-1, 0, 0], [10, 1, 0], [10, 0, 0], [11, 1, 0], [11, 0, 0]],
                   np.float32)
    e = np.array([[0, 1], [0, 2], [1, 2], [1, 3], [2, 3], [4, 5], [4, 6], [5, 6], [5, 7], [6, 7]], np.uint32)
    f = np.array([[0, 2, 1], [1, 2, 3], [4, 5, 6], [5, 7, 6]], np.uint32)
    return pos, e, f


@pytest.mark.parametrize("name,build", _builders())
def test_seam_two_hop_unfolded_rest_and_boundary_bits(name, build):
    """mesh.rs:2052-2138."""
    pos, e, f = _quads_10m()
    m = build(pos, e, f, sewing_springs=[[0, 4], [1, 5]], **REF_KW)
    for a, b in ((0, 4), (4, 0), (1, 5), (5, 1), (0, 5), (4, 1)):
        assert _two_hop(m, a, b) == 0.0, (a, b)
    d26, d62 = _two_hop(m, 2, 6), _two_hop(m, 6, 2)
    assert d26 is not None and d62 is not None
    assert abs(d26 - 2.0) < 1e-4 and abs(d62 - 2.0) < 1e-4
    assert d26 == d62
    ed = _edges_of(m, name)
    bits = m["edge_bbits"]
    for k, (a, b) in enumerate(ed.tolist()):
        pair = (min(a, b), max(a, b))
        is_b = (int(bits[k >> 5]) >> (k & 31)) & 1
        assert is_b == (0 if pair in ((0, 1), (4, 5), (1, 2), (5, 6)) else 1), pair


@pytest.mark.parametrize("name,build", _builders())
def test_point_seam_slit_two_hop_rest(name, build):
    """mesh.rs:2141-2200: one point seam (0, 3); 1 and 4 are a true two-hop slit of rest 1 + 1."""
    pos = np.array([[0, 0, 0], [0, 1, 0], [-1, 0, 0], [5, 0, 0], [5, 1, 0], [6, 0, 0]], np.float32)
    e = np.array([[0, 1], [1, 2], [2, 0], [3, 4], [4, 5], [5, 3]], np.uint32)
    f = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    m = build(pos, e, f, sewing_springs=[[0, 3]], **REF_KW)
    d14, d41 = _two_hop(m, 1, 4), _two_hop(m, 4, 1)
    assert d14 is not None and abs(d14 - 2.0) < 1e-4
    assert d14 == d41


# ------------------------------------------------------------------ CPU: invariants, product = oracle
def sewn_panels(n=8, spacing=0.02, gap=0.05, stride=1, seed=0):
    """Two n x n panels in the xy plane, gap apart along x; the facing columns sewn every stride rows."""
    pa, ea, fa = cases.grid(n, n, spacing)
    pb, eb, fb = cases.grid(n, n, spacing, origin=((n - 1) * spacing + gap, 0.0, 0.0))
    nv = len(pa)
    pos = np.concatenate([pa, pb]).astype(np.float32)
    pos += cases.uniform_jitter(len(pos), 1e-4, seed).astype(np.float32)
    e = np.concatenate([ea, eb + nv]).astype(np.uint32)
    f = np.concatenate([fa, fb + nv]).astype(np.uint32)
    right = [j * n + (n - 1) for j in range(0, n, stride)]
    left = [nv + j * n for j in range(0, n, stride)]
    sew = np.array(list(zip(right, left)), np.uint32)
    return pos, e, f, sew


def _grid_order_ok():
    p, _, _ = cases.grid(3, 3, 1.0)
    return np.allclose(p[1], [1.0, 0.0, 0.0]) and np.allclose(p[3], [0.0, 1.0, 0.0])


@pytest.mark.parametrize("n,stride,extra", [(6, 1, False), (9, 2, False), (7, 1, True)])
def test_sewn_mesh_invariants_and_product_equals_oracle(n, stride, extra):
    assert _grid_order_ok()
    pos, e, f, sew = sewn_panels(n, stride=stride, seed=n)
    if extra:   # a three-vertex group (a chain of two sewing pairs) and a repeated pair
        sew = np.concatenate([sew, [[sew[0, 1], 2 * n * n - 1]], sew[:1]]).astype(np.uint32)
    kw = dict(REF_KW, sewing_springs=sew, sewing_shrink_speed=0.7)
    a = T.build_mesh_arrays(pos, e, f, **kw)
    b = O.build_mesh(pos, e, f, **kw)
    for k in ("adj_off", "adj_idx", "th_off", "th_idx", "th_rest", "island", "edge_bbits", "bend_v0", "bend_v1",
              "bend_v2", "bend_v3", "bend_rest", "bend_pad0", "bend_pad1", "bend_col_off", "bend_col_cnt",
              "sew_v0", "sew_v1", "sew_rest", "sew_col_off", "sew_col_cnt", "dist_v0", "dist_v1"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    nv = len(pos)
    adj = [_adj(b, v) for v in range(nv)]
    for v in range(nv):
        for u in adj[v]:
            assert v in adj[u]
    for x, y in sew.tolist():
        if x != y:
            assert y in adj[x]
    for x, y in e.tolist():
        assert y in adj[x]
    th = {}
    for v in range(nv):
        for k in range(int(b["th_off"][v]), int(b["th_off"][v + 1])):
            th[(v, int(b["th_idx"][k]))] = float(b["th_rest"][k])
    for (v, u), r in th.items():
        if (u, v) in th:
            assert th[(u, v)] == r
    # the sewing constraints: one per valid input pair, repeats kept (mesh.rs:921-940), conflict-free
    # colours, rest = initial distance
    s0, s1 = b["sew_v0"], b["sew_v1"]
    assert sorted(zip(s0.tolist(), s1.tolist())) == sorted(tuple(x) for x in sew.tolist() if x[0] != x[1])
    assert np.allclose(b["sew_rest"], np.linalg.norm(pos[s0] - pos[s1], axis=1), atol=1e-6)
    off, cnt = b["sew_col_off"], b["sew_col_cnt"]
    for c in range(len(off)):
        used = set()
        for k in range(int(off[c]), int(off[c] + cnt[c])):
            assert int(s0[k]) not in used and int(s1[k]) not in used
            used |= {int(s0[k]), int(s1[k])}
    # a single island: the seam joins both panels
    assert len(set(b["island"].tolist())) == 1


def test_sewing_validation():
    pos, e, f, sew = sewn_panels(4)
    with pytest.raises(ValueError):
        T.ClothSimulator(pos, e, f, sewing_springs=np.array([[0, len(pos)]], np.uint32))
    with pytest.raises(ValueError):
        T.ClothSimulator(pos, e, f, sewing_springs=np.array([0, 1, 2], np.uint32))


# ------------------------------------------------------------------ CPU: solver closed forms
def _pair_sim(d0=1.0, precision="f64", thick=None, **kw):
    pos = np.array([[0.0, 0.0, 0.0], [d0, 0.0, 0.0]], np.float32)
    s = O.OracleSim(pos, np.zeros((0, 2), np.uint32), None, np.ones(2, np.float32), thicknesses=thick,
                    sewing_springs=[[0, 1]], precision=precision, **kw)
    s.set_gravity(0.0, 0.0, 0.0)
    s.set_damping(0.0)
    s.set_enable_self_collision(False)
    return s


def _pxy(s):
    return s.state(0)


def test_oracle_hard_sewing_tracks_shrinking_rest():
    """Zero compliance (stiffness >= 5000): distance = 1 - speed * t after every frame, midpoint fixed."""
    s = _pair_sim(sewing_shrink_speed=0.5)
    assert s.sewing_rest_lengths()[0] == pytest.approx(1.0, abs=1e-7)
    for fr in range(1, 5):
        s.step(0.1, 4, 3)
        r = s.sewing_rest_lengths()[0]
        assert r == pytest.approx(1.0 - 0.05 * fr, abs=1e-6)
        p = _pxy(s)
        assert np.linalg.norm(p[1] - p[0]) == pytest.approx(r, abs=1e-6)
        assert 0.5 * (p[0, 0] + p[1, 0]) == pytest.approx(0.5, abs=1e-6)
        assert np.abs(p[:, 1:]).max() < 1e-9
    # the shrink stops at the target length 0 ... unless the lock closes the pair first
    s.set_sewing_shrink_speed(2.0)
    s.step(0.1, 4, 3)
    assert s.sewing_rest_lengths()[0] == pytest.approx(0.6, abs=1e-6)


@pytest.mark.parametrize("iters", [0, 1, 5])
def test_oracle_compliant_sewing_fixed_point(iters):
    """One substep from rest: the projection leaves c * a / (2 + a), a = compliance / dt^2, for any count
    of passes (lambda accumulates inside the substep)."""
    s = _pair_sim(sewing_shrink_speed=1.0)
    s.set_sewing_stiffness(200.0)
    dt = float(np.float32(0.01))      # the time step and the compliance enter as fp32 values
    s.step(dt, 1, iters)
    c0 = 1.0 * dt                     # the rest length shrank by speed * dt; the particles had not moved
    a = float(np.float32(1.0 / 200.0)) / (dt * dt)
    p = _pxy(s)
    d = float(np.linalg.norm(p[1] - p[0]))
    assert d - (1.0 - dt) == pytest.approx(c0 * a / (2.0 + a), rel=1e-9, abs=1e-12)


def _xpbd_1d(d0, speed, stiffness, dt, substeps, frames):
    """One sewing spring between two unit masses along a line: shrink, predict, one projection."""
    d, v, r = d0, 0.0, d0
    comp = 0.0 if stiffness >= 5000.0 else float(np.float32(1.0 / max(stiffness, 1.0)))
    h = float(np.float32(dt)) / substeps       # fp32 frame step, substep and speed as the simulator takes them
    speed = float(np.float32(speed))
    a = comp / (h * h)
    for _ in range(frames * substeps):
        r = max(0.0, r - speed * h)
        dp = d + v * h
        dl = -(dp - r) / (2.0 + a)
        dn = dp + 2.0 * dl
        v = (dn - d) / h
        d = dn
    return d, r


@pytest.mark.parametrize("stiffness", [300.0, 2000.0, 10000.0])
def test_oracle_sewing_follows_the_1d_recursion(stiffness):
    """Several substeps and frames: the distance follows the one-dimensional XPBD recursion with lambda
    cleared each substep (a lambda kept across substeps would lag behind it)."""
    s = _pair_sim(sewing_shrink_speed=0.8)
    s.set_sewing_stiffness(stiffness)
    s.step(1.0 / 30.0, 6, 4)
    s.step(1.0 / 30.0, 6, 4)
    d_ref, r_ref = _xpbd_1d(1.0, 0.8, stiffness, 1.0 / 30.0, 6, 2)
    p = _pxy(s)
    assert float(np.linalg.norm(p[1] - p[0])) == pytest.approx(d_ref, rel=1e-9)
    assert s.sewing_rest_lengths()[0] == pytest.approx(r_ref, rel=1e-12)


@pytest.mark.parametrize("d0,thick,lock_dist,lock,closes", [
    (0.015, None, None, True, True),      # within the default 0.02 m lock distance
    (0.015, None, 0.01, True, False),     # the lock distance lowered below the gap
    (0.015, None, None, False, False),    # lock off
    (0.08, 0.05, None, True, True),       # within the thickness sum 0.1 m
    (0.12, 0.05, None, True, False),      # outside both
])
def test_oracle_sewing_lock(d0, thick, lock_dist, lock, closes):
    th = None if thick is None else np.full(2, thick, np.float32)
    s = _pair_sim(d0, thick=th, sewing_shrink_speed=0.1, sewing_lock_distance=lock_dist)
    s.set_enable_sewing_lock(lock)
    s.step(0.01, 2, 2)
    p = _pxy(s)
    d = float(np.linalg.norm(p[1] - p[0]))
    r = s.sewing_rest_lengths()[0]
    if closes:
        assert d < 1e-7 and r == 0.0
    else:
        assert r == pytest.approx(d0 - 0.001, abs=1e-7)   # fp32 inputs
        assert d == pytest.approx(r, abs=1e-7)


def test_oracle_rest_length_setter_round_trip():
    s = _pair_sim(sewing_shrink_speed=0.5)
    s.step(0.1, 4, 2)
    s.sewing_rest_lengths(np.array([0.3]))
    assert s.sewing_rest_lengths()[0] == pytest.approx(0.3)
    s.step(0.1, 4, 2)
    assert s.sewing_rest_lengths()[0] == pytest.approx(0.25, abs=1e-6)


def test_closure_ratio_pure():
    """tests.rs:637-654 through the C ABI (host only)."""
    pos = np.array([[0, 0, 0], [0.001, 0, 0], [0, 0, 0], [1.0, 0, 0]], np.float32)
    assert T.sewing_closure_ratio(pos, [[0, 1], [2, 3]], 0.005) == pytest.approx(0.5, abs=1e-6)
    assert T.sewing_closure_ratio(pos, np.zeros((0, 2)), 0.005) == 1.0
    assert T.sewing_closure_ratio(pos, [[0, 99]], 0.005) == 1.0
    # brute force on random pairs, including out-of-range ones
    rng = np.random.default_rng(5)
    p = rng.uniform(-0.01, 0.01, (50, 3)).astype(np.float32)
    q = rng.integers(0, 60, (200, 2))
    ok = (q < 50).all(axis=1)
    d2 = ((p[q[ok, 0]] - p[q[ok, 1]]) ** 2).sum(axis=1)
    want = float(np.float32((d2 <= np.float32(0.01) ** 2).sum()) / np.float32(ok.sum()))
    assert T.sewing_closure_ratio(p, q, 0.01) == pytest.approx(want, abs=1e-6)


# ------------------------------------------------------------------ GPU: product against the oracle
def _gpos(sim):
    out = np.empty(3 * sim.num_vertices, np.float32)
    sim.get_positions(out)
    return out.reshape(-1, 3)


def _gpair(d0=1.0, thick=None, **kw):
    pos = np.array([[0.0, 0.0, 0.0], [d0, 0.0, 0.0]], np.float32)
    g = T.ClothSimulator(pos, np.zeros((0, 2), np.uint32), None, np.ones(2, np.float32), thicknesses=thick,
                         sewing_springs=np.array([[0, 1]], np.uint32), **kw)
    g.set_gravity(0.0, 0.0, 0.0)
    g.set_damping(0.0)
    g.set_enable_self_collision(False)
    return g


@pytest.mark.gpu
def test_gpu_dynamic_shrink_speed_preserves_rest_length(require_gpu):
    """tests.rs:2998-3026 with the product: 1.0 -> 0.95 at 0.5 m/s, then 0.95 -> 0.75 at 2.0 m/s."""
    g = _gpair(sewing_shrink_speed=0.5, sewing_stiffness=1000.0, enable_sewing_lock=False)
    g.step(0.1, 4)
    r1 = g.get_sewing_current_rest_lengths()[0]
    _say(f"shrink: r1 = {r1:.6f}")
    assert abs(r1 - 0.95) < 1e-3
    g.set_sewing_shrink_speed(2.0)
    assert abs(g.get_sewing_shrink_speed() - 2.0) < 1e-6
    assert abs(g.get_sewing_current_rest_lengths()[0] - 0.95) < 1e-3
    g.step(0.1, 4)
    r2 = g.get_sewing_current_rest_lengths()[0]
    _say(f"shrink: r2 = {r2:.6f}")
    assert abs(r2 - 0.75) < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("path", [1, 2])   # one launch per colour; one launch per iteration
@pytest.mark.parametrize("stiffness,lock,d0,thick", [(10000.0, True, 1.0, None), (300.0, True, 1.0, None),
                                                     (2000.0, False, 0.015, None), (10000.0, True, 0.015, None),
                                                     (10000.0, True, 0.08, 0.05)])
def test_gpu_pair_bit_exact(require_gpu, stiffness, lock, d0, thick, path):
    th = None if thick is None else np.full(2, thick, np.float32)
    g = _gpair(d0, thick=th, sewing_shrink_speed=0.3)
    g.set_solver_path(path)
    assert g.solver_info()["path"] == path
    o = _pair_sim(d0, precision="f32", thick=th, sewing_shrink_speed=0.3)
    for s in (g, o):
        s.set_sewing_stiffness(stiffness)
        s.set_enable_sewing_lock(lock)
    for fr in range(4):
        g.step(1.0 / 60.0, 5, 3)
        o.step(1.0 / 60.0, 5, 3)
        assert np.array_equal(_gpos(g), o.get_positions().reshape(-1, 3)), fr
        rg = np.asarray(g.get_sewing_current_rest_lengths(), np.float32)
        ro = o.sewing_rest_lengths().astype(np.float32)
        assert np.array_equal(rg, ro), fr


def _setup_drape(s):
    s.set_gravity(0.0, 0.0, -9.81)
    s.set_compliances_raw(1e-8, 1e-6, 1e-5)
    s.set_enable_self_collision(True)


@pytest.mark.gpu
@pytest.mark.parametrize("path", [1, 2])   # one launch per colour; one launch per iteration
@pytest.mark.parametrize("n,stride,damp", [(10, 1, False), (24, 3, False), (12, 1, True)])
def test_gpu_sewn_panels_drape_bit_exact(require_gpu, n, stride, damp, path):
    """stride 1 sews every boundary vertex, so the seam edges carry seam-gated bending (mesh.rs:706-831);
    both solver paths give the oracle's bits, with and without constraint damping."""
    pos, e, f, sew = sewn_panels(n, spacing=0.015, gap=0.04, stride=stride, seed=n)
    inv = np.ones(len(pos), np.float32)
    inv[[0, n * (n - 1), n * n + n - 1, 2 * n * n - 1]] = 0.0   # the outer corners (away from the seam) held
    kw = dict(thickness=0.003, sewing_shrink_speed=0.4)
    g = T.ClothSimulator(pos, e, f, inv, sewing_springs=sew, **kw)
    o = O.OracleSim(pos, e, f, inv, sewing_springs=sew, **kw)
    for s in (g, o):
        _setup_drape(s)
        if damp:
            s.set_damping_all(0.3, 0.2, 0.25, 0.4)
    g.set_solver_path(path)
    info = g.solver_info()
    assert info["path"] == path
    if path == 2:
        assert info["plan_available"] == 1.0, info["plan_note"]
    if stride == 1:
        assert int(np.count_nonzero((o.mesh()["bend_pad0"] > 0.5) & (o.mesh()["bend_pad0"] < 1.5))) > 0
    _say(f"sewn panels {n}x{n} x2, {len(sew)} sewing pairs, self-collision on, path {path}, damping {damp}")
    for fr in range(6):
        g.step(1.0 / 60.0, 5, 6)
        o.step(1.0 / 60.0, 5, 6)
        diff = float(np.abs(_gpos(g) - o.get_positions().reshape(-1, 3)).max())
        _say(f"  frame {fr}: max |hip - oracle| = {diff:.3e}")
        assert diff == 0.0
    rg = np.asarray(g.get_sewing_current_rest_lengths(), np.float32)
    assert np.array_equal(rg, o.sewing_rest_lengths().astype(np.float32))
    p = _gpos(g)
    gap = np.linalg.norm(p[sew[:, 0]] - p[sew[:, 1]], axis=1)
    _say(f"  seam gap after 6 frames: max {gap.max() * 1e3:.3f} mm, median {np.median(gap) * 1e3:.3f} mm "
         f"(from 40 mm)")
    assert gap.max() < 0.02


@pytest.mark.gpu
def test_gpu_flap_closure_no_inversion(require_gpu):
    """tests.rs:1377-1450: two facing flaps, anchors fixed, top corners sewn; rims do not cross."""
    pos = np.array([[-0.05, 0.0, 0.0], [-0.05, 0.1, 0.0], [-0.05, 0.2, 0.0],
                    [0.05, 0.0, 0.0], [0.05, 0.1, 0.0], [0.05, 0.2, 0.0]], np.float32)
    e = np.array([[0, 1], [1, 2], [2, 0], [3, 4], [4, 5], [5, 3]], np.uint32)
    f = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    inv = np.array([0, 1, 1, 0, 1, 1], np.float32)
    kw = dict(thickness=0.005, stiffness=10000.0, compression_stiffness=10000.0, shear_stiffness=5000.0,
              bending_stiffness=0.0)
    g = T.ClothSimulator(pos, e, f, inv, sewing_springs=np.array([[2, 5]], np.uint32), **kw)
    o = O.OracleSim(pos, e, f, inv, sewing_springs=np.array([[2, 5]], np.uint32), **kw)
    for s in (g, o):
        s.set_enable_self_collision(True)
        s.set_gravity(0.0, 0.0, 0.0)
    it = 10
    for _ in range(30):
        g.step(1.0 / 60.0, 10, it)
        o.step(1.0 / 60.0, 10, it)
    p = _gpos(g)
    assert np.array_equal(p, o.get_positions().reshape(-1, 3))
    _say(f"flap closure: L = {p[1, 0]:.4f}, R = {p[4, 0]:.4f}")
    assert p[1, 0] <= p[4, 0] + 1e-4


@pytest.mark.gpu
def test_gpu_flat_seam_closure_no_stagnation(require_gpu):
    """tests.rs:1453-1532: two flat quads 5 cm apart, seam edges sewn; both pairs close below 1 mm."""
    pos = np.array([[-0.05, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.1, 0.0], [-0.05, 0.1, 0.0],
                    [0.05, 0.0, 0.0], [0.05, 0.1, 0.0], [0.10, 0.1, 0.0], [0.10, 0.0, 0.0]], np.float32)
    e = np.array([[0, 1], [1, 2], [2, 3], [3, 0], [0, 2], [4, 5], [5, 6], [6, 7], [7, 4], [4, 6]], np.uint32)
    f = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]], np.uint32)
    sew = np.array([[1, 4], [2, 5]], np.uint32)
    kw = dict(thickness=0.005, stiffness=10000.0, compression_stiffness=10000.0, shear_stiffness=5000.0,
              bending_stiffness=0.0)
    g = T.ClothSimulator(pos, e, f, None, sewing_springs=sew, **kw)
    o = O.OracleSim(pos, e, f, None, sewing_springs=sew, **kw)
    for s in (g, o):
        s.set_enable_self_collision(True)
        s.set_gravity(0.0, 0.0, 0.0)
    for _ in range(30):
        g.step(1.0 / 60.0, 15, 10)
        o.step(1.0 / 60.0, 15, 10)
    p = _gpos(g)
    assert np.array_equal(p, o.get_positions().reshape(-1, 3))
    d14, d25 = float(np.linalg.norm(p[1] - p[4])), float(np.linalg.norm(p[2] - p[5]))
    _say(f"flat seam closure: d(1,4) = {d14:.5f}, d(2,5) = {d25:.5f}")
    assert d14 < 0.001 and d25 < 0.001


@pytest.mark.gpu
def test_gpu_sewing_debug_record_replay(require_gpu, tmp_path):
    """A sewn simulator's recording holds its sewing pairs and rest lengths and replays bit for bit."""
    from taremincloth_amd import debug_recorder as DR
    pos, e, f, sew = sewn_panels(6, seed=3)
    g = T.ClothSimulator(pos, e, f, None, sewing_springs=sew, sewing_shrink_speed=0.6)
    g.set_gravity(0.0, 0.0, -9.81)
    g.set_solver_iterations(5)
    g.start_debug_recording("sewn", max_frames=10)
    for k in range(4):
        if k == 2:
            g.set_sewing_shrink_speed(1.5)
        g.step(1.0 / 60.0, 4)
    g.stop_debug_recording()
    path = g.save_debug_recording(str(tmp_path / "sewn.jsonl.gz"))
    meta, frames = DR.load(path)
    assert meta["sewing_springs"] == sew.tolist()
    assert np.allclose(frames[-1]["sewing_rest_lengths"], g.get_sewing_current_rest_lengths())
    n = 0
    for fr, p in DR.replay(path):
        assert np.array_equal(p, np.asarray(fr["positions"], np.float32)), fr["frame_index"]
        n += 1
    assert n == 4


# ------------------------------------------------------------------ GPU: sewing priority (mod.rs:1262-1420)
def _prio_pair():
    """tests.rs:657-665: two vertices, one sewing pair, 1 m apart."""
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]], np.float32)
    return pos, T.ClothSimulator(pos, np.zeros((0, 2), np.uint32), None, None,
                                 sewing_springs=np.array([[0, 1]], np.uint32), stiffness=10000.0,
                                 compression_stiffness=10.0, shear_stiffness=5000.0, bending_stiffness=0.0)


@pytest.mark.gpu
def test_gpu_sewing_priority_latch_and_ramp(require_gpu):
    pos, g = _prio_pair()
    assert g.num_sewing_constraints == 1
    g.set_sewing_priority_options(True, 0.9, 0.005, 2, 0)
    r0, s0, lat0 = g.update_sewing_priority(pos)
    assert (r0, s0, lat0) == (0.0, 0.0, False)
    assert g.is_sewing_priority_active()
    closed = np.array([[0.0, 0.0, 0.0], [0.001, 0.0, 0.0]], np.float32)
    r1, s1, lat1 = g.update_sewing_priority(closed)
    assert r1 == 1.0 and lat1 and abs(s1 - 0.5) < 1e-6      # smoothstep(0.5)
    _, s2, _ = g.update_sewing_priority(closed)
    assert abs(s2 - 1.0) < 1e-6
    _, s3, lat3 = g.update_sewing_priority(pos)              # one-way latch
    assert lat3 and s3 == 1.0
    assert g.get_sewing_priority_state() == (True, 4, 1.0, 0.0)
    # an unchanged setting keeps the latch, a changed one resets it
    g.set_sewing_priority_options(True, 0.9, 0.005, 2, 0)
    assert g.get_sewing_priority_state()[0]
    g.set_sewing_priority_options(True, 0.8, 0.005, 2, 0)
    assert g.get_sewing_priority_state() == (False, 0, 0.0, 0.0)
    # reset clears it too
    g.update_sewing_priority(closed)
    g.reset()
    assert g.get_sewing_priority_state() == (False, 0, 0.0, 0.0)


@pytest.mark.gpu
def test_gpu_sewing_priority_max_frames_and_disabled(require_gpu):
    pos, g = _prio_pair()
    g.set_sewing_priority_options(True, 1.0, 0.005, 0, 2)       # tests.rs:698-718
    _, s0, lat0 = g.update_sewing_priority(pos)
    assert not lat0 and s0 == 0.0
    _, s1, lat1 = g.update_sewing_priority(pos)
    assert lat1 and s1 == 1.0
    pos, g = _prio_pair()                                       # tests.rs:721-738
    assert g.get_sewing_priority_scale() == 1.0 and not g.is_sewing_priority_active()
    assert g.update_sewing_priority(pos) == (1.0, 1.0, True)
    assert g.get_sewing_closure_ratio() == 1.0
    # clamps (mod.rs:1270-1273) and the configuration record
    g.set_sewing_priority_options(True, 1.5, 1.0, 500, 10 ** 6)
    c = json.loads(g.get_config_json())
    assert (c["sewing_priority_threshold"], c["sewing_priority_ramp_frames"], c["sewing_priority_max_frames"]) == \
        (1.0, 120, 100000)
    assert abs(c["sewing_priority_merge_dist"] - 0.1) < 1e-7
    with pytest.raises(OverflowError):
        g.set_sewing_priority_options(True, 0.9, 0.005, -1, 0)


@pytest.mark.gpu
def test_gpu_sewing_priority_scales_gravity_bit_exact(require_gpu):
    """Gravity times the latch's scale, frame by frame: the product with the priority on equals the FP32
    oracle stepped with that scaled gravity; while the latch holds, the panels do not fall."""
    n = 10
    pos, e, f, sew = sewn_panels(n, spacing=0.015, gap=0.03, seed=7)
    kw = dict(thickness=0.003, sewing_shrink_speed=0.6)
    g = T.ClothSimulator(pos, e, f, None, sewing_springs=sew, **kw)
    o = O.OracleSim(pos, e, f, None, sewing_springs=sew, **kw)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_enable_self_collision(True)
    g.set_sewing_priority_options(True, 0.9, 0.004, 3, 0)
    prev = pos.copy()
    scales = []
    for fr in range(10):
        _, sc, _ = g.update_sewing_priority(prev)
        scales.append(sc)
        gz = np.float32(-9.81) * np.float32(sc)
        o.set_gravity(0.0 * np.float32(sc), 0.0 * np.float32(sc), float(gz))
        g.step(1.0 / 60.0, 5, 6)
        o.step(1.0 / 60.0, 5, 6)
        prev = _gpos(g)
        assert np.array_equal(prev, o.get_positions().reshape(-1, 3)), fr
        if sc == 0.0:
            assert float(np.abs(prev[:, 2] - pos[:, 2]).max()) < 1e-3, fr
    _say(f"priority scales per frame: {scales}")
    assert scales[0] == 0.0 and scales[-1] == 1.0 and any(0.0 < x < 1.0 for x in scales)


@pytest.mark.gpu
def test_gpu_sewing_priority_suppresses_wrinkle_field(require_gpu):
    """tests.rs:741-776 made exact: with the wrinkle field on, the unlatched frames skip the field and the
    ramping frames scale its blend by the latch's scale; the product equals the FP32 oracle stepped
    without the field, then with the blend scaled, frame by frame."""
    pos = np.array([[0.07, 0.0, 0.0], [-0.07, 0.0, 0.0]], np.float32)   # both inside the field's range
    sew = np.array([[0, 1]], np.uint32)
    g = T.ClothSimulator(pos, np.zeros((0, 2), np.uint32), None, None, sewing_springs=sew)
    o = O.OracleSim(pos, np.zeros((0, 2), np.uint32), None, None, sewing_springs=sew)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, 0.0)
        s.set_enable_self_collision(False)
    tw, th = 16, 8
    tex = np.random.default_rng(3).integers(0, 256, size=(th, tw, 4), dtype=np.uint8)
    g.set_enable_wrinkle_field(True)
    g.set_wrinkle_field_params([0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                               influence_radius=0.05, bone_radius=0.05, stiffness=3.0, blend_weight=0.9)
    g.set_wrinkle_field_texture_2d(tw, th, tex.tobytes(), -0.2, 0.2, 0.02, 0.16)
    g.set_sewing_priority_options(True, 0.9, 0.005, 2, 0)
    wp = g._wrinkle_params()
    closed = np.array([[0.0, 0.0, 0.0], [0.001, 0.0, 0.0]], np.float32)
    seen = []
    for fr, upd in enumerate([pos, pos, closed, closed, closed]):
        _, sc, _ = g.update_sewing_priority(upd)
        seen.append(sc)
        w = wp.copy()
        w[15] = np.float32(w[15]) * np.float32(sc)
        o.set_wrinkle(sc > 0.0, w, tex.tobytes(), tw, th)
        g.step(0.016, 2, 3)
        o.step(0.016, 2, 3)
        assert np.array_equal(_gpos(g), o.get_positions().reshape(-1, 3)), fr
    assert seen == [0.0, 0.0, 0.5, 1.0, 1.0]
    # the field acts on these vertices: the same run with the priority off ends elsewhere
    h = T.ClothSimulator(pos, np.zeros((0, 2), np.uint32), None, None, sewing_springs=sew)
    h.set_gravity(0.0, 0.0, 0.0)
    h.set_enable_self_collision(False)
    h.set_enable_wrinkle_field(True)
    h.set_wrinkle_field_params([0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                               influence_radius=0.05, bone_radius=0.05, stiffness=3.0, blend_weight=0.9)
    h.set_wrinkle_field_texture_2d(tw, th, tex.tobytes(), -0.2, 0.2, 0.02, 0.16)
    for _ in range(5):
        h.step(0.016, 2, 3)
    assert not np.array_equal(_gpos(h), _gpos(g))


@pytest.mark.gpu
def test_gpu_sewing_priority_record_replay(require_gpu, tmp_path):
    from taremincloth_amd import debug_recorder as DR
    pos, e, f, sew = sewn_panels(6, gap=0.02, seed=4)
    g = T.ClothSimulator(pos, e, f, None, sewing_springs=sew, sewing_shrink_speed=0.8)
    g.set_sewing_priority_options(True, 0.9, 0.004, 2, 0)
    g.set_solver_iterations(5)
    g.start_debug_recording("sewn", max_frames=20)
    prev = pos
    for _ in range(6):
        g.update_sewing_priority(prev)
        g.step(1.0 / 60.0, 4)
        prev = _gpos(g)
    g.stop_debug_recording()
    path = g.save_debug_recording(str(tmp_path / "prio.jsonl.gz"))
    meta, frames = DR.load(path)
    assert meta["config"]["sewing_priority_enabled"] is True
    assert all(fr["sewing_priority"] is not None for fr in frames)
    n = 0
    for fr, p in DR.replay(path):
        assert np.array_equal(p, np.asarray(fr["positions"], np.float32)), fr["frame_index"]
        n += 1
    assert n == 6
