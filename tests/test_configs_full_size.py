"""The workload configurations at their defining sizes, on the GPU, against the oracle or properties.

- 256 x 256 cloth at 4 mm with self-collision (thickness 2 mm, hash cell 4 mm, device hash tables of
  2^20 buckets), sphere r 0.3 + capsule r 0.1, friction 0.3, 20 iterations: folded over itself so
  that every self-collision pass has contacts, and falling onto the sphere; 3 substeps against the
  fp32 oracle, and the hash tables bit for bit at the device table size.
- 128 x 128 cloth (1 cm) on the icosphere-6 collider (81,920 triangles, 5% noise), friction 0.4:
  substeps in contact against the oracle's scan of every triangle.
- The 128^3 dense (with its fp32 brick atlas, 16^3 bricks) and hierarchical bakes of that
  icosphere: the dense bake and the atlas (within 1e-3 voxel) on 10,000 sampled voxels
  against an independent FP64 brute force to half a binary16 ulp; the hierarchical bake bit for bit
  against the oracle's two passes, with its signs and error bound against the same brute force.
- 512 x 512 learned warm start: the network run inside the substep equals the standalone pass on the
  recorded features, and the recorded features equal the oracle's on a vertex sample.
"""
import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O
from oracle import warmstart_oracle as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[full-size] {msg}", flush=True)
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


def folded_256(spacing=0.004, gap=0.003, height=0.3035):
    """256 x 256 vertices: rows 0..127 flat at `height`, rows 128..255 folded back over them `gap`
    higher, the fold centred over the sphere's top; seeded 1e-4 m jitter breaks symmetry."""
    pos, edges, faces = cases.grid(256, 256, spacing, "xy")
    pos = pos.astype(np.float64)
    half = 127.5 * spacing
    y = pos[:, 1]
    upper = y > half
    pos[upper, 1] = 2 * half - y[upper] + spacing * 0.37
    pos[:, 2] = np.where(upper, height + gap, height)
    pos[:, 0] -= 127.5 * spacing
    pos[:, 1] -= half * 0.5
    pos = pos + cases.uniform_jitter(len(pos), 1e-4, 11)
    pos = pos.astype(np.float32)
    inv = cases.areal_inv_masses_np(pos, faces, 0.15)
    return pos, edges, faces, inv


def test_cfg2_256_self_collision_matches_oracle():
    T = _T()
    pos, e, f, inv = folded_256()
    _say(f"cfg2: {len(pos)} vertices, folded, falling on sphere + capsule")
    vel = np.zeros_like(pos)
    vel[:, 2] = -0.5
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.002, stiffness=1000.0, bending_stiffness=10.0)
    o = O.OracleSim(pos, e, f, inv, thickness=0.002, stiffness=1000.0, bending_stiffness=10.0)
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_compliances_raw(1e-8, 1e-6, 1e-5)
        s.add_sphere_collider([0.0, 0.0, 0.0], 0.3, 0.3, 0.0)
        s.add_capsule_collider([-0.3, 0.35, 0.1], [0.3, 0.35, 0.1], 0.1, 0.3, 0.0)
        s.set_enable_self_collision(True)
        s.set_hash_cell_size(0.004)
        s.set_solver_iterations(20)
    g.set_positions_and_velocities(pos, vel)
    o.set_state(x=pos.astype(np.float64), p=pos.astype(np.float64), v=vel.astype(np.float64))
    dt_sub = float(np.float32(np.float32(1.0 / 60.0) / np.float32(10.0)))
    for k in range(3):
        g.step_single_substep(dt_sub)
        o.step_single_substep(dt_sub)
        diff = float(np.abs(_pos(g) - o.get_positions().reshape(-1, 3)).max())
        _say(f"  substep {k + 1}: max |hip - oracle| = {diff:.3e} m")
        assert diff < 1e-5, diff
    st = o.contact_stats()
    _say(f"  oracle contacts {st}")
    assert st["vv"] + st["vt_prox"] + st["vt_ccd"] > 0 and st["ee_prox"] + st["ee_ccd"] > 0, st
    ht = g.hash_tables()
    assert ht["table_size"] == 1 << 20 and abs(ht["cell_size"] - 0.004) < 1e-9
    pts = g.predicted_positions()
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


def _icosphere6():
    v, f = cases.icosphere(6, radius=0.3, noise_amp=0.05, seed=0)
    return v, f


def test_cfg3_cloth_on_icosphere6_matches_oracle_full_scan():
    T = _T()
    v, fc = _icosphere6()
    tris = v[fc.astype(np.int64)]
    top = float(v[:, 2].max())
    pos, e, f, inv = cases.drape_grid(128, 0.01, top + 0.006, seed=3)
    _say(f"cfg3 collider: {len(pos)} vertices on {len(fc)} triangles, top at {top:.4f}")
    vel = np.zeros_like(pos)
    vel[:, 2] = -1.0
    g = T.ClothSimulator(pos, e, f, inv, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0)
    o = O.OracleSim(pos, e, f, inv, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0, precision="f32_omp")
    for s in (g, o):
        s.set_gravity(0.0, 0.0, -9.81)
        s.set_compliances_raw(1e-8, 1e-6, 1e-5)
        s.set_enable_self_collision(False)
        s.set_mesh_collider_triangles(tris, friction=0.4, thickness=0.002, restitution=0.0)
        s.set_solver_iterations(20)
    g.set_positions_and_velocities(pos, vel)
    o.set_state(x=pos.astype(np.float64), p=pos.astype(np.float64), v=vel.astype(np.float64))
    dt_sub = float(np.float32(np.float32(1.0 / 60.0) / np.float32(10.0)))
    for k in range(6):
        g.step_single_substep(dt_sub)
        o.step_single_substep(dt_sub)
        diff = float(np.abs(_pos(g) - o.get_positions().reshape(-1, 3)).max())
        _say(f"  substep {k + 1}: max |hip - oracle| = {diff:.3e} m")
        assert diff == 0.0, diff
    # in contact: the same cloth without the collider ends elsewhere
    free = T.ClothSimulator(pos, e, f, inv, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0)
    free.set_gravity(0.0, 0.0, -9.81)
    free.set_compliances_raw(1e-8, 1e-6, 1e-5)
    free.set_enable_self_collision(False)
    free.set_solver_iterations(20)
    free.set_positions_and_velocities(pos, vel)
    for _ in range(6):
        free.step_single_substep(dt_sub)
    pushed = (_pos(g)[:, 2] - _pos(free)[:, 2])
    _say(f"  {int((pushed > 1e-4).sum())} vertices held up by the mesh (up to {pushed.max() * 1e3:.2f} mm)")
    assert (pushed > 1e-4).sum() > 100


def _nearest_fp64(points, v, t):
    """Distance and signed side of the nearest triangle for each point, in FP64: the candidates are
    the triangles whose centroid lies within (nearest-centroid distance + 2 x the largest
    centroid-to-vertex radius), a superset of every triangle that can be nearest."""
    from scipy.spatial import cKDTree
    sys.path.insert(0, cases.ROOT + "/tests")
    from test_sdf import brute_force
    A = v[t[:, 0]].astype(np.float64)
    B = v[t[:, 1]].astype(np.float64)
    Cc = v[t[:, 2]].astype(np.float64)
    cen = (A + B + Cc) / 3.0
    rad = np.max([np.linalg.norm(X - cen, axis=1) for X in (A, B, Cc)], axis=0).max()
    tree = cKDTree(cen)
    d0, _ = tree.query(points)
    dist = np.empty(len(points))
    sign = np.empty(len(points))
    ambiguous = np.zeros(len(points), bool)
    for i, p in enumerate(points):
        cand = np.asarray(tree.query_ball_point(p, d0[i] + 2 * rad), np.int64)
        d, q, nn = brute_force(p[None], v, t[cand])
        d, q, nn = d[0], q[0], nn[0]
        j = int(np.argmin(d))
        dist[i] = d[j]
        s = np.sign(np.einsum("tk,tk->t", p - q, nn))
        s = np.where(s == 0, 1, s)
        near = d <= d[j] + 1e-5
        sign[i] = s[j]
        ambiguous[i] = not np.all(s[near] == s[j]) or d[j] < 1e-6
    return dist, sign, ambiguous


def test_cfg3_bakes_128_against_fp64_brute_force():
    T = _T()
    v, fc = _icosphere6()
    fi = fc.astype(np.int32)
    margin, thickness = 0.02, 0.005
    lo, hi = v.min(0), v.max(0)
    size = np.maximum(hi - lo, 0.05)
    padded = size + 2 * (np.maximum(size * margin, 0.04) + thickness + 0.02)
    voxel = float(padded.max()) / 128.0 * (1.0 + 1e-6)
    dense = T.bake_mesh_sdf_atlas_fp32(v, fi, voxel_size=voxel, margin=margin, thickness=thickness)
    hier = T.bake_mesh_sdf_hierarchical_gpu(v, fi, voxel_size=voxel, margin=margin, thickness=thickness)
    dims = np.array([dense["width"], dense["height"], dense["depth"]])
    _say(f"cfg3 bakes: dims {dims.tolist()}, voxel {voxel * 1e3:.4f} mm, {len(fc)} triangles")
    assert dims.max() == 128
    sys.path.insert(0, cases.ROOT + "/tests")
    from test_sdf import check_atlas, domain, f16_decode, f16_halfulp
    assert dense["brick_dims"] == (16, 16, 16)  # 4,096 bricks of 8^3
    atlas = check_atlas(dense, dims).ravel()
    dlo, dhi, _, ddims = domain(v, voxel, margin, thickness)
    assert np.array_equal(np.asarray(ddims), dims)
    rng = np.random.default_rng(0)
    idx = rng.choice(int(np.prod(dims)), 10000, replace=False)
    iz, rem = np.divmod(idx, dims[0] * dims[1])
    iy, ix = np.divmod(rem, dims[0])
    ijk = np.stack([ix, iy, iz], 1).astype(np.float64)
    p = np.asarray(dlo, np.float64) + (ijk + 0.5) / dims * (np.asarray(dhi, np.float64) - np.asarray(dlo, np.float64))
    dist, sign, amb = _nearest_fp64(p, v.astype(np.float64), fc.astype(np.int64))
    got = f16_decode(np.frombuffer(dense["texture_bytes"], np.uint32)[idx])
    err = np.abs(np.abs(got) - dist)
    _say(f"  dense: max | |d| - fp64 | = {err.max():.3e} m, half-ulp bound held by all; "
         f"{amb.sum()} sign-ambiguous voxels")
    assert np.all(err <= f16_halfulp(dist) + 1e-6)
    gs = np.where(got >= 0, 1, -1)
    assert np.array_equal(gs[~amb], sign[~amb].astype(int))
    vsz = float(np.min((np.asarray(dhi, np.float64) - np.asarray(dlo, np.float64)) / dims))
    aerr = np.abs(np.abs(atlas[idx].astype(np.float64)) - dist)
    _say(f"  fp32 atlas: max | |d| - fp64 | = {aerr.max():.3e} m = {aerr.max() / vsz:.2e} voxel")
    assert aerr.max() <= 1e-3 * vsz
    assert np.array_equal(np.where(atlas[idx] >= 0, 1, -1)[~amb], sign[~amb].astype(int))
    # hierarchical: bit-equal to the oracle's restatement of the reference's two passes at full size.
    # Its near field is not the dense bake's on this mesh: the fine pass sees only the 8 nearest
    # triangles of each coarse cell within +-2 cells, which on 81,920 small triangles can miss the
    # nearest one (bake_mesh_sdf_hier_fine_main.wgsl:5-9 accepts this); what the reference's
    # equivalence tests pin (cube, double wall) is in tests/test_sdf.py.
    want = O.bake_mesh_sdf(v, fi, voxel, margin, 0.5, thickness, 0.0, True, 1, precision="f32_omp")
    ht_ = np.frombuffer(hier["texture_bytes"], np.uint32)
    mism = np.flatnonzero(ht_ != want["texels"])
    assert mism.size == 0, f"{mism.size} texels differ from the oracle, first at {mism[:5]}"
    hv = f16_decode(ht_[idx])
    herr = np.abs(np.abs(hv) - dist)
    exact = herr <= f16_halfulp(dist) + 1e-6
    band = dist <= 4 * voxel
    _say(f"  hierarchical: bit-equal to the oracle; vs fp64 within half an ulp at {exact.mean():.3f} of the "
         f"sampled voxels ({exact[band].mean():.3f} within 4 voxels of the surface), max error {herr.max():.3e} m")
    assert np.array_equal(np.where(hv >= 0, 1, -1)[~amb], sign[~amb].astype(int))
    assert herr.max() <= 4 * voxel * np.sqrt(3.0)


def test_cfg5_512_network_in_substep_matches_standalone_and_features_match_oracle():
    from taremincloth_amd.warmstart import WarmStartNet
    T = _T()
    pos, e, f, inv = cases.drape_grid(512, 0.01, 0.5, seed=0)
    _say(f"cfg5: {len(pos)} vertices")
    sim = T.ClothSimulator(pos, e, f, inv, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0)
    sim.set_gravity(0.0, 0.0, -9.81)
    sim.set_compliances_raw(1e-8, 1e-6, 1e-5)
    sim.set_enable_self_collision(False)
    sim.add_sphere_collider([0.0, 0.0, 0.0], 0.3, 0.3, 0.0)
    sim.add_capsule_collider([-0.3, 0.35, 0.1], [0.3, 0.35, 0.1], 0.1, 0.3, 0.0)
    net = WarmStartNet.load(cases.ROOT + "/taremincloth_amd/weights/warmstart_r3.npz")
    sim.set_warm_start(net, 2, 8, 0, 1e9)
    for _ in range(24):   # 0.4 s: the cloth is on the colliders
        sim.step(1.0 / 60.0, 10)
    sim.ws_record(1, 1, net.history_weight)
    dt_sub = 1.0 / 600.0
    sim.step_single_substep(dt_sub)
    r = sim.ws_get_record(0)
    y_sim = sim.ws_last_output()
    m = O.build_mesh(pos, e, f, inv, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0)
    stretch = m["dist_type"] == 0
    off, nbr, _ = W.csr_pairs(len(pos), m["dist_v0"][stretch], m["dist_v1"][stretch])
    y_ref, _, _ = net.forward_backward(r["features"], off, nbr)
    assert np.array_equal(y_sim, y_ref)
    assert np.abs(y_sim[:, :3]).max() > 0
    # features of a vertex sample against the oracle (FP64, rounded to BF16)
    s = 9.81 * dt_sub * dt_sub
    cols = [(0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.3), (1, (-0.3, 0.35, 0.1), (0.3, 0.35, 0.1), 0.1)]
    want = W.bf16_round(W.features(r["prediction"], r["previous"], np.full(len(pos), 0.005), m, cols, s,
                                   m["mean_edge"], net.history_weight, fp32_points=True))
    sample = np.random.default_rng(1).choice(len(pos), 20000, replace=False)
    err = np.abs(r["features"][sample] - want[sample])
    _say(f"  features: max error {err.max():.3e} on {len(sample)} vertices; "
         f"{int((np.abs(want[sample, 12]) < W.soft(4.0) - 1e-3).sum())} near a collider")
    bad = np.argwhere(err > 1e-2 * np.maximum(np.abs(want[sample]), 1.0))
    for r_, c_ in bad[:8]:
        i_ = sample[r_]
        _say(f"  vertex {i_} column {c_}: want {want[i_, c_]!r} got {r['features'][i_, c_]!r}")
    assert len(bad) == 0, err.max()
    assert (np.abs(want[sample, 12]) < W.soft(4.0) - 1e-3).sum() > 100  # the sample includes contact
