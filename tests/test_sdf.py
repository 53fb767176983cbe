"""Mesh SDF bake (dense, coarse top-8, hierarchical), open-boundary capping and SDF colliders.

CPU tests pin the oracle (oracle/sdf_oracle.c):
  * binary16 conversion against numpy's float16 (round to nearest even) on every class of input;
  * the dense bake of the reference's unit cube against the closed-form box distance, and the
    domain / collider info against their closed forms (sdf_baker.rs:939-1013);
  * dense bakes of irregular meshes against an independent FP64 brute force (projection and
    edge distances, not Ericson's regions), sign where the nearest triangles agree;
  * the reference's coarse-cube expectations (sdf_baker.rs:1997-2055) plus the top-8 lists
    against the independent brute force;
  * the reference's hierarchical == dense property: near field bit-equal, far field same sign
    and bounded difference (sdf_baker.rs:2058-2133, tests/core/test_mesh_sdf_hierarchical.py);
  * the reference's open-cylinder capping counts (sdf_baker.rs:2135-2171), watertightness and
    orientation of the capped mesh, and a flat sheet whose rim stays open;
  * the reference's falling-vertex-on-cube SDF collider scene (tests/collider/
    test_mesh_sdf_baker.py:172-235) on the oracle simulator.
GPU tests: every bake of the HIP path bit-equal to the oracle's (texels and nearest lists), the
fp32 brick atlas within 1e-3 voxel of an FP64 brute force at every voxel (its values rounding to
the dense texels, its padding clamped to the edge), the capping host code equal to the oracle's, a >256 MiB bake against the closed form, and the SDF
collider in the simulator against the oracle (1e-5 m) and the reference scene's bounds.
"""
import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O

CUBE_V = np.array([[-0.5, -0.5, -0.5], [0.5, -0.5, -0.5], [0.5, 0.5, -0.5], [-0.5, 0.5, -0.5],
                   [-0.5, -0.5, 0.5], [0.5, -0.5, 0.5], [0.5, 0.5, 0.5], [-0.5, 0.5, 0.5]], np.float32)
CUBE_T = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                   [2, 3, 7], [2, 7, 6], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], np.int32)


def _say(msg):
    print(f"[sdf] {msg}", flush=True)
    sys.stdout.flush()


def double_wall():
    """Two 2 mm plates 3 mm apart (tests/core/test_mesh_sdf_hierarchical.py:70-96)."""
    def plate(z0):
        t, s = 0.002, 0.15
        v = np.array([[-s, -s, z0], [s, -s, z0], [s, s, z0], [-s, s, z0],
                      [-s, -s, z0 + t], [s, -s, z0 + t], [s, s, z0 + t], [-s, s, z0 + t]], np.float32)
        return v, CUBE_T.copy()
    v1, t1 = plate(0.0)
    v2, t2 = plate(0.005)
    return np.vstack([v1, v2]), np.vstack([t1, t2 + len(v1)]).astype(np.int32)


def open_cylinder(n_seg=8, r_bot=0.045, r_top=0.055, z_top=0.10):
    """The reference's flared open cylinder (sdf_baker.rs:2142-2164)."""
    th = np.arange(n_seg, dtype=np.float32) * np.float32(2 * np.pi / n_seg)
    bot = np.stack([r_bot * np.cos(th), r_bot * np.sin(th), np.zeros(n_seg)], 1)
    top = np.stack([r_top * np.cos(th), r_top * np.sin(th), np.full(n_seg, z_top)], 1)
    v = np.vstack([bot, top]).astype(np.float32)
    t = []
    for i in range(n_seg):
        b0, b1, t0, t1 = i, (i + 1) % n_seg, n_seg + i, n_seg + (i + 1) % n_seg
        t += [[b0, b1, t1], [b0, t1, t0]]
    return v, np.array(t, np.int32)


def lumpy_sphere(subdiv=2, seed=0):
    v, f = cases.icosphere(subdiv, radius=0.25, noise_amp=0.3, seed=seed)
    return v, f.astype(np.int32)


# ------------------------------------------------------------------ independent FP64 references
def domain(verts, voxel, margin, thickness):
    """mesh_sdf_domain (sdf_baker.rs:939-980) in closed form, f32 arithmetic."""
    v = np.asarray(verts, np.float32)
    lo, hi = v.min(0), v.max(0)
    size = np.maximum(hi - lo, np.float32(0.05))
    pad = (np.maximum(size * np.float32(margin), np.float32(0.04)) + np.float32(max(thickness, 0.0))
           + np.float32(0.02)).astype(np.float32)
    lo, hi = (lo - pad).astype(np.float32), (hi + pad).astype(np.float32)
    size = (hi - lo).astype(np.float32)
    dims = np.clip(np.ceil((size / np.float32(max(voxel, 0.001))).astype(np.float64)), 8, 2048).astype(int)
    return lo, hi, size, dims


def voxel_centres(lo, hi, dims):
    axes = [(lo[a] + ((np.arange(dims[a]) + 0.5) / dims[a]) * (hi[a] - lo[a])) for a in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float64)


def seg_dist(p, a, b):
    ab = b - a
    t = np.clip(np.einsum("...k,...k->...", p - a, ab) / np.maximum(np.einsum("...k,...k->...", ab, ab), 1e-300),
                0.0, 1.0)
    q = a + t[..., None] * ab
    return np.linalg.norm(p - q, axis=-1), q


def brute_force(points, verts, tris):
    """(M, T) FP64 distances and closest points: plane projection when it lands inside the
    triangle, else the nearest of the three edges."""
    P = np.asarray(points, np.float64)[:, None, :]
    A = np.asarray(verts, np.float64)[tris[:, 0]][None]
    B = np.asarray(verts, np.float64)[tris[:, 1]][None]
    Cc = np.asarray(verts, np.float64)[tris[:, 2]][None]
    n = np.cross(B - A, Cc - A)
    nn = n / np.maximum(np.linalg.norm(n, axis=-1, keepdims=True), 1e-300)
    h = np.einsum("mtk,mtk->mt", P - A, nn)
    proj = P - h[..., None] * nn
    # inside test by the signs of the three sub-triangle normals
    def side(u, v):
        return np.einsum("mtk,mtk->mt", np.cross(v - u, proj - u), n)
    inside = (side(A, B) >= 0) & (side(B, Cc) >= 0) & (side(Cc, A) >= 0)
    best_d = np.where(inside, np.abs(h), np.inf)
    best_q = np.where(inside[..., None], proj, 0.0)
    for u, v in ((A, B), (B, Cc), (Cc, A)):
        d, q = seg_dist(P, u, v)
        take = d < best_d
        best_d = np.where(take, d, best_d)
        best_q = np.where(take[..., None], q, best_q)
    return best_d, best_q, nn


def f16_decode(texels):
    return (np.asarray(texels, np.uint32) & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64)


def f16_halfulp(x):
    """Half a binary16 ulp at |x| (normal range), for a rounding-error bound."""
    ax = np.maximum(np.abs(x), 2.0 ** -14)
    return 2.0 ** (np.floor(np.log2(ax)) - 11)


def box_sdf(p, half=0.5):
    q = np.abs(p) - half
    out = np.linalg.norm(np.maximum(q, 0.0), axis=1)
    inside = np.minimum(q.max(1), 0.0)
    return out + inside


# ------------------------------------------------------------------ CPU: oracle pinned
def test_f16_conversion_matches_numpy():
    rng = np.random.default_rng(1)
    bits = rng.integers(0, 2 ** 32, 200_000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    special = np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e-8, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                        2.0 ** -14, 2.0 ** -14 * (1 - 2.0 ** -11), 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11,
                        -2.5 * 2.0 ** -24, np.inf, -np.inf], np.float32)
    # values near binary16 rounding ties over the whole range
    ties = (np.arange(-24, 16, dtype=np.float64)[:, None] + 0 * np.arange(4))
    tie_vals = (2.0 ** ties * (1 + (np.arange(4) * 2 + 1) * 2.0 ** -11)).astype(np.float32).ravel()
    for vals in (x, special, tie_vals, -tie_vals, np.linspace(-3, 3, 100_001, dtype=np.float32)):
        with np.errstate(over="ignore"):
            want = vals.astype(np.float16).view(np.uint16)
        got = O.f32_to_f16_bits(vals)
        np.testing.assert_array_equal(got, want)
    allh = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    want = allh.view(np.float16).astype(np.float32)
    got = O.f16_bits_to_f32(allh)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    np.testing.assert_array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


@pytest.mark.parametrize("voxel,margin,friction,thickness,restitution",
                         [(0.1, 0.2, 0.5, 0.01, 0.0), (0.05, 0.2, 0.6, 0.01, 0.1), (0.07, 0.02, 0.5, 0.005, 0.0)])
def test_cube_dense_closed_form(voxel, margin, friction, thickness, restitution):
    r = O.bake_mesh_sdf(CUBE_V, CUBE_T, voxel, margin, friction, thickness, restitution, True, 0)
    lo, hi, size, dims = domain(CUBE_V, voxel, margin, thickness)
    assert (r["width"], r["height"], r["depth"]) == tuple(dims)
    assert len(r["texture_bytes"]) == 4 * dims.prod()
    bi = r["bone_info"]
    np.testing.assert_array_equal(bi[0:3], lo)
    np.testing.assert_array_equal(bi[4:7], hi)
    np.testing.assert_array_equal(bi[8:11], (np.float32(1) / size).astype(np.float32))
    np.testing.assert_array_equal(bi[12:15], (-lo / size).astype(np.float32))
    assert bi[3] == 0.0 and bi[7] == np.float32(dims[0]) and bi[15] == 0.0 and bi[19] == 0.0
    assert (bi[16], bi[17], bi[18]) == (np.float32(friction), np.float32(thickness), np.float32(restitution))
    tex = r["texels"]
    assert np.all((tex >> 16) == 0x3C00)  # second channel: 1.0
    p = voxel_centres(lo.astype(np.float64), hi.astype(np.float64), dims)
    exact = box_sdf(p)
    got = f16_decode(tex)
    err = np.abs(got - exact)
    assert np.all(err <= f16_halfulp(exact) + 1e-6), float((err - f16_halfulp(exact)).max())
    c = (dims // 2)
    assert got[(c[2] * dims[1] + c[1]) * dims[0] + c[0]] < 0 < got[0]


@pytest.mark.parametrize("mesh", ["lumpy", "double_wall", "cylinder"])
def test_dense_against_independent_brute_force(mesh):
    if mesh == "lumpy":
        v, t = lumpy_sphere(2)
        voxel, close = 0.03, True
    elif mesh == "double_wall":
        v, t = double_wall()
        voxel, close = 0.012, True
    else:
        v, t = open_cylinder()
        voxel, close = 0.01, False
    r = O.bake_mesh_sdf(v, t, voxel, 0.02, 0.5, 0.005, 0.0, close, 0)
    lo, hi, _, dims = domain(v, voxel, 0.02, 0.005)
    p = voxel_centres(lo.astype(np.float64), hi.astype(np.float64), dims)
    d, q, nn = brute_force(p, v, t)
    order = np.argsort(d, axis=1, kind="stable")
    best = order[:, 0]
    m = np.arange(len(p))
    dmin = d[m, best]
    got = f16_decode(r["texels"])
    assert np.all(np.abs(np.abs(got) - dmin) <= f16_halfulp(dmin) + 1e-6)
    # sign where every triangle within 1e-5 of the minimum gives the same sign
    sgn = np.sign(np.einsum("mtk,mtk->mt", p[:, None, :] - q, nn))
    sgn[sgn == 0] = 1
    near = d <= dmin[:, None] + 1e-5
    agree = np.all(np.where(near, sgn == sgn[m, best][:, None], True), axis=1) & (dmin > 1e-6)
    assert agree.mean() > 0.9
    gs = np.where(got >= 0, 1, -1)
    assert np.array_equal(gs[agree], sgn[m, best][agree].astype(int))


def test_coarse_cube_reference_expectations():
    r = O.bake_mesh_sdf(CUBE_V, CUBE_T, 0.05, 0.2, 0.5, 0.01, 0.0, True, 2)
    assert (r["width"], r["height"], r["depth"]) == (8, 8, 8)
    assert abs(r["coarse_voxel"] - 0.2) < 1e-6
    near = r["nearest_tri"]
    assert r["texels"].size == 512 and near.size == 512 * 8
    assert np.all(near < 12)
    flat = lambda x, y, z: (z * 8 + y) * 8 + x  # noqa: E731
    assert (r["texels"][flat(0, 0, 0)] & 0x8000) == 0
    assert (r["texels"][flat(4, 4, 4)] & 0x8000) != 0
    # the eight lists are the eight nearest triangles, nearest first
    lo, hi, _, dims = domain(CUBE_V, 0.2, 0.2, 0.01)
    p = voxel_centres(lo.astype(np.float64), hi.astype(np.float64), dims)
    d, _, _ = brute_force(p, CUBE_V, CUBE_T)
    dl = np.take_along_axis(d, near.reshape(-1, 8).astype(np.int64), 1)
    assert np.all(np.diff(dl, axis=1) >= -1e-6)
    np.testing.assert_allclose(dl, np.sort(d, axis=1)[:, :8], atol=1e-6)
    assert all(len(set(row)) == 8 for row in near.reshape(-1, 8).tolist())
    # Tie order, traced by hand through the reference's update (strict "<", a pushed-down entry
    # passes equal ones) at cell (0, 0, 0): triangles 0,1,4,5,8,9 are at the near distance s,
    # 2,3,6,10,11 at L and 7 at XL.  After 0..7: [0,1,4,5,2,3,6,7]; 8 pushes 2 past 3 and 6 into
    # the last slot (7 drops): [0,1,4,5,8,3,6,2]; 9 pushes 3 past 6 and 2 (3 drops):
    # [0,1,4,5,8,9,6,2]; 10 and 11 are not below the last entry.
    assert near.reshape(-1, 8)[flat(0, 0, 0)].tolist() == [0, 1, 4, 5, 8, 9, 6, 2]


def _near_far(dense, hier):
    ld, hd = f16_decode(dense["texels"]), f16_decode(hier["texels"])
    lp, hp = dense["texels"] & 0xFFFF, hier["texels"] & 0xFFFF
    return ld, hd, lp, hp


@pytest.mark.parametrize("mesh,voxel,margin,thickness,near_b,far_b", [
    ("cube", 0.05, 0.2, 0.01, 0.21, 0.18),      # sdf_baker.rs:2058-2133
    ("cube", 0.05, 0.02, 0.005, 0.21, 0.18),    # test_mesh_sdf_hierarchical.py:147-150
    ("double_wall", 0.004, 0.02, 0.005, 0.02, 0.03),  # test_mesh_sdf_hierarchical.py:152-156
])
def test_hierarchical_matches_dense(mesh, voxel, margin, thickness, near_b, far_b):
    v, t = (CUBE_V, CUBE_T) if mesh == "cube" else double_wall()
    dense = O.bake_mesh_sdf(v, t, voxel, margin, 0.5, thickness, 0.0, True, 0)
    hier = O.bake_mesh_sdf(v, t, voxel, margin, 0.5, thickness, 0.0, True, 1)
    assert (hier["width"], hier["height"], hier["depth"]) == (dense["width"], dense["height"], dense["depth"])
    np.testing.assert_array_equal(hier["bone_info"], dense["bone_info"])
    ld, hd, lp, hp = _near_far(dense, hier)
    near = np.abs(ld) <= near_b
    assert np.array_equal(lp[near], hp[near])
    assert np.array_equal(ld[~near] < 0, hd[~near] < 0)
    assert np.abs(ld[~near] - hd[~near]).max(initial=0.0) <= far_b


def _directed_edges(t):
    return np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])


def test_open_cylinder_capping():
    v, t = open_cylinder()
    cv, ct, loops = O.close_open_boundary_loops(v, t)
    assert loops == 2
    assert len(cv) == len(v) + 2 and len(ct) == len(t) + 16
    np.testing.assert_array_equal(cv[: len(v)], v)
    np.testing.assert_array_equal(ct[: len(t)], t)
    # watertight and consistently oriented: each directed edge once, and its reverse present
    de = _directed_edges(ct)
    keys = set(map(tuple, de.tolist()))
    assert len(keys) == len(de)
    assert all((b, a) in keys for a, b in keys)
    # cap centres are the loop centroids
    np.testing.assert_allclose(cv[len(v)], v[:8].mean(0), atol=1e-7)
    np.testing.assert_allclose(cv[len(v) + 1], v[8:].mean(0), atol=1e-7)
    # a closed surface has no loops left to close
    _, _, again = O.close_open_boundary_loops(cv, ct)
    assert again == 0


def test_flat_sheet_rim_stays_open():
    pos, _, faces = cases.grid(6, 6, 0.1)
    _, ct, loops = O.close_open_boundary_loops(pos, faces.astype(np.int32))
    assert loops == 0 and len(ct) == len(faces)


def test_open_cylinder_closed_signs():
    v, t = open_cylinder()
    r = O.bake_mesh_sdf(v, t, 0.005, 0.35, 0.5, 0.01, 0.0, True, 0)
    lo, hi, _, dims = domain(v, 0.005, 0.35, 0.01)
    p = voxel_centres(lo.astype(np.float64), hi.astype(np.float64), dims)
    d = f16_decode(r["texels"])
    # face-normal signs (no pseudonormals in the mesh bake): inside the capped neck negative,
    # beside it positive
    outside = (np.hypot(p[:, 0], p[:, 1]) > 0.08) & (p[:, 2] > 0.02) & (p[:, 2] < 0.08)
    assert outside.any() and np.all(d[outside] > 0)
    inside = (np.hypot(p[:, 0], p[:, 1]) < 0.03) & (p[:, 2] > 0.02) & (p[:, 2] < 0.08)
    assert inside.any() and np.all(d[inside] < 0)


def _falling_vertex(sim_cls, sdf):
    sim = sim_cls(np.array([[0.0, 0.0, 0.8]], np.float32), np.empty((0, 2), np.uint32), thickness=0.01)
    sim.set_bone_sdf_colliders(sdf["width"], sdf["height"], sdf["depth"], sdf["texture_bytes"],
                               sdf["bone_info"].reshape(1, 20))
    eye = np.eye(4, dtype=np.float32).reshape(1, 4, 4)
    sim.update_bone_transforms(eye, eye)
    return sim


def test_oracle_vertex_rests_on_cube_sdf():
    """tests/collider/test_mesh_sdf_baker.py:172-235 on the oracle: rest height in (0.50, 0.55)."""
    sdf = O.bake_mesh_sdf(CUBE_V, CUBE_T, 0.05, 0.2, 0.3, 0.01, 0.0, True, 0)
    sim = _falling_vertex(O.OracleSim, sdf)
    for _ in range(60):
        sim.step(1.0 / 60.0, 20)
    z = float(sim.get_positions()[2])
    assert 0.50 < z < 0.55, z


# ------------------------------------------------------------------ GPU: HIP path vs oracle
def _T():
    from taremincloth_amd import taremin_cloth_core as T
    return T


def _product_bake(T, method, v, t, voxel, margin, friction, thickness, restitution, close):
    fn = {0: T.bake_mesh_sdf_gpu, 1: T.bake_mesh_sdf_hierarchical_gpu, 2: T.bake_mesh_sdf_coarse}[method]
    return fn(v, t, voxel_size=voxel, margin=margin, friction=friction, thickness=thickness,
              restitution=restitution, close_open_boundaries=close)


GPU_BAKES = [
    ("cube", 0.05, 0.2, 0.01, True),
    ("cube", 0.013, 0.02, 0.005, True),
    ("double_wall", 0.004, 0.02, 0.005, True),
    ("lumpy", 0.02, 0.02, 0.005, True),
    ("cylinder", 0.004, 0.35, 0.01, True),
    ("cylinder", 0.004, 0.35, 0.01, False),
]


def _mesh(name):
    return {"cube": lambda: (CUBE_V, CUBE_T), "double_wall": double_wall, "lumpy": lambda: lumpy_sphere(3),
            "cylinder": open_cylinder}[name]()


@pytest.mark.gpu
@pytest.mark.parametrize("method", [0, 2, 1])
@pytest.mark.parametrize("mesh,voxel,margin,thickness,close", GPU_BAKES)
def test_gpu_bake_bit_exact(require_gpu, method, mesh, voxel, margin, thickness, close):
    T = _T()
    v, t = _mesh(mesh)
    _say(f"bake method={method} mesh={mesh} voxel={voxel} close={close}")
    want = O.bake_mesh_sdf(v, t, voxel, margin, 0.4, thickness, 0.1, close, method)
    if want is None:
        with pytest.raises(RuntimeError):
            _product_bake(T, method, v, t, voxel, margin, 0.4, thickness, 0.1, close)
        return
    got = _product_bake(T, method, v, t, voxel, margin, 0.4, thickness, 0.1, close)
    _say(f"  fallback voxels {got.get('fallback_voxels')}, kernel ms {got.get('kernel_ms')}")
    assert (got["width"], got["height"], got["depth"]) == (want["width"], want["height"], want["depth"])
    np.testing.assert_array_equal(np.asarray(got["bone_info"], np.float32), want["bone_info"])
    gt = np.frombuffer(got["texture_bytes"], np.uint32)
    mism = np.flatnonzero(gt != want["texels"])
    assert mism.size == 0, f"{mism.size} texels differ, first at {mism[:5]}"
    if method == 2:
        np.testing.assert_array_equal(got["nearest_tri"], want["nearest_tri"])
        assert abs(got["coarse_voxel"] - want["coarse_voxel"]) == 0.0


def check_atlas(r, voxel_dims):
    """The fp32 brick atlas of a method-3 bake: brick counts, clamp-to-edge padding, and every voxel's
    value rounding (binary16, nearest even) to the dense texel beside it.  Returns the atlas as a
    dense (d, h, w) array."""
    from taremincloth_amd.taremin_cloth_core import atlas_to_dense
    w, h, d = r["width"], r["height"], r["depth"]
    bd = r["brick_dims"]
    assert bd == ((w + 7) // 8, (h + 7) // 8, (d + 7) // 8)
    a = np.asarray(r["atlas"], np.float32)
    assert a.size == bd[0] * bd[1] * bd[2] * 512 and np.isfinite(a).all()
    full = a.reshape(bd[2], bd[1], bd[0], 8, 8, 8).transpose(0, 3, 1, 4, 2, 5).reshape(bd[2] * 8, bd[1] * 8,
                                                                                     bd[0] * 8)
    dense = atlas_to_dense(a, bd, (w, h, d))
    zi = np.minimum(np.arange(bd[2] * 8), d - 1)
    yi = np.minimum(np.arange(bd[1] * 8), h - 1)
    xi = np.minimum(np.arange(bd[0] * 8), w - 1)
    np.testing.assert_array_equal(full, dense[zi][:, yi][:, :, xi])
    tex = np.frombuffer(r["texture_bytes"], np.uint32)
    assert np.all((tex >> 16) == 0x3C00)
    packed = dense.ravel().astype(np.float16).view(np.uint16).astype(np.uint32)
    mism = np.flatnonzero(packed != (tex & 0xFFFF))
    assert mism.size == 0, f"{mism.size} atlas values do not round to their texel, first at {mism[:5]}"
    return dense


@pytest.mark.gpu
@pytest.mark.parametrize("mesh,voxel,close", [("lumpy", 0.02, True), ("double_wall", 0.004, True),
                                              ("cylinder", 0.004, False), ("cube", 0.013, True)])
def test_gpu_atlas_fp32_against_fp64(require_gpu, mesh, voxel, close):
    """The fp32 atlas within 1e-3 voxel of the FP64 distance to the nearest triangle at every voxel,
    with the FP64 sign wherever the nearest triangles agree on it; its texels equal the dense bake's."""
    T = _T()
    v, t = _mesh(mesh)
    _say(f"atlas mesh={mesh} voxel={voxel} close={close}")
    r = T.bake_mesh_sdf_atlas_fp32(v, t, voxel_size=voxel, margin=0.02, thickness=0.005,
                                   close_open_boundaries=close)
    d0 = T.bake_mesh_sdf_gpu(v, t, voxel_size=voxel, margin=0.02, thickness=0.005, close_open_boundaries=close)
    assert r["texture_bytes"] == d0["texture_bytes"]
    lo, hi, size, dims = domain(v, voxel, 0.02, 0.005)
    assert (r["width"], r["height"], r["depth"]) == tuple(dims)
    dense = check_atlas(r, dims).ravel().astype(np.float64)
    vsz = float(np.min((hi.astype(np.float64) - lo.astype(np.float64)) / dims))
    vv, tt = v, t  # closed meshes, or the open cylinder baked uncapped: the bake sees exactly (v, t)
    p = voxel_centres(lo.astype(np.float64), hi.astype(np.float64), dims)
    err = np.empty(len(p))
    amb = np.zeros(len(p), bool)
    ok_sign = np.ones(len(p), bool)
    for c0 in range(0, len(p), 4096):
        sl = slice(c0, c0 + 4096)
        d, q, nn = brute_force(p[sl], vv, tt)
        m = np.arange(d.shape[0])
        best = np.argmin(d, axis=1)
        dmin = d[m, best]
        err[sl] = np.abs(np.abs(dense[sl]) - dmin)
        sgn = np.sign(np.einsum("mtk,mtk->mt", p[sl][:, None, :] - q, nn))
        sgn[sgn == 0] = 1
        near = d <= dmin[:, None] + 1e-5
        amb[sl] = ~np.all(np.where(near, sgn == sgn[m, best][:, None], True), axis=1) | (dmin <= 1e-6)
        ok_sign[sl] = np.where(dense[sl] >= 0, 1, -1) == sgn[m, best]
    _say(f"  {len(p)} voxels, max | |d| - fp64 | = {err.max():.3e} m = {err.max() / vsz:.2e} voxel; "
         f"{amb.sum()} sign-ambiguous")
    assert err.max() <= 1e-3 * vsz
    assert ok_sign[~amb].all()


@pytest.mark.gpu
@pytest.mark.parametrize("mesh", ["cylinder", "sheet", "lumpy_open"])
def test_gpu_capping_matches_oracle(require_gpu, mesh):
    T = _T()
    if mesh == "cylinder":
        v, t = open_cylinder()
    elif mesh == "sheet":
        pos, _, faces = cases.grid(6, 6, 0.1)
        v, t = pos, faces.astype(np.int32)
    else:  # a sphere with two holes cut out
        v, t = lumpy_sphere(2)
        c = v[t].mean(1)
        t = t[(c[:, 2] < 0.2) & (c[:, 0] > -0.2)]
    _say(f"capping {mesh}")
    gv, gt, gl = T.close_open_boundary_loops(v, t)
    ov, ot, ol = O.close_open_boundary_loops(v, t)
    assert gl == ol
    np.testing.assert_array_equal(np.asarray(gv, np.float32).reshape(-1, 3), ov)
    np.testing.assert_array_equal(np.asarray(gt, np.int32).reshape(-1, 3), ot)


@pytest.mark.gpu
def test_gpu_bake_over_256mib_closed_form(require_gpu):
    """tests/collider/test_mesh_sdf_baker.py:237-262: a 0.8 m cube at 2 mm voxels (> 256 MiB)."""
    T = _T()
    s = 0.4
    v = CUBE_V * np.float32(2 * s)
    _say("bake 0.8 m cube at 2 mm")
    r = T.bake_mesh_sdf_gpu(v, CUBE_T, voxel_size=0.002, margin=0.02)
    lo, hi, _, dims = domain(v, 0.002, 0.02, 0.005)
    assert (r["width"], r["height"], r["depth"]) == tuple(dims)
    assert len(r["texture_bytes"]) / 2 ** 20 > 256.0
    tex = np.frombuffer(r["texture_bytes"], np.uint32)
    idx = np.random.default_rng(0).choice(tex.size, 400_000, replace=False)
    z, rem = np.divmod(idx, dims[0] * dims[1])
    y, x = np.divmod(rem, dims[0])
    c = np.stack([lo[a] + ((np.array([x, y, z][a]) + 0.5) / dims[a]) * (hi[a] - lo[a]) for a in range(3)], 1)
    exact = box_sdf(c.astype(np.float64), s)
    got = f16_decode(tex[idx])
    assert np.all(np.abs(got - exact) <= f16_halfulp(exact) + 1e-6)


@pytest.mark.gpu
def test_gpu_vertex_rests_on_cube_sdf(require_gpu):
    T = _T()
    _say("falling vertex on cube SDF")
    sdf = T.bake_mesh_sdf_gpu(CUBE_V, CUBE_T, voxel_size=0.05, margin=0.2, friction=0.3, thickness=0.01,
                              restitution=0.0)
    g = _falling_vertex(T.ClothSimulator, sdf)
    o = _falling_vertex(O.OracleSim, O.bake_mesh_sdf(CUBE_V, CUBE_T, 0.05, 0.2, 0.3, 0.01, 0.0, True, 0))
    out = np.empty(3, np.float32)
    for f in range(60):
        g.step(1.0 / 60.0, 20)
        o.step(1.0 / 60.0, 20)
        g.get_positions(out)
        assert np.abs(out - o.get_positions()).max() < 1e-5, f"frame {f}"
    assert 0.50 < float(out[2]) < 0.55


def _rigid(angle, axis, t):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = t
    return M


@pytest.mark.gpu
def test_gpu_cloth_on_transformed_sdf_matches_oracle(require_gpu):
    """A 24 x 24 sheet dropped on a rotated, translated lumpy-sphere SDF (with a second, distant
    collider so that the multi-collider slice clamp is used); positions within 1e-5 m of the oracle."""
    T = _T()
    v, t = lumpy_sphere(3, seed=3)
    sdf = O.bake_mesh_sdf(v, t, 0.015, 0.05, 0.4, 0.004, 0.1, True, 0)
    g_sdf = T.bake_mesh_sdf_gpu(v, t, voxel_size=0.015, margin=0.05, friction=0.4, thickness=0.004, restitution=0.1)
    assert g_sdf["texture_bytes"] == sdf["texture_bytes"]
    infos = np.stack([sdf["bone_info"], sdf["bone_info"]])
    infos[1, 15] = 0.3  # bounding-sphere radius: prefilter path
    infos[1, 3], infos[1, 7], infos[1, 11] = 5.0, 5.0, 5.0
    W0 = _rigid(0.4, (1, 2, 0.5), (0.05, -0.02, 0.0))
    W1 = _rigid(0.0, (0, 0, 1), (5.0, 5.0, 5.0))
    world = np.stack([W0.T, W1.T]).astype(np.float32)   # uploaded as given: result[k] = sum_j m[j][k] v[j]
    inv = np.stack([np.linalg.inv(W0).T, np.linalg.inv(W1).T]).astype(np.float32)
    pos, e, f = cases.grid(24, 24, 0.03, origin=(-0.345, -0.345, 0.4))
    sims = []
    for cls in (T.ClothSimulator, O.OracleSim):
        s = cls(pos, e, f, thickness=0.004)
        s.set_bone_sdf_colliders(sdf["width"], sdf["height"], sdf["depth"], sdf["texture_bytes"], infos)
        s.update_bone_transforms(world, inv)
        sims.append(s)
    _say("sheet on transformed SDF")
    out = np.empty(3 * len(pos), np.float32)
    for fi in range(12):
        for s in sims:
            s.step(1.0 / 60.0, 10, 2)
        sims[0].get_positions(out)
        diff = np.abs(out - sims[1].get_positions()).max()
        assert diff < 1e-5, f"frame {fi}: {diff:.3e}"
    # the sheet was caught: nothing fell through the collider's centre region
    p = out.reshape(-1, 3)
    assert p[:, 2].max() > 0.1


@pytest.mark.parametrize("method", [0, 1, 2])
def test_openmp_oracle_bake_equals_the_serial_bake(method):
    """The multithreaded oracle build bakes voxel slabs on several threads; texels (and the coarse
    pass's nearest lists) are the serial bake's bit for bit."""
    v, t = lumpy_sphere(2)
    a = O.bake_mesh_sdf(v, t, 0.02, 0.02, 0.5, 0.005, 0.0, True, method)
    b = O.bake_mesh_sdf(v, t, 0.02, 0.02, 0.5, 0.005, 0.0, True, method, precision="f32_omp")
    assert np.array_equal(a["texels"], b["texels"])
    if method == 2:
        assert np.array_equal(a["nearest_tri"], b["nearest_tri"])
