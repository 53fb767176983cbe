"""Bone SDF atlas bake: per bone, the signed distance to its skinned triangles in bone space on one
res^3 tile of a 3D atlas, with the bone weight interpolated at the nearest point
(sdf_baker.rs:494-930, bake_sdf.wgsl; Python entry lib.rs:1190-1348).

CPU tests pin the oracle (oracle/sdf_oracle.c):
  * compute_3d_atlas_layout: the reference's test (sdf_baker.rs:1885-1892) and closed forms,
    including the clamp of the tile size to 2048 texels per axis and its floor of 16;
  * a cube's texels against the closed-form box distance, in FP64 and FP32;
  * the weight channel against the linear weight field it interpolates (exact on planar triangles);
  * the bind matrix: a translated cube under the inverse translation bakes bit-equal to the cube at
    the origin, and the homogeneous divide happens only when |w - 1| > 1e-5;
  * the info row against its closed form (local box, margin and thickness pad, 5 cm floor, tile
    scale and offset), active-bone selection by the weight threshold;
  * loop capping with weights: the cap centre carries the loop's mean weight;
  * the open-rim sign rule: below a single sheet, voxels past the rim read outside;
  * the reference's ten-bone bake (sdf_baker.rs:2236-2353): centre voxels inside, tile corners
    outside, for every batch.
GPU tests: the HIP bake bit-equal to the FP32 oracle on all of these meshes (texels, infos, active
list), including a tile size that is not a multiple of the 4^3 brick, and a timed bake.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

H = 0.0625
BOX_V = np.array([[-H, -H, -H], [H, -H, -H], [H, H, -H], [-H, H, -H],
                  [-H, -H, H], [H, -H, H], [H, H, H], [-H, H, H]], np.float32)
BOX_T = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [1, 2, 6], [1, 6, 5], [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7]], np.int32)
EYE = np.eye(4, dtype=np.float32)


def _say(msg):
    print(f"[bone-sdf] {msg}", flush=True)
    sys.stdout.flush()


def _dist(tex):
    return (np.asarray(tex, np.uint32) & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64)


def _alpha(tex):
    return (np.asarray(tex, np.uint32) >> 16).astype(np.uint16).view(np.float16).astype(np.float64)


def _voxel_centres(info, res):
    """Bone-space voxel centres of a tile, (res, res, res, 3) in z, y, x order."""
    lmin = info[0:3].astype(np.float64)
    lmax = info[4:7].astype(np.float64)
    c = (np.arange(res) + 0.5) / res
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    u = np.stack([x, y, z], -1)
    return lmin + u * (lmax - lmin)


def _tile(r, k, res):
    """Texels of active bone k's tile as (res, res, res) in z, y, x order."""
    w, h, d = r["width"], r["height"], r["depth"]
    cols, rows = w // res, h // res
    l, rem = divmod(k, cols * rows)
    row, col = divmod(rem, cols)
    t = np.asarray(r["texels"]).reshape(d, h, w)
    return t[l * res:(l + 1) * res, row * res:(row + 1) * res, col * res:(col + 1) * res]


def _box_sdf(p, h):
    q = np.abs(p) - h
    out = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(q.max(-1), 0.0)
    return out + inside


def octagon_tube(r=0.05, z0=-0.08, z1=0.08):
    """An open tube of two symmetric octagon rings (8 quads, both ends open)."""
    a = np.float32(r) * np.float32(0.70710677)
    ring = np.array([[r, 0], [a, a], [0, r], [-a, a], [-r, 0], [-a, -a], [0, -r], [a, -a]], np.float32)
    v = np.vstack([np.c_[ring, np.full(8, z0, np.float32)], np.c_[ring, np.full(8, z1, np.float32)]])
    t = []
    for i in range(8):
        j = (i + 1) % 8
        t += [[i, j, 8 + j], [i, 8 + j, 8 + i]]
    return v.astype(np.float32), np.array(t, np.int32)


def ten_bones():
    """sdf_baker.rs:2241-2293: ten 0.1 m boxes 1 m apart, each fully weighted to its own bone."""
    h = 0.05
    lb = np.array([[-h, -h, -h], [h, -h, -h], [h, h, -h], [-h, h, -h],
                   [-h, -h, h], [h, -h, h], [h, h, h], [-h, h, h]], np.float32)
    v, t, w, m = [], [], np.zeros((10, 80), np.float32), []
    for b in range(10):
        v.append(lb + np.float32([b, 0.0, 0.0]))
        t.append(BOX_T + 8 * b)
        w[b, 8 * b:8 * b + 8] = 1.0
        mb = np.eye(4, dtype=np.float32)
        mb[0, 3] = -float(b)
        m.append(mb)
    return np.vstack(v), np.vstack(t).astype(np.int32), w, np.stack(m)


def _rot(ax, ay, az):
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ ry @ rx


def _bind(rot, shift):
    """Inverse of the placement x -> rot x + shift, as a row-major 4x4."""
    m = np.eye(4)
    m[:3, :3] = rot.T
    m[:3, 3] = -rot.T @ shift
    return m.astype(np.float32)


# ------------------------------------------------------------------ CPU: layout
def test_atlas_layout_reference_and_closed_forms():
    cols, rows, layers, res, w, h, d = O.atlas_layout(8, 32, 2048)  # sdf_baker.rs:1885-1892
    assert res == 32 and cols * rows * layers >= 8 and (w, h, d) == (cols * 32, rows * 32, layers * 32)
    assert O.atlas_layout(0, 64) == (0, 0, 0, 64, 0, 0, 0)
    assert O.atlas_layout(1, 64) == (1, 1, 1, 64, 64, 64, 64)
    assert O.atlas_layout(10, 16) == (3, 2, 2, 16, 48, 32, 32)     # ceil(10^(1/3)) = 3, ceil(sqrt(4)) = 2
    assert O.atlas_layout(200, 64) == (6, 6, 6, 64, 384, 384, 384)
    assert O.atlas_layout(8, 2048) == (2, 2, 2, 1024, 2048, 2048, 2048)  # 2 x 2048 > 2048: 2048 / 2
    assert O.atlas_layout(3_000_000, 64)[3] == 16                         # 2048 / 145 = 14, floored at 16
    for n in range(1, 300):
        c, r_, l, s, w, h, d = O.atlas_layout(n, 32)
        assert c * r_ * l >= n and c * r_ * (l - 1) < n and s == 32 and (w, h, d) == (32 * c, 32 * r_, 32 * l), n
        assert (c - 1) ** 3 < n <= c ** 3 + 1, n   # ceil of the cube root (pow may land an ulp high)


# ------------------------------------------------------------------ CPU: the oracle against closed forms
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_oracle_cube_matches_box_distance(precision):
    w = np.ones((1, 8), np.float32)
    r = O.bake_bone_sdf(BOX_V, BOX_T, w, EYE[None], resolution=16, margin=0.25, thickness=0.0,
                        close_open_boundaries=False, precision=precision)
    assert (r["width"], r["height"], r["depth"]) == (16, 16, 16)
    p = _voxel_centres(r["bone_infos"][0], 16)
    want = _box_sdf(p, H).astype(np.float16).astype(np.float64)
    got = _dist(r["texels"]).reshape(16, 16, 16)
    tol = 2.0 ** -10 * np.abs(want) + 1e-7
    assert np.all(np.abs(got - want) <= tol), np.abs(got - want).max()
    assert np.all(np.abs(_alpha(r["texels"]) - 1.0) <= 2.0 ** -10)


def test_oracle_weight_channel_interpolates_linear_field():
    """Weight 0.5 + 2x at each corner: on planar triangles barycentric interpolation is exact, so the
    channel reads 0.5 + 2 x_nearest, where outside the box the nearest point is the clamp of p."""
    w = (0.5 + 2.0 * BOX_V[:, 0]).astype(np.float32)[None]
    r = O.bake_bone_sdf(BOX_V, BOX_T, w, EYE[None], resolution=16, margin=0.3, thickness=0.0,
                        close_open_boundaries=False, precision="f64")
    p = _voxel_centres(r["bone_infos"][0], 16).reshape(-1, 3)
    outside = np.abs(p).max(1) > H + 1e-9
    want = 0.5 + 2.0 * np.clip(p[:, 0], -H, H)
    got = _alpha(r["texels"])
    assert outside.sum() > 2000
    assert np.abs(got[outside] - want[outside]).max() <= 2.0 ** -10
    # inside, the nearest face decides; the channel stays within the corner weights
    assert got[~outside].min() >= 0.375 - 1e-3 and got[~outside].max() <= 0.625 + 1e-3


def test_oracle_bind_matrix_translation_and_homogeneous_divide():
    w = np.ones((1, 8), np.float32)
    base = O.bake_bone_sdf(BOX_V, BOX_T, w, EYE[None], resolution=12, margin=0.2, close_open_boundaries=False)
    shift = np.float32([0.5, -0.25, 0.125])
    m = EYE.copy()
    m[:3, 3] = -shift
    moved = O.bake_bone_sdf(BOX_V + shift, BOX_T, w, m[None], resolution=12, margin=0.2,
                            close_open_boundaries=False)
    assert np.array_equal(moved["texels"], base["texels"])
    assert np.array_equal(moved["bone_infos"], base["bone_infos"])
    # w = 2 (every entry doubled): divided back, bit-equal
    dbl = O.bake_bone_sdf(BOX_V, BOX_T, w, (2.0 * EYE)[None], resolution=12, margin=0.2,
                          close_open_boundaries=False)
    assert np.array_equal(dbl["texels"], base["texels"])
    # |w - 1| <= 1e-5: no divide (x, y, z untouched, so bit-equal); |w - 1| > 1e-5: divided
    near = EYE.copy()
    near[3, 3] = 1.000004
    assert np.array_equal(O.bake_bone_sdf(BOX_V, BOX_T, w, near[None], resolution=12, margin=0.2,
                                          close_open_boundaries=False)["texels"], base["texels"])
    far = EYE.copy()
    far[3, 3] = 1.25
    rf = O.bake_bone_sdf(BOX_V, BOX_T, w, far[None], resolution=12, margin=0.2, close_open_boundaries=False)
    hw = H / 1.25
    assert np.allclose(rf["bone_infos"][0][4:7], hw + 2 * hw * 0.2 + 0.005, rtol=1e-6)


def test_oracle_info_row_and_active_bones():
    v, t = BOX_V * np.float32([1.0, 2.0, 0.1]), BOX_T   # z extent 0.0125 < 5 cm floor
    w = np.stack([np.ones(8), np.zeros(8), np.full(8, 0.02), np.where(np.arange(8) < 4, 0.5, 0.01)]).astype(np.float32)
    m = np.stack([EYE] * 4)
    margin, thick, wth = 0.2, 0.003, 0.02
    r = O.bake_bone_sdf(v, t, w, m, resolution=8, margin=margin, weight_threshold=wth, blend_k=0.07,
                        friction=0.4, thickness=thick, restitution=0.1, close_open_boundaries=False)
    assert list(r["active"]) == [0, 3]                 # 0.02 is not above the threshold
    cols, rows, layers, res, W, Hh, D = O.atlas_layout(2, 8)
    assert (r["width"], r["height"], r["depth"]) == (W, Hh, D)
    for k, (b, sel) in enumerate([(0, np.ones(8, bool)), (3, np.arange(8) < 4)]):
        lo = v[sel].astype(np.float64).min(0)
        hi = v[sel].astype(np.float64).max(0)
        size = np.maximum(hi - lo, 0.05)
        pad = size * margin + thick
        lo, hi = lo - pad, hi + pad
        size = hi - lo
        row = r["bone_infos"][k].astype(np.float64)
        assert np.allclose(row[0:3], lo, rtol=1e-6) and np.allclose(row[4:7], hi, rtol=1e-6)
        assert row[3] == k and row[7] == res and row[11] == np.float32(0.07) and row[15] == 0.0
        l_, rem = divmod(k, cols * rows)
        rr, cc = divmod(rem, cols)
        n = np.array([cols, rows, layers], np.float64)
        assert np.allclose(row[8:11], 1.0 / size / n, rtol=1e-5)
        assert np.allclose(row[12:15], (-lo / size + np.array([cc, rr, l_])) / n, rtol=1e-5, atol=1e-7)
        assert np.allclose(row[16:20], [0.4, thick, 0.1, wth], rtol=1e-6)
    assert O.bake_bone_sdf(v, t, np.zeros((2, 8), np.float32), m[:2], resolution=8) is None


def test_oracle_cap_centre_carries_mean_weight():
    """Tube along z with both ends capped: on the axis beyond the top cap the nearest point is the
    fan apex, whose weight is the mean of the ring's weights."""
    v, t = octagon_tube()
    rng = np.random.default_rng(5)
    w = np.zeros(16, np.float32)
    w[8:] = rng.uniform(0.2, 1.0, 8)
    w[:8] = 0.3
    r = O.bake_bone_sdf(v, t, w[None], EYE[None], resolution=15, margin=0.3, thickness=0.0,
                        close_open_boundaries=True, precision="f64")
    tile = _tile(r, 0, 15)
    p = _voxel_centres(r["bone_infos"][0], 15)
    assert abs(p[0, 7, 7, 0]) < 1e-9 and abs(p[0, 7, 7, 1]) < 1e-9
    top = [z for z in range(15) if p[z, 7, 7, 2] > 0.08 + 1e-3]
    assert top
    for z in top:
        a = _alpha(tile[z, 7, 7])
        assert abs(a - float(w[8:].astype(np.float64).mean())) < 2e-3, (z, a)
    # the capped tube is solid: the axis between the caps reads inside
    mid = [z for z in range(15) if abs(p[z, 7, 7, 2]) < 0.07]
    assert mid and all(_dist(tile[z, 7, 7]) < 0 for z in mid)
    # uncapped, axis voxels above the tube are nearest to the open rim: the face sign says inside,
    # and the rim rule turns it outside once the height above the rim exceeds a quarter of the
    # radius (0.0125 m)
    ro = O.bake_bone_sdf(v, t, w[None], EYE[None], resolution=15, margin=0.3, thickness=0.0,
                         close_open_boundaries=False, precision="f64")
    tro = _tile(ro, 0, 15)
    high = [z for z in top if p[z, 7, 7, 2] > 0.08 + 0.015]
    low = [z for z in top if p[z, 7, 7, 2] < 0.08 + 0.010]
    assert high and low
    assert all(_dist(tro[z, 7, 7]) > 0 for z in high) and all(_dist(tro[z, 7, 7]) < 0 for z in low)


def test_oracle_open_rim_sign_rule():
    """A 0.2 m square sheet at z = 0 facing +z: below the interior reads inside (face sign); below
    and past the rim by more than a quarter of the depth reads outside (bake_sdf.wgsl:150-172)."""
    s = 0.1
    v = np.array([[-s, -s, 0], [s, -s, 0], [s, s, 0], [-s, s, 0]], np.float32)
    t = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    r = O.bake_bone_sdf(v, t, np.ones((1, 4), np.float32), EYE[None], resolution=20, margin=0.5,
                        thickness=0.0, close_open_boundaries=True, precision="f64")
    p = _voxel_centres(r["bone_infos"][0], 20).reshape(-1, 3)
    d = _dist(r["texels"])
    below = p[:, 2] < -1e-6
    lat = np.abs(p[:, 0]) - s                         # past the x rim, y inside the sheet
    edge = below & (np.abs(p[:, 1]) < s - 1e-6)
    out = edge & (lat > 0.3 * np.abs(p[:, 2]))
    ins = below & (np.abs(p[:, 0]) < s - 1e-6) & (np.abs(p[:, 1]) < s - 1e-6)
    assert out.sum() > 50 and ins.sum() > 50
    assert np.all(d[out] > 0) and np.all(d[ins] < 0)
    assert np.all(d[edge & (lat < -1e-6)] < 0)


def test_oracle_reference_ten_bone_bake():
    v, t, w, m = ten_bones()
    r = O.bake_bone_sdf(v, t, w, m, resolution=16, margin=0.25, weight_threshold=0.02, blend_k=0.0,
                        friction=0.5, thickness=0.0, restitution=0.0, close_open_boundaries=False)
    assert len(r["active"]) == 10
    cols, rows = r["width"] // 16, r["height"] // 16
    d = _dist(r["texels"]).reshape(r["depth"], r["height"], r["width"])
    for b in [0, 1, 7, 8, 9]:
        l, rem = divmod(b, cols * rows)
        rr, cc = divmod(rem, cols)
        assert d[l * 16 + 8, rr * 16 + 8, cc * 16 + 8] < -0.02
        assert d[l * 16, rr * 16, cc * 16] > 0.0


def test_product_rejects_bad_input():
    from taremincloth_amd import taremin_cloth_core as T
    with pytest.raises(ValueError):
        T.bake_bone_sdf_gpu(BOX_V, BOX_T, ["a"], [np.ones(8)], EYE[None], resolution=0)
    with pytest.raises(ValueError):
        T.bake_bone_sdf_gpu(BOX_V, BOX_T, ["a"], [np.ones(7)], EYE[None])
    with pytest.raises(ValueError):
        T.bake_bone_sdf_gpu(BOX_V, BOX_T, ["a", "b"], [np.ones(8)], EYE[None])
    with pytest.raises(ValueError):
        T.bake_bone_sdf_gpu(BOX_V, BOX_T, ["a"], [np.ones(8)], np.zeros((1, 3, 4), np.float32))
    with pytest.raises(RuntimeError):     # no bone above the threshold: fails before any device work
        T.bake_bone_sdf_gpu(BOX_V, BOX_T, ["a"], [np.full(8, 0.01)], EYE[None])


# ------------------------------------------------------------------ GPU: the product against the oracle
def _tube_bones():
    v, t = octagon_tube()
    z = v[:, 2]
    w = np.stack([np.clip(0.5 - z * 6, 0, 1), np.clip(0.5 + z * 6, 0, 1), np.full(16, 0.25)]).astype(np.float32)
    m = np.stack([_bind(_rot(0.3, -0.2, 0.9), np.array([0.01, 0.02, -0.05])),
                  _bind(_rot(-1.1, 0.4, 0.2), np.array([0.0, -0.03, 0.05])),
                  2.0 * _bind(_rot(0.0, 0.0, 0.5), np.array([0.0, 0.0, 0.0]))]).astype(np.float32)
    return v, t, w, m


def _blob(n_lat=14, n_lon=22, seed=3):
    """A closed lumpy sphere (UV-sphere topology) with smooth weights to four bones."""
    rng = np.random.default_rng(seed)
    v = [[0, 0, 0.1]]
    for i in range(1, n_lat):
        th = np.pi * i / n_lat
        for j in range(n_lon):
            ph = 2 * np.pi * j / n_lon
            rad = 0.1 * (1 + 0.15 * rng.standard_normal())
            v.append([rad * np.sin(th) * np.cos(ph), rad * np.sin(th) * np.sin(ph), 0.1 * np.cos(th)])
    v.append([0, 0, -0.1])
    v = np.array(v, np.float32)
    t = []
    for j in range(n_lon):
        t.append([0, 1 + j, 1 + (j + 1) % n_lon])
    for i in range(n_lat - 2):
        for j in range(n_lon):
            a, b = 1 + i * n_lon + j, 1 + i * n_lon + (j + 1) % n_lon
            t += [[a, a + n_lon, b], [b, a + n_lon, b + n_lon]]
    last = len(v) - 1
    base = 1 + (n_lat - 2) * n_lon
    for j in range(n_lon):
        t.append([last, base + (j + 1) % n_lon, base + j])
    c = np.array([[0.06, 0, 0], [-0.06, 0, 0], [0, 0.06, 0], [0, 0, -0.06]])
    d2 = ((v[:, None, :] - c[None]) ** 2).sum(-1)
    w = np.exp(-d2 / 0.004)
    w = (w / w.sum(1, keepdims=True)).T.astype(np.float32)
    m = np.stack([_bind(_rot(0.1 * k, 0.2, -0.3 * k), c[k]) for k in range(4)]).astype(np.float32)
    return v, np.array(t, np.int32), w, m


CASES = {
    "single_triangle": lambda: (np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.int32),
                                np.ones((1, 3), np.float32), EYE[None], dict(resolution=16, margin=0.1, thickness=0.01)),
    "ten_bones": lambda: (*ten_bones(), dict(resolution=16, margin=0.25, blend_k=0.0, thickness=0.0,
                                              close_open_boundaries=False)),
    "tube_capped_res22": lambda: (*_tube_bones(), dict(resolution=22, margin=0.2, thickness=0.004)),
    "sheet_rim": lambda: (np.array([[-0.1, -0.1, 0], [0.1, -0.1, 0], [0.1, 0.1, 0], [-0.1, 0.1, 0]], np.float32),
                          np.array([[0, 1, 2], [0, 2, 3]], np.int32), np.ones((1, 4), np.float32), EYE[None],
                          dict(resolution=20, margin=0.5, thickness=0.0)),
    "blob_4_bones_res32": lambda: (*_blob(), dict(resolution=32, margin=0.2)),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_bone_bake_bit_exact_to_oracle(require_gpu, name):
    from taremincloth_amd import taremin_cloth_core as T
    v, t, w, m, kw = CASES[name]()
    names = [f"bone{i}" for i in range(len(w))]
    _say(f"{name}: {len(v)} verts, {len(t)} tris, {len(w)} bones, {kw}")
    g = T.bake_bone_sdf_gpu(v, t, names, list(w), m, **kw)
    o = O.bake_bone_sdf(v, t, w, m, **kw)
    assert (g["width"], g["height"], g["depth"]) == (o["width"], o["height"], o["depth"])
    tex = np.frombuffer(g["texture_bytes"], np.uint32)
    assert len(g["texture_bytes"]) == 4 * g["width"] * g["height"] * g["depth"]
    nd = int((tex != o["texels"]).sum())
    _say(f"  atlas {g['width']}x{g['height']}x{g['depth']}, kernel {g['kernel_ms']:.3f} ms, texels differing {nd}")
    assert nd == 0
    assert np.array_equal(g["bone_infos"], o["bone_infos"])
    assert g["active_bones"] == [names[i] for i in o["active"]]
    assert np.array_equal(g["bind_matrices"], o["bind_matrices"])
    if name == "ten_bones":                 # sdf_baker.rs:2325-2351
        cols, rows = g["width"] // 16, g["height"] // 16
        d = _dist(tex).reshape(g["depth"], g["height"], g["width"])
        for b in [0, 1, 7, 8, 9]:
            l, rem = divmod(b, cols * rows)
            rr, cc = divmod(rem, cols)
            assert d[l * 16 + 8, rr * 16 + 8, cc * 16 + 8] < -0.02
            assert d[l * 16, rr * 16, cc * 16] > 0.0
    if name == "single_triangle":           # sdf_baker.rs:1895-1937
        assert (g["width"], g["height"], g["depth"]) == (16, 16, 16)
        assert len(g["texture_bytes"]) == 16 * 16 * 16 * 4


@pytest.mark.gpu
def test_gpu_bone_bake_large_against_closed_form(require_gpu):
    """64 cubes on 64 bones at res 64 (a 256^3 atlas): every tile against the box distance."""
    from taremincloth_amd import taremin_cloth_core as T
    n = 64
    v = np.vstack([BOX_V + np.float32([b, 0, 0]) for b in range(n)])
    t = np.vstack([BOX_T + 8 * b for b in range(n)]).astype(np.int32)
    w = np.zeros((n, 8 * n), np.float32)
    m = np.stack([EYE] * n)
    for b in range(n):
        w[b, 8 * b:8 * b + 8] = 1.0
        m[b, 0, 3] = -float(b)
    g = T.bake_bone_sdf_gpu(v, t, [str(b) for b in range(n)], list(w), m, resolution=64, margin=0.25, thickness=0.0,
                            close_open_boundaries=False)
    _say(f"64 bones res 64: atlas {g['width']}x{g['height']}x{g['depth']}, kernel {g['kernel_ms']:.3f} ms")
    tex = np.frombuffer(g["texture_bytes"], np.uint32)
    r = {"texels": tex, "width": g["width"], "height": g["height"], "depth": g["depth"]}
    p = _voxel_centres(g["bone_infos"][0], 64)
    want = _box_sdf(p, H).astype(np.float16).astype(np.float64)
    for k in range(n):
        got = _dist(_tile(r, k, 64))
        assert np.all(np.abs(got - want) <= 2.0 ** -10 * np.abs(want) + 1e-6), k
