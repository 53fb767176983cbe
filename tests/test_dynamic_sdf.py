"""Dynamic bone SDF: a skinned collider mesh whose bone-SDF atlas is re-baked on the device at the
start of a frame (setup_dynamic_bone_sdf, lib.rs:2362-2588; dispatch_dynamic_bone_sdf_update,
simulation/mod.rs:2320-2457; skinning.wgsl, prep_bone_triangles.wgsl, bake_sdf.wgsl), and the
bounding spheres update_bone_transforms writes into the collider infos (simulation/mod.rs:1886-1925).

CPU tests pin the oracle (oracle/xpbd_oracle.c, oracle/sdf_oracle.c):
  * at the rest pose (world = bind^-1) the dynamic atlas equals the static bone bake of the same mesh,
    weights, boxes and tiles bit for bit: skinning is the identity there, the bone-space corners are
    the static bake's, and the open-edge masks of the triangle sources are the mesh's;
  * a rigid translation of every bone moves the skinned mesh by exactly that translation (dyadic
    coordinates) and leaves the atlas unchanged: the SDF is stored in bone space;
  * a vertex whose weights sum to 1e-6 or less keeps its rest position;
  * the bounding spheres: identity and uniformly scaled world matrices against the closed form;
  * the update policy: the first update bakes every bone, interval k updates frames 1, k+1, 2k+1...,
    a dirty list re-bakes only its bones' tiles and an empty one none.
GPU tests: the product's skinned positions, bone-space triangles, atlas texels and infos bit-equal to
the fp32 oracle over a posed, blended sequence with dirty lists and an interval; the rest-pose atlas
equal to the product's own static bake; a cloth falling on the posed dynamic collider bit-equal to
the oracle.
"""
import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O

EYE = np.eye(4, dtype=np.float32)


def _say(msg):
    print(f"[dyn-sdf] {msg}", flush=True)
    sys.stdout.flush()


def square_tube(levels=(-0.125, 0.0, 0.125), a=0.0625):
    """An open tube of square rings (dyadic coordinates) along z: 4 quads per segment, ends open."""
    ring = np.array([[a, a], [-a, a], [-a, -a], [a, -a]], np.float32)
    v = np.vstack([np.c_[ring, np.full(4, z, np.float32)] for z in levels]).astype(np.float32)
    t = []
    for k in range(len(levels) - 1):
        for i in range(4):
            j = (i + 1) % 4
            a0, a1, b0, b1 = 4 * k + i, 4 * k + j, 4 * (k + 1) + i, 4 * (k + 1) + j
            t += [[a0, a1, b1], [a0, b1, b0]]
    return v, np.array(t, np.int32)


def tube_scene():
    """Three rings, two bones: the lower ring on bone 0, the upper on bone 1, the middle ring
    half-and-half (weights 0, 0.5, 1 only, so blends of equal positions stay exact)."""
    v, t = square_tube()
    w = np.zeros((2, len(v)), np.float32)
    w[0, 0:4] = 1.0
    w[0, 4:8] = 0.5
    w[1, 4:8] = 0.5
    w[1, 8:12] = 1.0
    binds = np.stack([EYE.copy(), EYE.copy()])
    binds[0, :3, 3] = [0.25, -0.125, 0.5]      # world -> bone: pure dyadic translations
    binds[1, :3, 3] = [-0.375, 0.0625, -0.25]
    return v, t, w, binds


def dynamic_inputs(v, t, w, binds, static):
    """setup_dynamic_bone_sdf arguments equivalent to a static bake: its infos, the atlas tiles it
    used, every triangle with a weighted corner listed under each bone, four influences per vertex."""
    nb = len(w)
    res = int(static["bone_infos"][0][7])
    src, tw = [], []
    for b in range(nb):
        for tri in t:
            cw = w[b, tri]
            if (cw > 0).any():
                src.append([tri[0], tri[1], tri[2], b])
                tw.append(cw)
    bi = np.zeros((len(v), 4), np.uint32)
    bw = np.zeros((len(v), 4), np.float32)
    for i in range(len(v)):
        k = 0
        for b in range(nb):
            if w[b, i] > 0:
                bi[i, k], bw[i, k] = b, w[b, i]
                k += 1
    return dict(width=static["width"], height=static["height"], depth=static["depth"], res=res,
                bone_infos=static["bone_infos"], rest_verts=v, bone_indices=bi, bone_weights=bw,
                tri_sources=np.array(src, np.uint32), tri_weights=np.array(tw, np.float32),
                bone_bind_inv_matrices=binds)


def as_given(m):
    """A row-major 4x4 as update_bone_transforms takes it (WGSL reads each given row as a column)."""
    return np.ascontiguousarray(np.asarray(m, np.float64).T.astype(np.float32))


def rest_pose(binds):
    world = np.stack([as_given(np.linalg.inv(b.astype(np.float64))) for b in binds])
    inv = np.stack([as_given(b) for b in binds])
    return world, inv


def translated_pose(binds, d):
    world, inv = [], []
    for b in binds:
        tr = np.eye(4)
        tr[:3, 3] = d
        wm = tr @ np.linalg.inv(b.astype(np.float64))
        world.append(as_given(wm))
        inv.append(as_given(np.linalg.inv(wm)))
    return np.stack(world), np.stack(inv)


def _rot(ax, ang):
    ax = np.asarray(ax, np.float64) / np.linalg.norm(ax)
    k = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    return np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * k @ k


def bent_pose(binds, ang, shift):
    """Bone 0 at rest; bone 1 rotated by `ang` about the x axis through the middle ring, then shifted."""
    world, inv = [], []
    for b, bm in enumerate(binds):
        wm = np.linalg.inv(bm.astype(np.float64))
        if b == 1:
            r = np.eye(4)
            r[:3, :3] = _rot((1, 0, 0), ang)
            r[:3, 3] = shift
            wm = r @ wm
        world.append(as_given(wm))
        inv.append(as_given(np.linalg.inv(wm)))
    return np.stack(world), np.stack(inv)


def host_sim(cls):
    """A small cloth far above the collider (the dynamic SDF needs a simulator to live in)."""
    pos, e, f = cases.grid(6, 6, 0.04, origin=(-0.1, -0.1, 2.0))
    return cls(pos, e, f, thickness=0.004)


def oracle_atlas(sim, inp):
    return sim.sdf_arrays(inp["width"] * inp["height"] * inp["depth"], len(inp["bone_infos"]))


def _static(v, t, w, binds):
    return O.bake_bone_sdf(v, t, w, binds, resolution=16, margin=0.25, thickness=0.004, close_open_boundaries=False)


# ------------------------------------------------------------------ CPU: the oracle pinned
def test_oracle_rest_pose_equals_static_bake():
    v, t, w, binds = tube_scene()
    st = _static(v, t, w, binds)
    inp = dynamic_inputs(v, t, w, binds, st)
    o = host_sim(O.OracleSim)
    o.setup_dynamic_bone_sdf(**inp)
    o.update_bone_transforms(*rest_pose(binds))
    o.step(1.0 / 60.0, 2)
    tex, info = oracle_atlas(o, inp)
    sk, _ = o.dynamic_arrays(len(v), len(inp["tri_sources"]))
    assert np.array_equal(sk, v)                       # skinning is the identity at rest
    assert np.array_equal(tex, st["texels"])           # and the atlas is the static bake's
    assert ((tex & 0xFFFF) != 0).all()


def test_oracle_rigid_translation_moves_the_mesh_and_keeps_the_atlas():
    v, t, w, binds = tube_scene()
    st = _static(v, t, w, binds)
    inp = dynamic_inputs(v, t, w, binds, st)
    o = host_sim(O.OracleSim)
    o.setup_dynamic_bone_sdf(**inp)
    d = np.array([0.5, -0.25, 0.125])
    o.update_bone_transforms(*translated_pose(binds, d))
    o.step(1.0 / 60.0, 2)
    tex, info = oracle_atlas(o, inp)
    sk, tris = o.dynamic_arrays(len(v), len(inp["tri_sources"]))
    assert np.array_equal(sk, v + d.astype(np.float32))
    assert np.array_equal(tex, st["texels"])
    # the bounding-sphere centre followed the translation: box centre (local) -> world
    for b in range(2):
        lo, hi = info[b, 0:3].astype(np.float64), info[b, 4:7].astype(np.float64)
        c_world = (lo + hi) / 2 - binds[b, :3, 3] + d
        assert np.allclose([info[b, 3], info[b, 7], info[b, 11]], c_world, atol=1e-6)
        assert abs(info[b, 15] - (0.5 * np.linalg.norm(hi - lo) + 0.1)) < 1e-6


def test_oracle_skinning_weight_floor_keeps_the_rest_position():
    v, t, w, binds = tube_scene()
    st = _static(v, t, w, binds)
    inp = dynamic_inputs(v, t, w, binds, st)
    inp["bone_weights"] = inp["bone_weights"].copy()
    inp["bone_weights"][0] = [1e-6, 0, 0, 0]           # sums to 1e-6: rest position kept
    inp["bone_weights"][1] = [2e-6, 0, 0, 0]           # above the floor: skinned with weight 2e-6
    o = host_sim(O.OracleSim)
    o.setup_dynamic_bone_sdf(**inp)
    o.update_bone_transforms(*translated_pose(binds, np.array([1.0, 0, 0])))
    o.step(1.0 / 60.0, 1)
    sk, _ = o.dynamic_arrays(len(v), len(inp["tri_sources"]))
    assert np.array_equal(sk[0], v[0])
    assert np.allclose(sk[1], np.float32(2e-6) * (v[1] + np.float32([1, 0, 0])), rtol=1e-6)


def test_oracle_bounding_sphere_closed_form():
    """update_bone_transforms on static SDF colliders: centre = the box centre through the world
    matrix, radius = half the box diagonal x the largest column length + 0.1."""
    o = host_sim(O.OracleSim)
    info = np.zeros((2, 20), np.float32)
    info[:, 0:3] = [-0.1, -0.2, -0.3]
    info[:, 4:7] = [0.3, 0.2, 0.1]
    info[:, 8:11] = 1.0
    o.set_bone_sdf_colliders(2, 2, 2, np.zeros(8, np.uint32).tobytes(), info)
    sc = np.diag([2.0, 2.0, 2.0, 1.0])
    sc[:3, 3] = [1.0, 2.0, 3.0]
    world = np.stack([EYE, as_given(sc)])
    o.update_bone_transforms(world, world)
    _, got = o.sdf_arrays(8, 2)
    half_diag = 0.5 * np.linalg.norm([0.4, 0.4, 0.4])
    assert np.allclose(got[0, [3, 7, 11, 15]], [0.1, 0.0, -0.1, half_diag + 0.1], atol=1e-6)
    assert np.allclose(got[1, [3, 7, 11, 15]], [1.2, 2.0, 2.8, 2 * half_diag + 0.1], atol=1e-6)


def _tile(tex, inp, b):
    w, h, d, res = inp["width"], inp["height"], inp["depth"], inp["res"]
    cols, rows = w // res, h // res
    l, rem = divmod(b, cols * rows)
    r, c = divmod(rem, cols)
    return tex.reshape(d, h, w)[l * res:(l + 1) * res, r * res:(r + 1) * res, c * res:(c + 1) * res]


def test_oracle_update_policy_interval_and_dirty_bones():
    v, t, w, binds = tube_scene()
    inp = dynamic_inputs(v, t, w, binds, _static(v, t, w, binds))
    o = host_sim(O.OracleSim)
    inp2 = dict(inp, update_interval=2)
    o.setup_dynamic_bone_sdf(**inp2)
    o.update_bone_transforms(*bent_pose(binds, 0.0, (0, 0, 0)), dirty_bone_indices=[1])
    o.step(1.0 / 60.0, 1)                    # frame 1: first update, every bone baked despite the list
    a1, _ = oracle_atlas(o, inp)
    assert all(((_tile(a1, inp, b) & 0xFFFF) != 0).all() for b in range(2))
    o.update_bone_transforms(*bent_pose(binds, 0.6, (0, 0.03, 0)), dirty_bone_indices=[1])
    o.step(1.0 / 60.0, 1)                    # frame 2: skipped by the interval
    a2, _ = oracle_atlas(o, inp)
    assert np.array_equal(a1, a2)
    o.step(1.0 / 60.0, 1)                    # frame 3: bone 1's tile only
    a3, _ = oracle_atlas(o, inp)
    assert np.array_equal(_tile(a3, inp, 0), _tile(a1, inp, 0))
    assert not np.array_equal(_tile(a3, inp, 1), _tile(a1, inp, 1))
    o.set_dynamic_bone_sdf_dirty_bones([])   # empty list: nothing re-baked
    o.update_bone_transforms(*bent_pose(binds, -0.4, (0, 0, 0)), dirty_bone_indices=[])
    o.step(1.0 / 60.0, 1)
    o.step(1.0 / 60.0, 1)
    a5, _ = oracle_atlas(o, inp)
    assert np.array_equal(a5, a3)
    o.set_dynamic_bone_sdf_enabled(False)    # off: nothing re-baked even with every bone dirty
    o.update_bone_transforms(*bent_pose(binds, 0.9, (0, 0, 0)))
    o.step(1.0 / 60.0, 1)
    o.step(1.0 / 60.0, 1)
    a7, _ = oracle_atlas(o, inp)
    assert np.array_equal(a7, a3)


def test_oracle_rejects_out_of_range_inputs():
    v, t, w, binds = tube_scene()
    inp = dynamic_inputs(v, t, w, binds, _static(v, t, w, binds))
    o = host_sim(O.OracleSim)
    bad = dict(inp, tri_sources=inp["tri_sources"].copy())
    bad["tri_sources"][0, 3] = 7
    with pytest.raises(ValueError):
        o.setup_dynamic_bone_sdf(**bad)
    with pytest.raises(ValueError):
        o.setup_dynamic_bone_sdf(**dict(inp, res=64))          # tiles outside a 16-texel atlas


# ------------------------------------------------------------------ GPU: the product against the oracle
def _pair():
    from taremincloth_amd import taremin_cloth_core as T
    return host_sim(T.ClothSimulator), host_sim(O.OracleSim)


def _g_arrays(g, inp):
    a = g.sdf_collider_arrays()
    return a["texels"], a["infos"], a["skinned"], a["bake_triangles"]


@pytest.mark.gpu
def test_gpu_rest_pose_equals_the_static_bake(require_gpu):
    from taremincloth_amd import taremin_cloth_core as T
    v, t, w, binds = tube_scene()
    st = T.bake_bone_sdf_gpu(v, t, ["a", "b"], list(w), binds, resolution=16, margin=0.25, thickness=0.004,
                             close_open_boundaries=False)
    st = dict(st, texels=np.frombuffer(st["texture_bytes"], np.uint32))
    inp = dynamic_inputs(v, t, w, binds, st)
    g, _ = _pair()
    g.setup_dynamic_bone_sdf(**inp)
    g.update_bone_transforms(*rest_pose(binds))
    g.step(1.0 / 60.0, 2)
    tex, _, sk, _ = _g_arrays(g, inp)
    _say(f"rest pose: atlas {inp['width']}x{inp['height']}x{inp['depth']}, stats {g.dynamic_bone_sdf_stats()}")
    assert np.array_equal(sk, v)
    assert np.array_equal(tex, st["texels"])


@pytest.mark.gpu
def test_gpu_posed_sequence_bit_exact_to_oracle(require_gpu):
    v, t, w, binds = tube_scene()
    inp = dict(dynamic_inputs(v, t, w, binds, _static(v, t, w, binds)), update_interval=2)
    g, o = _pair()
    for s in (g, o):
        s.setup_dynamic_bone_sdf(**inp)
    plan = [(0.0, None), (0.3, [1]), (0.5, [1]), (0.7, None), (-0.2, []), (0.4, [0, 1, 1]), (0.9, [1, 5])]
    for fi, (ang, dirty) in enumerate(plan):
        pose = bent_pose(binds, ang, (0.0, 0.01 * fi, -0.005 * fi))
        for s in (g, o):
            s.update_bone_transforms(*pose, dirty_bone_indices=dirty)
            s.step(1.0 / 60.0, 2)
        tex, info, sk, tris = _g_arrays(g, inp)
        otex, oinfo = oracle_atlas(o, inp)
        osk, otris = o.dynamic_arrays(len(v), len(inp["tri_sources"]))
        nd = int((tex != otex).sum())
        _say(f"frame {fi + 1}: angle {ang}, dirty {dirty}: texels differing {nd}, stats {g.dynamic_bone_sdf_stats()}")
        assert np.array_equal(sk, osk)
        assert np.array_equal(tris, otris)
        assert nd == 0
        assert np.array_equal(info, oinfo)


@pytest.mark.gpu
def test_gpu_cloth_on_the_dynamic_collider_bit_exact(require_gpu):
    """A sheet dropped onto the posed tube (collider thickness 4 mm): positions bit-equal to the oracle
    while the tube bends between frames."""
    from taremincloth_amd import taremin_cloth_core as T
    v, t, w, binds = tube_scene()
    v = v * np.float32(2.0)
    binds = binds.copy()
    st = _static(v, t, w, binds)
    inp = dynamic_inputs(v, t, w, binds, st)
    pos, e, f = cases.grid(16, 16, 0.025, origin=(-0.1875, -0.1875, 0.2))
    sims = [T.ClothSimulator(pos, e, f, thickness=0.004), O.OracleSim(pos, e, f, thickness=0.004)]
    for s in sims:
        s.setup_dynamic_bone_sdf(**inp)
        s.set_gravity(0.0, 0.0, -9.81)
    # the tube lies along y under the sheet: rotate both bones' world placement about x by 90 degrees
    out = np.empty(3 * len(pos), np.float32)
    for fi in range(14):
        world, inv = [], []
        for b, bm in enumerate(binds):
            wm = np.linalg.inv(bm.astype(np.float64))
            r = np.eye(4)
            r[:3, :3] = _rot((1, 0, 0), np.pi / 2 - (0.05 * fi if b == 1 else 0.0))  # bone 1's half lifts
            wm = r @ wm
            world.append(as_given(wm))
            inv.append(as_given(np.linalg.inv(wm)))
        for s in sims:
            s.update_bone_transforms(np.stack(world), np.stack(inv))
            s.step(1.0 / 60.0, 10, 4)
        sims[0].get_positions(out)
        diff = float(np.abs(out - sims[1].get_positions()).max())
        _say(f"frame {fi}: max |hip - oracle| {diff:.3e}")
        assert diff == 0.0
    # the sheet was caught by the tube: bone 0's half keeps its top face at z = 0.125 in the world
    # (world y in [0, 0.25]), where a free fall would have reached z = -0.07
    p = out.reshape(-1, 3)
    over = (np.abs(p[:, 0]) < 0.1) & (p[:, 1] > 0.02) & (p[:, 1] < 0.2)
    _say(f"lowest vertex over bone 0's half of the tube: z = {p[over, 2].min():.4f}")
    assert over.any() and p[over, 2].min() > 0.1
