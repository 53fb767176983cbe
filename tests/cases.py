"""Shared test inputs: synthetic grids with a fixed edge order and the golden cases.

The reference's demo grid builds its edge list from a HashSet (order not
stable across runs, crates/cloth_gui/src/demo_scenes.rs:110-134); the grids
here fix the order: horizontal lattice edges row-major, then vertical edges,
each as (min, max).
"""
from __future__ import annotations

import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_golden")


def splitmix64(seed: int):
    state = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def uniform_jitter(n: int, amp: float, seed: int) -> np.ndarray:
    """n x 3 values uniform in [-amp, amp] from splitmix64 (float32)."""
    g = splitmix64(seed)
    u = np.array([next(g) >> 11 for _ in range(3 * n)], dtype=np.float64) / float(1 << 53)
    return ((2.0 * u - 1.0) * amp).astype(np.float32).reshape(n, 3)


def lattice_edges(nx: int, ny: int) -> np.ndarray:
    idx = np.arange(nx * ny, dtype=np.uint32).reshape(ny, nx)
    h = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], axis=1)
    v = np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], axis=1)
    return np.concatenate([h, v]).astype(np.uint32)


def grid(nx: int, ny: int, spacing: float, plane: str = "xy", origin=(0.0, 0.0, 0.0)):
    """Quad lattice split into 2 triangles per quad along (v00, v11)
    (demo_scenes.rs:112-131); returns positions, edges, faces."""
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    u = (i * spacing).astype(np.float32).ravel()
    w = (j * spacing).astype(np.float32).ravel()
    z = np.zeros_like(u)
    if plane == "xy":
        pos = np.stack([u, w, z], axis=1)
    else:  # vertical curtain in XZ
        pos = np.stack([u, z, w], axis=1)
    pos = pos + np.asarray(origin, np.float32)
    idx = np.arange(nx * ny, dtype=np.uint32).reshape(ny, nx)
    v00, v10, v01, v11 = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
    faces = np.empty((2 * v00.size, 3), np.uint32)
    faces[0::2] = np.stack([v00, v10, v11], axis=1)
    faces[1::2] = np.stack([v00, v11, v01], axis=1)
    return pos.astype(np.float32), lattice_edges(nx, ny), faces


def poked_grid(nx: int, ny: int, spacing: float, plane: str = "xz"):
    """Quad lattice with a centre vertex per quad and 4 fan triangles (config "poked to 4 tris/quad")."""
    pos, edges, _ = grid(nx, ny, spacing, plane)
    idx = np.arange(nx * ny, dtype=np.uint32).reshape(ny, nx)
    v00, v10, v01, v11 = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
    centres = ((pos[v00].astype(np.float64) + pos[v10] + pos[v01] + pos[v11]) / 4.0).astype(np.float32)
    c = np.arange(nx * ny, nx * ny + v00.size, dtype=np.uint32)
    faces = np.empty((4 * v00.size, 3), np.uint32)
    faces[0::4] = np.stack([v00, v10, c], axis=1)
    faces[1::4] = np.stack([v10, v11, c], axis=1)
    faces[2::4] = np.stack([v11, v01, c], axis=1)
    faces[3::4] = np.stack([v01, v00, c], axis=1)
    return np.concatenate([pos, centres]), edges, faces


def areal_inv_masses_np(pos, faces, density):
    """Independent numpy restatement of the area/3 mass split (float64 sums)."""
    m = np.zeros(len(pos))
    f = faces.astype(np.int64)
    ab = pos[f[:, 1]].astype(np.float64) - pos[f[:, 0]]
    ac = pos[f[:, 2]].astype(np.float64) - pos[f[:, 0]]
    share = 0.5 * np.linalg.norm(np.cross(ab, ac), axis=1) * density / 3.0
    for k in range(3):
        np.add.at(m, f[:, k], share)
    return (1.0 / np.maximum(m, 1e-9)).astype(np.float32)


def config1(seed: int = 0):
    """32x32 lattice at 1 cm in the XZ plane, poked (1,985 vertices); top two corners pinned,
    masses from areal density 0.15, jitter +-1e-4 m."""
    pos, edges, faces = poked_grid(32, 32, 0.01, "xz")
    pos = pos + uniform_jitter(len(pos), 1e-4, seed)
    inv = areal_inv_masses_np(pos, faces, 0.15)
    inv[32 * 31] = 0.0  # top row is j = 31 (z = 0.31)
    inv[32 * 32 - 1] = 0.0
    return pos.astype(np.float32), edges, faces, inv


def drape_grid(n: int, spacing: float, height: float, seed: int = 0, jitter: float = 1e-3):
    """Horizontal n x n cloth centred on the z axis at `height` with seeded jitter."""
    pos, edges, faces = grid(n, n, spacing, "xy")
    half = (n - 1) * spacing * 0.5
    pos = pos + np.array([-half, -half, height], np.float32)
    pos = pos + uniform_jitter(len(pos), jitter, seed)
    inv = areal_inv_masses_np(pos, faces, 0.15)
    return pos.astype(np.float32), edges, faces, inv


GOLDEN_CASES = {
    # parameters from tests/regression/test_single_step_exact.py and test_golden_regression.py
    "case1": dict(file="case1_cloth_body.npz", step="step_case1_f1.npz", stiffness=1500.0, bending_stiffness=15.0,
                  density=None, colliders=[]),
    "case2": dict(file="case2_plane_sphere.npz", step="step_case2_f1.npz", stiffness=1200.0, bending_stiffness=5.0,
                  density=None, colliders=[("sphere", [0.0, 0.0, 0.16], 0.14, 0.4, 0.0)]),
    "case3": dict(file="case3_multi_collider.npz", step="step_case3_f1.npz", stiffness=1000.0,
                  bending_stiffness=10.0, density=0.15,
                  colliders=[("capsule", [-0.3, -0.3, 0.4], [0.3, -0.3, 0.5], 0.08, 0.4, 0.0),
                             ("capsule", [-0.3, 0.3, 0.5], [0.3, 0.3, 0.4], 0.08, 0.4, 0.0),
                             ("sphere", [0.0, 0.0, 0.3], 0.12, 0.5, 0.0),
                             ("plane", [0.0, 0.0, 0.0], [0.0, 0.0, 1.0], 0.6, 0.0)]),
    "case4": dict(file="case4_self_collision_drape.npz", step=None, stiffness=800.0, bending_stiffness=5.0,
                  density=0.10, colliders=[("sphere", [0.0, 0.0, 0.2], 0.20, 0.6, 0.0)],
                  sc_options=(1.0, 0.5, True, True, 256)),
}


def load_npz(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def golden_inputs(case: str):
    c = GOLDEN_CASES[case]
    d = load_npz(c["file"])
    pos = d["initial_pos"]
    if "inv_masses" in d:
        inv = d["inv_masses"]
    else:
        inv = areal_inv_masses_np(pos, d["faces"], c["density"])
    return d, pos, inv


def add_colliders(sim, colliders):
    for col in colliders:
        kind = col[0]
        if kind == "sphere":
            sim.add_sphere_collider(col[1], radius=col[2], friction=col[3], restitution=col[4])
        elif kind == "capsule":
            sim.add_capsule_collider(col[1], col[2], radius=col[3], friction=col[4], restitution=col[5])
        else:
            sim.add_plane_collider(col[1], col[2], friction=col[3], restitution=col[4])


def make_golden_sim(factory, case: str, positions=None):
    """Construct `factory` (ClothSimulator or OracleSim) exactly as the reference tests do."""
    c = GOLDEN_CASES[case]
    d, pos, inv = golden_inputs(case)
    if positions is not None:
        pos = positions
    sim = factory(pos, d["edges"], d["faces"], inv, thickness=0.005, stiffness=c["stiffness"],
                  bending_stiffness=c["bending_stiffness"])
    sim.set_gravity(0.0, 0.0, -9.81)
    if "sc_options" in c:
        sim.set_enable_self_collision(True)
        sim.set_self_collision_options(*c["sc_options"])
    add_colliders(sim, c["colliders"])
    return sim, d, pos


def folded_cloth(n: int = 12, spacing: float = 0.02, gap: float = 0.004):
    """One cloth folded over itself (a single island): lower half flat, upper half
    returning above it at `gap`, so vertex-vertex, vertex-triangle and edge-edge
    contacts all occur in the first substep."""
    pos, edges, faces = grid(n, 2 * n, spacing, "xy")
    pos = pos.astype(np.float64)
    half = (n - 1) * spacing
    y = pos[:, 1]
    upper = y > half
    # fold around y = half: mirror the upper part back over the lower part, lifted by `gap`
    pos[upper, 1] = 2 * half - y[upper] + spacing * 0.37
    pos[upper, 2] = gap
    pos[:, 0] += (np.arange(len(pos)) % 7) * 1e-4  # break exact symmetry
    inv = areal_inv_masses_np(pos.astype(np.float32), faces, 0.15)
    return pos.astype(np.float32), edges, faces, inv


def icosphere(subdiv: int, radius: float = 0.3, noise_amp: float = 0.05, seed: int = 0, center=(0.0, 0.0, 0.0)):
    """Icosphere (subdivision level `subdiv`: 10*4^s + 2 vertices, 20*4^s triangles) with the
    radius perturbed by seeded 3-octave value noise of relative amplitude `noise_amp`."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    V = np.array(verts)
    rng = np.random.default_rng(seed)
    lat = [rng.uniform(-1, 1, (8 * 2 ** o,) * 3) for o in range(3)]

    def value_noise(p):
        out = np.zeros(len(p))
        for o, g in enumerate(lat):
            n = g.shape[0]
            q = (p * 0.5 + 0.5) * (n - 1)
            i0 = np.clip(np.floor(q).astype(int), 0, n - 2)
            fr = q - i0
            s = fr * fr * (3 - 2 * fr)
            acc = 0.0
            for dx in (0, 1):
                for dy in (0, 1):
                    for dz in (0, 1):
                        w = ((s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1]) *
                             (s[:, 2] if dz else 1 - s[:, 2]))
                        acc = acc + w * g[i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz]
            out += acc / (2 ** o)
        return out / 1.75
    r = radius * (1.0 + noise_amp * value_noise(V))
    pos = (V * r[:, None] + np.asarray(center)).astype(np.float32)
    return pos, np.array(f, np.uint32)


def batch_instances(count: int, side_lo: int, side_hi: int, seed: int = 0, spacing: float = 0.01,
                    jitter: float = 1e-4):
    """`count` vertical curtains (XZ lattices, 2 triangles per quad) with sides drawn uniformly
    from [side_lo, side_hi] by splitmix64(seed), laid side by side along x; the top two corners
    of each are pinned, masses from areal density 0.15, positions jittered by +-`jitter`."""
    g = splitmix64(seed)
    out = []
    x0 = 0.0
    for k in range(count):
        side = side_lo + int(next(g) % (side_hi - side_lo + 1))
        pos, edges, faces = grid(side, side, spacing, "xz", origin=(x0, 0.0, 0.0))
        pos = pos + uniform_jitter(len(pos), jitter, seed * 1000 + k)
        inv = areal_inv_masses_np(pos, faces, 0.15)
        inv[side * (side - 1)] = 0.0
        inv[side * side - 1] = 0.0
        out.append({"positions": pos.astype(np.float32), "edges": edges, "faces": faces, "inv_masses": inv})
        x0 += side * spacing + 0.05
    return out
