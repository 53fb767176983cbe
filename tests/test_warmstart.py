"""Learned warm start: the FP64 oracle pinned by finite differences and closed forms (CPU),
then the HIP forward / backward / Adam / features / gate checked against it (GPU)."""
import numpy as np
import pytest

from oracle import warmstart_oracle as W

import cases


def ring_graph(n, extra=7, seed=0):
    """Symmetric neighbour lists: a ring plus seeded chords (no self loops)."""
    rng = np.random.default_rng(seed)
    pairs = {(i, (i + 1) % n) for i in range(n)}
    for _ in range(extra):
        a, b = rng.integers(0, n, 2)
        if a != b:
            pairs.add((int(a), int(b)))
    a = np.array([p[0] for p in pairs])
    b = np.array([p[1] for p in pairs])
    off, nbr, _ = W.csr_pairs(n, a, b)
    return off, nbr


def random_case(rows, rounds, seed, free_frac=0.8):
    rng = np.random.default_rng(seed)
    feat = rng.normal(0, 1, (rows, W.F)).astype(np.float32)
    feat[:, 16:] = 0.0
    t = np.zeros((rows, 4), np.float32)
    t[:, :3] = rng.normal(0, 1, (rows, 3))
    t[:, 3] = rng.random(rows) < free_frac
    off, nbr = ring_graph(rows, extra=rows // 2, seed=seed)
    return W.init_params(rounds, seed), feat, off, nbr, t


# ---------------------------------------------------------------- CPU: pin the oracle
def test_oracle_backward_matches_finite_differences():
    rounds = 2
    p, feat, off, nbr, t = random_case(12, rounds, 3)
    loss, g, _ = W.backward(p, rounds, feat, off, nbr, t, bf16=False)
    rng = np.random.default_rng(5)
    lay, _ = W.layout(rounds)
    picks = []
    for o, K, N in lay:  # a few weights and a bias of every layer
        picks += list(o + rng.integers(0, K * N, 3)) + [o + K * N + int(rng.integers(0, N))]
    pd = p.astype(np.float64)
    for i in picks:
        h = 1e-6
        a, b = pd.copy(), pd.copy()
        a[i] += h
        b[i] -= h
        la = W.loss_grad(W.forward(a, rounds, feat, off, nbr, bf16=False)[0], t)[0]
        lb = W.loss_grad(W.forward(b, rounds, feat, off, nbr, bf16=False)[0], t)[0]
        fd = (la - lb) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_oracle_graph_transpose_is_the_adjoint():
    off, nbr = ring_graph(9, extra=6, seed=1)
    n = 9
    A = np.zeros((n, n))
    deg = np.diff(off)
    for i in range(n):
        for k in range(off[i], off[i + 1]):
            A[i, nbr[k]] += 1.0 / deg[i]
    x = np.random.default_rng(0).normal(size=(n, 4))
    assert np.allclose(W.mean_agg(x, off, nbr), A @ x)
    assert np.allclose(W.mean_agg_transpose(x, off, nbr), A.T @ x)


def test_oracle_features_closed_form():
    """Two free vertices 0.012 m apart with rest 0.010 m: the stretch term pulls each toward the other
    by (len - L) / s; a sphere 0.05 m below gives the signed gap and the outward normal."""
    s, L = 1e-4, 0.010
    ppred = np.array([[0.0, 0.0, 0.1, 1.0], [0.012, 0.0, 0.1, 1.0]])
    mesh = {"dist_v0": [0], "dist_v1": [1], "dist_rest": [L], "bend_v2": [], "bend_v3": [], "bend_rest": []}
    cols = [(0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.05)]
    f = W.features(ppred, np.zeros((2, 3)), np.full(2, 0.005), mesh, cols, s, 0.01)
    pull = (0.012 - L) / s
    assert np.allclose(f[0, 3:6], [W.soft(pull), 0, 0]) and np.allclose(f[1, 3:6], [-W.soft(pull), 0, 0])
    gap0 = (np.hypot(0.0, 0.1) - 0.05 - 0.005) / 0.01
    assert np.isclose(f[0, 12], W.soft(min(gap0, 4.0)))
    assert f[0, 9] == 1 and f[0, 10] == 0.125 and np.allclose(f[:, 16:], 0)
    assert np.allclose(f[0, 13:16], 0)  # gap of 4.5 lengths: beyond the range, no normal
    near = W.features(ppred - [0, 0, 0.04, 0], np.zeros((2, 3)), np.full(2, 0.005), mesh, cols, s, 0.01)
    assert np.allclose(near[0, 13:16], [0, 0, 1])


def _torch_forward(params, rounds, feat, off, nbr):
    """The network restated with torch on the CPU in float64: torch.nn.functional.linear / relu
    and an index_add_ neighbour mean, from the documented layer layout (W stored K x N)."""
    import torch
    import torch.nn.functional as Fn
    p = torch.as_tensor(np.asarray(params, np.float64))
    layers, o = [], 0
    dims = [(W.F, W.H), (W.H, W.H)] + [(2 * W.H, W.H), (W.H, W.H)] * rounds + [(W.H, W.H), (W.H, W.O)]
    for K, N in dims:
        layers.append((p[o:o + K * N].reshape(K, N).T.contiguous(), p[o + K * N:o + K * N + N]))
        o += K * N + N
    x = torch.as_tensor(np.asarray(feat, np.float64))
    offt = torch.as_tensor(np.asarray(off, np.int64))
    deg = offt[1:] - offt[:-1]
    rows = torch.repeat_interleave(torch.arange(len(deg)), deg)
    cols = torch.as_tensor(np.asarray(nbr, np.int64))

    def mlp(v, a, b):
        return Fn.linear(Fn.relu(Fn.linear(v, *layers[a])), *layers[b])

    h = mlp(x, 0, 1)
    for r in range(rounds):
        s_ = torch.zeros_like(h).index_add_(0, rows, h[cols])
        mean = s_ / deg.clamp(min=1).unsqueeze(1).to(h.dtype)
        h = h + mlp(torch.cat([h, mean], dim=1), 2 + 2 * r, 3 + 2 * r)
    return mlp(h, 2 + 2 * rounds, 3 + 2 * rounds).numpy()


@pytest.mark.parametrize("rounds", [1, 3])
def test_oracle_forward_matches_a_torch_restatement(rounds):
    p, feat, off, nbr, _ = random_case(40, rounds, 11 + rounds)
    # one isolated vertex: its neighbour mean is zero
    deg = np.diff(off).astype(np.int64)
    keep = np.repeat(np.arange(40), deg) != 7
    keep &= nbr != 7
    rows = np.repeat(np.arange(40), deg)[keep]
    nbr2 = nbr[keep]
    off2 = np.zeros(41, np.uint32)
    off2[1:] = np.cumsum(np.bincount(rows, minlength=40))
    y, _ = W.forward(p, rounds, feat, off2, nbr2, bf16=False)
    yt = _torch_forward(p, rounds, feat, off2, nbr2)
    assert np.abs(y - yt).max() <= 1e-10 * max(1.0, np.abs(yt).max())


def test_oracle_forward_is_permutation_equivariant():
    """Relabelling the vertices permutes the output rows and nothing else."""
    rounds = 2
    p, feat, off, nbr, _ = random_case(30, rounds, 4)
    perm = np.random.default_rng(9).permutation(30)
    inv = np.argsort(perm)
    deg = np.diff(off).astype(np.int64)
    rows = np.repeat(np.arange(30), deg)
    a, b = inv[rows], inv[nbr]          # new labels of each directed pair
    order = np.lexsort((np.arange(len(a)), a))
    off_p = np.zeros(31, np.uint32)
    off_p[1:] = np.cumsum(np.bincount(a, minlength=30))
    y, _ = W.forward(p, rounds, feat, off, nbr, bf16=False)
    yp, _ = W.forward(p, rounds, feat[perm], off_p, b[order].astype(np.uint32), bf16=False)
    assert np.allclose(yp, y[perm], rtol=1e-12, atol=1e-12)


def _feat(ppred, mesh, cols, dprev=None, s=1e-4, lscale=0.01, th=0.005, beta=0.5):
    n = len(ppred)
    return W.features(np.asarray(ppred, np.float64), np.zeros((n, 3)) if dprev is None else np.asarray(dprev),
                      np.full(n, th), mesh, cols, s, lscale, beta)


_NO_PAIRS = {"dist_v0": [], "dist_v1": [], "dist_rest": [], "bend_v2": [], "bend_v3": [], "bend_rest": []}


def test_oracle_features_capsule_and_plane_closed_forms():
    # capsule along x from -0.3 to 0.3, radius 0.1: above its middle, and beyond its end cap
    cap = [(1, (-0.3, 0.0, 0.0), (0.3, 0.0, 0.0), 0.1)]
    f = _feat([[0.1, 0.0, 0.13, 1.0], [0.35, 0.0, 0.05, 1.0]], _NO_PAIRS, cap)
    r0 = (0.13 - 0.1 - 0.005) / 0.01
    assert np.isclose(f[0, 12], W.soft(r0)) and np.allclose(f[0, 13:16], [0, 0, 1])
    r1 = (np.hypot(0.05, 0.05) - 0.1 - 0.005) / 0.01
    assert np.isclose(f[1, 12], W.soft(r1)) and np.allclose(f[1, 13:16], [np.sqrt(0.5), 0, np.sqrt(0.5)])
    # plane through (0, 0, 0.1) with an unnormalised normal (0, 0, 2): signed height, also below it
    pl = [(2, (0.0, 0.0, 0.1), (0.0, 0.0, 2.0), 0.0)]
    f = _feat([[0.3, -0.2, 0.12, 1.0], [0.0, 0.0, 0.09, 1.0]], _NO_PAIRS, pl)
    assert np.isclose(f[0, 12], W.soft((0.02 - 0.005) / 0.01)) and np.allclose(f[0, 13:16], [0, 0, 1])
    assert np.isclose(f[1, 12], W.soft((-0.01 - 0.005) / 0.01)) and np.allclose(f[1, 13:16], [0, 0, 1])
    # nearest of two colliders wins: the plane is nearer than the capsule here
    f = _feat([[0.1, 0.0, 0.13, 1.0]], _NO_PAIRS, cap + [(2, (0.0, 0.0, 0.12), (0.0, 0.0, 1.0), 0.0)])
    assert np.isclose(f[0, 12], W.soft((0.01 - 0.005) / 0.01))
    # no collider: the range cap
    f = _feat([[0.0, 0.0, 0.0, 1.0]], _NO_PAIRS, [])
    assert np.isclose(f[0, 12], W.soft(4.0)) and np.allclose(f[0, 13:16], 0)


def test_oracle_features_bending_pull_history_and_degrees():
    s, Lb = 1e-4, 0.02
    mesh = {"dist_v0": [0, 1], "dist_v1": [1, 2], "dist_rest": [0.0125, 0.0125],
            "bend_v2": [0], "bend_v3": [2], "bend_rest": [Lb]}
    pp = [[0.0, 0.0, 0.0, 1.0], [0.0125, 0.0, 0.0, 1.0], [0.025, 0.0, 0.0, 0.0]]
    f = _feat(pp, mesh, [], s=s)
    pull = (0.025 - Lb) / s          # bending pair stretched by 5 mm, toward each other
    assert np.allclose(f[0, 6:9], [W.soft(pull), 0, 0]) and np.allclose(f[2, 6:9], [-W.soft(pull), 0, 0])
    assert np.allclose(f[1, 6:9], 0) and np.allclose(f[:, 3:6], 0)        # distance pairs at rest
    assert np.allclose(f[:, 10], [0.125, 0.25, 0.125]) and np.allclose(f[:, 11], [0.125, 0.0, 0.125])
    assert np.allclose(f[:, 9], [1, 1, 0])                                 # pinned vertex 2
    # the strains are taken at prediction + beta * previous correction; the correction itself is a feature
    d = np.array([[0.0, 0.0, 0.0], [0.002, 0.0, 0.0], [0.0, 0.0, 0.0]])
    for beta in (0.5, 1.0):
        f = _feat(pp, mesh, [], dprev=d, s=s, beta=beta)
        assert np.allclose(f[1, 0:3], [W.soft(0.002 / s), 0, 0])
        # vertex 1 moved by beta * 2 mm: pair (0,1) stretched, pair (1,2) compressed by the same amount
        assert np.allclose(f[0, 3:6], [W.soft(beta * 0.002 / s), 0, 0])
        assert np.allclose(f[1, 3:6], W.soft(np.array([-2 * beta * 0.002 / s, 0.0, 0.0])))


def test_layout_matches_documented_parameter_count():
    assert W.layout(3)[1] == 187536
    assert W.layout(1)[1] == 187536 - 2 * (256 * 128 + 128 + 128 * 128 + 128)


# ---------------------------------------------------------------- GPU
def _net(rounds, params):
    from taremincloth_amd.warmstart import WarmStartNet
    net = WarmStartNet(rounds)
    net.set_params(params)
    return net


@pytest.mark.gpu
@pytest.mark.parametrize("rows,rounds", [(200, 3), (1000, 2), (64, 1)])
def test_gpu_forward_and_gradients_match_oracle(require_gpu, rows, rounds):
    p, feat, off, nbr, t = random_case(rows, rounds, rows)
    net = _net(rounds, p)
    y, g, loss = net.forward_backward(feat, off, nbr, t)
    lo, go, yo = W.backward(p, rounds, feat, off, nbr, t, bf16=True)
    ey = np.abs(y - yo).max() / np.abs(yo).max()
    assert ey < 2e-2, ey
    assert abs(loss - lo) <= 2e-2 * lo
    lay, _ = W.layout(rounds)
    for o, K, N in lay:
        for sl in (slice(o, o + K * N), slice(o + K * N, o + K * N + N)):
            num = np.linalg.norm(g[sl] - go[sl])
            den = np.linalg.norm(go[sl]) + 1e-12
            assert num / den < 5e-2, (o, K, N, num / den)


def hub_graph(n, seed=0):
    """A ring over the first n - 1 vertices, hubs 0..4 joined to 10..60 others, vertex n - 1
    isolated: degrees from 0 to about 60, beyond what the round kernel prefetches per row."""
    rng = np.random.default_rng(seed)
    pairs = {(i, (i + 1) % (n - 1)) for i in range(n - 1)}
    for hub, k in zip(range(5), (10, 20, 30, 45, 60)):
        for j in rng.choice(np.arange(5, n - 1), k, replace=False):
            pairs.add((hub, int(j)))
    a = np.array([p[0] for p in pairs])
    b = np.array([p[1] for p in pairs])
    off, nbr, _ = W.csr_pairs(n, a, b)
    return off, nbr


def lattice_graph(w, h):
    """4-neighbour lattice of w x h rows: degree 2 to 4, the shallow prefetch of the round kernel."""
    idx = np.arange(w * h).reshape(h, w)
    a = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    b = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    off, nbr, _ = W.csr_pairs(w * h, a, b)
    return off, nbr


@pytest.mark.gpu
def test_gpu_forward_on_a_degree_4_lattice_matches_oracle(require_gpu):
    rows, rounds = 12 * 13, 3
    p, feat, _, _, t = random_case(rows, rounds, 31)
    off, nbr = lattice_graph(12, 13)
    assert np.diff(off).max() == 4
    net = _net(rounds, p)
    y, g, loss = net.forward_backward(feat, off, nbr, t)
    lo, go, yo = W.backward(p, rounds, feat, off, nbr, t, bf16=True)
    assert np.abs(y - yo).max() / np.abs(yo).max() < 2e-2
    assert abs(loss - lo) <= 2e-2 * lo
    assert np.linalg.norm(g - go) / np.linalg.norm(go) < 5e-2


@pytest.mark.gpu
def test_gpu_forward_with_hub_and_isolated_vertices_matches_oracle(require_gpu):
    """Rows of degree above the prefetched neighbours finish their mean in a loop; an isolated
    row's mean is zero; 301 rows leave a ragged last tile."""
    rows, rounds = 301, 3
    p, feat, _, _, t = random_case(rows, rounds, 77)
    off, nbr = hub_graph(rows, seed=77)
    deg = np.diff(off)
    assert deg.max() > 40 and deg[-1] == 0
    net = _net(rounds, p)
    y, g, loss = net.forward_backward(feat, off, nbr, t)
    lo, go, yo = W.backward(p, rounds, feat, off, nbr, t, bf16=True)
    for rows_sel in (slice(None), slice(0, 5), slice(rows - 1, rows)):
        ey = np.abs(y[rows_sel] - yo[rows_sel]).max() / np.abs(yo).max()
        assert ey < 2e-2, (rows_sel, ey)
    assert abs(loss - lo) <= 2e-2 * lo
    num = np.linalg.norm(g - go)
    assert num / np.linalg.norm(go) < 5e-2


@pytest.mark.gpu
def test_gpu_adam_first_step_closed_form(require_gpu):
    p, feat, off, nbr, t = random_case(128, 1, 9)
    net = _net(1, p)
    _, g, _ = net.forward_backward(feat, off, nbr, t)
    net.adam_step(1e-3)
    got = net.get_params()
    want = W.adam_first_step(p.astype(np.float64), g.astype(np.float64), 1e-3)
    assert np.abs(got - want).max() < 1e-6


def _drape(n=24, jitter_seed=0):
    from taremincloth_amd import taremin_cloth_core as T
    pos, edges, faces = cases.grid(n, n, 0.02)
    pos = pos + np.array([-(n - 1) * 0.01, -(n - 1) * 0.01, 0.25], np.float32)
    inv = cases.areal_inv_masses_np(pos, faces, 0.15)
    sim = T.ClothSimulator(pos, edges, faces, inv, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0)
    sim.set_enable_self_collision(False)
    sim.set_compliances_raw(1e-8, 1e-6, 1e-5)
    sim.add_sphere_collider([0.0, 0.0, 0.0], 0.15, 0.3, 0.0)
    sim.set_solver_iterations(16)
    return sim, pos, edges, faces, inv


@pytest.mark.gpu
def test_gpu_recorded_features_and_targets_match_oracle(require_gpu):
    from oracle import oracle as O
    sim, pos, edges, faces, inv = _drape()
    dt_sub = 1.0 / 600.0
    for _ in range(12):  # fall onto the sphere first
        sim.step(1.0 / 60.0, 10, 16)
    sim.ws_record(4, 1)
    sim.step(dt_sub * 2, 2, 16)
    assert sim.ws_num_records() == 2
    mesh = O.build_mesh(pos, edges, faces, inv, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0)
    s = 9.81 * dt_sub * dt_sub
    cols = [(0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.15)]
    r0, r1 = sim.ws_get_record(0), sim.ws_get_record(1)
    for r in (r0, r1):
        want = W.bf16_round(W.features(r["prediction"], r["previous"], np.full(len(pos), 0.005), mesh, cols, s,
                                       mesh["mean_edge"]))
        err = np.abs(r["features"] - want)
        assert (err <= 1e-2 * np.maximum(np.abs(want), 1.0)).all(), err.max()
    # the second sample's previous correction is the first sample's correction
    prev1 = 0.5 * r0["previous"] + r0["target"][:, :3] * s
    assert np.abs(r1["previous"] - prev1).max() < 1e-7
    assert (r0["target"][:, 3] == (inv > 0)).all()


@pytest.mark.gpu
def test_gpu_gate_fallback_equals_plain_iterations(require_gpu):
    """Previous-correction mode on the first substep (correction 0): 4 iterations + gate + 6 fallback
    iterations equal 10 plain iterations bit for bit when the gate fails (threshold 0), and 4 plain
    iterations when it passes (threshold 1e9)."""
    dt_sub = 1.0 / 600.0
    outs = {}
    for key, mode, th, its in (("plain10", 0, 0.0, 10), ("plain4", 0, 0.0, 4), ("fail", 1, 0.0, 4),
                               ("pass", 1, 1e9, 4)):
        sim, pos, *_ = _drape()
        sim.set_solver_iterations(its)
        if mode:
            sim.set_warm_start(None, mode, its, 6, th, history_weight=0.5)
        sim.step_single_substep(dt_sub)
        out = np.zeros(pos.size, np.float32)
        sim.get_positions(out)
        outs[key] = out
        if mode:
            st = sim.warm_start_stats()
            assert st["checked"] == 1 and st["fallbacks"] == (1 if th == 0.0 else 0)
            # chunks of 4, then the remaining 2
            assert st["fallback_chunks"] == (2 if th == 0.0 else 0)
    assert np.array_equal(outs["fail"], outs["plain10"])
    assert np.array_equal(outs["pass"], outs["plain4"])


@pytest.mark.gpu
@pytest.mark.parametrize("damped", [False, True])
def test_gpu_gate_stops_after_the_chunk_that_passes(require_gpu, damped):
    """With the threshold between the residuals after 4 and after 8 plain iterations, the first
    check fails, the chained check after one 4-iteration chunk passes, and the remaining chunks do
    not run: the substep equals 8 plain iterations bit for bit.  `damped` runs it with constraint
    damping (5, 5, 5, 0.5), where the residual includes the damping term."""
    dt_sub = 1.0 / 600.0

    def settled():  # cloth in contact with the sphere, identical in every call
        sim, pos, *_ = _drape()
        if damped:
            sim.set_damping_all(5.0, 5.0, 5.0, 0.5)
        for _ in range(12):
            sim.step(1.0 / 60.0, 10, 16)
        return sim, pos

    res = {}
    for its in (4, 8):
        sim, pos = settled()
        sim.set_solver_iterations(its)
        sim.substep_parts(dt_sub, collide=False, self_collide=False, update=False)
        res[its], _ = sim.residual(dt_sub)
    assert res[4] > res[8]
    th = float(np.sqrt(res[4] * res[8]))
    sim, pos = settled()
    sim.set_solver_iterations(8)
    sim.step_single_substep(dt_sub)
    want = np.zeros(pos.size, np.float32)
    sim.get_positions(want)
    sim, pos = settled()
    sim.set_solver_iterations(4)
    sim.set_warm_start(None, 1, 4, 12, th, history_weight=0.5)
    sim.step_single_substep(dt_sub)
    got = np.zeros(pos.size, np.float32)
    sim.get_positions(got)
    st = sim.warm_start_stats()
    assert st["checked"] == 1 and st["fallbacks"] == 1 and st["fallback_chunks"] == 1
    assert np.array_equal(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [24, 128])
def test_gpu_host_read_gate_equals_the_device_gated_frame(require_gpu, n, monkeypatch):
    """With graph replay, a warm-started frame reads the gate on the host and launches fallback
    chunks only while it stays open; without (TCLOTH_USE_GRAPH=0) every chunk is enqueued and
    gated on the device. Both run the same kernels in the same order: 3 frames of 10 substeps
    from a settled drape, with the threshold between the residuals after 4 and 8 iterations,
    give the same positions bit for bit and the same gate statistics. n = 128 runs the
    one-launch-per-iteration schedule, n = 24 the per-colour one."""
    dt_sub = 1.0 / 600.0

    def settled():
        sim, pos, *_ = _drape(n)
        for _ in range(6):
            sim.step(1.0 / 60.0, 10, 16)
        return sim, pos

    monkeypatch.setenv("TCLOTH_USE_GRAPH", "0")
    res = {}
    for its in (4, 8):
        sim, pos = settled()
        sim.set_solver_iterations(its)
        sim.substep_parts(dt_sub, collide=False, self_collide=False, update=False)
        res[its], _ = sim.residual(dt_sub)
    th = float(np.sqrt(res[4] * res[8]))
    outs, stats = {}, {}
    for graphs in ("1", "0"):
        monkeypatch.setenv("TCLOTH_USE_GRAPH", graphs)
        sim, pos = settled()
        sim.set_solver_iterations(4)
        sim.set_warm_start(None, 1, 4, 12, th, history_weight=0.5)
        for _ in range(3):
            sim.step(1.0 / 60.0, 10)
        out = np.zeros(pos.size, np.float32)
        sim.get_positions(out)
        outs[graphs] = out
        st = sim.warm_start_stats()
        stats[graphs] = (st["checked"], st["fallbacks"], st["fallback_chunks"])
    assert stats["1"] == stats["0"] and stats["1"][0] == 30 and stats["1"][1] > 0
    assert np.array_equal(outs["1"], outs["0"])


def _stretch_graph(pos, edges, faces, inv):
    from oracle import oracle as O
    m = O.build_mesh(pos, edges, faces, inv)
    stretch = m["dist_type"] == 0
    off, nbr, _ = W.csr_pairs(len(pos), m["dist_v0"][stretch], m["dist_v1"][stretch])
    return off, nbr


@pytest.mark.gpu
def test_gpu_training_reduces_loss_on_recorded_samples(require_gpu):
    from taremincloth_amd.warmstart import WarmStartNet
    sim, pos, edges, faces, inv = _drape()
    for _ in range(10):
        sim.step(1.0 / 60.0, 10, 16)
    sim.ws_record(40, 1)
    for _ in range(4):
        sim.step(1.0 / 60.0, 10, 16)
    assert sim.ws_num_records() == 40
    off, nbr = _stretch_graph(pos, edges, faces, inv)
    recs = [sim.ws_get_record(k) for k in range(0, 40, 4)]
    net = WarmStartNet(3)
    net.init(seed=1)

    def mean_loss():
        return np.mean([net.forward_backward(r["features"], off, nbr, r["target"])[2] for r in recs])

    before = mean_loss()
    losses = net.train([sim], steps=400, lr=1e-3, seed=2)
    assert np.isfinite(losses).all()
    after = mean_loss()
    assert after < 0.8 * before, (before, after)


@pytest.mark.gpu
def test_gpu_network_inside_the_substep_matches_the_standalone_pass(require_gpu):
    from taremincloth_amd.warmstart import WarmStartNet
    sim, pos, edges, faces, inv = _drape()
    for _ in range(8):
        sim.step(1.0 / 60.0, 10, 16)
    net = WarmStartNet(3)
    net.init(seed=4)
    p = net.get_params()
    p[-(128 * 16 + 16):] = np.random.default_rng(0).normal(0, 0.05, 128 * 16 + 16)  # non-zero last layer
    net.set_params(p)
    sim.ws_record(1, 1)
    sim.set_warm_start(net, 2, 8, 0, 1e9, history_weight=0.5)
    sim.step_single_substep(1.0 / 600.0)
    r = sim.ws_get_record(0)
    y_sim = sim.ws_last_output()
    off, nbr = _stretch_graph(pos, edges, faces, inv)
    y_ref, _, _ = net.forward_backward(r["features"], off, nbr)
    assert np.array_equal(y_sim, y_ref)
    assert np.abs(y_sim[:, :3]).max() > 0


@pytest.mark.gpu
def test_gpu_calibrated_network_gains_over_the_history_alone(require_gpu):
    """What the trained network adds to the warm start it is built on: on a 128 x 128 drape that
    neither training, validation nor the gain calibration saw (seed 9104), with the gate at that
    scene's plain 64-iteration median residual, the network + previous correction needs fewer
    fallback chunks than the previous correction alone."""
    from taremincloth_amd import train_warmstart as TW
    from taremincloth_amd.warmstart import WarmStartNet

    net = WarmStartNet.load(TW.DEFAULT_WEIGHTS)
    seed, warm, frames = 9104, 25, 20
    thr = TW.scene_threshold(seed, 128, warm, frames, 10, net.history_weight)

    def run(mode):
        sim = TW.make_scene(seed, 128)
        sim.set_warm_start(net if mode == 2 else None, mode, 8, 56, thr, net.history_weight)
        for _ in range(warm):
            sim.step(1.0 / 60.0, 10)
        sim.synchronize()
        sim.ws_reset_stats()
        for _ in range(frames):
            sim.step(1.0 / 60.0, 10)
        st = sim.warm_start_stats()
        return st["fallback_chunks"], st["fallbacks"], st["checked"]

    net_chunks, net_fb, checked = run(2)
    hist_chunks, hist_fb, _ = run(1)
    print(f"[warm start] threshold {thr:.3e}: network {net_fb:.0f}/{checked:.0f} fallbacks, {net_chunks:.0f} chunks; "
          f"history alone {hist_fb:.0f} fallbacks, {hist_chunks:.0f} chunks", flush=True)
    assert checked == frames * 10
    assert net_chunks < hist_chunks
