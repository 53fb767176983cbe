"""CPU ORACLE of the learned warm start -- TEST INFRASTRUCTURE ONLY.

numpy FP64 restatement of the product's warm-start network (forward, loss,
backward), of its input features and of one Adam step.  Only tests/ and the
smoke check import it, as the checker of the HIP kernels.

The warm start has no counterpart in the reference: it sits between predict
and the projection loop of the reference substep
(crates/cloth_core/src/simulation/dispatch.rs:137 -> 169), so nothing in the
reference pins it.  What pins this oracle instead:
  * backward: central finite differences of the FP64 loss (exact arithmetic,
    no BF16 rounding) on small graphs, and a dense-matrix statement of the
    graph transpose (tests/test_warmstart.py);
  * features: closed-form two- and three-vertex configurations;
  * Adam: the closed form of the first step (theta - lr * g / (|g| + eps)).

Where the GPU rounds an operand to BF16 (inputs, weights, stored activations,
the gradients entering MFMA), `bf16=True` rounds the same value the same way,
so GPU and oracle differ only by FP32 accumulation order and the rare BF16
rounding-boundary flip that follows from it.

Network (H = 128, F = 32 inputs, O = 16 outputs, R rounds):
    h0      = W_e2 relu(W_e1 f + b_e1) + b_e2
    h_{r+1} = h_r + W_r2 relu(W_r1 [h_r, mean_{j in N(i)} h_r[j]] + b_r1) + b_r2
    y       = W_d2 relu(W_d1 h_R + b_d1) + b_d2
    loss    = sum_{free i} |y_i[0:3] - t_i|^2 / n_free
"""
from __future__ import annotations

import numpy as np

H, F, O = 128, 32, 16


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(rounds: int):
    """(offset, K, N) of every dense layer in the flat parameter vector."""
    dims = [(F, H), (H, H)] + [(2 * H, H), (H, H)] * rounds + [(H, H), (H, O)]
    out, off = [], 0
    for K, N in dims:
        out.append((off, K, N))
        off += K * N + N
    return out, off


def unpack(params: np.ndarray, rounds: int):
    lay, n = layout(rounds)
    assert params.size == n
    p = np.asarray(params, np.float64)
    return [(p[o:o + K * N].reshape(K, N), p[o + K * N:o + K * N + N]) for o, K, N in lay]


def init_params(rounds: int, seed: int, scale: float = 1.0) -> np.ndarray:
    """Random parameters for tests (numpy generator; not the library's initialiser)."""
    lay, n = layout(rounds)
    rng = np.random.default_rng(seed)
    p = np.zeros(n, np.float32)
    for o, K, N in lay:
        p[o:o + K * N] = rng.uniform(-1, 1, K * N) * np.sqrt(6.0 / K) * 0.5 * scale
        p[o + K * N:o + K * N + N] = rng.uniform(-0.1, 0.1, N) * scale
    return p


def mean_agg(h: np.ndarray, off: np.ndarray, nbr: np.ndarray) -> np.ndarray:
    deg = np.diff(off).astype(np.int64)
    rows = np.repeat(np.arange(len(deg)), deg)
    s = np.zeros_like(h)
    np.add.at(s, rows, h[nbr])
    return np.where(deg[:, None] > 0, s / np.maximum(deg, 1)[:, None], 0.0)


def mean_agg_transpose(g: np.ndarray, off: np.ndarray, nbr: np.ndarray) -> np.ndarray:
    """Adjoint of mean_agg: out[j] = sum over i with j in N(i) of g[i] / deg(i)."""
    deg = np.diff(off).astype(np.int64)
    rows = np.repeat(np.arange(len(deg)), deg)
    out = np.zeros_like(g)
    np.add.at(out, nbr, g[rows] / np.maximum(deg[rows], 1)[:, None])
    return out


def forward(params, rounds, feat, off, nbr, bf16=True):
    """Returns y (rows x 16) and the saved activations for backward."""
    r_ = bf16_round if bf16 else (lambda x: np.asarray(x, np.float64))
    L = unpack(params, rounds)
    b_ = (lambda b: b.astype(np.float32).astype(np.float64)) if bf16 else (lambda b: b)
    Wq = [(r_(W), b_(b)) for W, b in L]
    f = r_(feat)
    (W1, b1), (W2, b2) = Wq[0], Wq[1]
    a_enc = r_(np.maximum(f @ W1 + b1, 0.0))
    h = [r_(a_enc @ W2 + b2)]
    xs, as_ = [], []
    for r in range(rounds):
        (W1, b1), (W2, b2) = Wq[2 + 2 * r], Wq[3 + 2 * r]
        x = np.concatenate([h[r], r_(mean_agg(h[r], off, nbr))], axis=1)
        a = r_(np.maximum(x @ W1 + b1, 0.0))
        h.append(r_(h[r] + (a @ W2 + b2)))
        xs.append(x)
        as_.append(a)
    (W1, b1), (W2, b2) = Wq[2 + 2 * rounds], Wq[3 + 2 * rounds]
    a_dec = r_(np.maximum(h[rounds] @ W1 + b1, 0.0))
    y = a_dec @ W2 + b2
    return y, dict(f=f, a_enc=a_enc, h=h, xs=xs, as_=as_, a_dec=a_dec, W=Wq)


def loss_grad(y, t):
    """Loss and dLoss/dy for targets t (rows x 4: xyz, weight)."""
    m = t[:, 3] > 0
    nfree = max(int(m.sum()), 1)
    d = np.zeros_like(y)
    d[m, :3] = y[m, :3] - t[m, :3]
    return float((d ** 2).sum() / nfree), 2.0 * d / nfree


def backward(params, rounds, feat, off, nbr, t, bf16=True):
    """Loss and the gradient of every parameter (flat, layer layout of `layout`)."""
    r_ = bf16_round if bf16 else (lambda x: np.asarray(x, np.float64))
    y, s = forward(params, rounds, feat, off, nbr, bf16)
    loss, dy = loss_grad(y, t)
    W = s["W"]
    grads = [None] * len(W)

    def block(li, x, a, dy_):
        """Two-layer block li (layers li, li+1): returns dX, fills the weight gradients."""
        dyb = r_(dy_)
        W1, W2 = W[li][0], W[li + 1][0]
        dz = np.where(a > 0, dyb @ W2.T, 0.0)
        dzb = r_(dz)
        grads[li + 1] = (a.T @ dyb, dyb.sum(0))
        grads[li] = (x.T @ dzb, dzb.sum(0))
        return dzb @ W1.T

    dh = block(2 + 2 * rounds, s["h"][rounds], s["a_dec"], dy)
    for r in reversed(range(rounds)):
        dx = block(2 + 2 * r, s["xs"][r], s["as_"][r], dh)
        dh = dh + dx[:, :H] + mean_agg_transpose(dx[:, H:], off, nbr)
    block(0, s["f"], s["a_enc"], dh)
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat, y


def adam_first_step(theta, g, lr, eps=1e-8):
    """Closed form of Adam's first step (bias-corrected m = g, v = g^2)."""
    return theta - lr * g / (np.abs(g) + eps)


# ------------------------------------------------------------------ features
def csr_pairs(nv, a, b, rest=None, keep=None):
    """CSR of the pairs (a[k], b[k]) in both directions, in pair order."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    idx = np.arange(len(a)) if keep is None else np.nonzero(keep)[0]
    lists = [[] for _ in range(nv)]
    for k in idx:
        L = 0.0 if rest is None else float(rest[k])
        lists[a[k]].append((b[k], L))
        lists[b[k]].append((a[k], L))
    off = np.zeros(nv + 1, np.uint32)
    off[1:] = np.cumsum([len(x) for x in lists])
    nbr = np.array([j for x in lists for j, _ in x], np.uint32)
    rr = np.array([L for x in lists for _, L in x], np.float64)
    return off, nbr, rr


def soft(x):
    return x / (1.0 + np.abs(x) / 16.0)


def features(ppred, dprev, thickness, mesh, colliders, s, lscale, beta=0.5, fp32_points=False):
    """Input features of every vertex (rows x 32, FP64 before BF16 rounding).

    ppred: N x 4 (predicted xyz, inverse mass); dprev: N x 3 previous correction;
    thickness: N; mesh: the mesh arrays (dist_v0/v1/rest, bend_v2/v3/rest);
    colliders: list of (type, A(3), B(3), radius); s = max(|g|, 1) dt^2;
    lscale = mean edge length; beta = history weight (strains at prediction + beta * dprev).
    fp32_points: form the points prediction + beta * dprev as the GPU stores them, in fp32
    (beta * dprev rounded, then the sum rounded); on a large cloth the coordinates' fp32 spacing
    (2.4e-7 m at 2.5 m) is not small against s, so FP64 points would differ by that much.
    """
    nv = len(ppred)
    if fp32_points:
        b32 = np.float32(beta)
        p0 = (ppred[:, :3].astype(np.float32) + (b32 * dprev.astype(np.float32))).astype(np.float64)
    else:
        p0 = ppred[:, :3].astype(np.float64) + beta * dprev.astype(np.float64)
    f = np.zeros((nv, F))
    f[:, 0:3] = soft(dprev / s)
    for col, (a, b, rest) in ((3, ("dist_v0", "dist_v1", "dist_rest")), (6, ("bend_v2", "bend_v3", "bend_rest"))):
        off, nbr, rr = csr_pairs(nv, mesh[a], mesh[b], mesh[rest])
        acc = np.zeros((nv, 3))
        rows = np.repeat(np.arange(nv), np.diff(off).astype(np.int64))
        d = p0[nbr] - p0[rows]
        ln = np.linalg.norm(d, axis=1)
        ok = ln > 1e-12
        contrib = np.zeros_like(d)
        contrib[ok] = d[ok] * ((ln[ok] - rr[ok]) / ln[ok])[:, None]
        np.add.at(acc, rows, contrib)
        f[:, col:col + 3] = soft(acc / s)
        if col == 3:
            f[:, 10] = np.diff(off) * 0.125
        else:
            f[:, 11] = np.diff(off) * 0.125
    f[:, 9] = (ppred[:, 3] > 0).astype(np.float64)
    best = np.full(nv, np.inf)
    bn = np.zeros((nv, 3))
    for typ, A, B, radius in colliders:
        A = np.asarray(A, np.float64)
        B = np.asarray(B, np.float64)
        if typ == 2:
            n = np.broadcast_to(B / np.linalg.norm(B), (nv, 3))
            sd = (p0 - A) @ (B / np.linalg.norm(B))
        else:
            q = np.broadcast_to(A, (nv, 3))
            if typ == 1:
                ab = B - A
                l2 = ab @ ab
                tt = np.clip(((p0 - A) @ ab) / l2, 0.0, 1.0) if l2 > 1e-7 else np.zeros(nv)
                q = A + tt[:, None] * ab
            d = p0 - q
            dist = np.linalg.norm(d, axis=1)
            n = np.where(dist[:, None] > 1e-7, d / np.maximum(dist, 1e-30)[:, None], np.array([0.0, 0.0, 1.0]))
            sd = dist - radius
        sd = sd - thickness
        better = sd < best
        best = np.where(better, sd, best)
        bn = np.where(better[:, None], n, bn)
    r = np.minimum(best / lscale, 4.0) if colliders else np.full(nv, 4.0)
    f[:, 12] = soft(r)
    near = r < 4.0
    f[near, 13:16] = bn[near]
    return f
