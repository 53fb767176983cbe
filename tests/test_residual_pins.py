"""The relative XPBD residual r = ||h|| / ||L0|| of the oracle, pinned by closed forms (CPU).

h_j = C_j(p) + (alpha_j / dt^2) lambda_j + gamma_j grad C_j . (p - x) over distance, shear and
bending constraints, the numerator of the damped projection's update (distance.wgsl:60-124,
bending.wgsl:59-126), so the projection drives it to zero.  Each case builds a configuration whose
constraint set and lengths are known by construction (3-4-5 rectangles give exact float lengths),
so a wrong alpha~ (alpha instead of alpha / dt^2), a wrong sign, a wrong normaliser, a dropped
bending term or a missing damping term changes the value well beyond the tolerance.  The GPU
residual is compared with this oracle in tests/test_gpu_parity.py.
"""
import numpy as np
import pytest

from oracle import oracle as O

DT = np.float32(1.0 / 600.0)


def f32(x):
    return float(np.float32(x))


def chain(xs, precision="f64", inv_masses=None):
    """Vertices on the x axis at xs, consecutive ones joined by a stretch edge (no faces).  An
    edge-only mesh also gets a bending pair across each interior vertex (here (0, 2) for three
    vertices); the tests that want distance constraints alone give it compliance 1e10."""
    pos = np.zeros((len(xs), 3), np.float32)
    pos[:, 0] = xs
    edges = np.array([[i, i + 1] for i in range(len(xs) - 1)], np.uint32)
    s = O.OracleSim(pos, edges, inv_masses=inv_masses, precision=precision)
    s.set_gravity(0, 0, 0)
    s.set_damping(0)
    s.set_enable_self_collision(False)
    return s, pos


def quad(diagonal_is_edge, precision="f64"):
    """A 0.75 x 1.0 rectangle of two triangles (0,1,2), (0,2,3); every pair distance is 0.75, 1 or 1.25."""
    pos = np.array([[0, 0, 0], [0.75, 0, 0], [0.75, 1.0, 0], [0, 1.0, 0]], np.float32)
    edges = [[0, 1], [1, 2], [2, 3], [3, 0]] + ([[0, 2]] if diagonal_is_edge else [])
    s = O.OracleSim(pos, np.array(edges, np.uint32), faces=np.array([[0, 1, 2], [0, 2, 3]], np.uint32),
                    precision=precision)
    s.set_gravity(0, 0, 0)
    s.set_damping(0)
    s.set_enable_self_collision(False)
    return s, pos


def pairs(s):
    m = s.mesh()
    d = [(int(a), int(b)) for a, b in zip(m["dist_v0"], m["dist_v1"])]
    b = [(int(a), int(c)) for a, c in zip(m["bend_v2"], m["bend_v3"])]
    return d, b, m


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_rest_state_has_zero_residual(precision):
    s, pos = quad(True, precision)
    d, b, _ = pairs(s)
    assert len(d) == 5 and len(b) == 1 and sorted(b[0]) == [1, 3]
    s.set_compliances_raw(1e-4, 1e-4, 1e-3)
    s.set_state(p=pos.astype(np.float64))
    r, mx = s.residual(DT)
    assert r == 0.0 and mx == 0.0


def test_one_stretched_constraint():
    """|delta + (alpha / dt^2) lambda| / L0 for one constraint of rest 1 stretched by 0.125."""
    s, pos = chain([0.0, 1.0])
    alpha = f32(1e-4)
    s.set_compliances_raw(alpha, alpha, alpha)
    delta = 0.125
    lam = f32(-0.05 * float(DT) ** 2 / alpha)  # alpha~ lambda = -0.05: h = 0.075
    p = pos.astype(np.float64)
    p[1, 0] += delta
    s.set_state(p=p)
    s.set_lambdas([lam])
    h = delta + alpha / float(DT) ** 2 * lam
    r, mx = s.residual(DT)
    assert abs(h - 0.075) < 1e-6
    assert r == pytest.approx(abs(h), rel=1e-12) and mx == pytest.approx(abs(h), rel=1e-12)
    # what the plausible mistakes would give instead
    assert abs(r - abs(delta + alpha * lam)) > 0.04       # alpha without the 1/dt^2
    assert abs(r - abs(delta - alpha / float(DT) ** 2 * lam)) > 0.09  # sign of the lambda term


def test_two_constraints_use_the_root_sum_of_rest_lengths():
    """sqrt(h1^2 + h2^2) / sqrt(L1^2 + L2^2), max_j |h_j| / L_j."""
    s, pos = chain([0.0, 1.0, 1.75])
    alpha = f32(1e-5)
    s.set_compliances_raw(alpha, alpha, 1e10)
    p = pos.astype(np.float64)
    p[1, 0] += 0.125           # first stretched by 0.125, second compressed by 0.125
    p[2, 0] += 0.0625          # second: 0.75 - 0.0625 = 0.6875
    s.set_state(p=p)
    d, b, m = pairs(s)
    assert len(d) == 2 and b == [(0, 2)]
    lam = {(0, 1): f32(2e-4), (1, 2): f32(-3e-4)}
    lam_sorted = [lam[tuple(sorted(k))] for k in d]
    s.set_lambdas(lam_sorted)
    at = alpha / float(DT) ** 2
    h1 = 0.125 + at * lam[(0, 1)]
    h2 = -0.0625 + at * lam[(1, 2)]
    r, mx = s.residual(DT)
    assert r == pytest.approx(np.hypot(h1, h2) / np.hypot(1.0, 0.75), rel=1e-12)
    assert mx == pytest.approx(max(abs(h1) / 1.0, abs(h2) / 0.75), rel=1e-12)
    assert abs(r - np.hypot(h1, h2) / 1.75) > 1e-3   # sum of rest lengths as the normaliser


def test_bending_and_shear_terms_enter():
    """Out-of-plane lift of vertex 1 by dz: the bending pair (1, 3) and the four sides change."""
    dz = 0.5
    s, pos = quad(True)
    s.set_compliances_raw(0.0, 0.0, 0.0)
    p = pos.astype(np.float64)
    p[1, 2] = dz
    s.set_state(p=p)
    hs = [np.hypot(0.75, dz) - 0.75, np.hypot(1.0, dz) - 1.0, 0.0, 0.0, 0.0]  # sides 01, 12, 23, 30, diag 02
    hb = np.hypot(1.25, dz) - 1.25
    l2 = 2 * 0.75 ** 2 + 2 * 1.0 + 1.25 ** 2 + 1.25 ** 2
    r, mx = s.residual(DT)
    full = np.sqrt(sum(h * h for h in hs) + hb * hb) / np.sqrt(l2)
    assert r == pytest.approx(full, rel=1e-12)
    assert abs(r - np.sqrt(sum(h * h for h in hs)) / np.sqrt(l2 - 1.25 ** 2)) > 1e-3  # bending dropped
    assert mx == pytest.approx(max(hs[0] / 0.75, hs[1] / 1.0, hb / 1.25), rel=1e-12)
    # without the diagonal as an input edge it becomes a shear pair, with its cross (1, 3) as well
    s2, _ = quad(False)
    d2, b2, m2 = pairs(s2)
    assert len(d2) == 6 and len(b2) == 0 and int(np.sum(m2["dist_type"] == 1)) == 2
    s2.set_compliances_raw(0.0, 0.0, 0.0)
    s2.set_state(p=p)
    r2, _ = s2.residual(DT)
    assert r2 == pytest.approx(full, rel=1e-12)


def test_skipped_constraints_are_excluded():
    # both ends pinned: w0 + w1 = 0 <= 1e-7; only the second constraint counts
    s, pos = chain([0.0, 1.0, 1.75], inv_masses=np.array([0.0, 0.0, 1.0], np.float32))
    s.set_compliances_raw(0.0, 0.0, 1e10)
    p = pos.astype(np.float64)
    p[1, 0] += 0.125
    s.set_state(p=p)
    r, _ = s.residual(DT)
    assert r == pytest.approx(0.125 / 0.75, rel=1e-12)
    # compliance >= 1e9 excludes the distance constraints; the bending pair alone remains
    s, pos = quad(True)
    s.set_compliances_raw(1e10, 1e10, 0.0)
    p = pos.astype(np.float64)
    p[1, 2] = 0.5
    s.set_state(p=p)
    r, _ = s.residual(DT)
    assert r == pytest.approx((np.hypot(1.25, 0.5) - 1.25) / 1.25, rel=1e-12)
    # coincident endpoints (|p0 - p1| <= 1e-7) are skipped
    s, pos = chain([0.0, 1.0, 1.75])
    s.set_compliances_raw(0.0, 0.0, 1e10)
    p = pos.astype(np.float64)
    p[1, 0] = 0.0
    s.set_state(p=p)
    r, _ = s.residual(DT)
    assert r == pytest.approx(1.0 / 0.75, rel=1e-12)


def test_damping_term_closed_form():
    """gamma grad C . (p - x) with gamma = damping * 1e-5 / dt: tension damping when stretched,
    compression damping when compressed."""
    alpha = f32(1e-4)
    for delta, damp_key in ((0.125, "t"), (-0.125, "c")):
        s, pos = chain([0.0, 1.0])
        s.set_compliances_raw(alpha, alpha, alpha)
        s.set_damping_all(5.0, 2.0, 0.0, 0.0)
        p = pos.astype(np.float64)
        p[1, 0] += delta
        s.set_state(p=p)            # x stays at rest: p - x = (0, delta) per endpoint
        lam = f32(1e-6)
        s.set_lambdas([lam])
        gamma = (5.0 if damp_key == "t" else 2.0) * float(np.float32(1e-5)) / float(DT)
        h = delta + alpha / float(DT) ** 2 * lam + gamma * delta
        r, _ = s.residual(DT)
        assert r == pytest.approx(abs(h), rel=1e-6)
        assert abs(r - abs(delta + alpha / float(DT) ** 2 * lam)) > 0.5 * gamma * 0.125


def test_projection_drives_the_damped_residual_to_zero():
    """One constraint alone: a single damped projection solves it, so the residual after one
    iteration is at the rounding floor with the damping term, and gamma * v_proj without it."""
    s, pos = chain([0.0, 1.0])
    s.set_compliances_raw(f32(1e-6), f32(1e-6), f32(1e-6))
    s.set_damping_all(5.0, 5.0, 5.0, 0.5)
    s.set_solver_iterations(1)
    s.set_state(v=np.array([[0, 0, 0], [3.0, 0, 0]], np.float64))  # predicted stretch of 3 * dt
    s.substep_parts(DT, False, False, False)
    r, _ = s.residual(DT)
    assert r < 1e-12
    s.set_damping_all(0.0, 0.0, 0.0, 0.0)  # the same state measured without the damping term
    r_plain, _ = s.residual(DT)
    assert r_plain > 1e-5
