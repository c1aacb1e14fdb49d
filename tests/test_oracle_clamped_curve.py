"""Pinning tests for the oracle's clamped closed-curve velocity map and its
control-point projection (CPU only).

The curve is evaluated here independently of the oracle (the textbook
Cox-de Boor recursion on the clamped uniform knot vector of 9 control points
P0..P5, P0, P0, P0), densely sampled, and the oracle's outputs are checked
against what geometry fixes: the unsigned distance equals the brute-force
nearest distance to the dense curve, the sign equals the winding number of the
dense polygon, the returned parameter lands on a nearest point, the returned
normal is the spatial gradient of the signed distance, the velocity is the
two-level blend with the mean value on the curve, and the projection equals
-d/dctrl of sum_i a_i v_i by central differences.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle as O

VIN, VOUT, TAU = 2000.0, 3000.0, 10.0


def _ctrl(seed=3):
    rng = np.random.default_rng(seed)
    th = np.linspace(0.0, 2 * np.pi, 7)[:-1] + 0.3
    return np.stack([500 + 260 * np.cos(th), 480 + 220 * np.sin(th)], 1) + rng.normal(0, 15, (6, 2))


def _cox_de_boor(u, knots, i, k):
    """N_{i,k}(u) by the recursion, right end included in the last span."""
    if k == 0:
        last = knots[i + 1] == knots[-1] and knots[i] < knots[i + 1]
        return ((knots[i] <= u) & ((u < knots[i + 1]) | (last & (u == knots[-1])))).astype(float)
    out = np.zeros_like(u)
    d1 = knots[i + k] - knots[i]
    if d1 > 0:
        out += (u - knots[i]) / d1 * _cox_de_boor(u, knots, i, k - 1)
    d2 = knots[i + k + 1] - knots[i + 1]
    if d2 > 0:
        out += (knots[i + k + 1] - u) / d2 * _cox_de_boor(u, knots, i + 1, k - 1)
    return out


def _curve(ctrl6, u):
    """Clamped cubic B-spline through control points P0..P5, P0, P0, P0 on
    knots [0]*4 + [1/6 .. 5/6] + [1]*4 (closed: C(0) = C(1) = P0)."""
    P = np.vstack([ctrl6, ctrl6[0], ctrl6[0], ctrl6[0]])
    knots = np.concatenate([np.zeros(4), np.arange(1, 6) / 6.0, np.ones(4)])
    B = np.stack([_cox_de_boor(u, knots, i, 3) for i in range(9)], axis=1)
    return B @ P


def _winding(X, poly):
    """Winding number of the closed polyline about each point of X."""
    out = np.empty(len(X))
    for s in range(0, len(X), 64):
        d = poly[None, :, :] - X[s:s + 64, None, :]
        ang = np.arctan2(d[..., 1], d[..., 0])
        da = np.diff(np.concatenate([ang, ang[:, :1]], axis=1), axis=1)
        da = (da + np.pi) % (2 * np.pi) - np.pi
        out[s:s + 64] = da.sum(1) / (2 * np.pi)
    return out


def test_independent_curve_is_closed_through_p0():
    ctrl = _ctrl()
    C = _curve(ctrl, np.array([0.0, 1.0]))
    assert np.allclose(C, [ctrl[0], ctrl[0]], atol=1e-12)
    # partition of unity of the 9 basis functions on the whole interval
    knots = np.concatenate([np.zeros(4), np.arange(1, 6) / 6.0, np.ones(4)])
    u = np.linspace(0.0, 1.0, 101)
    assert np.allclose(sum(_cox_de_boor(u, knots, i, 3) for i in range(9)), 1.0, atol=1e-12)


def test_signed_distance_parameter_and_sign_against_brute_force():
    ctrl = _ctrl()
    rng = np.random.default_rng(11)
    X = np.concatenate([rng.uniform([120, 130], [880, 830], (600, 2)), ctrl[:1]])
    v, sd, ust, nout = O.closed_curve_velocity(X, ctrl, VIN, VOUT, TAU)
    dense = _curve(ctrl, np.linspace(0.0, 1.0, 200001))
    dist, _ = cKDTree(dense).query(X)
    # dense spacing <= 0.02 m, so the sampled distance overestimates by < 1e-2 m
    assert np.max(np.abs(np.abs(sd) - dist)) < 1e-2
    assert abs(sd[-1]) < 1e-9  # P0 lies on the curve
    # the returned parameter is a nearest point of the independently evaluated curve
    assert np.max(np.abs(np.linalg.norm(X - _curve(ctrl, ust), axis=1) - np.abs(sd))) < 1e-6
    # sign: negative inside (winding number 1 or -1), positive outside (0)
    far = dist > 1.0
    wn = np.rint(np.abs(_winding(X[far], dense[::20])))
    assert set(np.unique(wn)) <= {0.0, 1.0}
    assert np.array_equal(np.sign(sd[far]), np.where(wn == 1.0, -1.0, 1.0))
    assert np.sum(wn == 1.0) > 50 and np.sum(wn == 0.0) > 50
    # two-level blend: mean value on the curve, the plateaus far from it
    assert abs(v[-1] - 0.5 * (VIN + VOUT)) < 1e-6
    assert np.allclose(v[sd < -20 * TAU], VIN, atol=1e-3)
    assert np.allclose(v[sd > 20 * TAU], VOUT, atol=1e-3)
    order = np.argsort(sd)
    assert np.all(np.diff(v[order]) >= 0)


def test_normal_is_the_gradient_of_the_signed_distance():
    ctrl = _ctrl()
    rng = np.random.default_rng(5)
    X = rng.uniform([150, 150], [850, 810], (300, 2))
    _, sd, ust, nout = O.closed_curve_velocity(X, ctrl, VIN, VOUT, TAU)
    # stay away from the kink at P0 (u* = 0 or 1) and from points with two
    # nearest points (the medial axis), where sd is not differentiable
    keep = (ust > 0.02) & (ust < 0.98) & (np.abs(sd) > 5.0) & (np.abs(sd) < 120.0)
    X, nout = X[keep], nout[keep]
    h = 1e-3
    g = np.zeros_like(X)
    for q in range(2):
        e = np.zeros(2)
        e[q] = h
        sp = O.closed_curve_velocity(X + e, ctrl, VIN, VOUT, TAU)[1]
        sm = O.closed_curve_velocity(X - e, ctrl, VIN, VOUT, TAU)[1]
        g[:, q] = (sp - sm) / (2 * h)
    assert len(X) > 40
    assert np.max(np.abs(g - nout)) < 1e-5
    assert np.allclose(np.linalg.norm(nout, axis=1), 1.0, atol=1e-12)


@pytest.mark.parametrize("seed", [3, 8])
def test_clamped_projection_matches_finite_differences(seed):
    """d/dctrl of sum_i a_i v_i(ctrl) against the analytic projection, nodes
    whose nearest point is the kink at P0 included.  Nodes within 5 cm of the
    curve get no weight: the sign comes from a point-in-polygon test on the
    800-sample polygon, whose chords sag about R (pi/800)^2 / 2 ~ 2 mm off the
    smooth curve, so a node inside that band can take the other side's sign
    and the reversed normal (seed 8 places one 2.4 mm away); the reference
    decides the sign the same way."""
    ctrl = _ctrl(seed)
    xs = np.linspace(0.0, 1000.0, 91)
    Zg, Xg = np.meshgrid(xs, xs, indexing="ij")
    X = np.stack([Xg.ravel(), Zg.ravel()], axis=1)
    a = np.random.default_rng(seed + 100).normal(size=len(X))
    _, sd, ust, nout = O.closed_curve_velocity(X, ctrl, VIN, VOUT, TAU)
    a[np.abs(sd) < 0.05] = 0.0
    assert np.any((ust == 0.0) & (a != 0.0)) or np.any((ust == 1.0) & (a != 0.0))
    g = O.closed_curve_projection(a, sd, ust, nout, VIN, VOUT, TAU)
    h = 1e-2
    fd = np.zeros(12)
    for j in range(12):
        cp, cm = ctrl.ravel().copy(), ctrl.ravel().copy()
        cp[j] += h
        cm[j] -= h
        vp = O.closed_curve_velocity(X, cp.reshape(6, 2), VIN, VOUT, TAU)[0]
        vm = O.closed_curve_velocity(X, cm.reshape(6, 2), VIN, VOUT, TAU)[0]
        fd[j] = a @ (vp - vm) / (2 * h)
    # the projection is -(d/dctrl) sum a_i v_i (the reference feeds it -dJ/dv)
    assert np.linalg.norm(g + fd) / np.linalg.norm(fd) < 1e-6
