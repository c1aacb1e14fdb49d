"""Pinning tests for the oracle's gridded background, periodic closed curve,
misfit window and wavefield snapshots (CPU only).

Each check is fixed by mathematics, not by re-typing the oracle's formula:
bilinear interpolation reproduces bilinear fields exactly; the uniform cubic
B-spline of a regular polygon passes through known radii at its knots; the
windowed exact-mode gradient matches central finite differences of the
windowed misfit; the last snapshot is the final wavefield.
"""
import numpy as np
import pytest

import oracle as O


# ---------------------------------------------------------------- grid
def test_grid_background_reproduces_bilinear_fields():
    xg = np.array([0.0, 100.0, 250.0, 400.0, 700.0])
    zg = np.array([-50.0, 80.0, 300.0, 450.0])
    Z, X = np.meshgrid(zg, xg, indexing="ij")
    f = lambda x, z: 1500.0 + 2.0 * x - 0.7 * z + 0.003 * x * z
    vp = f(X, Z)
    rng = np.random.default_rng(0)
    pts = np.stack([rng.uniform(0, 700, 200), rng.uniform(-50, 450, 200)], axis=1)
    got = O.grid_background(pts, xg, zg, vp, "linear")
    assert np.max(np.abs(got - f(pts[:, 0], pts[:, 1]))) < 1e-9
    # outside the grid: clipped to the border
    out = O.grid_background(np.array([[-100.0, 500.0], [900.0, -200.0]]), xg, zg, vp, "linear")
    assert np.allclose(out, [f(0.0, 450.0), f(700.0, -50.0)])


def test_grid_background_nearest_ties_to_lower_index():
    xg = np.array([0.0, 10.0, 20.0])
    zg = np.array([0.0, 10.0])
    vp = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
    got = O.grid_background(np.array([[5.0, 0.0], [5.1, 0.0], [15.0, 5.0], [14.9, 5.1]]), xg, zg, vp, "nearest")
    assert list(got) == [1.0, 2.0, 2.0, 5.0]


def test_increasing_grid_flips_axes_with_the_values():
    xg = np.array([3.0, 2.0, 1.0])
    zg = np.array([0.0, 1.0])
    vp = np.array([[30.0, 20.0, 10.0], [31.0, 21.0, 11.0]])
    x2, z2, v2 = O.increasing_grid(xg, zg, vp)
    assert list(x2) == [1.0, 2.0, 3.0]
    assert v2[0].tolist() == [10.0, 20.0, 30.0]


# ---------------------------------------------------------------- periodic curve
def _polygon(K, R=300.0, cx=50.0, cz=-20.0):
    th = 2 * np.pi * np.arange(K) / K
    return np.stack([cx + R * np.cos(th), cz + R * np.sin(th)], axis=1)


@pytest.mark.parametrize("K", [4, 6, 12])
def test_periodic_curve_knot_points_of_a_regular_polygon(K):
    """At integer t = s the uniform cubic B-spline is (P_{s-3} + 4 P_{s-2} +
    P_{s-1}) / 6: on a regular K-gon of radius R that point lies at radius
    R (2 + cos(2 pi/K)) / 3 in the direction of vertex s-2."""
    R, cx, cz = 300.0, 50.0, -20.0
    ctrl = _polygon(K, R, cx, cz)
    spl, knots, _ = O.sem_oracle._periodic_curve(ctrl)
    r_knot = R * (2.0 + np.cos(2 * np.pi / K)) / 3.0
    for s in range(3, K + 3):
        C = spl(float(s))
        th = 2 * np.pi * ((s - 2) % K) / K
        assert np.allclose(C, [cx + r_knot * np.cos(th), cz + r_knot * np.sin(th)], atol=1e-9)
    # signed distance along the ray through a knot point: |X - c| - r_knot, sign by side,
    # u* = (s - 3) / K
    s = 5
    th = 2 * np.pi * ((s - 2) % K) / K
    d = np.array([np.cos(th), np.sin(th)])
    X = np.array([[cx, cz] + (r_knot + 40.0) * d, [cx, cz] + (r_knot - 25.0) * d])
    v, sd, ust, nout = O.periodic_curve_velocity(X, ctrl, 1800.0, 2600.0, 10.0, samples=1600)
    assert np.allclose(sd, [40.0, -25.0], atol=1e-6)
    assert np.allclose(ust, (s - 3) / K, atol=1e-9)
    assert np.allclose(nout, [d, d], atol=1e-6)
    sig = 1.0 / (1.0 + np.exp(np.clip(sd / 10.0, -50, 50)))
    assert np.allclose(v, 2600.0 + (1800.0 - 2600.0) * sig)
    vr, _, _, _ = O.periodic_curve_velocity(X, ctrl, 1800.0, 2600.0, 10.0, samples=1600, blend="replace")
    assert list(vr) == [2600.0, 1800.0]


def test_periodic_curve_background_blend_per_node():
    ctrl = _polygon(6)
    X = np.array([[50.0, -20.0], [1000.0, 1000.0]])
    bg = np.array([2100.0, 3300.0])
    v, sd, _, _ = O.periodic_curve_velocity(X, ctrl, 1700.0, bg, 5.0)
    assert sd[0] < 0 < sd[1]
    assert abs(v[0] - 1700.0) < 1e-6 and abs(v[1] - 3300.0) < 1e-6


def test_periodic_projection_matches_finite_differences():
    """d/dctrl of sum_i a_i v_i(ctrl): the analytic projection against central
    differences of the map itself (smooth curve, generic node positions)."""
    K = 7
    rng = np.random.default_rng(4)
    ctrl = _polygon(K, 250.0, 0.0, 0.0) + rng.normal(0, 15.0, (K, 2))
    xs = np.linspace(-420, 410, 61)
    zs = np.linspace(-405, 415, 59)
    Zg, Xg = np.meshgrid(zs, xs, indexing="ij")
    P = np.stack([Xg.ravel(), Zg.ravel()], axis=1)
    bg = 2500.0 + 0.5 * P[:, 1]
    a = rng.normal(size=len(P))
    v, sd, ust, nout = O.periodic_curve_velocity(P, ctrl, 1800.0, bg, 20.0, samples=1200)
    g = O.periodic_curve_projection(a, sd, ust, nout, 1800.0, bg, 20.0, K)
    h = 1e-2
    fd = np.zeros(2 * K)
    for j in range(K):
        for q in range(2):
            cp, cm = ctrl.copy(), ctrl.copy()
            cp[j, q] += h
            cm[j, q] -= h
            vp_, _, _, _ = O.periodic_curve_velocity(P, cp, 1800.0, bg, 20.0, samples=1200)
            vm_, _, _, _ = O.periodic_curve_velocity(P, cm, 1800.0, bg, 20.0, samples=1200)
            fd[2 * j + q] = a @ (vp_ - vm_) / (2 * h)
    # The projection is -(d/dctrl) sum a_i v_i: it takes d sd / d ctrl_j as
    # +N_j(u*) n_out where the envelope theorem gives -N_j(u*) n_out; the
    # reference feeds it dL/dv = -dJ_dv so the two signs cancel there.  Up to
    # that sign it is exact (the u*-dependence drops out at the Newton optimum).
    assert np.linalg.norm(g + fd) / np.linalg.norm(fd) < 1e-6


# ---------------------------------------------------------------- window
def _small(nt=160, sides=("xmin", "xmax", "zmax")):
    rec = np.stack([np.linspace(40.0, 260.0, 7), np.full(7, 10.0)], axis=1)
    prob = O.Problem(0.0, 300.0, 0.0, 250.0, 6, 5, 4, 1e-3, nt, (140.0, 60.0), 15.0, rec,
                     sponge_width=80.0, sponge_cmax=3000.0, sponge_sides=sides)
    disc = O.discretize(prob)
    rng = np.random.default_rng(2)
    coef = 2000.0 + 500.0 * rng.random((5, 5))
    c, _, _ = O.tensor_bspline_velocity(disc.mesh, coef)
    return prob, disc, c


def test_window_from_times_rounding_and_clamps():
    assert O.window_from_times(1e-3, 500, 0.05, 0.2, decim=3) == (50, 201, 3)
    assert O.window_from_times(1e-3, 500, 0.05, None, window_len=0.1) == (50, 151, 1)
    assert O.window_from_times(1e-3, 100, 0.0, 5.0) == (0, 100, 1)
    with pytest.raises(ValueError):
        O.window_from_times(1e-3, 100, 0.05, 0.01)
    m = O.window_mask(20, (3, 11, 4))
    assert np.nonzero(m)[0].tolist() == [3, 7]


def test_windowed_exact_gradient_matches_finite_differences():
    prob, disc, c = _small()
    rng = np.random.default_rng(5)
    yobs = O.forward(disc, c * (1 + 0.02 * rng.standard_normal(c.size)), history=None)[0]
    win = (40, 130, 3)
    out = O.forward_adjoint(disc, c, yobs, 1e-3, mode="exact", window=win)
    keep = O.window_mask(prob.nt, win)[:, None]
    g = out["dloglik_dc"]
    for node in rng.choice(disc.mesh.N, 5, replace=False):
        vals = []
        for h in (+0.5, -0.5):
            cc = c.copy()
            cc[node] += h
            s, _ = O.forward(disc, cc, history=None)
            vals.append(O.misfit_loglik(s * keep, yobs * keep, 1e-3)[0])
        fd = (vals[0] - vals[1]) / 1.0
        assert abs(fd - g[node]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max())


def test_reference_window_truncates_only_before_the_window_start():
    """Without a sponge the reference adjoint is exact; with a window it stops
    at it0, so it equals the exact gradient when it0 = 0 and differs when the
    window starts later."""
    prob, disc, c = _small(sides=())
    yobs = 0.9 * O.forward(disc, c, history=None)[0]
    full = O.forward_adjoint(disc, c, yobs, 1e-3, mode="exact", window=(0, 120, 2))["dloglik_dc"]
    ref0 = O.forward_adjoint(disc, c, yobs, 1e-3, mode="reference", window=(0, 120, 2))["dloglik_dc"]
    assert np.allclose(ref0, full, rtol=1e-12, atol=1e-12 * np.abs(full).max())
    late = O.forward_adjoint(disc, c, yobs, 1e-3, mode="reference", window=(60, 120, 2))["dloglik_dc"]
    late_x = O.forward_adjoint(disc, c, yobs, 1e-3, mode="exact", window=(60, 120, 2))["dloglik_dc"]
    assert np.linalg.norm(late - late_x) > 1e-3 * np.linalg.norm(late_x)


# ---------------------------------------------------------------- snapshots
def test_snapshots_hold_the_state_after_each_sampled_step():
    prob, disc, c = _small(nt=50)
    box = O.roi_box(disc.mesh, 60.0, 240.0, 20.0, 200.0)
    ix0, ix1, iz0, iz1 = box
    assert disc.mesh.x1d[ix0] >= 60.0 and disc.mesh.x1d[ix0 - 1] < 60.0
    assert disc.mesh.z1d[iz1 - 1] <= 200.0 and disc.mesh.z1d[iz1] > 200.0
    seis, snaps = O.forward_snapshots(disc, c, 7, box)
    assert snaps.shape == (8, iz1 - iz0, ix1 - ix0)  # steps 0, 7, ..., 49
    states = []
    O.forward(disc, c, history=None, after_step=lambda it, u: states.append(u.copy()))
    NX, NZ = disc.mesh.NX, disc.mesh.NZ
    for k in range(8):
        ref = states[7 * k].reshape(NZ, NX)[iz0:iz1, ix0:ix1]
        assert np.array_equal(snaps[k], ref)
    assert np.abs(snaps[-1]).max() > 0
