"""Pins the oracle's time stepping, adjoint and gradient.

* energy identity of the undamped, source-free leapfrog
* lambda^T K(c) u is quadratic in c: its central difference is exact, so the
  correlation kernel is checked to rounding
* the sponge-aware ('exact') adjoint equals central finite differences of the
  full log-likelihood; with no sponge the reference recurrence equals it too
* checkpoint/recompute is bit-identical to the stored history
"""
import numpy as np
import pytest

import oracle as O


def _prob(sides=("xmin", "xmax", "zmax"), nt=160, dt=1e-3, nex=5, nez=4, gate=0.3):
    rec = np.array([[40.0 + 45 * i, 10.0] for i in range(7)])
    p = O.Problem(0.0, 100.0 * nex, 0.0, 100.0 * nez, nex, nez, 4, dt, nt, (230.0, 60.0), 15.0, rec,
                  sponge_width=80.0, sponge_cmax=3000.0, sponge_sides=sides, gate_time=gate)
    return p, O.discretize(p)


def test_energy_identity_without_sponge():
    """After the source stops, E = 1/2 |(u^{n+1}-u^n)/dt|_M^2 + 1/2 u^{n+1}.K u^n
    is conserved by leapfrog (to rounding)."""
    p, d = _prob(sides=(), nt=400, gate=0.02)
    rng = np.random.default_rng(0)
    c = 2000 + 500 * rng.random(d.mesh.N)
    K = O.stiffness_csr(c, d.mesh, d.ops)
    _, hist = O.forward(d, c, history="full", K=K)
    H = hist.full
    E = []
    for n in range(100, p.nt - 1):
        v = (H[n + 1] - H[n]) / p.dt
        E.append(0.5 * v @ (d.M * v) + 0.5 * H[n + 1] @ (K @ H[n]))
    E = np.array(E)
    assert np.abs(E - E[0]).max() < 1e-9 * abs(E[0])


def test_sensitivity_is_exact_derivative_of_quadratic_form():
    p, d = _prob()
    rng = np.random.default_rng(1)
    c = 2000 + 500 * rng.random(d.mesh.N)
    u, lam = rng.standard_normal(d.mesh.N), rng.standard_normal(d.mesh.N)
    g = 2 * c * O.sensitivity_dc2(u, lam, d.mesh, d.ops)
    for k in rng.choice(d.mesh.N, 6, replace=False):
        h = 37.0  # any h: the form is quadratic in c_k
        cp, cm = c.copy(), c.copy()
        cp[k] += h
        cm[k] -= h
        fd = (lam @ (O.stiffness_csr(cp, d.mesh, d.ops) @ u) - lam @ (O.stiffness_csr(cm, d.mesh, d.ops) @ u)) / (2 * h)
        assert abs(fd - g[k]) < 1e-8 * max(abs(fd), np.abs(g).max() * 1e-3)


def _fd_check(d, c, yobs, mode, nodes, h=0.05, noise=1e-3):
    out = O.forward_adjoint(d, c, yobs, noise, mode=mode)
    errs = []
    for k in nodes:
        cp, cm = c.copy(), c.copy()
        cp[k] += h
        cm[k] -= h
        lp = O.misfit_loglik(O.forward(d, cp, history=None)[0], yobs, noise)[0]
        lm = O.misfit_loglik(O.forward(d, cm, history=None)[0], yobs, noise)[0]
        fd = (lp - lm) / (2 * h)
        errs.append(abs(out["dloglik_dc"][k] - fd) / np.abs(out["dloglik_dc"]).max())
    return max(errs)


def test_exact_adjoint_matches_finite_differences_with_sponge():
    p, d = _prob()
    rng = np.random.default_rng(2)
    c = 2000 + 400 * rng.random(d.mesh.N)
    yobs, _ = O.forward(d, c + 30 * rng.standard_normal(d.mesh.N), history=None)
    assert _fd_check(d, c, yobs, "exact", rng.choice(d.mesh.N, 6, replace=False)) < 1e-6


def test_reference_adjoint_equals_exact_without_sponge_and_matches_fd():
    p, d = _prob(sides=())
    rng = np.random.default_rng(3)
    c = 2000 + 400 * rng.random(d.mesh.N)
    yobs, _ = O.forward(d, c + 30 * rng.standard_normal(d.mesh.N), history=None)
    a = O.forward_adjoint(d, c, yobs, 1e-3, mode="reference")["dloglik_dc"]
    b = O.forward_adjoint(d, c, yobs, 1e-3, mode="exact")["dloglik_dc"]
    np.testing.assert_array_equal(a, b)
    assert _fd_check(d, c, yobs, "reference", rng.choice(d.mesh.N, 4, replace=False)) < 1e-6


def test_reference_adjoint_differs_with_sponge():
    """The reference recurrence ignores the sponge (core.py:1252-1263): with a
    sponge it is not the gradient of the damped forward."""
    p, d = _prob()
    rng = np.random.default_rng(4)
    c = 2000 + 400 * rng.random(d.mesh.N)
    yobs, _ = O.forward(d, c + 30 * rng.standard_normal(d.mesh.N), history=None)
    a = O.forward_adjoint(d, c, yobs, 1e-3, mode="reference")["dloglik_dc"]
    b = O.forward_adjoint(d, c, yobs, 1e-3, mode="exact")["dloglik_dc"]
    assert np.linalg.norm(a - b) > 1e-3 * np.linalg.norm(b)


@pytest.mark.parametrize("every", [1, 7, 40, 159, 160, 500])
def test_checkpoint_recompute_bit_identical(every):
    p, d = _prob()
    rng = np.random.default_rng(5)
    c = 2000 + 400 * rng.random(d.mesh.N)
    yobs, _ = O.forward(d, c * 1.01, history=None)
    a = O.forward_adjoint(d, c, yobs, 1e-3, mode="exact")
    b = O.forward_adjoint(d, c, yobs, 1e-3, mode="exact", checkpoint_every=every)
    np.testing.assert_array_equal(a["dloglik_dc"], b["dloglik_dc"])


def test_record_conventions_shift_by_one_sample():
    """run() samples R u^n (pre), run_forward_and_adjoint R u^{n+1} (post)."""
    p, d = _prob()
    c = np.full(d.mesh.N, 2200.0)
    pre, _ = O.forward(d, c, record="pre", history=None)
    post, _ = O.forward(d, c, record="post", history=None)
    np.testing.assert_array_equal(pre[1:], post[:-1])
    assert np.all(pre[0] == 0)


def test_misfit_definition():
    s = np.arange(12.0).reshape(4, 3)
    y = s + 0.5
    ll, r, s2 = O.misfit_loglik(s, y, 0.1)
    assert abs(ll + 0.5 * 12 * 0.25 / (0.01 + 1e-24)) < 1e-12
    np.testing.assert_array_equal(r, -0.5 * np.ones((4, 3)))
