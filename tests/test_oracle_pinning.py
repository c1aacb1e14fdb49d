"""Properties the mathematics fixes for the oracle pieces that are otherwise
only restated from the reference: the forward's sponge placement (D applied
twice, Bspline_Coupling/SEM/sem_waveform/core.py:1207-1210, boundary.py:85-106),
the reference adjoint's missing sponge (core.py:1252-1263), the cubic-ramp
sponge profile (boundary.py:26-52) and the wavelet's DC removal
(sources.py:7-21).  A plausible misreading of any of them fails one of these.
"""
import math

import numpy as np
import pytest

import oracle as O

SIDES = ("xmin", "xmax", "zmax")


def _problem(nt, src, rec, dt=1e-4, gate=0.3, sides=SIDES):
    prob = O.Problem(0.0, 300.0, 0.0, 300.0, 6, 6, 4, dt, nt, src, 25.0, np.asarray(rec, dtype=np.float64),
                     gate_time=gate, sponge_width=150.0, sponge_cmax=3000.0, sponge_sides=sides)
    return O.discretize(prob)


def test_forward_sponge_node_has_double_root():
    """c = 0 gives K = 0, so after the source is gated off the source node
    follows u^{n+1} = D (2 u^n - D u^{n-1}).  D applied to prev, curr and next
    before the rotation (the reference's in-place damping) makes D a double
    root: u^n = (a + b n) D^n, i.e. q_n = u^n / D^n is linear in n.  Damping
    only the new state, u^{n+1} = D (2 u^n - u^{n-1}), has the complex roots
    D +- i sqrt(D - D^2) and fails this."""
    disc = _problem(260, (20.0, 150.0), [[100.0, 0.0]], gate=0.002)
    src = disc.src_node
    D = disc.decay[src]
    assert 0.5 < D < 0.99, D
    trace = []
    O.forward(disc, np.zeros(disc.mesh.N), history=None, after_step=lambda it, u: trace.append(u[src]))
    u = np.concatenate([[0.0], np.asarray(trace)])          # u^0 .. u^nt
    ng = int(math.ceil(0.002 / disc.prob.dt))               # first step without source
    n = np.arange(ng, u.size)
    q = u[ng:] / D ** n
    d2 = q[2:] - 2 * q[1:-1] + q[:-2]
    assert np.abs(q).max() > 0
    assert np.abs(d2).max() <= 1e-10 * np.abs(q).max()
    # the once-damped recurrence from the same two starting values departs from it
    v = u[:ng + 2].tolist()
    for k in range(ng + 1, u.size - 1):
        v.append(D * (2 * v[k] - v[k - 1]))
    assert np.abs(np.asarray(v[ng:]) / D ** n - q).max() > 1e-3 * np.abs(q).max()


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_adjoint_sponge_growth_at_zero_stiffness(mode):
    """c = 0 (K = 0) and a residual at one time index k0 on a receiver inside
    the sponge.  Every node then evolves alone: the reference adjoint
    lam_{k-1} = 2 lam_k - lam_{k+1} grows linearly, lam_{k0-j} = j lam_{k0-1},
    with no decay in the sponge; the exact adjoint lam_{k-1} = D (2 lam_k -
    D lam_{k+1}) gives lam_{k0-j} = j D^{j-1} lam_{k0-1}."""
    nt, k0 = 120, 100
    disc = _problem(nt, (150.0, 100.0), [[37.5, 150.0]])
    N = disc.mesh.N
    _, hist = O.forward(disc, np.zeros(N), record="post", history="full")
    res = np.zeros((nt, 1))
    res[k0, 0] = 1.0
    lams = {}
    O.adjoint(disc, np.zeros(N), hist, res, 1.0, mode=mode, after_step=lambda it, lam: lams.setdefault(it, lam))
    first = lams[k0]                                          # lam_{k0-1}
    assert np.abs(first).max() > 0
    hit = np.abs(first) > 1e-6 * np.abs(first).max()
    assert disc.decay[hit].max() < 0.999                      # the injection lands in the sponge
    for j in range(2, k0 + 1):
        lam = lams[k0 - j + 1]                                # lam_{k0-j}
        want = j * first if mode == "reference" else j * disc.decay ** (j - 1) * first
        assert np.abs(lam - want).max() <= 1e-12 * np.abs(want).max(), j
    if mode == "reference":
        assert np.abs(lams[1]).max() > 90 * np.abs(first).max()   # grown, not damped


@pytest.mark.parametrize("side", ["xmin", "xmax", "zmin", "zmax"])
def test_sponge_ramp_integral_gives_the_reflection_target(side):
    """The cubic ramp sigma = d0 (s/delta)^3, d0 = -(m+1) c_max ln R / (2 delta),
    integrates across the layer to -c_max ln R / 2, so the round-trip
    amplitude exp(-2 int sigma ds / c_max) equals the target R.  The ramp is a
    cubic on each element of a layer that is a whole number of elements thick,
    so GLL quadrature (exact to degree 7) integrates it exactly."""
    R, cmax, dt, width = 1e-6, 2750.0, 1e-3, 300.0
    m = O.StructuredMesh(0.0, 1000.0, 0.0, 1200.0, 10, 12, 4)
    D = O.sponge_decay(m, width, dt, cmax, R, (side,)).reshape(m.NZ, m.NX)
    sigma = -np.log(D) / dt
    w = O.gll_weights(O.gll_points(4))
    if side in ("xmin", "xmax"):
        prof, h, n_el = sigma[m.NZ // 2], m.dx, m.nex
    else:
        prof, h, n_el = sigma[:, m.NX // 2], m.dz, m.nez
    integral = sum(np.dot(w, prof[4 * e:4 * e + 5]) * h / 2 for e in range(n_el))
    assert abs(math.exp(-2.0 * integral / cmax) / R - 1.0) < 1e-10


def test_dc_removed_wavelet_has_zero_trapezoid_integral():
    """Without the gate, subtracting trapz(w) / (t_end - t_0) from every
    sample leaves a trace whose trapezoid integral is zero; the same holds for
    the product's wavelet (mun_amd.solver.source_wavelet), which must equal the
    oracle's."""
    from mun_amd.solver import source_wavelet
    for nt, dt, f0 in ((1000, 5e-4, 10.0), (777, 2.5e-4, 15.0), (4000, 4e-4, 10.0)):
        t = np.arange(nt) * dt
        raw = 3.0 * O.ricker(t, f0, 0.1)
        tol = 1e-13 * np.abs(raw).max() * t[-1]
        assert abs(np.trapezoid(raw, t)) > 1e3 * tol   # the truncated Ricker has a DC term to remove
        for w in (O.source_wavelet(nt, dt, f0, gate_time=1e9, amp=3.0),
                  source_wavelet(nt, dt, f0, gate_time=1e9, amp=3.0)):
            assert abs(np.trapezoid(w, t)) <= tol
        np.testing.assert_allclose(source_wavelet(nt, dt, f0), O.source_wavelet(nt, dt, f0), rtol=0, atol=1e-15)
