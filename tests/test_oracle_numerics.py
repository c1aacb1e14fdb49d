"""Pins the oracle's L0/L1 pieces against what the mathematics fixes.

Each check is chosen so that a plausible slip (wrong node, transposed D,
missing Jacobian, wrong sponge exponent, dropped DC term) fails it.
"""
import math

import numpy as np
import pytest

import oracle as O


def test_gll_p4_closed_form():
    x = O.gll_points(4)
    s = math.sqrt(3.0 / 7.0)
    np.testing.assert_allclose(x, [-1.0, -s, 0.0, s, 1.0], atol=1e-15)
    w = O.gll_weights(x)
    np.testing.assert_allclose(w, [0.1, 49 / 90, 32 / 45, 49 / 90, 0.1], rtol=1e-14)


@pytest.mark.parametrize("p", [2, 3, 4, 5, 6, 7])
def test_gll_quadrature_exact_to_degree_2p_minus_1(p):
    x = O.gll_points(p)
    w = O.gll_weights(x)
    for k in range(2 * p):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) < 1e-13
    # one degree higher is no longer exact (GLL is exact to 2p-1 only)
    k = 2 * p
    assert abs(np.sum(w * x ** k) - 2.0 / (k + 1)) > 1e-6


@pytest.mark.parametrize("p", [3, 4, 5])
def test_differentiation_matrix_exact_on_polynomials(p):
    x = O.gll_points(p)
    D = O.differentiation_matrix(x)
    for k in range(p + 1):
        np.testing.assert_allclose(D @ x ** k, k * x ** max(k - 1, 0) if k else 0 * x, atol=1e-12)
    # rows differentiate, they are not transposed: D[i, j] = l_j'(x_i)
    h = 1e-6
    for j in range(p + 1):
        d_num = (O.lagrange_basis(x + h, j, x, 0.0) - O.lagrange_basis(x - h, j, x, 0.0)) / (2 * h)
        np.testing.assert_allclose(D[:, j], d_num, atol=1e-6)


def test_lagrange_reference_epsilon_quirk():
    x = O.gll_points(4)
    for i in range(5):
        li = O.lagrange_basis(x, i, x, 1e-12)
        off = np.delete(li, i)
        assert np.all(off == 0.0)            # a factor (x_j - x_j) is exactly zero
        assert abs(1.0 - li[i]) < 1e-10      # 1 + O(1e-11) on the diagonal
        np.testing.assert_array_equal(O.lagrange_basis(x, i, x, 0.0), np.eye(5)[i])
    # the end nodes' factors do not cancel pairwise: the quirk is visible there
    assert O.lagrange_basis(x[0], 0, x, 1e-12) != 1.0


def test_sem_operators_structure():
    dx, dz = 30.0, 20.0
    Gx, Gz, S, w2D = O.sem_operators(dx, dz, 4, lagrange_eps=0.0)
    x = O.gll_points(4)
    X, Z = np.meshgrid(x, x, indexing="xy")      # local order j*5 + i, i along x
    f = (X ** 3 * Z + 2 * Z ** 2).reshape(-1)
    np.testing.assert_allclose(Gx @ f, (2 / dx) * (3 * X ** 2 * Z).reshape(-1), atol=1e-12)
    np.testing.assert_allclose(Gz @ f, (2 / dz) * (X ** 3 + 4 * Z).reshape(-1), atol=1e-12)
    np.testing.assert_allclose(S, np.eye(25), atol=0)
    assert abs(w2D.sum() - dx * dz) < 1e-10   # integrates 1 over the element


def test_mass_closed_form_and_area():
    m = O.StructuredMesh(-100.0, 300.0, 0.0, 250.0, 8, 5, 4)
    M = O.mass_lumped(m)
    assert abs(M.sum() - 400.0 * 250.0) < 1e-8
    w = O.gll_weights(O.gll_points(4))
    Wx = np.zeros(m.NX)
    for e in range(m.nex):
        Wx[4 * e:4 * e + 5] += w
    Wz = np.zeros(m.NZ)
    for e in range(m.nez):
        Wz[4 * e:4 * e + 5] += w
    J = m.dx * m.dz / 4
    np.testing.assert_allclose(M, J * np.outer(Wz, Wx).reshape(-1), rtol=1e-14)


def _dense_K_bruteforce(c, coords_elems, Gx, Gz, S, w2D, N):
    """Literal element loop of Bspline_Coupling/SEM/sem_waveform/operators.py:86-122."""
    K = np.zeros((N, N))
    for en in coords_elems:
        c2q = S @ (c[en] ** 2)
        wc = w2D * c2q
        Ke = (Gx * wc[:, None]).T @ Gx + (Gz * wc[:, None]).T @ Gz
        for a, ga in enumerate(en):
            for b, gb in enumerate(en):
                K[ga, gb] += Ke[a, b]
    return K


@pytest.mark.parametrize("nex,nez", [(1, 1), (2, 2), (3, 2)])
def test_stiffness_vs_bruteforce_and_invariants(nex, nez):
    rng = np.random.default_rng(nex + 10 * nez)
    m = O.StructuredMesh(0.0, 40.0 * nex, 0.0, 30.0 * nez, nex, nez, 4)
    ops = O.sem_operators(m.dx, m.dz, 4)
    c = 1500 + 1500 * rng.random(m.N)
    K = O.stiffness_csr(c, m, ops).toarray()
    Kb = _dense_K_bruteforce(c, m.conn, *ops, m.N)
    np.testing.assert_allclose(K, Kb, rtol=1e-12, atol=1e-9 * np.abs(Kb).max())
    np.testing.assert_allclose(K, K.T, atol=1e-9 * np.abs(K).max())
    assert np.abs(K @ np.ones(m.N)).max() < 1e-9 * np.abs(K).max()
    # positive semi-definite with a one-dimensional null space (constants)
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert ev[0] > -1e-8 * ev[-1] and ev[1] > 1e-8 * ev[-1]


def test_stiffness_energy_of_linear_field():
    """u = a x + b z with constant c: u^T K u = c^2 (a^2 + b^2) * area, and
    K u vanishes on rows away from the boundary."""
    m = O.StructuredMesh(0.0, 200.0, 0.0, 150.0, 5, 4, 4)
    ops = O.sem_operators(m.dx, m.dz, 4, lagrange_eps=0.0)
    c = np.full(m.N, 2000.0)
    K = O.stiffness_csr(c, m, ops)
    X = m.coords
    a, b = 0.3, -0.7
    u = a * X[:, 0] + b * X[:, 1]
    assert abs(u @ (K @ u) / (2000.0 ** 2 * (a * a + b * b) * 200 * 150) - 1) < 1e-12
    Ku = (K @ u).reshape(m.NZ, m.NX)
    assert np.abs(Ku[1:-1, 1:-1]).max() < 1e-7 * np.abs(Ku).max()


def test_sponge_formula_and_sides():
    m = O.StructuredMesh(0.0, 1000.0, 0.0, 800.0, 10, 8, 4)
    dt, cmax, d = 1e-3, 3000.0, 200.0
    D = O.sponge_decay(m, d, dt, cmax, 1e-8, ("xmin", "zmax"))
    d0 = 4 * cmax * math.log(1e8) / (2 * d)
    X = m.coords
    sx = np.where(X[:, 0] < d, d0 * ((d - X[:, 0]) / d) ** 3, 0.0)
    sz = np.where(X[:, 1] > 800 - d, d0 * ((X[:, 1] - 600) / d) ** 3, 0.0)
    np.testing.assert_allclose(D, np.exp(-(sx + sz) * dt), rtol=1e-14)
    assert D.reshape(m.NZ, m.NX)[0, -1] == 1.0          # undamped corner (xmax, zmin)
    assert abs(D.reshape(m.NZ, m.NX)[-1, 0] - math.exp(-2 * d0 * dt)) < 1e-14


def test_source_wavelet_dc_and_gate():
    nt, dt = 1000, 5e-4
    w = O.source_wavelet(nt, dt, f0=10.0, amp=2.0)
    t = np.arange(nt) * dt
    raw = 2.0 * O.ricker(t, 10.0, 0.1)
    dc = np.trapezoid(raw, t) / (t[-1] - t[0])
    np.testing.assert_allclose(w[t < 0.3], raw[t < 0.3] - dc, rtol=0, atol=1e-15)
    assert np.all(w[t >= 0.3] == 0.0)
    assert abs(O.ricker(0.1, 10.0, 0.1) - 1.0) < 1e-15


def test_receiver_interpolation_exact_for_degree4():
    m = O.StructuredMesh(0.0, 500.0, 0.0, 400.0, 5, 4, 4)
    rng = np.random.default_rng(0)
    pos = np.column_stack([rng.uniform(0, 500, 30), rng.uniform(0, 400, 30)])
    pos[0] = [500.0, 400.0]      # domain corner (clamped element)
    pos[1] = [100.0, 100.0]      # element corner -> right/upper element
    R = O.receiver_matrix(pos, m, lagrange_eps=0.0)
    X = m.coords
    for f in (lambda x, z: np.ones_like(x), lambda x, z: x ** 4 - 3 * z ** 2 * x,
              lambda x, z: (x * z) ** 2 + z ** 4):
        np.testing.assert_allclose(R @ f(X[:, 0], X[:, 1]), f(pos[:, 0], pos[:, 1]), rtol=1e-9,
                                   atol=1e-9 * np.abs(f(X[:, 0], X[:, 1])).max())
    # the element choice follows int((x - xmin)/dx): (100, 100) sits in element (1, 1)
    cols = R[1].indices
    e11 = m.conn[1 * m.nex + 1]
    assert set(cols) <= set(e11)


def test_reference_numbering_permutation():
    nex, nez = 3, 2
    coords_ref, conn_ref = O.create_global_mesh_reference(0.0, 300.0, 0.0, 200.0, nex, nez, 4)
    m = O.StructuredMesh(0.0, 300.0, 0.0, 200.0, nex, nez, 4)
    perm = O.reference_node_order(nex, nez, 4)
    np.testing.assert_allclose(coords_ref, m.coords[perm], atol=1e-12)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    np.testing.assert_array_equal(inv[m.conn], conn_ref)


def test_closest_node_matches_bruteforce():
    m = O.StructuredMesh(0.0, 500.0, 0.0, 400.0, 5, 4, 4)
    rng = np.random.default_rng(1)
    X = m.coords
    for _ in range(50):
        x, z = rng.uniform(-10, 510), rng.uniform(-10, 410)
        k = int(np.argmin(np.hypot(X[:, 0] - x, X[:, 1] - z)))
        assert m.find_closest_node(x, z) == k


def test_tensor_bspline_partition_of_unity_and_linear_reproduction():
    m = O.StructuredMesh(0.0, 800.0, 0.0, 600.0, 8, 6, 4)
    c1, Bz, Bx = O.tensor_bspline_velocity(m, np.ones((7, 9)))
    np.testing.assert_allclose(c1, 1.0, atol=1e-14)
    gz = O.greville(0.0, 600.0, 7)
    gx = O.greville(0.0, 800.0, 9)
    coef = 1500.0 + 2.0 * gz[:, None] + 0.5 * gx[None, :]
    c, _, _ = O.tensor_bspline_velocity(m, coef)
    X = m.coords
    np.testing.assert_allclose(c, 1500 + 2.0 * X[:, 1] + 0.5 * X[:, 0], rtol=1e-13)
    assert np.all(Bx >= 0) and np.all((Bx > 0).sum(1) <= 4)


def test_layered_background_searchsorted_right():
    v = O.layered_background(np.array([-700.0, -600.0, -1.0, 0.0, 299.0, 300.0]), [-600.0, 0.0, 300.0],
                             [1500.0, 2000.0, 2500.0, 3000.0])
    np.testing.assert_array_equal(v, [1500, 2000, 2000, 2500, 2500, 3000])
