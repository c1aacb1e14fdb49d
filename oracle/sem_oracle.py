"""FP64 CPU restatement of the reference SEM forward/adjoint solver.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  It is the checker for
the HIP path and the timed CPU baseline of ``bench.py``; nothing in the product
package imports it.

Every function cites the reference file and lines it restates.  Short names:

    S/   = Bspline_Coupling/SEM/sem_waveform/
    C2D/ = CO2_plume/2D_model/inversion/sem_waveform/
    CLY/ = CO2_plume/layered_model/inversion/sem_waveform/

Structure follows the reference: an assembled global CSR stiffness matrix, a
sparse matrix-vector product per step, the reference's leapfrog / sponge /
source recurrences, and its element-by-element gradient correlation with dense
(p+1)^2 x (p+1)^2 operators.  Two deliberate, documented differences:

* Nodes are numbered lexicographically on the structured GLL grid
  (``g = jz * NX + ix``) instead of the reference's element-blocked
  first-appearance numbering (S/mesh.py:4-43).  ``reference_node_order`` gives
  the permutation between the two, so node-indexed outputs can be compared in
  either numbering.
* ``dt`` is used as given (the reference's CFL shrink, S/utils.py:96-111, is
  available as ``reference_cfl_dt`` but not applied automatically).

Pinning (what checks each function; tests live in ``tests/test_oracle_*.py``):
    gll_points / gll_weights / differentiation_matrix: closed forms for p=4,
        exactness of GLL quadrature, D exact on polynomials.
    sem_operators / stiffness_csr / mass_lumped: K 1 = 0, symmetry, brute
        force dense assembly on 1x1 and 2x2 meshes, energy identity, quadrature
        of a known integral, K u = 0 for linear u (interior rows).
    sponge_decay / source_wavelet / receiver_matrix: formulas, polynomial
        exactness of the receiver interpolation (degree <= p).
    forward / adjoint / sensitivity: discrete energy conservation (no sponge,
        no source), exact FD of lambda^T K(c) u (quadratic in c), central FD
        of the full misfit in the sponge-aware ('exact') adjoint mode, dot
        product test of the adjoint recurrence.
    tensor B-spline map: partition of unity, linear reproduction, FD.
    closed_curve_velocity / closed_curve_projection (clamped 6-point curve):
        unsigned distance against brute force on a dense, independently
        evaluated curve (Cox-de Boor), sign against its winding number, u*
        on a nearest point, n_out = grad_X sd by FD, mean velocity on the
        curve, projection = -d/dctrl (a^T v) by central differences.
    periodic_curve_velocity / periodic_curve_projection: knot points of a
        regular K-gon at radius R (2 + cos 2 pi/K) / 3 with sd, u* and normal
        along the ray through them; projection = -d/dctrl (a^T v) by FD.
    grid_background / increasing_grid: bilinear fields reproduced exactly,
        nearest-neighbour ties, axis flips carry the values.
    window_mask / window_from_times / forward_adjoint windows: rounding and
        clamps, central FD of the windowed misfit (exact mode), the reference
        mode's truncation before the window start.
    forward_snapshots: each snapshot equals the state after its step.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import scipy.sparse as sp

__all__ = [
    "gll_points", "gll_weights", "lagrange_basis", "barycentric_weights",
    "differentiation_matrix", "sem_operators", "reference_cfl_dt",
    "StructuredMesh", "create_global_mesh_reference", "reference_node_order",
    "mass_lumped", "element_stiffness", "stiffness_csr", "sponge_decay",
    "ricker", "source_wavelet", "circular_receivers", "receiver_matrix",
    "Problem", "Discretization", "discretize", "forward", "adjoint",
    "sensitivity_dc2", "misfit_loglik", "forward_adjoint",
    "bspline_basis_matrix", "tensor_bspline_velocity", "layered_background", "greville", "clamped_knots",
    "closed_curve_velocity", "closed_curve_projection",
    "forward_snapshots", "roi_box", "window_mask", "window_from_times",
    "grid_background", "increasing_grid", "periodic_curve_velocity", "periodic_curve_projection",
]


# ---------------------------------------------------------------------------
# L0 numerics  (S/utils.py)
# ---------------------------------------------------------------------------
def _legendre(n: int, x):
    """P_n(x) and P_n'(x) by the three-term recurrence."""
    x = np.asarray(x, dtype=np.float64)
    p0 = np.ones_like(x)
    if n == 0:
        return p0, np.zeros_like(x)
    p1 = x.copy()
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    # derivative from (1-x^2) P_n' = n (P_{n-1} - x P_n)   (interior points only)
    with np.errstate(divide="ignore", invalid="ignore"):
        dp = n * (p0 - x * p1) / (1.0 - x * x)
    return p1, dp


def gll_points(p: int) -> np.ndarray:
    """GLL nodes: endpoints +-1 and the roots of P_p'.

    Restates S/utils.py:5-21: Chebyshev-Gauss-Lobatto initial guesses
    cos(pi k / p), sorted, each interior one refined by Newton on P_p'.
    """
    if p == 0:
        return np.array([0.0])
    if p == 1:
        return np.array([-1.0, 1.0])
    cheb = np.polynomial.legendre.Legendre.basis(p).deriv()
    dcheb = cheb.deriv()
    x = np.sort(np.cos(np.pi * np.arange(p + 1) / p))
    for i in range(1, p):
        xi = x[i]
        for _ in range(100):
            step = cheb(xi) / dcheb(xi)
            xi -= step
            if abs(step) < 1e-16:
                break
        x[i] = xi
    x[0], x[-1] = -1.0, 1.0
    return x


def gll_weights(nodes: np.ndarray) -> np.ndarray:
    """w_i = 2 / (p (p+1) P_p(x_i)^2)  (S/utils.py:23-29)."""
    p = len(nodes) - 1
    pn, _ = _legendre(p, nodes)
    return 2.0 / (p * (p + 1) * pn ** 2)


def lagrange_basis(x, i: int, nodes: np.ndarray, eps: float = 1e-12):
    """i-th Lagrange polynomial at x, with the reference's +eps in every
    denominator (S/utils.py:31-37; quirk Q7).  ``eps=0`` gives the exact one."""
    x = np.asarray(x, dtype=np.float64)
    out = np.ones_like(x)
    for j in range(len(nodes)):
        if j != i:
            out = out * ((x - nodes[j]) / (nodes[i] - nodes[j] + eps))
    return out


def barycentric_weights(x: np.ndarray) -> np.ndarray:
    """S/utils.py:39-47."""
    n = len(x)
    w = np.ones(n)
    for j in range(n):
        for k in range(n):
            if k != j:
                w[j] /= (x[j] - x[k])
    return w


def differentiation_matrix(x: np.ndarray) -> np.ndarray:
    """D[i,j] = l_j'(x_i) by barycentric weights, negative-sum diagonal
    (S/utils.py:49-61)."""
    w = barycentric_weights(x)
    n = len(x)
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                D[i, j] = w[j] / (w[i] * (x[i] - x[j]))
        D[i, i] = -np.sum(D[i, :])
    return D


def reference_cfl_dt(c_max: float, p: int, dx: float, dz: float) -> float:
    """The reference's empirical SEM limit (S/utils.py:64-93; quirk Q1):
    0.35 * (min(dx,dz)/(p+1)) / (c_max (p+1) sqrt(2))."""
    h = min(dx, dz) / (p + 1)
    return 0.35 * h / (c_max * (p + 1) * math.sqrt(2.0))


def sem_operators(dx: float, dz: float, p: int, lagrange_eps: float = 1e-12):
    """Gx = (2/dx) kron(L, D), Gz = (2/dz) kron(D, L), S = kron(L, L),
    w2D = (w x w) dx dz / 4   (S/operators.py:6-42).

    Local node order within an element is ``j*(p+1) + i`` with i along x."""
    nodes = gll_points(p)
    w = gll_weights(nodes)
    D = differentiation_matrix(nodes)
    ng = p + 1
    L = np.zeros((ng, ng))
    for j in range(ng):
        for n in range(ng):
            L[j, n] = lagrange_basis(nodes[j], n, nodes, lagrange_eps)
    J = dx * dz * 0.25
    Gx = (2.0 / dx) * np.kron(L, D)
    Gz = (2.0 / dz) * np.kron(D, L)
    S = np.kron(L, L)
    w2D = ((w[:, None] * w[None, :]) * J).reshape(-1)
    return Gx, Gz, S, w2D


# ---------------------------------------------------------------------------
# L1 mesh  (S/mesh.py)
# ---------------------------------------------------------------------------
@dataclass
class StructuredMesh:
    """Structured quad mesh of order-p GLL nodes, lexicographic numbering.

    Global node (jz, ix) has id ``jz * NX + ix``; element (ez, ex) has id
    ``ez * nex + ex`` and local node (j, i) maps to (p*ez + j, p*ex + i)."""
    xmin: float
    xmax: float
    zmin: float
    zmax: float
    nex: int
    nez: int
    p: int
    x1d: np.ndarray = field(init=False)
    z1d: np.ndarray = field(init=False)
    conn: np.ndarray = field(init=False)

    def __post_init__(self):
        p = self.p
        self.NX = self.nex * p + 1
        self.NZ = self.nez * p + 1
        self.N = self.NX * self.NZ
        self.dx = (self.xmax - self.xmin) / self.nex
        self.dz = (self.zmax - self.zmin) / self.nez
        nodes = gll_points(p)
        self.x1d = self._axis(self.xmin, self.xmax, self.nex, nodes)
        self.z1d = self._axis(self.zmin, self.zmax, self.nez, nodes)
        ez, ex, j, i = np.meshgrid(np.arange(self.nez), np.arange(self.nex),
                                   np.arange(p + 1), np.arange(p + 1), indexing="ij")
        self.conn = ((p * ez + j) * self.NX + (p * ex + i)).reshape(self.nez * self.nex, (p + 1) ** 2)

    @staticmethod
    def _axis(a, b, ne, nodes):
        # same arithmetic as S/mesh.py:22-30: x = 0.5((1-xi) x0 + (1+xi) x1)
        corners = np.linspace(a, b, ne + 1)
        p = len(nodes) - 1
        out = np.empty(ne * p + 1)
        for e in range(ne):
            x0, x1 = corners[e], corners[e + 1]
            for i in range(p + 1):
                if e > 0 and i == 0:
                    continue  # shared node already set by the previous element
                out[e * p + i] = 0.5 * ((1 - nodes[i]) * x0 + (1 + nodes[i]) * x1)
        return out

    @property
    def coords(self) -> np.ndarray:
        X, Z = np.meshgrid(self.x1d, self.z1d, indexing="xy")
        return np.stack([X.reshape(-1), Z.reshape(-1)], axis=1)

    def find_closest_node(self, x: float, z: float) -> int:
        """Nearest GLL node (S/mesh.py:45-54).  On a tensor grid the squared
        distance separates, so the argmin is taken per axis."""
        ix = int(np.argmin(np.abs(self.x1d - x)))
        jz = int(np.argmin(np.abs(self.z1d - z)))
        return jz * self.NX + ix


def create_global_mesh_reference(xmin, xmax, zmin, zmax, nex, nez, p):
    """Element-blocked, rounded-coordinate de-duplicated mesh exactly as
    S/mesh.py:4-43 builds it.  Python loop: use on small meshes only."""
    nodes = gll_points(p)
    cx = np.linspace(xmin, xmax, nex + 1)
    cz = np.linspace(zmin, zmax, nez + 1)
    coords, conn, node_map = [], [], {}
    for ez in range(nez):
        for ex in range(nex):
            x0, x1, z0, z1 = cx[ex], cx[ex + 1], cz[ez], cz[ez + 1]
            en = []
            for j in range(p + 1):
                for i in range(p + 1):
                    x = 0.5 * ((1 - nodes[i]) * x0 + (1 + nodes[i]) * x1)
                    z = 0.5 * ((1 - nodes[j]) * z0 + (1 + nodes[j]) * z1)
                    key = (round(x, 6), round(z, 6))
                    if key not in node_map:
                        node_map[key] = len(coords)
                        coords.append((x, z))
                    en.append(node_map[key])
            conn.append(en)
    return np.asarray(coords), np.asarray(conn, dtype=np.int64)


def reference_node_order(nex: int, nez: int, p: int) -> np.ndarray:
    """perm[k] = lexicographic id of the reference's node k (S/mesh.py:19-38
    visit order: ez, ex, j, i; first appearance wins)."""
    NX = nex * p + 1
    seen = np.full((nez * p + 1) * NX, -1, dtype=np.int64)
    order = np.empty(seen.size, dtype=np.int64)
    k = 0
    for ez in range(nez):
        for ex in range(nex):
            for j in range(p + 1):
                row = (p * ez + j) * NX + p * ex
                for i in range(p + 1):
                    g = row + i
                    if seen[g] < 0:
                        seen[g] = k
                        order[k] = g
                        k += 1
    return order


# ---------------------------------------------------------------------------
# L1 operators  (S/operators.py)
# ---------------------------------------------------------------------------
def mass_lumped(mesh: StructuredMesh) -> np.ndarray:
    """M_i = sum_e w_a w_b dx dz / 4   (S/operators.py:65-84)."""
    nodes = gll_points(mesh.p)
    w = gll_weights(nodes)
    J = (mesh.dx / 2.0) * (mesh.dz / 2.0)
    wl = (w[:, None] * w[None, :] * J).reshape(-1)
    M = np.zeros(mesh.N)
    np.add.at(M, mesh.conn.reshape(-1), np.tile(wl, mesh.conn.shape[0]))
    return M


def element_stiffness(c: np.ndarray, mesh: StructuredMesh, ops) -> np.ndarray:
    """Ke = Gx^T diag(w2D * S c_e^2) Gx + Gz^T diag(w2D * S c_e^2) Gz for every
    element, shape (nelem, ng, ng)   (S/operators.py:193-227)."""
    Gx, Gz, S, w2D = ops
    c2e = (c[mesh.conn] ** 2)                      # (nelem, ng)
    wc = w2D[None, :] * (c2e @ S.T)                # (nelem, ng)
    Ke = np.einsum("qa,eq,qb->eab", Gx, wc, Gx, optimize=True)
    Ke += np.einsum("qa,eq,qb->eab", Gz, wc, Gz, optimize=True)
    return Ke


def stiffness_csr(c: np.ndarray, mesh: StructuredMesh, ops) -> sp.csr_matrix:
    """Global assembled K (COO -> CSR, duplicates summed), S/operators.py:193-242."""
    Ke = element_stiffness(c, mesh, ops)
    ng = mesh.conn.shape[1]
    rows = np.repeat(mesh.conn, ng, axis=1).reshape(-1)
    cols = np.tile(mesh.conn, (1, ng)).reshape(-1)
    K = sp.coo_matrix((Ke.reshape(-1), (rows, cols)), shape=(mesh.N, mesh.N)).tocsr()
    K.sum_duplicates()
    return K


SIDES_ALL = ("xmin", "xmax", "zmin", "zmax")


def sponge_decay(mesh: StructuredMesh, width: float, dt: float, c_max: float,
                 R: float = 1e-8, sides: Sequence[str] = SIDES_ALL) -> np.ndarray:
    """Per-node decay exp(-sigma dt) of the cubic-ramp sponge
    (S/boundary.py:5-58).  sigma = sigma_x + sigma_z,
    d0 = -(m+1) c_max ln(R) / (2 delta), m = 3.  The reference damps all four
    sides; ``sides`` selects a subset (the solver's ``absorbing_sides``)."""
    m = 3
    d = float(width)
    d0 = -(m + 1) * float(c_max) * np.log(R) / (2.0 * d)
    X = mesh.coords
    x, z = X[:, 0], X[:, 1]
    sx = np.zeros(mesh.N)
    sz = np.zeros(mesh.N)
    if "xmin" in sides:
        msk = x < mesh.xmin + d
        sx[msk] = d0 * ((mesh.xmin + d - x[msk]) / d) ** m
    if "xmax" in sides:
        msk = x > mesh.xmax - d
        sx[msk] = d0 * ((x[msk] - (mesh.xmax - d)) / d) ** m
    if "zmin" in sides:
        msk = z < mesh.zmin + d
        sz[msk] = d0 * ((mesh.zmin + d - z[msk]) / d) ** m
    if "zmax" in sides:
        msk = z > mesh.zmax - d
        sz[msk] = d0 * ((z[msk] - (mesh.zmax - d)) / d) ** m
    return np.exp(-(sx + sz) * dt)


# ---------------------------------------------------------------------------
# source / receivers  (S/sources.py, S/receivers.py, C2D/core.py)
# ---------------------------------------------------------------------------
def ricker(t, f0=20.0, t0=0.1):
    """S/sources.py:3-5."""
    a = (np.pi * f0 * (t - t0)) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def source_wavelet(nt, dt, f0=20.0, t0=0.1, gate_time=0.3, amp=1.0):
    """Ricker * amp, minus trapz(w)/(t_end - t_0), zeroed for t >= gate_time
    (S/sources.py:7-21; quirk Q6)."""
    t = np.arange(nt) * dt
    w = ricker(t, f0, t0) * amp
    w = w - np.trapezoid(w, t) / (t[-1] - t[0])
    w[~(t < gate_time)] = 0.0
    return w.astype(np.float64)


def circular_receivers(src_x, src_z, radius, n):
    """S/receivers.py:14-18."""
    ang = np.linspace(0, 2 * np.pi, n, endpoint=False)
    return np.stack([src_x + radius * np.cos(ang), src_z + radius * np.sin(ang)], axis=1)


def receiver_matrix(positions: np.ndarray, mesh: StructuredMesh,
                    lagrange_eps: float = 1e-12) -> sp.csr_matrix:
    """R (nrec x N): tensor-product Lagrange weights in the element found by
    int((x - xmin)/dx) clamped (S/receivers.py:30-61; C2D/core.py:120-160)."""
    p = mesh.p
    nodes = gll_points(p)
    rows, cols, vals = [], [], []
    for r, (rx, rz) in enumerate(np.asarray(positions, dtype=np.float64)):
        ex = int((rx - mesh.xmin) / mesh.dx)
        ez = int((rz - mesh.zmin) / mesh.dz)
        ex = max(0, min(ex, mesh.nex - 1))
        ez = max(0, min(ez, mesh.nez - 1))
        x0 = mesh.xmin + ex * mesh.dx
        z0 = mesh.zmin + ez * mesh.dz
        xi = 2.0 * (rx - x0) / mesh.dx - 1.0
        eta = 2.0 * (rz - z0) / mesh.dz - 1.0
        en = mesh.conn[ez * mesh.nex + ex]
        for j in range(p + 1):
            lz = lagrange_basis(eta, j, nodes, lagrange_eps)
            for i in range(p + 1):
                rows.append(r)
                cols.append(en[j * (p + 1) + i])
                vals.append(lagrange_basis(xi, i, nodes, lagrange_eps) * lz)
    return sp.csr_matrix((vals, (rows, cols)), shape=(len(positions), mesh.N))


# ---------------------------------------------------------------------------
# problem description + discretisation
# ---------------------------------------------------------------------------
@dataclass
class Problem:
    """One shot on one mesh.  Field names follow the reference config dict
    (S/core.py:151-325)."""
    xmin: float
    xmax: float
    zmin: float
    zmax: float
    nelem_x: int
    nelem_z: int
    p: int
    dt: float
    nt: int
    src_xz: tuple
    f0: float
    rec_xz: np.ndarray
    amp: float = 1.0
    t0: float = 0.1
    gate_time: float = 0.3
    sponge_width: float = 200.0
    sponge_cmax: float = 3000.0
    sponge_R: float = 1e-8
    sponge_sides: tuple = SIDES_ALL
    lagrange_eps: float = 1e-12


@dataclass
class Discretization:
    prob: Problem
    mesh: StructuredMesh
    ops: tuple
    M: np.ndarray
    decay: np.ndarray
    R: sp.csr_matrix
    src_node: int
    wavelet: np.ndarray


def discretize(prob: Problem, wavelet: Optional[np.ndarray] = None) -> Discretization:
    mesh = StructuredMesh(prob.xmin, prob.xmax, prob.zmin, prob.zmax,
                          prob.nelem_x, prob.nelem_z, prob.p)
    ops = sem_operators(mesh.dx, mesh.dz, prob.p, prob.lagrange_eps)
    M = mass_lumped(mesh)
    decay = sponge_decay(mesh, prob.sponge_width, prob.dt, prob.sponge_cmax,
                         prob.sponge_R, prob.sponge_sides)
    R = receiver_matrix(prob.rec_xz, mesh, prob.lagrange_eps)
    src = mesh.find_closest_node(*prob.src_xz)
    if wavelet is None:
        wavelet = source_wavelet(prob.nt, prob.dt, prob.f0, prob.t0, prob.gate_time, prob.amp)
    return Discretization(prob, mesh, ops, M, decay, R, src, np.asarray(wavelet, dtype=np.float64))


# ---------------------------------------------------------------------------
# time stepping  (S/core.py:598-659 run(), S/core.py:1197-1215 forward+adjoint)
# ---------------------------------------------------------------------------
class _History:
    """u~^{n+1} for n = 0..nt-1, either stored in full or rebuilt from FP64
    checkpoints every ``every`` steps (bit-identical recompute)."""

    def __init__(self, nt, N, every):
        self.nt, self.N, self.every = nt, N, every
        self.full = np.zeros((nt, N)) if every is None else None
        self.ckpt = {}
        self._seg = None
        self._seg_id = None
        self._stepper = None

    def put(self, it, u_prev, u_curr):
        if self.full is not None:
            self.full[it] = u_curr
        elif (it + 1) % self.every == 0 or it == -1:
            self.ckpt[it] = (u_prev.copy(), u_curr.copy())

    def get(self, it):
        if self.full is not None:
            return self.full[it]
        seg = (it + 1) // self.every  # history index it+1 lives in segment seg
        start = seg * self.every - 1  # checkpoint taken after step `start`
        if self._seg_id != seg:
            up, uc = self.ckpt[start]
            up, uc = up.copy(), uc.copy()
            buf = {start: uc.copy()}
            for n in range(start + 1, min(start + self.every, self.nt)):
                up, uc = self._stepper(n, up, uc)
                buf[n] = uc.copy()
            self._seg, self._seg_id = buf, seg
        return self._seg[it]


def forward(disc: Discretization, c: np.ndarray, record: str = "post",
            history: Optional[str] = "full", checkpoint_every: Optional[int] = None,
            K: Optional[sp.csr_matrix] = None, after_step=None):
    """Damped leapfrog with point source.

    u~^{n+1} = D o [2u~^n - D o u~^{n-1} + dt^2 M^-1 (s_n M_src e_src - K u~^n)]

    The sponge multiplies prev, curr and next in place before the rotation, so
    D enters twice (S/core.py:1207-1210, S/boundary.py:85-106; quirk Q3).
    ``record='post'`` samples R u~^{n+1} after the rotation as
    run_forward_and_adjoint does (S/core.py:1211,1218-1219); ``'pre'`` samples
    R u~^n before the update as run() does (S/core.py:633-639; quirk Q4).
    ``after_step(it, u)`` (optional) sees u~^{it+1} after every step.
    Returns (seismograms (nt, nrec), history or None)."""
    p = disc.prob
    if K is None:
        K = stiffness_csr(c, disc.mesh, disc.ops)
    N = disc.mesh.N
    dt2 = p.dt * p.dt
    Minv = 1.0 / disc.M
    D = disc.decay
    s = disc.wavelet
    src = disc.src_node
    Msrc = disc.M[src]
    R = disc.R
    seis = np.zeros((p.nt, R.shape[0]))

    def step(it, u_prev, u_curr):
        rhs = -(K @ u_curr)
        rhs[src] += s[it] * Msrc
        u_next = 2.0 * u_curr - u_prev + dt2 * (Minv * rhs)
        return D * u_curr, D * u_next

    hist = None
    if history is not None:
        hist = _History(p.nt, N, checkpoint_every)
        hist._stepper = step
    u_prev = np.zeros(N)
    u_curr = np.zeros(N)
    if hist is not None and hist.full is None:
        hist.put(-1, u_prev, u_curr)
    for it in range(p.nt):
        if record == "pre":
            seis[it] = R @ u_curr
        u_prev, u_curr = step(it, u_prev, u_curr)
        if record == "post":
            seis[it] = R @ u_curr
        if hist is not None:
            hist.put(it, u_prev, u_curr)
        if after_step is not None:
            after_step(it, u_curr)
    return seis, hist


def forward_snapshots(disc: Discretization, c: np.ndarray, every: int, box, record: str = "pre"):
    """Wavefield snapshots of run() (S/core.py:318-325, 400-401, 649-651):
    after step it with it % every == 0 the state u~^{it+1} restricted to the
    ROI.  The ROI mask (x, z inside a rectangle) is a node box on the
    structured grid: box = (ix0, ix1, iz0, iz1), half-open.  Returns
    (seis, snapshots (nsnap, iz1-iz0, ix1-ix0))."""
    ix0, ix1, iz0, iz1 = box
    NX, NZ = disc.mesh.NX, disc.mesh.NZ
    snaps = []

    def grab(it, u):
        if it % every == 0:
            snaps.append(u.reshape(NZ, NX)[iz0:iz1, ix0:ix1].copy())

    seis, _ = forward(disc, c, record=record, history=None, after_step=grab)
    return seis, np.stack(snaps)


def roi_box(mesh: StructuredMesh, xmin: float, xmax: float, zmin: float, zmax: float):
    """Node box of the ROI mask (x >= xmin) & (x <= xmax) & (z >= zmin) &
    (z <= zmax) (S/core.py:318-325); None if empty."""
    ix = np.nonzero((mesh.x1d >= xmin) & (mesh.x1d <= xmax))[0]
    iz = np.nonzero((mesh.z1d >= zmin) & (mesh.z1d <= zmax))[0]
    if ix.size == 0 or iz.size == 0:
        return None
    return int(ix[0]), int(ix[-1]) + 1, int(iz[0]), int(iz[-1]) + 1


def sensitivity_dc2(u: np.ndarray, lam: np.ndarray, mesh: StructuredMesh, ops) -> np.ndarray:
    """d(lam^T K(c) u)/d(c^2) at the nodes: assemble_e S^T(w2D (Gx u.Gx lam +
    Gz u.Gz lam)).  The chain-rule factor 2c of S/core.py:49-85 is applied by
    the caller once the time sum is complete (the sum is linear)."""
    Gx, Gz, S, w2D = ops
    ul = u[mesh.conn]
    ll = lam[mesh.conn]
    gq = w2D[None, :] * ((ul @ Gx.T) * (ll @ Gx.T) + (ul @ Gz.T) * (ll @ Gz.T))
    loc = gq @ S  # S^T applied per element: (S^T g)_a = sum_b S[b,a] g_b
    out = np.zeros(mesh.N)
    np.add.at(out, mesh.conn.reshape(-1), loc.reshape(-1))
    return out


def misfit_loglik(seis: np.ndarray, y_obs: np.ndarray, noise_std: float):
    """r = d_syn - y_obs; loglik = -0.5 ||r||^2 / sigma^2, sigma^2 = std^2 +
    1e-24  (S/core.py:1085-1086, 1224-1230)."""
    sigma2 = float(noise_std) ** 2 + 1e-24
    r = seis - y_obs
    return -0.5 * float(np.sum(r * r)) / sigma2, r, sigma2


def adjoint(disc: Discretization, c: np.ndarray, hist: _History, residual: np.ndarray,
            sigma2: float, mode: str = "reference", K: Optional[sp.csr_matrix] = None, stop: int = 0,
            after_step=None):
    """Reverse-time adjoint and gradient correlation.

    reference mode (S/core.py:1247-1275; quirk Q5, no sponge):
        lam_{k-1} = 2 lam_k - lam_{k+1} + dt^2 M^-1 (R^T r_k / sigma^2 - K lam_k)
    exact mode (the discrete adjoint of ``forward`` with its sponge):
        lam_{k-1} = D o [2 lam_k - D o lam_{k+1} + dt^2 M^-1 (R^T r_k/sigma^2 - K lam_k)]
    both with lam_{nt-1} = lam_nt = 0 and
        dJ_dv = 2c o sum_k sens(u~^{k+1}, lam_k)      (S/core.py:1265-1270).
    The loop runs it = nt-1 .. stop; ``stop`` > 0 is the misfit-window
    truncation of C1S/core.py:1530-1560, whose adjoint covers the window only.
    ``after_step(it, lam)`` (optional) sees lam_{it-1} after every step.
    Returns dJ_dv, which equals d loglik / dc (node order of ``disc.mesh``)."""
    p = disc.prob
    if K is None:
        K = stiffness_csr(c, disc.mesh, disc.ops)
    N = disc.mesh.N
    dt2 = p.dt * p.dt
    Minv = 1.0 / (disc.M + 1e-30)  # S/core.py:1172 (quirk Q8)
    D = disc.decay
    RT = disc.R.T.tocsr()
    rs = residual / sigma2
    lam_curr = np.zeros(N)
    lam_next = np.zeros(N)
    g = np.zeros(N)
    exact = (mode == "exact")
    for it in range(p.nt - 1, stop - 1, -1):
        rt = RT @ rs[it]
        rhs = rt - K @ lam_curr
        if exact:
            lam_prev = D * (2.0 * lam_curr - D * lam_next + dt2 * (Minv * rhs))
        else:
            lam_prev = 2.0 * lam_curr - lam_next + dt2 * (Minv * rhs)
        if it < p.nt - 1:  # lam_{nt-1} = 0 contributes nothing
            g += sensitivity_dc2(hist.get(it), lam_curr, disc.mesh, disc.ops)
        lam_next, lam_curr = lam_curr, lam_prev
        if after_step is not None:
            after_step(it, lam_curr)
    return 2.0 * c * g


def window_mask(nt: int, window) -> np.ndarray:
    """Rows of the misfit window (C1S/core.py:1433-1470): it0 <= t < it1 and
    (t - it0) % decim == 0; window = (it0, it1, decim) or None (all rows)."""
    m = np.zeros(nt, dtype=bool)
    if window is None:
        m[:] = True
        return m
    it0, it1, decim = window
    m[it0:it1:max(int(decim), 1)] = True
    return m


def window_from_times(dt: float, nt: int, tmin: float = 0.0, tmax: Optional[float] = None,
                      window_len: Optional[float] = None, decim: int = 1):
    """Step window of method.misfit_window (C1S/core.py:1440-1466)."""
    if tmax is None and window_len is not None:
        tmax = tmin + float(window_len)
    tmax = float(tmax) if tmax is not None else float(dt * (nt - 1))
    if tmax <= tmin:
        raise ValueError(f"misfit_window invalid: tmax({tmax}) <= tmin({tmin})")
    decim = max(int(decim), 1)
    it0 = int(round(tmin / dt))
    it1 = int(round(tmax / dt)) + 1
    it0 = max(0, min(it0, nt - 1))
    it1 = max(it0 + 1, min(it1, nt))
    return it0, it1, decim


def forward_adjoint(disc: Discretization, c: np.ndarray, y_obs: np.ndarray, noise_std: float,
                    mode: str = "reference", checkpoint_every: Optional[int] = None, window=None):
    """run_forward_and_adjoint's node-level part (S/core.py:1067-1278):
    returns dict(loglik, dloglik_dc (N,), seis (nt,nrec), residual).

    ``window`` = (it0, it1, decim) restricts the misfit to the windowed,
    decimated rows (C1S/core.py:1433-1520; y_obs is given on all nt rows and
    read on the windowed ones); in the reference mode the adjoint then stops
    at it0 as the reference's does."""
    K = stiffness_csr(c, disc.mesh, disc.ops)
    seis, hist = forward(disc, c, record="post", checkpoint_every=checkpoint_every, K=K)
    keep = window_mask(disc.prob.nt, window)
    loglik, r, sigma2 = misfit_loglik(seis * keep[:, None], y_obs * keep[:, None], noise_std)
    stop = window[0] if (window is not None and mode == "reference") else 0
    g = adjoint(disc, c, hist, r, sigma2, mode=mode, K=K, stop=stop)
    return {"loglik": loglik, "dloglik_dc": g, "seis": seis, "residual": r, "sigma2": sigma2}


# ---------------------------------------------------------------------------
# velocity parameterisations
# ---------------------------------------------------------------------------
def _bspline_1d(t: np.ndarray, knots: np.ndarray, k: int, i: int) -> np.ndarray:
    """Cox-de Boor recursion (as S/core.py:781-792), right end inclusive."""
    t = np.asarray(t, dtype=np.float64)
    if k == 0:
        last = knots[-1]
        hi = knots[i + 1]
        inside = (knots[i] <= t) & (t < hi)
        # the last non-degenerate interval also owns the right end point
        if hi == last and knots[i] < hi:
            inside |= (t == last)
        return inside.astype(np.float64)
    out = np.zeros_like(t)
    d1 = knots[i + k] - knots[i]
    if abs(d1) > 1e-10:
        out += (t - knots[i]) / d1 * _bspline_1d(t, knots, k - 1, i)
    d2 = knots[i + k + 1] - knots[i + 1]
    if abs(d2) > 1e-10:
        out += (knots[i + k + 1] - t) / d2 * _bspline_1d(t, knots, k - 1, i + 1)
    return out


def clamped_knots(a: float, b: float, n_ctrl: int, k: int = 3) -> np.ndarray:
    inner = np.linspace(a, b, n_ctrl - k + 1)
    return np.concatenate([np.full(k, a), inner, np.full(k, b)])


def bspline_basis_matrix(coord: np.ndarray, a: float, b: float, n_ctrl: int, k: int = 3) -> np.ndarray:
    """B[m, q] = N_q(coord[m]) for a clamped uniform degree-k B-spline with
    n_ctrl control values on [a, b] (the tensor velocity grid's 1-D factor)."""
    knots = clamped_knots(a, b, n_ctrl, k)
    return np.stack([_bspline_1d(coord, knots, k, q) for q in range(n_ctrl)], axis=1)


def greville(a: float, b: float, n_ctrl: int, k: int = 3) -> np.ndarray:
    knots = clamped_knots(a, b, n_ctrl, k)
    return np.array([knots[q + 1:q + k + 1].mean() for q in range(n_ctrl)])


def tensor_bspline_velocity(mesh: StructuredMesh, coef: np.ndarray, base: Optional[np.ndarray] = None):
    """c[jz, ix] = base + sum_pq Bz[jz,p] Bx[ix,q] coef[p,q]  (lexicographic).

    No reference counterpart (the reference's velocity is a closed-curve
    anomaly); its nearest analogue is the gridded background of
    C1S/velocity.py:212-274.  Returns (c (N,), Bz, Bx)."""
    nz, nx = coef.shape
    Bz = bspline_basis_matrix(mesh.z1d, mesh.zmin, mesh.zmax, nz)
    Bx = bspline_basis_matrix(mesh.x1d, mesh.xmin, mesh.xmax, nx)
    c = Bz @ coef @ Bx.T
    if base is not None:
        c = c + base.reshape(mesh.NZ, mesh.NX)
    return c.reshape(-1), Bz, Bx


def layered_background(z: np.ndarray, interfaces_z, velocities) -> np.ndarray:
    """v = velocities[searchsorted(interfaces, z, 'right')]  (CLY/velocity.py:187-219)."""
    idx = np.searchsorted(np.asarray(interfaces_z, dtype=np.float64), z, side="right")
    return np.asarray(velocities, dtype=np.float64)[idx]


# -- closed-curve anomaly (S/velocity.py:8-202, S/core.py:902-970) ----------
def _closed_curve_knots(k: int = 3):
    n = 6 + k - 1
    knots = np.zeros(n + k + 2)
    knots[-(k + 1):] = 1.0
    knots[k + 1:-(k + 1)] = np.linspace(0.0, 1.0, n - k + 2)[1:-1]
    return knots


def closed_curve_velocity(coords: np.ndarray, ctrl6: np.ndarray, v_in: float, v_out: float,
                          tau: float, samples: int = 800, newton_steps: int = 7):
    """Signed distance to the closed cubic B-spline (6 control points, P0
    repeated 3 times, clamped knots) by nearest sample + Newton, sign from a
    point-in-polygon test of the sampled curve, then
    v = v_out + (v_in - v_out) sigmoid(-sd/tau)   (S/velocity.py:8-202).
    Returns (v, sd, u_star, n_out)."""
    from scipy.interpolate import BSpline
    from matplotlib.path import Path
    ctrl6 = np.asarray(ctrl6, dtype=np.float64).reshape(6, 2)
    k = 3
    P = np.vstack([ctrl6, np.tile(ctrl6[0], (k, 1))])
    knots = _closed_curve_knots(k)
    spl = BSpline(knots, P, k, extrapolate=False)
    s1, s2 = spl.derivative(1), spl.derivative(2)
    u0 = np.linspace(knots[3], knots[-4], samples)
    C0 = spl(u0)
    path = Path(C0, closed=True)
    X = np.asarray(coords, dtype=np.float64)
    sd = np.empty(len(X))
    ust = np.empty(len(X))
    nout = np.empty((len(X), 2))
    for s in range(0, len(X), 20000):
        Xc = X[s:s + 20000]
        d2 = ((Xc[:, None, :] - C0[None, :, :]) ** 2).sum(-1)
        u = u0[np.argmin(d2, axis=1)]
        for _ in range(newton_steps):
            r = spl(u) - Xc
            cp, cpp = s1(u), s2(u)
            f1 = (r * cp).sum(-1)
            f2 = (cp * cp).sum(-1) + (r * cpp).sum(-1)
            f2 = np.where(np.abs(f2) < 1e-14, 1e-14, f2)
            u = np.clip(u - f1 / f2, knots[3], knots[-4])
        rv = Xc - spl(u)
        dist = np.linalg.norm(rv, axis=1)
        sign = np.where(path.contains_points(Xc, radius=1e-12), -1.0, 1.0)
        sd[s:s + 20000] = sign * dist
        ust[s:s + 20000] = u
        nout[s:s + 20000] = sign[:, None] * rv / (dist[:, None] + 1e-15)
    x = np.clip(sd / max(float(tau), 1e-12), -50.0, 50.0)
    sig = 1.0 / (1.0 + np.exp(x))
    v = v_out + (v_in - v_out) * sig
    return v, sd, ust, nout


def closed_curve_projection(dL_dv: np.ndarray, sd, ust, nout, v_in, v_out, tau):
    """Analytic projection of a nodal gradient onto the 6 control points,
    returned interleaved [x0, z0, x1, z1, ...]  (S/core.py:902-970,
    1357-1363)."""
    from scipy.interpolate import BSpline
    ts = max(float(tau), 1e-12)
    x = np.clip(sd / ts, -50.0, 50.0)
    sig = 1.0 / (1.0 + np.exp(x))
    dv = (v_in - v_out) * sig * (1.0 - sig) * (-1.0 / ts)
    mask = np.abs(dv) > 1e-8 * (v_out - v_in)
    out = np.zeros(12)
    if not np.any(mask):
        return out
    w = (dL_dv[mask] * dv[mask])[:, None]
    knots = _closed_curve_knots(3)
    B9 = np.stack([BSpline(knots, np.eye(9)[j], 3, extrapolate=False)(ust[mask]) for j in range(9)], axis=1)
    gx = (w * B9 * nout[mask, 0:1]).sum(0)
    gz = (w * B9 * nout[mask, 1:2]).sum(0)
    gx6 = np.concatenate([[gx[0] + gx[6] + gx[7] + gx[8]], gx[1:6]])
    gz6 = np.concatenate([[gz[0] + gz[6] + gz[7] + gz[8]], gz[1:6]])
    out[0::2] = gx6
    out[1::2] = gz6
    return out


# -- gridded background (C1S/velocity.py:212-274, C1S/core.py:353-415) -------
def grid_background(coords: np.ndarray, xg: np.ndarray, zg: np.ndarray, vp_grid: np.ndarray,
                    method: str = "linear", clip: bool = True) -> np.ndarray:
    """Background velocity on nodes from vp_grid[iz, ix] on increasing axes:
    scipy RegularGridInterpolator over (z, x) at the node coordinates, clipped
    into the grid (build_background_from_grid, C1S/velocity.py:212-274)."""
    from scipy.interpolate import RegularGridInterpolator
    X = np.asarray(coords, dtype=np.float64).reshape(-1, 2)
    xg = np.asarray(xg, dtype=np.float64).reshape(-1)
    zg = np.asarray(zg, dtype=np.float64).reshape(-1)
    vp_grid = np.asarray(vp_grid, dtype=np.float64)
    if vp_grid.shape != (zg.size, xg.size):
        raise ValueError(f"vp_grid shape {vp_grid.shape} must be (len(zg), len(xg))")
    x, z = X[:, 0].copy(), X[:, 1].copy()
    if clip:
        x = np.clip(x, xg[0], xg[-1])
        z = np.clip(z, zg[0], zg[-1])
    rgi = RegularGridInterpolator((zg, xg), vp_grid, method=str(method), bounds_error=False, fill_value=None)
    return rgi(np.column_stack([z, x])).astype(np.float64)


def increasing_grid(xg, zg, vp_grid):
    """Flip decreasing axes together with the grid (C1S/core.py:396-403)."""
    xg = np.asarray(xg, dtype=np.float64).reshape(-1)
    zg = np.asarray(zg, dtype=np.float64).reshape(-1)
    vp_grid = np.asarray(vp_grid, dtype=np.float64)
    if xg[0] > xg[-1]:
        xg, vp_grid = xg[::-1], vp_grid[:, ::-1]
    if zg[0] > zg[-1]:
        zg, vp_grid = zg[::-1], vp_grid[::-1, :]
    return xg, zg, vp_grid


# -- periodic closed curve over a background (C1S/velocity.py:9-152, 277-332) --
def _periodic_curve(ctrl: np.ndarray, k: int = 3):
    """ctrl extended by its first k points, knots 0, 1, ..., len + k
    (C1S/velocity.py:9-35)."""
    from scipy.interpolate import BSpline
    ctrl = np.asarray(ctrl, dtype=np.float64).reshape(-1, 2)
    P = np.vstack([ctrl, ctrl[:k]])
    knots = np.arange(0.0, len(P) + k + 1, dtype=np.float64)
    return BSpline(knots, P, k, extrapolate=False), knots, P


def periodic_curve_velocity(coords: np.ndarray, ctrl: np.ndarray, v_in: float, v_bg, tau: float,
                            samples: int = 800, newton_steps: int = 7, blend: str = "smooth"):
    """Signed distance to the periodic closed cubic B-spline by nearest of
    `samples` samples on [3, K+3) (endpoint excluded) + Newton (clipped to
    [3, nextafter(K+3, 3)]), sign from the sampled polygon, u* normalised to
    [0, 1) (C1S/velocity.py:54-152); blended over the background v_bg (scalar
    or per node): smooth v = v_bg + (v_in - v_bg) sigmoid(-sd/tau), replace
    v = v_in inside (C1S/velocity.py:277-332).  Returns (v, sd, u_star, n_out)."""
    from matplotlib.path import Path
    spl, knots, _ = _periodic_curve(ctrl)
    s1, s2 = spl.derivative(1), spl.derivative(2)
    k = 3
    lo, hi = float(knots[k]), float(knots[-(k + 1)])
    hi_safe = np.nextafter(hi, lo)
    u0 = np.linspace(lo, hi, samples, endpoint=False, dtype=np.float64)
    C0 = spl(u0)
    path = Path(C0, closed=True)
    X = np.asarray(coords, dtype=np.float64).reshape(-1, 2)
    n = len(X)
    sd, ust, nout = np.empty(n), np.empty(n), np.empty((n, 2))
    for s in range(0, n, 20000):
        Xc = X[s:s + 20000]
        d2 = ((Xc[:, None, :] - C0[None, :, :]) ** 2).sum(-1)
        u = u0[np.argmin(d2, axis=1)]
        for _ in range(newton_steps):
            r = spl(u) - Xc
            cp, cpp = s1(u), s2(u)
            f1 = (r * cp).sum(-1)
            f2 = (cp * cp).sum(-1) + (r * cpp).sum(-1)
            f2 = np.where(np.abs(f2) < 1e-14, 1e-14, f2)
            u = np.clip(u - f1 / f2, lo, hi_safe)
        rv = Xc - spl(u)
        dist = np.linalg.norm(rv, axis=1)
        sign = np.where(path.contains_points(Xc, radius=1e-12), -1.0, 1.0)
        sd[s:s + 20000] = sign * dist
        ust[s:s + 20000] = (u - lo) / max(hi - lo, 1e-300)
        nout[s:s + 20000] = sign[:, None] * rv / (dist[:, None] + 1e-15)
    v_bg = np.broadcast_to(np.asarray(v_bg, dtype=np.float64), (n,))
    if blend == "replace":
        v = v_bg.copy()
        v[sd < 0.0] = float(v_in)
    else:
        x = np.clip(sd / max(float(tau), 1e-12), -50.0, 50.0)
        sig = 1.0 / (1.0 + np.exp(x))
        v = v_bg + (float(v_in) - v_bg) * sig
    return v, sd, ust, nout


def periodic_curve_projection(dL_dv: np.ndarray, sd, ust, nout, v_in: float, v_bg, tau: float, K: int):
    """Projection onto the K control points of the periodic curve, interleaved
    [x0, z0, ...]: periodic basis at t = 3 + (u* mod 1) K, the 3 appended
    controls folded back onto ctrl0..ctrl2, nodes kept where |dv/dsd| > 1e-8
    max(1, |v_in - v_bg|)  (C1S/core.py:1075-1185)."""
    from scipy.interpolate import BSpline
    n = len(sd)
    v_bg = np.broadcast_to(np.asarray(v_bg, dtype=np.float64), (n,))
    ts = max(float(tau), 1e-12)
    x = np.clip(np.asarray(sd) / ts, -50.0, 50.0)
    sig = 1.0 / (1.0 + np.exp(x))
    dv = (float(v_in) - v_bg) * sig * (1.0 - sig) * (-1.0 / ts)
    mask = np.abs(dv) > 1e-8 * np.maximum(1.0, np.abs(float(v_in) - v_bg))
    out = np.zeros(2 * K)
    if not np.any(mask):
        return out
    w = (np.asarray(dL_dv)[mask] * dv[mask])[:, None]
    knots = np.arange(0.0, K + 3 + 3 + 1, dtype=np.float64)
    lo, hi = float(knots[3]), float(knots[-4])
    t = np.clip(lo + np.mod(np.asarray(ust)[mask], 1.0) * (hi - lo), lo, np.nextafter(hi, lo))
    Bc = np.stack([BSpline(knots, np.eye(K + 3)[j], 3, extrapolate=False)(t) for j in range(K + 3)], axis=1)
    B = Bc[:, :K].copy()
    B[:, :3] += Bc[:, K:]
    gx = (w * B * nout[mask, 0:1]).sum(0)
    gz = (w * B * nout[mask, 1:2]).sum(0)
    out[0::2] = gx
    out[1::2] = gz
    return out
