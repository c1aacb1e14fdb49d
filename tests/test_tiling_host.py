"""Host emulation of the step kernel's work decomposition.

Restates, in Python, the index arithmetic of ``step_kernel`` in
mun_amd/csrc/sem_kernels.h (17-wide tiles of 15, 7 or 3 element rows ->
element threads, halo column/row,
owner-computes assembly through the 9-value exchange) and checks on the CPU
that (1) every global node is written exactly once, (2) every read is in
bounds, and (3) the assembled element contributions equal the oracle's
assembled K u.  Meshes include partial tiles and 1-element edges.
"""
import numpy as np
import pytest

import oracle as O

EX = 17
TX = EX - 1
P = 4


def emulate_Ku(nex, nez, u, c2, dx, dz, EZ=15):
    TZ = EZ - 1
    NX, NZ = nex * P + 1, nez * P + 1
    xi = O.gll_points(P)
    w = O.gll_weights(xi)
    D = O.differentiation_matrix(xi)
    cx, cz = 4.0 / dx ** 2, 4.0 / dz ** 2
    J = dx * dz / 4
    ntx, ntz = -(-nex // TX), -(-nez // TZ)
    out = np.zeros(NX * NZ)
    writes = np.zeros(NX * NZ, dtype=int)
    for tz in range(ntz):
        for tx in range(ntx):
            col = np.zeros((EZ, EX, P + 1))
            row = np.zeros((EZ, EX, P))
            rr = {}
            for ezl in range(EZ):
                for exl in range(EX):
                    ex, ez = tx * TX - 1 + exl, tz * TZ - 1 + ezl
                    valid = 0 <= ex < nex and 0 <= ez < nez
                    r = np.zeros((P + 1, P + 1))
                    if valid:
                        rows = P * ez + np.arange(P + 1)
                        cols = P * ex + np.arange(P + 1)
                        assert rows.max() < NZ and cols.max() < NX
                        ue = u.reshape(NZ, NX)[np.ix_(rows, cols)]
                        cw = c2.reshape(NZ, NX)[np.ix_(rows, cols)] * np.outer(w, w)
                        for j in range(P + 1):
                            f = cw[j] * cx * (D @ ue[j])
                            r[j] = D.T @ f
                        for i in range(P + 1):
                            f = cw[:, i] * cz * (D @ ue[:, i])
                            r[:, i] += D.T @ f
                    col[ezl, exl] = r[:, P]
                    row[ezl, exl] = r[P, :P]
                    rr[(ezl, exl)] = (r, ex, ez, valid)
            for (ezl, exl), (r, ex, ez, valid) in rr.items():
                if not (valid and exl >= 1 and ezl >= 1):
                    continue
                jmax = P if ez == nez - 1 else P - 1
                imax = P if ex == nex - 1 else P - 1
                for j in range(jmax + 1):
                    for i in range(imax + 1):
                        acc = r[j, i]
                        if i == 0:
                            acc += col[ezl, exl - 1, j]
                        if j == 0:
                            acc += row[ezl - 1, exl, i] if i < P else col[ezl - 1, exl, P]
                        if i == 0 and j == 0:
                            acc += col[ezl - 1, exl - 1, P]
                        g = (P * ez + j) * NX + P * ex + i
                        out[g] = acc * J
                        writes[g] += 1
    return out, writes


@pytest.mark.parametrize("ez", [15, 7, 3])
@pytest.mark.parametrize("nex,nez", [(1, 1), (2, 3), (16, 14), (17, 15), (33, 29), (31, 2)])
def test_tile_decomposition_matches_assembled_K(nex, nez, ez):
    rng = np.random.default_rng(nex * 100 + nez)
    mesh = O.StructuredMesh(0.0, 25.0 * nex, 0.0, 20.0 * nez, nex, nez, P)
    ops = O.sem_operators(mesh.dx, mesh.dz, P, lagrange_eps=0.0)
    c = 1500 + 2000 * rng.random(mesh.N)
    u = rng.standard_normal(mesh.N)
    K = O.stiffness_csr(c, mesh, ops)
    ref = K @ u
    got, writes = emulate_Ku(nex, nez, u, c * c, mesh.dx, mesh.dz, ez)
    assert writes.min() == 1 and writes.max() == 1, "every node must be written exactly once"
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
