"""Host-side geometry of the structured GLL mesh (node coordinates only).

The device library builds its own FP64 tables (GLL nodes, weights,
derivative matrix, sponge, receiver weights) at plan creation; this module
gives Python callers the node coordinates for laying out velocity models and
for the reference-compatible ``global_coords`` attribute.
"""
import numpy as np


def gll_nodes(p: int) -> np.ndarray:
    """Gauss-Lobatto-Legendre nodes on [-1, 1]: +-1 and the roots of P_p'."""
    if p < 1:
        raise ValueError("polynomial order must be >= 1")
    dP = np.polynomial.legendre.Legendre.basis(p).deriv()
    inner = np.sort(np.real(dP.roots()))
    # polish the roots with Newton steps on P_p'
    d2P = dP.deriv()
    for _ in range(3):
        inner = inner - dP(inner) / d2P(inner)
    return np.concatenate([[-1.0], inner, [1.0]])


def gll_axis(a: float, b: float, nelem: int, p: int) -> np.ndarray:
    """Node coordinates along one axis: x = 0.5((1-xi) x0 + (1+xi) x1) per
    element, shared element-edge nodes taken from the element corners."""
    xi = gll_nodes(p)
    corners = np.linspace(a, b, nelem + 1)
    out = np.empty(nelem * p + 1)
    for e in range(nelem):
        x0, x1 = corners[e], corners[e + 1]
        out[e * p:(e + 1) * p + 1] = 0.5 * ((1 - xi) * x0 + (1 + xi) * x1)
    return out


def node_coords(x1d: np.ndarray, z1d: np.ndarray) -> np.ndarray:
    """(N, 2) coordinates in lexicographic order (x fastest)."""
    X, Z = np.meshgrid(x1d, z1d, indexing="xy")
    return np.stack([X.reshape(-1), Z.reshape(-1)], axis=1)
