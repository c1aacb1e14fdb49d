"""Clamped uniform B-spline helpers used to lay out velocity control grids.

The tensor-product velocity map itself (c = base + Bz C Bx^T) runs on the GPU
(``sem_bspline_map`` in the C-ABI library); this module only provides the knot
vector and the Greville abscissae that workload generators place control
values at.
"""
import numpy as np


def clamped_knots(a: float, b: float, n_ctrl: int, degree: int = 3) -> np.ndarray:
    """Open uniform knot vector with ``degree`` repeated end knots."""
    if n_ctrl <= degree:
        raise ValueError(f"need more than {degree} control values, got {n_ctrl}")
    inner = np.linspace(a, b, n_ctrl - degree + 1)
    return np.concatenate([np.full(degree, a), inner, np.full(degree, b)])


def greville(a: float, b: float, n_ctrl: int, degree: int = 3) -> np.ndarray:
    """Greville abscissae: mean of ``degree`` consecutive interior knots."""
    t = clamped_knots(a, b, n_ctrl, degree)
    return np.array([t[q + 1:q + degree + 1].mean() for q in range(n_ctrl)])
