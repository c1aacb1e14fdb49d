"""mun_amd: MI355X-native SEM forward/adjoint solver.

A drop-in for the 2-D acoustic spectral-element solver that the xingdaj/MUN
inversion loops call (Bspline_Coupling/SEM/sem_waveform): same Python call
surface (``SEMSimulation``, ``SEMLikelihoodAdjointFn``), computed by
hand-written HIP kernels for gfx950 behind a C ABI (include/sem_capi.h).
"""
__version__ = "0.1.0"

__all__ = ["Plan", "build"]


def __getattr__(name):
    if name == "Plan":
        from .solver import Plan
        return Plan
    if name == "build":
        from .build import build
        return build
    raise AttributeError(name)
