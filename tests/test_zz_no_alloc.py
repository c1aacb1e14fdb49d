"""sem_bspline_project_noalloc (the projection that allocates nothing) gives
the bits of sem_bspline_project, through the C ABI, on two calls in a row --
the second on a smaller batch -- including a batch that spans several chunks
of the library's scratch array."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _plan(nex, nez, L=4000.0):
    from mun_amd.solver import Plan
    return Plan(nex, nez, 0.0, L, 0.0, L, 2.5e-4, 10, 4, 300.0, 3500.0, 1e-8, ("xmin", "xmax", "zmax"),
                precision="fp32", checkpoint_every=0, device="cuda:0")


def _project(plan, fn, g, nbz, nbx):
    out = torch.full((g.shape[0], nbz, nbx), float("nan"), dtype=torch.float64, device=plan.device)
    rc = fn(plan.h, g.shape[0], nbz, nbx, C.c_void_p(g.data_ptr()), C.c_void_p(out.data_ptr()),
            C.c_void_p(torch.cuda.current_stream(plan.device).cuda_stream))
    return rc, out


def test_noalloc_projection_equals_the_allocating_one():
    from mun_amd import _lib as L
    lib = L.lib()
    plan = _plan(200, 200)   # 801 node rows: 81 models of 16 columns per 8 MiB chunk
    nbz = nbx = 16
    plan.bspline_map(torch.zeros((1, nbz, nbx), dtype=torch.float64))   # builds the basis tables
    rng = np.random.default_rng(17)
    for M in (170, 7):      # three chunks, then a smaller batch in the same scratch
        print(f"[noalloc] {M} models", flush=True)
        g = torch.as_tensor(rng.standard_normal((M, plan.N)), device="cuda:0")
        rc_a, ref = _project(plan, lib.sem_bspline_project, g, nbz, nbx)
        rc_b, got = _project(plan, lib.sem_bspline_project_noalloc, g, nbz, nbx)
        torch.cuda.synchronize()
        assert rc_a == 0 and rc_b == 0, lib.sem_last_error()
        assert bool(torch.isfinite(got).all())
        assert torch.equal(got, ref)
        # and against a dense projection Bz^T g Bx on the host for one model
        import oracle as O
        Bz = O.bspline_basis_matrix(plan.z1d, 0.0, 4000.0, nbz)
        Bx = O.bspline_basis_matrix(plan.x1d, 0.0, 4000.0, nbx)
        k = M - 1
        dense = Bz.T @ g[k].cpu().numpy().reshape(plan.NZ, plan.NX) @ Bx
        assert np.abs(got[k].cpu().numpy() - dense).max() <= 1e-12 * np.abs(dense).max()


def test_noalloc_projection_errors():
    from mun_amd import _lib as L
    lib = L.lib()
    plan = _plan(12, 10, L=1000.0)
    g = torch.zeros((2, plan.N), dtype=torch.float64, device="cuda:0")
    # no basis tables for 9 x 9 nodes yet: the call would have to allocate them
    rc, _ = _project(plan, lib.sem_bspline_project_noalloc, g, 9, 9)
    assert rc == -1
    assert "basis tables" in lib.sem_last_error().decode()
    # one model larger than the scratch: 1 x 8200 elements, 32801 node rows x 32 columns
    tall = _plan(1, 8200, L=1000.0)
    tall.bspline_map(torch.zeros((1, 4, 32), dtype=torch.float64))
    gt = torch.zeros((1, tall.N), dtype=torch.float64, device="cuda:0")
    rc, _ = _project(tall, lib.sem_bspline_project_noalloc, gt, 4, 32)
    assert rc == -3 and "scratch" in lib.sem_last_error().decode()
