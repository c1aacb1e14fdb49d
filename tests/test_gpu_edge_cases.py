"""Edge cases on the GPU through the C ABI: meshes smaller than one tile or
one tile wide/high, receivers on element edges and domain corners, a source
on a corner node, solves without receivers, and non-finite models (rejected
with SEM_E_NONFINITE before any step)."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
from tests.gpu_helpers import gpu_setup, make_plan, rel_l2

pytestmark = pytest.mark.gpu


def tiny_problem(nex, nez, nt=120, src=None, corner_receivers=True):
    L_x, L_z = 40.0 * nex, 40.0 * nez
    rec = [[0.5 * L_x, 0.0], [40.0 * (nex // 2), 0.5 * L_z]]  # on an element edge
    if corner_receivers:
        rec += [[0.0, 0.0], [L_x, L_z], [L_x, 0.0], [0.0, L_z]]
    src = src if src is not None else (0.5 * L_x + 3.0, 0.5 * L_z - 2.0)
    prob = O.Problem(0.0, L_x, 0.0, L_z, nex, nez, 4, 4e-4, nt, src, 25.0, np.array(rec),
                     sponge_width=min(L_x, L_z) * 0.3, sponge_cmax=3000.0, sponge_sides=("xmin", "xmax", "zmax"))
    disc = O.discretize(prob)
    rng = np.random.default_rng(nex * 100 + nez)
    coef = 1800.0 + 800.0 * rng.random((4, 4))
    c, _, _ = O.tensor_bspline_velocity(disc.mesh, coef)
    return prob, disc, c


@pytest.mark.parametrize("nex,nez", [(1, 1), (2, 3), (17, 1), (1, 16), (33, 2), (16, 14), (16, 15)])
def test_tiny_and_tile_edge_meshes_match_oracle(nex, nez):
    print(f"[mesh {nex}x{nez}]", flush=True)
    prob, disc, c = tiny_problem(nex, nez)
    yobs = 0.7 * O.forward(disc, c, history=None)[0]
    ref = O.forward_adjoint(disc, c, yobs, 1e-9, mode="exact")
    for tile_rows in (15, 3):
        plan = make_plan(prob, "fp64", adjoint_mode="exact", checkpoint_every=50, tile_rows=tile_rows)
        rs, src, w, c2 = gpu_setup(prob, disc, c, plan)
        obs = torch.as_tensor(yobs, dtype=torch.float64, device=plan.device)[None]
        out = plan.forward_adjoint(c2, w, [src], obs, 1e-9, [rs])
        torch.cuda.synchronize()
        assert rel_l2(out["seis"][0].cpu().numpy(), ref["seis"]) < 1e-9
        assert rel_l2(out["grad_c"][0].cpu().numpy(), ref["dloglik_dc"]) < 1e-9


def test_source_on_a_corner_node():
    print("[corner source]", flush=True)
    prob, disc, c = tiny_problem(5, 4, src=(200.0, 160.0))  # the top-right corner node
    assert disc.src_node == disc.mesh.N - 1
    ref, _ = O.forward(disc, c, history=None)
    plan = make_plan(prob, "fp64")
    rs, src, w, c2 = gpu_setup(prob, disc, c, plan)
    seis = plan.forward(c2, w, [src], [rs])
    assert rel_l2(seis[0].cpu().numpy(), ref) < 1e-9


def test_solves_without_receivers():
    print("[no receivers]", flush=True)
    prob, disc, c = tiny_problem(6, 5)
    plan = make_plan(prob, "fp32", adjoint_mode="exact", checkpoint_every=40)
    rs = plan.add_receivers(np.zeros((0, 2)))
    src = plan.source_node(*prob.src_xz)
    w = torch.as_tensor(disc.wavelet, dtype=plan.dtype, device=plan.device)
    c2 = torch.as_tensor(c * c, dtype=plan.dtype, device=plan.device)[None]
    seis = plan.forward(c2, w, [src], [rs])
    assert tuple(seis.shape) == (1, prob.nt, 0)
    obs = torch.zeros((1, prob.nt, 0), dtype=torch.float64, device=plan.device)
    out = plan.forward_adjoint(c2, w, [src], obs, 1e-3, [rs])
    assert float(out["misfit"][0]) == 0.0
    assert float(out["grad_c"].abs().max()) == 0.0
    _, snaps = plan.forward_snapshots(c2, w, [src], 40, (0, plan.NX, 0, plan.NZ), [rs])
    assert float(snaps.abs().max()) > 0.0


@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
def test_non_finite_model_is_rejected_by_the_c_abi(bad):
    from mun_amd import _lib as L
    print(f"[non-finite {bad}]", flush=True)
    prob, disc, c = tiny_problem(4, 4)
    plan = make_plan(prob, "fp32")
    rs, src, w, c2 = gpu_setup(prob, disc, c, plan)
    c2 = c2.clone()
    c2[0, 17] = bad
    seis = torch.empty((1, prob.nt, plan.rec_sizes[rs]), dtype=plan.dtype, device=plan.device)
    srct = torch.tensor([src], dtype=torch.int32, device=plan.device)
    bt = plan._batch(c2, w, srct, [rs], None, seis, seis.shape[2])
    ws = plan.workspace(1, seis.shape[2], False)
    rc = plan.lib.sem_forward(plan.h, C.byref(bt), C.c_void_p(ws.data_ptr()), ws.numel(), None)
    assert rc == -5
    assert "NaN or Inf" in plan.lib.sem_last_error().decode()
    # the Python layer refuses before calling in
    with pytest.raises(ValueError):
        plan.forward(c2, w, [src], [rs])
    # a finite model on the same plan still runs
    good = plan.forward(c2.nan_to_num(posinf=4e6, nan=4e6), w, [src], [rs])
    assert bool(torch.isfinite(good).all())
