"""The torch-CPU form of the oracle (timed as the CPU baseline) computes what
the oracle computes: same seismograms, log-likelihood and gradient."""
import numpy as np
import pytest

import oracle as O
from oracle.torch_cpu import forward_adjoint_torch
from tests.gpu_helpers import rel_l2, small_problem


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_torch_cpu_matches_oracle(mode):
    prob, disc, coef, c = small_problem(nex=9, nez=8, nt=220)
    rng = np.random.default_rng(2)
    cobs, _, _ = O.tensor_bspline_velocity(disc.mesh, coef + 30.0 * rng.standard_normal(coef.shape))
    yobs, _ = O.forward(disc, cobs, history=None)
    ref = O.forward_adjoint(disc, c, yobs, 1e-9, mode=mode)
    got = forward_adjoint_torch(disc, c, yobs, 1e-9, mode=mode, threads=2)
    assert got["steps"] == prob.nt
    assert rel_l2(got["seis"], ref["seis"]) < 1e-12
    assert abs(got["loglik"] / ref["loglik"] - 1.0) < 1e-12
    assert rel_l2(got["dloglik_dc"], ref["dloglik_dc"]) < 1e-11
    assert set(got["times"]) == {"assembly", "forward", "adjoint"}


def test_torch_cpu_step_sample():
    """A bounded sample runs the first nsteps steps only."""
    prob, disc, coef, c = small_problem(nex=6, nez=5, nt=300)
    got = forward_adjoint_torch(disc, c, np.zeros((prob.nt, prob.rec_xz.shape[0])), 1e-9, nsteps=40, threads=1)
    assert got["steps"] == 40 and got["seis"].shape[0] == 40
    ref = O.forward(disc, c, history=None)[0][:40]
    assert rel_l2(got["seis"], ref) < 1e-12
