"""Facade behaviour that needs no GPU: config validation order and the
reference node numbering."""
import numpy as np
import pytest

import oracle as O
from mun_amd.simulation import SEMSimulation, cfl_dt, reference_node_order


@pytest.mark.parametrize("missing", ["domain", "time", "source", "receivers", "method", "velocity", "output"])
def test_missing_config_key_raises_valueerror(missing):
    cfg = {k: {} for k in SEMSimulation.REQUIRED}
    del cfg[missing]
    with pytest.raises(ValueError, match=f"Missing required config key: {missing}"):
        SEMSimulation(cfg)


@pytest.mark.parametrize("nex,nez", [(1, 1), (2, 3), (7, 5)])
def test_reference_node_order_matches_oracle_and_reference_mesh(nex, nez):
    perm = reference_node_order(nex, nez, 4)
    np.testing.assert_array_equal(perm, O.reference_node_order(nex, nez, 4))
    coords_ref, _ = O.create_global_mesh_reference(0.0, 10.0 * nex, 0.0, 10.0 * nez, nex, nez, 4)
    m = O.StructuredMesh(0.0, 10.0 * nex, 0.0, 10.0 * nez, nex, nez, 4)
    np.testing.assert_allclose(m.coords[perm], coords_ref, atol=1e-12)


def test_cfl_formula_matches_oracle():
    assert abs(cfl_dt(3000.0, 4, 50.0, 40.0) - O.reference_cfl_dt(3000.0, 4, 50.0, 40.0)) < 1e-18


def test_driver_noise_variance():
    """The drivers give the solver sqrt(noise_std^2 + EPS): EPS = 1e-12 in
    Bspline_Coupling/RQS_Coupling/RQS_Coupling.py:67-70, 1e-30 in
    CO2_plume/one_stage/single_cubic/RQS_NF.py:161.  At the Bspline driver's
    obs_noise_std = 5e-9 (RQS_Coupling.py:82) that is sqrt(1.000025e-12)."""
    import torch
    from mun_amd.autograd import EPS_BSPLINE_DRIVER, EPS_CO2_DRIVER, effective_noise_std
    assert EPS_BSPLINE_DRIVER == 1e-12 and EPS_CO2_DRIVER == 1e-30
    v = effective_noise_std(5e-9, EPS_BSPLINE_DRIVER)
    assert abs(v * v / (2.5e-17 + 1e-12) - 1.0) < 1e-14
    assert effective_noise_std(torch.tensor(5e-9, dtype=torch.float64), 1e-12) == v
    assert abs(effective_noise_std(1e-4, EPS_CO2_DRIVER) / 1e-4 - 1.0) < 1e-15


def test_source_positions_schema():
    from mun_amd.likelihood import source_positions
    np.testing.assert_array_equal(source_positions({"source": {"positions": [[1, 2], [3, 4]]}}),
                                  [[1.0, 2.0], [3.0, 4.0]])
    np.testing.assert_array_equal(source_positions({"source": {"position": (5, 6)}}), [[5.0, 6.0]])
    with pytest.raises(ValueError, match="non-empty"):
        source_positions({"source": {"positions": []}})
