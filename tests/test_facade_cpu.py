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
