"""GPU test of the matrix-core (MFMA) variant of the elastic element kernel
against the VALU kernel that assembly uses (itself checked against the
oracle element by element): the same K_e within FP64 rounding of a
different summation order, and both timed."""
import numpy as np
import pytest

import geo_simulator_amd as G
from geo_simulator_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("per_elem", [False, True])
def test_mfma_elastic_elements_match_valu(per_elem):
    mesh, _, _ = W.elastic_box(G, 8, 7, 5)
    mat = G.IsotropicElasticity(1e11, 0.25)
    E = None
    if per_elem:
        rng = np.random.default_rng(12)
        E = np.exp(np.log(1e11) + 0.3 * rng.standard_normal(mesh.num_elements()))
    ms_v, d_v = G.time_elastic_elements(mesh, mat, E, "valu", reps=3)
    ms_m, d_m = G.time_elastic_elements(mesh, mat, E, "mfma", reps=3)
    assert d_v == 0.0
    assert d_m < 1e-13
    assert ms_v > 0.0 and ms_m > 0.0


def test_element_variant_errors():
    mesh, _, _ = W.elastic_box(G, 2, 2, 2)
    with pytest.raises(ValueError):
        G.time_elastic_elements(mesh, G.IsotropicElasticity(1e11, 0.25), variant="tensor")
    with pytest.raises(G.GeoFemError):
        G.time_elastic_elements(mesh, G.IsotropicElasticity(1e11, 0.25), reps=0)
