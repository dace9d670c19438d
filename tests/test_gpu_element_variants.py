"""GPU test of the elastic element kernel variants against the FP64 VALU
kernel that assembly uses (itself checked against the oracle element by
element, bit for bit): the FP64 matrix-core (MFMA) kernel and the FP64
row-per-lane kernel within FP64 rounding, the FP32 B^T D B kernels (VALU and
MFMA) within FP32 rounding, all timed."""
import numpy as np
import pytest

import geo_simulator_amd as G
from geo_simulator_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("per_elem", [False, True])
def test_elastic_element_variants_match_valu(per_elem):
    mesh, _, _ = W.elastic_box(G, 8, 7, 5)
    mat = G.IsotropicElasticity(1e11, 0.25)
    E = None
    if per_elem:
        rng = np.random.default_rng(12)
        E = np.exp(np.log(1e11) + 0.3 * rng.standard_normal(mesh.num_elements()))
    out = {v: G.time_elastic_elements(mesh, mat, E, v, reps=3)
           for v in ("valu", "mfma", "valu_rows", "valu_f32", "mfma_f32")}
    print({k: (f"{ms:.3f} ms", f"{d:.1e}") for k, (ms, d) in out.items()}, flush=True)
    assert out["valu"][1] == 0.0
    assert out["valu_rows"][1] == 0.0  # same products in the same order
    assert out["mfma"][1] < 1e-13
    for v in ("valu_f32", "mfma_f32"):
        assert 0.0 < out[v][1] < 2e-6, v  # FP32 rounding of the entries and the sums
    assert all(ms > 0.0 for ms, _ in out.values())


def test_element_variant_errors():
    mesh, _, _ = W.elastic_box(G, 2, 2, 2)
    with pytest.raises(ValueError):
        G.time_elastic_elements(mesh, G.IsotropicElasticity(1e11, 0.25), variant="tensor")
    with pytest.raises(G.GeoFemError):
        G.time_elastic_elements(mesh, G.IsotropicElasticity(1e11, 0.25), reps=0)
