"""The library's quadrature tables (GaussQuadrature, src/fem/quadrature.rs:
13-104) equal the oracle's bit for bit; the oracle's are pinned by exactness
on monomials (tests/test_oracle_basis.py).  Host only."""
import numpy as np
import pytest

import geo_simulator_amd as G


@pytest.mark.parametrize("n", [1, 4, 5, 11])
def test_tables_equal_oracle(oracle, n):
    q = {1: G.GaussQuadrature.tet_1point, 4: G.GaussQuadrature.tet_4point,
         5: G.GaussQuadrature.tet_5point, 11: G.GaussQuadrature.tet_11point}[n]()
    pts, wts = oracle.quad_rule(n)
    assert q.num_points() == n
    assert np.array_equal(np.array(q.points), pts) and np.array_equal(np.array(q.weights), wts)
    # the reference's own KAT: the weights sum to the reference volume 1/6
    assert abs(sum(q.weights) - 1.0 / 6.0) < 1e-15


def test_bad_rule():
    with pytest.raises(ValueError):
        G.GaussQuadrature(3)
