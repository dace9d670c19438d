"""CPU checks of the benchmark workload generators (geo_simulator_amd.workloads):
the boundary sets match the oracle's restatement of the reference benchmarks'
selections, the closed forms solve their equations, the Gaussian random field
has the stated statistics and correlation length, and the adjoint cases are
reproducible.  No compute entry point of the library is called."""
import math

import numpy as np

import geo_simulator_amd as G
from geo_simulator_amd import workloads as W


def test_thermal_box_matches_oracle_bcs(oracle):
    mesh, dof, exact = W.thermal_box(G, 4, 3, 5)
    xyz, _ = oracle.generate_cube(4, 3, 5, 1000.0, 1000.0, 600.0)
    mask, vals = oracle.thermal_dirichlet(xyz, 600.0)
    assert np.array_equal(dof.dirichlet_mask.astype(np.uint8), mask)
    assert np.array_equal(dof.dirichlet_values[mask == 1], vals[mask == 1])
    z = xyz[:, 2]
    # exact solves -k T'' = Q with the boundary values
    k, q = 3.0, 1e-6
    T2 = -q / k  # second derivative of the quadratic
    zz = np.linspace(0, 600, 7)
    Tz = 1280.0 * (1 - zz / 600.0) + (q / (2 * k)) * zz * (600.0 - zz)
    assert abs(Tz[0] - 1280.0) < 1e-12 and abs(Tz[-1]) < 1e-12
    assert abs(np.polyfit(zz, Tz, 2)[0] * 2 - T2) < 1e-15
    assert np.allclose(exact, oracle.thermal_analytic(z, 600.0, 1280.0, 0.0, 3.0, 1e-6),
                       rtol=0, atol=1e-9)


def test_elastic_box_matches_oracle_bcs(oracle):
    mesh, dof, uz = W.elastic_box(G, 3, 4, 2)
    xyz, _ = oracle.generate_cube(3, 4, 2, 1.0e6, 1.0e6, 6.0e5)
    mask, _ = oracle.elastic_dirichlet(xyz, 1.0e6, 1.0e6)
    assert np.array_equal(dof.dirichlet_mask.astype(np.uint8), mask)
    assert np.all(dof.dirichlet_values == 0.0)
    ref = oracle.elastic_analytic_uz(xyz[:, 2], 6.0e5, 3300.0, 9.81, 1e11, 0.25)
    assert np.allclose(uz, ref, rtol=1e-14, atol=0)
    assert np.all(uz[np.abs(xyz[:, 2]) < 1.0] == 0.0)


def test_gaussian_field_statistics():
    f = W.gaussian_field(64, 64, 32, 1000.0, 1000.0, 600.0, 100.0, seed=5)
    assert f.shape == (64 * 64 * 32,)
    assert abs(f.mean()) < 1e-12 and abs(f.std() - 1.0) < 1e-12
    g = f.reshape(32, 64, 64)
    # correlation at one correlation length along x (cell 1000/64 km) is
    # exp(-1/2) for a squared-exponential covariance, up to sampling noise
    lag = int(round(100.0 / (1000.0 / 64)))
    c = np.mean(g * np.roll(g, lag, axis=2))
    assert abs(c - math.exp(-0.5 * (lag * 1000.0 / 64 / 100.0) ** 2)) < 0.1
    # reproducible, and different seeds differ
    assert np.array_equal(f, W.gaussian_field(64, 64, 32, 1000.0, 1000.0, 600.0, 100.0, 5))
    assert not np.allclose(f, W.gaussian_field(64, 64, 32, 1000.0, 1000.0, 600.0, 100.0, 6))


def test_conductivity_map_monotone_in_range():
    x = np.linspace(-6, 6, 101)
    k = W.conductivity_from_field(x)
    assert np.all(np.diff(k) > 0) and k.min() > 1.0 and k.max() < 5.0
    assert abs(W.conductivity_from_field(np.array([0.0]))[0] - 3.0) < 1e-15


def test_adjoint_cases_reproducible():
    m1, d1, lk1, lt1, o1 = W.thermal_adjoint_case(G, 8, 8, 4, n_obs=50, seed=3)
    m2, d2, lk2, lt2, o2 = W.thermal_adjoint_case(G, 8, 8, 4, n_obs=50, seed=3)
    assert np.array_equal(lk1, lk2) and np.array_equal(o1, o2)
    assert len(np.unique(o1)) == 50
    assert not np.any(d1.dirichlet_mask[o1])
    # one value per hex, shared by its six tets
    assert np.all(lk1.reshape(-1, 6) == lk1.reshape(-1, 6)[:, :1])
    k = np.exp(lk1)
    assert k.min() > 1.0 and k.max() < 5.0
    me, de, lE, lEt, oe = W.elastic_adjoint_case(G, 4, 4, 2, seed=9)
    assert len(oe) == 9 * 9 and np.all(oe % 3 == 2)
    assert abs(np.median(lE) - math.log(1e11)) < 0.3
