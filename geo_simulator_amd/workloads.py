"""Synthetic inputs of the benchmark configurations.

thermal_box / elastic_box build the reference benchmarks' boxes and boundary
conditions with their closed-form solutions:
  - thermal: -div(k grad T) = Q on lx x ly x lz (km), T = t_bot on z = 0 and
    t_top on z = lz, insulated sides (benchmarks/thermal_1d_improved_mesh.rs:
    18-52, selection |z - z_b| < 1.0); exact T(z) = t_bot (1 - z/lz) +
    t_top z/lz + Q/(2k) z (lz - z), which P2 elements reproduce.
  - elastic: gravity column in metres, bottom fixed, u_x = 0 on the x faces,
    u_y = 0 on the y faces (benchmarks/elasticity_halfspace_gravity.rs:33-103
    with the bottom fully fixed); exact u_z(z) = -(rho g / M)(H z - z^2/2),
    M = E(1-nu)/((1+nu)(1-2nu)) (:23-27), u_x = u_y = 0.

gaussian_field draws a stationary Gaussian random field on the generator's
hex grid by spectral (FFT) synthesis from seeded white noise (numpy PCG64),
with squared-exponential covariance of correlation length `corr`, standardised
to zero mean and unit variance.  Element e lies in hex e // 6, whose index is
(iz * ny + iy) * nx + ix -- the generator's hex order -- so the field value of
a hex is its six tets' value (the value at the hex centre stands for the
centroid value).

thermal_adjoint_case / elastic_adjoint_case assemble the inversion-gradient
workloads: per-element log-parameter fields, observations from a forward
solve with a second seeded field, and the least-squares misfit.
"""
from __future__ import annotations

import math

import numpy as np


def thermal_box(G, nx, ny, nz, lx=1000.0, ly=1000.0, lz=600.0, k=3.0, q=1e-6,
                t_bot=1280.0, t_top=0.0):
    mesh = G.ImprovedMeshGenerator.generate_cube(nx, ny, nz, lx, ly, lz)
    z = mesh.nodes[:, 2]
    dof = G.DofManager(mesh.num_nodes(), 1)
    bot = np.nonzero(np.abs(z) < 1.0)[0]
    top = np.nonzero((np.abs(z) >= 1.0) & (np.abs(z - lz) < 1.0))[0]
    dof.set_dirichlet_many(bot, t_bot)
    dof.set_dirichlet_many(top, t_top)
    exact = t_bot * (1.0 - z / lz) + t_top * z / lz + (q / (2.0 * k)) * z * (lz - z)
    return mesh, dof, exact


def elastic_mask(xyz, lx, ly):
    """Constrained DOFs (3 per node): bottom all, x faces u_x, y faces u_y."""
    n = len(xyz)
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    mask = np.zeros(3 * n, dtype=bool)
    bot = np.nonzero(np.abs(z) < 1.0)[0]
    for c in range(3):
        mask[3 * bot + c] = True
    xf = np.nonzero((np.abs(x) < 1.0) | (np.abs(x - lx) < 1.0))[0]
    yf = np.nonzero((np.abs(y) < 1.0) | (np.abs(y - ly) < 1.0))[0]
    mask[3 * xf + 0] = True
    mask[3 * yf + 1] = True
    return mask


def elastic_box(G, nx, ny, nz, lx=1.0e6, ly=1.0e6, lz=6.0e5, E=1e11, nu=0.25,
                rho=3300.0, g=9.81):
    mesh = G.ImprovedMeshGenerator.generate_cube(nx, ny, nz, lx, ly, lz)
    xyz = mesh.nodes
    dof = G.DofManager(mesh.num_nodes(), 3)
    dof.set_dirichlet_many(np.nonzero(elastic_mask(xyz, lx, ly))[0], 0.0)
    M = E * (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    z = xyz[:, 2]
    uz = -(rho * g / M) * (lz * z - z * z / 2.0)
    return mesh, dof, uz


def gaussian_field(nx, ny, nz, lx, ly, lz, corr, seed):
    """Standardised Gaussian random field on the (nz, ny, nx) hex grid,
    flattened in the generator's hex order."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.standard_normal((nz, ny, nx))
    kz = 2.0 * np.pi * np.fft.fftfreq(nz, d=lz / nz)
    ky = 2.0 * np.pi * np.fft.fftfreq(ny, d=ly / ny)
    kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=lx / nx)
    k2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    # square root of the squared-exponential spectral density exp(-k^2 l^2 / 2)
    f = np.fft.ifftn(np.fft.fftn(w) * np.exp(-0.25 * k2 * corr * corr)).real
    f = (f - f.mean()) / f.std()
    return f.ravel()


def per_element(hex_values, n_elems):
    return np.repeat(hex_values, 6)[:n_elems]


def conductivity_from_field(f):
    """Monotone map of a standard-normal field to k in (1, 5):
    k = 1 + 4 Phi(f), Phi the standard normal CDF."""
    phi = 0.5 * (1.0 + np.vectorize(math.erf)(f / math.sqrt(2.0)))
    return 1.0 + 4.0 * phi


def thermal_adjoint_case(G, nx=64, ny=64, nz=32, n_obs=4096, seed=2024):
    """log k_e from a Gaussian field (correlation 100 km), observations T at
    n_obs seeded free nodes from a second field's forward solve."""
    mesh, dof, _ = thermal_box(G, nx, ny, nz)
    ne = mesh.num_elements()
    f1 = gaussian_field(nx, ny, nz, 1000.0, 1000.0, 600.0, 100.0, seed)
    f2 = gaussian_field(nx, ny, nz, 1000.0, 1000.0, 600.0, 100.0, seed + 1)
    logk = np.log(per_element(conductivity_from_field(f1), ne))
    logk_true = np.log(per_element(conductivity_from_field(f2), ne))
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    free = np.nonzero(~dof.dirichlet_mask.astype(bool))[0]
    obs = np.sort(rng.choice(free, size=min(n_obs, len(free)), replace=False))
    return mesh, dof, logk, logk_true, obs


def elastic_adjoint_case(G, nx=48, ny=48, nz=24, seed=4048):
    """log E_e log-normal (median 1e11 Pa, sigma_log 0.5, iid per element);
    observations u_z at every top-face node from a second draw."""
    mesh, dof, _ = elastic_box(G, nx, ny, nz)
    ne = mesh.num_elements()
    rng = np.random.Generator(np.random.PCG64(seed))
    logE = math.log(1e11) + 0.5 * rng.standard_normal(ne)
    logE_true = math.log(1e11) + 0.5 * rng.standard_normal(ne)
    top = np.nonzero(np.abs(mesh.nodes[:, 2] - 6.0e5) < 1.0)[0]
    obs = 3 * top + 2
    return mesh, dof, logE, logE_true, obs
