"""Synthetic problem definitions (mesh + fields + boundary conditions) for the
heat-conduction and low-Mach workloads.

Every random draw comes from the machine-independent LCG of the native
generator, so a (size, seed) pair defines the same inputs on every platform.
"""
import numpy as np

from . import lattice, lognormal_field, normal_field

PI = np.pi


def lattice_ijk(n):
    """Lattice indices (i, j, k) of every node, node index = i + (n+1)(j + (n+1)k)."""
    n1 = n + 1
    idx = np.arange(n1 ** 3)
    return idx % n1, (idx // n1) % n1, idx // (n1 * n1)


def heat_problem(n, jitter=0.1, sigma=0.5, seed=1234):
    """Steady heat conduction on the unit cube: jittered lattice, lognormal(0, sigma)
    nodal conductivity, Dirichlet T = 0 at x = 0 and T = 1 at x = 1, adiabatic
    elsewhere, initial T = 0."""
    coords, conn = lattice(n, jitter=jitter, seed=seed)
    nn = coords.shape[0]
    k = lognormal_field(nn, sigma, seed + 1)
    i, _, _ = lattice_ijk(n)
    bc_nodes = np.nonzero((i == 0) | (i == n))[0].astype(np.int32)
    bc_values = np.zeros(nn)
    bc_values[i == n] = 1.0
    return dict(n=n, coords=coords, conn=conn, k=k, T=np.zeros(nn), bc_nodes=bc_nodes, bc_values=bc_values)


def taylor_green(coords):
    """3-D Taylor-Green vortex velocity, pressure and pressure gradient on [0,1]^3."""
    x, y, z = coords[:, 0], coords[:, 1], coords[:, 2]
    a = 2.0 * PI
    u = np.stack([np.sin(a * x) * np.cos(a * y) * np.cos(a * z),
                  -np.cos(a * x) * np.sin(a * y) * np.cos(a * z),
                  np.zeros_like(x)], axis=1)
    cz = np.cos(2 * a * z) + 2.0
    p = (np.cos(2 * a * x) + np.cos(2 * a * y)) * cz / 16.0
    gp = np.stack([-(a / 8.0) * np.sin(2 * a * x) * cz,
                   -(a / 8.0) * np.sin(2 * a * y) * cz,
                   -(a / 8.0) * (np.cos(2 * a * x) + np.cos(2 * a * y)) * np.sin(2 * a * z)], axis=1)
    return u, p, gp


def lowmach_problem(n, jitter=0.1, noise=0.05, rho_sigma=0.0, mu=1e-3, seed=4321, stretch=1.0, max_aspect=0.0):
    """Low-Mach momentum + continuity inputs: Taylor-Green velocity plus N(0, noise^2)
    per component, analytic pressure and its gradient as Gpdx, density 1 (or
    lognormal(0, rho_sigma)), constant viscosity."""
    coords, conn = lattice(n, jitter=jitter, stretch=stretch, max_aspect=max_aspect, seed=seed)
    nn = coords.shape[0]
    u, p, gp = taylor_green(coords)
    if noise > 0:
        u = u + normal_field(3 * nn, noise, seed + 7).reshape(nn, 3)
    rho = np.ones(nn) if rho_sigma <= 0 else lognormal_field(nn, rho_sigma, seed + 11)
    return dict(n=n, coords=coords, conn=conn, u=u, p=p, gpdx=gp, rho=rho, mu=np.full(nn, mu), tau=1.0)


def poisson_rhs(nn, seed, zero_mean=True):
    """Unit-variance Gaussian right-hand side, projected to zero mean."""
    b = normal_field(nn, 1.0, seed)
    if zero_mean:
        b = b - b.mean()
    return b
