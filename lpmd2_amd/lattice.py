"""Synthetic inputs: crystal3d fcc lattices and seeded Maxwell-Boltzmann velocities."""
from __future__ import annotations

import numpy as np

# liblpmd/support/elements.cc:20-35 for the species the workloads use
ELEMENT_MASS = {
    1: 1.00794,
    11: 22.989768,
    13: 26.981539,
    17: 35.4527,
    18: 39.948,
    26: 55.847,
    28: 58.6934,
    29: 63.546,
    42: 95.94,
    47: 107.8682,
    79: 196.96654,
}
KIN2EV = 103.64269  # liblpmd/lpmd/properties.h:10
KBOLTZMANN = 8.6173422e-05  # liblpmd/lpmd/properties.h:11


def fcc(nx: int, a: float, ny: int | None = None, nz: int | None = None):
    """fcc lattice with crystal3d's offsets (plugins/crystal3d.cc:84-104): the
    basis (0.51,.5,.5) (0.51,0,1) (1.01,0,.5) (1.01,.5,1) in cell units, i.e. a
    rigid +0.01 a shift along x; positions f*L are not wrapped (x may equal L)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    L = np.array([nx * a, ny * a, nz * a], dtype=np.float64)
    off = np.array([(0.51, 0.5, 0.5), (0.51, 0.0, 1.0), (1.01, 0.0, 0.5), (1.01, 0.5, 1.0)])
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cells = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    frac = (cells[:, None, :] + off[None, :, :]).reshape(-1, 3)
    inv = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    x = (frac * inv) * L
    return x, L


def temperature(v, mass):
    """Temperature with 3N degrees of freedom (liblpmd/lpmd/properties.h:22-53)."""
    ke = 0.0
    for i in range(v.shape[0]):
        ke += 0.5 * mass[i] * (v[i, 0] * v[i, 0] + v[i, 1] * v[i, 1] + v[i, 2] * v[i, 2])
    ke *= KIN2EV
    return (2.0 / 3.0) * ke / (KBOLTZMANN * float(v.shape[0]))


def maxwell_boltzmann(mass, T: float, seed: int):
    """Gaussian velocities (A/fs) with sigma^2 = kB T / (m kin2ev), zero total
    momentum, rescaled so the 3N-DOF temperature is exactly T."""
    mass = np.asarray(mass, dtype=np.float64)
    n = mass.shape[0]
    rng = np.random.default_rng(seed)
    sig = np.sqrt(KBOLTZMANN * T / (mass * KIN2EV))
    v = rng.standard_normal((n, 3)) * sig[:, None]
    p = (v * mass[:, None]).sum(axis=0)
    v -= p / mass.sum()
    ke = 0.5 * (mass[:, None] * v * v).sum() * KIN2EV
    t0 = (2.0 / 3.0) * ke / (KBOLTZMANN * n)
    if t0 > 0:
        v *= np.sqrt(T / t0)
    return v
