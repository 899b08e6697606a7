"""Synthetic inputs: crystal3d fcc lattices and seeded Maxwell-Boltzmann velocities."""
from __future__ import annotations

import numpy as np

# ElemMass[Z] of liblpmd/support/elements.cc:20-35, Z = 0 .. 118 (an entry of 0
# means no mass is known)
_MASS = (
    0.00054859, 1.00794, 4.002602, 6.941, 9.01218, 10.811, 12.011, 14.00674,
    15.9994, 18.9984303, 20.1797, 22.989768, 24.305, 26.981539, 28.0855, 30.973762,
    32.066, 35.4527, 39.948, 39.0983, 40.078, 44.95591, 47.88, 50.9415,
    51.9961, 54.93805, 55.847, 58.9332, 58.6934, 63.546, 65.39, 69.723,
    72.61, 74.92159, 78.96, 79.904, 83.8, 85.4678, 87.62, 88.90585,
    91.224, 92.90638, 95.94, 97.9072, 102.07, 102.9055, 106.42, 107.8682,
    112.411, 114.818, 118.71, 121.76, 127.6, 126.90447, 131.29, 132.90543,
    137.327, 138.9055, 140.115, 140.90765, 144.24, 144.9127, 150.36, 151.965,
    157.25, 158.92534, 162.5, 164.93032, 167.26, 168.93421, 173.04, 174.967,
    178.49, 180.9479, 183.84, 186.207, 190.23, 192.22, 195.08, 196.96654,
    200.59, 204.3833, 207.2, 208.98037, 208.9824, 209.9871, 222.0176, 223.0197,
    226.0254, 227.0278, 232.0381, 231.03588, 238.0289, 237.048, 244.0642, 243.0614,
    247.0703, 247.0703, 251.0796, 252.083, 257.0951, 258.1, 259.1009, 262.11,
    261.0, 262.0, 266.0, 264.0, 269.0, 268.0, 269.0, 272.0,
    277.0, 0.0, 289.0, 0.0, 0.0, 0.0, 0.0,
)
ELEMENT_MASS = {z: m for z, m in enumerate(_MASS)}
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
