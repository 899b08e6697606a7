"""Seeded synthetic workloads (fcc lattices in the crystal3d convention,
Maxwell-Boltzmann velocities) for the single-system, batched and served runs."""
from __future__ import annotations

import math

import numpy as np

from .lattice import ELEMENT_MASS, fcc, maxwell_boltzmann

SIGMA_AR = 3.41
EPS_AR = 0.0103

# Sutton-Chen sets (n, m, epsilon eV, c, a A) for the fcc metals
SUTTON_CHEN = {
    "Cu": dict(z=29, n=9, m=6, e=1.2382e-2, c=39.432, a=3.61),
    "Ni": dict(z=28, n=9, m=6, e=1.5707e-2, c=39.432, a=3.52),
    "Ag": dict(z=47, n=12, m=6, e=2.5415e-3, c=144.41, a=4.09),
}
# Gupta (Cleri-Rosato) sets (A eV, r0 A, p, B eV, q) and lattice constants
GUPTA = {
    "Cu": dict(z=29, A=0.0855, r0=2.556, p=10.960, B=1.2240, q=2.278, a=3.615),
    "Al": dict(z=13, A=0.1221, r0=2.864, p=8.612, B=1.3160, q=2.516, a=4.05),
    "Ni": dict(z=28, A=0.0376, r0=2.491, p=16.999, B=1.0700, q=1.189, a=3.524),
}
# Finnis-Sinclair parameter sets for the served mix: synthetic, not fitted to the
# metals (they bind the fcc lattice at these lattice constants, nothing more)
FINNIS_SINCLAIR = {
    "Cu": dict(z=29, c0=0.6, c1=-0.3, c2=0.04, A=1.2, B=0.5, c=4.9, d=4.6, a=3.61),
    "Al": dict(z=13, c0=0.5, c1=-0.25, c2=0.035, A=1.1, B=0.4, c=5.5, d=5.2, a=4.05),
    "Ni": dict(z=28, c0=0.65, c1=-0.32, c2=0.045, A=1.3, B=0.5, c=4.8, d=4.5, a=3.52),
}
# short-range Buckingham terms for rocksalt NaCl (Na-Na, Cl-Cl, Na-Cl)
BUCKINGHAM_NACL = [
    ("B1 424.1 B2 1.05 Ro 0.3170 cutoff 8.0", 11, 11),
    ("B1 1256.3 B2 2.0 Ro 0.3321 cutoff 8.0", 17, 17),
    ("B1 1254.5 B2 0.0 Ro 0.3249 cutoff 8.0", 11, 17),
]


def fcc_cells(n_target: int):
    """(nx, ny, nz) with 4 nx ny nz close to n_target."""
    c = max(1.0, (n_target / 4.0) ** (1.0 / 3.0))
    nx = max(1, int(round(c)))
    ny = max(1, int(round(math.sqrt(n_target / 4.0 / nx))))
    nz = max(1, int(round(n_target / 4.0 / nx / ny)))
    return nx, ny, nz


def _finish(rng, x, L, z, T, noise, potentials, cm, dt, steps):
    if noise > 0:
        x = np.mod(x + rng.normal(0.0, noise, x.shape), L)
    if np.ndim(z) == 0:
        zz = np.full(x.shape[0], int(z), dtype=np.int32)
    else:
        zz = np.asarray(z, dtype=np.int32)
    m = np.array([ELEMENT_MASS[int(q)] for q in zz])
    v = maxwell_boltzmann(m, T, seed=int(rng.integers(1 << 31)))
    return dict(x=x, box=L, z=zz, potentials=potentials, cm=cm, v=v, dt=dt, steps=steps)


def _system(rng, nx, ny, nz, a, z, T, noise, potentials, cm, dt, steps):
    x, L = fcc(nx, a, ny, nz)
    return _finish(rng, x, L, z, T, noise, potentials, cm, dt, steps)


def rocksalt(nx: int, ny: int, nz: int, a: float):
    """NaCl rocksalt: a simple-cubic grid of spacing a/2 (2nx x 2ny x 2nz
    sites), Na (11) where i+j+k is even, Cl (17) where it is odd."""
    h = a / 2.0
    k, j, i = np.meshgrid(np.arange(2 * nz), np.arange(2 * ny), np.arange(2 * nx), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    x = (idx.astype(np.float64) + 0.25) * h
    z = np.where(idx.sum(axis=1) % 2 == 0, 11, 17).astype(np.int32)
    L = np.array([2 * nx * h, 2 * ny * h, 2 * nz * h])
    return x, L, z


def lj_system(rng, n, rho_star, T, steps, dt=1.0, skin=1.0):
    nx, ny, nz = fcc_cells(n)
    a = (4.0 / (rho_star / SIGMA_AR**3)) ** (1.0 / 3.0)
    pot = ("lennardjones", f"sigma {SIGMA_AR} epsilon {EPS_AR} cutoff 8.5")
    return _system(rng, nx, ny, nz, a, 18, T, 0.02, [pot], f"cutoff 8.5 skin {skin}", dt, steps)


def sc_system(rng, n, metal, T, steps, dt=2.0, skin=0.3):
    p = SUTTON_CHEN[metal]
    nx, ny, nz = fcc_cells(n)
    rc = 2.0 * p["a"]
    pot = ("suttonchen", f"e {p['e']} a {p['a']} n {p['n']} m {p['m']} c {p['c']} cutoff {rc}")
    return _system(rng, nx, ny, nz, p["a"], p["z"], T, 0.01, [pot], f"cutoff {rc} skin {skin}", dt, steps)


def batched_ensemble(seed: int = 3, n_systems: int = 512, steps: int = 1000):
    """512 periodic systems, sizes log-uniform in [256, 16384]: half LJ argon at
    reduced density 0.8-1.0, half Sutton-Chen Cu/Ni/Ag fcc; temperatures
    100-900 K; each runs `steps` velocity-Verlet steps."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_systems):
        n = int(round(math.exp(rng.uniform(math.log(256), math.log(16384)))))
        T = float(rng.uniform(100.0, 900.0))
        if k % 2 == 0:
            out.append(lj_system(rng, n, float(rng.uniform(0.8, 1.0)), min(T, 150.0), steps))
        else:
            metal = ("Cu", "Ni", "Ag")[int(rng.integers(3))]
            out.append(sc_system(rng, n, metal, T, steps))
    return out


def served_jobs(seed: int = 5, n_jobs: int = 64, rate_hz: float = 200.0, min_atoms=500, max_atoms=65536,
                min_steps=100, max_steps=5000):
    """Poisson arrivals of MD jobs: one periodic system each, 500-65536 atoms
    (log-uniform), Gupta or Finnis-Sinclair Cu/Al/Ni fcc or Buckingham rocksalt
    NaCl, 300-1200 K, 100-5000 steps.  Returns (arrival_time_s, job) pairs."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for _ in range(n_jobs):
        t += float(rng.exponential(1.0 / rate_hz))
        n = int(round(math.exp(rng.uniform(math.log(min_atoms), math.log(max_atoms)))))
        T = float(rng.uniform(300.0, 1200.0))
        steps = int(round(math.exp(rng.uniform(math.log(min_steps), math.log(max_steps)))))
        kind = int(rng.integers(3))
        nx, ny, nz = fcc_cells(n)
        if kind == 0:
            metal = ("Cu", "Al", "Ni")[int(rng.integers(3))]
            p = GUPTA[metal]
            rc = 2.0 * p["a"]
            pot = [("gupta", f"A {p['A']} r0 {p['r0']} p {p['p']} B {p['B']} q {p['q']} cutoff {rc}")]
            job = _system(rng, nx, ny, nz, p["a"], p["z"], T, 0.01, pot, f"cutoff {rc} skin 0.3", 2.0, steps)
        elif kind == 1:
            metal = ("Cu", "Al", "Ni")[int(rng.integers(3))]
            p = FINNIS_SINCLAIR[metal]
            pot = [("finnissinclair", " ".join(f"{k} {p[k]}" for k in ("c0", "c1", "c2", "A", "B", "c", "d")))]
            rc = max(p["c"], p["d"])
            job = _system(rng, nx, ny, nz, p["a"], p["z"], T, 0.01, pot, f"cutoff {rc} skin 0.3", 2.0, steps)
        else:
            # 8 ions per conventional cell
            x, L, zz = rocksalt(*fcc_cells(n // 2), 5.64)
            pot = [("buckingham", args, s1, s2) for args, s1, s2 in BUCKINGHAM_NACL]
            job = _finish(rng, x, L, zz, T, 0.01, pot, "cutoff 8.0 skin 0.4", 1.0, steps)
        jobs.append((t, job))
    return jobs
