"""Write the golden fixtures under tests/golden/ from the reference's own code.

Runs in a container that has the reference sources: `make -f oracle/ref.mk`
first (a partial build of the reference's plugins, see oracle/ref.mk), then

    python tests/golden/gen_golden.py

Every fixture holds inputs and the reference's outputs for them: neighbour
lists from plugins/linkedcell.cc, plugins/verletlist.cc, plugins/minimumimage.cc
and plugins/lcbinary.cc, per-pair values of
the eleven potential plugins (a pair table file for tabulatedpair), crystal3d lattices and velocityverlet
trajectories.  Large lists are stored as SHA-256 digests of their exact bytes.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import refcheck as R  # noqa: E402

POTS = {
    # Ar, the workload's LJ parameters
    "lennardjones": dict(sigma=3.41, epsilon=0.0103, cutoff=8.5),
    "fastlj": dict(sigma=3.41, epsilon=0.0103, cutoff=8.5, bins=500),
    # Cu Morse (Girifalco-Weizer style numbers), no cutoff keyword
    "morse": dict(depth=0.3429, a=1.3588, re=2.866),
    # Na-Cl short-range Buckingham
    "buckingham": dict(B1=1254.5, B2=0.0, Ro=0.3249, cutoff=8.0),
    # Sutton-Chen Cu
    "suttonchen": dict(e=1.2382e-2, a=3.61, n=9, m=6, c=39.432, cutoff=7.22),
    # Finnis-Sinclair (Fe-like numbers)
    "finnissinclair": dict(c0=1.2371147, c1=-0.3592185, c2=-0.0385607, A=1.828905, B=1.8, c=3.4, d=3.569745),
    # Gupta Cu (Cleri-Rosato)
    "gupta": dict(A=0.0855, r0=2.556, p=10.960, B=1.2240, q=2.278, cutoff=7.22),
    # the plugin's help example, no cutoff keyword
    "harmonic": dict(k=3.4, a=3.0),
    # Ar sigma/epsilon with exponents 10 and 5 (m/2 integral, n/2 not) and partial cohesion
    "lennardjonesMod": dict(sigma=3.41, epsilon=0.0103, m=10.0, n=5.0, cohesive=0.8, cutoff=8.5),
    # Mo (the plugin's help example) with B = 1.2 so that the rho term's B is exercised
    "finnissinclair-ext": dict(c0=47.98066, c1=-34.09924, c2=5.832293, c3=0.017494, c4=0.020393,
                               A=1.848648, B=1.2, c=3.2572, d=4.1472),
}
METALS = ("suttonchen", "finnissinclair", "gupta", "finnissinclair-ext")
TABLE = "table_morse_cu.dat"


def write_table(path, bins=400, cutoff=7.0):
    """A Cu Morse pair table in the tabulatedpair format (r E ff per row, r up
    to the cutoff), with a comment line, a blank line and a row of another width
    that the reader skips."""
    D, a, re_ = 0.3429, 1.3588, 2.866
    with open(path, "w") as f:
        f.write("# r  E  ff   (Morse Cu, depth 0.3429 a 1.3588 re 2.866)\n\n")
        f.write("0.0 0.0\n")
        for k in range(bins):
            r = cutoff * (k + 1) / bins
            ex = np.exp(-a * (r - re_))
            f.write(f"{r:.10g}\t{D * (1 - ex) ** 2:.12g} {2 * a * D / r * (1 - ex) * ex:.12g}\n")


def args_of(kw):
    return " ".join(f"{k} {v!r}" if isinstance(v, float) else f"{k} {v}" for k, v in kw.items())


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def fcc(nx, a, symbol="Ar"):
    n = 4 * nx**3
    L = np.array([nx * a] * 3)
    return R.crystal(f"type fcc symbol {symbol} nx {nx} ny {nx} nz {nx}", n, L), L


def jitter(x, L, sigma, seed):
    rng = np.random.default_rng(seed)
    return np.mod(x + rng.normal(0.0, sigma, x.shape), L)


def list_case(name, x, L, cm_name, cm_args, rcut, store_full):
    rc = R.RefConf(x, L)
    cm = R.RefCellManager(cm_name, cm_args)
    cm.update(rc, attach=True)
    out = {"x": x, "L": L}
    meta = {"cm": cm_name, "cm_args": cm_args, "rcut": rcut}
    for full in (1, 0):
        counts, js, r2, rij = cm.lists(rc, bool(full), rcut)
        tag = "full" if full else "half"
        out[f"{tag}_counts"] = counts.astype(np.int32)
        if store_full:
            out[f"{tag}_j"] = js.astype(np.int32)
            out[f"{tag}_r2"] = r2
        meta[f"{tag}_digest"] = digest(counts.astype(np.int64), js.astype(np.int64), r2, rij)
    np.savez_compressed(os.path.join(HERE, f"lists_{name}.npz"), **out)
    return meta


def main():
    meta = {"lists": {}, "pots": {}, "crystal": {}, "vv": {}}

    # ---- neighbour lists ------------------------------------------------
    x, L = fcc(8, 5.26)  # 2048 Ar, the LJ argon workload's lattice
    meta["lists"]["ar2048_lc_auto"] = list_case("ar2048_lc_auto", x, L, "linkedcell", "cutoff 8.5 mode auto", 8.5, False)
    xj = jitter(x, L, 0.05, 11)
    meta["lists"]["ar2048j_lc_auto"] = list_case("ar2048j_lc_auto", xj, L, "linkedcell", "cutoff 8.5 mode auto", 8.5, False)
    meta["lists"]["ar2048j_lc_n4"] = list_case("ar2048j_lc_n4", xj, L, "linkedcell", "cutoff 8.5 nx 4 ny 4 nz 4", 8.5, False)
    meta["lists"]["ar2048j_vl"] = list_case("ar2048j_vl", xj, L, "verletlist", "cutoff 8.5 maxneighbors 200 each 5", 8.5, False)
    # 256-atom Cu: auto n = 4 with side 2, the wrapped stencil visits cells twice
    x, L = fcc(4, 3.61)
    xj = jitter(x, L, 0.01, 12)
    meta["lists"]["cu256j_lc_dup"] = list_case("cu256j_lc_dup", xj, L, "linkedcell", "cutoff 7.22 mode auto", 7.22, True)
    # cutoff exactly on the 8th fcc shell (rc = 2a): membership decided by FP64 rounding
    meta["lists"]["cu256_shell_vl"] = list_case("cu256_shell_vl", x, L, "verletlist", "cutoff 7.22 maxneighbors 200 each 5", 7.22, True)
    x6, L6 = fcc(6, 3.61)
    meta["lists"]["cu864_shell_lc"] = list_case("cu864_shell_lc", x6, L6, "linkedcell", "cutoff 7.22 mode auto", 7.22, False)
    # a ragged random box (different edges) with explicit subdivisions
    rng = np.random.default_rng(13)
    Lr = np.array([23.7, 31.1, 19.3])
    xr = rng.uniform(0, 1, (300, 3)) * Lr
    meta["lists"]["rand300_lc_explicit"] = list_case("rand300_lc_explicit", xr, Lr, "linkedcell", "cutoff 6.0 nx 5 ny 6 nz 4", 6.0, True)
    meta["lists"]["rand300_vl"] = list_case("rand300_vl", xr, Lr, "verletlist", "cutoff 6.0 maxneighbors 120 each 5", 6.0, True)
    # minimumimage: one image per pair by brute force; the second radius is
    # wider than half the smallest edge (19.3 / 2)
    meta["lists"]["rand300_mi"] = list_case("rand300_mi", xr, Lr, "minimumimage", "cutoff 0.0", 6.0, True)
    meta["lists"]["rand300_mi_wide"] = list_case("rand300_mi_wide", xr, Lr, "minimumimage", "cutoff 12.0", 12.0, True)
    # lcbinary: subcells of about cutoff/5 holding at most one atom; random
    # atoms collide and grow the grid; the 256-atom Cu box has a stencil wider
    # than the grid along no axis but wider than half of it (entries repeat)
    meta["lists"]["rand300_lcb"] = list_case("rand300_lcb", xr, Lr, "lcbinary", "cutoff 6.0", 6.0, True)
    x, L = fcc(8, 5.26)
    xj = jitter(x, L, 0.05, 11)
    meta["lists"]["ar2048j_lcb_auto"] = list_case("ar2048j_lcb_auto", xj, L, "lcbinary", "cutoff 8.5 mode auto", 8.5, False)
    x, L = fcc(4, 3.61)
    xj = jitter(x, L, 0.01, 12)
    meta["lists"]["cu256j_lcb"] = list_case("cu256j_lcb", xj, L, "lcbinary", "cutoff 7.22 mode auto", 7.22, True)

    # ---- potentials ---------------------------------------------------------
    rng = np.random.default_rng(21)
    write_table(os.path.join(HERE, TABLE))
    pots = dict(POTS)
    pots["tabulatedpair"] = {"file": TABLE}  # resolved in tests/golden/
    for name, kw in pots.items():
        if name == "tabulatedpair":
            rp = R.RefPot(name, f"file {os.path.join(HERE, TABLE)}")
        else:
            rp = R.RefPot(name, args_of(kw))
        rmax = min(rp.cutoff, 9.0)
        rij = rng.normal(0, 1, (256, 3))
        rij *= (rng.uniform(1.9, rmax, 256) / np.linalg.norm(rij, axis=1))[:, None]
        rec = {"rij": rij}
        if name in METALS:
            rhoi = rng.uniform(2.0, 80.0, 256)
            rhoj = rng.uniform(2.0, 80.0, 256)
            ep, rho, fr, pf, mb = rp.metal(rij, rhoi, rhoj)
            rec.update(rhoi=rhoi, rhoj=rhoj, epair=ep, rho=rho, frho=fr, pf=pf, mb=mb)
        else:
            e, f = rp.pair(rij)
            rec.update(e=e, f=f)
        np.savez_compressed(os.path.join(HERE, f"pot_{name}.npz"), **rec)
        meta["pots"][name] = {"params": kw, "cutoff": rp.cutoff}

    # ---- crystal3d -------------------------------------------------------------
    for nx, a in ((3, 5.26), (8, 5.26), (4, 3.61)):
        x, L = fcc(nx, a)
        key = f"fcc{nx}_{a}"
        np.savez_compressed(os.path.join(HERE, f"crystal_{key}.npz"), x=x, L=L)
        meta["crystal"][key] = {"nx": nx, "a": a}

    # ---- velocityverlet with replayed accelerations ----------------------------
    x, L = fcc(3, 5.26)
    rng = np.random.default_rng(31)
    n = x.shape[0]
    v = rng.normal(0, 2e-3, (n, 3))
    steps = 12
    acc = rng.normal(0, 1e-3, (2 * steps, n, 3))
    # one atom pushed across the box edge to exercise FittedInside
    v[0] = [-0.5, 0.7, -40.0]
    rc = R.RefConf(x, L)
    rc.set_velocities(v)
    cm = R.RefCellManager("linkedcell", "cutoff 8.5 nx 4 ny 4 nz 4")
    R.vv_run("dt 1.0", rc, cm, steps, acc.reshape(2 * steps, -1))
    xf, vf, af = rc.get()
    np.savez_compressed(os.path.join(HERE, "vv_ar108.npz"), x0=x, v0=v, L=L, acc=acc, xf=xf, vf=vf, af=af)
    meta["vv"]["ar108"] = {"dt": 1.0, "steps": steps}

    # ---- leapfrog with replayed accelerations (plugins/leapfrog.cc) ------------
    # Initialize's force call at the old configuration, then one per Advance
    x, L = fcc(3, 5.26)
    rng = np.random.default_rng(33)
    n = x.shape[0]
    v = rng.normal(0, 2e-3, (n, 3))
    steps = 12
    acc = rng.normal(0, 1e-3, (steps + 1, n, 3))
    v[0] = [0.5, -0.7, 40.0]  # across the box edge: FittedInside, also in GoBack
    rc = R.RefConf(x, L)
    rc.set_velocities(v)
    cm = R.RefCellManager("linkedcell", "cutoff 8.5 nx 4 ny 4 nz 4")
    R.vv_run("dt 2.0", rc, cm, steps, acc.reshape(steps + 1, -1), integrator="leapfrog")
    xf, vf, af = rc.get()
    np.savez_compressed(os.path.join(HERE, "lf_ar108.npz"), x0=x, v0=v, L=L, acc=acc, xf=xf, vf=vf, af=af)
    meta["lf"] = {"ar108": {"dt": 2.0, "steps": steps}}

    # ---- beeman with replayed accelerations (plugins/beeman.cc) ----------------
    # Initialize's force call at the old configuration, then two per Advance
    x, L = fcc(3, 5.26)
    rng = np.random.default_rng(34)
    n = x.shape[0]
    v = rng.normal(0, 2e-3, (n, 3))
    steps = 12
    acc = rng.normal(0, 1e-3, (2 * steps + 1, n, 3))
    v[0] = [0.5, -0.7, 40.0]  # across the box edge: FittedInside, also in GoBack
    rc = R.RefConf(x, L)
    rc.set_velocities(v)
    cm = R.RefCellManager("linkedcell", "cutoff 8.5 nx 4 ny 4 nz 4")
    R.vv_run("dt 2.0", rc, cm, steps, acc.reshape(2 * steps + 1, -1), integrator="beeman")
    xf, vf, af = rc.get()
    np.savez_compressed(os.path.join(HERE, "bm_ar108.npz"), x0=x, v0=v, L=L, acc=acc, xf=xf, vf=vf, af=af)
    meta["beeman"] = {"ar108": {"dt": 2.0, "steps": steps}}

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
