"""ctypes front end of the FP64 oracle (oracle/lpmd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblpmd_oracle.so")

LJ, FASTLJ, MORSE, BUCKINGHAM, SUTTONCHEN, FINNISSINCLAIR, GUPTA, HARMONIC, LJMOD, FINNISSINCLAIREXT, TABULATED = \
    range(1, 12)
METAL = (SUTTONCHEN, FINNISSINCLAIR, GUPTA, FINNISSINCLAIREXT)
KIND = {
    "lennardjones": LJ,
    "fastlj": FASTLJ,
    "morse": MORSE,
    "buckingham": BUCKINGHAM,
    "suttonchen": SUTTONCHEN,
    "finnissinclair": FINNISSINCLAIR,
    "gupta": GUPTA,
    "harmonic": HARMONIC,
    "lennardjonesMod": LJMOD,
    "finnissinclair-ext": FINNISSINCLAIREXT,
    "tabulatedpair": TABULATED,
}
# keyword order of each plugin (plugins/<name>.cc constructors)
KEYWORDS = {
    "lennardjones": ("sigma", "epsilon", "cutoff"),
    "fastlj": ("sigma", "epsilon", "cutoff", "bins"),
    "morse": ("depth", "a", "re"),
    "buckingham": ("B1", "B2", "Ro", "cutoff"),
    "suttonchen": ("e", "a", "n", "m", "c", "cutoff"),
    "finnissinclair": ("c0", "c1", "c2", "A", "B", "c", "d"),
    "gupta": ("A", "r0", "p", "B", "q", "cutoff"),
    "harmonic": ("k", "a"),
    "lennardjonesMod": ("sigma", "epsilon", "m", "n", "cohesive", "cutoff"),
    "finnissinclair-ext": ("c0", "c1", "c2", "c3", "c4", "A", "B", "c", "d"),
}
DEFAULTS = {"fastlj": {"bins": 500.0}, "lennardjonesMod": {"m": 12.0, "n": 6.0, "cohesive": 1.0}}

FIX_VEL, FIX_POS, FIX_TAGGED = 1, 2, 4  # the ORC_FIX_* flags
FORCEFACTOR = 0.00964853
KIN2EV = 103.64269
KBOLTZMANN = 8.6173422e-05


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (-O2, no FMA contraction)."""
    src = os.path.join(_HERE, "lpmd_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        import subprocess

        subprocess.check_call(
            ["gcc", "-O2", "-std=c11", "-fPIC", "-fno-semantic-interposition", "-ffp-contract=off", "-shared",
             "-o", LIB_PATH, src, "-lm"]
        )
    return LIB_PATH


class _Conf(C.Structure):
    _fields_ = [
        ("n", C.c_long),
        ("L", C.c_double * 3),
        ("per", C.c_int * 3),
        ("x", C.c_void_p),
        ("v", C.c_void_p),
        ("a", C.c_void_p),
        ("z", C.c_void_p),
        ("m", C.c_void_p),
        ("fix", C.c_void_p),
    ]


class _Pot(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("spc_sum", C.c_int),
        ("cutoff", C.c_double),
        ("p", C.c_double * 10),
        ("bins", C.c_long),
        ("etable", C.c_void_p),
        ("fftable", C.c_void_p),
        ("an", C.c_double),
        ("am", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, d, l, i = C.c_void_p, C.c_double, C.c_long, C.c_int
        L.orc_cm_linkedcell.restype = vp
        L.orc_cm_linkedcell.argtypes = [d, i, i, i, i]
        L.orc_cm_verletlist.restype = vp
        L.orc_cm_verletlist.argtypes = [d, l, l]
        L.orc_cm_minimage.restype = vp
        L.orc_cm_minimage.argtypes = [d]
        L.orc_cm_lcbinary.restype = vp
        L.orc_cm_lcbinary.argtypes = [d, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_lcb_auto_n.restype = C.c_int
        L.orc_lcb_auto_n.argtypes = [d, d]
        L.orc_cm_free.argtypes = [vp]
        L.orc_cm_update.argtypes = [vp, vp]
        L.orc_cm_grid.argtypes = [vp, vp]
        L.orc_cm_all_lists.restype = l
        L.orc_cm_all_lists.argtypes = [vp, vp, i, d, vp, l, vp, vp, vp]
        L.orc_pot_init.argtypes = [vp, i, vp, i, i]
        L.orc_pot_init_table.argtypes = [vp, vp, vp, l, d, i, i]
        L.orc_pot_free.argtypes = [vp]
        L.orc_pair_energy.restype = d
        L.orc_pair_energy.argtypes = [vp, d]
        L.orc_pair_force.argtypes = [vp, vp, vp]
        L.orc_rhoij.restype = d
        L.orc_rhoij.argtypes = [vp, d]
        L.orc_embed.restype = d
        L.orc_embed.argtypes = [vp, d]
        L.orc_metal_pairforce.argtypes = [vp, vp, d, vp]
        L.orc_metal_manybodies.argtypes = [vp, vp, d, d, d, vp]
        L.orc_update_forces.argtypes = [vp, vp, i, vp, vp, vp, vp]
        L.orc_md_run.argtypes = [vp, vp, i, vp, d, l, vp, vp, vp]
        L.orc_vv_position.argtypes = [vp, d, vp]
        L.orc_vv_velocity.argtypes = [vp, d, vp]
        L.orc_displacement.argtypes = [vp, vp, vp, vp, vp]
        L.orc_fitted_inside.argtypes = [vp, vp, vp]
        L.orc_kinetic_energy.restype = d
        L.orc_kinetic_energy.argtypes = [vp]
        L.orc_temperature.restype = d
        L.orc_temperature.argtypes = [vp]
        L.orc_momentum.argtypes = [vp, vp]
        L.orc_kinetic_energy_tag.restype = d
        L.orc_kinetic_energy_tag.argtypes = [vp, i]
        L.orc_temperature_tag.restype = d
        L.orc_temperature_tag.argtypes = [vp, i]
        L.orc_momentum_tag.argtypes = [vp, i, vp]
        L.orc_srand48.restype = C.c_ulonglong
        L.orc_srand48.argtypes = [l]
        L.orc_drand48.restype = d
        L.orc_drand48.argtypes = [vp]
        L.orc_init_velocities.argtypes = [vp, vp]
        L.orc_tags.argtypes = [vp, d, d, vp, vp]
        L.orc_atom_energies.argtypes = [vp, vp, vp, vp]
        L.orc_sample_forces.argtypes = [vp, i, vp, d, l, vp, vp]
        L.orc_lc_auto_n.restype = i
        L.orc_lc_auto_n.argtypes = [d, d]
        L.orc_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


class Conf:
    """Atoms in a periodic orthogonal box, FixedSizeParticleSet-style AoS FP64 arrays."""

    def __init__(self, x, L, z=18, mass=None, v=None, per=(1, 1, 1)):
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3).copy()
        n = self.x.shape[0]
        self.L = np.asarray(L, dtype=np.float64).reshape(3).copy()
        self.v = np.zeros((n, 3)) if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3).copy()
        self.a = np.zeros((n, 3))
        self.z = np.broadcast_to(np.asarray(z, dtype=np.int32), (n,)).copy()
        if mass is None:
            from oracle import elements

            mass = elements.mass_of(self.z)
        self.m = np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,)).copy()
        self.per = tuple(per)
        # fixed-atom flags (FIX_*) of the fixedvel / fixedpos tags, or None
        self.fix = None
        # SimulationEngine's velocitiesSet and the drand48 state of its draws
        # (an unseeded drand48 starts at 0x1234ABCD330E)
        self.velocities_set = False
        self.rng48 = C.c_ulonglong(0x1234ABCD330E)
        self._s = _Conf()

    def set_fixed(self, fixedvel=None, fixedpos=None, fixedpos_true=None):
        """Boolean masks of the atoms tagged fixedvel / fixedpos (fixedpos_true:
        which fixedpos tags read "true", default all); no argument clears."""
        if fixedvel is None and fixedpos is None:
            self.fix = None
            return
        fv = np.zeros(self.n, bool) if fixedvel is None else np.asarray(fixedvel, bool)
        fp = np.zeros(self.n, bool) if fixedpos is None else np.asarray(fixedpos, bool)
        ft = fp if fixedpos_true is None else (fp & np.asarray(fixedpos_true, bool))
        self.fix = np.ascontiguousarray(np.where(fv, FIX_VEL, 0) | np.where(ft, FIX_POS, 0) |
                                        np.where(fv | fp, FIX_TAGGED, 0), dtype=np.int32)

    @property
    def tagged(self):
        """Whether any atom carries fixedvel or fixedpos (the tag-aware sums apply)."""
        return self.fix is not None and bool(np.any(self.fix & FIX_TAGGED))

    @property
    def n(self):
        return self.x.shape[0]

    def struct(self):
        s = self._s
        s.n = self.n
        for q in range(3):
            s.L[q] = self.L[q]
            s.per[q] = int(self.per[q])
        s.x, s.v, s.a, s.z, s.m = _p(self.x), _p(self.v), _p(self.a), _p(self.z), _p(self.m)
        s.fix = None if self.fix is None else _p(self.fix)
        return C.byref(s)


class CellManager:
    """linkedcell / verletlist / minimumimage / lcbinary neighbour provider of the oracle."""

    def __init__(self, kind: str, cutoff: float, nx=0, ny=0, nz=0, mode="auto", maxneighbors=80, each=5):
        L = lib()
        self.kind = kind
        if kind == "linkedcell":
            self.h = L.orc_cm_linkedcell(cutoff, nx, ny, nz, 1 if mode in ("auto", "AUTO") else 0)
        elif kind == "verletlist":
            self.h = L.orc_cm_verletlist(cutoff, maxneighbors, each)
        elif kind == "minimumimage":
            self.h = L.orc_cm_minimage(cutoff)
        elif kind == "lcbinary":
            self.h = L.orc_cm_lcbinary(cutoff, nx, ny, nz, 1 if mode in ("auto", "AUTO") else 0)
        else:
            raise ValueError(kind)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_cm_free(self.h)
            self.h = None

    def update(self, conf: Conf):
        _check(lib().orc_cm_update(self.h, conf.struct()))

    def grid(self):
        g = np.zeros(5, dtype=np.int64)
        lib().orc_cm_grid(self.h, _p(g))
        return g

    def lists(self, conf: Conf, full: bool, rcut: float, cap: int | None = None):
        """CSR lists (counts, j, r2, rij) in each atom's own list order."""
        n = conf.n
        counts = np.zeros(n, dtype=np.int64)
        cap = cap or max(1024, n * 64)
        while True:
            js = np.zeros(cap, dtype=np.int64)
            r2 = np.zeros(cap)
            rij = np.zeros((cap, 3))
            tot = lib().orc_cm_all_lists(self.h, conf.struct(), int(full), rcut, _p(counts), cap, _p(js), _p(r2), _p(rij))
            if tot == -1:
                raise OracleError(lib().orc_last_error().decode())
            if tot >= 0:
                return counts, js[:tot], r2[:tot], rij[:tot]
            cap = -tot


def read_pair_table(path: str):
    """The table of plugins/tabulatedpair.cc:34-79: rows of exactly three
    whitespace-separated fields (r, E, ff) not starting with '#', in file
    order; the cutoff is the last r larger than every r before it.  Fields are
    converted like atof (a leading number, else 0)."""
    import re

    def atof(t):
        m = re.match(r"[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?)", t)
        return float(m.group(0)) if m else 0.0

    e, ff = [], []
    cutoff, previo = 0.0, 0.0
    with open(path) as f:
        for line in f:
            w = line.replace("\t", " ").split()
            if len(w) != 3 or w[0].startswith("#"):
                continue
            r = atof(w[0])
            if r > previo:
                cutoff = previo = r
            e.append(atof(w[1]))
            ff.append(atof(w[2]))
    return np.array(e), np.array(ff), cutoff


class Pot:
    def __init__(self, name: str, s1: int = -1, s2: int = -1, **kw):
        self.name = name
        self.s = _Pot()
        if name == "tabulatedpair":
            e, ff, cutoff = read_pair_table(kw["file"])
            e = np.ascontiguousarray(e, dtype=np.float64)
            ff = np.ascontiguousarray(ff, dtype=np.float64)
            _check(lib().orc_pot_init_table(C.byref(self.s), _p(e), _p(ff), len(e), cutoff, s1, s2))
            return
        vals = dict(DEFAULTS.get(name, {}))
        vals.update(kw)
        params = np.zeros(10)
        for q, k in enumerate(KEYWORDS[name]):
            params[q] = float(vals[k])
        self.name = name
        self.s = _Pot()
        _check(lib().orc_pot_init(C.byref(self.s), KIND[name], _p(params), s1, s2))

    def __del__(self):
        if getattr(self, "s", None) is not None:
            lib().orc_pot_free(C.byref(self.s))

    @property
    def cutoff(self):
        return self.s.cutoff

    @property
    def metal(self):
        return self.s.kind in METAL

    def pair_energy(self, r):
        return np.array([lib().orc_pair_energy(C.byref(self.s), float(x)) for x in np.atleast_1d(r)])

    def pair_force(self, rij):
        rij = np.ascontiguousarray(np.atleast_2d(rij), dtype=np.float64)
        out = np.zeros_like(rij)
        for k in range(rij.shape[0]):
            lib().orc_pair_force(C.byref(self.s), _p(rij[k]), _p(out[k]))
        return out

    def rhoij(self, r):
        return np.array([lib().orc_rhoij(C.byref(self.s), float(x)) for x in np.atleast_1d(r)])

    def embed(self, rho):
        return np.array([lib().orc_embed(C.byref(self.s), float(x)) for x in np.atleast_1d(rho)])

    def metal_forces(self, rij, rhoi, rhoj):
        """(PairForce, ManyBodies) for each displacement, with r and r-hat formed as
        metalpotential.cc:146-149 forms them."""
        rij = np.ascontiguousarray(np.atleast_2d(rij), dtype=np.float64)
        pf = np.zeros_like(rij)
        mb = np.zeros_like(rij)
        for k in range(rij.shape[0]):
            v = rij[k]
            r = np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
            nrm = np.ascontiguousarray(v / r)
            lib().orc_metal_pairforce(C.byref(self.s), _p(nrm), r, _p(pf[k]))
            lib().orc_metal_manybodies(C.byref(self.s), _p(nrm), float(np.atleast_1d(rhoi)[k]), float(np.atleast_1d(rhoj)[k]), r, _p(mb[k]))
        return pf, mb


def _pot_array(pots):
    arr = (_Pot * max(1, len(pots)))()
    for k, p in enumerate(pots):
        arr[k] = p.s
    return arr


def update_forces(conf: Conf, cm: CellManager, pots):
    """CombinedPotential::UpdateForces: fills conf.a; returns (epot, virial, stress[3,3])."""
    arr = _pot_array(pots)
    ep = np.zeros(1)
    vir = np.zeros(1)
    S = np.zeros(9)
    _check(lib().orc_update_forces(conf.struct(), cm.h, len(pots), arr, _p(ep), _p(vir), _p(S)))
    return float(ep[0]), float(vir[0]), S.reshape(3, 3)


def sample_forces(conf: Conf, pots, rlist: float, idx):
    """The accelerations CombinedPotential::UpdateForces gives the atoms idx
    when every pair within rlist is listed by its minimum image (verletlist /
    minimumimage sets); (len(idx), 3).  For boxes too large for a full
    evaluation: each sampled atom's neighbours come from a cell grid."""
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = np.zeros((len(idx), 3))
    arr = _pot_array(pots)
    _check(lib().orc_sample_forces(conf.struct(), len(pots), arr, float(rlist), len(idx), _p(idx), _p(out)))
    return out


def atom_energies(conf: Conf, cm: CellManager, pot):
    """Potential::AtomEnergy of every atom (pairpotential.cc:21-33,
    metalpotential.cc:47-60), the cell manager updated first."""
    cm.update(conf)
    e = np.zeros(conf.n)
    _check(lib().orc_atom_energies(conf.struct(), cm.h, C.byref(pot.s), _p(e)))
    return e


def md_run(conf: Conf, cm: CellManager, pots, dt: float, steps: int):
    """The lpmd application loop with velocityverlet; returns per-step (epot, ekin, virial)."""
    arr = _pot_array(pots)
    ep = np.zeros(steps)
    ek = np.zeros(steps)
    vr = np.zeros(steps)
    _check(lib().orc_md_run(conf.struct(), cm.h, len(pots), arr, dt, steps, _p(ep), _p(ek), _p(vr)))
    return ep, ek, vr


def fitted_inside_all(x, L):
    """OrthogonalCell::FittedInside (orthogonalcell.h:102-114) on every row of
    x (n, 3), the arithmetic of orc_fitted_inside: outside [0, L] the position
    moves by L floor(x / L); x == L stays."""
    x = np.array(x, dtype=np.float64)
    L = np.asarray(L, dtype=np.float64)
    out = (x < 0.0) | (x > L)
    x[out] = (x - L * np.floor(x / L))[out]
    return x


def leapfrog_run(conf: Conf, forces, dt: float, steps: int, after_step=None):
    """Leapfrog::Initialize then `steps` OneStepIntegrator::Advance calls
    (plugins/leapfrog.cc:58-77, liblpmd/simulation/onestepintegrator.cc:13-46).
    forces(c) fills c.a for configuration c and returns its potential energy
    (update_forces with a cell manager and potentials, or replayed tables).
    Set-up: the old configuration FittedInside(x - v dt) with velocities v - a dt
    (Integrator::GoBack, integrator.cc:31-42), forces there, and half-step
    velocities vh = (v - a dt) + (a_old * 0.5) dt.  Each step: forces at x,
    vhalf = vh + a dt, x = FittedInside(x + vhalf dt), v = 0.5 (vhalf + vh),
    vh = vhalf.  Returns the per-step potential energy (at the positions the
    step's forces saw) and kinetic energy (of the velocities after the step)."""
    L = np.asarray(conf.L, dtype=np.float64)
    old = Conf(fitted_inside_all(conf.x - conf.v * dt, L), L, z=conf.z, mass=conf.m, v=conf.v - conf.a * dt)
    forces(old)
    vh = old.v + old.a * 0.5 * dt
    ep = np.zeros(steps)
    ek = np.zeros(steps)
    # OneStepIntegrator::Advance's tags (onestepintegrator.cc:18-38): fixedvel
    # atoms get a = 0 after the force call, fixedpos ("true") atoms are not advanced
    fv = None if conf.fix is None else (conf.fix & FIX_VEL) != 0
    mv = slice(None) if conf.fix is None else (conf.fix & FIX_POS) == 0
    for s in range(steps):
        ep[s] = forces(conf)
        if fv is not None:
            conf.a[fv] = 0.0
        vn = vh + conf.a * dt
        conf.x[mv] = fitted_inside_all(conf.x + vn * dt, L)[mv]
        conf.v[mv] = (0.5 * (vn + vh))[mv]
        vh[mv] = vn[mv]
        if after_step:
            after_step(s + 1, conf)  # the modifiers of step s + 1 (lpmd/lpmd.cc:92-93)
        ek[s] = kinetic_energy_tag(conf, conf.tagged)
    return ep, ek


def beeman_run(conf: Conf, forces, dt: float, steps: int, after_step=None):
    """Beeman::Initialize then `steps` TwoStepIntegrator::Advance calls
    (plugins/beeman.cc:55-84, liblpmd/simulation/twostepintegrator.cc:12-68).
    forces(c) fills c.a for configuration c and returns its potential energy.
    Set-up: forces at the old configuration FittedInside(x - v dt)
    (Integrator::GoBack, integrator.cc:31-42) give a_old.  Each step: forces
    at x (a_n); aux = a_old; x = FittedInside(x + v dt + (2/3) a_n dt^2
    - (1/6) a_old dt^2); a_old = a_n; forces at the new x (a_n+1);
    v += (1/3) a_n+1 dt + (5/6) a_old dt - (1/6) aux dt.  Returns the per-step
    potential energy and kinetic energy at the end of each step."""
    L = np.asarray(conf.L, dtype=np.float64)
    old = Conf(fitted_inside_all(conf.x - conf.v * dt, L), L, z=conf.z, mass=conf.m, v=conf.v - conf.a * dt)
    forces(old)
    a_old = old.a.copy()
    aux = np.zeros_like(a_old)
    ep = np.zeros(steps)
    ek = np.zeros(steps)
    # TwoStepIntegrator::Advance's tags (twostepintegrator.cc:17-65): fixedvel
    # atoms get a = 0 after each force call, fixedpos ("true") atoms skip
    # AdvancePosition and AdvanceVelocity (so their history stays as it was)
    fv = None if conf.fix is None else (conf.fix & FIX_VEL) != 0
    mv = np.ones(conf.n, bool) if conf.fix is None else (conf.fix & FIX_POS) == 0
    for s in range(steps):
        forces(conf)
        if fv is not None:
            conf.a[fv] = 0.0
        aux[mv] = a_old[mv]
        xn = conf.x + conf.v * dt + (2.0 / 3.0) * conf.a * dt * dt - (1.0 / 6.0) * a_old * dt * dt
        a_old[mv] = conf.a[mv]
        conf.x[mv] = fitted_inside_all(xn, L)[mv]
        ep[s] = forces(conf)
        if fv is not None:
            conf.a[fv] = 0.0
        vn = conf.v + (1.0 / 3.0) * conf.a * dt + (5.0 / 6.0) * a_old * dt - (1.0 / 6.0) * aux * dt
        conf.v[mv] = vn[mv]
        if after_step:
            after_step(s + 1, conf)
        ek[s] = kinetic_energy_tag(conf, conf.tagged)
    return ep, ek


def nosehoover_run(conf: Conf, forces, dt: float, steps: int, fmass: float, temp: float,
                   kboltzmann: float = 0.00008617, after_step=None):
    """NoseHoover::Initialize then `steps` TwoStepIntegrator::Advance calls
    (plugins/nosehoover.cc:62-96, liblpmd/simulation/twostepintegrator.cc:12-68).
    forces(c) fills c.a and returns the potential energy.  Set-up: forces at
    the old configuration FittedInside(x - v dt) give old.a; friction = 0.
    AdvancePosition, atom by atom in order: aux = old.a;
    friction += ((3 N / Q) k_B (T - temp)) dt with T the Temperature of the
    current velocities (k_B the session's "0.00008617"); newacc = a - v friction;
    a = newacc; x = FittedInside(x + v dt + (2/3) newacc dt^2 - (1/6) old.a dt^2);
    old.a = newacc.  After the second force call AdvanceVelocity:
    v += (1/3) a dt + (5/6) old.a dt - (1/6) aux dt.  fixedvel atoms get a = 0
    after each force call; fixedpos ("true") atoms skip both updates, so the
    friction grows once per atom that moves.  after_step(k, conf) runs the
    modifiers of step k after it.  Returns the per-step potential and kinetic
    energy."""
    L = np.asarray(conf.L, dtype=np.float64)
    old = Conf(fitted_inside_all(conf.x - conf.v * dt, L), L, z=conf.z, mass=conf.m, v=conf.v - conf.a * dt)
    forces(old)
    old_a = old.a.copy()
    aux = np.zeros_like(old_a)
    friction = 0.0
    n = conf.n
    fv = None if conf.fix is None else (conf.fix & FIX_VEL) != 0
    mv = np.ones(n, bool) if conf.fix is None else (conf.fix & FIX_POS) == 0
    ep = np.zeros(steps)
    ek = np.zeros(steps)
    for s in range(steps):
        forces(conf)
        if fv is not None:
            conf.a[fv] = 0.0
        T = temperature(conf)  # the velocities do not change during the position loop
        incr = ((3.0 * n / fmass) * kboltzmann * (T - temp) * dt)
        fr = np.zeros(n)
        for i in range(n):  # the friction grows once per moving atom, in atom order
            if mv[i]:
                friction += incr
                fr[i] = friction
        aux[mv] = old_a[mv]
        newacc = conf.a[mv] - conf.v[mv] * fr[mv][:, None]
        conf.a[mv] = newacc
        xn = conf.x[mv] + conf.v[mv] * dt + (2.0 / 3.0) * newacc * dt * dt - (1.0 / 6.0) * old_a[mv] * dt * dt
        old_a[mv] = newacc
        conf.x[mv] = fitted_inside_all(xn, L)
        ep[s] = forces(conf)
        if fv is not None:
            conf.a[fv] = 0.0
        conf.v[mv] = (conf.v[mv] + (1.0 / 3.0) * conf.a[mv] * dt + (5.0 / 6.0) * old_a[mv] * dt
                      - (1.0 / 6.0) * aux[mv] * dt)
        if after_step:
            after_step(s + 1, conf)  # the modifiers of step s + 1 (lpmd/lpmd.cc:92-93)
        ek[s] = kinetic_energy_tag(conf, conf.tagged)
    return ep, ek


def vv_position(conf: Conf, dt: float):
    aold = np.zeros((conf.n, 3))
    lib().orc_vv_position(conf.struct(), dt, _p(aold))
    return aold


def vv_velocity(conf: Conf, dt: float, aold: np.ndarray):
    aold = np.ascontiguousarray(aold, dtype=np.float64)
    lib().orc_vv_velocity(conf.struct(), dt, _p(aold))


def displacement(L, a, b, per=(1, 1, 1)):
    L = np.asarray(L, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    per = np.asarray(per, dtype=np.int32)
    d = np.zeros(3)
    lib().orc_displacement(_p(L), _p(per), _p(a), _p(b), _p(d))
    return d


def fitted_inside(L, x, per=(1, 1, 1)):
    L = np.asarray(L, dtype=np.float64)
    x = np.array(x, dtype=np.float64)
    per = np.asarray(per, dtype=np.int32)
    lib().orc_fitted_inside(_p(L), _p(per), _p(x))
    return x


def kinetic_energy(conf: Conf):
    return lib().orc_kinetic_energy(conf.struct())


def temperature(conf: Conf):
    return lib().orc_temperature(conf.struct())


def set_temperature(conf: Conf, temp: float, tag=None):
    """SimulationEngine::SetTemperature(temp, tag) (liblpmd/simulation/simulationbuilder.cc:53-71):
    InitVelocities first while no velocities were set; then every velocity
    times sqrt(temp / T_now), T_now the Temperature of the sequential kinetic
    sum (properties.h:22-53).  tag (default: whether any atom is tagged, as
    tempscaling.cc:70 passes it): tagged atoms are neither counted nor scaled."""
    if not conf.velocities_set:
        init_velocities(conf)
    tag = conf.tagged if tag is None else tag
    ti = lib().orc_temperature_tag(conf.struct(), int(tag))
    f = np.sqrt(temp / ti)
    if tag and conf.fix is not None:
        keep = (conf.fix & FIX_TAGGED) != 0
        conf.v[~keep] *= f
    else:
        conf.v *= f


def seed48(conf: Conf, seed: int):
    """srand48(seed) for the configuration's drand48 draws."""
    conf.rng48 = C.c_ulonglong(lib().orc_srand48(seed))


def init_velocities(conf: Conf):
    """SimulationEngine::InitVelocities (simulationbuilder.cc:167-185) from
    the configuration's drand48 stream; marks the velocities as set."""
    lib().orc_init_velocities(conf.struct(), C.byref(conf.rng48))
    conf.velocities_set = True


def drand48(state: C.c_ulonglong) -> float:
    return lib().orc_drand48(C.byref(state))


def tags(conf: Conf, epot: float, virial: float, stress) -> np.ndarray:
    """SetSimulationTags (simulationbuilder.cc:84-124): out[28] in the
    lpmdx_observables layout."""
    out = np.zeros(28)
    S = np.ascontiguousarray(np.asarray(stress, dtype=np.float64).reshape(9))
    lib().orc_tags(conf.struct(), float(epot), float(virial), _p(S), _p(out))
    return out


class Berendsen:
    """The berendsen modifier (plugins/berendsen.cc:14-80) as the lpmd
    application runs it after each step (lpmd/application.cc:404-459): the
    first active step arms it (stop = step + tau / dt, each = 1), the steps up
    to stop scale every velocity by sqrt(1 + (dt / tau) (T_set / T - 1)) with
    T the sequential Temperature, the step after disarms it (each restored)."""

    def __init__(self, frm, to, tau=400.0, start=0, end=-1, each=1):
        self.frm, self.to, self.tau = frm, to, tau
        self.st = Stepper(start, end, each)
        self.stop = -1
        self.old_each = each

    def apply(self, k, conf, dt):
        """Returns True when step k scaled the velocities."""
        if not self.st.active(k):
            return False
        set_temp = self.st.value(k, self.frm, self.to)
        if self.stop == -1:
            self.stop = int(float(k) + self.tau / dt)
            self.old_each = self.st.each
            self.st.each = 1
            return False
        if k <= self.stop:
            ti = temperature(conf)
            xi = np.sqrt(1.0 + (float(dt) / self.tau) * (set_temp / ti - 1.0))
            conf.v *= xi
            return True
        self.stop = -1
        self.st.each = self.old_each
        return False


class Stepper:
    """lpmd::Stepper (liblpmd/lpmd/stepper.h:14-35): when a modifier runs and
    the value it ramps to."""

    def __init__(self, start=0, end=-1, each=1):
        self.start, self.end, self.each = start, end, each

    def active(self, k):
        if k < self.start:
            return False
        if self.end != -1 and k > self.end:
            return False
        return (k - self.start) % self.each == 0

    def value(self, k, frm, to):
        if self.start == self.end:
            return frm
        return frm + float(k - self.start) / float(self.end - self.start) * (to - frm)


def momentum(conf: Conf, tag: bool = False):
    """Momentum(atoms, tag), properties.h:64-81."""
    p = np.zeros(3)
    lib().orc_momentum_tag(conf.struct(), int(tag), _p(p))
    return p


def kinetic_energy_tag(conf: Conf, tag: bool):
    """KineticEnergy(atoms, tag), properties.h:22-37."""
    return lib().orc_kinetic_energy_tag(conf.struct(), int(tag))


def temperature_tag(conf: Conf, tag: bool):
    """Temperature(atoms, tag), properties.h:39-53."""
    return lib().orc_temperature_tag(conf.struct(), int(tag))


def lc_auto_n(edge: float, cutoff: float) -> int:
    return lib().orc_lc_auto_n(edge, cutoff)


def lcb_auto_n(edge: float, cutoff: float) -> int:
    """lcbinary's automatic subdivision count (plugins/lcbinary.cc:88-145)."""
    return lib().orc_lcb_auto_n(edge, cutoff)
