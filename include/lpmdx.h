/*
 * lpmdx.h -- C ABI of the MI355X force-and-integration engine for LPMD.
 *
 * Plain C types only: pointers, sizes, doubles.  Every entry point returns an
 * int status (LPMDX_OK = 0, negative on error) and leaves a message for
 * lpmdx_last_error(), following the reference's only FFI precedent
 * (liblpmd/rust/src/lib.rs:26-73: null checks -> -1, non-finite input -> -2).
 *
 * Host arrays are AoS FP64 [3*n] in the reference's FixedSizeParticleSet
 * layout (liblpmd/lpmd/fixedsizeparticleset.h:33-43); atoms keep the caller's
 * index order at this boundary whatever order the device keeps internally.
 *
 * Each entry point names the reference interface a maintainer's plugin would
 * forward to it (file:line).  A system is one Configuration (atoms + orthogonal
 * periodic cell) resident in HBM; calls on one system are serialised on its own
 * HIP stream, calls on different systems may overlap.
 */
#ifndef LPMDX_H
#define LPMDX_H

#ifdef __cplusplus
extern "C" {
#endif

#define LPMDX_ABI_VERSION 2

#if defined(__GNUC__)
#define LPMDX_API __attribute__((visibility("default")))
#else
#define LPMDX_API
#endif

/* status codes */
#define LPMDX_OK 0
#define LPMDX_EINVAL (-1)     /* null pointer, bad size or unknown keyword value */
#define LPMDX_ENONFINITE (-2) /* non-finite input value */
#define LPMDX_EDEVICE (-3)    /* no HIP device, or a HIP call failed */
#define LPMDX_ECAPACITY (-4)  /* a neighbour list reached maxneighbors (verletlist.cc:96-97) */
#define LPMDX_ESTATE (-5)     /* missing cell manager / potential (configuration.h:59-63) */

/* Potential kinds 1 .. LPMDX_NKINDS, as lpmdx_batch_set_lanes / lanes_out index them:
 * 1 lennardjones, 2 fastlj, 3 morse, 4 buckingham, 5 suttonchen, 6 finnissinclair,
 * 7 gupta, 8 harmonic, 9 lennardjonesMod, 10 finnissinclair-ext. */
#define LPMDX_NKINDS 10
#define LPMDX_ENOMEM (-6)     /* device allocation or HBM budget exceeded */

/* potential kinds (plugins/<name>.cc) */
#define LPMDX_LENNARDJONES 1
#define LPMDX_FASTLJ 2
#define LPMDX_MORSE 3
#define LPMDX_BUCKINGHAM 4
#define LPMDX_SUTTONCHEN 5
#define LPMDX_FINNISSINCLAIR 6
#define LPMDX_GUPTA 7

/* observables written by lpmdx_observables, SimulationEngine::SetSimulationTags
 * (liblpmd/simulation/simulationbuilder.cc:84-124); the kinetic energy,
 * temperature and momentum leave out atoms carrying fixedvel or fixedpos
 * when any atom does (properties.h:22-81 with tag = true) */
#define LPMDX_OBS_KINETIC 0   /* kinetic-energy (eV) */
#define LPMDX_OBS_TEMP 1      /* temperature (K) */
#define LPMDX_OBS_PX 2        /* px, py, pz */
#define LPMDX_OBS_POTENTIAL 5 /* potential-energy (eV) */
#define LPMDX_OBS_VIRIAL 6    /* virial (eV) */
#define LPMDX_OBS_PRESSURE 7  /* pressure (MPa) */
#define LPMDX_OBS_STRESS 8    /* 9 entries, s_ab = pressfactor/V * stress[a][b] */
#define LPMDX_OBS_TOTAL 17            /* total-energy (eV) */
#define LPMDX_OBS_MOMENTUM 18         /* momentum |p| */
#define LPMDX_OBS_VOLUME 19           /* volume (A^3) */
#define LPMDX_OBS_VOLUME_PER_ATOM 20  /* volume-per-atom */
#define LPMDX_OBS_PARTICLE_DENSITY 21 /* particle-density (1/A^3) */
#define LPMDX_OBS_DENSITY 22          /* density (g/cm^3) */
#define LPMDX_OBS_VIRIAL_PRESSURE 23  /* virial-pressure (MPa) */
#define LPMDX_OBS_KINETIC_PRESSURE 24 /* kinetic-pressure (MPa) */
#define LPMDX_OBS_CELL_A 25           /* cell-a, cell-b, cell-c (A) */
#define LPMDX_NOBS 28

/* per-atom flags of lpmdx_system_set_fixed */
#define LPMDX_FIX_VEL 1    /* the fixedvel tag: a = 0 after each force evaluation of a step */
#define LPMDX_FIX_POS 2    /* fixedpos with the value "true": not advanced by a step */
#define LPMDX_FIX_TAGGED 4 /* carries fixedvel or fixedpos, whatever the value */

typedef struct lpmdx_system lpmdx_system;

/* Message of the last failed call on this thread. */
LPMDX_API const char* lpmdx_last_error(void);
LPMDX_API int lpmdx_abi_version(void);
/* Number of visible HIP devices (0 on a host without one). */
LPMDX_API int lpmdx_device_count(int* count);

/* Configuration: SimulationBuilder::CreateFixedOrthogonal
 * (liblpmd/simulation/simulationbuilder.cc:321-326) with the cell set to
 * box[q]*e_q.  z[n] are atomic numbers; mass may be NULL, then ElemMass[z]
 * (liblpmd/support/elements.cc:20-35, BasicAtom::operator=, basicatom.h:57-67). */
LPMDX_API int lpmdx_system_create(lpmdx_system** out, long n, const double box[3], const int* z,
                        const double* mass);
LPMDX_API int lpmdx_system_destroy(lpmdx_system* sys);
LPMDX_API long lpmdx_system_size(const lpmdx_system* sys);

/* Positions / velocities / accelerations, AoS [3n] FP64; any pointer may be
 * NULL to leave that field.  BasicAtom::Position()/Velocity()/Acceleration()
 * (liblpmd/lpmd/basicatom.h:37-52). */
LPMDX_API int lpmdx_system_set_state(lpmdx_system* sys, const double* x, const double* v, const double* a);
LPMDX_API int lpmdx_system_get_state(lpmdx_system* sys, double* x, double* v, double* a);

/* Cell manager, as Configuration::SetCellManager (configuration.h:54-57) after
 * PluginManager::LoadPluginAs<CellManager>(name, args) (pluginmanager.h:60):
 *   "linkedcell": keywords cutoff nx ny nz mode warn-outside
 *                 (plugins/linkedcell.cc:17-41; same defaults);
 *   "verletlist": keywords cutoff extcutoff maxneighbors each skin
 *                 (plugins/verletlist.cc:12-29; `skin`, when given, makes the
 *                 list radius cutoff+skin and rebuilds only when the skin is
 *                 exceeded);
 *   "minimumimage": keyword cutoff (plugins/minimumimage.cc:12-19; 0 means the
 *                 largest half cell edge), each potential listed with its own
 *                 cutoff, one image per pair; `skin` as for verletlist;
 *   "lcbinary":   keywords cutoff nx ny nz mode (plugins/lcbinary.cc:9-33):
 *                 linked subcells of at most one atom, grown by two per axis
 *                 while two atoms share one.
 * Runs one UpdateCell, like SetCellManager does. */
LPMDX_API int lpmdx_set_cellmanager(lpmdx_system* sys, const char* name, const char* args);

/* CellManager::UpdateCell (liblpmd/lpmd/cellmanager.h:22). */
LPMDX_API int lpmdx_update_cell(lpmdx_system* sys);

/* CellManager::BuildNeighborList(conf, i, nlist, full, rcut) for every atom i
 * (cellmanager.h:24-25): counts[n]; while room remains (cap entries) the
 * neighbour indices js and r2 values, in the order the reference's manager
 * lists them.  *total receives the number of entries. */
LPMDX_API int lpmdx_neighbor_lists(lpmdx_system* sys, int full, double rcut, long* counts, long cap,
                         long* js, double* r2, long* total);

/* CombinedPotential::Append + Potential::SetValidSpecies(s1, s2)
 * (liblpmd/lpmd/combinedpotential.h, liblpmd/simulation/potential.cc:23), the
 * potential being LoadPluginAs<Potential>(name, args): lennardjones fastlj
 * morse buckingham suttonchen finnissinclair gupta with their plugin keywords
 * (plugins/<name>.cc constructors).  s1 = s2 = -1 is the wildcard. */
LPMDX_API int lpmdx_potential_append(lpmdx_system* sys, const char* name, const char* args, int s1, int s2);
LPMDX_API int lpmdx_potential_clear(lpmdx_system* sys);

/* CombinedPotential::UpdateForces (combinedpotential.h:39-52): zeroes the
 * accelerations, virial and stress, runs UpdateCell and every potential's
 * UpdateForces (pairpotential.cc:35-84, metalpotential.cc:62-180); returns
 * the potential energy, virial and 3x3 stress (row-major stress[a][b]). */
LPMDX_API int lpmdx_update_forces(lpmdx_system* sys, double* epot, double* virial, double stress[9]);

/* One potential's UpdateForces (pairpotential.cc:35-84 for pair potentials,
 * metalpotential.cc:62-180 for metals) on its own: potential k (the k-th
 * appended, from 0) evaluated on the lists of the last UpdateCell (run first
 * if the positions changed since).  acc[3n] receives that potential's
 * accelerations F_k * forcefactor / m in caller order (what the reference adds
 * to each atom's acceleration); its energy, virial and stress go to epot,
 * virial, stress (each may be NULL).  The system's own accelerations are left
 * as they are.  This is the call a plugin forwards one Potential::UpdateForces
 * of a CombinedPotential to. */
LPMDX_API int lpmdx_potential_forces(lpmdx_system* sys, int k, double* acc, double* epot, double* virial,
                                     double stress[9]);

/* Potential::AtomEnergy(conf, i) of potential k for every atom i, e[n] in
 * caller order: PairPotential::AtomEnergy (pairpotential.cc:21-33) sums the
 * pair energy over atom i's full neighbour list (species rule and cutoff as
 * in UpdateForces); MetalPotential::AtomEnergy (metalpotential.cc:47-60) adds
 * F(rho_i) to the pair energies where the species rule applies.  Lists of the
 * last UpdateCell (run first if the positions changed since). */
LPMDX_API int lpmdx_atom_energies(lpmdx_system* sys, int k, double* e);

/* The application loop with the velocityverlet integrator (lpmd/lpmd.cc:70-103,
 * plugins/velocityverlet.cc:55-75, twostepintegrator.cc:12-68): `steps`
 * DoSteps of length dt (fs).  State stays device-resident between calls.
 * Every obs_every steps (0 = never) the potential and kinetic energy are
 * written to epot_t[k], ekin_t[k] (k = step/obs_every - 1; either may be NULL). */
LPMDX_API int lpmdx_velocityverlet(lpmdx_system* sys, double dt, long steps, long obs_every, double* epot_t,
                         double* ekin_t);

/* The leapfrog integrator (plugins/leapfrog.cc:58-77).  lpmdx_leapfrog_init is
 * Leapfrog::Initialize: the old configuration x - v dt, v - a dt
 * (Integrator::GoBack, integrator.cc:31-42), forces there, and the half-step
 * velocities v - a dt + (a_old / 2) dt; it leaves x, v and a as they were.
 * lpmdx_leapfrog then runs `steps` OneStepIntegrator::Advance calls
 * (onestepintegrator.cc:13-46) with that dt; after each, x is the next
 * positions and v the velocities of the step's force evaluation, as in the
 * reference.  Energies every obs_every steps as in lpmdx_velocityverlet.
 * Without a prior lpmdx_leapfrog_init: LPMDX_ESTATE; another dt: LPMDX_EINVAL. */
LPMDX_API int lpmdx_leapfrog_init(lpmdx_system* sys, double dt);
LPMDX_API int lpmdx_leapfrog(lpmdx_system* sys, double dt, long steps, long obs_every, double* epot_t,
                             double* ekin_t);

/* The Beeman integrator (plugins/beeman.cc:55-84).  lpmdx_beeman_init is
 * Beeman::Initialize: forces at the old configuration FittedInside(x - v dt)
 * (Integrator::GoBack, integrator.cc:31-42) become the first step's previous
 * acceleration; x, v and a are left as they were.  lpmdx_beeman then runs
 * `steps` TwoStepIntegrator::Advance calls (twostepintegrator.cc:12-68) with
 * Beeman's AdvancePosition / AdvanceVelocity and that dt, one force
 * evaluation per step as in lpmdx_velocityverlet.  Energies every obs_every
 * steps as there.  Without a prior lpmdx_beeman_init: LPMDX_ESTATE; another
 * dt: LPMDX_EINVAL. */
LPMDX_API int lpmdx_beeman_init(lpmdx_system* sys, double dt);
LPMDX_API int lpmdx_beeman(lpmdx_system* sys, double dt, long steps, long obs_every, double* epot_t,
                           double* ekin_t);

/* The nosehoover integrator (plugins/nosehoover.cc:62-96), a two-step scheme
 * over Beeman's updates with a friction term.  lpmdx_nosehoover_init is
 * NoseHoover::Initialize with the plugin's keywords dt, fmass (Q) and t and
 * the session's kboltzmann (the string "0.00008617" by default,
 * liblpmd/runtime/session.cc:17-23,32, read with double()): the
 * forces at the old configuration FittedInside(x - v dt)
 * (Integrator::GoBack, integrator.cc:31-42) become the first old
 * acceleration, the friction starts at zero; x, v and a are left as they
 * were.  lpmdx_nosehoover then runs `steps` TwoStepIntegrator::Advance calls
 * (twostepintegrator.cc:12-68): each position update adds to the friction
 * D = (3 N / Q) k_B (T - t) dt once per atom in atom order (T the temperature
 * of the velocities, k_B = kboltzmann), so atom i
 * integrates with newacc = a - v f_i, f_i the friction after its r_i + 1
 * increments, r_i the number of moving atoms before it (formed as
 * F + (r_i + 1) D rather than as the reference's sequential sum: the same
 * value to FP64 rounding).  One force evaluation per
 * step as in lpmdx_velocityverlet; energies every obs_every steps as there.
 * A fixedpos "true" atom is not advanced and adds no increment
 * (twostepintegrator.cc:35-45); the tempscaling / berendsen modifiers act
 * after each step, before the next position update reads the temperature.
 * Without a prior init: LPMDX_ESTATE; another dt: LPMDX_EINVAL. */
LPMDX_API int lpmdx_nosehoover_init(lpmdx_system* sys, double dt, double fmass, double t, double kboltzmann);
LPMDX_API int lpmdx_nosehoover(lpmdx_system* sys, double dt, long steps, long obs_every, double* epot_t,
                               double* ekin_t);

/* SimulationEngine::SetTemperature (liblpmd/simulation/simulationbuilder.cc:53-71)
 * on the device.  While no velocities were set (a new system: velocitiesSet
 * is false until InitVelocities or Restore) it first runs InitVelocities
 * (lpmdx_init_velocities), as the reference does; then every velocity is
 * multiplied by sqrt(temp / T_now), T_now from the kinetic energy
 * (properties.h:22-53) summed on the device (FP64, in a fixed order that
 * differs from the reference's sequential sum, so the factor agrees to FP64
 * rounding).  With tagged atoms (lpmdx_system_set_fixed) the sum, the atom
 * count and the scaling leave them out, as SetTemperature(temp, true) does
 * when tempscaling passes HaveAny(fixedvel) || HaveAny(fixedpos)
 * (plugins/tempscaling.cc:70).  A system at rest is left unchanged (the
 * reference would divide by zero). */
LPMDX_API int lpmdx_set_temperature(lpmdx_system* sys, double temp);

/* SimulationEngine::InitVelocities (simulationbuilder.cc:167-185): random
 * directions from the system's drand48 stream (RandomVector,
 * liblpmd/lpmd/vector.h:161-167), speed sqrt(3 k_B (1 K) / m), the mean
 * momentum taken out; marks the velocities as set.  lpmdx_seed48 seeds the
 * stream as srand48(seed) seeds drand48 (lpmd/application.cc:38 seeds it with
 * the time); an unseeded stream starts where an unseeded drand48 does.
 * lpmdx_velocities_set marks the velocities as set (or not) without drawing,
 * as SimulationEngine::Restore does (simulationbuilder.cc:290). */
LPMDX_API int lpmdx_init_velocities(lpmdx_system* sys);
LPMDX_API int lpmdx_seed48(lpmdx_system* sys, long seed);
LPMDX_API int lpmdx_velocities_set(lpmdx_system* sys, int set);

/* The fixedvel / fixedpos atom tags (TwoStepIntegrator::Advance,
 * twostepintegrator.cc:12-68; OneStepIntegrator::Advance,
 * onestepintegrator.cc:13-46): flags[n] in caller order, LPMDX_FIX_* bits
 * (NULL clears).  The integrator entry points then zero a fixedvel atom's
 * acceleration after each force evaluation of a step and leave a fixedpos
 * ("true") atom's position and velocity as they are; forces on the other
 * atoms are unchanged.  Tagged atoms drop out of the tag-aware observables and
 * SetTemperature.  An atom with LPMDX_FIX_VEL or LPMDX_FIX_POS must also
 * carry LPMDX_FIX_TAGGED. */
LPMDX_API int lpmdx_system_set_fixed(lpmdx_system* sys, const int* flags);

/* The tempscaling modifier (plugins/tempscaling.cc:14-63) as the lpmd
 * application runs it (lpmd/lpmd.cc:89-93, application.cc:404-459): after
 * every integrator step whose step number k (steps counted since the system
 * was created, the simulation's CurrentStep) is active for
 * Stepper(start, end, each) (stepper.h:17-25), the velocities are scaled to
 * the temperature ValueAtStep(k, from, to) (stepper.h:27-31), on the device
 * inside lpmdx_velocityverlet / lpmdx_beeman / lpmdx_leapfrog.  end == -1
 * means no end and is allowed only for from == to (the application replaces
 * it by the run's length); each == 0 switches the modifier off. */
LPMDX_API int lpmdx_tempscaling(lpmdx_system* sys, double from, double to, long start, long end, long each);

/* The berendsen modifier (plugins/berendsen.cc:14-80) on the same device
 * schedule as lpmdx_tempscaling (applied after it when both are set): its
 * first active step arms it (stop step = that step + tau / dt of the running
 * integrator, each = 1), each step up to the stop step scales the velocities
 * by sqrt(1 + (dt / tau) (T / T_now - 1)) with T = ValueAtStep(step, from,
 * to), and the step after disarms it.  each == 0 switches it off. */
LPMDX_API int lpmdx_berendsen(lpmdx_system* sys, double from, double to, double tau, long start, long end,
                              long each);

/* Device-side SetSimulationTags (simulationbuilder.cc:84-124, properties.h:22-81):
 * out[LPMDX_NOBS], see LPMDX_OBS_*.  The potential terms are those of the
 * last force evaluation. */
LPMDX_API int lpmdx_observables(lpmdx_system* sys, double out[LPMDX_NOBS]);

/* lpmdx_observables without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * kinetic partials and their sums live in a static device array of the
 * library (512 KiB + 64 B per GPU, held until the process ends), so systems of
 * at most 4194304 atoms are accepted (more: LPMDX_EINVAL); out[] is
 * bit-identical to lpmdx_observables'.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (several threads, or systems on
 * different streams) race and must be serialised by the caller. */
LPMDX_API int lpmdx_observables_noalloc(lpmdx_system* sys, double out[LPMDX_NOBS]);

/* Counters of the system's engine: [0] list builds, [1] force evaluations,
 * [2] steps, [3] list capacity per atom, [4] largest list, [5] lanes per atom,
 * [6] brick edge in cells of the LDS-staged force kernels (0: unstaged),
 * [7] total list entries of the last build, [8] step chunks cut short by the
 * skin test (their speculative steps past the cut ran no kernel). */
LPMDX_API int lpmdx_system_stats(const lpmdx_system* sys, long out[9]);

/* Per-kernel timing with HIP events recorded on the system's own stream:
 * enable (k >= 1) or disable (0); enabling clears the totals.  With k > 1
 * only the launches of every k-th integrator step are bracketed (list builds
 * always are), which keeps the events' own cost out of the other steps.
 * read returns the summed milliseconds and launch counts of [0] force pass
 * (pair or EAM), [1] EAM rho pass, [2] list build (sort + list), [3] other. */
LPMDX_API int lpmdx_system_profile(lpmdx_system* sys, int enable);
LPMDX_API int lpmdx_system_profile_read(lpmdx_system* sys, double ms[4], long launches[4]);

/* Lanes per atom used by the force kernels (1, 2, 4, 8 or 16); 0 = automatic. */
LPMDX_API int lpmdx_system_set_lanes(lpmdx_system* sys, int lanes_per_atom);

/* ------------------------------------------------------------------------
 * Batched engine: many independent periodic systems of uneven size in shared
 * HBM arenas, advanced together (one HIP graph per step; every per-system
 * decision is taken on the device).  No reference counterpart: it is the
 * engine behind running many lpmd jobs (lpmd/lpmd.cc:58-108, one Configuration
 * each) at once.  Each system uses verletlist semantics with a skin (the exact
 * minimum-image set, plugins/verletlist.cc:135-167) and the velocityverlet
 * integrator; nothing a system computes depends on the others.
 * ---------------------------------------------------------------------- */
typedef struct lpmdx_batch lpmdx_batch;

/* Arenas: atom slots, neighbour-list words and cell slots. */
LPMDX_API int lpmdx_batch_create(lpmdx_batch** out, long atom_capacity, long list_words, long cell_slots);
LPMDX_API int lpmdx_batch_destroy(lpmdx_batch* b);
/* Adds one system (AoS FP64 x, v; v may be NULL) with 1-4 potentials of one
 * kind (names[k], args[k] as lpmdx_potential_append, species s1[k], s2[k] or
 * NULL for wildcards) and cell-manager keywords "cutoff C skin S
 * [maxneighbors M]"; it will run `steps` velocity-Verlet steps of dt (fs).
 * *sid receives its id.  LPMDX_ENOMEM when an arena has no room. */
LPMDX_API int lpmdx_batch_add(lpmdx_batch* b, long n, const double box[3], const int* z, const double* mass,
                              const double* x, const double* v, const char* cm_args, int npot,
                              const char** names, const char** args, const int* s1, const int* s2,
                              double dt, long steps, long* sid);
LPMDX_API int lpmdx_batch_remove(lpmdx_batch* b, long sid);
/* Advances every system by up to `steps` steps (a new system's first step is
 * its initial force evaluation); returns when the device is done. */
LPMDX_API int lpmdx_batch_run(lpmdx_batch* b, long steps);
/* State of one system in its caller's atom order; info[8] = steps done, steps
 * left, list builds, started, longest list, list capacity, lanes, overflow. */
LPMDX_API int lpmdx_batch_get(lpmdx_batch* b, long sid, double* x, double* v, double* a, long info[8]);
/* Potential energy, virial and kinetic energy of one system at its last step. */
LPMDX_API int lpmdx_batch_energies(lpmdx_batch* b, long sid, double out[3]);
/* No counterpart in the reference (its runs have no launch shape to tune).
 * Times trial_steps steps for each potential kind on a snapshot of the batch
 * per candidate pair of lanes per atom (1, 2, 4) of its force launches and
 * cells per list radius (1, 2) of its systems' cell grids and keeps the
 * fastest pair, then the same for the tiles of 64 atoms per workgroup (1, 2,
 * 4) of its force and rho launches; lanes_out[kind], cells_out[kind] and
 * tiles_out[kind] for kind 0 .. LPMDX_NKINDS (LPMDX_NKINDS + 1 entries each;
 * any may be NULL).  The snapshot is restored: trajectories are those of the
 * chosen shape. */
LPMDX_API int lpmdx_batch_autotune(lpmdx_batch* b, long trial_steps, int lanes_out[LPMDX_NKINDS + 1],
                                   int cells_out[LPMDX_NKINDS + 1], int tiles_out[LPMDX_NKINDS + 1]);
/* Tiles of 64 atoms per workgroup (1, 2 or 4, default 1) of one kind's force
 * and rho launches.  No counterpart in the reference. */
LPMDX_API int lpmdx_batch_set_tiles(lpmdx_batch* b, int kind, int tiles_per_group);
LPMDX_API int lpmdx_batch_set_lanes(lpmdx_batch* b, int kind, int lanes);
/* Cells per list radius (1 or 2, default 1) of the grids of one kind's
 * systems: 2 applies where five half-radius cells fit along every axis (the
 * list build then scans the rows of a 5 x 5 x 5 stencil trimmed per atom);
 * lists are built again at the next step.  No counterpart in the reference. */
LPMDX_API int lpmdx_batch_set_cells(lpmdx_batch* b, int kind, int per_radius);
/* [0] live systems, [1] atoms, [2] atom capacity, [3] largest free atom run,
 * [4] largest free list run (words), [5] largest free cell run, [6] steps run,
 * [7] system slots. */
LPMDX_API int lpmdx_batch_info(lpmdx_batch* b, long out[8]);
/* HBM bytes one atom slot takes in the per-atom arenas (for admission budgets). */
LPMDX_API long lpmdx_batch_bytes_per_atom(void);
/* Per system slot s < nslots: left[s] = graph steps still to run (its steps
 * plus the initial force evaluation while that is pending), flags[s] bit 0
 * live, bit 1 list overflow, bit 2 initial force evaluation pending. */
LPMDX_API int lpmdx_batch_progress(lpmdx_batch* b, long nslots, long* left, int* flags);

/* ------------------------------------------------------------------------
 * Serving loop: MD jobs (one system, its potentials, dt and step count each)
 * are submitted at any time; a worker thread admits queued jobs into a batch
 * whose arenas are sized from an HBM budget whenever they fit, advances all
 * admitted jobs together in rounds of graph-replayed steps, and retires each
 * job when its steps are done, so arrivals join work already on the device
 * (continuous batching).  Job state stays in HBM from admission to
 * completion.  No reference counterpart: it runs many lpmd invocations
 * (lpmd/lpmd.cc:58-108) as requests.  Results of a job do not depend on the
 * other jobs it ran beside.
 * ---------------------------------------------------------------------- */
typedef struct lpmdx_server lpmdx_server;

/* hbm_budget: bytes for the atom, list and cell arenas; max_round_steps:
 * most steps between two admission points. */
LPMDX_API int lpmdx_server_create(lpmdx_server** out, long hbm_budget, long max_round_steps);
/* Stops the worker (jobs still queued or running are dropped) and frees. */
LPMDX_API int lpmdx_server_destroy(lpmdx_server* s);
/* Queues one job (arguments as lpmdx_batch_add; inputs are copied). */
LPMDX_API int lpmdx_server_submit(lpmdx_server* s, long n, const double box[3], const int* z,
                                  const double* mass, const double* x, const double* v,
                                  const char* cm_args, int npot, const char** names,
                                  const char** args, const int* s1, const int* s2, double dt,
                                  long steps, long* job);
/* Waits up to timeout_s seconds (< 0: no limit) for one job (job >= 0) or for
 * every job submitted so far (job < 0); *done = 1 when it finished. */
LPMDX_API int lpmdx_server_wait(lpmdx_server* s, long job, double timeout_s, int* done);
/* Final x, v, a of a finished job in its atom order (each may be NULL) and
 * out[8] = potential energy, virial, kinetic energy, seconds queued (submit to
 * admission), seconds to completion (submit to done), list builds, steps
 * run, status (0 done, else the LPMDX_E* code the job failed with).  Returns
 * LPMDX_ESTATE while the job is unfinished. */
LPMDX_API int lpmdx_server_result(lpmdx_server* s, long job, double* x, double* v, double* a, double out[8]);
/* Frees the stored result of a finished job. */
LPMDX_API int lpmdx_server_release(lpmdx_server* s, long job);
/* out[LPMDX_NSERVER_STATS] = jobs done, jobs failed, p50 and p99 completion
 * latency (s), mean latency (s), atom-steps done, seconds from the first
 * submission to the last completion, rounds, most jobs resident at once, jobs
 * queued, jobs running, graph steps replayed, then the worker's seconds spent
 * admitting jobs, running steps and retiring jobs, and 0. */
#define LPMDX_NSERVER_STATS 16
LPMDX_API int lpmdx_server_stats(lpmdx_server* s, double out[LPMDX_NSERVER_STATS]);
/* Lanes per atom of one potential kind's force launch (1, 2 or 4). */
LPMDX_API int lpmdx_server_set_lanes(lpmdx_server* s, int kind, int lanes);

/* ---- snapshots and configuration files -------------------------------------
 * Cells are nine numbers, the three cell vectors row by row (the reference's
 * cell[i][q]); the writers take orthogonal cells.  Positions at this boundary
 * are Cartesian and in the caller's atom order, as everywhere above. */

/* A snapshot of a system's state for checkpoints and trajectory files.
 * lpmdx_snapshot_take enqueues the copy and returns at once: x, v and a are
 * gathered on the system's stream into the snapshot's device buffer, then
 * copied to its pinned host buffer on a stream of its own, so the system's
 * next steps overlap the copy.  lpmdx_snapshot_wait (and info / state, which
 * wait first) block until it has landed.  Taking a new snapshot into the same
 * object waits for the previous one.  No reference counterpart: it feeds the
 * writers below without stopping the step loop. */
typedef struct lpmdx_snapshot lpmdx_snapshot;
LPMDX_API int lpmdx_snapshot_create(lpmdx_snapshot** out);
LPMDX_API int lpmdx_snapshot_destroy(lpmdx_snapshot* snap);
LPMDX_API int lpmdx_snapshot_take(lpmdx_system* sys, lpmdx_snapshot* snap, long step);
LPMDX_API int lpmdx_snapshot_wait(lpmdx_snapshot* snap);
LPMDX_API int lpmdx_snapshot_info(lpmdx_snapshot* snap, long* n, long* step, double cell[9]);
/* z[n], x / v / a [3n] and the fixed-atom flags [n] (0 where none were set);
 * any pointer may be NULL. */
LPMDX_API int lpmdx_snapshot_state(lpmdx_snapshot* snap, int* z, double* x, double* v, double* a, int* flags);

/* SimulationEngine::Dump (liblpmd/simulation/simulationbuilder.cc:195-239):
 * the step, the cell rows, then per atom "Z x y z vx vy vz ax ay az", a colour
 * line ("false": colours are not kept) and a tag line, numbers in the
 * stream's default format (%g, six significant digits, as the reference
 * writes them).  The tags written are the fixed-atom flags (may be NULL):
 * fixedvel true, fixedpos true, or fixedpos false for an atom that is only
 * tagged. */
LPMDX_API int lpmdx_dump_write(const char* path, long step, const double cell[9], long n, const int* z,
                               const double* x, const double* v, const double* a, const int* flags);
/* SimulationEngine::Restore (simulationbuilder.cc:241-290): *n receives the
 * atom count, *step and cell (may be NULL) the step and cell; the arrays
 * (each may be NULL) are filled only when cap >= *n, so a first call with
 * cap 0 sizes them.  Colours are skipped; fixedvel / fixedpos tags become
 * LPMDX_FIX_* flags. */
LPMDX_API int lpmdx_dump_read(const char* path, long* step, double cell[9], long* n, long cap, int* z, double* x,
                              double* v, double* a, int* flags);

/* The lpmd / zlp configuration format (plugins/lpmd.cc).  Writing
 * (LPMDFormat::WriteHeader / WriteCell, lpmd.cc:525-642): args are the
 * plugin's keywords level (0 positions, 1 + velocities, 2 + accelerations),
 * type (lpmd or zlp; a file name ending in "lpmd" or ".zlp" decides, as
 * there), blocksize, compression and extra (TYPE only: colours are not kept);
 * header "LPMD 2.0 L|Z" and "HDR SYM X Y Z ...", then per frame the atom
 * count, the cell rows and per atom the symbol, the scaled position
 * x / |cell row| and the velocity / acceleration, all in %f.  zlp frames are
 * deflated on one stream, in blocksize pieces each ending with a sync flush,
 * and prefixed by a byte giving sizeof(unsigned long) and the size in
 * network order in that many bytes. */
typedef struct lpmdx_cfgfile lpmdx_cfgfile;
LPMDX_API int lpmdx_lpmd_open_write(lpmdx_cfgfile** out, const char* path, const char* args);
LPMDX_API int lpmdx_lpmd_write(lpmdx_cfgfile* f, long n, const double cell[9], const int* z, const double* x,
                               const double* v, const double* a);
/* Reading (LPMDFormat::ReadHeader / ReadCell, lpmd.cc:125-430): LPMD 1.0
 * and 2.0 text, LPMD 2.0 compressed, and the legacy ZLP 1 stream (a "Z4L2P"
 * header).  *level receives the header's level (-1 for ZLP 1, whose frames
 * carry their own).  lpmdx_lpmd_next reads the next frame and sets *n to its
 * atom count, or -1 at the end of the file; lpmdx_lpmd_frame then copies its
 * cell, atomic numbers (ElemNum: 0 for an unknown symbol), scaled positions s
 * as written in the file (the reader's Cartesian(s) is left to the caller,
 * who knows which cell applies), velocities, accelerations and level. */
LPMDX_API int lpmdx_lpmd_open_read(lpmdx_cfgfile** out, const char* path, int* level);
LPMDX_API int lpmdx_lpmd_next(lpmdx_cfgfile* f, long* n);
LPMDX_API int lpmdx_lpmd_frame(lpmdx_cfgfile* f, double cell[9], int* z, double* s, double* v, double* a,
                               int* level);
LPMDX_API int lpmdx_lpmd_close(lpmdx_cfgfile* f);

/* XYZFormat::WriteCell (plugins/xyz.cc:218-249): the atom count, an empty
 * line, then per atom the symbol and the position (velocity, acceleration at
 * level 1, 2) in %.<precout>f; args are the plugin's keywords (level,
 * precout, inside: positions folded into the cell as FittedInside does).
 * append = 1 adds a frame to an existing file. */
LPMDX_API int lpmdx_xyz_write(const char* path, int append, long n, const double cell[9], const int* z,
                              const double* x, const double* v, const double* a, const char* args);

#ifdef __cplusplus
}
#endif
#endif
