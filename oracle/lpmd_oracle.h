/*
 * lpmd_oracle.h -- FP64 CPU restatement of the LPMD force and integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline), never as part of the product path.
 *
 * Each function follows a reference file (cited at its definition in
 * lpmd_oracle.c) with the same loop order and the same order of IEEE FP64
 * operations, so that neighbour sets and r2 values are bit-exact and sums
 * agree with the reference to rounding of libm only.
 */
#ifndef LPMD_ORACLE_H
#define LPMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Unit constants as the reference's session hands them to the potentials:
 * liblpmd/runtime/units.cc:6 gives 0.0096485341, liblpmd/runtime/session.cc:17-23,31
 * prints it fixed with 8 decimals and re-parses it, so the potentials see 0.00964853. */
#define ORC_FORCEFACTOR 0.00964853
#define ORC_KIN2EV 103.64269        /* liblpmd/lpmd/properties.h:10 */
#define ORC_KBOLTZMANN 8.6173422E-05 /* liblpmd/lpmd/properties.h:11 */
#define ORC_PRESSFACTOR 160217.65    /* units.cc:9 (survives the 8-decimal round trip) */

typedef struct {
  long n;
  double L[3];   /* orthogonal box edges */
  int per[3];    /* periodicity flags (OrthogonalCell::p) */
  double* x;     /* AoS [3n] positions (FixedSizeParticleSet storepos) */
  double* v;     /* AoS [3n] velocities */
  double* a;     /* AoS [3n] accelerations */
  int* z;        /* [n] atomic numbers */
  double* m;     /* [n] masses */
  const int* fix; /* [n] fixed-atom flags ORC_FIX_* (the fixedvel / fixedpos tags), or NULL */
} orc_conf;

#define ORC_FIX_VEL 1    /* carries fixedvel */
#define ORC_FIX_POS 2    /* carries fixedpos with the value "true" */
#define ORC_FIX_TAGGED 4 /* carries fixedvel or fixedpos, whatever the value */

typedef struct {
  long i;
  long size, cap;
  long* j;
  double* r2;
  double* rij; /* 3*size */
} orc_nlist;

enum { ORC_CM_LINKEDCELL = 1, ORC_CM_VERLETLIST = 2, ORC_CM_MINIMAGE = 3, ORC_CM_LCBINARY = 4 };

typedef struct orc_cm orc_cm;

orc_cm* orc_cm_linkedcell(double cutoff, int nx, int ny, int nz, int mode_auto);
orc_cm* orc_cm_lcbinary(double cutoff, int nx, int ny, int nz, int mode_auto);
int orc_lcb_auto_n(double edge, double cutoff);
orc_cm* orc_cm_verletlist(double cutoff, long maxneighbors, long each);
orc_cm* orc_cm_minimage(double cutoff);
void orc_cm_free(orc_cm* cm);
int orc_cm_update(orc_cm* cm, const orc_conf* c);
int orc_cm_list(orc_cm* cm, const orc_conf* c, long i, int full, double rcut, orc_nlist* out);
/* linkedcell grid after the first update: nx, ny, nz, side, cells_inside */
int orc_cm_grid(const orc_cm* cm, long out[5]);
void orc_nlist_free(orc_nlist* l);

/* All lists at once in CSR form; returns total entries, or -(needed) if cap is too small. */
long orc_cm_all_lists(orc_cm* cm, const orc_conf* c, int full, double rcut, long* counts,
                      long cap, long* js, double* r2, double* rij);

enum {
  ORC_LJ = 1,
  ORC_FASTLJ = 2,
  ORC_MORSE = 3,
  ORC_BUCKINGHAM = 4,
  ORC_SUTTONCHEN = 5,
  ORC_FINNISSINCLAIR = 6,
  ORC_GUPTA = 7,
  ORC_HARMONIC = 8,
  ORC_LJMOD = 9,
  ORC_FINNISSINCLAIREXT = 10,
  ORC_TABULATED = 11
};

typedef struct {
  int kind;
  int spc_sum;    /* Potential::SetValidSpecies: s1+s2, negative = wildcard */
  double cutoff;  /* Potential::rcutoff */
  double p[10];   /* kind-specific parameters in keyword order */
  long bins;      /* fastlj */
  double* etable; /* fastlj */
  double* fftable;
  double an, am;  /* suttonchen a^n, a^m */
} orc_pot;

/* Parameters in the plugin's keyword order:
 *  LJ: sigma epsilon cutoff          FASTLJ: sigma epsilon cutoff bins
 *  MORSE: depth a re                 BUCKINGHAM: B1 B2 Ro cutoff
 *  SUTTONCHEN: e a n m c cutoff      FINNISSINCLAIR: c0 c1 c2 A B c d
 *  GUPTA: A r0 p B q cutoff          HARMONIC: k a
 *  LJMOD: sigma epsilon m n cohesive cutoff
 *  FINNISSINCLAIREXT: c0 c1 c2 c3 c4 A B c d */
int orc_pot_init(orc_pot* p, int kind, const double* params, int s1, int s2);
/* tabulatedpair: the table's rows (energies, force factors) and cutoff */
int orc_pot_init_table(orc_pot* p, const double* e, const double* ff, long bins, double cutoff, int s1, int s2);
void orc_pot_free(orc_pot* p);
int orc_pot_is_metal(const orc_pot* p);
double orc_pair_energy(const orc_pot* p, double r);
void orc_pair_force(const orc_pot* p, const double rij[3], double out[3]);
double orc_rhoij(const orc_pot* p, double r);
double orc_embed(const orc_pot* p, double rho);
void orc_metal_pairforce(const orc_pot* p, const double nrm[3], double r, double out[3]);
void orc_metal_manybodies(const orc_pot* p, const double nrm[3], double rhoi, double rhoj,
                          double r, double out[3]);

/* CombinedPotential::UpdateForces: zero a, virial, stress; UpdateCell; each potential. */
int orc_update_forces(orc_conf* c, orc_cm* cm, int npot, orc_pot* pots, double* epot,
                      double* virial, double stress[9]);

/* The lpmd application loop (lpmd/lpmd.cc:70-103) with velocityverlet:
 * initial force call, then `steps` DoSteps (2 force calls each).  Per step it
 * records potential energy, kinetic energy, virial (arrays of length steps,
 * any may be NULL). */
int orc_md_run(orc_conf* c, orc_cm* cm, int npot, orc_pot* pots, double dt, long steps,
               double* epot_t, double* ekin_t, double* vir_t);

/* One velocityverlet position / velocity half step for all atoms, given the
 * accelerations: plugins/velocityverlet.cc:60-75. */
void orc_vv_position(orc_conf* c, double dt, double* aold);
void orc_vv_velocity(orc_conf* c, double dt, const double* aold);

/* Geometry: liblpmd/lpmd/orthogonalcell.h. */
void orc_displacement(const double L[3], const int per[3], const double a[3], const double b[3],
                      double d[3]);
void orc_fitted_inside(const double L[3], const int per[3], double x[3]);

/* Properties: liblpmd/lpmd/properties.h:22-81. */
double orc_kinetic_energy(const orc_conf* c);
double orc_temperature(const orc_conf* c);
void orc_momentum(const orc_conf* c, double p[3]);
/* CombinedPotential::UpdateForces' accelerations of the atoms idx[0..m) on
 * the exact minimum-image neighbour set within rlist, into acc[3m] */
int orc_sample_forces(const orc_conf* c, int npot, const orc_pot* pots, double rlist, long m, const long* idx,
                      double* acc);
/* Potential::AtomEnergy of every atom into e[n] (cell manager updated) */
int orc_atom_energies(orc_conf* c, orc_cm* cm, const orc_pot* p, double* e);
/* the tag-aware forms (tag = 1 leaves out ORC_FIX_TAGGED atoms) */
double orc_kinetic_energy_tag(const orc_conf* c, int tag);
double orc_temperature_tag(const orc_conf* c, int tag);
void orc_momentum_tag(const orc_conf* c, int tag, double p[3]);
/* drand48's generator: state X (48 bits); orc_srand48 returns the state
 * srand48(seed) sets, orc_drand48 advances it and returns the draw */
unsigned long long orc_srand48(long seed);
double orc_drand48(unsigned long long* x);
/* SimulationEngine::InitVelocities from the generator state *x */
void orc_init_velocities(orc_conf* c, unsigned long long* x);
/* SimulationEngine::SetSimulationTags: out[28] in the LPMDX_OBS_* layout */
void orc_tags(const orc_conf* c, double epot, double virial, const double stress[9], double out[28]);

/* linkedcell's auto subdivision for one axis (plugins/linkedcell.cc:146-165). */
int orc_lc_auto_n(double edge, double cutoff);

const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
