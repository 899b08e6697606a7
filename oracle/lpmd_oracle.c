/*
 * lpmd_oracle.c -- FP64 CPU restatement of the LPMD force and integration path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  The product
 * never links or calls this file.
 *
 * Every routine below restates a reference routine, cited by file:line, keeping
 * its loop order and the order of its FP64 operations (the reference is built
 * -O2 without -march on x86-64, so there is no FMA contraction; this file is
 * compiled with -ffp-contract=off for the same reason).
 *
 * Where the reference reads uninitialised memory the choice made here is
 * stated at the spot:
 *   - verletlist's extcut/step_count/old_size (plugins/verletlist.h:27-28):
 *     taken as zero, i.e. extcut = 1.08*cutoff and the first UpdateCell rebuilds;
 *   - MetalPotential's local stress[3][3] (metalpotential.cc:68): zero.
 */
#include "lpmd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[256];
static int fail(const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return -1;
}
const char* orc_last_error(void) { return g_err; }

/* ------------------------------------------------------------------------ */
/* geometry: liblpmd/lpmd/orthogonalcell.h                                   */
/* ------------------------------------------------------------------------ */

/* OrthogonalCell::Displacement, orthogonalcell.h:131-149: d = b - a, one wrap per periodic axis. */
void orc_displacement(const double L[3], const int per[3], const double a[3], const double b[3],
                      double d[3]) {
  for (int q = 0; q < 3; ++q)
    d[q] = b[q] - a[q];
  for (int q = 0; q < 3; ++q) {
    if (!per[q])
      continue;
    double m = L[q];
    double mhalf = 0.5 * m;
    if (d[q] >= mhalf)
      d[q] -= m;
    else if (d[q] < -mhalf)
      d[q] += m;
  }
}

/* GenericVector::SquareModule, vector.h:110: (x*x + y*y) + z*z. */
static inline double sqmod(const double v[3]) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }

/* Module of the cell vector L*e_q (vector.h:108), as FittedInside and linkedcell read it. */
static inline double edge_module(double Lq) { return sqrt(Lq * Lq + 0.0 * 0.0 + 0.0 * 0.0); }

/* OrthogonalCell::FittedInside, orthogonalcell.h:102-114 (x == L is left as is). */
void orc_fitted_inside(const double L[3], const int per[3], double x[3]) {
  for (int q = 0; q < 3; ++q) {
    if (!per[q])
      continue;
    double ll = edge_module(L[q]);
    if (x[q] < 0.0)
      x[q] -= ll * floor(x[q] / ll);
    else if (x[q] > ll)
      x[q] -= ll * floor(x[q] / ll);
  }
}

/* ------------------------------------------------------------------------ */
/* neighbour lists                                                            */
/* ------------------------------------------------------------------------ */

static void nl_clear(orc_nlist* l) { l->size = 0; }

static int nl_push(orc_nlist* l, long j, const double rij[3], double r2) {
  if (l->size == l->cap) {
    long nc = l->cap ? 2 * l->cap : 64;
    long* nj = realloc(l->j, nc * sizeof(long));
    double* nr = realloc(l->r2, nc * sizeof(double));
    double* nv = realloc(l->rij, 3 * nc * sizeof(double));
    if (!nj || !nr || !nv)
      return fail("out of memory");
    l->j = nj;
    l->r2 = nr;
    l->rij = nv;
    l->cap = nc;
  }
  l->j[l->size] = j;
  l->r2[l->size] = r2;
  memcpy(l->rij + 3 * l->size, rij, 3 * sizeof(double));
  l->size++;
  return 0;
}

static void nl_copy(orc_nlist* dst, const orc_nlist* src) {
  nl_clear(dst);
  for (long k = 0; k < src->size; ++k)
    nl_push(dst, src->j[k], src->rij + 3 * k, src->r2[k]);
}

void orc_nlist_free(orc_nlist* l) {
  free(l->j);
  free(l->r2);
  free(l->rij);
  memset(l, 0, sizeof *l);
}

struct orc_cm {
  int kind;
  double cutoff;
  /* linkedcell */
  int nx, ny, nz, mode_auto, side;
  long cells_inside, last_n;
  int* head;
  int* tail;
  int* subcell;
  long* atomlist;
  long* indexc;
  orc_nlist* cache_half;
  orc_nlist* cache_full;
  /* lcbinary: the atom of each subcell, or -1 */
  long* occ;
  /* verletlist */
  long maxneighbors, each, step_count, old_size;
  double extcut;
  long* nv;
  double* vlist; /* n * maxneighbors, stored as double like verletlist.cc:71-84 */
};

/* WrapAround, plugins/linkedcell.cc:9-15 (note i == -n maps to n). */
static inline int wrap_around(int i, int n) {
  if (i < 0)
    i += ((-i / n) + 1) * n;
  else if (i > n - 1)
    i -= (i / n) * n;
  return i;
}

orc_cm* orc_cm_linkedcell(double cutoff, int nx, int ny, int nz, int mode_auto) {
  orc_cm* cm = calloc(1, sizeof *cm);
  cm->kind = ORC_CM_LINKEDCELL;
  cm->cutoff = cutoff;
  cm->nx = nx;
  cm->ny = ny;
  cm->nz = nz;
  /* linkedcell.cc:29-41: the final value only depends on the mode keyword */
  cm->mode_auto = mode_auto;
  cm->last_n = -1;
  return cm;
}

orc_cm* orc_cm_verletlist(double cutoff, long maxneighbors, long each) {
  orc_cm* cm = calloc(1, sizeof *cm);
  cm->kind = ORC_CM_VERLETLIST;
  cm->cutoff = cutoff;
  cm->maxneighbors = maxneighbors;
  cm->each = each;
  /* verletlist.cc:26-27 with extcut read as 0 */
  cm->extcut = 1.08 * cutoff;
  cm->step_count = 0;
  cm->old_size = 0;
  return cm;
}

orc_cm* orc_cm_lcbinary(double cutoff, int nx, int ny, int nz, int mode_auto) {
  orc_cm* cm = calloc(1, sizeof *cm);
  cm->kind = ORC_CM_LCBINARY;
  cm->cutoff = cutoff;
  cm->nx = nx;
  cm->ny = ny;
  cm->nz = nz;
  /* lcbinary.cc:22-28: automatic when a subdivision is 0 or with mode auto */
  cm->mode_auto = mode_auto || nx == 0 || ny == 0 || nz == 0;
  return cm;
}

orc_cm* orc_cm_minimage(double cutoff) {
  orc_cm* cm = calloc(1, sizeof *cm);
  cm->kind = ORC_CM_MINIMAGE;
  cm->cutoff = cutoff;
  return cm;
}

static void lc_free_lists(orc_cm* cm) {
  if (cm->cache_half) {
    for (long i = 0; i < cm->last_n; ++i) {
      orc_nlist_free(&cm->cache_half[i]);
      orc_nlist_free(&cm->cache_full[i]);
    }
  }
  free(cm->cache_half);
  free(cm->cache_full);
  cm->cache_half = cm->cache_full = NULL;
}

void orc_cm_free(orc_cm* cm) {
  if (!cm)
    return;
  lc_free_lists(cm);
  free(cm->head);
  free(cm->tail);
  free(cm->subcell);
  free(cm->atomlist);
  free(cm->indexc);
  free(cm->nv);
  free(cm->vlist);
  free(cm->occ);
  free(cm);
}

/* linkedcell.cc:146-165: step n upward while |edge/n - cutoff/2.5| decreases; then make it even. */
int orc_lc_auto_n(double edge, double cutoff) {
  double fnn = (cutoff / 2.5);
  int n = 1;
  while (1) {
    double previo = fabs(edge / n - fnn);
    n++;
    double actual = fabs(edge / n - fnn);
    if (actual < 1E-3)
      break;
    else if (actual > previo) {
      n--;
      break;
    }
  }
  if ((n % 2) != 0)
    n--;
  return n;
}

static int lc_bin_one(orc_cm* cm, const orc_conf* c, long r) {
  /* Fractional (orthogonalcell.h:90-94), then WrapAround(int(floor(n*frac))) (linkedcell.cc:249-255) */
  const double* x = c->x + 3 * r;
  double f0 = x[0] / c->L[0], f1 = x[1] / c->L[1], f2 = x[2] / c->L[2];
  int i = wrap_around((int)floor(cm->nx * f0), cm->nx);
  int j = wrap_around((int)floor(cm->ny * f1), cm->ny);
  int k = wrap_around((int)floor(cm->nz * f2), cm->nz);
  long ncell = (long)cm->nx * cm->ny * cm->nz;
  long q = (long)k * (cm->nx * cm->ny) + (long)j * cm->nx + i;
  cm->indexc[r] = ((q >= 0) && (q < ncell)) ? q : -1;
  q = cm->indexc[r];
  if (q >= 0) {
    if (cm->head[q] == -1)
      cm->head[q] = cm->tail[q] = (int)r;
    else {
      cm->atomlist[cm->tail[q]] = r;
      cm->tail[q] = (int)r;
    }
  }
  cm->atomlist[r] = -1;
  return 0;
}

/* LinkedCell::UpdateCell, plugins/linkedcell.cc:112-267 */
static int lc_update(orc_cm* cm, const orc_conf* c) {
  const long n = c->n;
  if (n != cm->last_n) {
    lc_free_lists(cm);
    free(cm->atomlist);
    free(cm->indexc);
    cm->atomlist = malloc(n * sizeof(long));
    cm->indexc = malloc(n * sizeof(long));
    cm->cache_half = calloc(n, sizeof(orc_nlist));
    cm->cache_full = calloc(n, sizeof(orc_nlist));
  }
  cm->last_n = n;
  if (cm->mode_auto) {
    cm->nx = orc_lc_auto_n(edge_module(c->L[0]), cm->cutoff);
    cm->ny = orc_lc_auto_n(edge_module(c->L[1]), cm->cutoff);
    cm->nz = orc_lc_auto_n(edge_module(c->L[2]), cm->cutoff);
    cm->mode_auto = 0;
  }
  if (cm->nx <= 0 || cm->ny <= 0 || cm->nz <= 0)
    return fail("linkedcell: a subdivision count is zero");
  long ncell = (long)cm->nx * cm->ny * cm->nz;
  free(cm->head);
  free(cm->tail);
  cm->head = malloc(ncell * sizeof(int));
  cm->tail = malloc(ncell * sizeof(int));
  if (cm->subcell == NULL) {
    /* built once, on the first call (linkedcell.cc:215-241) */
    double d = edge_module(c->L[0]) / (double)cm->nx;
    if (edge_module(c->L[1]) / (double)cm->ny < d)
      d = edge_module(c->L[1]) / (double)cm->ny;
    if (edge_module(c->L[2]) / (double)cm->nz < d)
      d = edge_module(c->L[2]) / (double)cm->nz;
    int side = (int)ceil((cm->cutoff / d) - 0.5);
    cm->side = side;
    cm->cells_inside = (long)(2 * side + 1) * (2 * side + 1) * (2 * side + 1);
    cm->subcell = malloc(ncell * cm->cells_inside * sizeof(int));
    if (!cm->subcell)
      return fail("linkedcell: stencil table allocation failed");
    for (int k = 0; k < cm->nz; ++k)
      for (int j = 0; j < cm->ny; ++j)
        for (int i = 0; i < cm->nx; ++i) {
          long r = 0;
          long q = (long)k * (cm->nx * cm->ny) + (long)j * cm->nx + i;
          for (int dz = -side; dz <= side; ++dz)
            for (int dy = -side; dy <= side; ++dy)
              for (int dx = -side; dx <= side; ++dx) {
                int z0 = wrap_around(i + dx, cm->nx);
                int z1 = wrap_around(j + dy, cm->ny);
                int z2 = wrap_around(k + dz, cm->nz);
                long p = (long)z2 * (cm->nx * cm->ny) + (long)z1 * cm->nx + z0;
                cm->subcell[q * cm->cells_inside + (r++)] = (int)p;
              }
        }
  }
  for (long q = 0; q < ncell; ++q)
    cm->head[q] = cm->tail[q] = -1;
  for (long r = 0; r < n; ++r) {
    nl_clear(&cm->cache_half[r]);
    nl_clear(&cm->cache_full[r]);
    lc_bin_one(cm, c, r);
  }
  return 0;
}

/* LinkedCell::BuildNeighborList, plugins/linkedcell.cc:293-360, including its
 * per-atom caches: the first rcut asked for after an UpdateCell is kept. */
static int lc_list(orc_cm* cm, const orc_conf* c, long i, int full, double rcut, orc_nlist* out) {
  nl_clear(out);
  out->i = i;
  long cind = cm->indexc[i];
  if (cind < 0) {
    fprintf(stderr, "[Warning] linkedcell: atom %ld detected far outside the simulation box\n", i);
    return 0;
  }
  orc_nlist* cache = full ? &cm->cache_full[i] : &cm->cache_half[i];
  if (cache->size == 0) {
    const double* xi = c->x + 3 * i;
    for (long cc = 0; cc < cm->cells_inside; ++cc) {
      int nc = cm->subcell[cind * cm->cells_inside + cc];
      for (long zz = cm->head[nc]; zz != -1; zz = cm->atomlist[zz]) {
        if (zz == i)
          continue;
        if (!full && zz > i)
          continue;
        double rij[3];
        orc_displacement(c->L, c->per, xi, c->x + 3 * zz, rij);
        double r2 = sqmod(rij);
        if (r2 < rcut * rcut)
          if (nl_push(cache, zz, rij, r2))
            return -1;
      }
    }
  }
  nl_copy(out, cache);
  out->i = i;
  return 0;
}

/* lcbinary.cc:88-145: like linkedcell's choice but toward cutoff/5, and an odd n is rounded up. */
int orc_lcb_auto_n(double edge, double cutoff) {
  double fnn = cutoff / 5.0;
  int n = 1;
  while (1) {
    double previo = fabs(edge / n - fnn);
    n++;
    double actual = fabs(edge / n - fnn);
    if (actual < 1E-3)
      break;
    else if (actual > previo) {
      n--;
      break;
    }
  }
  if ((n % 2) != 0)
    n++;
  return n;
}

/* Subcell of position x as lcbinary.cc:198-212 bins it: floor(n * fraction),
 * one wrap of n - 1 < i.  A negative index (an atom below the box) indexes
 * outside the reference's tables; it is an error here. */
static long lcb_cell(const orc_cm* cm, const orc_conf* c, const double* x) {
  int i = (int)floor(cm->nx * (x[0] / c->L[0]));
  int j = (int)floor(cm->ny * (x[1] / c->L[1]));
  int k = (int)floor(cm->nz * (x[2] / c->L[2]));
  if (i > cm->nx - 1)
    i -= cm->nx;
  if (j > cm->ny - 1)
    j -= cm->ny;
  if (k > cm->nz - 1)
    k -= cm->nz;
  if (i < 0 || j < 0 || k < 0 || i > cm->nx - 1 || j > cm->ny - 1 || k > cm->nz - 1)
    return -1;
  return (long)k * (cm->nx * cm->ny) + (long)j * cm->nx + i;
}

/* LCBinary::UpdateCell, plugins/lcbinary.cc:85-241: the stencil table of the
 * current subdivision (offsets wrapped once), then every atom into its
 * subcell; a subcell reached twice grows every axis by two subcells and
 * everything is done again. */
static int lcb_update(orc_cm* cm, const orc_conf* c) {
  if (cm->mode_auto) {
    cm->nx = orc_lcb_auto_n(edge_module(c->L[0]), cm->cutoff);
    cm->ny = orc_lcb_auto_n(edge_module(c->L[1]), cm->cutoff);
    cm->nz = orc_lcb_auto_n(edge_module(c->L[2]), cm->cutoff);
    cm->mode_auto = 0;
  }
  for (;;) {
    if (cm->nx <= 0 || cm->ny <= 0 || cm->nz <= 0)
      return fail("lcbinary: a subdivision count is zero");
    const long ncell = (long)cm->nx * cm->ny * cm->nz;
    if (cm->subcell == NULL) {
      double d = edge_module(c->L[0]) / (double)cm->nx;
      if (edge_module(c->L[1]) / (double)cm->ny < d)
        d = edge_module(c->L[1]) / (double)cm->ny;
      if (edge_module(c->L[2]) / (double)cm->nz < d)
        d = edge_module(c->L[2]) / (double)cm->nz;
      const int side = (int)ceil((cm->cutoff / d) - 0.5);
      if (side >= cm->nx || side >= cm->ny || side >= cm->nz)
        return fail("lcbinary: the stencil is wider than the subcell grid");
      cm->side = side;
      cm->cells_inside = (long)(2 * side + 1) * (2 * side + 1) * (2 * side + 1);
      cm->subcell = malloc(ncell * cm->cells_inside * sizeof(int));
      if (!cm->subcell)
        return fail("lcbinary: stencil table allocation failed");
      for (int k = 0; k < cm->nz; ++k)
        for (int j = 0; j < cm->ny; ++j)
          for (int i = 0; i < cm->nx; ++i) {
            long r = 0;
            long q = (long)k * (cm->nx * cm->ny) + (long)j * cm->nx + i;
            for (int dz = -side; dz <= side; ++dz)
              for (int dy = -side; dy <= side; ++dy)
                for (int dx = -side; dx <= side; ++dx) {
                  int z0 = i + dx, z1 = j + dy, z2 = k + dz;
                  if (z0 < 0)
                    z0 += cm->nx;
                  else if (z0 > cm->nx - 1)
                    z0 -= cm->nx;
                  if (z1 < 0)
                    z1 += cm->ny;
                  else if (z1 > cm->ny - 1)
                    z1 -= cm->ny;
                  if (z2 < 0)
                    z2 += cm->nz;
                  else if (z2 > cm->nz - 1)
                    z2 -= cm->nz;
                  long p = (long)z2 * (cm->nx * cm->ny) + (long)z1 * cm->nx + z0;
                  cm->subcell[q * cm->cells_inside + (r++)] = (int)p;
                }
          }
    }
    if (cm->occ == NULL)
      cm->occ = malloc(ncell * sizeof(long));
    for (long q = 0; q < ncell; ++q)
      cm->occ[q] = -1;
    int grow = 0;
    for (long r = 0; r < c->n; ++r) {
      long q = lcb_cell(cm, c, c->x + 3 * r);
      if (q < 0)
        return fail("lcbinary: an atom lies below the cell");
      if (cm->occ[q] != -1)
        grow = 1;
      cm->occ[q] = r;
    }
    if (!grow)
      return 0;
    cm->nx += 2;
    cm->ny += 2;
    cm->nz += 2;
    free(cm->subcell);
    cm->subcell = NULL;
    free(cm->occ);
    cm->occ = NULL;
  }
}

/* LCBinary::BuildNeighborList, plugins/lcbinary.cc:264-306: the occupied
 * subcells of the stencil in stencil order (no cache); half keeps j < i. */
static int lcb_list(orc_cm* cm, const orc_conf* c, long i, int full, double rcut, orc_nlist* out) {
  nl_clear(out);
  out->i = i;
  const double* xi = c->x + 3 * i;
  const long cind = lcb_cell(cm, c, xi);
  if (cind < 0)
    return fail("lcbinary: an atom lies below the cell");
  for (long s = 0; s < cm->cells_inside; ++s) {
    const long z = cm->occ[cm->subcell[cind * cm->cells_inside + s]];
    if (z < 0 || z == i)
      continue;
    if (z > i && !full)
      continue;
    double rij[3];
    orc_displacement(c->L, c->per, xi, c->x + 3 * z, rij);
    const double r2 = sqmod(rij);
    if (r2 < rcut * rcut)
      if (nl_push(out, z, rij, r2))
        return -1;
  }
  return 0;
}

/* VerletListCellManager::UpdateVerletList, plugins/verletlist.cc:67-104 */
static int vl_rebuild(orc_cm* cm, const orc_conf* c) {
  if (fabs(cm->cutoff) < 1E-10)
    return fail("verletlist: No cutoff was defined");
  const long n = c->n;
  if (n != cm->old_size || cm->nv == NULL) {
    free(cm->nv);
    free(cm->vlist);
    cm->nv = malloc(n * sizeof(long));
    cm->vlist = malloc((size_t)n * cm->maxneighbors * sizeof(double));
    if (!cm->nv || !cm->vlist)
      return fail("verletlist: allocation failed");
  }
  for (long i = 0; i < n; ++i)
    cm->nv[i] = 0;
  const long mx = cm->maxneighbors;
  for (long i = 0; i < n - 1; ++i)
    for (long j = i + 1; j < n; ++j) {
      double rij[3];
      orc_displacement(c->L, c->per, c->x + 3 * i, c->x + 3 * j, rij);
      if (sqmod(rij) < cm->extcut * cm->extcut) {
        if ((cm->nv[j] == mx - 1) || (cm->nv[i] == mx - 1))
          return fail("verletlist: maxneighbors is too small!");
        cm->vlist[j * mx + cm->nv[j]++] = (double)i;
        cm->vlist[i * mx + cm->nv[i]++] = (double)j;
      }
    }
  return 0;
}

/* VerletListCellManager::UpdateCell, verletlist.cc:124-131 */
static int vl_update(orc_cm* cm, const orc_conf* c) {
  int must = ((cm->step_count % cm->each) == 0) || (c->n != cm->old_size);
  if (must && vl_rebuild(cm, c))
    return -1;
  cm->step_count++;
  cm->old_size = c->n;
  return 0;
}

/* VerletListCellManager::BuildNeighborList, verletlist.cc:135-167 (half keeps j > i) */
static int vl_list(orc_cm* cm, const orc_conf* c, long i, int full, double rcu, orc_nlist* out) {
  nl_clear(out);
  out->i = i;
  const double* xi = c->x + 3 * i;
  for (long jp = 0; jp < cm->nv[i]; ++jp) {
    long j = (long)cm->vlist[i * cm->maxneighbors + jp];
    if (!full && j < i)
      continue;
    double rij[3];
    orc_displacement(c->L, c->per, xi, c->x + 3 * j, rij);
    double r2 = sqmod(rij);
    if (r2 < rcu * rcu)
      if (nl_push(out, j, rij, r2))
        return -1;
  }
  return 0;
}

/* Exact minimum-image set by brute force, ascending j (the pair set
 * plugins/minimumimage.cc:55-100 produces; half keeps j > i like verletlist). */
static int mi_list(const orc_conf* c, long i, int full, double rcut, orc_nlist* out) {
  nl_clear(out);
  out->i = i;
  const double* xi = c->x + 3 * i;
  for (long j = full ? 0 : i + 1; j < c->n; ++j) {
    if (j == i)
      continue;
    double rij[3];
    orc_displacement(c->L, c->per, xi, c->x + 3 * j, rij);
    double r2 = sqmod(rij);
    if (r2 < rcut * rcut)
      if (nl_push(out, j, rij, r2))
        return -1;
  }
  return 0;
}

int orc_cm_update(orc_cm* cm, const orc_conf* c) {
  switch (cm->kind) {
  case ORC_CM_LINKEDCELL:
    return lc_update(cm, c);
  case ORC_CM_VERLETLIST:
    return vl_update(cm, c);
  case ORC_CM_LCBINARY:
    return lcb_update(cm, c);
  default:
    return 0;
  }
}

int orc_cm_list(orc_cm* cm, const orc_conf* c, long i, int full, double rcut, orc_nlist* out) {
  switch (cm->kind) {
  case ORC_CM_LINKEDCELL:
    return lc_list(cm, c, i, full, rcut, out);
  case ORC_CM_VERLETLIST:
    return vl_list(cm, c, i, full, rcut, out);
  case ORC_CM_LCBINARY:
    return lcb_list(cm, c, i, full, rcut, out);
  default:
    return mi_list(c, i, full, rcut, out);
  }
}

int orc_cm_grid(const orc_cm* cm, long out[5]) {
  out[0] = cm->nx;
  out[1] = cm->ny;
  out[2] = cm->nz;
  out[3] = cm->side;
  out[4] = cm->cells_inside;
  return 0;
}

long orc_cm_all_lists(orc_cm* cm, const orc_conf* c, int full, double rcut, long* counts,
                      long cap, long* js, double* r2, double* rij) {
  orc_nlist l = {0};
  long k = 0;
  for (long i = 0; i < c->n; ++i) {
    if (orc_cm_list(cm, c, i, full, rcut, &l)) {
      orc_nlist_free(&l);
      return -1;
    }
    counts[i] = l.size;
    for (long p = 0; p < l.size; ++p, ++k)
      if (k < cap) {
        js[k] = l.j[p];
        if (r2)
          r2[k] = l.r2[p];
        if (rij)
          memcpy(rij + 3 * k, l.rij + 3 * p, 3 * sizeof(double));
      }
  }
  orc_nlist_free(&l);
  return k <= cap ? k : -k;
}

/* ------------------------------------------------------------------------ */
/* potentials                                                                 */
/* ------------------------------------------------------------------------ */

int orc_pot_is_metal(const orc_pot* p) {
  return p->kind == ORC_SUTTONCHEN || p->kind == ORC_FINNISSINCLAIR || p->kind == ORC_GUPTA ||
         p->kind == ORC_FINNISSINCLAIREXT;
}

/* Potential::SetValidSpecies / AppliesTo, liblpmd/simulation/potential.cc:23-27 */
static inline int applies_to(const orc_pot* p, int s1, int s2) {
  return (p->spc_sum < 0) || ((s1 + s2) == p->spc_sum);
}

int orc_pot_init(orc_pot* p, int kind, const double* a, int s1, int s2) {
  memset(p, 0, sizeof *p);
  p->kind = kind;
  p->spc_sum = s1 + s2;
  p->cutoff = HUGE_VAL; /* Potential::Potential, potential.cc:13 */
  switch (kind) {
  case ORC_LJ: /* lennardjones.cc:11-22 */
    memcpy(p->p, a, 3 * sizeof(double));
    p->cutoff = a[2];
    break;
  case ORC_FASTLJ: { /* fastlj.cc:11-26, Tabulate fastlj.cc:64-80 */
    memcpy(p->p, a, 4 * sizeof(double));
    p->cutoff = a[2];
    p->bins = (long)a[3];
    if (p->bins <= 0)
      return fail("fastlj: bins must be positive");
    p->etable = malloc(p->bins * sizeof(double));
    p->fftable = malloc(p->bins * sizeof(double));
    const double sigma = a[0], epsilon = a[1], cutoff = a[2];
    for (long i = 0; i < p->bins; ++i) {
      double r = cutoff * (double)i / (double)p->bins;
      double r6 = pow(sigma / r, 6.0e0);
      double r12 = r6 * r6;
      p->etable[i] = 4.0e0 * epsilon * (r12 - r6);
      p->fftable[i] = -48.0e0 * (epsilon / (r * r)) * (r12 - 0.50e0 * r6);
    }
    break;
  }
  case ORC_MORSE: /* morse.cc:11-21: no SetCutoff, rcutoff stays HUGE_VAL */
    memcpy(p->p, a, 3 * sizeof(double));
    break;
  case ORC_BUCKINGHAM: /* buckingham.cc:11-24 */
    memcpy(p->p, a, 4 * sizeof(double));
    p->cutoff = a[3];
    break;
  case ORC_SUTTONCHEN: /* suttonchen.cc:11-32 */
    memcpy(p->p, a, 6 * sizeof(double));
    p->an = pow(a[1], a[2]);
    p->am = pow(a[1], a[3]);
    p->cutoff = a[5];
    break;
  case ORC_FINNISSINCLAIR: /* finnissinclair.cc:10-31: cutoff = max(c, d) */
    memcpy(p->p, a, 7 * sizeof(double));
    p->cutoff = (a[6] >= a[5]) ? a[6] : a[5];
    break;
  case ORC_GUPTA: /* gupta.cc:10-27 */
    memcpy(p->p, a, 6 * sizeof(double));
    p->cutoff = a[5];
    break;
  case ORC_HARMONIC: /* harmonic.cc:11-19: no SetCutoff, rcutoff stays HUGE_VAL */
    memcpy(p->p, a, 2 * sizeof(double));
    break;
  case ORC_LJMOD: /* lennardjonesMod.cc:11-28 */
    memcpy(p->p, a, 6 * sizeof(double));
    p->cutoff = a[5];
    break;
  case ORC_FINNISSINCLAIREXT: /* finnissinclair-ext.cc:11-36: cutoff = max(c, d) */
    memcpy(p->p, a, 9 * sizeof(double));
    p->cutoff = (a[8] >= a[7]) ? a[8] : a[7];
    break;
  default:
    return fail("unknown potential kind");
  }
  return 0;
}

/* tabulatedpair.cc:12-79: bins = rows, cutoff = the last increasing r; the
 * lookup is fastlj's (tabulatedpair.cc:81-96) */
int orc_pot_init_table(orc_pot* p, const double* e, const double* ff, long bins, double cutoff, int s1, int s2) {
  memset(p, 0, sizeof *p);
  p->kind = ORC_TABULATED;
  p->spc_sum = s1 + s2;
  if (bins <= 0 || !(cutoff > 0.0))
    return fail("tabulatedpair: empty table");
  p->cutoff = cutoff;
  p->p[2] = cutoff; /* fastlj_bin reads the cutoff here */
  p->bins = bins;
  p->etable = malloc(bins * sizeof(double));
  p->fftable = malloc(bins * sizeof(double));
  memcpy(p->etable, e, bins * sizeof(double));
  memcpy(p->fftable, ff, bins * sizeof(double));
  return 0;
}

void orc_pot_free(orc_pot* p) {
  free(p->etable);
  free(p->fftable);
  p->etable = p->fftable = NULL;
}

/* fastlj table index, fastlj.cc:83-84,90-91.  The reference reads past the
 * table when sqrt(r2) rounds up to the cutoff; the index is clamped here. */
static inline long fastlj_bin(const orc_pot* p, double r) {
  long k = (long)floor(p->bins * r / p->p[2]);
  if (k > p->bins - 1)
    k = p->bins - 1;
  return k;
}

double orc_pair_energy(const orc_pot* p, double r) {
  const double* a = p->p;
  switch (p->kind) {
  case ORC_LJ: { /* lennardjones.cc:53-58 */
    double rtmp = a[0] / r;
    double r6 = rtmp * rtmp * rtmp * rtmp * rtmp * rtmp;
    double r12 = r6 * r6;
    return 4.0e0 * a[1] * (r12 - r6);
  }
  case ORC_FASTLJ: /* fastlj.cc:81-85 */
  case ORC_TABULATED: /* tabulatedpair.cc:81-85 */
    return p->etable[fastlj_bin(p, r)];
  case ORC_MORSE: /* morse.cc:53-55 */
    return a[0] * (1.0 - exp(-a[1] * (r - a[2]))) * (1.0 - exp(-a[1] * (r - a[2])));
  case ORC_BUCKINGHAM: { /* buckingham.cc:58-63 */
    double T1 = a[0] * exp(-r / a[2]);
    double r6 = r * r * r * r * r * r;
    double T2 = a[1] / r6;
    return T1 - T2;
  }
  case ORC_SUTTONCHEN: { /* suttonchen.cc:73-76 */
    double ir = 1 / r;
    return a[0] * p->an * pow(ir, a[2]);
  }
  case ORC_FINNISSINCLAIR: /* finnissinclair.cc:75-80 */
    if (r <= a[5])
      return (r - a[5]) * (r - a[5]) * (a[0] + a[1] * r + a[2] * r * r);
    return 0.0e0;
  case ORC_GUPTA: /* gupta.cc:67 */
    return a[0] * exp(-a[2] * (r - a[1]) / a[1]);
  case ORC_HARMONIC: /* harmonic.cc:50-52 */
    return 0.5e0 * a[0] * (fabs(r) - a[1]) * (fabs(r) - a[1]);
  case ORC_LJMOD: { /* lennardjonesMod.cc:65-70 */
    double rtmp = a[0] / r;
    double rn = pow(rtmp, a[3]);
    double rm = pow(rtmp, a[2]);
    return 4.0e0 * a[1] * (rm - a[4] * rn);
  }
  case ORC_FINNISSINCLAIREXT: /* finnissinclair-ext.cc:85-91 */
    if (r <= a[7]) {
      double r2 = r * r;
      return (r - a[7]) * (r - a[7]) * (a[0] + a[1] * r + a[2] * r2 + a[3] * r2 * r + a[4] * r2 * r2);
    }
    return 0.0e0;
  }
  return NAN;
}

void orc_pair_force(const orc_pot* p, const double rij[3], double out[3]) {
  const double* a = p->p;
  double ff = NAN;
  switch (p->kind) {
  case ORC_LJ: { /* lennardjones.cc:60-67 */
    double rr2 = sqmod(rij);
    double r6 = pow(a[0] * a[0] / rr2, 3.0e0);
    double r12 = r6 * r6;
    ff = -48.0e0 * (a[1] / rr2) * (r12 - 0.50e0 * r6);
    break;
  }
  case ORC_FASTLJ: /* fastlj.cc:87-95 */
  case ORC_TABULATED: { /* tabulatedpair.cc:87-96 */
    double rr2 = sqmod(rij);
    double rr = sqrt(rr2);
    ff = p->fftable[fastlj_bin(p, rr)];
    break;
  }
  case ORC_MORSE: { /* morse.cc:57-62 */
    double rr = sqrt(sqmod(rij));
    ff = 2.0 * a[1] * (a[0] / rr) * (1.0 - exp(-a[1] * (rr - a[2]))) * exp(-a[1] * (rr - a[2]));
    break;
  }
  case ORC_BUCKINGHAM: { /* buckingham.cc:65-75 */
    double rr2 = sqmod(rij);
    double rr = sqrt(rr2);
    double r8 = rr * rr * rr * rr * rr * rr * rr * rr;
    double f6 = 6 * a[1] / r8;
    double pe = a[0] * exp(-rr / a[2]) / (a[2] * rr);
    ff = f6 - pe;
    break;
  }
  case ORC_HARMONIC: { /* harmonic.cc:54-60 */
    double rr = sqrt(sqmod(rij));
    ff = a[0] * (rr - a[1]) / rr;
    break;
  }
  case ORC_LJMOD: { /* lennardjonesMod.cc:72-79 */
    double rr2 = sqmod(rij);
    double rn = pow(a[0] * a[0] / rr2, a[3] / 2.0);
    double rm = pow(a[0] * a[0] / rr2, a[2] / 2.0);
    ff = -4.0e0 * (a[1] / rr2) * (a[2] * rm - a[4] * a[3] * rn);
    break;
  }
  }
  for (int q = 0; q < 3; ++q)
    out[q] = rij[q] * ff;
}

double orc_rhoij(const orc_pot* p, double r) {
  const double* a = p->p;
  switch (p->kind) {
  case ORC_SUTTONCHEN: { /* suttonchen.cc:78-81 */
    double ir = 1 / r;
    return p->am * pow(ir, a[3]);
  }
  case ORC_FINNISSINCLAIR: /* finnissinclair.cc:82-87 */
    if (r <= a[6])
      return ((r - a[6]) * (r - a[6])) * (1.0e0 + a[4] * (r - a[6]) / a[6]);
    return 0.0e0;
  case ORC_GUPTA: /* gupta.cc:69 */
    return exp(-2.0e0 * a[4] * (r - a[1]) / a[1]);
  case ORC_FINNISSINCLAIREXT: /* finnissinclair-ext.cc:93-99 */
    if (r <= a[8]) {
      double rd = r - a[8];
      return (rd * rd) * (1.0e0 + a[6] * a[6] * rd * rd);
    }
    return 0.0e0;
  }
  return NAN;
}

double orc_embed(const orc_pot* p, double rho) {
  const double* a = p->p;
  switch (p->kind) {
  case ORC_SUTTONCHEN: /* suttonchen.cc:83 */
    return -a[4] * a[0] * sqrt(rho);
  case ORC_FINNISSINCLAIR: /* finnissinclair.cc:89 */
    return -a[3] * sqrt(rho);
  case ORC_GUPTA: /* gupta.cc:71 */
    return -a[3] * sqrt(rho);
  case ORC_FINNISSINCLAIREXT: /* finnissinclair-ext.cc:101 */
    return -a[5] * sqrt(rho);
  }
  return NAN;
}

void orc_metal_pairforce(const orc_pot* p, const double nrm[3], double r, double out[3]) {
  const double* a = p->p;
  double s = 0.0;
  switch (p->kind) {
  case ORC_SUTTONCHEN: { /* suttonchen.cc:85-88 */
    double ir = 1.0e0 / r;
    s = -a[2] * a[0] * p->an * pow(ir, a[2] + 1);
    break;
  }
  case ORC_FINNISSINCLAIR: /* finnissinclair.cc:91-98 */
    if (r <= a[5]) {
      double t1 = 2.0e0 * (r - a[5]) * (a[0] + a[1] * r + a[2] * r * r);
      double t2 = (r - a[5]) * (r - a[5]) * (a[1] + 2.0e0 * a[2] * r);
      s = (t1 + t2);
    } else {
      out[0] = out[1] = out[2] = 0.0;
      return;
    }
    break;
  case ORC_GUPTA: /* gupta.cc:73-75 */
    s = -(a[0] * a[2] / a[1]) * exp(-a[2] * (r - a[1]) / a[1]);
    break;
  case ORC_FINNISSINCLAIREXT: /* finnissinclair-ext.cc:103-112 */
    if (r <= a[7]) {
      double r2 = r * r;
      double rc = r - a[7];
      double t1 = 2.0e0 * rc * (a[0] + a[1] * r + a[2] * r2 + a[3] * r2 * r + a[4] * r2 * r2);
      double t2 = (a[1] + 2.0e0 * a[2] * r + 3.0e0 * a[3] * r2 + 4.0e0 * a[4] * r2 * r) * rc * rc;
      s = t1 + t2;
    } else {
      out[0] = out[1] = out[2] = 0.0;
      return;
    }
    break;
  }
  for (int q = 0; q < 3; ++q)
    out[q] = nrm[q] * s;
}

void orc_metal_manybodies(const orc_pot* p, const double nrm[3], double rhoi, double rhoj,
                          double r, double out[3]) {
  const double* a = p->p;
  double s = 0.0;
  switch (p->kind) {
  case ORC_SUTTONCHEN: { /* suttonchen.cc:90-95 */
    double ir = 1.0e0 / r;
    s = 0.5 * a[3] * a[4] * a[0] * ((1 / sqrt(rhoi)) + 1 / (sqrt(rhoj))) * p->am *
        pow(ir, a[3] + 1);
    break;
  }
  case ORC_FINNISSINCLAIR: /* finnissinclair.cc:100-108 */
    if (r <= a[6]) {
      double t1 = ((1.0e0 / sqrt(rhoi)) + (1.0e0 / sqrt(rhoj))) * a[3] / 2.0e0;
      double t2 = 2.0e0 * (r - a[6]) + 3.0e0 * a[4] * (r - a[6]) * (r - a[6]) / a[6];
      s = -t1 * t2;
    } else {
      out[0] = out[1] = out[2] = 0.0;
      return;
    }
    break;
  case ORC_GUPTA: /* gupta.cc:77-82 */
    s = (a[3] * a[4] / a[1]) * ((1.0e0 / sqrt(rhoi)) + (1.0e0 / sqrt(rhoj))) *
        exp(-2.0e0 * a[4] * (r - a[1]) / a[1]);
    break;
  case ORC_FINNISSINCLAIREXT: /* finnissinclair-ext.cc:114-123 */
    if (r <= a[8]) {
      double rd = r - a[8];
      double t1 = a[5] * 0.5e0 * ((1.0e0 / sqrt(rhoi)) + (1.0e0 / sqrt(rhoj)));
      double t2 = 2.0e0 * rd + 4.0e0 * a[6] * a[6] * rd * rd * rd;
      s = -t1 * t2;
    } else {
      out[0] = out[1] = out[2] = 0.0;
      return;
    }
    break;
  }
  for (int q = 0; q < 3; ++q)
    out[q] = nrm[q] * s;
}

/* ------------------------------------------------------------------------ */
/* force composition                                                          */
/* ------------------------------------------------------------------------ */

/* PairPotential::UpdateForces, liblpmd/simulation/pairpotential.cc:35-84 */
static int pair_update(orc_conf* c, orc_cm* cm, orc_pot* p, double* ecache, double* virial,
                       double S[9]) {
  const double forcefactor = ORC_FORCEFACTOR;
  const long n = c->n;
  double stress[3][3] = {{0}};
  double etmp = 0.0e0, tmpvir = 0.0e0;
  const double cutoff = p->cutoff;
  orc_nlist nl = {0};
  for (long i = 0; i < n; ++i) {
    if (orc_cm_list(cm, c, i, 0, cutoff, &nl)) {
      orc_nlist_free(&nl);
      return -1;
    }
    for (long k = 0; k < nl.size; ++k) {
      long j = nl.j[k];
      const double* rij = nl.rij + 3 * k;
      if (applies_to(p, c->z[i], c->z[j]) && nl.r2[k] < cutoff * cutoff) {
        etmp += orc_pair_energy(p, sqrt(nl.r2[k]));
        double ff[3];
        orc_pair_force(p, rij, ff);
        double fi = forcefactor / c->m[i], fj = forcefactor / c->m[j];
        for (int q = 0; q < 3; ++q)
          c->a[3 * i + q] += ff[q] * fi;
        for (int q = 0; q < 3; ++q)
          c->a[3 * j + q] -= ff[q] * fj;
        tmpvir -= (rij[0] * ff[0] + rij[1] * ff[1] + rij[2] * ff[2]);
        for (int l = 0; l < 3; ++l) {
          stress[0][l] += -(rij[0]) * ff[l];
          stress[1][l] += -(rij[1]) * ff[l];
          stress[2][l] += -(rij[2]) * ff[l];
        }
      }
    }
  }
  orc_nlist_free(&nl);
  *ecache = 0.0;
  *ecache += etmp;
  *virial += tmpvir;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      S[3 * a + b] += stress[a][b];
  return 0;
}

/* Potential::AtomEnergy(conf, i) for every atom (the cell manager updated):
 * PairPotential::AtomEnergy, liblpmd/simulation/pairpotential.cc:21-33 (pair
 * energies over i's full list where the species rule applies and r2 < rc^2);
 * MetalPotential::AtomEnergy, metalpotential.cc:47-60 (rho_i over the full
 * list, the pair energies where the rule applies, plus F(rho_i)). */
int orc_atom_energies(orc_conf* c, orc_cm* cm, const orc_pot* p, double* e) {
  orc_nlist nl = {0};
  const double cutoff = p->cutoff;
  for (long i = 0; i < c->n; ++i) {
    double etmp = 0.0, rhoi = 0.0;
    if (orc_cm_list(cm, c, i, 1, cutoff, &nl)) {
      orc_nlist_free(&nl);
      return -1;
    }
    for (long k = 0; k < nl.size; ++k) {
      const long j = nl.j[k];
      if (orc_pot_is_metal(p)) {
        rhoi += orc_rhoij(p, sqrt(nl.r2[k]));
        if (applies_to(p, c->z[i], c->z[j]))
          etmp += orc_pair_energy(p, sqrt(nl.r2[k]));
      } else if (applies_to(p, c->z[i], c->z[j]) && nl.r2[k] < cutoff * cutoff) {
        etmp += orc_pair_energy(p, sqrt(nl.r2[k]));
      }
    }
    if (orc_pot_is_metal(p))
      etmp += orc_embed(p, rhoi);
    e[i] = etmp;
  }
  orc_nlist_free(&nl);
  return 0;
}

/* MetalPotential::UpdateForces, liblpmd/simulation/metalpotential.cc:62-180.
 * The first-call corrections pass (:86-116) only prints, so it is omitted. */
static int metal_update(orc_conf* c, orc_cm* cm, orc_pot* p, double* ecache, double* virial,
                        double S[9]) {
  const double forcefactor = ORC_FORCEFACTOR;
  const long n = c->n;
  double stress[3][3] = {{0}}; /* uninitialised in the reference */
  double tmpvir = 0.0e0, etmp = 0.0e0, etmp2 = 0.0e0;
  const double mpd = 0.0;
  double* rho = calloc(n, sizeof(double));
  orc_nlist nl = {0};
  for (long i = 0; i < n; ++i) {
    double rhoi = 0.0e0;
    if (orc_cm_list(cm, c, i, 1, p->cutoff, &nl))
      goto err;
    for (long k = 0; k < nl.size; ++k)
      rhoi += orc_rhoij(p, sqrt(nl.r2[k]));
    rho[i] = rhoi + mpd;
  }
  for (long i = 0; i < n; ++i) {
    if (orc_cm_list(cm, c, i, 0, p->cutoff, &nl))
      goto err;
    for (long k = 0; k < nl.size; ++k) {
      long j = nl.j[k];
      const double* rij = nl.rij + 3 * k;
      if (applies_to(p, c->z[i], c->z[j])) {
        double r = sqrt(nl.r2[k]);
        double nrm[3] = {rij[0], rij[1], rij[2]};
        double mod = sqrt(sqmod(nrm)); /* Vector::Normalize, vector.h:102-106 */
        for (int q = 0; q < 3; ++q)
          nrm[q] /= mod;
        etmp += orc_pair_energy(p, r);
        double pf[3], mb[3], t[3];
        orc_metal_pairforce(p, nrm, r, pf);
        orc_metal_manybodies(p, nrm, rho[i], rho[j], r, mb);
        for (int q = 0; q < 3; ++q)
          t[q] = pf[q] + mb[q];
        double fi = forcefactor / c->m[i], fj = forcefactor / c->m[j];
        for (int q = 0; q < 3; ++q)
          c->a[3 * i + q] += t[q] * fi;
        for (int q = 0; q < 3; ++q)
          c->a[3 * j + q] -= t[q] * fj;
        tmpvir -= (rij[0] * t[0] + rij[1] * t[1] + rij[2] * t[2]);
        for (int q = 0; q < 3; q++) {
          stress[0][q] += -(rij[0]) * t[q];
          stress[1][q] += -(rij[1]) * t[q];
          stress[2][q] += -(rij[2]) * t[q];
        }
      }
    }
    etmp2 += orc_embed(p, rho[i]);
  }
  orc_nlist_free(&nl);
  free(rho);
  *ecache = 0.0;
  *ecache += etmp + 0.0 + etmp2;
  *virial += tmpvir + 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      S[3 * a + b] += stress[a][b];
  return 0;
err:
  orc_nlist_free(&nl);
  free(rho);
  return -1;
}

/* CombinedPotential::UpdateForces, liblpmd/lpmd/combinedpotential.h:39-52 */
int orc_update_forces(orc_conf* c, orc_cm* cm, int npot, orc_pot* pots, double* epot,
                      double* virial, double S[9]) {
  for (long i = 0; i < 3 * c->n; ++i)
    c->a[i] = 0.0;
  *virial = 0.0;
  for (int k = 0; k < 9; ++k)
    S[k] = 0.0;
  if (orc_cm_update(cm, c))
    return -1;
  double en = 0.0;
  double* ec = malloc((npot ? npot : 1) * sizeof(double));
  for (int p = 0; p < npot; ++p) {
    int rc = orc_pot_is_metal(&pots[p]) ? metal_update(c, cm, &pots[p], &ec[p], virial, S)
                                        : pair_update(c, cm, &pots[p], &ec[p], virial, S);
    if (rc) {
      free(ec);
      return -1;
    }
  }
  /* CombinedPotential::energy, combinedpotential.h:24-29 */
  for (int p = 0; p < npot; ++p)
    en += ec[p];
  free(ec);
  *epot = en;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* forces of selected atoms on the exact neighbour set                        */
/* ------------------------------------------------------------------------ */

/* Neighbours of atom i within rlist by their single minimum image
 * (OrthogonalCell::Displacement, orthogonalcell.h:131-149), in any order:
 * from a cell grid of at least three cells a side with edges >= rlist, or by
 * brute force when the box is too small for one. */
typedef struct {
  int n[3];
  long* head;
  long* next;
} orc_grid;

static long grid_cell(const orc_conf* c, const orc_grid* g, const double* x, int q3[3]) {
  for (int q = 0; q < 3; ++q) {
    long k = (long)floor(x[q] / c->L[q] * g->n[q]);
    k %= g->n[q];
    if (k < 0)
      k += g->n[q];
    q3[q] = (int)k;
  }
  return ((long)q3[2] * g->n[1] + q3[1]) * g->n[0] + q3[0];
}

static long neighbours(const orc_conf* c, const orc_grid* g, long i, double rlist, long* js, double* rij,
                       double* r2) {
  long m = 0;
  const double rl2 = rlist * rlist;
  if (!g->head) {
    for (long j = 0; j < c->n; ++j) {
      if (j == i)
        continue;
      double d[3];
      orc_displacement(c->L, c->per, c->x + 3 * i, c->x + 3 * j, d);
      const double s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (s < rl2) {
        js[m] = j;
        memcpy(rij + 3 * m, d, sizeof d);
        r2[m++] = s;
      }
    }
    return m;
  }
  int q3[3];
  grid_cell(c, g, c->x + 3 * i, q3);
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int a = (q3[0] + dx + g->n[0]) % g->n[0];
        const int b = (q3[1] + dy + g->n[1]) % g->n[1];
        const int e = (q3[2] + dz + g->n[2]) % g->n[2];
        for (long j = g->head[((long)e * g->n[1] + b) * g->n[0] + a]; j >= 0; j = g->next[j]) {
          if (j == i)
            continue;
          double d[3];
          orc_displacement(c->L, c->per, c->x + 3 * i, c->x + 3 * j, d);
          const double s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          if (s < rl2) {
            js[m] = j;
            memcpy(rij + 3 * m, d, sizeof d);
            r2[m++] = s;
          }
        }
      }
  return m;
}

/* CombinedPotential::UpdateForces' acceleration of the atoms idx[0..m) when
 * the cell manager lists every pair within rlist by its minimum image (the
 * set verletlist and minimumimage deliver, verletlist.cc:135-167): pair
 * terms where the species rule applies and r2 < rc^2 (pairpotential.cc:48-72);
 * EAM rho over every listed neighbour within the cutoff, pair and many-body
 * terms where the rule applies (metalpotential.cc:121-169).  Summed per atom
 * over its neighbours rather than in the reference's half-list order, so it
 * agrees with orc_update_forces to FP64 rounding.  acc[3m]. */
int orc_sample_forces(const orc_conf* c, int npot, const orc_pot* pots, double rlist, long m, const long* idx,
                      double* acc) {
  orc_grid g = {{1, 1, 1}, NULL, NULL};
  int grid = 1;
  for (int q = 0; q < 3; ++q) {
    g.n[q] = (int)floor(c->L[q] / rlist);
    if (g.n[q] < 3)
      grid = 0;
  }
  if (grid) {
    const long ncell = (long)g.n[0] * g.n[1] * g.n[2];
    g.head = malloc(ncell * sizeof(long));
    g.next = malloc(c->n * sizeof(long));
    for (long k = 0; k < ncell; ++k)
      g.head[k] = -1;
    for (long j = c->n - 1; j >= 0; --j) {
      int q3[3];
      const long k = grid_cell(c, &g, c->x + 3 * j, q3);
      g.next[j] = g.head[k];
      g.head[k] = j;
    }
  }
  const long cap = c->n;
  long* js = malloc(cap * sizeof(long));
  long* js2 = malloc(cap * sizeof(long));
  double* rij = malloc(3 * cap * sizeof(double));
  double* rij2 = malloc(3 * cap * sizeof(double));
  double* r2 = malloc(cap * sizeof(double));
  double* r22 = malloc(cap * sizeof(double));
  for (long s = 0; s < m; ++s) {
    const long i = idx[s];
    double f[3] = {0.0, 0.0, 0.0};
    const long nn = neighbours(c, &g, i, rlist, js, rij, r2);
    for (int p = 0; p < npot; ++p) {
      const orc_pot* P = &pots[p];
      const double rc2 = P->cutoff * P->cutoff;
      if (!orc_pot_is_metal(P)) {
        for (long k = 0; k < nn; ++k) {
          const long j = js[k];
          if (applies_to(P, c->z[i], c->z[j]) && r2[k] < rc2) {
            double ff[3];
            orc_pair_force(P, rij + 3 * k, ff);
            for (int q = 0; q < 3; ++q)
              f[q] += ff[q];
          }
        }
        continue;
      }
      /* rho of i and of each neighbour (rho ignores species, metalpotential.cc:121-132) */
      double rhoi = 0.0;
      for (long k = 0; k < nn; ++k)
        if (r2[k] < rc2)
          rhoi += orc_rhoij(P, sqrt(r2[k]));
      for (long k = 0; k < nn; ++k) {
        const long j = js[k];
        if (!(r2[k] < rc2) || !applies_to(P, c->z[i], c->z[j]))
          continue;
        const long n2 = neighbours(c, &g, j, rlist, js2, rij2, r22);
        double rhoj = 0.0;
        for (long q = 0; q < n2; ++q)
          if (r22[q] < rc2)
            rhoj += orc_rhoij(P, sqrt(r22[q]));
        const double r = sqrt(r2[k]);
        double nrm[3] = {rij[3 * k], rij[3 * k + 1], rij[3 * k + 2]};
        const double mod = sqrt(sqmod(nrm));
        for (int q = 0; q < 3; ++q)
          nrm[q] /= mod;
        double pf[3], mb[3];
        orc_metal_pairforce(P, nrm, r, pf);
        orc_metal_manybodies(P, nrm, rhoi, rhoj, r, mb);
        for (int q = 0; q < 3; ++q)
          f[q] += pf[q] + mb[q];
      }
    }
    const double fi = ORC_FORCEFACTOR / c->m[i];
    for (int q = 0; q < 3; ++q)
      acc[3 * s + q] = f[q] * fi;
  }
  free(js);
  free(js2);
  free(rij);
  free(rij2);
  free(r2);
  free(r22);
  free(g.head);
  free(g.next);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* integration                                                                */
/* ------------------------------------------------------------------------ */

/* VelocityVerlet::AdvancePosition, plugins/velocityverlet.cc:60-67 */
void orc_vv_position(orc_conf* c, double dt, double* aold) {
  for (long i = 0; i < c->n; ++i) {
    double* x = c->x + 3 * i;
    const double* v = c->v + 3 * i;
    const double* a = c->a + 3 * i;
    double np[3];
    for (int q = 0; q < 3; ++q)
      np[q] = (x[q] + v[q] * dt) + a[q] * 0.5 * dt * dt;
    for (int q = 0; q < 3; ++q)
      aold[3 * i + q] = a[q];
    orc_fitted_inside(c->L, c->per, np);
    memcpy(x, np, sizeof np);
  }
}

/* VelocityVerlet::AdvanceVelocity, plugins/velocityverlet.cc:69-75 */
void orc_vv_velocity(orc_conf* c, double dt, const double* aold) {
  const double h = 0.5 * dt;
  for (long i = 0; i < c->n; ++i)
    for (int q = 0; q < 3; ++q)
      c->v[3 * i + q] = c->v[3 * i + q] + (aold[3 * i + q] + c->a[3 * i + q]) * h;
}

/* TwoStepIntegrator::Advance's tag handling (twostepintegrator.cc:17-26,45-54):
 * after a force evaluation every fixedvel atom's acceleration is set to zero */
static void zero_fixedvel(orc_conf* c) {
  if (!c->fix)
    return;
  for (long i = 0; i < c->n; ++i)
    if (c->fix[i] & ORC_FIX_VEL)
      c->a[3 * i] = c->a[3 * i + 1] = c->a[3 * i + 2] = 0.0;
}

/* lpmd/lpmd.cc:70-103 with SimulationEngine::DoStep (simulationbuilder.cc:126-136),
 * Integrator::UseOldConfig (integrator.cc:16-22) and TwoStepIntegrator::Advance
 * (twostepintegrator.cc:12-68): fixedvel atoms get a = 0 after each force
 * evaluation, fixedpos ("true") atoms skip AdvancePosition and
 * AdvanceVelocity (c->fix). */
int orc_md_run(orc_conf* c, orc_cm* cm, int npot, orc_pot* pots, double dt, long steps,
               double* epot_t, double* ekin_t, double* vir_t) {
  double ep, vir, S[9];
  double* aold = malloc(3 * c->n * sizeof(double));
  if (orc_update_forces(c, cm, npot, pots, &ep, &vir, S))
    goto err;
  for (long s = 0; s < steps; ++s) {
    for (long i = 0; i < 3 * c->n; ++i)
      c->a[i] = 0.0;
    if (s == 0) {
      /* the old-configuration copy is attached to the same cell manager,
       * which costs one UpdateCell (configuration.h:54-57) */
      if (orc_cm_update(cm, c))
        goto err;
    }
    if (orc_update_forces(c, cm, npot, pots, &ep, &vir, S))
      goto err;
    zero_fixedvel(c);
    if (!c->fix) {
      orc_vv_position(c, dt, aold);
    } else {
      /* AdvancePosition for every atom but the fixedpos ones */
      orc_conf one = *c;
      for (long i = 0; i < c->n; ++i) {
        if (c->fix[i] & ORC_FIX_POS)
          continue;
        one.n = 1;
        one.x = c->x + 3 * i;
        one.v = c->v + 3 * i;
        one.a = c->a + 3 * i;
        orc_vv_position(&one, dt, aold + 3 * i);
      }
    }
    if (orc_update_forces(c, cm, npot, pots, &ep, &vir, S))
      goto err;
    zero_fixedvel(c);
    if (!c->fix) {
      orc_vv_velocity(c, dt, aold);
    } else {
      orc_conf one = *c;
      for (long i = 0; i < c->n; ++i) {
        if (c->fix[i] & ORC_FIX_POS)
          continue;
        one.n = 1;
        one.x = c->x + 3 * i;
        one.v = c->v + 3 * i;
        one.a = c->a + 3 * i;
        orc_vv_velocity(&one, dt, aold + 3 * i);
      }
    }
    if (epot_t)
      epot_t[s] = ep;
    if (ekin_t)
      ekin_t[s] = orc_kinetic_energy_tag(c, c->fix != NULL);
    if (vir_t)
      vir_t[s] = vir;
  }
  free(aold);
  return 0;
err:
  free(aold);
  return -1;
}

/* ------------------------------------------------------------------------ */
/* properties: liblpmd/lpmd/properties.h                                      */
/* ------------------------------------------------------------------------ */

/* KineticEnergy, properties.h:22-37 */
double orc_kinetic_energy(const orc_conf* c) {
  double K = 0.0;
  for (long i = 0; i < c->n; ++i)
    K += 0.5 * c->m[i] * sqmod(c->v + 3 * i);
  return K * ORC_KIN2EV;
}

/* Temperature, properties.h:39-53 (no tags) */
double orc_temperature(const orc_conf* c) {
  if (c->n == 0)
    return 0.0;
  return (2.0 / 3.0) * orc_kinetic_energy(c) / (ORC_KBOLTZMANN * (double)c->n);
}

/* Momentum, properties.h:64-81 */
void orc_momentum(const orc_conf* c, double p[3]) { orc_momentum_tag(c, 0, p); }

static inline int tagged(const orc_conf* c, int tag, long i) {
  return tag && c->fix && (c->fix[i] & ORC_FIX_TAGGED);
}

/* KineticEnergy(atoms, tag), properties.h:22-37 */
double orc_kinetic_energy_tag(const orc_conf* c, int tag) {
  double K = 0.0;
  for (long i = 0; i < c->n; ++i) {
    if (tagged(c, tag, i))
      continue;
    K += 0.5 * c->m[i] * sqmod(c->v + 3 * i);
  }
  return K * ORC_KIN2EV;
}

/* Temperature(atoms, tag), properties.h:39-53: the tagged atoms leave the count */
double orc_temperature_tag(const orc_conf* c, int tag) {
  if (c->n == 0)
    return 0.0;
  long tmp = 0;
  if (tag)
    for (long i = 0; i < c->n; ++i)
      tmp += tagged(c, tag, i);
  return (2.0 / 3.0) * orc_kinetic_energy_tag(c, tag) / (ORC_KBOLTZMANN * (double)(c->n - tmp));
}

/* Momentum(atoms, tag), properties.h:64-81: p += Mass() * Velocity() */
void orc_momentum_tag(const orc_conf* c, int tag, double p[3]) {
  p[0] = p[1] = p[2] = 0.0;
  for (long i = 0; i < c->n; ++i) {
    if (tagged(c, tag, i))
      continue;
    for (int q = 0; q < 3; ++q)
      p[q] += c->v[3 * i + q] * c->m[i];
  }
}

/* drand48 as POSIX defines it: X' = (0x5DEECE66D X + 0xB) mod 2^48, the draw
 * X' / 2^48; srand48(s) sets X = (low 32 bits of s) 2^16 + 0x330E */
unsigned long long orc_srand48(long seed) {
  return ((unsigned long long)(unsigned int)seed << 16) | 0x330EULL;
}

double orc_drand48(unsigned long long* x) {
  *x = (0x5DEECE66DULL * *x + 0xBULL) & ((1ULL << 48) - 1);
  return ldexp((double)*x, -48);
}

/* SimulationEngine::InitVelocities, simulationbuilder.cc:167-185, with
 * RandomVector (liblpmd/lpmd/vector.h:161-167) and the unit system's
 * kboltzmann (units.cc:7) */
void orc_init_velocities(orc_conf* c, unsigned long long* x) {
  const double kboltzmann = 8.6173422E-05;
  const double targettemp = 1.0;
  const double Kpart = (3.0 / 2.0) * kboltzmann * targettemp;
  double totalp[3] = {0.0, 0.0, 0.0};
  for (long i = 0; i < c->n; ++i) {
    const double v2 = 2.0 * Kpart / c->m[i];
    const double m = sqrt(v2);
    double w[3];
    for (int q = 0; q < 3; ++q)
      w[q] = 2.0 * orc_drand48(x) - 1.0;
    const double r = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double vel[3] = {m * w[0] / r, m * w[1] / r, m * w[2] / r};
    for (int q = 0; q < 3; ++q) {
      totalp[q] += vel[q] * c->m[i];
      c->v[3 * i + q] = vel[q];
    }
  }
  for (int q = 0; q < 3; ++q)
    totalp[q] = totalp[q] / (double)c->n;
  for (long i = 0; i < c->n; ++i)
    for (int q = 0; q < 3; ++q)
      c->v[3 * i + q] -= totalp[q] / c->m[i];
}

/* SimulationEngine::SetSimulationTags, simulationbuilder.cc:84-124, with the
 * session's pressfactor and ua3togrcm3 ("160217.65000000", "1.66050000") and
 * OrthogonalCell::Volume (orthogonalcell.h:168); tag-aware when any atom
 * carries fixedvel or fixedpos.  out[] in the LPMDX_OBS_* layout. */
void orc_tags(const orc_conf* c, double epot, double virial, const double stress[9], double out[28]) {
  const double pressfactor = 160217.65, densityfact = 1.6605;
  const double v = edge_module(c->L[0]) * edge_module(c->L[1]) * edge_module(c->L[2]);
  int tag = 0;
  if (c->fix)
    for (long i = 0; i < c->n; ++i)
      tag = tag || (c->fix[i] & ORC_FIX_TAGGED);
  double P[3];
  orc_momentum_tag(c, tag, P);
  for (int k = 0; k < 28; ++k)
    out[k] = 0.0;
  const double kin = orc_kinetic_energy_tag(c, tag);
  out[0] = kin;
  out[1] = orc_temperature_tag(c, tag);
  out[2] = P[0];
  out[3] = P[1];
  out[4] = P[2];
  out[5] = epot;
  out[6] = virial;
  out[7] = (pressfactor / v) * ((2.0 / 3.0) * kin + (1.0 / 3.0) * virial);
  for (int q = 0; q < 3; ++q)
    for (int p = 0; p < 3; ++p)
      out[8 + 3 * q + p] = (pressfactor / v) * stress[3 * q + p];
  out[17] = epot + kin;
  out[18] = sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]);
  out[19] = v;
  out[20] = v / (double)c->n;
  out[21] = (double)c->n / v;
  double total_mass = 0.0; /* Density, properties.h:55-62 */
  for (long i = 0; i < c->n; ++i)
    total_mass += c->m[i];
  out[22] = densityfact * (total_mass / v);
  out[23] = (pressfactor / v) * (1.0 / 3.0) * virial;
  out[24] = (pressfactor / v) * (2.0 / 3.0) * kin;
  for (int q = 0; q < 3; ++q)
    out[25 + q] = edge_module(c->L[q]);
}
