/*
 * pib_dem_oracle.c -- TEST INFRASTRUCTURE ONLY (parity checker + CPU baseline).
 *
 * Serial FP64 restatement of the soft-sphere DEM of PUReIBM-PS2D-FD1D as
 * des_time_march(.FALSE.) runs it for a freely moving bed (paths relative to
 * PUReIBM_PS2D-FD1D/SRC/FLO).  The loops follow the reference routine named
 * in each comment: the coarse cell grid, the per-cell particle lists, the
 * neighbour lists NEIGHBOURS(l, 1..) built by the 27-cell scan, the pair loop
 * of CALC_FORCE_DES in neighbour-list order and CFNEWVALUES, with the
 * reference's expression association.  Build with -ffp-contract=off.
 *
 * Left out, with the reason:
 *   - the tangential force and the contact histories PN/PV/PFN/PFT: with the
 *     reference's MEW = 0 (collision_mod.F90:388) CFSLIDE sets FT to a signed
 *     zero (soft_sphere_coll.F90:862-866), the +PFN term is commented out
 *     (:426), so neither reaches FC; adding a signed zero leaves every sum;
 *   - V_ROT of CFRELVEL: OMEGA_NEW stays zero (CFNEWVALUES :1386), so the
 *     rotational slip is a signed zero too;
 *   - wall contacts (WALLDTSPLIT = .FALSE., :596), cage runs, shrink and the
 *     configuration-generation branches.
 *
 * Pinned by tests/test_dem_oracle.py: head-on pair against the closed-form
 * restitution coefficient and contact time of the linear spring-dashpot,
 * momentum balance, free flight with constant hydrodynamic terms, periodic
 * wrap, and the neighbour sets against brute force over periodic images.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pureibm_dem.h"

#define SMALL_NUMBER ((double)1.0E-35f) /* constants.F90:92 */
#define MN 30                            /* initialize_coarse_grid :358-362 (one phase) */
#define NWALLS 6
#define MAXNEIGHBORS (MN + 1 + NWALLS)   /* DES_ALLOCATE_ARRAYS :505 */

static const double PI = 3.14159265358979323846264338327950288; /* 4*atan(1) */

typedef struct orc_dem {
  pib_dem_config cfg;
  int n;
  double kn, en;
  /* state, (n,3) column-major as DES_POS_NEW(n,3) etc. */
  double *pos_new, *pos_old, *vel_new, *vel_old, *omega, *fc, *cforce;
  double *rad, *pmass, *pvol;
  /* coarse grid (initialize_coarse_grid) */
  int cx, cy, cz, imax1, jmax1, kmax1, imax2, jmax2, kmax2;
  double dxc, dyc, dzc, lx, ly, lz, max_radius;
  int *pijk;                   /* (n,3) */
  int *cnd, *pic_start, *pic;  /* per-cell counts, offsets, lists (ascending ip) */
  int *nbrs;                   /* NEIGHBOURS(n, MAXNEIGHBORS) */
  double dtsolid_orig, dtsolid, tcoll, s_time, overlap_max;
  int factor, ncontacts, substeps;
  char err[256];
} orc_dem;

#define P3(a, l, d) (a)[(size_t)(d) * D->n + (l)]
#define NB(l, k) D->nbrs[(size_t)(l) * MAXNEIGHBORS + (k) - 1] /* NEIGHBOURS(l+1, k) */
#define CELL(i, j, k) (((size_t)(k) - 1) * D->jmax2 * D->imax2 + ((size_t)(j) - 1) * D->imax2 + ((i) - 1))

static int fail(orc_dem* D, int code, const char* msg) {
  snprintf(D->err, sizeof D->err, "%s", msg);
  return code;
}

/* DES_DOTPRDCT (dem_mod.F90:199-218) */
static double dot3(const double* a, const double* b) {
  double s = 0.0;
  for (int i = 0; i < 3; ++i) s = s + a[i] * b[i];
  return s;
}

/* find_cell_index (collision_mod.F90:743-764) */
static void find_cell_index(orc_dem* D) {
  for (int ip = 0; ip < D->n; ++ip) {
    D->pijk[ip] = (int)(P3(D->pos_new, ip, 0) / D->dxc) + 2;
    if (D->pijk[ip] > D->imax1) D->pijk[ip] = D->imax1;
    D->pijk[D->n + ip] = (int)(P3(D->pos_new, ip, 1) / D->dyc) + 2;
    if (D->pijk[D->n + ip] > D->jmax1) D->pijk[D->n + ip] = D->jmax1;
    D->pijk[2 * D->n + ip] = (int)(P3(D->pos_new, ip, 2) / D->dzc) + 2;
    if (D->pijk[2 * D->n + ip] > D->kmax1) D->pijk[2 * D->n + ip] = D->kmax1;
  }
}

/* particles_in_cell (collision_mod.F90:766-805): per-cell lists in particle order */
static int particles_in_cell(orc_dem* D) {
  size_t nc = (size_t)D->imax2 * D->jmax2 * D->kmax2;
  memset(D->cnd, 0, nc * sizeof(int));
  for (int ip = 0; ip < D->n; ++ip) {
    int i = D->pijk[ip], j = D->pijk[D->n + ip], k = D->pijk[2 * D->n + ip];
    if (i < 1 || j < 1 || k < 1) return fail(D, PIB_ERR_ARG, "particle below the coarse grid");
    D->cnd[CELL(i, j, k)]++;
  }
  int acc = 0;
  for (size_t c = 0; c < nc; ++c) {
    D->pic_start[c] = acc;
    acc += D->cnd[c];
  }
  D->pic_start[nc] = acc;
  int* fillc = (int*)calloc(nc, sizeof(int));
  for (int ip = 0; ip < D->n; ++ip) {
    size_t c = CELL(D->pijk[ip], D->pijk[D->n + ip], D->pijk[2 * D->n + ip]);
    D->pic[D->pic_start[c] + fillc[c]++] = ip;
  }
  free(fillc);
  return PIB_OK;
}

/* GRID_BASED_NEIGHBOR_SEARCH (collision_mod.F90:807-1062), periodic in x, y, z */
static int neighbour_search(orc_dem* D) {
  for (int l = 0; l < D->n; ++l) {
    for (int k = 1; k <= MAXNEIGHBORS; ++k) NB(l, k) = -1;
    NB(l, 1) = 0;
  }
  for (int ll = 0; ll < D->n; ++ll) {
    NB(ll, 1) += 1;
    int nlim = NB(ll, 1) + 1;
    if (nlim > MAXNEIGHBORS) return fail(D, PIB_ERR_CAPACITY, "neighbour list overflow (MAXNEIGHBORS)");
    NB(ll, nlim) = ll + 1;
    int pc0 = D->pijk[ll], pc1 = D->pijk[D->n + ll], pc2 = D->pijk[2 * D->n + ll];
    for (int kk = pc2 - 1; kk <= pc2 + 1; ++kk)
      for (int jj = pc1 - 1; jj <= pc1 + 1; ++jj)
        for (int ii = pc0 - 1; ii <= pc0 + 1; ++ii) {
          int i = ii, j = jj, k = kk;
          double xf = 0, yf = 0, zf = 0;
          if (ii > D->imax1) { i = 2; xf = 1.0; }
          if (ii < 2) { i = D->imax1; xf = -1.0; }
          if (jj > D->jmax1) { j = 2; yf = 1.0; }
          if (jj < 2) { j = D->jmax1; yf = -1.0; }
          if (kk > D->kmax1) { k = 2; zf = 1.0; }
          if (kk < 2) { k = D->kmax1; zf = -1.0; }
          size_t c = CELL(i, j, k);
          for (int t = D->pic_start[c]; t < D->pic_start[c + 1]; ++t) {
            int pno = D->pic[t];
            if (pno <= ll) continue;
            double rlm = D->rad[ll] + D->rad[pno];
            rlm = 1.2 * rlm;  /* FACTOR_RLM (:356) */
            double cord[3], dist[3], pl[3];
            cord[0] = P3(D->pos_new, pno, 0) + xf * D->lx;
            cord[1] = P3(D->pos_new, pno, 1) + yf * D->ly;
            cord[2] = P3(D->pos_new, pno, 2) + zf * D->lz;
            for (int d = 0; d < 3; ++d) pl[d] = P3(D->pos_new, ll, d);
            for (int d = 0; d < 3; ++d) dist[d] = cord[d] - pl[d];
            double distmag = sqrt(dot3(dist, dist));
            int already = 0;
            for (int q = 2; q <= NB(ll, 1) + 1; ++q)
              if (pno + 1 == NB(ll, q)) already = 1;
            if (rlm - distmag > SMALL_NUMBER && !already) {
              NB(ll, 1) += 1;
              nlim = NB(ll, 1) + 1;
              if (nlim > MAXNEIGHBORS)
                return fail(D, PIB_ERR_CAPACITY, "neighbour list overflow (MAXNEIGHBORS)");
              NB(ll, nlim) = pno + 1;
              NB(pno, 1) += 1;
              nlim = NB(pno, 1) + 1;
              if (nlim > MAXNEIGHBORS)
                return fail(D, PIB_ERR_CAPACITY, "neighbour list overflow (MAXNEIGHBORS)");
              NB(pno, nlim) = ll + 1;
            }
          }
        }
  }
  return PIB_OK;
}

/* CALC_FORCE_DES (soft_sphere_coll.F90:52-551), normal spring-dashpot part */
static void calc_force_des(orc_dem* D) {
  const double alog = fabs(log(D->en));
  const double den = sqrt(PI * PI + log(D->en) * log(D->en));
  D->overlap_max = 0.0;
  D->ncontacts = 0;
  for (int ll = 0; ll < D->n; ++ll) {
    for (int ii = 2; ii <= NB(ll, 1) + 1; ++ii) {
      int i = NB(ll, ii) - 1;
      if (i <= ll) continue;
      double pi_[3] = {P3(D->pos_new, i, 0), P3(D->pos_new, i, 1), P3(D->pos_new, i, 2)};
      double pl[3] = {P3(D->pos_new, ll, 0), P3(D->pos_new, ll, 1), P3(D->pos_new, ll, 2)};
      const double box[3] = {D->cfg.xlength, D->cfg.ylength, D->cfg.zlength};
      for (int d = 0; d < 3; ++d) {  /* :309-349, EX2-WX1 = XLENGTH etc. */
        double tempd = fabs(pl[d] - pi_[d]);
        if (tempd > 4.0 * D->max_radius) {
          if (pi_[d] > pl[d])
            pi_[d] = pi_[d] - box[d];
          else
            pi_[d] = pi_[d] + box[d];
        }
      }
      double rlm = D->rad[ll] + D->rad[i];
      double dist[3];
      for (int d = 0; d < 3; ++d) dist[d] = pi_[d] - pl[d];
      double distmod = sqrt(dot3(dist, dist));
      if (!(rlm - distmod > SMALL_NUMBER)) continue;
      double ov = ((rlm - distmod) / rlm) * 100.0;
      if (ov > D->overlap_max) D->overlap_max = ov;
      D->ncontacts++;
      double normal[3], vrel[3];
      for (int d = 0; d < 3; ++d) normal[d] = dist[d] / distmod;
      /* CFRELVEL (:728-772) */
      for (int d = 0; d < 3; ++d) vrel[d] = P3(D->vel_new, ll, d) - P3(D->vel_new, i, d);
      double vrn = dot3(vrel, normal);
      double overlap_n = rlm - distmod;
      double mi = (PI * pow(2.0 * D->rad[ll], 3.0) * D->cfg.rhos) / 6.0;
      double mj = (PI * pow(2.0 * D->rad[i], 3.0) * D->cfg.rhos) / 6.0;
      double meff = (mi * mj) / (mi + mj);
      double eta = 2.0 * sqrt(D->kn * meff) * alog;
      eta = eta / den;
      for (int d = 0; d < 3; ++d) {
        double fns1 = -(D->kn * overlap_n * normal[d]);
        double fns2 = -(eta * vrn * normal[d]);
        double fn = fns1 + fns2;
        /* CFFCTOW (:873-904) */
        P3(D->fc, ll, d) = P3(D->fc, ll, d) + fn;
        P3(D->fc, i, d) = P3(D->fc, i, d) - fn;
      }
    }
  }
}

/* CFNEWVALUES (collision_mod.F90:1335-1462), soft-sphere, not generating */
static int cfnewvalues(orc_dem* D, int l, const double* pres, const double* visc, const double* mpg,
                       const double* fa) {
  for (int d = 0; d < 3; ++d) {
    double f = P3(D->fc, l, d);
    P3(D->cforce, l, d) = f;
    if (D->cfg.only_dem)
      f = f / D->pmass[l];
    else
      f = f / D->pmass[l] + D->cfg.rhof * (P3(pres, l, d) + P3(visc, l, d)) / D->pmass[l] - fa[d] -
          mpg[d] * D->pvol[l] / D->pmass[l];
    P3(D->vel_new, l, d) = P3(D->vel_old, l, d) + f * D->dtsolid;
    P3(D->omega, l, d) = P3(D->omega, l, d);
    P3(D->fc, l, d) = 0.0;
  }
  for (int d = 0; d < 3; ++d) P3(D->pos_new, l, d) = P3(D->pos_old, l, d) + P3(D->vel_new, l, d) * D->dtsolid;
  double dd[3];
  for (int d = 0; d < 3; ++d) dd[d] = P3(D->pos_new, l, d) - P3(D->pos_old, l, d);
  if (sqrt(dot3(dd, dd)) >= D->rad[l]) return fail(D, PIB_ERR_MOTION, "MOVEMENT UNDESIRED: a particle moved by its radius in one sub-step");
  const double box[3] = {D->cfg.xlength, D->cfg.ylength, D->cfg.zlength};
  for (int d = 0; d < 3; ++d) {
    if (P3(D->pos_new, l, d) > box[d])
      P3(D->pos_new, l, d) = P3(D->pos_new, l, d) - box[d];
    else if (P3(D->pos_new, l, d) < 0.0)
      P3(D->pos_new, l, d) = P3(D->pos_new, l, d) + box[d];
  }
  return PIB_OK;
}

static void cfupdateold(orc_dem* D) {
  memcpy(D->pos_old, D->pos_new, 3 * D->n * sizeof(double));
  memcpy(D->vel_old, D->vel_new, 3 * D->n * sizeof(double));
}

static int nint(double x) { return (int)(x >= 0 ? floor(x + 0.5) : -floor(-x + 0.5)); }

int orc_dem_create(const pib_dem_config* cfg, const double* xc, const double* radbdy, const double* velbdy,
                   void** out) {
  if (!cfg || !xc || !radbdy || !out) return PIB_ERR_ARG;
  if (cfg->nbody < 1) return PIB_ERR_ARG;
  orc_dem* D = (orc_dem*)calloc(1, sizeof(orc_dem));
  D->cfg = *cfg;
  int n = D->n = cfg->nbody;
  D->kn = cfg->kn > 0 ? cfg->kn : 800000.0;
  D->en = cfg->coeff_rest;
  double** arrs[] = {&D->pos_new, &D->pos_old, &D->vel_new, &D->vel_old, &D->omega, &D->fc, &D->cforce};
  for (size_t a = 0; a < sizeof arrs / sizeof arrs[0]; ++a) *arrs[a] = (double*)calloc(3 * (size_t)n, sizeof(double));
  D->rad = (double*)calloc(n, sizeof(double));
  D->pmass = (double*)calloc(n, sizeof(double));
  D->pvol = (double*)calloc(n, sizeof(double));
  D->pijk = (int*)calloc(3 * (size_t)n, sizeof(int));
  D->nbrs = (int*)calloc((size_t)n * MAXNEIGHBORS, sizeof(int));
  *out = D;
  if (!(D->en > 0.0 && D->en <= 1.0)) return fail(D, PIB_ERR_ARG, "coeff_rest must lie in (0, 1]");
  /* initialize_coarse_grid (:346-492) */
  double maxr = radbdy[0];
  for (int m = 1; m < n; ++m) maxr = radbdy[m] > maxr ? radbdy[m] : maxr;
  double dxeff = 2.0 * maxr * 1.5 * cfg->dx;
  D->cy = nint(cfg->ylength / dxeff); if (D->cy < 2) D->cy = 2;
  D->cx = nint(cfg->xlength / dxeff); if (D->cx < 2) D->cx = 2;
  D->cz = nint(cfg->zlength / dxeff); if (D->cz < 2) D->cz = 2;
  D->dyc = cfg->ylength / (double)(D->cy - 1);
  D->dxc = cfg->xlength / (double)(D->cx - 1);
  D->dzc = cfg->zlength / (double)(D->cz - 1);
  D->imax1 = D->cx; D->jmax1 = D->cy; D->kmax1 = D->cz;
  D->imax2 = D->cx + 1; D->jmax2 = D->cy + 1; D->kmax2 = D->cz + 1;
  double xe = 0, yn = 0, zt = 0, xe1 = 0, yn1 = 0, zt1 = 0;
  for (int i = 2; i <= D->imax2; ++i) { xe = xe + D->dxc; if (i == D->imax1) xe1 = xe; }
  for (int j = 2; j <= D->jmax2; ++j) { yn = yn + D->dyc; if (j == D->jmax1) yn1 = yn; }
  for (int k = 2; k <= D->kmax2; ++k) { zt = zt + D->dzc; if (k == D->kmax1) zt1 = zt; }
  D->lx = xe1 - 0.0; D->ly = yn1 - 0.0; D->lz = zt1 - 0.0;  /* :827-829 */
  size_t nc = (size_t)D->imax2 * D->jmax2 * D->kmax2;
  D->cnd = (int*)calloc(nc, sizeof(int));
  D->pic_start = (int*)calloc(nc + 1, sizeof(int));
  D->pic = (int*)calloc(n, sizeof(int));
  /* DES_INIT_ARRAYS (:1203-1316), not generating */
  for (int l = 0; l < n; ++l) {
    P3(D->pos_old, l, 0) = (xc[l] + cfg->foffset - 1.0) * cfg->dx;
    P3(D->pos_old, l, 1) = (xc[n + l] - 1.0) * cfg->dy;
    P3(D->pos_old, l, 2) = (xc[2 * n + l] - 1.0) * cfg->dz;
    D->rad[l] = radbdy[l] * cfg->dx;
    for (int d = 0; d < 3; ++d) P3(D->vel_old, l, d) = velbdy ? velbdy[(size_t)d * n + l] : 0.0;
  }
  memcpy(D->pos_new, D->pos_old, 3 * n * sizeof(double));
  memcpy(D->vel_new, D->vel_old, 3 * n * sizeof(double));
  /* CFASSIGN (:582-741) */
  D->max_radius = 0.0;
  for (int l = 0; l < n; ++l) {
    double rad2 = D->rad[l] * D->rad[l];
    D->pvol[l] = (4.0 / 3.0) * PI * rad2 * D->rad[l];
    D->pmass[l] = D->pvol[l] * cfg->rhos;
    if (D->rad[l] > D->max_radius) D->max_radius = D->rad[l];
  }
  const double box[3] = {cfg->xlength, cfg->ylength, cfg->zlength};
  const double cell[3] = {D->dxc, D->dyc, D->dzc};
  for (int d = 0; d < 3; ++d) {
    if (!(box[d] >= 8.0 * D->max_radius)) return fail(D, PIB_ERR_UNSUPPORTED, "box shorter than 8 particle radii");
    /* centres are wrapped once they move (CFNEWVALUES :1431-1459); one coarse
     * cell outside keeps find_cell_index inside the grid */
    for (int l = 0; l < n; ++l)
      if (!(P3(D->pos_new, l, d) > -cell[d] && P3(D->pos_new, l, d) < box[d] + cell[d]))
        return fail(D, PIB_ERR_ARG, "particle centre more than a coarse cell outside the periodic box");
  }
  {
    double mi = (PI * pow(cfg->dia_phys, 3.0) * cfg->rhos) / 6.0;
    double mj = (PI * pow(cfg->dia_phys, 3.0) * cfg->rhos) / 6.0;
    double meff = (mi * mj) / (mi + mj);
    double etan = 2.0 * sqrt(D->kn * meff) * fabs(log(D->en));
    etan = etan / sqrt(PI * PI + log(D->en) * log(D->en));
    double q = etan / meff;
    D->tcoll = PI / sqrt(D->kn / meff - (q * q) / 4.0);
    D->dtsolid_orig = D->tcoll / 50.0;
    D->dtsolid = D->dtsolid_orig;
  }
  find_cell_index(D);
  int rc = particles_in_cell(D);
  if (rc) return rc;
  return neighbour_search(D);
}

int orc_dem_destroy(void* h) {
  orc_dem* D = (orc_dem*)h;
  if (!D) return PIB_OK;
  double* arrs[] = {D->pos_new, D->pos_old, D->vel_new, D->vel_old, D->omega, D->fc, D->cforce,
                    D->rad, D->pmass, D->pvol};
  for (size_t a = 0; a < sizeof arrs / sizeof arrs[0]; ++a) free(arrs[a]);
  free(D->pijk); free(D->nbrs); free(D->cnd); free(D->pic_start); free(D->pic);
  free(D);
  return PIB_OK;
}

const char* orc_dem_last_error(void* h) { return h ? ((orc_dem*)h)->err : "null context"; }

/* des_time_march(.FALSE.), soft-sphere branch (:272-297), then IBMUPDATE (:1464-1474) */
int orc_dem_march(void* h, double dt, const double* pres, const double* visc, const double* mpg,
                  const double* fa, double* xc, double* velbdy, double* angv) {
  orc_dem* D = (orc_dem*)h;
  if (!D) return PIB_ERR_ARG;
  if (!D->cfg.only_dem && (!pres || !visc || !mpg || !fa))
    return fail(D, PIB_ERR_ARG, "pres, visc, mpg and frame_accln are required unless only_dem");
  if (!(dt > 0)) return fail(D, PIB_ERR_ARG, "dt must be positive");
  double fq = dt / D->dtsolid_orig;
  if (!(fq < 2.0e9)) return fail(D, PIB_ERR_ARG, "dt/DTSOLID too large");
  D->factor = (int)fq;
  D->dtsolid = D->dtsolid_orig;
  if (D->factor < 1) { D->factor = 1; D->dtsolid = dt; }
  for (int s = 0; s < D->factor; ++s) {
    /* SOFT_SPHERE_COLLISION (:1658-1761) */
    calc_force_des(D);
    for (int l = 0; l < D->n; ++l) {
      int rc = cfnewvalues(D, l, pres, visc, mpg, fa);
      if (rc) return rc;
    }
    D->s_time = D->s_time + D->dtsolid;
    cfupdateold(D);
    find_cell_index(D);
    int rc = particles_in_cell(D);
    if (rc) return rc;
    rc = neighbour_search(D);
    if (rc) return rc;
    D->substeps++;
  }
  int n = D->n;
  for (int m = 0; m < n; ++m) {
    if (xc) {
      xc[m] = P3(D->pos_new, m, 0) / D->cfg.dx + 1.0 - D->cfg.foffset;
      xc[n + m] = P3(D->pos_new, m, 1) / D->cfg.dy + 1.0;
      xc[2 * n + m] = P3(D->pos_new, m, 2) / D->cfg.dz + 1.0;
    }
    for (int d = 0; d < 3; ++d) {
      if (velbdy) velbdy[(size_t)d * n + m] = P3(D->vel_new, m, d);
      if (angv) angv[(size_t)d * n + m] = P3(D->omega, m, d);
    }
  }
  return PIB_OK;
}

int orc_dem_get_state(void* h, double* pos, double* vel, double* omega, double* cforce, pib_dem_info* info) {
  orc_dem* D = (orc_dem*)h;
  if (!D) return PIB_ERR_ARG;
  size_t b = 3 * (size_t)D->n * sizeof(double);
  if (pos) memcpy(pos, D->pos_new, b);
  if (vel) memcpy(vel, D->vel_new, b);
  if (omega) memcpy(omega, D->omega, b);
  if (cforce) memcpy(cforce, D->cforce, b);
  if (info) {
    memset(info, 0, sizeof *info);
    info->dtsolid = D->dtsolid_orig;
    info->tcoll = D->tcoll;
    info->dtsolid_used = D->dtsolid;
    info->s_time = D->s_time;
    info->overlap_max = D->overlap_max;
    info->box[0] = D->lx; info->box[1] = D->ly; info->box[2] = D->lz;
    info->factor = D->factor;
    info->ncontacts = D->ncontacts;
    info->cells[0] = D->cx - 1; info->cells[1] = D->cy - 1; info->cells[2] = D->cz - 1;
    info->substeps = D->substeps;
  }
  return PIB_OK;
}

/* Test access to NEIGHBOURS(l, 1..): nb receives, per particle, the count and
 * then the entries (1-based particle numbers), MAXNEIGHBORS ints a particle. */
int orc_dem_neighbours(void* h, int* nb) {
  orc_dem* D = (orc_dem*)h;
  if (!D) return PIB_ERR_ARG;
  memcpy(nb, D->nbrs, (size_t)D->n * MAXNEIGHBORS * sizeof(int));
  return MAXNEIGHBORS;
}

int orc_dem_set_state(void* h, const double* pos, const double* vel) {
  orc_dem* D = (orc_dem*)h;
  if (!D) return PIB_ERR_ARG;
  size_t b = 3 * (size_t)D->n * sizeof(double);
  if (pos) { memcpy(D->pos_new, pos, b); memcpy(D->pos_old, pos, b); }
  if (vel) { memcpy(D->vel_new, vel, b); memcpy(D->vel_old, vel, b); }
  find_cell_index(D);
  int rc = particles_in_cell(D);
  if (rc) return rc;
  return neighbour_search(D);
}
