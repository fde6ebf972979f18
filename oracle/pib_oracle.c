/*
 * pib_oracle.c -- TEST INFRASTRUCTURE ONLY (parity checker + CPU baseline).
 *
 * FP64 restatement of PUReIBM-PS2D-FD1D's per-time-step fluid solve in plain
 * C.  The plane transforms and the per-mode line solves, which are
 * independent of each other, run under OpenMP (orc_set_threads); every sum
 * over planes, points or bodies stays serial in the reference's order, so
 * the results are the same bits for any thread count.  Each routine follows the reference routine named in its comment
 * (paths relative to PUReIBM_PS2D-FD1D/SRC/FLO), keeping the reference's loop
 * order and expression association so that the restatement can be read
 * against the Fortran line by line.  Arrays are held in the reference's
 * column-major layout with 1-based index helpers.  Build with
 * -ffp-contract=off so that every operation rounds as written.
 *
 * Pinned by tests/test_oracle.py and tests/test_oracle_pins.py (closed forms,
 * library routines, brute force, and bcset's ferror history against the
 * norm_drag sequence it is formed from).
 *
 * The only product-independent choices are (a) the FFT, a recursive
 * mixed-radix DIT DFT with directly evaluated twiddles standing in for FFTW
 * (whose 2-D c2r uses the row-column convention: inverse c2c along z, then
 * c2r along y ignoring Im of the ky=0 and ky=my/2 entries), and (b) the
 * neighbour search, evaluated by brute force over the 27 periodic images
 * instead of the coarse-grid cell list (same pair set for R_LM < cell size).
 */
#include "pib_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef struct { double re, im; } cpx;

/* constants.F90:92-93 are single-precision literals widened to real(8) */
#define SMALL_NUMBER ((double)1.0E-35f)
#define LARGE_NUMBER ((double)1.0E+32f)

static const double PI = 3.14159265358979323846264338327950288;

static inline cpx C(double r, double i) { cpx z; z.re = r; z.im = i; return z; }
static inline cpx cadd(cpx a, cpx b) { return C(a.re + b.re, a.im + b.im); }
static inline cpx csub(cpx a, cpx b) { return C(a.re - b.re, a.im - b.im); }
static inline cpx cscal(double s, cpx a) { return C(s * a.re, s * a.im); }
static inline cpx cdivr(cpx a, double s) { return C(a.re / s, a.im / s); }
/* (i*k) * z, the product of a pure-imaginary wavenumber with z */
static inline cpx cik(double k, cpx z) { return C(-k * z.im, k * z.re); }
static inline cpx cmul(cpx a, cpx b) {
  return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}

typedef struct orc {
  pib_config cfg;
  int nl_pending; /* rks of an orc_nonlinear_begin not yet finished */
  int nx, my, mz, my2, mx, mxf;
  double dx, dx2, dx3, vis;
  double coef[3][4];
  int itrmax;
  /* bcset's drag-error history (boundary_condition.F90:828-833) */
  int nerr_steps;
  double ferror_hist;
  double ferror_array[PIB_NERR_MAX];
  double *wy, *wz, *w2; /* wy, wz hold the imaginary parts; w2(my2,mz) */
  cpx *u, *nl, *onl, *ff; /* (nx+1, my2, mz, 3) */
  /* scalar transport, one species: phif, nlphif, onlphif, ffphi (nx+1, my2, mz) */
  cpx *phif, *nlphif, *onlphif, *ffphi;
  double sc_gamma, heat_ratio, heat_ratio_new, heat_ratio_old;
  /* the physical scalar arrays of the whole scalstep, planes 0..nx+1 (IS0):
   * phi, its nl and onl, diffn, the three gradient components, the forcing
   * and the x-flux of form_nlphi; per-body surface value and flux */
  double *phir, *nlphir, *onlphir, *dphi, *gphi, *frphi, *fdd;
  double *phisurfall, *flux_body;
  int nsurf;
  double phi_fluid_mean, phi_solid_mean, flux_global;
  cpx *p;                 /* (nx, my2, mz) */
  double *ubcp, *nlbc, *onlbc, *diffn, *fr; /* (nx+1, my, mz, 3); fr doubles as omega */
  double *omega_copy;     /* last omega, kept for parity dumps */
  double *pbcp;           /* (nx, my, mz) */
  double *ppr;            /* (nx, my, mz, 3) */
  unsigned char* fluid;   /* (0:nx+1, my, mz) */
  int nbody, nbnd;
  double *xc, *radbdy, *radibdy, *radobdy, *velbdy, *angv;
  double *bnd, *cd;
  unsigned char* if_rev; /* (nbnd, nbody) */
  int* nrpr_active;
  double *pres, *visc, *torq, *force, *force_chem; /* (nbody, 3) */
  long long count_solid, count_fluid;
  pib_scalars s;
  int have_points, have_particles;
  /* adjoint state and the checkpoint stack (orc_adjoint_*, orc_checkpoint_*) */
  void* adj;
  void** ck;
  int nck, ckcap;
  void** rec; /* recorded forward fields of steps to be reversed (orc_adjoint_record) */
  int nrec, reccap;
  /* plane work arrays */
  double *ra, *rb;
  cpx *ca, *cb;
  char err[256];
} orc;

/* ---- index helpers (1-based, column-major as in the reference) -------- */
#define IU(o, i, j, k, n) ((((size_t)((n)-1) * (o)->mz + ((k)-1)) * (o)->my2 + ((j)-1)) * ((o)->nx + 1) + ((i)-1))
#define IP(o, i, j, k) ((((size_t)((k)-1)) * (o)->my2 + ((j)-1)) * (o)->nx + ((i)-1))
#define IR(o, i, j, k, n) ((((size_t)((n)-1) * (o)->mz + ((k)-1)) * (o)->my + ((j)-1)) * ((o)->nx + 1) + ((i)-1))
#define IPR(o, i, j, k) ((((size_t)((k)-1)) * (o)->my + ((j)-1)) * (o)->nx + ((i)-1))
#define IPPR(o, i, j, k, n) ((((size_t)((n)-1) * (o)->mz + ((k)-1)) * (o)->my + ((j)-1)) * (o)->nx + ((i)-1))
#define IFL(o, i, j, k) ((((size_t)((k)-1)) * (o)->my + ((j)-1)) * ((o)->nx + 2) + (i))
#define XC(o, m, n) ((o)->xc[(size_t)((n)-1) * (o)->nbody + ((m)-1)])
#define PB(o, a, m, n) ((o)->a[(size_t)((n)-1) * (o)->nbody + ((m)-1)])
#define BND(o, n, l) ((o)->bnd[(size_t)((l)-1) * 3 + ((n)-1)])
#define CD(o, n, l) ((o)->cd[(size_t)((l)-1) * 3 + ((n)-1)])

static int fail(orc* o, int code, const char* msg) {
  snprintf(o->err, sizeof o->err, "%s", msg);
  return code;
}

/* ======================================================================
 * FFT: recursive mixed-radix decimation in time.  out[k] = sum_n in[n*s]
 * exp(sign*2*pi*i*n*k/n).  Twiddles are evaluated directly from the exact
 * integer phase (r*k mod n) so that no error accumulates across levels.
 * ====================================================================== */
static int small_factor(int n) {
  for (int p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}
/* exp(i*2*pi*t/n) for t < n, one table per length (cached) */
typedef struct { int n; cpx* tab; } twcache;
static twcache g_tw[64];
static int g_ntw = 0; /* entries are complete before the count is published */
static const cpx* tw_find(int n) {
  int cnt = __atomic_load_n(&g_ntw, __ATOMIC_ACQUIRE);
  for (int q = 0; q < cnt; ++q)
    if (g_tw[q].n == n) return g_tw[q].tab;
  return NULL;
}
static const cpx* tw_get(int n) {
  const cpx* found = tw_find(n);
  if (found) return found;
  cpx* t = (cpx*)malloc(sizeof(cpx) * (size_t)n);
  for (int k = 0; k < n; ++k) {
    long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
    t[k] = C((double)cosl(a), (double)sinl(a));
  }
#pragma omp critical(orc_twiddles)
  {
    found = tw_find(n);
    if (!found && g_ntw < 64) {
      g_tw[g_ntw].n = n;
      g_tw[g_ntw].tab = t;
      __atomic_store_n(&g_ntw, g_ntw + 1, __ATOMIC_RELEASE);
      found = t;
      t = NULL;
    }
  }
  if (!found) return t; /* table full: this call keeps its own copy (leaked) */
  free(t);
  return found;
}

/* threads for the plane and mode loops (0: OpenMP's default) */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : omp_get_num_procs()); }
int orc_get_threads(void) { return omp_get_max_threads(); }
/* out[k] = sum_r in[r*s] exp(sign*2*pi*i*r*k/m); twiddles from the length-ntop
 * table (m divides ntop), indexed by the exact integer phase r*k mod m */
static void dft_rec(const cpx* in, size_t s, cpx* out, int m, int sign, cpx* tmp, const cpx* tab, int ntop) {
  if (m == 1) { out[0] = in[0]; return; }
  int p = small_factor(m);
  int q = m / p;
  for (int r = 0; r < p; ++r) dft_rec(in + r * s, s * p, tmp + (size_t)r * q, q, sign, out, tab, ntop);
  int scale = ntop / m;
  for (int g = 0; g < p; ++g) {
    for (int k = 0; k < q; ++k) {
      int kk = k + g * q;
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < p; ++r) {
        long long ph = ((long long)r * kk) % m;
        cpx w = tab[ph * scale];
        double wi = sign < 0 ? -w.im : w.im;
        cpx v = tmp[(size_t)r * q + k];
        sr += v.re * w.re - v.im * wi;
        si += v.re * wi + v.im * w.re;
      }
      out[kk] = C(sr, si);
    }
  }
}
static void fft_c(const cpx* in, cpx* out, int n, int sign) {
  cpx* tmp = (cpx*)malloc(sizeof(cpx) * (size_t)n);
  cpx* tmp2 = (cpx*)malloc(sizeof(cpx) * (size_t)n);
  /* dft_rec uses its out argument as scratch for deeper levels */
  memcpy(tmp2, in, sizeof(cpx) * (size_t)n);
  dft_rec(tmp2, 1, out, n, sign, tmp, tw_get(n), n);
  free(tmp);
  free(tmp2);
}
void orc_fft1d(int n, int sign, const double* in, double* out) {
  fft_c((const cpx*)in, (cpx*)out, n, sign);
}

/* ff2rc, fftw_interface.F90:256-297: b = DFT2(a)/(my*mz), y is the halved axis */
void orc_ff2rc(int my, int mz, const double* a, double* bout) {
  int my2 = my / 2 + 1;
  cpx* b = (cpx*)bout;
  cpx* row = (cpx*)malloc(sizeof(cpx) * (size_t)(my > mz ? my : mz));
  cpx* res = (cpx*)malloc(sizeof(cpx) * (size_t)(my > mz ? my : mz));
  for (int k = 0; k < mz; ++k) {
    for (int j = 0; j < my; ++j) row[j] = C(a[(size_t)k * my + j], 0.0);
    fft_c(row, res, my, -1);
    for (int j = 0; j < my2; ++j) b[(size_t)k * my2 + j] = res[j];
  }
  for (int j = 0; j < my2; ++j) {
    for (int k = 0; k < mz; ++k) row[k] = b[(size_t)k * my2 + j];
    fft_c(row, res, mz, -1);
    for (int k = 0; k < mz; ++k) b[(size_t)k * my2 + j] = res[k];
  }
  double scale = (double)(float)(my * mz); /* float(mylo*mzlo) */
  for (size_t q = 0; q < (size_t)my2 * mz; ++q) b[q] = cdivr(b[q], scale);
  free(row);
  free(res);
}

/* ff2cr, fftw_interface.F90:299-338: unscaled inverse, input preserved */
void orc_ff2cr(int my, int mz, const double* bin, double* a) {
  int my2 = my / 2 + 1;
  const cpx* b = (const cpx*)bin;
  cpx* t = (cpx*)malloc(sizeof(cpx) * (size_t)my2 * mz);
  cpx* row = (cpx*)malloc(sizeof(cpx) * (size_t)(my > mz ? my : mz));
  cpx* res = (cpx*)malloc(sizeof(cpx) * (size_t)(my > mz ? my : mz));
  for (int j = 0; j < my2; ++j) {
    for (int k = 0; k < mz; ++k) row[k] = b[(size_t)k * my2 + j];
    fft_c(row, res, mz, +1);
    for (int k = 0; k < mz; ++k) t[(size_t)k * my2 + j] = res[k];
  }
  for (int k = 0; k < mz; ++k) {
    const cpx* tk = t + (size_t)k * my2;
    row[0] = C(tk[0].re, 0.0);
    row[my / 2] = C(tk[my / 2].re, 0.0);
    for (int j = 1; j < my / 2; ++j) {
      row[j] = tk[j];
      row[my - j] = C(tk[j].re, -tk[j].im);
    }
    fft_c(row, res, my, +1);
    for (int j = 0; j < my; ++j) a[(size_t)k * my + j] = res[j].re;
  }
  free(t);
  free(row);
  free(res);
}

/* ---- per-thread plane buffers for the parallel plane loops ----------- */
typedef struct { double *ra, *rb, *rc; cpx *ca, *cb, *cc; } planebuf;
static planebuf pb_alloc(const orc* o);
static void pb_free(planebuf* b) { free(b->ra); free(b->rb); free(b->rc); free(b->ca); free(b->cb); free(b->cc); }

/* ---- plane gather/scatter around the transforms ---------------------- */
/* real plane i of a (nx+1 or nx, my, mz) array with leading dim ld */
static void get_rplane(const orc* o, const double* f, int ld, int i, double* out) {
  for (int k = 1; k <= o->mz; ++k)
    for (int j = 1; j <= o->my; ++j)
      out[(size_t)(k - 1) * o->my + (j - 1)] = f[((size_t)(k - 1) * o->my + (j - 1)) * ld + (i - 1)];
}
static void put_rplane(const orc* o, double* f, int ld, int i, const double* in) {
  for (int k = 1; k <= o->mz; ++k)
    for (int j = 1; j <= o->my; ++j)
      f[((size_t)(k - 1) * o->my + (j - 1)) * ld + (i - 1)] = in[(size_t)(k - 1) * o->my + (j - 1)];
}

/* ======================================================================
 * Thomas algorithm: tridag / tridag3, tridiagonal.F90:47-150
 * ====================================================================== */
int orc_tridag(const double* a, const double* b, const double* c, const double* r,
               double* u, int n) {
  double* gam = (double*)calloc((size_t)n + 1, sizeof(double));
  int st = 0;
  for (int i = 0; i < n; ++i) u[i] = 0.0;
  if (b[0] == 0.0) st = PIB_ERR_PIVOT;
  double bet = b[0];
  u[0] = r[0] / bet;
  for (int j = 1; j < n; ++j) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j] * gam[j];
    if (bet == 0.0) { bet = 1.0e-30; st = PIB_ERR_PIVOT; }
    u[j] = (r[j] - a[j] * u[j - 1]) / bet;
  }
  for (int j = n - 2; j >= 0; --j) u[j] = u[j] - gam[j + 1] * u[j + 1];
  free(gam);
  return st;
}
static int tridag3(const double* a, const double* b, const double* c, const double* r1,
                   const double* r2, const double* r3, double* u1, double* u2, double* u3,
                   int n) {
  double* gam = (double*)calloc((size_t)n + 1, sizeof(double));
  int st = 0;
  for (int i = 0; i < n; ++i) u1[i] = u2[i] = u3[i] = 0.0;
  if (b[0] == 0.0) st = PIB_ERR_PIVOT;
  double bet = b[0];
  u1[0] = r1[0] / bet;
  u2[0] = r2[0] / bet;
  u3[0] = r3[0] / bet;
  for (int j = 1; j < n; ++j) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j] * gam[j];
    if (bet == 0.0) { bet = 1.0e-30; st = PIB_ERR_PIVOT; }
    u1[j] = (r1[j] - a[j] * u1[j - 1]) / bet;
    u2[j] = (r2[j] - a[j] * u2[j - 1]) / bet;
    u3[j] = (r3[j] - a[j] * u3[j - 1]) / bet;
  }
  for (int j = n - 2; j >= 0; --j) {
    u1[j] = u1[j] - gam[j + 1] * u1[j + 1];
    u2[j] = u2[j] - gam[j + 1] * u2[j + 1];
    u3[j] = u3[j] - gam[j + 1] * u3[j + 1];
  }
  free(gam);
  return st;
}

/* Sherman-Morrison wrapper: alpha = beta = 1, auxu = e_1 + e_n,
 * auxv(1) = c(n)/auxu(n), auxv(n) = a(1)/auxu(1) (usteptime.F90:97-108,
 * poisson.F90:150-161).  Complex rhs given as re/im; b is modified in place. */
static int sm_solve(const double* a, double* b, const double* c, const double* rr,
                    const double* ri, double* xr, double* xi, int n) {
  double* auxu = (double*)calloc((size_t)n, sizeof(double));
  double* auxv = (double*)calloc((size_t)n, sizeof(double));
  double* auxz = (double*)calloc((size_t)n, sizeof(double));
  double* yr = (double*)calloc((size_t)n, sizeof(double));
  double* yi = (double*)calloc((size_t)n, sizeof(double));
  auxu[0] = 1.0;
  auxu[n - 1] = 1.0;
  auxv[0] = c[n - 1] / auxu[n - 1];
  auxv[n - 1] = a[0] / auxu[0];
  b[0] = b[0] - auxu[0] * auxv[0];
  b[n - 1] = b[n - 1] - auxu[n - 1] * auxv[n - 1];
  int st = tridag3(a, b, c, rr, auxu, ri, yr, auxz, yi, n);
  double vdotz = auxv[0] * auxz[0] + auxv[n - 1] * auxz[n - 1];
  double vdotyr = auxv[0] * yr[0] + auxv[n - 1] * yr[n - 1];
  double vdotyi = auxv[0] * yi[0] + auxv[n - 1] * yi[n - 1];
  for (int i = 0; i < n; ++i) {
    xr[i] = yr[i] - vdotyr * auxz[i] / (1.0 + vdotz);
    xi[i] = yi[i] - vdotyi * auxz[i] / (1.0 + vdotz);
  }
  free(auxu); free(auxv); free(auxz); free(yr); free(yi);
  return st;
}
int orc_cyclic_solve(const double* a, const double* b, const double* c, const double* r,
                     double* x, int n) {
  double* bb = (double*)malloc(sizeof(double) * (size_t)n);
  double* zi = (double*)calloc((size_t)n, sizeof(double));
  double* xi = (double*)calloc((size_t)n, sizeof(double));
  memcpy(bb, b, sizeof(double) * (size_t)n);
  int st = sm_solve(a, bb, c, r, zi, x, xi, n);
  free(bb); free(zi); free(xi);
  return st;
}

/* ======================================================================
 * quadgener / generbnd / dcgen, geom_init.F90:44-369, 573-586
 * ====================================================================== */
int orc_quadgener(double r, double dr, double f1, double f2, int maxpts, double* bnd,
                  double* cd) {
  (void)dr; (void)f1;
  double twopi = 2.0 * PI;
  int imax = (int)lround(twopi * r / 4.0 * f2);
  double dtheta = twopi / imax / 4.0;
  double theta = 0.0;
  int npt = 0;
  int cap = 0;
  /* first sweep: count */
  {
    double th = 0.0;
    for (int i = 1; i <= imax + 1; ++i) {
      int jmax = (int)lround(twopi * r * sin(th) / 4 * f2);
      cap += jmax + 1;
      th = th + dtheta;
    }
  }
  double* nxb = (double*)malloc(sizeof(double) * (size_t)cap);
  double* nyb = (double*)malloc(sizeof(double) * (size_t)cap);
  double* nzb = (double*)malloc(sizeof(double) * (size_t)cap);
  for (int i = 1; i <= imax + 1; ++i) {
    int jmax = (int)lround(twopi * r * sin(theta) / 4 * f2);
    double dphi = twopi / jmax / 4;
    double phi = 0.0;
    for (int j = 1; j <= jmax + 1; ++j) {
      double z = r * cos(theta) / r;
      if (fabs(z) < 1.0e-5) z = 0.0;
      double y = r * sin(theta) * cos(phi) / r;
      if (fabs(y) < 1.0e-5) y = 0.0;
      double x = r * sin(theta) * sin(phi) / r;
      if (fabs(x) < 1.0e-5) x = 0.0;
      nxb[npt] = x;
      nyb[npt] = y;
      nzb[npt] = z;
      ++npt;
      phi = phi + dphi;
    }
    theta = theta + dtheta;
  }
  /* generbnd: octant then the seven mirror images in the reference order */
  static const int sx[8] = {1, -1, 1, 1, -1, -1, 1, -1};
  static const int sy[8] = {1, 1, -1, 1, -1, 1, -1, -1};
  static const int sz[8] = {1, 1, 1, -1, 1, -1, -1, -1};
  int count = 0;
  for (int g = 0; g < 8; ++g) {
    for (int i = 0; i < npt; ++i) {
      int need = 1;
      if (sx[g] < 0 && nxb[i] == 0.0) need = 0;
      if (sy[g] < 0 && nyb[i] == 0.0) need = 0;
      if (sz[g] < 0 && nzb[i] == 0.0) need = 0;
      if (!need) continue;
      double x = sx[g] * nxb[i], y = sy[g] * nyb[i], z = sz[g] * nzb[i];
      if (count < maxpts && bnd) {
        bnd[3 * count + 0] = x;
        bnd[3 * count + 1] = y;
        bnd[3 * count + 2] = z;
        double rr = sqrt(x * x + y * y + z * z);
        if (cd) {
          cd[3 * count + 0] = x / rr;
          cd[3 * count + 1] = y / rr;
          cd[3 * count + 2] = z / rr;
        }
      }
      ++count;
    }
  }
  free(nxb); free(nyb); free(nzb);
  return count;
}

/* ======================================================================
 * context
 * ====================================================================== */
static void* xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

int orc_create(const pib_config* cfg, void** ctx) {
  if (!cfg || !ctx) return PIB_ERR_ARG;
  if (cfg->nx < 3 || cfg->my < 4 || cfg->mz < 4 || (cfg->my & 1) || (cfg->mz & 1)) return PIB_ERR_ARG;
  orc* o = (orc*)calloc(1, sizeof(orc));
  o->cfg = *cfg;
  o->nx = cfg->nx; o->my = cfg->my; o->mz = cfg->mz;
  o->my2 = cfg->my / 2 + 1;
  o->mx = cfg->nx + 1;
  o->mxf = o->mx;
  o->dx = cfg->dx;
  o->dx2 = cfg->dx * cfg->dx;
  o->dx3 = pow(cfg->dx, 3.0);
  o->vis = cfg->vis;
  /* coef, initialize_flo.F90:861-885 */
  if (cfg->rk3) {
    double c[3][4] = {{4.0 / 15.0, 4.0 / 15.0, 8.0 / 15.0, 0.0},
                      {1.0 / 15.0, 1.0 / 15.0, 5.0 / 12.0, -17.0 / 60.0},
                      {1.0 / 6.0, 1.0 / 6.0, 3.0 / 4.0, -5.0 / 12.0}};
    memcpy(o->coef, c, sizeof c);
    o->itrmax = 3;
  } else {
    o->coef[0][0] = 0.5; o->coef[0][1] = 0.5; o->coef[0][2] = 3.0 / 2.0; o->coef[0][3] = -0.5;
    o->itrmax = 1;
  }
  /* initialize_flo.F90:695-696 */
  o->nerr_steps = PIB_NERR_DEFAULT;
  for (int k = 0; k < PIB_NERR_MAX; ++k) o->ferror_array[k] = 1.0;
  o->ferror_hist = 1.0;
  /* wavenumbers, initialize_flo.F90:735-763 (dy = dz = dx) */
  int my = o->my, mz = o->mz, my2 = o->my2;
  double twopi = 2.0 * PI;
  o->wy = (double*)xcalloc((size_t)my2, sizeof(double));
  o->wz = (double*)xcalloc((size_t)mz, sizeof(double));
  o->w2 = (double*)xcalloc((size_t)my2 * mz, sizeof(double));
  for (int i = 1; i <= my2; ++i) o->wy[i - 1] = twopi * (double)(i - 1) / ((double)my * o->dx);
  for (int i = 1; i <= mz / 2; ++i) {
    o->wz[i - 1] = twopi * (double)(i - 1) / ((double)mz * o->dx);
    o->wz[mz + 1 - i - 1] = -twopi * (double)i / ((double)mz * o->dx);
  }
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my2; ++j) {
      /* w2 = -dreal(wy*wy + wz*wz) with wy, wz pure imaginary */
      double wyy = -(o->wy[j] * o->wy[j]);
      double wzz = -(o->wz[k] * o->wz[k]);
      o->w2[(size_t)k * my2 + j] = -(wyy + wzz);
    }
  size_t nc = (size_t)(o->nx + 1) * my2 * mz * 3;
  size_t nr = (size_t)(o->nx + 1) * my * mz * 3;
  o->u = (cpx*)xcalloc(nc, sizeof(cpx));
  o->nl = (cpx*)xcalloc(nc, sizeof(cpx));
  o->onl = (cpx*)xcalloc(nc, sizeof(cpx));
  o->ff = (cpx*)xcalloc(nc, sizeof(cpx));
  o->p = (cpx*)xcalloc((size_t)o->nx * my2 * mz, sizeof(cpx));
  o->ubcp = (double*)xcalloc(nr, sizeof(double));
  o->nlbc = (double*)xcalloc(nr, sizeof(double));
  o->onlbc = (double*)xcalloc(nr, sizeof(double));
  o->diffn = (double*)xcalloc(nr, sizeof(double));
  o->fr = (double*)xcalloc(nr, sizeof(double));
  o->omega_copy = (double*)xcalloc(nr, sizeof(double));
  o->pbcp = (double*)xcalloc((size_t)o->nx * my * mz, sizeof(double));
  o->ppr = (double*)xcalloc((size_t)o->nx * my * mz * 3, sizeof(double));
  o->fluid = (unsigned char*)xcalloc((size_t)(o->nx + 2) * my * mz, 1);
  memset(o->fluid, 1, (size_t)(o->nx + 2) * my * mz);
  o->ra = (double*)xcalloc((size_t)my * mz, sizeof(double));
  o->rb = (double*)xcalloc((size_t)my * mz, sizeof(double));
  o->ca = (cpx*)xcalloc((size_t)my2 * mz, sizeof(cpx));
  o->cb = (cpx*)xcalloc((size_t)my2 * mz, sizeof(cpx));
  o->count_fluid = (long long)o->nx * my * mz;
  *ctx = o;
  return PIB_OK;
}

static planebuf pb_alloc(const orc* o) {
  planebuf b;
  size_t nr = (size_t)o->my * o->mz, nc = (size_t)o->my2 * o->mz;
  b.ra = (double*)xcalloc(nr, sizeof(double));
  b.rb = (double*)xcalloc(nr, sizeof(double));
  b.rc = (double*)xcalloc(nr, sizeof(double));
  b.ca = (cpx*)xcalloc(nc, sizeof(cpx));
  b.cb = (cpx*)xcalloc(nc, sizeof(cpx));
  b.cc = (cpx*)xcalloc(nc, sizeof(cpx));
  return b;
}

static void free_particles(orc* o) {
  free(o->xc); free(o->radbdy); free(o->radibdy); free(o->radobdy);
  free(o->velbdy); free(o->angv); free(o->if_rev); free(o->nrpr_active);
  free(o->pres); free(o->visc); free(o->torq); free(o->force); free(o->force_chem);
  o->xc = o->radbdy = o->radibdy = o->radobdy = o->velbdy = o->angv = NULL;
  o->pres = o->visc = o->torq = o->force = o->force_chem = NULL;
  o->if_rev = NULL;
  o->nrpr_active = NULL;
}

static void adj_free(orc* o);
int orc_destroy(void* ctx) {
  orc* o = (orc*)ctx;
  if (!o) return PIB_OK;
  free(o->wy); free(o->wz); free(o->w2);
  free(o->u); free(o->nl); free(o->onl); free(o->ff); free(o->p);
  free(o->phif); free(o->nlphif); free(o->onlphif); free(o->ffphi);
  free(o->phir); free(o->nlphir); free(o->onlphir); free(o->dphi); free(o->gphi); free(o->frphi);
  free(o->fdd); free(o->phisurfall); free(o->flux_body);
  free(o->ubcp); free(o->nlbc); free(o->onlbc); free(o->diffn); free(o->fr); free(o->omega_copy);
  free(o->pbcp); free(o->ppr); free(o->fluid);
  free(o->ra); free(o->rb); free(o->ca); free(o->cb);
  free(o->bnd); free(o->cd);
  free_particles(o);
  adj_free(o);
  free(o);
  return PIB_OK;
}
const char* orc_last_error(void* ctx) { return ctx ? ((orc*)ctx)->err : "null context"; }

int orc_set_points(void* ctx, int nbnd, const double* bndpts, const double* cd) {
  orc* o = (orc*)ctx;
  if (!o || nbnd <= 0 || !bndpts || !cd) return PIB_ERR_ARG;
  free(o->bnd); free(o->cd);
  o->nbnd = nbnd;
  o->bnd = (double*)malloc(sizeof(double) * 3 * (size_t)nbnd);
  o->cd = (double*)malloc(sizeof(double) * 3 * (size_t)nbnd);
  memcpy(o->bnd, bndpts, sizeof(double) * 3 * (size_t)nbnd);
  memcpy(o->cd, cd, sizeof(double) * 3 * (size_t)nbnd);
  o->have_points = 1;
  return PIB_OK;
}

/* grid_nodes_insphere, dependent_functions.F90:783-963 (serial, xperiodic) */
static void grid_nodes_insphere(orc* o) {
  int nx = o->nx, my = o->my, mz = o->mz, mxf = o->mxf;
  for (int k = 1; k <= mz; ++k)
    for (int j = 1; j <= my; ++j)
      for (int i = 0; i <= nx + 1; ++i) o->fluid[IFL(o, i, j, k)] = 1;
  long long count_solid_tmp = 0, count_fluid_tmp = (long long)nx * my * mz;
  for (int m = 1; m <= o->nbody; ++m) {
    double r = o->radbdy[m - 1];
    int imin = (int)ceil(XC(o, m, 1) - r), imax = (int)floor(XC(o, m, 1) + r);
    int jmin = (int)ceil(XC(o, m, 2) - r), jmax = (int)floor(XC(o, m, 2) + r);
    int kmin = (int)ceil(XC(o, m, 3) - r), kmax = (int)floor(XC(o, m, 3) + r);
    for (int i = imin; i <= imax; ++i) {
      int ii = i;
      if (i < 1) ii = mxf + i - 1;
      if (i > mxf - 1) ii = i - (mxf - 1);
      for (int j = jmin; j <= jmax; ++j) {
        int jj = j;
        if (j < 1) jj = my + j;
        if (j > my) jj = j - my;
        for (int k = kmin; k <= kmax; ++k) {
          int kk = k;
          if (k < 1) kk = mz + k;
          if (k > mz) kk = k - mz;
          double dist = (i - XC(o, m, 1)) * (i - XC(o, m, 1));
          dist = dist + (j - XC(o, m, 2)) * (j - XC(o, m, 2)) + (k - XC(o, m, 3)) * (k - XC(o, m, 3));
          dist = sqrt(dist);
          if ((dist - r) <= SMALL_NUMBER) {
            o->fluid[IFL(o, ii, jj, kk)] = 0;
            count_solid_tmp += 1;
            count_fluid_tmp -= 1;
          }
        }
      }
    }
  }
  o->count_solid = count_solid_tmp;
  o->count_fluid = count_fluid_tmp;
  for (int k = 1; k <= mz; ++k)
    for (int j = 1; j <= my; ++j) {
      o->fluid[IFL(o, mxf, j, k)] = o->fluid[IFL(o, 1, j, k)];
      o->fluid[IFL(o, 0, j, k)] = o->fluid[IFL(o, mxf - 1, j, k)];
    }
  o->s.maxvolfrac = (double)o->count_solid / (double)((long long)nx * my * mz);
  o->s.count_solid = (double)o->count_solid;
  o->s.count_fluid = (double)o->count_fluid;
}

/* neighbour pairs of the soft-sphere search (collision_mod.F90:807-1062):
 * a, b neighbours iff some periodic image has 1.2*(r_a+r_b) - |d| > SMALL */
static int are_neighbours(const orc* o, int a, int b) {
  double dx = o->dx;
  double L[3] = {o->nx * dx, o->my * dx, o->mz * dx};
  double pa[3], pb[3];
  for (int n = 1; n <= 3; ++n) {
    pa[n - 1] = (XC(o, a, n) - 1.0) * dx;
    pb[n - 1] = (XC(o, b, n) - 1.0) * dx;
  }
  double rlm = 1.2 * (o->radbdy[a - 1] * dx + o->radbdy[b - 1] * dx);
  for (int sx = -1; sx <= 1; ++sx)
    for (int sy = -1; sy <= 1; ++sy)
      for (int sz = -1; sz <= 1; ++sz) {
        double d0 = pb[0] + sx * L[0] - pa[0];
        double d1 = pb[1] + sy * L[1] - pa[1];
        double d2 = pb[2] + sz * L[2] - pa[2];
        double dm = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        if (rlm - dm > SMALL_NUMBER) return 1;
      }
  return 0;
}

/* update_nrpr_array + check_external_pt, dependent_functions.F90:966-1051 */
static void update_nrpr(orc* o) {
  int nb = o->nbody, nbnd = o->nbnd;
  unsigned char* nbr = (unsigned char*)xcalloc((size_t)nb * nb, 1);
  for (int a = 1; a <= nb; ++a)
    for (int b = a + 1; b <= nb; ++b)
      if (are_neighbours(o, a, b)) nbr[(size_t)(a - 1) * nb + (b - 1)] = nbr[(size_t)(b - 1) * nb + (a - 1)] = 1;
  for (int m = 1; m <= nb; ++m) {
    int active = nbnd;
    for (int l = 1; l <= nbnd; ++l) {
      double xlo[3];
      for (int n = 1; n <= 3; ++n) xlo[n - 1] = XC(o, m, n) + BND(o, n, l) * o->radobdy[m - 1];
      if (xlo[0] < 1.0) xlo[0] = o->mxf + xlo[0];
      if (xlo[0] > o->mxf) xlo[0] = xlo[0] - o->mxf;
      if (xlo[1] < 1.0) xlo[1] = o->my + 1.0 + xlo[1];
      if (xlo[1] > o->my) xlo[1] = xlo[1] - o->my - 1.0;
      if (xlo[2] < 1.0) xlo[2] = o->mz + 1.0 + xlo[2];
      if (xlo[2] > o->mz) xlo[2] = xlo[2] - o->mz - 1.0;
      int rev = 1;
      for (int q = 1; q <= nb && rev; ++q) {
        if (q == m || !nbr[(size_t)(m - 1) * nb + (q - 1)]) continue;
        double dist = 0.0;
        for (int n = 1; n <= 3; ++n) dist = dist + (XC(o, q, n) - xlo[n - 1]) * (XC(o, q, n) - xlo[n - 1]);
        dist = sqrt(dist);
        if (dist < o->radbdy[q - 1]) rev = 0;
      }
      o->if_rev[(size_t)(m - 1) * nbnd + (l - 1)] = (unsigned char)rev;
      if (!rev) active -= 1;
    }
    o->nrpr_active[m - 1] = active;
  }
  free(nbr);
}

int orc_set_particles(void* ctx, int nbody, const double* xc, const double* radbdy,
                      const double* velbdy, const double* angv, long long* count_solid,
                      int* nrpr_active) {
  orc* o = (orc*)ctx;
  if (!o || nbody < 0 || (nbody > 0 && (!xc || !radbdy))) return PIB_ERR_ARG;
  if (nbody > 0 && !o->have_points) return fail(o, PIB_ERR_STATE, "set_points first");
  /* a moved bed keeps the step's per-particle force sums (grid_nodes_insphere
   * leaves them alone) */
  double* keep = NULL;
  if (o->have_particles && nbody == o->nbody && nbody > 0) {
    size_t n3k = (size_t)nbody * 3;
    keep = (double*)xcalloc(5 * n3k, sizeof(double));
    memcpy(keep, o->pres, n3k * sizeof(double));
    memcpy(keep + n3k, o->visc, n3k * sizeof(double));
    memcpy(keep + 2 * n3k, o->torq, n3k * sizeof(double));
    memcpy(keep + 3 * n3k, o->force, n3k * sizeof(double));
    memcpy(keep + 4 * n3k, o->force_chem, n3k * sizeof(double));
  }
  free_particles(o);
  o->nbody = nbody;
  size_t n3 = (size_t)nbody * 3;
  o->xc = (double*)xcalloc(n3, sizeof(double));
  o->radbdy = (double*)xcalloc((size_t)nbody, sizeof(double));
  o->radibdy = (double*)xcalloc((size_t)nbody, sizeof(double));
  o->radobdy = (double*)xcalloc((size_t)nbody, sizeof(double));
  o->velbdy = (double*)xcalloc(n3, sizeof(double));
  o->angv = (double*)xcalloc(n3, sizeof(double));
  o->pres = (double*)xcalloc(n3, sizeof(double));
  o->visc = (double*)xcalloc(n3, sizeof(double));
  o->torq = (double*)xcalloc(n3, sizeof(double));
  o->force = (double*)xcalloc(n3, sizeof(double));
  o->force_chem = (double*)xcalloc(n3, sizeof(double));
  o->if_rev = (unsigned char*)xcalloc((size_t)nbody * (o->nbnd ? o->nbnd : 1), 1);
  o->nrpr_active = (int*)xcalloc((size_t)nbody, sizeof(int));
  if (nbody > 0) {
    memcpy(o->xc, xc, sizeof(double) * n3);
    memcpy(o->radbdy, radbdy, sizeof(double) * (size_t)nbody);
    if (velbdy) memcpy(o->velbdy, velbdy, sizeof(double) * n3);
    if (angv) memcpy(o->angv, angv, sizeof(double) * n3);
  }
  if (keep) {
    memcpy(o->pres, keep, n3 * sizeof(double));
    memcpy(o->visc, keep + n3, n3 * sizeof(double));
    memcpy(o->torq, keep + 2 * n3, n3 * sizeof(double));
    memcpy(o->force, keep + 3 * n3, n3 * sizeof(double));
    memcpy(o->force_chem, keep + 4 * n3, n3 * sizeof(double));
    free(keep);
  }
  for (int m = 0; m < nbody; ++m) {
    o->radibdy[m] = o->radbdy[m] - o->cfg.dr; /* initialize_flo.F90:501-505 */
    o->radobdy[m] = o->radbdy[m] + o->cfg.dr;
  }
  grid_nodes_insphere(o);
  if (nbody > 0) update_nrpr(o);
  o->have_particles = 1;
  if (count_solid) *count_solid = o->count_solid;
  if (nrpr_active)
    for (int m = 0; m < nbody; ++m) nrpr_active[m] = o->nrpr_active[m];
  return PIB_OK;
}

int orc_get_if_rev(void* ctx, unsigned char* if_rev) {
  orc* o = (orc*)ctx;
  if (!o || !if_rev) return PIB_ERR_ARG;
  memcpy(if_rev, o->if_rev, (size_t)o->nbody * o->nbnd);
  return PIB_OK;
}

/* ---- field transfer ---------------------------------------------------- */
static int adj_field_info(orc* o, int field, void** ptr, size_t* ndoubles);
static void adj_fix_copy_plane(orc* o, cpx* f, int copy);
static int field_info(orc* o, int field, void** ptr, size_t* ndoubles) {
  size_t nc = (size_t)(o->nx + 1) * o->my2 * o->mz * 3;
  size_t nr = (size_t)(o->nx + 1) * o->my * o->mz * 3;
  switch (field) {
    case PIB_F_U: *ptr = o->u; *ndoubles = 2 * nc; return 0;
    case PIB_F_P: *ptr = o->p; *ndoubles = 2 * (size_t)o->nx * o->my2 * o->mz; return 0;
    case PIB_F_NL: *ptr = o->nl; *ndoubles = 2 * nc; return 0;
    case PIB_F_ONL: *ptr = o->onl; *ndoubles = 2 * nc; return 0;
    case PIB_F_FF: *ptr = o->ff; *ndoubles = 2 * nc; return 0;
    case PIB_F_UBCP: *ptr = o->ubcp; *ndoubles = nr; return 0;
    case PIB_F_PBCP: *ptr = o->pbcp; *ndoubles = (size_t)o->nx * o->my * o->mz; return 0;
    case PIB_F_NLBCP: *ptr = o->nlbc; *ndoubles = nr; return 0;
    case PIB_F_ONLBCP: *ptr = o->onlbc; *ndoubles = nr; return 0;
    case PIB_F_DIFFN: *ptr = o->diffn; *ndoubles = nr; return 0;
    case PIB_F_PPR: *ptr = o->ppr; *ndoubles = (size_t)o->nx * o->my * o->mz * 3; return 0;
    case PIB_F_FR: *ptr = o->fr; *ndoubles = nr; return 0;
    case PIB_F_OMEGA: *ptr = o->omega_copy; *ndoubles = nr; return 0;
    case PIB_F_PHI: *ptr = o->phif; *ndoubles = 2 * nc / 3; return o->phif ? 0 : 1;
    case PIB_F_NLPHI: *ptr = o->nlphif; *ndoubles = 2 * nc / 3; return o->phif ? 0 : 1;
    case PIB_F_ONLPHI: *ptr = o->onlphif; *ndoubles = 2 * nc / 3; return o->phif ? 0 : 1;
    case PIB_F_FFPHI: *ptr = o->ffphi; *ndoubles = 2 * nc / 3; return o->phif ? 0 : 1;
    default: return adj_field_info(o, field, ptr, ndoubles);
  }
}
int orc_set_field(void* ctx, int field, const double* host) {
  orc* o = (orc*)ctx;
  void* ptr;
  size_t n;
  if (!o || !host) return PIB_ERR_ARG;
  int r = field_info(o, field, &ptr, &n); /* 1: a scalar field before orc_scalar_set */
  if (r) return r > 0 ? fail(o, PIB_ERR_STATE, r == 2 ? "pib_adjoint_begin first" : "pib_scalar_set first")
                      : PIB_ERR_ARG;
  memcpy(ptr, host, n * sizeof(double));
  if (field == PIB_F_AU || field == PIB_F_ANL) adj_fix_copy_plane(o, (cpx*)ptr, 0);
  return PIB_OK;
}
int orc_get_field(void* ctx, int field, double* host) {
  orc* o = (orc*)ctx;
  if (!o || !host) return PIB_ERR_ARG;
  if (field == PIB_F_MASK) {
    for (int k = 1; k <= o->mz; ++k)
      for (int j = 1; j <= o->my; ++j)
        for (int i = 1; i <= o->nx; ++i)
          host[IPR(o, i, j, k)] = o->fluid[IFL(o, i, j, k)] ? 1.0 : 0.0;
    return PIB_OK;
  }
  void* ptr;
  size_t n;
  int r = field_info(o, field, &ptr, &n);
  if (r) return r > 0 ? fail(o, PIB_ERR_STATE, r == 2 ? "pib_adjoint_begin first" : "pib_scalar_set first")
                      : PIB_ERR_ARG;
  if (field == PIB_F_AU || field == PIB_F_ANL) adj_fix_copy_plane(o, (cpx*)ptr, 1);
  memcpy(host, ptr, n * sizeof(double));
  return PIB_OK;
}
int orc_set_scalars(void* ctx, const pib_scalars* s) {
  orc* o = (orc*)ctx;
  if (!o || !s) return PIB_ERR_ARG;
  double ms = o->s.maxvolfrac, cs = o->s.count_solid, cf = o->s.count_fluid;
  o->s = *s;
  if (o->have_particles) { o->s.maxvolfrac = ms; o->s.count_solid = cs; o->s.count_fluid = cf; }
  return PIB_OK;
}
int orc_get_ferror_history(void* ctx, int n, double* ferror_array, double* ferror_hist) {
  orc* o = (orc*)ctx;
  if (!o) return PIB_ERR_ARG;
  if (n != o->nerr_steps) return fail(o, PIB_ERR_ARG, "n must equal NERR_STEPS");
  if (ferror_array) memcpy(ferror_array, o->ferror_array, sizeof(double) * (size_t)n);
  if (ferror_hist) *ferror_hist = o->ferror_hist;
  return PIB_OK;
}
/* the history a run starts from: ones (initialize_flo.F90:695-696) or the
 * array of a restart file (restart_funcs.F90:270-274), whose mean is the
 * starting ferror_hist */
int orc_set_ferror_history(void* ctx, int n, const double* ferror_array) {
  orc* o = (orc*)ctx;
  if (!o) return PIB_ERR_ARG;
  if (n < 1 || n > PIB_NERR_MAX) return fail(o, PIB_ERR_ARG, "NERR_STEPS out of range");
  o->nerr_steps = n;
  double sum = 0.0;
  for (int k = 0; k < n; ++k) {
    o->ferror_array[k] = ferror_array ? ferror_array[k] : 1.0;
    sum += o->ferror_array[k];
  }
  o->ferror_hist = sum / (double)n;
  return PIB_OK;
}
int orc_get_scalars(void* ctx, pib_scalars* s) {
  orc* o = (orc*)ctx;
  if (!o || !s) return PIB_ERR_ARG;
  *s = o->s;
  return PIB_OK;
}
int orc_get_forces(void* ctx, double* force, double* pres, double* visc, double* torq) {
  orc* o = (orc*)ctx;
  if (!o) return PIB_ERR_ARG;
  size_t n = (size_t)o->nbody * 3 * sizeof(double);
  if (force) memcpy(force, o->force, n);
  if (pres) memcpy(pres, o->pres, n);
  if (visc) memcpy(visc, o->visc, n);
  if (torq) memcpy(torq, o->torq, n);
  return PIB_OK;
}

/* ======================================================================
 * compute_new_timestep + calc_pressure, dependent_functions.F90:1054-1288,
 * 298-351 (serial, not moving, movingcv = .false.)
 * ====================================================================== */
int orc_compute_new_timestep(void* ctx, int rks) {
  orc* o = (orc*)ctx;
  if (!o || rks < 1 || rks > o->itrmax) return PIB_ERR_ARG;
  pib_scalars* s = &o->s;
  int nx = o->nx, my = o->my, mz = o->mz, my2 = o->my2;
  double umax_loc = SMALL_NUMBER, vmax_loc = SMALL_NUMBER, wmax_loc = SMALL_NUMBER;
  double ufl[3] = {0, 0, 0}, usl[3] = {0, 0, 0}, uml[3] = {0, 0, 0};
  double mesh_vel[3] = {0, 0, 0};
#pragma omp parallel
  {
    planebuf w = pb_alloc(o);
#pragma omp for schedule(static) collapse(2)
    for (int i = 1; i <= nx; ++i)
      for (int n = 1; n <= 3; ++n) {
        for (int k = 1; k <= mz; ++k)
          for (int j = 1; j <= my2; ++j) w.ca[(size_t)(k - 1) * my2 + (j - 1)] = o->u[IU(o, i, j, k, n)];
        orc_ff2cr(my, mz, (const double*)w.ca, w.ra);
        for (int k = 1; k <= mz; ++k)
          for (int j = 1; j <= my; ++j) o->ubcp[IR(o, i, j, k, n)] = w.ra[(size_t)(k - 1) * my + (j - 1)];
      }
    pb_free(&w);
  }
  for (int i = 1; i <= nx; ++i) {
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j) {
        for (int n = 1; n <= 3; ++n) o->ubcp[IR(o, i, j, k, n)] = o->ubcp[IR(o, i, j, k, n)] + s->umean[n - 1];
        if (o->fluid[IFL(o, i, j, k)]) {
          for (int n = 1; n <= 3; ++n) ufl[n - 1] = ufl[n - 1] + o->ubcp[IR(o, i, j, k, n)];
        } else {
          for (int n = 1; n <= 3; ++n) usl[n - 1] = usl[n - 1] + o->ubcp[IR(o, i, j, k, n)];
        }
        for (int n = 1; n <= 3; ++n) uml[n - 1] = uml[n - 1] + o->ubcp[IR(o, i, j, k, n)];
        umax_loc = fmax(umax_loc, fabs(o->ubcp[IR(o, i, j, k, 1)] - mesh_vel[0]));
        vmax_loc = fmax(vmax_loc, fabs(o->ubcp[IR(o, i, j, k, 2)] - mesh_vel[1]));
        wmax_loc = fmax(wmax_loc, fabs(o->ubcp[IR(o, i, j, k, 3)] - mesh_vel[2]));
      }
  }
  for (int n = 1; n <= 3; ++n)
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j) o->ubcp[IR(o, o->mx, j, k, n)] = o->ubcp[IR(o, 1, j, k, n)];
  double umax_tmp = fmax(fmax(umax_loc, vmax_loc), wmax_loc);
  double dtnew;
  if (umax_tmp * s->dt / o->dx > o->cfg.cfl_max)
    dtnew = o->cfg.cfl_max * o->dx / umax_tmp;
  else
    dtnew = s->dt;
  /* moving particles (dependent_functions.F90:1156-1165): when the fastest
   * particle breaks the CFL bound the reference recomputes dtnew, still from
   * the fluid's umax_tmp */
  if (o->cfg.move_particles) {
    double up_max = SMALL_NUMBER;
    for (int m = 1; m <= o->nbody; ++m) {
      double v1 = PB(o, velbdy, m, 1), v2 = PB(o, velbdy, m, 2), v3 = PB(o, velbdy, m, 3);
      up_max = fmax(up_max, sqrt(v1 * v1 + v2 * v2 + v3 * v3));
    }
    if (up_max * s->dt / o->dx > o->cfg.cfl_max) dtnew = o->cfg.cfl_max * o->dx / umax_tmp;
  }
  s->dtnew = dtnew;
  s->umax_tmp = umax_tmp;
  for (int n = 0; n < 3; ++n) {
    s->ufmean_old[n] = s->ufmean[n];
    s->ufmean[n] = ufl[n] / (double)o->count_fluid;
    s->usmean_temp[n] = usl[n] / (double)o->count_solid;
    s->umean_temp[n] = uml[n] / (double)((long long)nx * my * mz);
  }
  for (int n = 1; n <= 3; ++n) {
    double sum = 0.0;
    for (int m = 1; m <= o->nbody; ++m) sum += PB(o, velbdy, m, n);
    s->usmean_act[n - 1] = o->nbody > 0 ? sum / (double)o->nbody : 0.0;
  }
  double meanslip[3], mixmeanslip[3];
  for (int n = 0; n < 3; ++n) {
    s->dufmeandt[n] = -s->cf * (s->ufmean[n] - s->ufmean_old[n]);
    meanslip[n] = (s->ufmean[n] - s->usmean_des[n]) * (1.0 - s->maxvolfrac);
  }
  s->meanslipmod = sqrt(meanslip[0] * meanslip[0] + meanslip[1] * meanslip[1] + meanslip[2] * meanslip[2]);
  /* single phase: usmean = volfrac*mean_spec_vel/mean_volfrac */
  for (int n = 0; n < 3; ++n) {
    double mean_spec_vel = 0.0;
    if (o->nbody > 0) {
      double sum = 0.0;
      for (int m = 1; m <= o->nbody; ++m) sum += PB(o, velbdy, m, n + 1);
      mean_spec_vel = sum / (double)o->nbody;
    }
    double us = 0.0 + s->mean_volfrac * mean_spec_vel;
    s->usmean[n] = us / s->mean_volfrac;
    mixmeanslip[n] = (1.0 - s->maxvolfrac) * (s->usmean[n] - s->ufmean[n]);
  }
  s->mixmeanslipmod = sqrt(mixmeanslip[0] * mixmeanslip[0] + mixmeanslip[1] * mixmeanslip[1] +
                           mixmeanslip[2] * mixmeanslip[2]);
  /* calc_pressure */
#pragma omp parallel
  {
    planebuf w = pb_alloc(o);
#pragma omp for schedule(static)
    for (int i = 1; i <= nx; ++i) {
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j) w.ca[(size_t)(k - 1) * my2 + (j - 1)] = o->p[IP(o, i, j, k)];
      orc_ff2cr(my, mz, (const double*)w.ca, w.ra);
      put_rplane(o, o->pbcp, nx, i, w.ra);
    }
    pb_free(&w);
  }
  if (rks == o->itrmax && o->cfg.adaptive) s->dt = dtnew;
  s->t = s->t + s->dt * (o->coef[rks - 1][0] + o->coef[rks - 1][1]);
  double tempt = s->ramp_frac_time * (double)(float)(o->mx - 1) * o->dx / o->cfg.ucharmod;
  if (o->cfg.soft_start) {
    if (s->t < tempt) {
      s->cf = -1.0 / s->dt;
      s->cforig = s->cf;
      s->cf = s->cf * (s->t / tempt);
    } else {
      s->cf = -1.0 / s->dt;
      s->cforig = s->cf;
    }
  } else {
    s->cf = -1.0 / s->dt;
    s->cforig = s->cf;
  }
  return PIB_OK;
}

/* ======================================================================
 * form_nl, nl_allflow.F90:71-327 (serial, mesh_vel = 0)
 * ====================================================================== */
static void form_nl(orc* o) {
  int nx = o->nx, my = o->my, mz = o->mz, my2 = o->my2, mxf = o->mxf;
  double dx = o->dx;
  int center = o->cfg.discr_scheme == PIB_SCHEME_CENTER;
  size_t pl = (size_t)my * mz;
  /* fdderiv(mx, my, mz, 3) */
  double* fd = (double*)xcalloc((size_t)(nx + 1) * pl * 3, sizeof(double));
#define FD(i, j, k, n) fd[((((size_t)(n)-1) * mz + ((k)-1)) * my + ((j)-1)) * (nx + 1) + ((i)-1)]
#pragma omp parallel
  {
  double* pr[5];
  cpx* pf[5];
  for (int q = 0; q < 5; ++q) {
    pr[q] = (double*)xcalloc(pl, sizeof(double));
    pf[q] = (cpx*)xcalloc((size_t)my2 * mz, sizeof(cpx));
  }
#pragma omp for schedule(static)
  for (int i = 1; i <= nx; ++i) {
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j) {
        double u1 = o->ubcp[IR(o, i, j, k, 1)], u2 = o->ubcp[IR(o, i, j, k, 2)], u3 = o->ubcp[IR(o, i, j, k, 3)];
        size_t q = (size_t)(k - 1) * my + (j - 1);
        pr[0][q] = u2 * u2; /* ur22 */
        pr[1][q] = u3 * u3; /* ur33 */
        pr[2][q] = u1 * u2; /* ur12 */
        pr[3][q] = u1 * u3; /* ur13 */
        pr[4][q] = u2 * u3; /* ur23 */
        if (center) {
          FD(i, j, k, 1) = u1 * u1;
          FD(i, j, k, 2) = u1 * u2;
          FD(i, j, k, 3) = u1 * u3;
        } else {
          FD(i, j, k, 1) = FD(i, j, k, 2) = FD(i, j, k, 3) = 0.0;
        }
      }
    for (int q = 0; q < 5; ++q) orc_ff2rc(my, mz, pr[q], (double*)pf[q]);
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my2; ++j) {
        size_t q = (size_t)(k - 1) * my2 + (j - 1);
        double wy = o->wy[j - 1], wz = o->wz[k - 1];
        o->nl[IU(o, i, j, k, 1)] = cadd(cik(wy, pf[2][q]), cik(wz, pf[3][q]));
        o->nl[IU(o, i, j, k, 2)] = cadd(cik(wy, pf[0][q]), cik(wz, pf[4][q]));
        o->nl[IU(o, i, j, k, 3)] = cadd(cik(wy, pf[4][q]), cik(wz, pf[1][q]));
      }
  }
  for (int q = 0; q < 5; ++q) { free(pr[q]); free(pf[q]); }
  }
  if (center) {
    for (int n = 1; n <= 3; ++n)
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my; ++j) FD(o->mx, j, k, n) = FD(1, j, k, n);
  } else {
    /* MUSCL reconstruction + Rusanov flux, nl_allflow.F90:142-294; the
     * (j, k) lines are independent */
#pragma omp parallel
    {
    double* uplus = (double*)xcalloc((size_t)nx + 1, sizeof(double));
    double* uminus = (double*)xcalloc((size_t)nx + 1, sizeof(double));
    double* Uplus = (double*)xcalloc((size_t)nx + 1, sizeof(double));
    double* Uminus = (double*)xcalloc((size_t)nx + 1, sizeof(double));
    double* slope = (double*)xcalloc((size_t)nx + 1, sizeof(double));
    double slope_factor = 1.0;
#pragma omp for schedule(static)
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j)
        for (int idim = 1; idim <= 3; ++idim) {
          for (int i = 1; i <= nx; ++i) {
            int im1 = i - 1, ip1 = i + 1;
            if (im1 < 1) im1 = mxf + im1 - 1;
            if (ip1 > mxf - 1) ip1 = ip1 - (mxf - 1);
            double ui = o->ubcp[IR(o, i, j, k, idim)];
            double uip = o->ubcp[IR(o, ip1, j, k, idim)];
            double uim = o->ubcp[IR(o, im1, j, k, idim)];
            double phi_ri = 1.0;
            slope[i] = (uip - uim) / (2.0 * dx);
            if (o->cfg.limiter == PIB_LIMITER_MINMOD) {
              phi_ri = (ui - uim) / (uip - ui + SMALL_NUMBER);
              if (phi_ri <= 0.0) phi_ri = 0.0;
              else phi_ri = fmin(1.0, phi_ri);
            }
            slope[i] = slope[i] * phi_ri * slope_factor;
            if (idim == 1) {
              uplus[im1] = ui - slope[i] * (dx / 2.0);
              uminus[i] = ui + slope[i] * (dx / 2.0);
              Uplus[im1] = uplus[im1];
              Uminus[i] = uminus[i];
            } else {
              Uplus[im1] = ui - slope[i] * (dx / 2.0);
              Uminus[i] = ui + slope[i] * (dx / 2.0);
            }
            FD(im1, j, k, idim) = FD(im1, j, k, idim) + 0.5 * (uplus[im1] - 0.0) * (Uplus[im1] - 0.0);
            FD(i, j, k, idim) = FD(i, j, k, idim) + 0.5 * (uminus[i] - 0.0) * (Uminus[i] - 0.0);
            if (i != 1) {
              double max_speed = fmax(fabs(uminus[im1] - 0.0), fabs(uplus[im1] - 0.0));
              FD(im1, j, k, idim) = FD(im1, j, k, idim) - 0.5 * max_speed * (Uplus[im1] - Uminus[im1]);
            }
          }
          double max_speed = fmax(fabs(uminus[nx] - 0.0), fabs(uplus[nx] - 0.0));
          FD(nx, j, k, idim) = FD(nx, j, k, idim) - 0.5 * max_speed * (Uplus[nx] - Uminus[nx]);
          FD(o->mx, j, k, idim) = FD(1, j, k, idim);
        }
    free(uplus); free(uminus); free(Uplus); free(Uminus); free(slope);
    }
  }
#pragma omp parallel
  {
  planebuf w = pb_alloc(o);
  for (int idim = 1; idim <= 3; ++idim) {
#pragma omp for schedule(static)
    for (int i = 1; i <= nx; ++i) {
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my; ++j) {
          int im1 = i - 1, ip1 = i + 1;
          if (im1 < 1) im1 = mxf + im1 - 1;
          if (ip1 > mxf - 1) ip1 = ip1 - (mxf - 1);
          double v;
          if (center)
            v = (FD(ip1, j, k, idim) - FD(im1, j, k, idim)) / (2.0 * dx);
          else
            v = (FD(i, j, k, idim) - FD(im1, j, k, idim)) / dx;
          w.ra[(size_t)(k - 1) * my + (j - 1)] = v;
        }
      orc_ff2rc(my, mz, w.ra, (double*)w.cb);
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j)
          o->nl[IU(o, i, j, k, idim)] = cadd(o->nl[IU(o, i, j, k, idim)], w.cb[(size_t)(k - 1) * my2 + (j - 1)]);
    }
#pragma omp single
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my2; ++j) o->nl[IU(o, nx + 1, j, k, idim)] = o->nl[IU(o, 1, j, k, idim)];
  }
  pb_free(&w);
  }
#undef FD
  free(fd);
}

/* ======================================================================
 * bcset pieces, boundary_condition.F90
 * ====================================================================== */
/* calc_visc :1767-1815 */
static void calc_visc(orc* o) {
  int nx = o->nx, my = o->my, mz = o->mz, my2 = o->my2;
#pragma omp parallel
  {
  planebuf w = pb_alloc(o);
#pragma omp for schedule(static) collapse(2)
  for (int n = 1; n <= 3; ++n)
    for (int i = 1; i <= nx + 1; ++i) {
      int im = (i == 1) ? nx : i - 1;
      int ip = (i == nx + 1) ? 2 : i + 1;
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j) {
          cpx um = o->u[IU(o, im, j, k, n)], uc = o->u[IU(o, i, j, k, n)], up = o->u[IU(o, ip, j, k, n)];
          cpx d2 = cadd(csub(um, cscal(2.0, uc)), up);
          cpx f = cscal(1.0 / o->dx2, d2);
          f = csub(f, cscal(o->w2[(size_t)(k - 1) * my2 + (j - 1)], uc));
          f = cscal(o->vis, f);
          w.ca[(size_t)(k - 1) * my2 + (j - 1)] = f;
        }
      orc_ff2cr(my, mz, (const double*)w.ca, w.ra);
      put_rplane(o, o->diffn + (size_t)(n - 1) * (nx + 1) * my * mz, nx + 1, i, w.ra);
    }
  pb_free(&w);
  }
}

/* compute_omega :1648-1765 (omega stored in fr) */
static void compute_omega(orc* o) {
  int nx = o->nx, my = o->my, mz = o->mz, my2 = o->my2, mxf = o->mxf;
  double dx = o->dx;
  double* om = o->fr;
#pragma omp parallel
  {
  planebuf w = pb_alloc(o);
#pragma omp for schedule(static)
  for (int i = 1; i <= nx + 1; ++i) {
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my2; ++j)
        w.ca[(size_t)(k - 1) * my2 + (j - 1)] = cik(o->wy[j - 1], o->u[IU(o, i, j, k, 1)]);
    orc_ff2cr(my, mz, (const double*)w.ca, w.ra); /* dudy */
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j) {
        double dudy = w.ra[(size_t)(k - 1) * my + (j - 1)];
        double v;
        if (i == 1)
          v = (o->ubcp[IR(o, i + 1, j, k, 2)] - o->ubcp[IR(o, mxf - 1, j, k, 2)]) / (2.0 * dx) - dudy;
        else if (i == mxf)
          v = (o->ubcp[IR(o, 2, j, k, 2)] - o->ubcp[IR(o, mxf - 1, j, k, 2)]) / (2.0 * dx) - dudy;
        else
          v = (o->ubcp[IR(o, i + 1, j, k, 2)] - o->ubcp[IR(o, i - 1, j, k, 2)]) / (2.0 * dx) - dudy;
        om[IR(o, i, j, k, 3)] = v;
      }
  }
#pragma omp for schedule(static)
  for (int i = 1; i <= nx + 1; ++i) {
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my2; ++j)
        w.ca[(size_t)(k - 1) * my2 + (j - 1)] = cik(o->wz[k - 1], o->u[IU(o, i, j, k, 1)]);
    orc_ff2cr(my, mz, (const double*)w.ca, w.ra); /* dudz */
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j) {
        double dudz = w.ra[(size_t)(k - 1) * my + (j - 1)];
        double v;
        if (i == 1)
          v = dudz - (o->ubcp[IR(o, i + 1, j, k, 3)] - o->ubcp[IR(o, mxf - 1, j, k, 3)]) / (2.0 * dx);
        else if (i == mxf)
          v = dudz - (o->ubcp[IR(o, 2, j, k, 3)] - o->ubcp[IR(o, mxf - 1, j, k, 3)]) / (2.0 * dx);
        else
          v = dudz - (o->ubcp[IR(o, i + 1, j, k, 3)] - o->ubcp[IR(o, i - 1, j, k, 3)]) / (2.0 * dx);
        om[IR(o, i, j, k, 2)] = v;
      }
  }
#pragma omp for schedule(static)
  for (int i = 1; i <= nx + 1; ++i) {
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my2; ++j) {
        w.ca[(size_t)(k - 1) * my2 + (j - 1)] = cik(o->wz[k - 1], o->u[IU(o, i, j, k, 2)]); /* dvdz */
        w.cb[(size_t)(k - 1) * my2 + (j - 1)] = cik(o->wy[j - 1], o->u[IU(o, i, j, k, 3)]); /* dwdy */
      }
    orc_ff2cr(my, mz, (const double*)w.ca, w.ra);
    orc_ff2cr(my, mz, (const double*)w.cb, w.rb);
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j) {
        size_t q = (size_t)(k - 1) * my + (j - 1);
        om[IR(o, i, j, k, 1)] = w.rb[q] - w.ra[q];
      }
  }
  pb_free(&w);
  }
  memcpy(o->omega_copy, o->fr, sizeof(double) * (size_t)(nx + 1) * my * mz * 3);
}

/* calc_pgrad :1817-1896 (xperiodic, nlcons) */
static void calc_pgrad(orc* o) {
  int nx = o->nx, my = o->my, mz = o->mz, my2 = o->my2;
  double dx = o->dx;
  pib_scalars* s = &o->s;
#pragma omp parallel
  {
  planebuf w = pb_alloc(o);
  cpx* f3 = w.cc;
  double* r3 = w.rc;
#pragma omp for schedule(static)
  for (int i = 1; i <= nx; ++i) {
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my2; ++j) {
        size_t q = (size_t)(k - 1) * my2 + (j - 1);
        cpx f1;
        if (i == 1) f1 = cdivr(csub(o->p[IP(o, i + 1, j, k)], o->p[IP(o, nx, j, k)]), 2.0 * dx);
        else if (i == nx) f1 = cdivr(csub(o->p[IP(o, 1, j, k)], o->p[IP(o, i - 1, j, k)]), 2.0 * dx);
        else f1 = cdivr(csub(o->p[IP(o, i + 1, j, k)], o->p[IP(o, i - 1, j, k)]), 2.0 * dx);
        w.ca[q] = f1;
        w.cb[q] = cik(o->wy[j - 1], o->p[IP(o, i, j, k)]);
        f3[q] = cik(o->wz[k - 1], o->p[IP(o, i, j, k)]);
      }
    orc_ff2cr(my, mz, (const double*)w.ca, w.ra);
    orc_ff2cr(my, mz, (const double*)w.cb, w.rb);
    orc_ff2cr(my, mz, (const double*)f3, r3);
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j) {
        size_t q = (size_t)(k - 1) * my + (j - 1);
        o->ppr[IPPR(o, i, j, k, 1)] = w.ra[q] + s->mpg[0];
        o->ppr[IPPR(o, i, j, k, 2)] = w.rb[q] + s->mpg[1];
        o->ppr[IPPR(o, i, j, k, 3)] = r3[q] + s->mpg[2];
      }
  }
  pb_free(&w);
  }
}

/* ---- trilinear ('lpi' order 2) interpolation -------------------------
 * set_interpolation_stencil/_pstencil reduce to ib = pc for order 2 with
 * periodic x, y, z (dependent_functions.F90:124-156, 241-273); weights
 * w(i,j,k) = xval(i)*yval(j)*zval(k), shape2 (interpolation.F90:740-751). */
typedef struct { int ib, jb, kb; double w[8]; /* w[(k*2+j)*2+i] */ } stencil;

static void make_weights(stencil* st, const double pos[3]) {
  double zx = pos[0] - (double)st->ib, zy = pos[1] - (double)st->jb, zz = pos[2] - (double)st->kb;
  double xv[2] = {1.0 - zx, zx}, yv[2] = {1.0 - zy, zy}, zv[2] = {1.0 - zz, zz};
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) st->w[(k * 2 + j) * 2 + i] = xv[i] * yv[j] * zv[k];
}
static inline int wrapx(const orc* o, int ii) {
  if (ii < 1) ii = o->mxf + ii - 1;
  if (ii > o->mxf - 1) ii = ii - o->mxf + 1;
  return ii;
}
static inline int wrapy(const orc* o, int jj) {
  if (jj < 1) jj = o->my + jj;
  if (jj > o->my) jj = jj - o->my;
  return jj;
}
static inline int wrapz(const orc* o, int kk) {
  if (kk < 1) kk = o->mz + kk;
  if (kk > o->mz) kk = kk - o->mz;
  return kk;
}
static int cellof(double x) { return x < 0.0 ? (int)(x - 1.0) : (int)x; }

/* interp_threed_scalar loop order (k, j, i) */
static double interp_kji(const stencil* st, const double v[8]) {
  double s = 0.0;
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) s = s + v[(k * 2 + j) * 2 + i] * st->w[(k * 2 + j) * 2 + i];
  return s;
}
/* array_dot_product_3d loop order (i, j, k) */
static double dot_ijk(const stencil* st, const double v[8]) {
  double s = 0.0;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) s = s + v[(k * 2 + j) * 2 + i] * st->w[(k * 2 + j) * 2 + i];
  return s;
}

/* interpolate_udata, dependent_functions.F90:612-708 */
static void interpolate_udata(const orc* o, const int pc[3], const double pos[3], stencil* st,
                              double ul[3], double nll[3], double onll[3], double dfll[3], int flag) {
  st->ib = pc[0]; st->jb = pc[1]; st->kb = pc[2];
  make_weights(st, pos);
  double vs[3][8], ns[3][8], os[3][8], ds[3][8];
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        int ii = wrapx(o, st->ib + i), jj = wrapy(o, st->jb + j), kk = wrapz(o, st->kb + k);
        int q = (k * 2 + j) * 2 + i;
        for (int n = 1; n <= 3; ++n) {
          vs[n - 1][q] = o->ubcp[IR(o, ii, jj, kk, n)];
          if (flag == 1) {
            ns[n - 1][q] = o->nlbc[IR(o, ii, jj, kk, n)];
            os[n - 1][q] = o->onlbc[IR(o, ii, jj, kk, n)];
            ds[n - 1][q] = o->diffn[IR(o, ii, jj, kk, n)];
          }
        }
      }
  for (int n = 0; n < 3; ++n) ul[n] = interp_kji(st, vs[n]);
  if (flag == 1)
    for (int n = 0; n < 3; ++n) {
      nll[n] = dot_ijk(st, ns[n]);
      onll[n] = dot_ijk(st, os[n]);
      dfll[n] = dot_ijk(st, ds[n]);
    }
}
/* interpolate_pdata :710-781 */
static void interpolate_pdata(const orc* o, const int pc[3], const double pos[3], double ppl[3],
                              double* pl) {
  stencil st;
  st.ib = pc[0]; st.jb = pc[1]; st.kb = pc[2];
  make_weights(&st, pos);
  double gs[3][8], ps[8];
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        int ii = wrapx(o, st.ib + i), jj = wrapy(o, st.jb + j), kk = wrapz(o, st.kb + k);
        int q = (k * 2 + j) * 2 + i;
        for (int n = 1; n <= 3; ++n) gs[n - 1][q] = o->ppr[IPPR(o, ii, jj, kk, n)];
        ps[q] = o->pbcp[IPR(o, ii, jj, kk)];
      }
  for (int n = 0; n < 3; ++n) ppl[n] = interp_kji(&st, gs[n]);
  *pl = dot_ijk(&st, ps);
}

/* calc_pres_visc_drag :1481-1645 */
static void calc_pres_visc_drag(orc* o, int m, double da1) {
  pib_scalars* s = &o->s;
  double rb = o->radbdy[m - 1];
  for (int l = 1; l <= o->nbnd; ++l) {
    double xl[3], xpb[3];
    for (int n = 1; n <= 3; ++n) xl[n - 1] = XC(o, m, n) + BND(o, n, l) * rb;
    xpb[0] = xl[0] - 0.5;
    xpb[1] = xl[1];
    xpb[2] = xl[2];
    int pcellb[3], vcellb[3];
    for (int n = 0; n < 3; ++n) {
      pcellb[n] = cellof(xpb[n]);
      vcellb[n] = cellof(xl[n]);
    }
    double ppll[3], pl;
    interpolate_pdata(o, pcellb, xpb, ppll, &pl);
    stencil st;
    double ul[3];
    interpolate_udata(o, vcellb, xl, &st, ul, NULL, NULL, NULL, 0);
    double vort[3] = {0, 0, 0};
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
          int ii = wrapx(o, st.ib + i), jj = wrapy(o, st.jb + j), kk = wrapz(o, st.kb + k);
          for (int n = 1; n <= 3; ++n)
            vort[n - 1] = vort[n - 1] + st.w[(k * 2 + j) * 2 + i] * o->fr[IR(o, ii, jj, kk, n)];
        }
    double df[3];
    df[0] = vort[2] * CD(o, 2, l);
    df[0] = df[0] - vort[1] * CD(o, 3, l);
    df[1] = vort[0] * CD(o, 3, l);
    df[1] = df[1] - vort[2] * CD(o, 1, l);
    df[2] = vort[1] * CD(o, 1, l);
    df[2] = df[2] - vort[0] * CD(o, 2, l);
    df[0] = -df[0]; df[1] = -df[1]; df[2] = -df[2];
    double tempforce[3];
    for (int d = 1; d <= 3; ++d) {
      PB(o, pres, m, d) = PB(o, pres, m, d) - pl * CD(o, d, l) * da1;
      s->pres_total[d - 1] = s->pres_total[d - 1] + pl * CD(o, d, l) * da1;
      PB(o, visc, m, d) = PB(o, visc, m, d) + df[d - 1] * o->vis * da1;
      s->visc_total[d - 1] = s->visc_total[d - 1] + df[d - 1] * o->vis * da1;
      tempforce[d - 1] = df[d - 1] * o->vis * da1;
    }
    double b1 = BND(o, 1, l), b2 = BND(o, 2, l), b3 = BND(o, 3, l);
    double cr[3] = {b2 * tempforce[2] - b3 * tempforce[1], b3 * tempforce[0] - b1 * tempforce[2],
                    b1 * tempforce[1] - b2 * tempforce[0]};
    for (int n = 1; n <= 3; ++n) PB(o, torq, m, n) = PB(o, torq, m, n) + cr[n - 1] * rb * o->dx;
  }
}

/* compute_mpg_at_n :302-464 (fixed particles) */
static void compute_mpg_at_n(orc* o, int rks, double norm_factor) {
  pib_scalars* s = &o->s;
  compute_omega(o);
  for (int n = 0; n < 3; ++n) {
    s->visc_total_old[n] = s->visc_total[n];
    s->pres_total[n] = 0.0;
    s->visc_total[n] = 0.0;
  }
  for (int m = 1; m <= o->nbody; ++m) {
    if (rks == 1)
      for (int n = 1; n <= 3; ++n) PB(o, pres, m, n) = PB(o, visc, m, n) = PB(o, torq, m, n) = 0.0;
    double rdx = o->radbdy[m - 1] * o->dx;
    double da1 = 4.0 * PI * (rdx * rdx) / (double)o->nbnd;
    calc_pres_visc_drag(o, m, da1);
  }
  double pres_avg[3], visc_avg[3];
  for (int n = 1; n <= 3; ++n) {
    double a = 0.0, b = 0.0;
    for (int m = 1; m <= o->nbody; ++m) { a += PB(o, pres, m, n); b += PB(o, visc, m, n); }
    pres_avg[n - 1] = a;
    visc_avg[n - 1] = b;
  }
  s->pres_drag = sqrt(pres_avg[0] * pres_avg[0] + pres_avg[1] * pres_avg[1] + pres_avg[2] * pres_avg[2]);
  s->visc_drag = sqrt(visc_avg[0] * visc_avg[0] + visc_avg[1] * visc_avg[1] + visc_avg[2] * visc_avg[2]);
  s->pres_drag = s->pres_drag / norm_factor / (1.0 - s->maxvolfrac);
  s->visc_drag = s->visc_drag / norm_factor / (1.0 - s->maxvolfrac);
  s->total_drag = s->pres_drag + s->visc_drag;
  const double rhos = o->cfg.rhos, rhof = o->cfg.rhof, mv = s->mean_volfrac, vd = o->cfg.voldom;
  if (!o->cfg.set_mpg) {
    for (int n = 0; n < 3; ++n) {
      if (o->cfg.move_particles) { /* :389-397 */
        s->mpg[n] = s->cf * (s->ufmean_des[n] - s->ufmean[n]) + s->cf * (s->usmean[n] - s->usmean_des[n]);
        s->mpg[n] = s->mpg[n] + (s->pres_total[n] - (s->visc_total[n])) *
                                    (1.0 / (rhof * (1.0 - mv)) + 1.0 / (mv * rhos)) / vd;
        s->mpg[n] = s->mpg[n] / (1.0 / rhof - 1.0 / rhos);
        s->frame_accln[n] = -s->mpg[n] / rhos - (s->pres_total[n] - s->visc_total[n]) / (mv * rhos * vd) +
                            s->cf * (s->usmean_des[n] - s->usmean[n]);
      } else {
        s->mpg[n] = s->cf * (s->ufmean_des[n] - s->ufmean[n]) +
                    (s->pres_total[n] / o->cfg.voldom - (s->visc_total[n]) / o->cfg.voldom) / (1.0 - s->maxvolfrac);
      }
      if (o->cfg.include_frmeanfluid) s->mpg[n] = s->mpg[n] + s->frmeanfluid[n];
    }
  } else if (o->cfg.move_particles) { /* :414-417 */
    for (int n = 0; n < 3; ++n)
      s->frame_accln[n] = -s->mpg[n] / rhos - (s->pres_total[n] - s->visc_total[n]) / (mv * rhos * vd) +
                          s->cf * (s->usmean_des[n] - s->usmean[n]);
  }
  double c12 = o->coef[rks - 1][0] + o->coef[rks - 1][1];
  for (int m = 1; m <= o->nbody; ++m) {
    double vol = pow(2.0 * o->radbdy[m - 1] * o->dx, 3.0);
    for (int n = 1; n <= 3; ++n) {
      PB(o, force, m, n) = c12 * (PB(o, pres, m, n) + PB(o, visc, m, n)) - c12 * s->mpg[n - 1] * PI * vol / 6.0;
      PB(o, force_chem, m, n) = c12 * (PB(o, pres, m, n) + PB(o, visc, m, n));
    }
  }
}

/* calc_inner_reversal_forcing :1104-1479 (serial) */
static void calc_inner_reversal_forcing(orc* o, int m, int rks, double da2) {
  pib_scalars* s = &o->s;
  double rb = o->radbdy[m - 1], ri = o->radibdy[m - 1], ro = o->radobdy[m - 1];
  double drm = o->cfg.dr;
  double c1 = o->coef[rks - 1][0], c2 = o->coef[rks - 1][1], c3 = o->coef[rks - 1][2], c4 = o->coef[rks - 1][3];
  for (int l = 1; l <= o->nbnd; ++l) {
    if (!o->if_rev[(size_t)(m - 1) * o->nbnd + (l - 1)]) continue;
    double xli[3], xlo[3], rad = 0.0;
    for (int n = 1; n <= 3; ++n) {
      xli[n - 1] = XC(o, m, n) + BND(o, n, l) * ri;
      xlo[n - 1] = XC(o, m, n) + BND(o, n, l) * ro;
      rad = rad + (BND(o, n, l) * ri) * (BND(o, n, l) * ri);
    }
    rad = sqrt(rad);
    double snorm[3];
    for (int n = 1; n <= 3; ++n) snorm[n - 1] = (BND(o, n, l) * ri) / rad;
    double xpi[3] = {xli[0] - 0.5, xli[1], xli[2]};
    int vcello[3], pcelli[3], vcelli[3];
    for (int n = 0; n < 3; ++n) {
      vcello[n] = cellof(xlo[n]);
      pcelli[n] = cellof(xpi[n]);
      vcelli[n] = cellof(xli[n]);
    }
    stencil st;
    double ulo[3], uli[3], nll[3], onll[3], dfll[3], ppll[3], pl;
    interpolate_udata(o, vcello, xlo, &st, ulo, NULL, NULL, NULL, 0);
    interpolate_pdata(o, pcelli, xpi, ppll, &pl);
    interpolate_udata(o, vcelli, xli, &st, uli, nll, onll, dfll, 1);
    double unmag = 0.0, ubnmag = 0.0;
    double w1 = PB(o, angv, m, 1), w2v = PB(o, angv, m, 2), w3 = PB(o, angv, m, 3);
    double crossp[3] = {w2v * snorm[2] - w3 * snorm[1], w3 * snorm[0] - w1 * snorm[2], w1 * snorm[1] - w2v * snorm[0]};
    double linvel[3];
    for (int n = 1; n <= 3; ++n) linvel[n - 1] = PB(o, velbdy, m, n) + crossp[n - 1] * rb * o->dx;
    for (int d = 0; d < 3; ++d) {
      unmag = unmag + snorm[d] * ulo[d];
      ubnmag = ubnmag + snorm[d] * linvel[d];
    }
    double ucar[3];
    for (int d = 0; d < 3; ++d) {
      double unorm = snorm[d] * unmag;
      double utang = ulo[d] - unorm;
      double ubnorm = snorm[d] * ubnmag;
      double ubtang = linvel[d] - ubnorm;
      double utangi = ubtang * (ro - ri) - utang * (rb - ri);
      utangi = utangi / (ro - rb);
      double unormi;
      if (o->cfg.revernorm == 1) {
        unormi = ubnorm * (ro - ri) - unorm * (rb - ri);
        unormi = unormi / (ro - rb);
      } else {
        unormi = 0.0;
      }
      ucar[d] = utangi + unormi;
    }
    double ft[3];
    for (int n = 0; n < 3; ++n) {
      ft[n] = s->cf * (-ucar[n] + uli[n]);
      ft[n] = ft[n] - c3 * nll[n] - c4 * onll[n];
      ft[n] = ft[n] + (c1 + c2) * (ppll[n] - dfll[n]);
      ft[n] = ft[n] * da2 * drm * o->dx;
    }
    /* spreading with redistribution of the fluid-node share, :1379-1462 */
    int count_fl = 0;
    double force_fl[3] = {0, 0, 0};
    for (int k = 0; k < 2; ++k) {
      int kk = wrapz(o, st.kb + k);
      for (int j = 0; j < 2; ++j) {
        int jj = wrapy(o, st.jb + j);
        for (int i = 0; i < 2; ++i) {
          int ii = wrapx(o, st.ib + i);
          if (o->fluid[IFL(o, ii, jj, kk)]) {
            count_fl = count_fl + 1;
            for (int n = 0; n < 3; ++n) {
              double tmppa = st.w[(k * 2 + j) * 2 + i] * ft[n];
              force_fl[n] = force_fl[n] + tmppa / o->dx3;
            }
          }
        }
      }
    }
    int count_so = 8 - count_fl;
    double force_dist[3];
    for (int n = 0; n < 3; ++n) force_dist[n] = force_fl[n] / (double)(float)count_so;
    for (int k = 0; k < 2; ++k) {
      int kk = wrapz(o, st.kb + k);
      for (int j = 0; j < 2; ++j) {
        int jj = wrapy(o, st.jb + j);
        for (int i = 0; i < 2; ++i) {
          int ii = wrapx(o, st.ib + i);
          for (int n = 1; n <= 3; ++n) {
            double tmppa = st.w[(k * 2 + j) * 2 + i] * ft[n - 1];
            if (o->cfg.include_frmeanfluid) {
              if (o->fluid[IFL(o, ii, jj, kk)]) s->frmeanfluid[n - 1] = s->frmeanfluid[n - 1] + tmppa / o->dx3;
              o->fr[IR(o, ii, jj, kk, n)] = o->fr[IR(o, ii, jj, kk, n)] + tmppa / o->dx3;
            } else {
              if (!o->fluid[IFL(o, ii, jj, kk)])
                o->fr[IR(o, ii, jj, kk, n)] = o->fr[IR(o, ii, jj, kk, n)] + tmppa / o->dx3 + force_dist[n - 1];
            }
            s->frmean[n - 1] = s->frmean[n - 1] + tmppa / o->dx3;
          }
        }
      }
    }
  }
}

/* calc_bnd_forcing :902-1102 (serial): the direct forcing at every surface
 * point, interpolated as at an inner reversal point and spread with its
 * stencil weights to all eight nodes (fluid or solid, no redistribution).
 * linvel = velbdy + (angv x n) rb dx with n the point's unit normal cd; the
 * reference takes n from the module variable SNORM, which this routine does
 * not set (it holds a normal of the last reversal point handled), so the two
 * agree only for angv = 0 -- the case of every reference run, whose DEM
 * never updates OMEGA (collision_mod.F90:1386). */
static void calc_bnd_forcing(orc* o, int m, int rks, double da1) {
  pib_scalars* s = &o->s;
  double rb = o->radbdy[m - 1];
  double drm = o->cfg.dr;
  double c1 = o->coef[rks - 1][0], c2 = o->coef[rks - 1][1], c3 = o->coef[rks - 1][2], c4 = o->coef[rks - 1][3];
  for (int l = 1; l <= o->nbnd; ++l) {
    double xl[3];
    for (int n = 1; n <= 3; ++n) xl[n - 1] = XC(o, m, n) + BND(o, n, l) * rb;
    double xpb[3] = {xl[0] - 0.5, xl[1], xl[2]};
    int pcell[3], vcell[3];
    for (int n = 0; n < 3; ++n) {
      pcell[n] = cellof(xpb[n]);
      vcell[n] = cellof(xl[n]);
    }
    double ppll[3], pl;
    interpolate_pdata(o, pcell, xpb, ppll, &pl);
    stencil st;
    double ul[3], nll[3], onll[3], dfll[3];
    interpolate_udata(o, vcell, xl, &st, ul, nll, onll, dfll, 1);
    double w1 = PB(o, angv, m, 1), w2v = PB(o, angv, m, 2), w3 = PB(o, angv, m, 3);
    double sn[3] = {CD(o, 1, l), CD(o, 2, l), CD(o, 3, l)};
    double crossp[3] = {w2v * sn[2] - w3 * sn[1], w3 * sn[0] - w1 * sn[2], w1 * sn[1] - w2v * sn[0]};
    double ft[3];
    for (int n = 0; n < 3; ++n) {
      double linvel = PB(o, velbdy, m, n + 1) + crossp[n] * rb * o->dx;
      ft[n] = s->cf * (-linvel + ul[n]);
      ft[n] = ft[n] - c3 * nll[n] - c4 * onll[n];
      ft[n] = ft[n] + (c1 + c2) * (ppll[n] - dfll[n]);
      ft[n] = ft[n] * da1 * drm * o->dx;
    }
    for (int k = 0; k < 2; ++k) {
      int kk = wrapz(o, st.kb + k);
      for (int j = 0; j < 2; ++j) {
        int jj = wrapy(o, st.jb + j);
        for (int i = 0; i < 2; ++i) {
          int ii = wrapx(o, st.ib + i);
          for (int n = 1; n <= 3; ++n) {
            double tmppa = st.w[(k * 2 + j) * 2 + i] * ft[n - 1];
            o->fr[IR(o, ii, jj, kk, n)] = o->fr[IR(o, ii, jj, kk, n)] + tmppa / o->dx3;
            if (o->fluid[IFL(o, ii, jj, kk)] && o->cfg.include_frmeanfluid)
              s->frmeanfluid[n - 1] = s->frmeanfluid[n - 1] + tmppa / o->dx3;
            s->frmean[n - 1] = s->frmean[n - 1] + tmppa / o->dx3;
          }
        }
      }
    }
  }
}

/* bcset :467-900, serial.  The first half (:517-538 with the diagnostics of
 * :723-884, which read only force, force_chem and mpg) runs before the point
 * where the reference moves particles (:539-554); the second half is
 * calc_pgrad and the reversal forcing. */
static int bcset_drag(orc* o, int rks) {
  pib_scalars* s = &o->s;
  double norm_factor = (3.0 * PI * o->vis * (s->mixmeanslipmod + SMALL_NUMBER) * o->cfg.char_length) * (double)o->nbody;
  calc_visc(o);
  if (rks == 1)
    for (int m = 1; m <= o->nbody; ++m)
      for (int n = 1; n <= 3; ++n) PB(o, force, m, n) = 0.0;
  compute_mpg_at_n(o, rks, norm_factor);
  /* diagnostics, :723-884 (single phase) */
  double avg_force[3], avg_chem[3];
  for (int n = 1; n <= 3; ++n) {
    double a = 0.0, b = 0.0;
    for (int m = 1; m <= o->nbody; ++m) { a += PB(o, force, m, n); b += PB(o, force_chem, m, n); }
    avg_force[n - 1] = a / o->nbody;
    avg_chem[n - 1] = b / o->nbody;
  }
  /* avg_force = volfrac*avg_force_spec/mean_volfrac with one phase */
  for (int n = 0; n < 3; ++n) {
    avg_force[n] = (0.0 + s->mean_volfrac * avg_force[n]) / s->mean_volfrac;
    avg_chem[n] = (0.0 + s->mean_volfrac * avg_chem[n]) / s->mean_volfrac;
  }
  double nd = sqrt(avg_force[0] * avg_force[0] + avg_force[1] * avg_force[1] + avg_force[2] * avg_force[2]);
  double ndc = sqrt(avg_chem[0] * avg_chem[0] + avg_chem[1] * avg_chem[1] + avg_chem[2] * avg_chem[2]);
  double mpg_avg = sqrt(s->mpg[0] * s->mpg[0] + s->mpg[1] * s->mpg[1] + s->mpg[2] * s->mpg[2]);
  s->total_drag_mpg = mpg_avg * o->cfg.voldom / norm_factor;
  double nf;
  if (o->cfg.set_mpg)
    nf = (3.0 * PI * o->vis * (s->mixmeanslipmod + SMALL_NUMBER) * o->cfg.dia);
  else
    nf = (3.0 * PI * o->vis * o->cfg.ucharmod * o->cfg.dia);
  s->norm_drag = nd / nf;
  s->norm_drag_chem = ndc / nf;
  if (s->norm_drag > 0.0) s->ferror = fabs(s->norm_drag - s->fold) / s->norm_drag;
  else s->ferror = 1.0;
  s->fold = s->norm_drag;
  /* :828-833, once a step: the array shifts by one, the new ferror enters
   * at its head, and ferror_hist is its mean */
  if (rks == o->itrmax) {
    int n = o->nerr_steps;
    double tmp[PIB_NERR_MAX];
    for (int k = 0; k < n; ++k) tmp[k] = o->ferror_array[k];
    for (int k = 1; k < n; ++k) o->ferror_array[k] = tmp[k - 1];
    o->ferror_array[0] = s->ferror;
    double sum = 0.0;
    for (int k = 0; k < n; ++k) sum += o->ferror_array[k];
    o->ferror_hist = sum / (double)n;
  }
  if (s->norm_drag > (double)1.0E+06f) return fail(o, PIB_ERR_BLOWUP, "norm_drag > 1e6 (blow-up)");
  return PIB_OK;
}

static int bcset_forcing(orc* o, int rks) {
  pib_scalars* s = &o->s;
  int nx = o->nx, my = o->my, mz = o->mz, my2 = o->my2;
  calc_pgrad(o);
  for (int n = 0; n < 3; ++n) { s->frmean[n] = 0.0; s->frmeanfluid[n] = 0.0; }
  memset(o->fr, 0, sizeof(double) * (size_t)(nx + 1) * my * mz * 3);
  for (int m = 1; m <= o->nbody; ++m) {
    double rbx = o->radbdy[m - 1] * o->dx;
    double da1 = 4.0 * PI * (rbx * rbx) / (double)o->nbnd;
    double rdx = o->radibdy[m - 1] * o->dx;
    double da2 = 4.0 * PI * (rdx * rdx) / (double)o->nrpr_active[m - 1];
    if (o->cfg.dobnd) calc_bnd_forcing(o, m, rks, da1);
    if (o->cfg.dorpr) calc_inner_reversal_forcing(o, m, rks, da2);
  }
  for (int n = 0; n < 3; ++n) {
    s->frmean[n] = s->frmean[n] / (double)((long long)nx * my * mz);
    s->frmeanfluid[n] = s->frmeanfluid[n] / (double)o->count_fluid;
  }
#pragma omp parallel
  {
    planebuf w = pb_alloc(o);
#pragma omp for schedule(static) collapse(2)
    for (int n = 1; n <= 3; ++n)
      for (int i = 1; i <= nx; ++i) {
        for (int j = 1; j <= my; ++j)
          for (int k = 1; k <= mz; ++k) o->fr[IR(o, i, j, k, n)] = o->fr[IR(o, i, j, k, n)] - s->frmean[n - 1];
        get_rplane(o, o->fr + (size_t)(n - 1) * (nx + 1) * my * mz, nx + 1, i, w.ra);
        orc_ff2rc(my, mz, w.ra, (double*)w.ca);
        for (int k = 1; k <= mz; ++k)
          for (int j = 1; j <= my2; ++j) o->ff[IU(o, i, j, k, n)] = w.ca[(size_t)(k - 1) * my2 + (j - 1)];
      }
    pb_free(&w);
  }
  for (int n = 1; n <= 3; ++n) {
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my2; ++j) o->ff[IU(o, nx + 1, j, k, n)] = o->ff[IU(o, 1, j, k, n)];
  }
  return PIB_OK;
}

static int bcset(orc* o, int rks) {
  int st = bcset_drag(o, rks);
  if (st) return st;
  return bcset_forcing(o, rks);
}

/* nonlinear, nl.F90:63-205; part 0 whole, 1 up to the drag, 2 the forcing */
static int nonlinear_part(orc* o, int rks, int part) {
  if (part == 2) return o->nbody > 0 ? bcset_forcing(o, rks) : PIB_OK;
  int nx = o->nx, my = o->my, mz = o->mz, my2 = o->my2;
  size_t nc = (size_t)(nx + 1) * my2 * mz * 3;
  for (size_t q = 0; q < nc; ++q) { o->onl[q] = o->nl[q]; o->nl[q] = C(0.0, 0.0); }
  form_nl(o);
  for (size_t q = 0; q < nc; ++q) o->nl[q] = C(-o->nl[q].re, -o->nl[q].im);
#pragma omp parallel
  {
    planebuf w = pb_alloc(o);
#pragma omp for schedule(static) collapse(2)
    for (int n = 1; n <= 3; ++n)
      for (int i = 1; i <= nx + 1; ++i) {
        for (int k = 1; k <= mz; ++k)
          for (int j = 1; j <= my2; ++j) {
            w.ca[(size_t)(k - 1) * my2 + (j - 1)] = o->nl[IU(o, i, j, k, n)];
            w.cb[(size_t)(k - 1) * my2 + (j - 1)] = o->onl[IU(o, i, j, k, n)];
          }
        orc_ff2cr(my, mz, (const double*)w.ca, w.ra);
        orc_ff2cr(my, mz, (const double*)w.cb, w.rb);
        put_rplane(o, o->nlbc + (size_t)(n - 1) * (nx + 1) * my * mz, nx + 1, i, w.ra);
        put_rplane(o, o->onlbc + (size_t)(n - 1) * (nx + 1) * my * mz, nx + 1, i, w.rb);
      }
    pb_free(&w);
  }
  if (o->nbody > 0) return part == 1 ? bcset_drag(o, rks) : bcset(o, rks);
  return PIB_OK;
}

int orc_nonlinear(void* ctx, int rks) {
  orc* o = (orc*)ctx;
  if (!o || rks < 1 || rks > o->itrmax) return PIB_ERR_ARG;
  return nonlinear_part(o, rks, 0);
}

int orc_nonlinear_begin(void* ctx, int rks) {
  orc* o = (orc*)ctx;
  if (!o || rks < 1 || rks > o->itrmax) return PIB_ERR_ARG;
  int st = nonlinear_part(o, rks, 1);
  o->nl_pending = st ? 0 : rks;
  return st;
}

int orc_nonlinear_finish(void* ctx, int rks) {
  orc* o = (orc*)ctx;
  if (!o) return PIB_ERR_ARG;
  if (o->nl_pending != rks) return fail(o, PIB_ERR_STATE, "pib_nonlinear_begin with this rks first");
  o->nl_pending = 0;
  return nonlinear_part(o, rks, 2);
}

/* ======================================================================
 * velstep: pressure / ustep / divcorr, poisson.F90:55-508,
 * usteptime.F90:45-258, steptheflow.F90:52-240
 * ====================================================================== */
static int pressure(orc* o, int rks, int j, int k) {
  int nx = o->nx;
  double dx = o->dx, dx2 = o->dx2;
  double w2 = o->w2[(size_t)(k - 1) * o->my2 + (j - 1)], wy = o->wy[j - 1], wz = o->wz[k - 1];
  double c1 = o->coef[rks - 1][0], c2 = o->coef[rks - 1][1], c3 = o->coef[rks - 1][2], c4 = o->coef[rks - 1][3];
  double *a = (double*)malloc(sizeof(double) * nx), *b = (double*)malloc(sizeof(double) * nx),
         *c = (double*)malloc(sizeof(double) * nx), *dv = (double*)calloc(nx, sizeof(double)),
         *dvi = (double*)calloc(nx, sizeof(double)), *dvt = (double*)calloc(nx, sizeof(double)),
         *dvti = (double*)calloc(nx, sizeof(double));
  for (int i = 0; i < nx; ++i) {
    a[i] = 1.0 / dx2 - 0.25 * w2;
    b[i] = -2.0 / dx2 - 0.5 * w2;
    c[i] = 1.0 / dx2 - 0.25 * w2;
  }
  for (int i = 1; i <= nx; ++i) {
    cpx tmpc = C(0.0, 0.0);
    cpx nl1 = csub(o->nl[IU(o, i + 1, j, k, 1)], o->nl[IU(o, i, j, k, 1)]);
    cpx ff1 = csub(o->ff[IU(o, i + 1, j, k, 1)], o->ff[IU(o, i, j, k, 1)]);
    tmpc = cadd(cadd(tmpc, cdivr(cscal(c3, nl1), dx)), cdivr(ff1, dx));
    cpx nl2 = cadd(o->nl[IU(o, i + 1, j, k, 2)], o->nl[IU(o, i, j, k, 2)]);
    cpx ff2 = cadd(o->ff[IU(o, i + 1, j, k, 2)], o->ff[IU(o, i, j, k, 2)]);
    tmpc = cadd(cadd(tmpc, cdivr(cik(c3 * wy, nl2), 2.0)), cdivr(cik(wy, ff2), 2.0));
    cpx nl3 = cadd(o->nl[IU(o, i + 1, j, k, 3)], o->nl[IU(o, i, j, k, 3)]);
    cpx ff3 = cadd(o->ff[IU(o, i + 1, j, k, 3)], o->ff[IU(o, i, j, k, 3)]);
    tmpc = cadd(cadd(tmpc, cdivr(cik(c3 * wz, nl3), 2.0)), cdivr(cik(wz, ff3), 2.0));
    cpx on1 = csub(o->onl[IU(o, i + 1, j, k, 1)], o->onl[IU(o, i, j, k, 1)]);
    tmpc = cadd(tmpc, cdivr(cscal(c4, on1), dx));
    cpx on2 = cadd(o->onl[IU(o, i + 1, j, k, 2)], o->onl[IU(o, i, j, k, 2)]);
    tmpc = cadd(tmpc, cdivr(cik(c4 * wy, on2), 2.0));
    cpx on3 = cadd(o->onl[IU(o, i + 1, j, k, 3)], o->onl[IU(o, i, j, k, 3)]);
    tmpc = cadd(tmpc, cdivr(cik(c4 * wz, on3), 2.0));
    tmpc = cscal(0.5 * (1.0 + 1.0), tmpc);
    dv[i - 1] = tmpc.re / (c1 + c2);
    dvi[i - 1] = tmpc.im / (c1 + c2);
  }
  int st = 0;
  if (j == 1 && k == 1) {
    st = orc_tridag(a, b, c, dv, dvt, nx - 1);
    dvt[nx - 1] = 0.0;
    for (int i = 0; i < nx; ++i) dvti[i] = 0.0;
  } else {
    st = sm_solve(a, b, c, dv, dvi, dvt, dvti, nx);
  }
  for (int i = 1; i <= nx; ++i) {
    cpx tmpp = C(dvt[i - 1], dvti[i - 1]);
    cpx pold = o->p[IP(o, i, j, k)];
    o->p[IP(o, i, j, k)] = cadd(pold, cscal(1.0, csub(tmpp, pold)));
  }
  free(a); free(b); free(c); free(dv); free(dvi); free(dvt); free(dvti);
  return st;
}

static int ustep(orc* o, int rks, int j, int k) {
  int nx = o->nx;
  double dx = o->dx, dx2 = o->dx2, dt = o->s.dt, vis = o->vis;
  double w2 = o->w2[(size_t)(k - 1) * o->my2 + (j - 1)], wy = o->wy[j - 1], wz = o->wz[k - 1];
  double c1 = o->coef[rks - 1][0], c2 = o->coef[rks - 1][1], c3 = o->coef[rks - 1][2], c4 = o->coef[rks - 1][3];
  double g = 1.0;
  double *a = (double*)malloc(sizeof(double) * nx), *b = (double*)malloc(sizeof(double) * nx),
         *c = (double*)malloc(sizeof(double) * nx), *bb = (double*)malloc(sizeof(double) * nx);
  double *tr = (double*)malloc(sizeof(double) * nx), *ti = (double*)malloc(sizeof(double) * nx);
  double *xr = (double*)malloc(sizeof(double) * nx), *xi = (double*)malloc(sizeof(double) * nx);
  cpx* r = (cpx*)malloc(sizeof(cpx) * nx);
  int st = 0;
  for (int i = 0; i < nx; ++i) {
    a[i] = (-g * dt * vis * c2 / dx2);
    b[i] = (1.0 + dt * vis * c2 * (g * 2.0 / dx2 + w2));
    c[i] = (-g * dt * vis * c2 / dx2);
  }
  for (int idim = 1; idim <= 3; ++idim) {
    for (int i = 1; i <= nx; ++i) {
      cpx uu = o->u[IU(o, i, j, k, idim)];
      cpx ri = cscal(1.0 - dt * vis * c1 * w2, uu);
      ri = cadd(cadd(ri, cscal(dt * c3, o->nl[IU(o, i, j, k, idim)])), cscal(dt, o->ff[IU(o, i, j, k, idim)]));
      ri = cadd(ri, cscal(dt * c4, o->onl[IU(o, i, j, k, idim)]));
      r[i - 1] = ri;
    }
    double c12 = c1 + c2;
    for (int i = 2; i <= nx; ++i) {
      cpx pi_ = o->p[IP(o, i, j, k)], pm = o->p[IP(o, i - 1, j, k)];
      if (idim == 1)
        r[i - 1] = csub(r[i - 1], cdivr(cscal(dt * c12, csub(pi_, pm)), dx));
      else if (idim == 2)
        r[i - 1] = csub(r[i - 1], cik(wy, cscal(dt * c12 * 0.5, cadd(pm, pi_))));
      else
        r[i - 1] = csub(r[i - 1], cik(wz, cscal(dt * c12 * 0.5, cadd(pm, pi_))));
      cpx um = o->u[IU(o, i - 1, j, k, idim)], up = o->u[IU(o, i + 1, j, k, idim)], uc = o->u[IU(o, i, j, k, idim)];
      cpx lap = csub(cadd(um, up), cscal(2.0, uc));
      r[i - 1] = cadd(r[i - 1], cdivr(cscal(g * dt * vis * c1, lap), dx2));
    }
    {
      cpx p1 = o->p[IP(o, 1, j, k)], pn = o->p[IP(o, nx, j, k)];
      if (idim == 1)
        r[0] = csub(r[0], cdivr(cscal(dt * c12, csub(p1, pn)), dx));
      else if (idim == 2)
        r[0] = csub(r[0], cik(wy, cscal(dt * c12 * 0.5, cadd(pn, p1))));
      else
        r[0] = csub(r[0], cik(wz, cscal(dt * c12 * 0.5, cadd(pn, p1))));
      cpx u2 = o->u[IU(o, 2, j, k, idim)], u1 = o->u[IU(o, 1, j, k, idim)], un = o->u[IU(o, nx, j, k, idim)];
      cpx lap = cadd(csub(u2, cscal(2.0, u1)), un);
      r[0] = cadd(r[0], cdivr(cscal(g * dt * vis * c1, lap), dx2));
    }
    for (int i = 0; i < nx; ++i) { tr[i] = r[i].re; ti[i] = r[i].im; bb[i] = b[i]; }
    int s2 = sm_solve(a, bb, c, tr, ti, xr, xi, nx);
    if (s2) st = s2;
    for (int i = 1; i <= nx; ++i) o->u[IU(o, i, j, k, idim)] = C(xr[i - 1], xi[i - 1]);
    o->u[IU(o, nx + 1, j, k, idim)] = o->u[IU(o, 1, j, k, idim)];
  }
  free(a); free(b); free(c); free(bb); free(tr); free(ti); free(xr); free(xi); free(r);
  return st;
}

static int divcorr(orc* o, int rks, int j, int k) {
  int nx = o->nx, mx1 = o->nx;
  double dx = o->dx, dx2 = o->dx2, dt = o->s.dt, vis = o->vis;
  double w2 = o->w2[(size_t)(k - 1) * o->my2 + (j - 1)], wy = o->wy[j - 1], wz = o->wz[k - 1];
  double c1 = o->coef[rks - 1][0], c2 = o->coef[rks - 1][1];
  double alp = ((c1 + c2) * dt);
  double *a = (double*)malloc(sizeof(double) * nx), *b = (double*)malloc(sizeof(double) * nx),
         *c = (double*)malloc(sizeof(double) * nx), *dv = (double*)calloc(nx, sizeof(double)),
         *dvi = (double*)calloc(nx, sizeof(double)), *dv1 = (double*)calloc(nx, sizeof(double)),
         *dv1i = (double*)calloc(nx, sizeof(double));
  cpx* div = (cpx*)malloc(sizeof(cpx) * (nx + 1));
  for (int i = 0; i < nx; ++i) {
    a[i] = 1.0 / dx2 - 0.25 * w2;
    b[i] = -2.0 / dx2 - 0.5 * w2;
    c[i] = 1.0 / dx2 - 0.25 * w2;
  }
  for (int i = 1; i <= nx; ++i) {
    cpx t = cdivr(csub(o->u[IU(o, i + 1, j, k, 1)], o->u[IU(o, i, j, k, 1)]), dx);
    t = cadd(t, cscal(0.5, cik(wy, cadd(o->u[IU(o, i, j, k, 2)], o->u[IU(o, i + 1, j, k, 2)]))));
    t = cadd(t, cscal(0.5, cik(wz, cadd(o->u[IU(o, i, j, k, 3)], o->u[IU(o, i + 1, j, k, 3)]))));
    dv[i - 1] = 0.5 * (1.0 + 1.0) * t.re;
    dvi[i - 1] = 0.5 * (1.0 + 1.0) * t.im;
    dv[i - 1] = dv[i - 1] / alp;
    dvi[i - 1] = dvi[i - 1] / alp;
  }
  int st;
  if (j == 1 && k == 1) {
    st = orc_tridag(a, b, c, dv, dv1, nx - 1);
    dv1[nx - 1] = 0.0;
    for (int i = 0; i < nx; ++i) dv1i[i] = 0.0;
  } else {
    st = sm_solve(a, b, c, dv, dvi, dv1, dv1i, nx);
  }
  for (int i = 1; i <= nx; ++i) div[i] = C(dv1[i - 1], dv1i[i - 1]);
  /* u(t+dt) = u* - grad(phi) */
  {
    cpx d1 = div[1], dn = div[mx1];
    o->u[IU(o, 1, j, k, 1)] = csub(o->u[IU(o, 1, j, k, 1)], cdivr(cscal(alp, csub(d1, dn)), dx));
    o->u[IU(o, 1, j, k, 2)] = csub(o->u[IU(o, 1, j, k, 2)], cscal(0.5, cik(alp * wy, cadd(d1, dn))));
    o->u[IU(o, 1, j, k, 3)] = csub(o->u[IU(o, 1, j, k, 3)], cscal(0.5, cik(alp * wz, cadd(d1, dn))));
  }
  for (int i = 2; i <= nx; ++i) {
    o->u[IU(o, i, j, k, 1)] = csub(o->u[IU(o, i, j, k, 1)], cdivr(cscal(alp, csub(div[i], div[i - 1])), dx));
    o->u[IU(o, i, j, k, 2)] = csub(o->u[IU(o, i, j, k, 2)], cik(alp * 0.5 * wy, cadd(div[i - 1], div[i])));
    o->u[IU(o, i, j, k, 3)] = csub(o->u[IU(o, i, j, k, 3)], cik(alp * 0.5 * wz, cadd(div[i - 1], div[i])));
  }
  for (int n = 1; n <= 3; ++n) o->u[IU(o, nx + 1, j, k, n)] = o->u[IU(o, 1, j, k, n)];
  double cv = (c2 * dt) * vis;
  for (int i = 1; i <= nx; ++i) {
    cpx dm = (i == 1) ? div[mx1] : div[i - 1];
    cpx dp = (i == nx) ? div[1] : div[i + 1];
    cpx lap = cadd(csub(dm, cscal(2.0, div[i])), dp);
    cpx corr = csub(cdivr(lap, dx2), cscal(w2, div[i]));
    o->p[IP(o, i, j, k)] = csub(cadd(o->p[IP(o, i, j, k)], div[i]), cscal(cv, corr));
  }
  free(a); free(b); free(c); free(dv); free(dvi); free(dv1); free(dv1i); free(div);
  return st;
}

int orc_velstep(void* ctx, int rks) {
  orc* o = (orc*)ctx;
  if (!o || rks < 1 || rks > o->itrmax) return PIB_ERR_ARG;
  int st = 0;
  /* the modes are independent; a zero pivot anywhere is reported */
#pragma omp parallel for schedule(static) collapse(2) reduction(min : st)
  for (int j = 1; j <= o->my2; ++j)
    for (int k = 1; k <= o->mz; ++k) {
      int s2 = pressure(o, rks, j, k);
      if (s2 < st) st = s2;
    }
#pragma omp parallel for schedule(static) collapse(2) reduction(min : st)
  for (int j = 1; j <= o->my2; ++j)
    for (int k = 1; k <= o->mz; ++k) {
      int s2 = ustep(o, rks, j, k);
      if (s2 < st) st = s2;
    }
#pragma omp parallel for schedule(static) collapse(2) reduction(min : st)
  for (int j = 1; j <= o->my2; ++j)
    for (int k = 1; k <= o->mz; ++k) {
      int s2 = divcorr(o, rks, j, k);
      if (s2 < st) st = s2;
    }
  /* check_divergence, steptheflow.F90:219-240 */
  double divmax = 0.0;
  for (int k = 1; k <= o->mz; ++k)
    for (int j = 1; j <= o->my2; ++j)
      for (int i = 1; i <= o->nx; ++i) {
        cpx t = cdivr(csub(o->u[IU(o, i + 1, j, k, 1)], o->u[IU(o, i, j, k, 1)]), o->dx);
        t = cadd(t, cscal(0.5, cik(o->wy[j - 1], cadd(o->u[IU(o, i, j, k, 2)], o->u[IU(o, i + 1, j, k, 2)]))));
        t = cadd(t, cscal(0.5, cik(o->wz[k - 1], cadd(o->u[IU(o, i, j, k, 3)], o->u[IU(o, i + 1, j, k, 3)]))));
        double m = sqrt(t.re * t.re + t.im * t.im);
        if (m > divmax) divmax = m;
      }
  o->s.divmax = divmax;
  for (int n = 0; n < 3; ++n)
    o->s.umean[n] = o->s.umean[n] + (-o->s.mpg[n] + o->s.frmean[n] - o->s.frame_accln[n]) * o->s.dt;
  if (o->cfg.move_particles) /* steptheflow.F90:101-104 */
    for (int n = 0; n < 3; ++n) {
      o->s.frame_vel[n] = o->s.frame_vel[n] + o->s.frame_accln[n] * o->s.dt;
      o->s.frame_pos[n] = o->s.frame_pos[n] + o->s.frame_vel[n] * o->s.dt;
    }
  if (st) return fail(o, st, "zero pivot in tridiagonal solve");
  return PIB_OK;
}

/* The pressure predictor alone (poisson.F90:55-239 for every mode), so that
 * its right-hand side and symbol can be pinned: in a periodic box the later
 * ustep/divcorr pair cancels whatever the predictor leaves in p, so the
 * predictor cannot be seen in velstep's results. */
int orc_pressure_modes(void* ctx, int rks) {
  orc* o = (orc*)ctx;
  if (!o || rks < 1 || rks > o->itrmax) return PIB_ERR_ARG;
  int st = 0;
#pragma omp parallel for schedule(static) collapse(2) reduction(min : st)
  for (int j = 1; j <= o->my2; ++j)
    for (int k = 1; k <= o->mz; ++k) {
      int s2 = pressure(o, rks, j, k);
      if (s2 < st) st = s2;
    }
  return st ? fail(o, st, "zero pivot in tridiagonal solve") : PIB_OK;
}

int orc_step(void* ctx, int nsteps) {
  orc* o = (orc*)ctx;
  if (!o || nsteps < 0) return PIB_ERR_ARG;
  for (int n = 0; n < nsteps; ++n)
    for (int rks = 1; rks <= o->itrmax; ++rks) {
      int st;
      if ((st = orc_compute_new_timestep(o, rks))) return st;
      if ((st = orc_nonlinear(o, rks))) return st;
      if ((st = orc_velstep(o, rks))) return st;
    }
  return PIB_OK;
}

/* ======================================================================
 * scalar transport, one species: the per-mode solve of scalstep
 * (stepthescalar.F90:51-241) without nlphi_graduphi, which supplies
 * nlphif, onlphif and ffphi
 * ====================================================================== */
int orc_scalar_set(void* ctx, double gamma, double heat_ratio, double heat_ratio_new) {
  orc* o = (orc*)ctx;
  if (!o || !(gamma > 0) || !(heat_ratio > 0) || !(heat_ratio_new > 0)) return PIB_ERR_ARG;
  if (!o->phif) {
    size_t n = (size_t)(o->nx + 1) * o->my2 * o->mz;
    cpx* f[4];
    for (int q = 0; q < 4; ++q) f[q] = (cpx*)calloc(n, sizeof(cpx));
    if (!f[0] || !f[1] || !f[2] || !f[3]) {
      for (int q = 0; q < 4; ++q) free(f[q]);
      return fail(o, PIB_ERR_ARG, "no host memory for the scalar fields");
    }
    o->nlphif = f[1];
    o->onlphif = f[2];
    o->ffphi = f[3];
    o->phif = f[0]; /* marks the fields as present */
    o->heat_ratio_old = heat_ratio;
  }
  o->sc_gamma = gamma;
  o->heat_ratio = heat_ratio;
  o->heat_ratio_new = heat_ratio_new;
  return PIB_OK;
}

int orc_scalar_ratios(void* ctx, const double* r) {
  orc* o = (orc*)ctx;
  if (!o || !r || !(r[0] > 0) || !(r[1] > 0) || !(r[2] > 0)) return PIB_ERR_ARG;
  if (!o->phif) return fail(o, PIB_ERR_STATE, "pib_scalar_set first");
  o->heat_ratio = r[0];
  o->heat_ratio_new = r[1];
  o->heat_ratio_old = r[2];
  return PIB_OK;
}

int orc_scalar_get(void* ctx, double* out) {
  orc* o = (orc*)ctx;
  if (!o || !out) return PIB_ERR_ARG;
  if (!o->phif) return fail(o, PIB_ERR_STATE, "pib_scalar_set first");
  out[0] = o->heat_ratio;
  out[1] = o->heat_ratio_new;
  out[2] = o->heat_ratio_old;
  out[3] = o->sc_gamma;
  return PIB_OK;
}

#define IS(o, i, j, k) ((size_t)((i) - 1) + (size_t)((o)->nx + 1) * ((size_t)((j) - 1) + (size_t)(o)->my2 * ((k) - 1)))

/* phistep (stepthescalar.F90:97-241), serial, x periodic, no source */
static int phistep(orc* o, int rks, int j, int k) {
  int nx = o->nx;
  double dt = o->s.dt, g = o->sc_gamma, dx2 = o->dx2;
  double w2 = o->w2[(size_t)(k - 1) * o->my2 + (j - 1)];
  double c1 = o->coef[rks - 1][0], c2 = o->coef[rks - 1][1], c3 = o->coef[rks - 1][2], c4 = o->coef[rks - 1][3];
  double hr = o->heat_ratio, hrn = o->heat_ratio_new;
  double alpha = hrn, beta = 1.0 / hrn;
  double *a = (double*)calloc(nx, sizeof(double)), *b = (double*)calloc(nx, sizeof(double)),
         *c = (double*)calloc(nx, sizeof(double)), *auxu = (double*)calloc(nx, sizeof(double)),
         *auxv = (double*)calloc(nx, sizeof(double)), *auxz = (double*)calloc(nx, sizeof(double)),
         *tmp = (double*)calloc(nx, sizeof(double)), *tmpi = (double*)calloc(nx, sizeof(double)),
         *ut = (double*)calloc(nx, sizeof(double)), *uti = (double*)calloc(nx, sizeof(double));
  cpx* r1 = (cpx*)calloc(nx, sizeof(cpx));
  cpx* phif = o->phif;
  for (int i = 1; i <= nx; ++i) {
    a[i - 1] = -(dt * g * c2 / dx2);
    b[i - 1] = 1.0 + dt * g * c2 * (2.0 / dx2 + w2);
    c[i - 1] = -(dt * g * c2 / dx2);
    r1[i - 1] = cscal(1.0 - dt * c1 * g * w2, phif[IS(o, i, j, k)]);
    r1[i - 1] = cadd(cadd(r1[i - 1], cscal(dt * c3, o->nlphif[IS(o, i, j, k)])), cscal(dt, o->ffphi[IS(o, i, j, k)]));
    r1[i - 1] = cadd(r1[i - 1], cscal(dt * c4, o->onlphif[IS(o, i, j, k)]));
  }
  a[0] = a[0] * hrn;
  c[nx - 1] = c[nx - 1] / hrn;
  for (int i = 2; i <= nx; ++i) {
    cpx lap = csub(cadd(phif[IS(o, i - 1, j, k)], phif[IS(o, i + 1, j, k)]), cscal(2.0, phif[IS(o, i, j, k)]));
    r1[i - 1] = cadd(r1[i - 1], cdivr(cscal(dt * g * c1, lap), dx2));
  }
  {
    cpx lap = cadd(csub(phif[IS(o, 2, j, k)], cscal(2.0, phif[IS(o, 1, j, k)])), cscal(hr, phif[IS(o, nx, j, k)]));
    r1[0] = cadd(r1[0], cdivr(cscal(dt * g * c1, lap), dx2));
  }
  for (int i = 0; i < nx; ++i) {
    tmp[i] = r1[i].re;
    tmpi[i] = r1[i].im;
  }
  auxu[0] = alpha;
  auxu[nx - 1] = beta;
  auxv[0] = c[nx - 1] / auxu[nx - 1];
  auxv[nx - 1] = a[0] / auxu[0];
  b[0] = b[0] - auxu[0] * auxv[0];
  b[nx - 1] = b[nx - 1] - auxu[nx - 1] * auxv[nx - 1];
  int st = tridag3(a, b, c, tmp, auxu, tmpi, ut, auxz, uti, nx);
  double vdotz = auxv[0] * auxz[0] + auxv[nx - 1] * auxz[nx - 1];
  double vdotyr = auxv[0] * ut[0] + auxv[nx - 1] * ut[nx - 1];
  double vdotyi = auxv[0] * uti[0] + auxv[nx - 1] * uti[nx - 1];
  for (int i = 1; i <= nx; ++i)
    phif[IS(o, i, j, k)] = C(ut[i - 1] - vdotyr * auxz[i - 1] / (1.0 + vdotz),
                             uti[i - 1] - vdotyi * auxz[i - 1] / (1.0 + vdotz));
  phif[IS(o, nx + 1, j, k)] = cdivr(phif[IS(o, 1, j, k)], hrn);
  free(a); free(b); free(c); free(auxu); free(auxv); free(auxz); free(tmp); free(tmpi); free(ut); free(uti);
  free(r1);
  return st;
}

/* scalstep's mode loop and heat_ratio update (stepthescalar.F90:67-91) */
int orc_scalstep_modes(void* ctx, int rks) {
  orc* o = (orc*)ctx;
  if (!o || rks < 1 || rks > o->itrmax) return PIB_ERR_ARG;
  if (!o->phif) return fail(o, PIB_ERR_STATE, "pib_scalar_set first");
  int st = 0;
#pragma omp parallel for schedule(static) collapse(2) reduction(min : st)
  for (int j = 1; j <= o->my2; ++j)
    for (int k = 1; k <= o->mz; ++k) {
      int r = phistep(o, rks, j, k);
      if (r < st) st = r;
    }
  o->heat_ratio_old = o->heat_ratio;
  o->heat_ratio = o->heat_ratio_new;
  if (st) return fail(o, st, "zero pivot in the scalar solve");
  return PIB_OK;
}

/* ======================================================================
 * The whole scalstep (stepthescalar.F90:51-95), serial, x periodic with the
 * heat-ratio twist phi(x + L) = phi(x) / heat_ratio, one species (the
 * reference's nscalmax = 1), no source term:
 *   nlphi_graduphi (nlphi.F90:56-155): onlphif <- nlphif, form_nlphi
 *     (nl_allphi.F90:59-430) with compute_phim_heat_ratio (:431-536), the
 *     physical nl and onl of phi;
 *   bcsetscal (bc_scalar.F90:65-383): calc_diffn (:935-1033), per body
 *     calc_bndrev_data (:385-556: the surface flux, and the surface forcing
 *     when DOBND) and calc_inner_rev_data (:558-932) when dorpr, then
 *     ffphi = ff2rc(fr);
 *   the mode solve and the heat-ratio roll (orc_scalstep_modes).
 * Physical arrays hold planes 0..nx+1 of the reference (IS0).  The velocity
 * is ubcp of the stage (compute_new_timestep), as in the reference, where
 * velstep leaves it in place.
 * ====================================================================== */
#define IS0(o, i, j, k) ((((size_t)((k)-1)) * (o)->my + ((j)-1)) * ((o)->nx + 2) + (i))

int orc_scalar_surface(void* ctx, const double* phisurfall) {
  orc* o = (orc*)ctx;
  if (!o || (!phisurfall && o->nbody > 0)) return PIB_ERR_ARG;
  if (!o->have_particles) return fail(o, PIB_ERR_STATE, "pib_set_particles first");
  free(o->phisurfall);
  free(o->flux_body);
  o->phisurfall = (double*)xcalloc((size_t)o->nbody, sizeof(double));
  o->flux_body = (double*)xcalloc((size_t)o->nbody, sizeof(double));
  for (int m = 0; m < o->nbody; ++m) o->phisurfall[m] = phisurfall[m];
  o->nsurf = o->nbody;
  return PIB_OK;
}

int orc_scalar_flux(void* ctx, double* flux_body, double* out4) {
  orc* o = (orc*)ctx;
  if (!o) return PIB_ERR_ARG;
  if (!o->phir) return fail(o, PIB_ERR_STATE, "no scalar step has run");
  if (flux_body)
    for (int m = 0; m < o->nsurf; ++m) flux_body[m] = o->flux_body[m];
  if (out4) {
    out4[0] = o->flux_global;
    out4[1] = o->phi_fluid_mean;
    out4[2] = o->phi_solid_mean;
    out4[3] = 0.0;
  }
  return PIB_OK;
}

/* the physical arrays of the last scalstep for the tests: which = 0 phi,
 * 1 nl, 2 onl, 3 diffn, 4..6 grad phi, 7 fr; (nx+2, my, mz), plane 0 first */
int orc_scalar_debug(void* ctx, int which, double* out) {
  orc* o = (orc*)ctx;
  if (!o || !out || which < 0 || which > 7) return PIB_ERR_ARG;
  if (!o->phir) return fail(o, PIB_ERR_STATE, "no scalar step has run");
  const size_t np = (size_t)(o->nx + 2) * o->my * o->mz;
  const double* src[8] = {o->phir, o->nlphir, o->onlphir, o->dphi, o->gphi, o->gphi + np, o->gphi + 2 * np, o->frphi};
  memcpy(out, src[which], np * sizeof(double));
  return PIB_OK;
}

/* scalar plane i (1..mx) of a spectral (nx+1, my2, mz) array to physical
 * plane i of an IS0 array */
static void phi_c2r(orc* o, const cpx* f, int i, double* dst, planebuf* w) {
  for (int k = 1; k <= o->mz; ++k)
    for (int j = 1; j <= o->my2; ++j) w->ca[(size_t)(k - 1) * o->my2 + (j - 1)] = f[IS(o, i, j, k)];
  orc_ff2cr(o->my, o->mz, (const double*)w->ca, w->ra);
  for (int k = 1; k <= o->mz; ++k)
    for (int j = 1; j <= o->my; ++j) dst[IS0(o, i, j, k)] = w->ra[(size_t)(k - 1) * o->my + (j - 1)];
}
/* ff2rc of the (my, mz) plane in w->ra into w->ca */
static void phi_r2c(orc* o, planebuf* w) { orc_ff2rc(o->my, o->mz, w->ra, (double*)w->ca); }

/* form_nlphi, nl_allphi.F90:59-430, with compute_phim_heat_ratio(.TRUE.) */
static void form_nlphi(orc* o) {
  int nx = o->nx, my = o->my, mz = o->mz, my2 = o->my2, mx = o->mx, mxf = o->mxf;
  const int center = o->cfg.discr_scheme == PIB_SCHEME_CENTER;
  const double h = o->heat_ratio, dx = o->dx;
  /* phi, the y/z products and their transforms, planes 1..mx (:97-160) */
#pragma omp parallel
  {
    planebuf w = pb_alloc(o);
#pragma omp for schedule(static)
    for (int i = 1; i <= mx; ++i) {
      phi_c2r(o, o->phif, i, o->phir, &w);
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my; ++j) {
          double ph = o->phir[IS0(o, i, j, k)];
          o->fdd[IS0(o, i, j, k)] = center ? o->ubcp[IR(o, i, j, k, 1)] * ph : 0.0;
          w.ra[(size_t)(k - 1) * my + (j - 1)] = o->ubcp[IR(o, i, j, k, 2)] * ph;
          w.rb[(size_t)(k - 1) * my + (j - 1)] = o->ubcp[IR(o, i, j, k, 3)] * ph;
        }
      orc_ff2rc(my, mz, w.ra, (double*)w.ca);
      orc_ff2rc(my, mz, w.rb, (double*)w.cb);
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j) {
          size_t q = (size_t)(k - 1) * my2 + (j - 1);
          o->nlphif[IS(o, i, j, k)] = cadd(cik(o->wy[j - 1], w.ca[q]), cik(o->wz[k - 1], w.cb[q]));
        }
    }
    pb_free(&w);
  }
  /* compute_phim_heat_ratio (:431-487): only plane mx sets heat_ratio_new */
  {
    double phim = 0.0, um = 0.0;
    int cnt = 0;
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j)
        if (o->fluid[IFL(o, mx, j, k)]) {
          phim = phim + o->ubcp[IR(o, mx, j, k, 1)] * o->phir[IS0(o, mx, j, k)];
          um = um + o->ubcp[IR(o, mx, j, k, 1)];
          cnt = cnt + 1;
        }
    um = um / (double)cnt;
    phim = phim / (um * (double)cnt);
    o->heat_ratio_new = 1.0 / phim;
  }
  if (center) {
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j) o->fdd[IS0(o, 0, j, k)] = o->fdd[IS0(o, nx, j, k)] * h;
  } else {
    /* MUSCL slopes and the Rusanov flux at the faces i+1/2, i = 0..nx (:185-344) */
    double *uplus = (double*)xcalloc((size_t)nx + 2, sizeof(double)), *uminus = (double*)xcalloc((size_t)nx + 2, sizeof(double));
    double *Up = (double*)xcalloc((size_t)nx + 2, sizeof(double)), *Um = (double*)xcalloc((size_t)nx + 2, sizeof(double));
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j) {
#define PH(i) o->phir[IS0(o, (i), j, k)]
#define U1(i) o->ubcp[IR(o, (i), j, k, 1)]
        double sphi0 = (PH(1) - PH(nx - 1) * h) / (2.0 * dx);
        Um[0] = PH(nx) * h + sphi0 * (dx / 2.0);
        double sphim = (PH(2) / h - PH(nx)) / (2.0 * dx);
        Up[nx] = PH(1) / h - sphim * (dx / 2.0);
        double su0 = (U1(1) - U1(nx - 1)) / (2.0 * dx);
        uminus[0] = U1(nx) + su0 * (dx / 2.0);
        double sum = (U1(2) - U1(nx)) / (2.0 * dx);
        uplus[nx] = U1(1) - sum * (dx / 2.0);
        for (int i = 1; i <= nx; ++i) {
          int im1 = i - 1, ip1 = i + 1;
          double pfm = 1.0, pfp = 1.0;
          if (im1 < 1) {
            im1 = mxf + im1 - 1;
            pfm = h;
          }
          if (ip1 > mxf) {
            ip1 = ip1 - (mxf - 1);
            pfp = 1.0 / h;
          }
          double sphi = (pfp * PH(ip1) - pfm * PH(im1)) / (2.0 * dx);
          double su = (U1(ip1) - U1(im1)) / (2.0 * dx);
          uplus[i - 1] = U1(i) - su * (dx / 2.0);
          uminus[i] = U1(i) + su * (dx / 2.0);
          Up[i - 1] = PH(i) - sphi * (dx / 2.0);
          Um[i] = PH(i) + sphi * (dx / 2.0);
        }
        for (int i = 0; i <= nx; ++i) {
          double ms = fmax(fabs(uminus[i]), fabs(uplus[i]));
          double f = 0.5 * (uplus[i] * Up[i]) + 0.5 * (uminus[i] * Um[i]);
          o->fdd[IS0(o, i, j, k)] = f - 0.5 * ms * (Up[i] - Um[i]);
        }
#undef PH
#undef U1
      }
    free(uplus); free(uminus); free(Up); free(Um);
  }
  /* x-derivative of the flux, transformed into nlphif (:347-375) */
#pragma omp parallel
  {
    planebuf w = pb_alloc(o);
#pragma omp for schedule(static)
    for (int i = 1; i <= nx; ++i) {
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my; ++j) {
          double d = center ? (o->fdd[IS0(o, i + 1, j, k)] - o->fdd[IS0(o, i - 1, j, k)]) / (2.0 * dx)
                            : (o->fdd[IS0(o, i, j, k)] - o->fdd[IS0(o, i - 1, j, k)]) / dx;
          w.ra[(size_t)(k - 1) * my + (j - 1)] = d;
        }
      phi_r2c(o, &w);
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j)
          o->nlphif[IS(o, i, j, k)] = cadd(o->nlphif[IS(o, i, j, k)], w.ca[(size_t)(k - 1) * my2 + (j - 1)]);
    }
    pb_free(&w);
  }
  for (int k = 1; k <= mz; ++k)
    for (int j = 1; j <= my2; ++j) o->nlphif[IS(o, mx, j, k)] = cdivr(o->nlphif[IS(o, 1, j, k)], h);
  /* the sign of nl, the phase means, phi of plane mx (:380-419) */
  double pf = 0.0, ps = 0.0;
  for (int i = 1; i <= nx; ++i)
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my; ++j) {
        if (j <= my2) o->nlphif[IS(o, i, j, k)] = cscal(-1.0, o->nlphif[IS(o, i, j, k)]);
        if (o->fluid[IFL(o, i, j, k)]) pf = pf + o->phir[IS0(o, i, j, k)];
        else ps = ps + o->phir[IS0(o, i, j, k)];
      }
  for (int k = 1; k <= mz; ++k) {
    for (int j = 1; j <= my2; ++j) o->nlphif[IS(o, mx, j, k)] = cscal(-1.0, o->nlphif[IS(o, mx, j, k)]);
    for (int j = 1; j <= my; ++j) o->phir[IS0(o, mx, j, k)] = o->phir[IS0(o, 1, j, k)] / h;
  }
  o->phi_fluid_mean = pf / (double)o->count_fluid;
  o->phi_solid_mean = o->count_solid > 0 ? ps / (double)o->count_solid : 0.0;
}

/* calc_diffn, bc_scalar.F90:935-1033: diffn and grad phi of planes 1..mx */
static void calc_diffn_phi(orc* o) {
  int nx = o->nx, my = o->my, mz = o->mz, my2 = o->my2, mx = o->mx;
  const double h = o->heat_ratio, dx = o->dx, g = o->sc_gamma;
#pragma omp parallel
  {
    planebuf w = pb_alloc(o);
    cpx* f4 = (cpx*)xcalloc((size_t)my2 * mz * 4, sizeof(cpx));
#pragma omp for schedule(static)
    for (int i = 1; i <= mx; ++i) {
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j) {
          const cpx* P = o->phif;
          cpx pm, pp, p0 = P[IS(o, i, j, k)];
          if (i == 1) {
            pm = cscal(h, P[IS(o, nx, j, k)]);
            pp = P[IS(o, 2, j, k)];
          } else if (i == nx) {
            pm = P[IS(o, i - 1, j, k)];
            pp = cdivr(P[IS(o, 1, j, k)], h);
          } else if (i == nx + 1) {
            pm = P[IS(o, i - 1, j, k)];
            pp = cdivr(P[IS(o, 2, j, k)], h);
          } else {
            pm = P[IS(o, i - 1, j, k)];
            pp = P[IS(o, i + 1, j, k)];
          }
          cpx f1 = cscal(1.0 / o->dx2, cadd(csub(pm, cscal(2.0, p0)), pp));
          f1 = csub(f1, cscal(o->w2[(size_t)(k - 1) * my2 + (j - 1)], p0));
          f1 = cscal(g, f1);
          size_t q = (size_t)(k - 1) * my2 + (j - 1);
          f4[q] = f1;
          f4[(size_t)my2 * mz + q] = cdivr(csub(pp, pm), 2.0 * dx);
          f4[(size_t)2 * my2 * mz + q] = cik(o->wy[j - 1], p0);
          f4[(size_t)3 * my2 * mz + q] = cik(o->wz[k - 1], p0);
        }
      double* dst[4] = {o->dphi, o->gphi, o->gphi + (size_t)(nx + 2) * my * mz, o->gphi + (size_t)2 * (nx + 2) * my * mz};
      for (int q = 0; q < 4; ++q) {
        orc_ff2cr(my, mz, (const double*)(f4 + (size_t)q * my2 * mz), w.ra);
        for (int k = 1; k <= mz; ++k)
          for (int j = 1; j <= my; ++j) dst[q][IS0(o, i, j, k)] = w.ra[(size_t)(k - 1) * my + (j - 1)];
      }
    }
    free(f4);
    pb_free(&w);
  }
}

/* interpolate_phidata, bc_scalar.F90:1036-1153 (order-2 stencil, serial):
 * x nodes below 1 wrap up a period with the factor heat_ratio (heat_ratio_old
 * for onl), above mxf down a period with 1/heat_ratio; ul by the
 * interpolator's (k, j, i) sum, the others by array_dot_product's (i, j, k) */
static void interp_phi(const orc* o, const int pc[3], const double pos[3], stencil* st, double* ul,
                       double* nll, double* onll, double* dfll, int flag) {
  st->ib = pc[0]; st->jb = pc[1]; st->kb = pc[2];
  make_weights(st, pos);
  double vs[8], ns[8], os[8], ds[8];
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        int ii = st->ib + i, jj = wrapy(o, st->jb + j), kk = wrapz(o, st->kb + k);
        double pf = 1.0, pfo = 1.0;
        if (ii < 1) {
          ii = o->mxf + ii - 1;
          pf = o->heat_ratio;
          pfo = o->heat_ratio_old;
        } else if (ii > o->mxf) {
          ii = ii - o->mxf + 1;
          pf = 1.0 / o->heat_ratio;
        }
        int q = (k * 2 + j) * 2 + i;
        vs[q] = o->phir[IS0(o, ii, jj, kk)] * pf;
        if (flag == 1) {
          ns[q] = o->nlphir[IS0(o, ii, jj, kk)] * pf;
          os[q] = o->onlphir[IS0(o, ii, jj, kk)] * pfo;
          ds[q] = o->dphi[IS0(o, ii, jj, kk)] * pf;
        }
      }
  *ul = interp_kji(st, vs);
  if (flag == 1) {
    *nll = dot_ijk(st, ns);
    *onll = dot_ijk(st, os);
    *dfll = dot_ijk(st, ds);
  }
}

/* calc_bndrev_data, bc_scalar.F90:385-556: the surface flux of body m and,
 * with DOBND, the surface forcing (its x nodes wrap into 1..mxf: a share on
 * plane mxf is not transformed) */
static void calc_bndrev_data(orc* o, int m, int rks, double da1) {
  const double drm = 1.0, rb = o->radbdy[m - 1], cf = o->s.cf, h = o->heat_ratio;
  const double c1 = o->coef[rks - 1][0], c2 = o->coef[rks - 1][1], c3 = o->coef[rks - 1][2], c4 = o->coef[rks - 1][3];
  const size_t gs = (size_t)(o->nx + 2) * o->my * o->mz;
  double fluxloc = 0.0;
  for (int l = 1; l <= o->nbnd; ++l) {
    double xl[3];
    int is[3];
    for (int n = 1; n <= 3; ++n) {
      xl[n - 1] = XC(o, m, n) + BND(o, n, l) * rb;
      is[n - 1] = cellof(xl[n - 1]);
    }
    double xlt = xl[0];
    int ct = is[0];
    if (!(1 <= ct && ct <= o->nx)) {
      if (is[0] < 1) { ct = o->mxf + is[0] - 1; xlt = o->mxf + xl[0] - 1; }
      if (is[0] >= o->mxf) { ct = is[0] - (o->mxf - 1); xlt = xl[0] - (o->mxf - 1); }
    }
    xl[0] = xlt;
    is[0] = ct;
    stencil st;
    double phil = 0.0, nlphil = 0.0, onlphil = 0.0, dfll = 0.0;
    interp_phi(o, is, xl, &st, &phil, NULL, NULL, NULL, 0);
    double tempor = 0.0;
    if (o->cfg.dobnd) {
      tempor = tempor + cf * (phil - o->phisurfall[m - 1]);
      tempor = tempor - c3 * nlphil - c4 * onlphil;
      tempor = tempor - (c1 + c2) * dfll;
      tempor = tempor * da1 * drm * o->dx;
    }
    double gsten[3][8];
    for (int k = 0; k < 2; ++k) {
      int kk = wrapz(o, st.kb + k);
      for (int j = 0; j < 2; ++j) {
        int jj = wrapy(o, st.jb + j);
        for (int i = 0; i < 2; ++i) {
          int ii = st.ib + i;
          double pf = 1.0;
          if (ii < 1) {
            ii = o->mxf + ii - 1;
            pf = h;
          }
          if (ii > o->mxf) {
            ii = ii - o->mxf + 1;
            pf = 1.0 / h;
          }
          int q = (k * 2 + j) * 2 + i;
          if (o->cfg.dobnd) o->frphi[IS0(o, ii, jj, kk)] = o->frphi[IS0(o, ii, jj, kk)] + (st.w[q] * tempor) / o->dx3;
          for (int n = 0; n < 3; ++n) gsten[n][q] = o->gphi[(size_t)n * gs + IS0(o, ii, jj, kk)] * pf;
        }
      }
    }
    double gb[3];
    for (int n = 0; n < 3; ++n) gb[n] = dot_ijk(&st, gsten[n]);
    fluxloc = fluxloc + (gb[0] * BND(o, 1, l) + gb[1] * BND(o, 2, l) + gb[2] * BND(o, 3, l));
  }
  o->flux_body[m - 1] = fluxloc;
}

/* calc_inner_rev_data, bc_scalar.F90:558-932 (serial): a reversal pair
 * whose inner point lies in cell nx is evaluated twice, the second time one
 * period down (cell 0); the stencil's x nodes stay in 0..nx+1 and the shares
 * on planes 0 and nx+1 are not transformed */
static void calc_inner_rev_data(orc* o, int m, int rks, double da2) {
  const double drm = 1.0, cf = o->s.cf;
  const double rb = o->radbdy[m - 1], ri = o->radibdy[m - 1], ro = o->radobdy[m - 1];
  const double c1 = o->coef[rks - 1][0], c2 = o->coef[rks - 1][1], c3 = o->coef[rks - 1][2], c4 = o->coef[rks - 1][3];
  for (int l = 1; l <= o->nbnd; ++l) {
    if (!o->if_rev[(size_t)(m - 1) * o->nbnd + (l - 1)]) continue;
    double xli[3];
    int iii[3];
    for (int n = 1; n <= 3; ++n) {
      xli[n - 1] = XC(o, m, n) + BND(o, n, l) * ri;
      iii[n - 1] = cellof(xli[n - 1]);
    }
    int ct = iii[0];
    double xlt = xli[0];
    if (!(1 <= ct && ct <= o->nx)) {
      if (iii[0] < 1) { ct = o->mxf + iii[0] - 1; xlt = o->mxf + xli[0] - 1; }
      if (iii[0] >= o->mxf) { ct = iii[0] - (o->mxf - 1); xlt = xli[0] - (o->mxf - 1); }
    }
    int nev = ct == o->nx ? 2 : 1;
    for (int ev = 1; ev <= nev; ++ev) {
      if (ev == 1) {
        iii[0] = ct;
        xli[0] = xlt;
      } else if (iii[0] >= o->mxf - 1) {
        iii[0] = iii[0] - (o->mxf - 1);
        xli[0] = xli[0] - (o->mxf - 1);
      }
      double xlo[3];
      int io[3];
      for (int n = 1; n <= 3; ++n) {
        xlo[n - 1] = xli[n - 1] + BND(o, n, l) * (ro - ri);
        io[n - 1] = cellof(xlo[n - 1]);
      }
      stencil st;
      double philo, phili, nlphil, onlphil, dfll;
      interp_phi(o, io, xlo, &st, &philo, NULL, NULL, NULL, 0);
      interp_phi(o, iii, xli, &st, &phili, &nlphil, &onlphil, &dfll, 1);
      philo = o->phisurfall[m - 1] * (ro - ri) - philo * (rb - ri);
      philo = philo / (ro - rb);
      double phil = cf * (-philo + phili);
      double tempor = 0.0;
      tempor = tempor + phil;
      tempor = tempor - c3 * nlphil - c4 * onlphil;
      tempor = tempor - (c2 + c1) * dfll;
      tempor = tempor * da2 * drm * o->dx;
      int count_fl = 0;
      double force_fl = 0.0;
      for (int k = 0; k < 2; ++k) {
        int kk = wrapz(o, st.kb + k);
        for (int j = 0; j < 2; ++j) {
          int jj = wrapy(o, st.jb + j);
          for (int i = 0; i < 2; ++i) {
            int ii = st.ib + i;
            if (ii < 0) ii = o->mxf + ii - 1;
            if (o->fluid[IFL(o, ii, jj, kk)]) {
              count_fl = count_fl + 1;
              force_fl = force_fl + st.w[(k * 2 + j) * 2 + i] * tempor / o->dx3;
            }
          }
        }
      }
      int count_so = 8 - count_fl;
      double force_dist = force_fl / (double)count_so;
      for (int k = 0; k < 2; ++k) {
        int kk = wrapz(o, st.kb + k);
        for (int j = 0; j < 2; ++j) {
          int jj = wrapy(o, st.jb + j);
          for (int i = 0; i < 2; ++i) {
            int ii = st.ib + i;
            if (ii < 0) ii = o->mxf + ii - 1;
            double add = st.w[(k * 2 + j) * 2 + i] * tempor / o->dx3;
            size_t q = IS0(o, ii, jj, kk);
            if (o->cfg.include_frmeanfluid) o->frphi[q] = o->frphi[q] + add;
            else if (!o->fluid[IFL(o, ii, jj, kk)]) o->frphi[q] = o->frphi[q] + add + force_dist;
          }
        }
      }
    }
  }
}

int orc_scalstep(void* ctx, int rks) {
  orc* o = (orc*)ctx;
  if (!o || rks < 1 || rks > o->itrmax) return PIB_ERR_ARG;
  if (!o->phif) return fail(o, PIB_ERR_STATE, "pib_scalar_set first");
  if (!o->have_particles) return fail(o, PIB_ERR_STATE, "pib_set_particles first");
  if (o->nbody > 0 && o->nsurf != o->nbody) return fail(o, PIB_ERR_STATE, "pib_scalar_surface first");
  if (o->nx < 3) return fail(o, PIB_ERR_ARG, "the scalar step needs nx >= 3");
  int nx = o->nx, my = o->my, mz = o->mz, my2 = o->my2, mx = o->mx;
  const size_t np = (size_t)(nx + 2) * my * mz;
  if (!o->phir) {
    o->phir = (double*)xcalloc(np, sizeof(double));
    o->nlphir = (double*)xcalloc(np, sizeof(double));
    o->onlphir = (double*)xcalloc(np, sizeof(double));
    o->dphi = (double*)xcalloc(np, sizeof(double));
    o->gphi = (double*)xcalloc(3 * np, sizeof(double));
    o->frphi = (double*)xcalloc(np, sizeof(double));
    o->fdd = (double*)xcalloc(np, sizeof(double));
  }
  /* nlphi_graduphi (nlphi.F90:73-132) */
  for (int k = 1; k <= mz; ++k)
    for (int j = 1; j <= my2; ++j)
      for (int i = 1; i <= nx + 1; ++i) {
        o->onlphif[IS(o, i, j, k)] = o->nlphif[IS(o, i, j, k)];
        o->nlphif[IS(o, i, j, k)] = C(0.0, 0.0);
      }
  form_nlphi(o);
#pragma omp parallel
  {
    planebuf w = pb_alloc(o);
#pragma omp for schedule(static)
    for (int i = 1; i <= mx; ++i) {
      phi_c2r(o, o->nlphif, i, o->nlphir, &w);
      phi_c2r(o, o->onlphif, i, o->onlphir, &w);
    }
    pb_free(&w);
  }
  /* bcsetscal (bc_scalar.F90:65-383) */
  memset(o->frphi, 0, np * sizeof(double));
  calc_diffn_phi(o);
  o->flux_global = 0.0;
  for (int m = 1; m <= o->nbody; ++m) {
    double rb = o->radbdy[m - 1], ri = o->radibdy[m - 1];
    double da1 = 4.0 * PI * (rb * o->dx) * (rb * o->dx) / (double)(float)o->nbnd;
    double da2 = 4.0 * PI * (ri * o->dx) * (ri * o->dx) / (double)(float)o->nrpr_active[m - 1];
    calc_bndrev_data(o, m, rks, da1);
    if (o->cfg.dorpr) calc_inner_rev_data(o, m, rks, da2);
    o->flux_global = o->flux_global + o->flux_body[m - 1] * da1;
  }
#pragma omp parallel
  {
    planebuf w = pb_alloc(o);
#pragma omp for schedule(static)
    for (int i = 1; i <= nx; ++i) {
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my; ++j) w.ra[(size_t)(k - 1) * my + (j - 1)] = o->frphi[IS0(o, i, j, k)];
      phi_r2c(o, &w);
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j) o->ffphi[IS(o, i, j, k)] = w.ca[(size_t)(k - 1) * my2 + (j - 1)];
    }
    pb_free(&w);
  }
  return orc_scalstep_modes(ctx, rks);
}

/* ======================================================================
 * Adjoint of the fixed-bed Euler step.  No counterpart in the reference,
 * which has no adjoint: this is the discrete transpose of the step restated
 * above (orc_compute_new_timestep, nonlinear_part, orc_velstep) for AB2/CN
 * with a fixed dt, set_mpg, the central x-flux, dorpr and a bed at rest.
 * The mask, the active reversal pairs, the stencil cells and nrpr_active
 * are piecewise constant in the centres and are held fixed.  The variables
 * are the re and im parts of the spectral entries of planes 1..nx (plane
 * nx+1 is the periodic copy), umean, xc and mpg.  Every operator of the step
 * is linear in the state except the products of form_nl, so the transpose
 * of each is formed from the same fields:
 *   c2r (row-column convention) x = Re sum c_ky Z e^{+i theta}, c_ky = 1 for
 *     ky = 0, my/2 and 2 otherwise -> Z_bar = c_ky * my*mz * ff2rc(x_bar)
 *   ff2rc  Y = (my*mz)^-1 sum x e^{-i theta} -> x_bar = ff2cr(Y_bar/c_ky)/(my*mz)
 *   (i k) z -> (-i k) z_bar;  the symmetric Toeplitz line matrices are their
 *   own transposes; the (0,0) pressure solve keeps rows 1..nx-1, re only.
 * Pinned by finite differences of the forward step and by the tangent of the
 * step in its state (tests/test_adjoint_oracle.py).
 * ====================================================================== */
typedef struct {
  cpx *au, *aunew, *ap, *apnew, *anl, *anlp, *aff; /* IU (planes 1..nx) / IP */
  double *aubcp, *anlbc, *aonlbc, *adiffn, *aom, *afr; /* IR (planes 1..nx) */
  double *apbcp, *appr;                                /* IPR, IPPR */
  double aumean[3], ampg[3];
  double* axc; /* (nbody, 3) */
  int nbody;
} orc_adj;

typedef struct {
  cpx *u, *p, *nl, *onl;
  pib_scalars s;
} orc_ck;

static void adj_free(orc* o) {
  orc_adj* A = (orc_adj*)o->adj;
  if (A) {
    free(A->au); free(A->aunew); free(A->ap); free(A->apnew); free(A->anl); free(A->anlp); free(A->aff);
    free(A->aubcp); free(A->anlbc); free(A->aonlbc); free(A->adiffn); free(A->aom); free(A->afr);
    free(A->apbcp); free(A->appr); free(A->axc);
    free(A);
    o->adj = NULL;
  }
  for (int q = 0; q < o->nck; ++q) {
    orc_ck* c = (orc_ck*)o->ck[q];
    free(c->u); free(c->p); free(c->nl); free(c->onl); free(c);
  }
  free(o->ck);
  o->ck = NULL;
  o->nck = o->ckcap = 0;
  for (int q = 0; q < o->nrec; ++q) free(o->rec[q]);
  free(o->rec);
  o->rec = NULL;
  o->nrec = o->reccap = 0;
}

/* After a whole step, ubcp, nlbc, onlbc, diffn and ppr still hold that step's
 * values (velstep writes only u and p), with its t, dt and cf in the
 * scalars: what the adjoint of the step reads besides the bed.  A record is
 * those five fields and the scalars. */
static size_t rec_doubles(const orc* o) {
  size_t nr = (size_t)(o->nx + 1) * o->my * o->mz * 3;
  return 4 * nr + (size_t)o->nx * o->my * o->mz * 3;
}
static void rec_copy(orc* o, double* b, int push) {
  size_t nr = (size_t)(o->nx + 1) * o->my * o->mz * 3, np = (size_t)o->nx * o->my * o->mz * 3;
  double* f[4] = {o->ubcp, o->nlbc, o->onlbc, o->diffn};
  for (int q = 0; q < 4; ++q) {
    if (push) memcpy(b + q * nr, f[q], nr * sizeof(double));
    else memcpy(f[q], b + q * nr, nr * sizeof(double));
  }
  if (push) memcpy(b + 4 * nr, o->ppr, np * sizeof(double));
  else memcpy(o->ppr, b + 4 * nr, np * sizeof(double));
  pib_scalars* sp = (pib_scalars*)(b + 4 * nr + np);
  if (push) *sp = o->s;
  else o->s = *sp;
}
int orc_adjoint_record(void* ctx, int nsteps) {
  orc* o = (orc*)ctx;
  if (!o || nsteps < 0) return PIB_ERR_ARG;
  if (!o->adj) return fail(o, PIB_ERR_STATE, "pib_adjoint_begin first");
  for (int n = 0; n < nsteps; ++n) {
    int st = orc_step(o, 1);
    if (st) return st;
    if (o->nrec == o->reccap) {
      o->reccap = o->reccap ? 2 * o->reccap : 16;
      o->rec = (void**)realloc(o->rec, sizeof(void*) * (size_t)o->reccap);
    }
    double* b = (double*)malloc(sizeof(double) * rec_doubles(o) + sizeof(pib_scalars));
    rec_copy(o, b, 1);
    o->rec[o->nrec++] = b;
  }
  return PIB_OK;
}

static int adj_field_info(orc* o, int field, void** ptr, size_t* ndoubles) {
  orc_adj* A = (orc_adj*)o->adj;
  size_t nc = (size_t)(o->nx + 1) * o->my2 * o->mz * 3;
  if (field != PIB_F_AU && field != PIB_F_AP && field != PIB_F_ANL) return -1;
  if (!A) return 2;
  if (field == PIB_F_AU) { *ptr = A->au; *ndoubles = 2 * nc; }
  else if (field == PIB_F_AP) { *ptr = A->ap; *ndoubles = 2 * (size_t)o->nx * o->my2 * o->mz; }
  else { *ptr = A->anl; *ndoubles = 2 * nc; }
  return 0;
}

/* plane nx+1 of an adjoint spectral field is not a variable: zero it after a
 * set, copy plane 1 into it before a get (as the forward fields hold it) */
static void adj_fix_copy_plane(orc* o, cpx* f, int copy) {
  for (int n = 1; n <= 3; ++n)
    for (int k = 1; k <= o->mz; ++k)
      for (int j = 1; j <= o->my2; ++j)
        f[IU(o, o->nx + 1, j, k, n)] = copy ? f[IU(o, 1, j, k, n)] : C(0.0, 0.0);
}

int orc_adjoint_begin(void* ctx) {
  orc* o = (orc*)ctx;
  if (!o) return PIB_ERR_ARG;
  if (!o->have_particles) return fail(o, PIB_ERR_STATE, "pib_set_particles first");
  orc_adj* A = (orc_adj*)o->adj;
  size_t nc = (size_t)(o->nx + 1) * o->my2 * o->mz * 3, np = (size_t)o->nx * o->my2 * o->mz;
  size_t nr = (size_t)(o->nx + 1) * o->my * o->mz * 3, npr = (size_t)o->nx * o->my * o->mz;
  if (A && A->nbody != o->nbody) {
    free(A->axc);
    A->axc = (double*)xcalloc((size_t)o->nbody * 3, sizeof(double));
    A->nbody = o->nbody;
  }
  if (!A) {
    A = (orc_adj*)calloc(1, sizeof(orc_adj));
    A->au = (cpx*)xcalloc(nc, sizeof(cpx)); A->aunew = (cpx*)xcalloc(nc, sizeof(cpx));
    A->anl = (cpx*)xcalloc(nc, sizeof(cpx)); A->anlp = (cpx*)xcalloc(nc, sizeof(cpx));
    A->aff = (cpx*)xcalloc(nc, sizeof(cpx));
    A->ap = (cpx*)xcalloc(np, sizeof(cpx)); A->apnew = (cpx*)xcalloc(np, sizeof(cpx));
    A->aubcp = (double*)xcalloc(nr, sizeof(double)); A->anlbc = (double*)xcalloc(nr, sizeof(double));
    A->aonlbc = (double*)xcalloc(nr, sizeof(double)); A->adiffn = (double*)xcalloc(nr, sizeof(double));
    A->aom = (double*)xcalloc(nr, sizeof(double)); A->afr = (double*)xcalloc(nr, sizeof(double));
    A->apbcp = (double*)xcalloc(npr, sizeof(double)); A->appr = (double*)xcalloc(3 * npr, sizeof(double));
    A->axc = (double*)xcalloc((size_t)o->nbody * 3, sizeof(double));
    A->nbody = o->nbody;
    o->adj = A;
  }
  for (int q = 0; q < o->nrec; ++q) free(o->rec[q]);
  o->nrec = 0;
  memset(A->au, 0, nc * sizeof(cpx));
  memset(A->anl, 0, nc * sizeof(cpx));
  memset(A->ap, 0, np * sizeof(cpx));
  memset(A->axc, 0, (size_t)o->nbody * 3 * sizeof(double));
  for (int n = 0; n < 3; ++n) A->aumean[n] = A->ampg[n] = 0.0;
  return PIB_OK;
}

int orc_adjoint_get(void* ctx, double* dxc, double* dmpg, double* dumean) {
  orc* o = (orc*)ctx;
  if (!o) return PIB_ERR_ARG;
  orc_adj* A = (orc_adj*)o->adj;
  if (!A) return fail(o, PIB_ERR_STATE, "pib_adjoint_begin first");
  if (dxc) memcpy(dxc, A->axc, (size_t)o->nbody * 3 * sizeof(double));
  for (int n = 0; n < 3; ++n) {
    if (dmpg) dmpg[n] = A->ampg[n];
    if (dumean) dumean[n] = A->aumean[n];
  }
  return PIB_OK;
}

int orc_adjoint_set_umean(void* ctx, const double* dumean) {
  orc* o = (orc*)ctx;
  if (!o || !dumean) return PIB_ERR_ARG;
  orc_adj* A = (orc_adj*)o->adj;
  if (!A) return fail(o, PIB_ERR_STATE, "pib_adjoint_begin first");
  for (int n = 0; n < 3; ++n) A->aumean[n] = dumean[n];
  return PIB_OK;
}

int orc_checkpoint_push(void* ctx) {
  orc* o = (orc*)ctx;
  if (!o) return PIB_ERR_ARG;
  size_t nc = (size_t)(o->nx + 1) * o->my2 * o->mz * 3, np = (size_t)o->nx * o->my2 * o->mz;
  if (o->nck == o->ckcap) {
    o->ckcap = o->ckcap ? 2 * o->ckcap : 16;
    o->ck = (void**)realloc(o->ck, sizeof(void*) * (size_t)o->ckcap);
  }
  orc_ck* c = (orc_ck*)calloc(1, sizeof(orc_ck));
  c->u = (cpx*)malloc(nc * sizeof(cpx)); memcpy(c->u, o->u, nc * sizeof(cpx));
  c->nl = (cpx*)malloc(nc * sizeof(cpx)); memcpy(c->nl, o->nl, nc * sizeof(cpx));
  c->onl = (cpx*)malloc(nc * sizeof(cpx)); memcpy(c->onl, o->onl, nc * sizeof(cpx));
  c->p = (cpx*)malloc(np * sizeof(cpx)); memcpy(c->p, o->p, np * sizeof(cpx));
  c->s = o->s;
  o->ck[o->nck++] = c;
  return PIB_OK;
}

int orc_checkpoint_pop(void* ctx) {
  orc* o = (orc*)ctx;
  if (!o) return PIB_ERR_ARG;
  if (o->nck == 0) return fail(o, PIB_ERR_STATE, "checkpoint stack is empty");
  size_t nc = (size_t)(o->nx + 1) * o->my2 * o->mz * 3, np = (size_t)o->nx * o->my2 * o->mz;
  orc_ck* c = (orc_ck*)o->ck[--o->nck];
  memcpy(o->u, c->u, nc * sizeof(cpx));
  memcpy(o->nl, c->nl, nc * sizeof(cpx));
  memcpy(o->onl, c->onl, nc * sizeof(cpx));
  memcpy(o->p, c->p, np * sizeof(cpx));
  o->s = c->s;
  free(c->u); free(c->p); free(c->nl); free(c->onl); free(c);
  return PIB_OK;
}

int orc_checkpoint_steps(void* ctx, int nsteps) {
  orc* o = (orc*)ctx;
  if (!o || nsteps < 0) return PIB_ERR_ARG;
  for (int n = 0; n < nsteps; ++n) {
    int st = orc_checkpoint_push(o);
    if (!st) st = orc_step(o, 1);
    if (st) return st;
  }
  return PIB_OK;
}

/* ---- transposed plane transforms ------------------------------------- */
static inline double ky_weight(const orc* o, int j) { return (j == 1 || j == o->my2) ? 1.0 : 2.0; }
/* out += (c2r)^T x for one real plane x(my, mz) */
static void adj_c2r_T(const orc* o, const double* x, cpx* tmp, cpx* out, int accumulate) {
  orc_ff2rc(o->my, o->mz, x, (double*)tmp);
  double s = (double)o->my * (double)o->mz;
  for (int k = 1; k <= o->mz; ++k)
    for (int j = 1; j <= o->my2; ++j) {
      size_t q = (size_t)(k - 1) * o->my2 + (j - 1);
      cpx v = cscal(s * ky_weight(o, j), tmp[q]);
      out[q] = accumulate ? cadd(out[q], v) : v;
    }
}
/* x = (ff2rc)^T Y for one spectral plane Y(my2, mz) */
static void adj_r2c_T(const orc* o, const cpx* y, cpx* tmp, double* x) {
  double s = (double)o->my * (double)o->mz;
  for (int k = 1; k <= o->mz; ++k)
    for (int j = 1; j <= o->my2; ++j) {
      size_t q = (size_t)(k - 1) * o->my2 + (j - 1);
      tmp[q] = cdivr(y[q], s * ky_weight(o, j));
    }
  orc_ff2cr(o->my, o->mz, (const double*)tmp, x);
}

/* ---- velstep, transposed per mode ------------------------------------ */
static inline int wrp(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
/* in place: (re, im) <- M^-1 (re, im), M the pressure/divcorr matrix of this
 * mode (poisson.F90:81-83); the (0,0) mode solves rows 1..nx-1 of re only */
static int adj_msolve(const orc* o, double w2, int zero_mode, double* re, double* im, double* w) {
  int nx = o->nx;
  double *a = w, *b = w + nx, *c = w + 2 * nx, *xr = w + 3 * nx, *xi = w + 4 * nx;
  for (int i = 0; i < nx; ++i) {
    a[i] = 1.0 / o->dx2 - 0.25 * w2;
    b[i] = -2.0 / o->dx2 - 0.5 * w2;
    c[i] = a[i];
  }
  int st;
  if (zero_mode) {
    st = orc_tridag(a, b, c, re, xr, nx - 1);
    xr[nx - 1] = 0.0;
    for (int i = 0; i < nx; ++i) xi[i] = 0.0;
  } else {
    st = sm_solve(a, b, c, re, im, xr, xi, nx);
  }
  memcpy(re, xr, sizeof(double) * nx);
  memcpy(im, xi, sizeof(double) * nx);
  return st;
}
static int adj_hsolve(const orc* o, double dt, double c2, double w2, double* re, double* im, double* w) {
  int nx = o->nx;
  double *a = w, *b = w + nx, *c = w + 2 * nx, *xr = w + 3 * nx, *xi = w + 4 * nx;
  for (int i = 0; i < nx; ++i) {
    a[i] = -dt * o->vis * c2 / o->dx2;
    b[i] = 1.0 + dt * o->vis * c2 * (2.0 / o->dx2 + w2);
    c[i] = a[i];
  }
  int st = sm_solve(a, b, c, re, im, xr, xi, nx);
  memcpy(re, xr, sizeof(double) * nx);
  memcpy(im, xi, sizeof(double) * nx);
  return st;
}

static int adj_velstep_mode(orc* o, orc_adj* A, int j, int k) {
  int nx = o->nx;
  double dx = o->dx, dx2 = o->dx2, dt = o->s.dt, vis = o->vis;
  double w2 = o->w2[(size_t)(k - 1) * o->my2 + (j - 1)], wy = o->wy[j - 1], wz = o->wz[k - 1];
  double c1 = o->coef[0][0], c2 = o->coef[0][1], c3 = o->coef[0][2], c4 = o->coef[0][3];
  double c12 = c1 + c2, alp = c12 * dt, cv = c2 * dt * vis;
  int zero = (j == 1 && k == 1);
  cpx* L = (cpx*)malloc(sizeof(cpx) * (size_t)nx * 8);
  double* R = (double*)malloc(sizeof(double) * (size_t)nx * 7);
  cpx *ub[3] = {L, L + nx, L + 2 * nx}, *pt = L + 3 * nx, *ph = L + 4 * nx, *s = L + 5 * nx, *q = L + 6 * nx;
  double *re = R, *im = R + nx, *w = R + 2 * nx;
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i < nx; ++i) ub[d][i] = A->au[IU(o, i + 1, j, k, d + 1)];
  /* divcorr^T: p = p~ + phi - cv (L(phi)/dx2 - w2 phi), u = u* - alp G(phi) */
  for (int i = 0; i < nx; ++i) {
    cpx pb = A->ap[IP(o, i + 1, j, k)];
    pt[i] = pb;
    cpx lap = cadd(csub(A->ap[IP(o, wrp(i - 1, nx) + 1, j, k)], cscal(2.0, pb)), A->ap[IP(o, wrp(i + 1, nx) + 1, j, k)]);
    ph[i] = csub(pb, cscal(cv, csub(cdivr(lap, dx2), cscal(w2, pb))));
  }
  for (int i = 0; i < nx; ++i) {
    int ip = wrp(i + 1, nx);
    cpx g = cdivr(csub(ub[0][i], ub[0][ip]), dx);
    g = cadd(g, cscal(0.5, cik(-wy, cadd(ub[1][i], ub[1][ip]))));
    g = cadd(g, cscal(0.5, cik(-wz, cadd(ub[2][i], ub[2][ip]))));
    ph[i] = csub(ph[i], cscal(alp, g));
  }
  for (int i = 0; i < nx; ++i) { re[i] = ph[i].re; im[i] = ph[i].im; }
  int st = adj_msolve(o, w2, zero, re, im, w);
  for (int i = 0; i < nx; ++i) s[i] = cdivr(C(re[i], im[i]), alp);
  /* u*_bar = u_bar + Dstag^T(s) */
  for (int i = 0; i < nx; ++i) {
    int im1 = wrp(i - 1, nx);
    ub[0][i] = cadd(ub[0][i], cdivr(csub(s[im1], s[i]), dx));
    ub[1][i] = cadd(ub[1][i], cscal(0.5, cik(-wy, cadd(s[im1], s[i]))));
    ub[2][i] = cadd(ub[2][i], cscal(0.5, cik(-wz, cadd(s[im1], s[i]))));
  }
  /* ustep^T */
  for (int d = 0; d < 3; ++d) {
    for (int i = 0; i < nx; ++i) { re[i] = ub[d][i].re; im[i] = ub[d][i].im; }
    int s2 = adj_hsolve(o, dt, c2, w2, re, im, w);
    if (s2) st = s2;
    for (int i = 0; i < nx; ++i) q[i] = C(re[i], im[i]);
    for (int i = 0; i < nx; ++i) {
      cpx lap = cadd(csub(q[wrp(i - 1, nx)], cscal(2.0, q[i])), q[wrp(i + 1, nx)]);
      cpx v = cadd(cscal(1.0 - dt * vis * c1 * w2, q[i]), cdivr(cscal(dt * vis * c1, lap), dx2));
      A->aunew[IU(o, i + 1, j, k, d + 1)] = v;
      size_t ix = IU(o, i + 1, j, k, d + 1);
      A->anl[ix] = cadd(A->anl[ix], cscal(dt * c3, q[i]));
      A->aff[ix] = cscal(dt, q[i]);
      A->anlp[ix] = cscal(dt * c4, q[i]);
    }
    /* p~_bar -= dt c12 G_d^T(q) */
    for (int i = 0; i < nx; ++i) {
      int ip = wrp(i + 1, nx);
      cpx g;
      if (d == 0) g = cdivr(csub(q[i], q[ip]), dx);
      else g = cscal(0.5, cik(d == 1 ? -wy : -wz, cadd(q[i], q[ip])));
      pt[i] = csub(pt[i], cscal(dt * c12, g));
    }
  }
  /* pressure^T: p~ = M^-1 [c3 Dstag(nl) + Dstag(ff) + c4 Dstag(onl)]/c12 */
  for (int i = 0; i < nx; ++i) { re[i] = pt[i].re; im[i] = pt[i].im; }
  int s3 = adj_msolve(o, w2, zero, re, im, w);
  if (s3) st = s3;
  for (int i = 0; i < nx; ++i) s[i] = cdivr(C(re[i], im[i]), c12);
  for (int i = 0; i < nx; ++i) {
    int im1 = wrp(i - 1, nx);
    cpx g[3];
    g[0] = cdivr(csub(s[im1], s[i]), dx);
    g[1] = cscal(0.5, cik(-wy, cadd(s[im1], s[i])));
    g[2] = cscal(0.5, cik(-wz, cadd(s[im1], s[i])));
    for (int d = 0; d < 3; ++d) {
      size_t ix = IU(o, i + 1, j, k, d + 1);
      A->anl[ix] = cadd(A->anl[ix], cscal(c3, g[d]));
      A->aff[ix] = cadd(A->aff[ix], g[d]);
      A->anlp[ix] = cadd(A->anlp[ix], cscal(c4, g[d]));
    }
  }
  free(L);
  free(R);
  return st;
}

/* ---- IBM pieces, transposed ------------------------------------------ */
/* gradient of the trilinear interpolant sum w f with respect to the position */
static void interp_grad(const stencil* st, const double pos[3], const double v[8], double g[3]) {
  double z[3] = {pos[0] - st->ib, pos[1] - st->jb, pos[2] - st->kb};
  double xv[2] = {1.0 - z[0], z[0]}, yv[2] = {1.0 - z[1], z[1]}, zv[2] = {1.0 - z[2], z[2]};
  g[0] = g[1] = g[2] = 0.0;
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        double f = v[(k * 2 + j) * 2 + i];
        g[0] += (i ? 1.0 : -1.0) * yv[j] * zv[k] * f;
        g[1] += xv[i] * (j ? 1.0 : -1.0) * zv[k] * f;
        g[2] += xv[i] * yv[j] * (k ? 1.0 : -1.0) * f;
      }
}
static void sten_at(const orc* o, const double pos[3], stencil* st, int idx[8]) {
  st->ib = cellof(pos[0]);
  st->jb = cellof(pos[1]);
  st->kb = cellof(pos[2]);
  make_weights(st, pos);
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i)
        idx[(k * 2 + j) * 2 + i] =
            (int)((((size_t)(wrapz(o, st->kb + k) - 1)) * o->my + (wrapy(o, st->jb + j) - 1)));
}
/* node (i,j,k) of plane ii in an IR-layout (nx+1 planes) or IPR (nx planes) field */
#define NODE_R(o, f, ld, q, ii, n) (f)[(((size_t)((n)-1) * (o)->mz * (o)->my) + (size_t)(q)) * (ld) + ((ii)-1)]
/* out = sum_q w_q f(node q), grad: d/dpos; and the scatter f(node q) += w_q s */
static double sten_get(const orc* o, const stencil* st, const int idx[8], const double* f, int ld, int n,
                       const double pos[3], double grad[3], double gscale) {
  double v[8], acc = 0.0;
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        int q = (k * 2 + j) * 2 + i;
        v[q] = NODE_R(o, f, ld, idx[q], wrapx(o, st->ib + i), n);
        acc += st->w[q] * v[q];
      }
  if (grad) {
    double g[3];
    interp_grad(st, pos, v, g);
    for (int d = 0; d < 3; ++d) grad[d] += gscale * g[d];
  }
  return acc;
}
static void sten_put(const orc* o, const stencil* st, const int idx[8], double* f, int ld, int n, double s) {
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        int q = (k * 2 + j) * 2 + i;
        NODE_R(o, f, ld, idx[q], wrapx(o, st->ib + i), n) += st->w[q] * s;
      }
}

/* calc_inner_reversal_forcing, transposed (one body) */
static void adj_forcing_body(orc* o, orc_adj* A, int m, const double sbar[3]) {
  pib_scalars* s = &o->s;
  int ld = o->nx + 1, ldp = o->nx;
  double rb = o->radbdy[m - 1], ri = o->radibdy[m - 1], ro = o->radobdy[m - 1];
  double c1 = o->coef[0][0], c2 = o->coef[0][1], c3 = o->coef[0][2], c4 = o->coef[0][3], c12 = c1 + c2;
  double rdx = ri * o->dx;
  double K = (4.0 * PI * (rdx * rdx) / (double)o->nrpr_active[m - 1]) * o->cfg.dr * o->dx;
  double kap = (rb - ri) / (ro - rb), rn = o->cfg.revernorm == 1 ? 1.0 : 0.0;
  double* gx = &A->axc[m - 1];
  for (int l = 1; l <= o->nbnd; ++l) {
    if (!o->if_rev[(size_t)(m - 1) * o->nbnd + (l - 1)]) continue;
    double xli[3], xlo[3], rad = 0.0;
    for (int n = 1; n <= 3; ++n) {
      xli[n - 1] = XC(o, m, n) + BND(o, n, l) * ri;
      xlo[n - 1] = XC(o, m, n) + BND(o, n, l) * ro;
      rad += (BND(o, n, l) * ri) * (BND(o, n, l) * ri);
    }
    rad = sqrt(rad);
    double nv[3];
    for (int n = 0; n < 3; ++n) nv[n] = (BND(o, n + 1, l) * ri) / rad;
    double xpi[3] = {xli[0] - 0.5, xli[1], xli[2]};
    stencil si, so, sp;
    int ii[8], io[8], ip[8];
    sten_at(o, xli, &si, ii);
    sten_at(o, xlo, &so, io);
    sten_at(o, xpi, &sp, ip);
    /* forward ft of this point (the same terms as the forward step) */
    double ulo[3], ft[3];
    for (int n = 0; n < 3; ++n) ulo[n] = sten_get(o, &so, io, o->ubcp, ld, n + 1, xlo, NULL, 0);
    double un = nv[0] * ulo[0] + nv[1] * ulo[1] + nv[2] * ulo[2];
    for (int d = 0; d < 3; ++d) {
      double unorm = nv[d] * un, utang = ulo[d] - unorm;
      double ucar = -kap * utang - rn * kap * unorm; /* bed at rest: linvel = 0 */
      double uli = sten_get(o, &si, ii, o->ubcp, ld, d + 1, xli, NULL, 0);
      double nll = sten_get(o, &si, ii, o->nlbc, ld, d + 1, xli, NULL, 0);
      double onll = sten_get(o, &si, ii, o->onlbc, ld, d + 1, xli, NULL, 0);
      double dfll = sten_get(o, &si, ii, o->diffn, ld, d + 1, xli, NULL, 0);
      double ppll = sten_get(o, &sp, ip, o->ppr, ldp, d + 1, xpi, NULL, 0);
      ft[d] = (s->cf * (uli - ucar) - c3 * nll - c4 * onll + c12 * (ppll - dfll)) * K;
    }
    /* spreading, transposed */
    int solid[8], count_so = 0;
    double wfl = 0.0;
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
          int q = (k * 2 + j) * 2 + i;
          int fl = o->fluid[IFL(o, wrapx(o, si.ib + i), wrapy(o, si.jb + j), wrapz(o, si.kb + k))];
          solid[q] = !fl;
          if (fl) wfl += si.w[q] / o->dx3;
          else count_so++;
        }
    double ftb[3], gq[8], Ad[3];
    for (int d = 0; d < 3; ++d) {
      double ssum = 0.0, direct = 0.0, wall = 0.0;
      for (int q = 0; q < 8; ++q) {
        int i = q & 1, j = (q >> 1) & 1, k = q >> 2;
        double fb = NODE_R(o, A->afr, ld, (((size_t)(wrapz(o, si.kb + k) - 1)) * o->my + (wrapy(o, si.jb + j) - 1)),
                           wrapx(o, si.ib + i), d + 1);
        if (solid[q]) { ssum += fb; direct += si.w[q] * fb / o->dx3; }
        wall += si.w[q] / o->dx3;
      }
      Ad[d] = count_so > 0 ? ssum / (double)count_so : 0.0;
      ftb[d] = direct + (count_so > 0 ? wfl * Ad[d] : 0.0) + sbar[d] * wall;
    }
    for (int q = 0; q < 8; ++q) {
      int i = q & 1, j = (q >> 1) & 1, k = q >> 2;
      gq[q] = 0.0;
      for (int d = 0; d < 3; ++d) {
        double fb = NODE_R(o, A->afr, ld, (((size_t)(wrapz(o, si.kb + k) - 1)) * o->my + (wrapy(o, si.jb + j) - 1)),
                           wrapx(o, si.ib + i), d + 1);
        gq[q] += ft[d] / o->dx3 * ((solid[q] ? fb : Ad[d]) + sbar[d]);
      }
    }
    double gpos[3];
    interp_grad(&si, xli, gq, gpos);
    for (int d = 0; d < 3; ++d) gx[(size_t)d * o->nbody] += gpos[d];
    /* ft, transposed onto the interpolated fields */
    double h[3], Ph[3];
    for (int d = 0; d < 3; ++d) h[d] = K * ftb[d];
    double nh = nv[0] * h[0] + nv[1] * h[1] + nv[2] * h[2];
    for (int d = 0; d < 3; ++d) Ph[d] = (h[d] - nv[d] * nh) + rn * nv[d] * nh;
    double gi[3] = {0, 0, 0}, go[3] = {0, 0, 0}, gp[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
      sten_get(o, &si, ii, o->ubcp, ld, d + 1, xli, gi, s->cf * h[d]);
      sten_get(o, &si, ii, o->nlbc, ld, d + 1, xli, gi, -c3 * h[d]);
      sten_get(o, &si, ii, o->onlbc, ld, d + 1, xli, gi, -c4 * h[d]);
      sten_get(o, &si, ii, o->diffn, ld, d + 1, xli, gi, -c12 * h[d]);
      sten_get(o, &so, io, o->ubcp, ld, d + 1, xlo, go, s->cf * kap * Ph[d]);
      sten_get(o, &sp, ip, o->ppr, ldp, d + 1, xpi, gp, c12 * h[d]);
      sten_put(o, &si, ii, A->aubcp, ld, d + 1, s->cf * h[d]);
      sten_put(o, &si, ii, A->anlbc, ld, d + 1, -c3 * h[d]);
      sten_put(o, &si, ii, A->aonlbc, ld, d + 1, -c4 * h[d]);
      sten_put(o, &si, ii, A->adiffn, ld, d + 1, -c12 * h[d]);
      sten_put(o, &so, io, A->aubcp, ld, d + 1, s->cf * kap * Ph[d]);
      sten_put(o, &sp, ip, A->appr, ldp, d + 1, c12 * h[d]);
    }
    for (int d = 0; d < 3; ++d) gx[(size_t)d * o->nbody] += gi[d] + go[d] + gp[d];
  }
}

/* compute_mpg_at_n's force(m) = c12 (pres + visc) - c12 mpg pi vol/6, transposed */
static void adj_drag_body(orc* o, orc_adj* A, int m, const double* wf) {
  int ld = o->nx + 1, ldp = o->nx;
  double c12 = o->coef[0][0] + o->coef[0][1];
  double rb = o->radbdy[m - 1], rdx = rb * o->dx;
  double da1 = 4.0 * PI * (rdx * rdx) / (double)o->nbnd;
  double w[3] = {wf[m - 1], wf[o->nbody + m - 1], wf[2 * o->nbody + m - 1]};
  double vol = pow(2.0 * rb * o->dx, 3.0);
  for (int d = 0; d < 3; ++d) A->ampg[d] -= c12 * w[d] * PI * vol / 6.0;
  double* gx = &A->axc[m - 1];
  for (int l = 1; l <= o->nbnd; ++l) {
    double xl[3], xpb[3];
    for (int n = 1; n <= 3; ++n) xl[n - 1] = XC(o, m, n) + BND(o, n, l) * rb;
    xpb[0] = xl[0] - 0.5; xpb[1] = xl[1]; xpb[2] = xl[2];
    double cd[3] = {CD(o, 1, l), CD(o, 2, l), CD(o, 3, l)};
    stencil sp, sv;
    int ip[8], iv[8];
    sten_at(o, xpb, &sp, ip);
    sten_at(o, xl, &sv, iv);
    double plb = -c12 * da1 * (w[0] * cd[0] + w[1] * cd[1] + w[2] * cd[2]);
    double cw[3] = {cd[1] * w[2] - cd[2] * w[1], cd[2] * w[0] - cd[0] * w[2], cd[0] * w[1] - cd[1] * w[0]};
    double g[3] = {0, 0, 0};
    /* pbcp is (nx, my, mz): one component */
    sten_get(o, &sp, ip, o->pbcp, ldp, 1, xpb, g, plb);
    sten_put(o, &sp, ip, A->apbcp, ldp, 1, plb);
    for (int c = 0; c < 3; ++c) {
      double vb = c12 * o->vis * da1 * cw[c];
      sten_get(o, &sv, iv, o->omega_copy, ld, c + 1, xl, g, vb);
      sten_put(o, &sv, iv, A->aom, ld, c + 1, vb);
    }
    for (int d = 0; d < 3; ++d) gx[(size_t)d * o->nbody] += g[d];
  }
}

int orc_adjoint_step(void* ctx, const double* wforce) {
  orc* o = (orc*)ctx;
  if (!o) return PIB_ERR_ARG;
  orc_adj* A = (orc_adj*)o->adj;
  if (!A) return fail(o, PIB_ERR_STATE, "pib_adjoint_begin first");
  if (A->nbody != o->nbody) return fail(o, PIB_ERR_STATE, "particle set changed since pib_adjoint_begin");
  const pib_config* c = &o->cfg;
  if (c->rk3 || c->adaptive || !c->set_mpg || c->discr_scheme != PIB_SCHEME_CENTER || !c->dorpr || c->dobnd ||
      c->include_frmeanfluid || c->move_particles)
    return fail(o, PIB_ERR_UNSUPPORTED, "adjoint: rk3=0, adaptive=0, set_mpg, center, dorpr, fixed bed only");
  int nx = o->nx, my = o->my, mz = o->mz, my2 = o->my2;
  size_t nc = (size_t)(nx + 1) * my2 * mz * 3, np = (size_t)nx * my2 * mz;
  size_t nr = (size_t)(nx + 1) * my * mz * 3, npr = (size_t)nx * my * mz;
  int st;
  /* the step's forward fields: recorded ones (no force weight), or run */
  if (!wforce && o->nrec > 0) {
    double* b = (double*)o->rec[--o->nrec];
    rec_copy(o, b, 0);
    free(b);
  } else {
    if ((st = orc_compute_new_timestep(o, 1))) return st;
    if ((st = nonlinear_part(o, 1, 0))) return st;
  }
  double dt = o->s.dt;
  memset(A->aunew, 0, nc * sizeof(cpx));
  memset(A->apnew, 0, np * sizeof(cpx));
  memset(A->anlp, 0, nc * sizeof(cpx));
  memset(A->aff, 0, nc * sizeof(cpx));
  memset(A->aubcp, 0, nr * sizeof(double));
  memset(A->anlbc, 0, nr * sizeof(double));
  memset(A->aonlbc, 0, nr * sizeof(double));
  memset(A->adiffn, 0, nr * sizeof(double));
  memset(A->aom, 0, nr * sizeof(double));
  memset(A->afr, 0, nr * sizeof(double));
  memset(A->apbcp, 0, npr * sizeof(double));
  memset(A->appr, 0, 3 * npr * sizeof(double));
  /* velstep */
  st = 0;
#pragma omp parallel for schedule(static) collapse(2) reduction(min : st)
  for (int j = 1; j <= my2; ++j)
    for (int k = 1; k <= mz; ++k) {
      int s2 = adj_velstep_mode(o, A, j, k);
      if (s2 < st) st = s2;
    }
  if (st) return fail(o, st, "zero pivot in the transposed line solves");
  /* umean += (-mpg + frmean) dt */
  double frmb[3];
  for (int n = 0; n < 3; ++n) {
    frmb[n] = A->aumean[n] * dt;
    A->ampg[n] -= A->aumean[n] * dt;
  }
  /* ff = ff2rc(fr - frmean) */
  cpx* tc = (cpx*)xcalloc((size_t)my2 * mz, sizeof(cpx));
  cpx* tc2 = (cpx*)xcalloc((size_t)my2 * mz, sizeof(cpx));
  double* tr = (double*)xcalloc((size_t)my * mz, sizeof(double));
  for (int n = 1; n <= 3; ++n)
    for (int i = 1; i <= nx; ++i) {
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j) tc2[(size_t)(k - 1) * my2 + (j - 1)] = A->aff[IU(o, i, j, k, n)];
      adj_r2c_T(o, tc2, tc, tr);
      put_rplane(o, A->afr + (size_t)(n - 1) * (nx + 1) * my * mz, nx + 1, i, tr);
      for (size_t q = 0; q < (size_t)my * mz; ++q) frmb[n - 1] -= tr[q];
    }
  /* frmean = (sum of the weighted shares)/(nx my mz) */
  double sbar[3];
  for (int n = 0; n < 3; ++n) sbar[n] = frmb[n] / (double)((long long)nx * my * mz);
  for (int m = 1; m <= o->nbody; ++m) adj_forcing_body(o, A, m, sbar);
  if (wforce)
    for (int m = 1; m <= o->nbody; ++m) adj_drag_body(o, A, m, wforce);
  /* calc_pgrad: ppr = c2r(G' p) + mpg */
  {
    cpx* Z = (cpx*)xcalloc(3 * np, sizeof(cpx));
    for (int n = 1; n <= 3; ++n)
      for (int i = 1; i <= nx; ++i) {
        get_rplane(o, A->appr + (size_t)(n - 1) * nx * my * mz, nx, i, tr);
        for (size_t q = 0; q < (size_t)my * mz; ++q) A->ampg[n - 1] += tr[q];
        adj_c2r_T(o, tr, tc, Z + ((size_t)(n - 1) * nx + (i - 1)) * my2 * mz, 0);
      }
#define ZZ(n, i, j, k) Z[(((size_t)((n)-1) * nx + ((i)-1)) * mz + ((k)-1)) * my2 + ((j)-1)]
    for (int i = 1; i <= nx; ++i) {
      int im = wrp(i - 2, nx) + 1, ip = wrp(i, nx) + 1;
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j) {
          cpx v = cdivr(csub(ZZ(1, im, j, k), ZZ(1, ip, j, k)), 2.0 * o->dx);
          v = cadd(v, cik(-o->wy[j - 1], ZZ(2, i, j, k)));
          v = cadd(v, cik(-o->wz[k - 1], ZZ(3, i, j, k)));
          A->apnew[IP(o, i, j, k)] = cadd(A->apnew[IP(o, i, j, k)], v);
        }
    }
    /* compute_omega */
    for (int n = 1; n <= 3; ++n)
      for (int i = 1; i <= nx; ++i) {
        get_rplane(o, A->aom + (size_t)(n - 1) * (nx + 1) * my * mz, nx + 1, i, tr);
        adj_c2r_T(o, tr, tc, Z + ((size_t)(n - 1) * nx + (i - 1)) * my2 * mz, 0);
      }
    for (int i = 1; i <= nx; ++i)
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j) {
          double wy = o->wy[j - 1], wz = o->wz[k - 1];
          size_t i1 = IU(o, i, j, k, 1), i2 = IU(o, i, j, k, 2), i3 = IU(o, i, j, k, 3);
          A->aunew[i1] = cadd(A->aunew[i1], cadd(cik(wy, ZZ(3, i, j, k)), cik(-wz, ZZ(2, i, j, k))));
          A->aunew[i3] = cadd(A->aunew[i3], cik(-wy, ZZ(1, i, j, k)));
          A->aunew[i2] = cadd(A->aunew[i2], cik(wz, ZZ(1, i, j, k)));
        }
    for (int i = 1; i <= nx; ++i) {
      int im = wrp(i - 2, nx) + 1, ip = wrp(i, nx) + 1;
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my; ++j) {
          A->aubcp[IR(o, i, j, k, 2)] += (A->aom[IR(o, im, j, k, 3)] - A->aom[IR(o, ip, j, k, 3)]) / (2.0 * o->dx);
          A->aubcp[IR(o, i, j, k, 3)] -= (A->aom[IR(o, im, j, k, 2)] - A->aom[IR(o, ip, j, k, 2)]) / (2.0 * o->dx);
        }
    }
    /* calc_visc */
    for (int n = 1; n <= 3; ++n)
      for (int i = 1; i <= nx; ++i) {
        get_rplane(o, A->adiffn + (size_t)(n - 1) * (nx + 1) * my * mz, nx + 1, i, tr);
        adj_c2r_T(o, tr, tc, Z + ((size_t)(n - 1) * nx + (i - 1)) * my2 * mz, 0);
      }
    for (int n = 1; n <= 3; ++n)
      for (int i = 1; i <= nx; ++i) {
        int im = wrp(i - 2, nx) + 1, ip = wrp(i, nx) + 1;
        for (int k = 1; k <= mz; ++k)
          for (int j = 1; j <= my2; ++j) {
            cpx lap = cadd(csub(ZZ(n, im, j, k), cscal(2.0, ZZ(n, i, j, k))), ZZ(n, ip, j, k));
            cpx v = cscal(o->vis, csub(cdivr(lap, o->dx2), cscal(o->w2[(size_t)(k - 1) * my2 + (j - 1)], ZZ(n, i, j, k))));
            size_t ix = IU(o, i, j, k, n);
            A->aunew[ix] = cadd(A->aunew[ix], v);
          }
      }
#undef ZZ
    free(Z);
  }
  /* nlbc = c2r(nl), onlbc = c2r(onl) */
  for (int n = 1; n <= 3; ++n)
    for (int i = 1; i <= nx; ++i) {
      get_rplane(o, A->anlbc + (size_t)(n - 1) * (nx + 1) * my * mz, nx + 1, i, tr);
      adj_c2r_T(o, tr, tc2, tc, 0);
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j) {
          size_t ix = IU(o, i, j, k, n);
          A->anl[ix] = cadd(A->anl[ix], tc[(size_t)(k - 1) * my2 + (j - 1)]);
        }
      get_rplane(o, A->aonlbc + (size_t)(n - 1) * (nx + 1) * my * mz, nx + 1, i, tr);
      adj_c2r_T(o, tr, tc2, tc, 0);
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j) {
          size_t ix = IU(o, i, j, k, n);
          A->anlp[ix] = cadd(A->anlp[ix], tc[(size_t)(k - 1) * my2 + (j - 1)]);
        }
    }
  /* form_nl (central) and the sign of nonlinear: nl = -[(i wy) F(u_c v) + (i wz) F(u_c w) + F(dFx_c)] */
  {
    double* Q = (double*)xcalloc(8 * npr, sizeof(double)); /* Q12 Q13 Q22 Q23 Q33 E1 E2 E3 */
#define QQ(t, i, j, k) Q[(((size_t)(t) * nx + ((i)-1)) * mz + ((k)-1)) * my + ((j)-1)]
    for (int i = 1; i <= nx; ++i) {
      for (int t = 0; t < 8; ++t) {
        for (int k = 1; k <= mz; ++k)
          for (int j = 1; j <= my2; ++j) {
            double wy = o->wy[j - 1], wz = o->wz[k - 1];
            cpx g1 = cscal(-1.0, A->anl[IU(o, i, j, k, 1)]), g2 = cscal(-1.0, A->anl[IU(o, i, j, k, 2)]),
                g3 = cscal(-1.0, A->anl[IU(o, i, j, k, 3)]);
            cpx v;
            switch (t) {
              case 0: v = cik(-wy, g1); break;
              case 1: v = cik(-wz, g1); break;
              case 2: v = cik(-wy, g2); break;
              case 3: v = cadd(cik(-wz, g2), cik(-wy, g3)); break;
              case 4: v = cik(-wz, g3); break;
              case 5: v = g1; break;
              case 6: v = g2; break;
              default: v = g3;
            }
            tc2[(size_t)(k - 1) * my2 + (j - 1)] = v;
          }
        adj_r2c_T(o, tc2, tc, &QQ(t, i, 1, 1));
      }
    }
    for (int i = 1; i <= nx; ++i) {
      int im = wrp(i - 2, nx) + 1, ip = wrp(i, nx) + 1;
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my; ++j) {
          double u1 = o->ubcp[IR(o, i, j, k, 1)], u2 = o->ubcp[IR(o, i, j, k, 2)], u3 = o->ubcp[IR(o, i, j, k, 3)];
          double F1 = (QQ(5, im, j, k) - QQ(5, ip, j, k)) / (2.0 * o->dx);
          double F2 = (QQ(6, im, j, k) - QQ(6, ip, j, k)) / (2.0 * o->dx);
          double F3 = (QQ(7, im, j, k) - QQ(7, ip, j, k)) / (2.0 * o->dx);
          double q12 = QQ(0, i, j, k), q13 = QQ(1, i, j, k), q22 = QQ(2, i, j, k), q23 = QQ(3, i, j, k),
                 q33 = QQ(4, i, j, k);
          A->aubcp[IR(o, i, j, k, 1)] += q12 * u2 + q13 * u3 + 2.0 * F1 * u1 + F2 * u2 + F3 * u3;
          A->aubcp[IR(o, i, j, k, 2)] += q12 * u1 + 2.0 * q22 * u2 + q23 * u3 + F2 * u1;
          A->aubcp[IR(o, i, j, k, 3)] += q13 * u1 + q23 * u2 + 2.0 * q33 * u3 + F3 * u1;
        }
    }
#undef QQ
    free(Q);
  }
  /* ubcp = c2r(u) + umean, pbcp = c2r(p) */
  for (int n = 1; n <= 3; ++n)
    for (int i = 1; i <= nx; ++i) {
      get_rplane(o, A->aubcp + (size_t)(n - 1) * (nx + 1) * my * mz, nx + 1, i, tr);
      for (size_t q = 0; q < (size_t)my * mz; ++q) A->aumean[n - 1] += tr[q];
      adj_c2r_T(o, tr, tc2, tc, 0);
      for (int k = 1; k <= mz; ++k)
        for (int j = 1; j <= my2; ++j) {
          size_t ix = IU(o, i, j, k, n);
          A->aunew[ix] = cadd(A->aunew[ix], tc[(size_t)(k - 1) * my2 + (j - 1)]);
        }
    }
  for (int i = 1; i <= nx; ++i) {
    get_rplane(o, A->apbcp, nx, i, tr);
    adj_c2r_T(o, tr, tc2, tc, 0);
    for (int k = 1; k <= mz; ++k)
      for (int j = 1; j <= my2; ++j) {
        size_t ix = IP(o, i, j, k);
        A->apnew[ix] = cadd(A->apnew[ix], tc[(size_t)(k - 1) * my2 + (j - 1)]);
      }
  }
  free(tc); free(tc2); free(tr);
  /* roll: the adjoint of this step's inputs */
  cpx* t;
  t = A->au; A->au = A->aunew; A->aunew = t;
  t = A->ap; A->ap = A->apnew; A->apnew = t;
  t = A->anl; A->anl = A->anlp; A->anlp = t;
  return PIB_OK;
}
