// pib_dem.hip -- soft-sphere DEM sub-cycle of PUReIBM-PS2D-FD1D on MI355X.
//
// des_time_march(.FALSE.) (collision_mod.F90:272-297) advances the particles
// by FACTOR sub-steps of DTSOLID; each sub-step computes the contact forces of
// CALC_FORCE_DES (soft_sphere_coll.F90:52-551) on the neighbour sets of the
// coarse-cell search (GRID_BASED_NEIGHBOR_SEARCH, collision_mod.F90:807-1062)
// and then moves every particle (CFNEWVALUES :1335-1462).
//
// Execution: one workgroup of 512 threads runs a whole chunk of sub-steps
// with the particle positions, velocities and radii resident in LDS (56 bytes
// a particle, up to 2048 particles) beside 16-bit candidate lists in the room
// left, and each particle's contact-free acceleration in registers; two
// barriers a sub-step.  Per-particle work is a few dozen dependent FP64
// operations, so a single CU with no round trip to the host between sub-steps
// beats any multi-workgroup split that would need a grid-wide barrier per
// sub-step.  Global memory is touched only for contacts (collision mass,
// hydrodynamic terms) and at the start and end of a launch.
//
// Exactness.  The reference sums FC(p) in a fixed order: first -FN(l,p) for
// the contacting l < p in increasing l (the outer particle loop of
// CALC_FORCE_DES), then +FN(p,i) for i > p in the order of p's neighbour list,
// which is the order of p's 27-cell scan (cells k, j, i ascending with the
// periodic wrap of :873-949, particles of a cell ascending).  Here each thread
// finds its contacts from a candidate list (all pairs within r_a+r_b+skin,
// rebuilt when a particle has moved skin/2), recovers the scan position of a
// contact from the two cells, and sums in that order; the pair force is
// evaluated from the lower particle's side with the reference's association,
// so both members of a pair get bit-identical FN.  FP contraction is off in
// this file: every operation rounds as written.
//
// The reference fixes MEW = 0 (collision_mod.F90:388): the tangential force is
// a signed zero (soft_sphere_coll.F90:862-866) and the contact histories never
// reach FC; OMEGA_NEW stays zero (CFNEWVALUES :1386).  Neither history is kept.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "pureibm_dem.h"

#pragma clang fp contract(off)

namespace {

constexpr int DEM_NT = 512;      // threads of the one workgroup (256 VGPRs a lane)
constexpr int DEM_NMAX = 2048;   // particles held in LDS
constexpr int DEM_PPT = DEM_NMAX / DEM_NT;  // particles a thread
constexpr int DEM_VCAP = 32;     // candidate partners per particle, at most
constexpr int DEM_LDS = 159 * 1024;  // LDS a workgroup may take on gfx950 (160 KB), less a margin
constexpr int DEM_HCAP = 16;     // contacts with higher-numbered partners per particle
constexpr int DEM_CHUNK = 4096;  // sub-steps per launch
constexpr double SMALL_NUMBER = (double)1.0E-35f;  // constants.F90:92
constexpr double PI = 3.14159265358979323846264338327950288;  // 4*atan(1)

enum { ERR_MOTION = 1, ERR_VCAP = 2, ERR_HCAP = 4 };

struct DemArgs {
  int n, nsub, only_dem, rebuild;
  int imax1, jmax1, kmax1, cap;  // cap: candidate slots a particle (LDS lists)
  double dxc, dyc, dzc;   // coarse cell sizes
  double lx, ly, lz;      // XE(IMAX1)-XE(1) etc., the shift of the cell scan
  double bx, by, bz;      // XLENGTH etc. (EX2-WX1), the shift of CALC_FORCE_DES and CFNEWVALUES
  double four_rmax;       // 4*MAX_RADIUS
  double kn, alog, den;   // KN, |ln e|, sqrt(pi^2 + ln^2 e)
  double dts;             // DTSOLID
  double skin, half_skin2;
  double fa[3];           // frame_accln
  double dx, dy, dz, foffset;
  const double* rad;      // DES_RADIUS
  const double* cmass;    // pi*(2r)^3*rhos/6, the collision mass of CALC_FORCE_DES
  const double* pmass;    // PMASS
  const double* t1;       // (3,n) RHOF*(PRES+VISC)/PMASS
  const double* t2;       // (3,n) mpg*PVOL/PMASS
  const double* a0;       // (3,n) acceleration without contact: ((+0 + t1) - frame_accln) - t2
  double* pos;            // (3,n) DES_POS_NEW
  double* vel;            // (3,n) DES_VEL_NEW
  double* disp;           // (3,n) displacement since the last candidate rebuild
  unsigned short* vl;     // (cap, n) candidate partners, ascending, kept between launches
  unsigned short* vcnt;   // n
  double* cforce;         // (3,n) FC of the last sub-step
  double* xc_out;         // (n,3) IBMUPDATE XC
  int* status;            // [0] error bits, [1] contacts, [2] rebuilds
  double* diag;           // [0] overlap_max (%)
};

__device__ inline int cell_axis(double x, double d, int imax1) {
  double t = x / d;
  int c = (t < 1.0e9 && t > -1.0e9) ? (int)t : 0;  // INT truncates toward zero
  c = c + 2;
  return c > imax1 ? imax1 : c;
}

// find_cell_index (collision_mod.F90:743-764), packed 10 bits an axis
__device__ inline int cell_of(const DemArgs& A, double x, double y, double z) {
  int i = cell_axis(x, A.dxc, A.imax1), j = cell_axis(y, A.dyc, A.jmax1), k = cell_axis(z, A.dzc, A.kmax1);
  return (i & 1023) | ((j & 1023) << 10) | ((k & 1023) << 20);
}

// periodic wrap of a scanned cell index (collision_mod.F90:873-949)
__device__ inline int wrap_cell(int ii, int imax1, double& f) {
  f = 0.0;
  if (ii > imax1) { f = 1.0; return 2; }
  if (ii < 2) { f = -1.0; return imax1; }
  return ii;
}

struct Lds {
  double *px, *py, *pz, *vx, *vy, *vz, *rad;
  unsigned short *cl, *cc;  // candidate lists (cap, n) and counts
  unsigned short* perm;     // particles by descending candidate count: lane j of a sub-step
                            // handles perm[j], so the lanes of a wave loop alike
};

// Order the particles by descending candidate count (counting sort; the
// order within a count does not matter, since no particle's arithmetic
// depends on the lane that runs it).
__device__ void make_perm(const DemArgs& A, const Lds& S, int* hist) {
  const int n = A.n, tid = threadIdx.x, nb = A.cap + 1;
  for (int i = tid; i < nb; i += DEM_NT) hist[i] = 0;
  __syncthreads();
  for (int p = tid; p < n; p += DEM_NT) atomicAdd(&hist[A.cap - S.cc[p]], 1);
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int i = 0; i < nb; ++i) {
      const int c = hist[i];
      hist[i] = acc;
      acc += c;
    }
  }
  __syncthreads();
  for (int p = tid; p < n; p += DEM_NT) {
    const int pos = atomicAdd(&hist[A.cap - S.cc[p]], 1);
    S.perm[pos] = (unsigned short)p;
  }
  __syncthreads();
}

// the particle of slot k of this lane, or n when there is none
__device__ inline int owned(const Lds& S, int n, int k) {
  const int j = threadIdx.x + k * DEM_NT;
  return j < n ? S.perm[j] : n;
}

// Position of b in a's 27-cell scan (0..26) with the first image that passes
// the neighbour test 1.2*(r_a+r_b) - |d| > SMALL_NUMBER (:955-1003), or -1
// when b is not a neighbour of a.
__device__ int scan_pos(const DemArgs& A, const Lds& S, int a, int b) {
  const int ca = cell_of(A, S.px[a], S.py[a], S.pz[a]), cb = cell_of(A, S.px[b], S.py[b], S.pz[b]);
  const int ai = ca & 1023, aj = (ca >> 10) & 1023, ak = ca >> 20;
  const int bi = cb & 1023, bj = (cb >> 10) & 1023, bk = cb >> 20;
  int o = 0;
  for (int kk = ak - 1; kk <= ak + 1; ++kk) {
    double zf;
    const int K = wrap_cell(kk, A.kmax1, zf);
    for (int jj = aj - 1; jj <= aj + 1; ++jj) {
      double yf;
      const int J = wrap_cell(jj, A.jmax1, yf);
      for (int ii = ai - 1; ii <= ai + 1; ++ii, ++o) {
        double xf;
        const int I = wrap_cell(ii, A.imax1, xf);
        if (I != bi || J != bj || K != bk) continue;
        double rlm = S.rad[a] + S.rad[b];
        rlm = 1.2 * rlm;
        const double d0 = (S.px[b] + xf * A.lx) - S.px[a];
        const double d1 = (S.py[b] + yf * A.ly) - S.py[a];
        const double d2 = (S.pz[b] + zf * A.lz) - S.pz[a];
        const double dm = sqrt(((0.0 + d0 * d0) + d1 * d1) + d2 * d2);
        if (rlm - dm > SMALL_NUMBER) return o;
      }
    }
  }
  return -1;
}

// Overlap test and normal force FN(a,b) on a, a < b (soft_sphere_coll.F90:
// 309-426 with CFRELVEL :728-772); false when the pair is not in contact.
__device__ inline bool pair_force(const DemArgs& A, const Lds& S, int a, int b, double& f0, double& f1, double& f2,
                                  double& ovpct) {
  const double pax = S.px[a], pay = S.py[a], paz = S.pz[a];
  double pbx = S.px[b], pby = S.py[b], pbz = S.pz[b];
  if (fabs(pax - pbx) > A.four_rmax) pbx = pbx > pax ? pbx - A.bx : pbx + A.bx;
  if (fabs(pay - pby) > A.four_rmax) pby = pby > pay ? pby - A.by : pby + A.by;
  if (fabs(paz - pbz) > A.four_rmax) pbz = pbz > paz ? pbz - A.bz : pbz + A.bz;
  const double rlm = S.rad[a] + S.rad[b];
  const double d0 = pbx - pax, d1 = pby - pay, d2 = pbz - paz;
  const double dsq = ((0.0 + d0 * d0) + d1 * d1) + d2 * d2;
  // |d| > r_a + r_b for certain: no contact, and no sqrt needed to say so
  if (dsq > rlm * rlm * (1.0 + 1.0e-9)) return false;
  const double dm = sqrt(dsq);
  if (!(rlm - dm > SMALL_NUMBER)) return false;
  ovpct = ((rlm - dm) / rlm) * 100.0;
  const double n0 = d0 / dm, n1 = d1 / dm, n2 = d2 / dm;
  const double vr0 = S.vx[a] - S.vx[b], vr1 = S.vy[a] - S.vy[b], vr2 = S.vz[a] - S.vz[b];
  const double vrn = ((0.0 + vr0 * n0) + vr1 * n1) + vr2 * n2;
  const double ov = rlm - dm;
  const double ma = A.cmass[a], mb = A.cmass[b];
  const double meff = (ma * mb) / (ma + mb);
  double eta = 2.0 * sqrt(A.kn * meff) * A.alog;
  eta = eta / A.den;
  f0 = -(A.kn * ov * n0) + -(eta * vrn * n0);
  f1 = -(A.kn * ov * n1) + -(eta * vrn * n1);
  f2 = -(A.kn * ov * n2) + -(eta * vrn * n2);
  return true;
}

// Candidate partners: every q within r_p + r_q + skin of p over the nearest
// periodic image (box lengths hold 8 radii, so that image is unique).
__device__ int build_candidates(const DemArgs& A, const Lds& S, int p) {
  const int n = A.n;
  int cnt = 0;
  const double hx = 0.5 * A.bx, hy = 0.5 * A.by, hz = 0.5 * A.bz;
  for (int q = 0; q < n; ++q) {
    if (q == p) continue;
    double d0 = S.px[q] - S.px[p], d1 = S.py[q] - S.py[p], d2 = S.pz[q] - S.pz[p];
    d0 = d0 > hx ? d0 - A.bx : (d0 < -hx ? d0 + A.bx : d0);
    d1 = d1 > hy ? d1 - A.by : (d1 < -hy ? d1 + A.by : d1);
    d2 = d2 > hz ? d2 - A.bz : (d2 < -hz ? d2 + A.bz : d2);
    const double r = S.rad[p] + S.rad[q] + A.skin;
    if (d0 * d0 + d1 * d1 + d2 * d2 < r * r) {
      if (cnt < A.cap) S.cl[cnt * n + p] = (unsigned short)q;
      ++cnt;
    }
  }
  S.cc[p] = (unsigned short)(cnt < A.cap ? cnt : A.cap);
  return cnt > A.cap ? ERR_VCAP : 0;
}

// Insert key into an ascending register array padded with INT_MAX.
__device__ inline void insert_key(int (&hk)[DEM_HCAP], int key) {
#pragma unroll
  for (int j = 0; j < DEM_HCAP; ++j) {
    const int a = hk[j];
    hk[j] = a < key ? a : key;
    key = a < key ? key : a;
  }
}

// FC(p) summed in the reference's order; counts the contacts p heads.
// Returns whether any contact force reached FC(p).
__device__ bool contact_sum(const DemArgs& A, const Lds& S, int p, double& f0, double& f1, double& f2, double& ovmax,
                            int& ncont, int& err) {
  const int n = A.n;
  const int cnt = S.cc[p];
  int hk[DEM_HCAP];
#pragma unroll
  for (int j = 0; j < DEM_HCAP; ++j) hk[j] = INT_MAX;
  int nh = 0;
  bool touched = false;
  f0 = 0.0; f1 = 0.0; f2 = 0.0;
  for (int t = 0; t < cnt; ++t) {
    const int q = S.cl[t * n + p];
    const int a = q < p ? q : p, b = q < p ? p : q;
    double g0, g1, g2, ov;
    if (!pair_force(A, S, a, b, g0, g1, g2, ov)) continue;
    const int o = scan_pos(A, S, a, b);
    if (o < 0) continue;
    touched = true;
    if (q < p) {  // FC(I) = FC(I) - FN(L) during particle q's turn (CFFCTOW :873-904)
      f0 = f0 - g0; f1 = f1 - g1; f2 = f2 - g2;
    } else {
      if (nh < DEM_HCAP) insert_key(hk, (o << 6) | t);
      else err |= ERR_HCAP;
      ++nh;
      ++ncont;
      ovmax = ov > ovmax ? ov : ovmax;
    }
  }
  // pop the keys in ascending order from the front of the register array
  while (hk[0] != INT_MAX) {
    const int q = S.cl[(hk[0] & 63) * n + p];
#pragma unroll
    for (int j = 0; j + 1 < DEM_HCAP; ++j) hk[j] = hk[j + 1];
    hk[DEM_HCAP - 1] = INT_MAX;
    double g0, g1, g2, ov;
    pair_force(A, S, p, q, g0, g1, g2, ov);
    f0 = f0 + g0; f1 = f1 + g1; f2 = f2 + g2;
  }
  return touched;
}

__global__ __launch_bounds__(DEM_NT) void k_dem_march(DemArgs A) {
  extern __shared__ double lds_d[];
  const int n = A.n, tid = threadIdx.x;
  Lds S;
  S.px = lds_d; S.py = S.px + n; S.pz = S.py + n;
  S.vx = S.pz + n; S.vy = S.vx + n; S.vz = S.vy + n;
  S.rad = S.vz + n;
  unsigned long long* s_ovmax = (unsigned long long*)(S.rad + n);
  S.cc = (unsigned short*)(s_ovmax + 2);
  S.cl = S.cc + n;
  S.perm = S.cl + (size_t)A.cap * n;
  __shared__ int s_hist[DEM_VCAP + 2];
  int* s_red = (int*)(s_ovmax + 1);  // [0] contacts, [1] rebuilds
  if (tid == 0) { *s_ovmax = 0ull; s_red[0] = 0; s_red[1] = 0; }
  for (int p = tid; p < n; p += DEM_NT) {
    S.px[p] = A.pos[p]; S.py[p] = A.pos[n + p]; S.pz[p] = A.pos[2 * n + p];
    S.vx[p] = A.vel[p]; S.vy[p] = A.vel[n + p]; S.vz[p] = A.vel[2 * n + p];
    S.rad[p] = A.rad[p];
    if (!A.rebuild) {
      S.cc[p] = A.vcnt[p];
      for (int t = 0; t < A.cap; ++t) S.cl[t * n + p] = A.vl[(size_t)t * n + p];
    }
  }
  __syncthreads();
  int err = 0, nreb = 0;
  if (A.rebuild) {
#pragma unroll 1
    for (int p = tid; p < n; p += DEM_NT) err |= build_candidates(A, S, p);
    ++nreb;
    __syncthreads();
  }
  make_perm(A, S, s_hist);
  double dsp[DEM_PPT][3], a0r[DEM_PPT][3];
#pragma unroll
  for (int k = 0; k < DEM_PPT; ++k) {
    const int p = owned(S, n, k);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      dsp[k][c] = p < n && !A.rebuild ? A.disp[c * n + p] : 0.0;
      a0r[k][c] = p < n ? A.a0[c * n + p] : 0.0;
    }
  }
  double ovmax = 0.0;
  int ncont = 0;
  for (int s = 0; s < A.nsub; ++s) {
    double f[DEM_PPT][3];
#pragma unroll
    for (int k = 0; k < DEM_PPT; ++k) f[k][0] = f[k][1] = f[k][2] = 0.0;
    ovmax = 0.0;
    ncont = 0;
    int touched = 0;  // bit k: a contact force reached particle k of this thread
#pragma unroll 1
    for (int k = 0; k < DEM_PPT; ++k) {
      const int p = owned(S, n, k);
      if (p >= n) break;
      double g0, g1, g2;
      if (contact_sum(A, S, p, g0, g1, g2, ovmax, ncont, err)) touched |= 1 << k;
#pragma unroll
      for (int kk = 0; kk < DEM_PPT; ++kk)
        if (kk == k) { f[kk][0] = g0; f[kk][1] = g1; f[kk][2] = g2; }
    }
    __syncthreads();
    int need = 0;
    const bool last = s + 1 == A.nsub;
#pragma unroll 1
    for (int k = 0; k < DEM_PPT; ++k) {
      const int p = owned(S, n, k);
      if (p >= n) break;
      double g[3], ds[3], a0[3];
#pragma unroll
      for (int kk = 0; kk < DEM_PPT; ++kk)
        if (kk == k)
          for (int c = 0; c < 3; ++c) { g[c] = f[kk][c]; ds[c] = dsp[kk][c]; a0[c] = a0r[kk][c]; }
      // CFNEWVALUES (collision_mod.F90:1361-1459)
      const bool hit = (touched >> k) & 1;
      const double pm = hit ? A.pmass[p] : 1.0;
      double* sp[3] = {S.px, S.py, S.pz};
      double* sv[3] = {S.vx, S.vy, S.vz};
      const double box[3] = {A.bx, A.by, A.bz};
      double dd[3], pn[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        // without a contact FC is +0, +0/PMASS is +0 and the sum is a0
        double acc = a0[c];
        if (hit) {
          acc = g[c] / pm;
          if (!A.only_dem) acc = acc + A.t1[c * n + p] - A.fa[c] - A.t2[c * n + p];
        }
        const double vn = sv[c][p] + acc * A.dts;
        const double po = sp[c][p];
        pn[c] = po + vn * A.dts;
        dd[c] = pn[c] - po;
        sv[c][p] = vn;
        if (last) A.cforce[c * n + p] = g[c];
      }
      const double dsq = ((0.0 + dd[0] * dd[0]) + dd[1] * dd[1]) + dd[2] * dd[2];
      const double rp = S.rad[p];
      if (dsq >= rp * rp * (1.0 - 1.0e-9) && sqrt(dsq) >= rp) err |= ERR_MOTION;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (pn[c] > box[c]) pn[c] = pn[c] - box[c];
        else if (pn[c] < 0.0) pn[c] = pn[c] + box[c];
        sp[c][p] = pn[c];
        ds[c] = ds[c] + dd[c];
      }
      need |= ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2] > A.half_skin2;
#pragma unroll
      for (int kk = 0; kk < DEM_PPT; ++kk)
        if (kk == k)
          for (int c = 0; c < 3; ++c) dsp[kk][c] = ds[c];
    }
    if (__syncthreads_or(need)) {
#pragma unroll 1
      for (int p = tid; p < n; p += DEM_NT) err |= build_candidates(A, S, p);
      ++nreb;
      __syncthreads();
      make_perm(A, S, s_hist);
#pragma unroll
      for (int k = 0; k < DEM_PPT; ++k) {
        const int p = owned(S, n, k);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          dsp[k][c] = 0.0;
          a0r[k][c] = p < n ? A.a0[c * n + p] : 0.0;
        }
      }
    }
  }
  // state back to HBM, IBMUPDATE (collision_mod.F90:1464-1474)
#pragma unroll
  for (int k = 0; k < DEM_PPT; ++k) {
    const int p = owned(S, n, k);
    if (p >= n) continue;
    A.pos[p] = S.px[p]; A.pos[n + p] = S.py[p]; A.pos[2 * n + p] = S.pz[p];
    A.vel[p] = S.vx[p]; A.vel[n + p] = S.vy[p]; A.vel[2 * n + p] = S.vz[p];
#pragma unroll
    for (int c = 0; c < 3; ++c) A.disp[c * n + p] = dsp[k][c];
    A.vcnt[p] = S.cc[p];
    for (int t = 0; t < A.cap; ++t) A.vl[(size_t)t * n + p] = S.cl[t * n + p];
    A.xc_out[p] = S.px[p] / A.dx + 1.0 - A.foffset;
    A.xc_out[n + p] = S.py[p] / A.dy + 1.0;
    A.xc_out[2 * n + p] = S.pz[p] / A.dz + 1.0;
  }
  // ovmax >= 0: its bit pattern orders as an unsigned integer
  atomicMax(s_ovmax, (unsigned long long)__double_as_longlong(ovmax));
  atomicAdd(&s_red[0], ncont);
  const int any_err = __syncthreads_or(err);
  const int err_or = (__syncthreads_or(err & ERR_MOTION) ? ERR_MOTION : 0) |
                     (__syncthreads_or(err & ERR_VCAP) ? ERR_VCAP : 0) |
                     (__syncthreads_or(err & ERR_HCAP) ? ERR_HCAP : 0);
  if (tid == 0) {
    A.status[0] = any_err ? (A.status[0] | err_or) : A.status[0];
    A.status[1] = s_red[0];
    A.status[2] = A.status[2] + nreb;
    A.diag[0] = __longlong_as_double((long long)*s_ovmax);
  }
}

struct DemCtx {
  pib_dem_config cfg{};
  int n = 0;
  char err[256] = {0};
  hipStream_t stream = nullptr;
  double kn = 0, en = 0, alog = 0, den = 0, tcoll = 0, dtsolid_orig = 0, dtsolid = 0, s_time = 0;
  double max_radius = 0, lx = 0, ly = 0, lz = 0, dxc = 0, dyc = 0, dzc = 0, skin = 0;
  int cx = 0, cy = 0, cz = 0, factor = 0, substeps = 0, rebuilds = 0, ncontacts = 0;
  double overlap_max = 0;
  bool need_rebuild = true;
  std::vector<double> pmass_h, pvol_h;
  double *rad = nullptr, *cmass = nullptr, *pmass = nullptr, *t1 = nullptr, *t2 = nullptr;
  double *pos = nullptr, *vel = nullptr, *disp = nullptr, *cforce = nullptr, *xco = nullptr, *diag = nullptr;
  double* a0 = nullptr;
  unsigned short *vl = nullptr, *vcnt = nullptr;
  int* status = nullptr;
  int cap = 0;
};

struct DemFail {
  int code;
  char msg[200];
};

#define DEMCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) {                                                                    \
      DemFail f_{PIB_ERR_HIP, {0}};                                                            \
      snprintf(f_.msg, sizeof f_.msg, "%s: %s", #x, hipGetErrorString(e_));                    \
      throw f_;                                                                                \
    }                                                                                          \
  } while (0)

[[noreturn]] void dem_fail(int code, const char* msg) {
  DemFail f{code, {0}};
  snprintf(f.msg, sizeof f.msg, "%s", msg);
  throw f;
}

int nint_(double x) { return (int)(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5)); }

// candidate slots that fit in LDS beside 7 doubles, a count and a slot of
// the order a particle
int cand_cap(int n) {
  long long room = (long long)DEM_LDS - 7LL * 8 * n - 4LL * n - 64;
  long long c = room / (2LL * n);
  return (int)(c < DEM_VCAP ? c : DEM_VCAP);
}

size_t lds_bytes(int n, int cap) { return (size_t)n * 7 * sizeof(double) + 16 + 2 * (size_t)n * (2 + cap) + 64; }

template <typename T>
T* dalloc(size_t count) {
  void* p = nullptr;
  DEMCHK(hipMalloc(&p, (count > 0 ? count : 1) * sizeof(T)));
  DEMCHK(hipMemset(p, 0, (count > 0 ? count : 1) * sizeof(T)));
  return (T*)p;
}

void dem_free(DemCtx* D) {
  void* ps[] = {D->rad, D->cmass, D->pmass, D->t1, D->t2, D->pos, D->vel, D->disp, D->cforce, D->xco,
                D->diag, D->vl, D->vcnt, D->status, D->a0};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (D->stream) (void)hipStreamDestroy(D->stream);
}

int run(DemCtx* D, void (*fn)(DemCtx*, void*), void* arg) {
  try {
    fn(D, arg);
    return PIB_OK;
  } catch (const DemFail& f) {
    snprintf(D->err, sizeof D->err, "%s", f.msg);
    return f.code;
  }
}

struct CreateArgs {
  const double *xc, *radbdy, *velbdy;
};

void do_create(DemCtx* D, void* va) {
  const CreateArgs* a = (const CreateArgs*)va;
  const pib_dem_config& c = D->cfg;
  const int n = D->n;
  if (!(c.dx > 0 && c.dy > 0 && c.dz > 0)) dem_fail(PIB_ERR_ARG, "dx, dy, dz must be positive");
  if (!(D->en > 0.0 && D->en <= 1.0)) dem_fail(PIB_ERR_ARG, "coeff_rest must lie in (0, 1]");
  // initialize_coarse_grid (collision_mod.F90:346-492)
  double maxr = a->radbdy[0];
  for (int m = 1; m < n; ++m) maxr = a->radbdy[m] > maxr ? a->radbdy[m] : maxr;
  const double dxeff = 2.0 * maxr * 1.5 * c.dx;
  D->cy = std::max(nint_(c.ylength / dxeff), 2);
  D->cx = std::max(nint_(c.xlength / dxeff), 2);
  D->cz = std::max(nint_(c.zlength / dxeff), 2);
  if (D->cx > 1023 || D->cy > 1023 || D->cz > 1023) dem_fail(PIB_ERR_UNSUPPORTED, "more than 1022 coarse cells a direction");
  D->dyc = c.ylength / (double)(D->cy - 1);
  D->dxc = c.xlength / (double)(D->cx - 1);
  D->dzc = c.zlength / (double)(D->cz - 1);
  double xe = 0, yn = 0, zt = 0;
  for (int i = 2; i <= D->cx; ++i) xe = xe + D->dxc;
  for (int j = 2; j <= D->cy; ++j) yn = yn + D->dyc;
  for (int k = 2; k <= D->cz; ++k) zt = zt + D->dzc;
  D->lx = xe - 0.0; D->ly = yn - 0.0; D->lz = zt - 0.0;
  // DES_INIT_ARRAYS (:1203-1316) and CFASSIGN (:582-741)
  std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n, 0.0), rad(n), cm(n);
  D->pmass_h.resize(n);
  D->pvol_h.resize(n);
  D->max_radius = 0.0;
  for (int l = 0; l < n; ++l) {
    pos[l] = (a->xc[l] + c.foffset - 1.0) * c.dx;
    pos[n + l] = (a->xc[n + l] - 1.0) * c.dy;
    pos[2 * n + l] = (a->xc[2 * n + l] - 1.0) * c.dz;
    rad[l] = a->radbdy[l] * c.dx;
    if (!(rad[l] > 0)) dem_fail(PIB_ERR_ARG, "radbdy must be positive");
    if (a->velbdy)
      for (int d = 0; d < 3; ++d) vel[(size_t)d * n + l] = a->velbdy[(size_t)d * n + l];
    const double rad2 = rad[l] * rad[l];
    D->pvol_h[l] = (4.0 / 3.0) * PI * rad2 * rad[l];
    D->pmass_h[l] = D->pvol_h[l] * c.rhos;
    cm[l] = (PI * std::pow(2.0 * rad[l], 3.0) * c.rhos) / 6.0;
    if (rad[l] > D->max_radius) D->max_radius = rad[l];
  }
  const double box[3] = {c.xlength, c.ylength, c.zlength};
  const double cell[3] = {D->dxc, D->dyc, D->dzc};
  for (int d = 0; d < 3; ++d) {
    if (!(box[d] >= 8.0 * D->max_radius)) dem_fail(PIB_ERR_UNSUPPORTED, "box shorter than 8 particle radii");
    // the reference wraps a centre only once it moves (CFNEWVALUES :1431-1459)
    for (int l = 0; l < n; ++l)
      if (!(pos[(size_t)d * n + l] > -cell[d] && pos[(size_t)d * n + l] < box[d] + cell[d]))
        dem_fail(PIB_ERR_ARG, "particle centre more than a coarse cell outside the periodic box");
  }
  {
    const double mi = (PI * std::pow(c.dia_phys, 3.0) * c.rhos) / 6.0;
    const double mj = (PI * std::pow(c.dia_phys, 3.0) * c.rhos) / 6.0;
    const double meff = (mi * mj) / (mi + mj);
    double etan = 2.0 * std::sqrt(D->kn * meff) * std::fabs(std::log(D->en));
    etan = etan / std::sqrt(PI * PI + std::log(D->en) * std::log(D->en));
    const double q = etan / meff;
    D->tcoll = PI / std::sqrt(D->kn / meff - (q * q) / 4.0);
    D->dtsolid_orig = D->tcoll / 50.0;
    D->dtsolid = D->dtsolid_orig;
    if (!(D->dtsolid_orig > 0)) dem_fail(PIB_ERR_ARG, "DTSOLID is not positive (dia_phys, rhos, kn)");
  }
  D->alog = std::fabs(std::log(D->en));
  D->den = std::sqrt(PI * PI + std::log(D->en) * std::log(D->en));
  // candidate skin: 0.4 of the largest radius (the cutoff stays below half a box)
  D->skin = 0.4 * D->max_radius;
  DEMCHK(hipStreamCreateWithFlags(&D->stream, hipStreamNonBlocking));
  D->rad = dalloc<double>(n);
  D->cmass = dalloc<double>(n);
  D->pmass = dalloc<double>(n);
  D->t1 = dalloc<double>(3 * (size_t)n);
  D->t2 = dalloc<double>(3 * (size_t)n);
  D->pos = dalloc<double>(3 * (size_t)n);
  D->vel = dalloc<double>(3 * (size_t)n);
  D->disp = dalloc<double>(3 * (size_t)n);
  D->cforce = dalloc<double>(3 * (size_t)n);
  D->xco = dalloc<double>(3 * (size_t)n);
  D->diag = dalloc<double>(4);
  D->cap = cand_cap(n);
  if (D->cap < 8) dem_fail(PIB_ERR_UNSUPPORTED, "too many particles for the LDS candidate lists");
  D->vl = dalloc<unsigned short>((size_t)D->cap * n);
  D->vcnt = dalloc<unsigned short>(n);
  D->a0 = dalloc<double>(3 * (size_t)n);
  D->status = dalloc<int>(4);
  DEMCHK(hipMemcpy(D->rad, rad.data(), n * sizeof(double), hipMemcpyHostToDevice));
  DEMCHK(hipMemcpy(D->cmass, cm.data(), n * sizeof(double), hipMemcpyHostToDevice));
  DEMCHK(hipMemcpy(D->pmass, D->pmass_h.data(), n * sizeof(double), hipMemcpyHostToDevice));
  DEMCHK(hipMemcpy(D->pos, pos.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice));
  DEMCHK(hipMemcpy(D->vel, vel.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice));
  DEMCHK(hipFuncSetAttribute((const void*)k_dem_march, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)DEM_LDS));
  D->need_rebuild = true;
}

struct MarchArgs {
  double dt;
  const double *pres, *visc, *mpg, *fa;
  double *xc, *velbdy, *angv;
};

void do_march(DemCtx* D, void* va) {
  const MarchArgs* a = (const MarchArgs*)va;
  const int n = D->n;
  const pib_dem_config& c = D->cfg;
  if (!c.only_dem && (!a->pres || !a->visc || !a->mpg || !a->fa))
    dem_fail(PIB_ERR_ARG, "pres, visc, mpg and frame_accln are required unless only_dem");
  if (!(a->dt > 0)) dem_fail(PIB_ERR_ARG, "dt must be positive");
  const double fq = a->dt / D->dtsolid_orig;
  if (!(fq < 2.0e9)) dem_fail(PIB_ERR_ARG, "dt/DTSOLID too large");
  // des_time_march (:274-281)
  D->factor = (int)fq;
  D->dtsolid = D->dtsolid_orig;
  if (D->factor < 1) {
    D->factor = 1;
    D->dtsolid = a->dt;
  }
  DemArgs A{};
  A.n = n;
  A.only_dem = c.only_dem ? 1 : 0;
  A.imax1 = D->cx; A.jmax1 = D->cy; A.kmax1 = D->cz;
  A.dxc = D->dxc; A.dyc = D->dyc; A.dzc = D->dzc;
  A.lx = D->lx; A.ly = D->ly; A.lz = D->lz;
  A.bx = c.xlength; A.by = c.ylength; A.bz = c.zlength;
  A.four_rmax = 4.0 * D->max_radius;
  A.kn = D->kn; A.alog = D->alog; A.den = D->den;
  A.dts = D->dtsolid;
  A.skin = D->skin;
  A.half_skin2 = 0.99 * (0.5 * D->skin) * (0.5 * D->skin);
  A.dx = c.dx; A.dy = c.dy; A.dz = c.dz; A.foffset = c.foffset;
  A.rad = D->rad; A.cmass = D->cmass; A.pmass = D->pmass; A.t1 = D->t1; A.t2 = D->t2;
  A.pos = D->pos; A.vel = D->vel; A.disp = D->disp; A.vl = D->vl; A.vcnt = D->vcnt;
  A.cap = D->cap; A.a0 = D->a0;
  A.cforce = D->cforce; A.xc_out = D->xco; A.status = D->status; A.diag = D->diag;
  // CFNEWVALUES :1374, the terms that stay fixed over the sub-cycle, and
  // the whole acceleration of a particle without contact
  std::vector<double> t1(3 * (size_t)n, 0.0), t2(3 * (size_t)n, 0.0), a0(3 * (size_t)n, 0.0);
  if (!c.only_dem) {
    for (int d = 0; d < 3; ++d) {
      A.fa[d] = a->fa[d];
      for (int l = 0; l < n; ++l) {
        const size_t i = (size_t)d * n + l;
        t1[i] = c.rhof * (a->pres[i] + a->visc[i]) / D->pmass_h[l];
        t2[i] = a->mpg[d] * D->pvol_h[l] / D->pmass_h[l];
        a0[i] = 0.0 + t1[i] - a->fa[d] - t2[i];
      }
    }
    DEMCHK(hipMemcpyAsync(D->t1, t1.data(), t1.size() * sizeof(double), hipMemcpyHostToDevice, D->stream));
    DEMCHK(hipMemcpyAsync(D->t2, t2.data(), t2.size() * sizeof(double), hipMemcpyHostToDevice, D->stream));
  }
  DEMCHK(hipMemcpyAsync(D->a0, a0.data(), a0.size() * sizeof(double), hipMemcpyHostToDevice, D->stream));
  int zero[4] = {0, 0, 0, 0};
  DEMCHK(hipMemcpyAsync(D->status, zero, sizeof zero, hipMemcpyHostToDevice, D->stream));
  for (int done = 0; done < D->factor; done += DEM_CHUNK) {
    A.nsub = std::min(DEM_CHUNK, D->factor - done);
    A.rebuild = D->need_rebuild ? 1 : 0;
    D->need_rebuild = false;
    hipLaunchKernelGGL(k_dem_march, dim3(1), dim3(DEM_NT), lds_bytes(n, D->cap), D->stream, A);
    DEMCHK(hipGetLastError());
  }
  int st[4];
  double dg[4];
  DEMCHK(hipMemcpyAsync(st, D->status, sizeof st, hipMemcpyDeviceToHost, D->stream));
  DEMCHK(hipMemcpyAsync(dg, D->diag, sizeof dg, hipMemcpyDeviceToHost, D->stream));
  if (a->xc) DEMCHK(hipMemcpyAsync(a->xc, D->xco, 3 * n * sizeof(double), hipMemcpyDeviceToHost, D->stream));
  if (a->velbdy) DEMCHK(hipMemcpyAsync(a->velbdy, D->vel, 3 * n * sizeof(double), hipMemcpyDeviceToHost, D->stream));
  DEMCHK(hipStreamSynchronize(D->stream));
  if (a->angv) std::memset(a->angv, 0, 3 * n * sizeof(double));
  for (int s = 0; s < D->factor; ++s) D->s_time = D->s_time + D->dtsolid;
  D->substeps += D->factor;
  D->ncontacts = st[1];
  D->rebuilds += st[2];
  D->overlap_max = dg[0];
  if (st[0] & ERR_MOTION)
    dem_fail(PIB_ERR_MOTION, "MOVEMENT UNDESIRED: a particle moved by its radius in one sub-step");
  if (st[0] & ERR_VCAP) dem_fail(PIB_ERR_CAPACITY, "more candidate partners of one particle than the LDS lists hold");
  if (st[0] & ERR_HCAP) dem_fail(PIB_ERR_CAPACITY, "more than 16 contacts of one particle with higher-numbered ones");
}

}  // namespace

extern "C" {

int pib_dem_create(const pib_dem_config* cfg, const double* xc, const double* radbdy, const double* velbdy,
                   void** out) {
  if (!cfg || !xc || !radbdy || !out) return PIB_ERR_ARG;
  DemCtx* D = new DemCtx();
  *out = D;
  D->cfg = *cfg;
  D->n = cfg->nbody;
  D->kn = cfg->kn > 0 ? cfg->kn : 800000.0;
  D->en = cfg->coeff_rest;
  if (cfg->nbody < 1 || cfg->nbody > DEM_NMAX) {
    snprintf(D->err, sizeof D->err, "nbody must lie in [1, %d]", DEM_NMAX);
    return cfg->nbody < 1 ? PIB_ERR_ARG : PIB_ERR_UNSUPPORTED;
  }
  CreateArgs a{xc, radbdy, velbdy};
  return run(D, do_create, &a);
}

int pib_dem_destroy(void* h) {
  DemCtx* D = (DemCtx*)h;
  if (!D) return PIB_OK;
  dem_free(D);
  delete D;
  return PIB_OK;
}

const char* pib_dem_last_error(void* h) { return h ? ((DemCtx*)h)->err : "null context"; }

int pib_dem_march(void* h, double dt, const double* pres, const double* visc, const double* mpg,
                  const double* frame_accln, double* xc, double* velbdy, double* angv) {
  DemCtx* D = (DemCtx*)h;
  if (!D || !D->stream) return PIB_ERR_STATE;
  MarchArgs a{dt, pres, visc, mpg, frame_accln, xc, velbdy, angv};
  return run(D, do_march, &a);
}

static void do_get(DemCtx* D, void* va) {
  void** p = (void**)va;
  const size_t b = 3 * (size_t)D->n * sizeof(double);
  if (p[0]) DEMCHK(hipMemcpy(p[0], D->pos, b, hipMemcpyDeviceToHost));
  if (p[1]) DEMCHK(hipMemcpy(p[1], D->vel, b, hipMemcpyDeviceToHost));
  if (p[2]) std::memset(p[2], 0, b);
  if (p[3]) DEMCHK(hipMemcpy(p[3], D->cforce, b, hipMemcpyDeviceToHost));
  pib_dem_info* info = (pib_dem_info*)p[4];
  if (info) {
    std::memset(info, 0, sizeof *info);
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
    info->rebuilds = D->rebuilds;
  }
}

int pib_dem_get_state(void* h, double* pos, double* vel, double* omega, double* cforce, pib_dem_info* info) {
  DemCtx* D = (DemCtx*)h;
  if (!D || !D->stream) return PIB_ERR_STATE;
  void* p[5] = {pos, vel, omega, cforce, info};
  return run(D, do_get, p);
}

static void do_set(DemCtx* D, void* va) {
  const double** p = (const double**)va;
  const size_t b = 3 * (size_t)D->n * sizeof(double);
  if (p[0]) {
    const pib_dem_config& c = D->cfg;
    const double box[3] = {c.xlength, c.ylength, c.zlength};
    const double cell[3] = {D->dxc, D->dyc, D->dzc};
    for (int d = 0; d < 3; ++d)
      for (int l = 0; l < D->n; ++l)
        if (!(p[0][(size_t)d * D->n + l] > -cell[d] && p[0][(size_t)d * D->n + l] < box[d] + cell[d]))
          dem_fail(PIB_ERR_ARG, "particle centre more than a coarse cell outside the periodic box");
    DEMCHK(hipMemcpy(D->pos, p[0], b, hipMemcpyHostToDevice));
  }
  if (p[1]) DEMCHK(hipMemcpy(D->vel, p[1], b, hipMemcpyHostToDevice));
  DEMCHK(hipMemset(D->disp, 0, b));
  D->need_rebuild = true;
}

int pib_dem_set_state(void* h, const double* pos, const double* vel) {
  DemCtx* D = (DemCtx*)h;
  if (!D || !D->stream) return PIB_ERR_STATE;
  const double* p[2] = {pos, vel};
  return run(D, do_set, p);
}

}  // extern "C"
