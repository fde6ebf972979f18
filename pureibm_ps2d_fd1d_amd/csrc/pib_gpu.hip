// pib_gpu.hip -- MI355X (gfx950) implementation of the C-ABI declared in
// include/pureibm_gpu.h: context and memory management, set-up kernels,
// the device-side scalar control logic of the reference's time step, and
// the orchestration of the per-step kernels of pib_kernels.h.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <map>
#include <mutex>
#include <vector>

#include "pib_kernels.h"
#include "pib_velstep.h"
#include "pib_scalar.h"
#include "pib_scalar_nl.h"
#include "pib_adjoint.h"

namespace {

constexpr int MAXNB = 96;  // neighbour-list capacity per body

// ===========================================================================
// set-up kernels
// ===========================================================================
__global__ void k_fill_u8(uint8_t* p, size_t n, uint8_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_zero_f64(double* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// grid_nodes_insphere (dependent_functions.F90:783-963): one workgroup per
// body walks its bounding box; duplicate solid flags are counted as the
// reference counts them.
__global__ void k_mask_body(uint8_t* mask, const double* xc, const double* rad, int nbody, int nx, int my,
                            int mz, long long* cnt) {
  __shared__ double sh[256];
  int m = blockIdx.x;
  double r = rad[m];
  double cx = xc[m], cy = xc[nbody + m], cz = xc[2 * nbody + m];
  int imin = (int)ceil(cx - r), imax = (int)floor(cx + r);
  int jmin = (int)ceil(cy - r), jmax = (int)floor(cy + r);
  int kmin = (int)ceil(cz - r), kmax = (int)floor(cz + r);
  int ni = imax - imin + 1, nj = jmax - jmin + 1, nk = kmax - kmin + 1;
  long long tot = (long long)ni * nj * nk;
  double c = 0.0;
  for (long long e = threadIdx.x; e < tot; e += blockDim.x) {
    int i = imin + (int)(e / ((long long)nj * nk));
    int rem = (int)(e % ((long long)nj * nk));
    int j = jmin + rem / nk, k = kmin + rem % nk;
    double dist = (i - cx) * (i - cx);
    dist = dist + (j - cy) * (j - cy) + (k - cz) * (k - cz);
    dist = sqrt(dist);
    if ((dist - r) <= PIB_SMALL_NUMBER) {
      int ii = wrap1(i, nx), jj = wrap1(j, my), kk = wrap1(k, mz);
      mask[((size_t)(ii - 1) * mz + (kk - 1)) * my + (jj - 1)] = 0;
      c += 1.0;
    }
  }
  double s = block_sum(c, sh);
  if (threadIdx.x == 0) cnt[m] = (long long)s;
}

// neighbour sets of the soft-sphere search (collision_mod.F90:807-1062):
// pairs with an image inside 1.2*(r_a + r_b) (physical units).  One
// workgroup a sphere; the set is used only as a set (check_external_pt ORs
// over it), so members are appended in any order.
__global__ void k_neighbours(const double* xc, const double* rad, int nbody, double dx, double lx, double ly,
                             double lz, int* nb, int* nbcnt) {
  __shared__ int scnt;
  const int m = blockIdx.x;
  if (threadIdx.x == 0) scnt = 0;
  __syncthreads();
  double pa0 = (xc[m] - 1.0) * dx, pa1 = (xc[nbody + m] - 1.0) * dx, pa2 = (xc[2 * nbody + m] - 1.0) * dx;
  for (int q = threadIdx.x; q < nbody; q += blockDim.x) {
    if (q == m) continue;
    double pb0 = (xc[q] - 1.0) * dx, pb1 = (xc[nbody + q] - 1.0) * dx, pb2 = (xc[2 * nbody + q] - 1.0) * dx;
    double rlm = 1.2 * (rad[m] * dx + rad[q] * dx);
    bool hit = false;
    for (int sx = -1; sx <= 1 && !hit; ++sx)
      for (int sy = -1; sy <= 1 && !hit; ++sy)
        for (int sz = -1; sz <= 1 && !hit; ++sz) {
          double d0 = pb0 + sx * lx - pa0, d1 = pb1 + sy * ly - pa1, d2 = pb2 + sz * lz - pa2;
          double dm = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
          if (rlm - dm > PIB_SMALL_NUMBER) hit = true;
        }
    if (hit) {
      const int slot = atomicAdd(&scnt, 1);
      if (slot < MAXNB) nb[(size_t)m * MAXNB + slot] = q;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) nbcnt[m] = scnt;
}

// check_external_pt (dependent_functions.F90:992-1051): a reversal pair is
// inactive when its outer point, wrapped as the reference wraps it, lies
// inside a neighbouring sphere (unshifted centre).
__global__ void k_if_rev(Particles P, int nx, int my, int mz, const int* nb, const int* nbcnt, uint8_t* if_rev) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P.nbody * P.nbnd) return;
  int m = idx / P.nbnd, l = idx - m * P.nbnd;
  double xlo[3];
  for (int n = 0; n < 3; ++n) xlo[n] = P.xc[n * P.nbody + m] + P.bnd[l * 3 + n] * P.rado[m];
  double mxf = (double)(nx + 1);
  if (xlo[0] < 1.0) xlo[0] = mxf + xlo[0];
  if (xlo[0] > mxf) xlo[0] = xlo[0] - mxf;
  if (xlo[1] < 1.0) xlo[1] = my + 1.0 + xlo[1];
  if (xlo[1] > my) xlo[1] = xlo[1] - my - 1.0;
  if (xlo[2] < 1.0) xlo[2] = mz + 1.0 + xlo[2];
  if (xlo[2] > mz) xlo[2] = xlo[2] - mz - 1.0;
  uint8_t rev = 1;
  int cnt = nbcnt[m] < MAXNB ? nbcnt[m] : MAXNB;
  for (int t = 0; t < cnt && rev; ++t) {
    int q = nb[(size_t)m * MAXNB + t];
    double dist = 0.0;
    for (int n = 0; n < 3; ++n) dist = dist + (P.xc[n * P.nbody + q] - xlo[n]) * (P.xc[n * P.nbody + q] - xlo[n]);
    dist = sqrt(dist);
    if (dist < P.rad[q]) rev = 0;
  }
  if_rev[idx] = rev;
}

// First target of each spreading range: range k holds the targets whose
// first record lies in [k*B, (k+1)*B); bt[nblk] = ntarget.  Target offsets
// are non-decreasing, so every range start has exactly one writer.
__global__ void k_bucket_starts(const int* toff, int ntarget, int B, int nblk, int* bt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntarget) return;
  const int b = t < ntarget ? toff[t] / B : nblk;
  const int bp = t > 0 ? toff[t - 1] / B : -1;
  for (int k = bp + 1; k <= b; ++k) bt[k] = t;
}

// The first record of each spreading range: rs[k] = toff[bt[k]], k = 0..nblk.
__global__ void k_range_starts(const int* toff, const int* bt, int nblk, int* rs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= nblk) rs[k] = toff[bt[k]];
}

__global__ void k_count_active(const uint8_t* if_rev, int nbnd, int* nrpr_active) {
  __shared__ double sh[256];
  int m = blockIdx.x;
  double c = 0.0;
  for (int l = threadIdx.x; l < nbnd; l += blockDim.x) c += if_rev[(size_t)m * nbnd + l];
  double s = block_sum(c, sh);
  if (threadIdx.x == 0) nrpr_active[m] = (int)s;
}

// Spreading records: for each active reversal pair and each of the 8 corners
// of its inner-point stencil, the grid node it deposits on (sentinel when it
// deposits nowhere), its weight and whether the redistributed fluid share is
// added (boundary_condition.F90:1379-1462).  With dobnd, each body's surface
// points come first (calc_bnd_forcing, all eight nodes, no redistribution):
// record ((m*2 + kind)*nbnd + l)*8 + corner, kind 0 surface, 1 reversal, so
// that the stable sort keeps the reference's (body, routine, point, corner)
// order within a node.  rec_point maps a record to its row of the per-point
// forces: reversal pairs first (m*nbnd + l), surface points after them.
__device__ __host__ inline int rec_point(int r8, int nbody, int nbnd, int dobnd) {
  if (!dobnd) return r8;
  const int m = r8 / (2 * nbnd), rest = r8 - m * 2 * nbnd, kind = rest / nbnd, l = rest - kind * nbnd;
  return kind ? m * nbnd + l : nbody * nbnd + m * nbnd + l;
}
__global__ void k_spread_records(Particles P, const uint8_t* mask, int nx, int my, int mz, int frmeanfluid,
                                 int dorpr, int dobnd, unsigned int sentinel, unsigned int* key, int* rec, double* w,
                                 uint8_t* dist) {
  // one lane per record (point idx, corner cidx = (k*2 + j)*2 + i), so that
  // the four record arrays are written coalesced
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int npk = P.nbody * P.nbnd;
  if (q >= (size_t)(dobnd ? 2 * npk : npk) * 8) return;
  const int idx = (int)(q >> 3), cidx = (int)(q & 7);
  const int i = cidx & 1, j = (cidx >> 1) & 1, k = cidx >> 2;
  int m, l, kind;
  if (dobnd) {
    m = idx / (2 * P.nbnd);
    const int rest = idx - m * 2 * P.nbnd;
    kind = rest / P.nbnd;
    l = rest - kind * P.nbnd;
  } else {
    m = idx / P.nbnd;
    l = idx - m * P.nbnd;
    kind = 1;
  }
  const bool act = kind == 0 || (dorpr && P.if_rev[(size_t)m * P.nbnd + l] != 0);
  const double r = kind ? P.radi[m] : P.rad[m];
  double x[3];
  for (int n = 0; n < 3; ++n) x[n] = P.xc[n * P.nbody + m] + P.bnd[l * 3 + n] * r;
  Sten s;
  make_sten(s, x, nx, my, mz);
  const unsigned int node = s.node(i, j, k);
  const bool fl = mask[node] != 0;
  const bool dep = act && (kind == 0 || frmeanfluid ? true : !fl);
  key[q] = dep ? node : sentinel;
  rec[q] = (int)q;
  w[q] = s.w(i, j, k);
  dist[q] = (uint8_t)(kind == 0 || frmeanfluid ? 0 : 1);
}
__global__ void k_gather_records(int n, const int* rec_sorted, const double* w, const uint8_t* dist, int nbody,
                                 int nbnd, int dobnd, int* rpt, double* rw, uint8_t* rd) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int r = rec_sorted[t];
  rpt[t] = rec_point(r / 8, nbody, nbnd, dobnd);
  rw[t] = w[r];
  rd[t] = (uint8_t)(dist[r] | ((r & 1) << 1));  // bit 1: the stencil's upper x node (corner i = 1)
}
// The run of the sentinel key (records that deposit nowhere) sorts last:
// its count is cleared so that the scan and the maximum over the whole count
// array see the target runs only; meta[0] = the number of target nodes.
__global__ void k_runs_fix(const unsigned int* ukeys, int* counts, const int* nruns, unsigned int sentinel, int* meta) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int n = *nruns;
  int nt = n;
  if (n > 0 && ukeys[n - 1] >= sentinel) {
    counts[n - 1] = 0;
    nt = n - 1;
  }
  meta[0] = nt;
}
// meta[1] = the number of target records (offset past the last target)
__global__ void k_runs_total(const int* toff, int* meta) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  meta[1] = toff[meta[0]];
}
// the mask-dependent scalars of the device state after a new particle set
__global__ void k_state_counts(DevState* st, int nb, int nbnd, double cs, double ntot) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  st->nbody = nb;
  st->nbnd = nbnd;
  st->s.count_solid = cs;
  st->s.count_fluid = ntot - cs;
  st->s.maxvolfrac = cs / ntot;
}

// ===========================================================================
// scalar control kernels: one workgroup of SCAL_T lanes; the sums are
// deterministic block trees, the scalar logic runs on lane 0
// ===========================================================================
#define SCAL_T 256
#define DRAG_T 1024

// The serial logic of a scalar kernel runs on lane 0 against a copy of the
// device state in LDS: one coalesced load of the whole state and one store
// back, instead of a global-memory round trip for each field it touches.
static_assert(sizeof(DevState) % sizeof(unsigned) == 0, "state copied in words");
constexpr int STATE_WORDS = (int)(sizeof(DevState) / sizeof(unsigned));
constexpr int STATE_PER_LANE = (STATE_WORDS + SCAL_T - 1) / SCAL_T;
// issue: the lane's words of the state (loads in flight until state_put)
__device__ inline void state_get(unsigned (&w)[STATE_PER_LANE], const DevState* st) {
#pragma unroll
  for (int r = 0; r < STATE_PER_LANE; ++r) {
    int q = threadIdx.x + r * blockDim.x;
    w[r] = q < STATE_WORDS ? ((const unsigned*)st)[q] : 0u;
  }
}
__device__ inline void state_put(DevState* L, const unsigned (&w)[STATE_PER_LANE]) {
#pragma unroll
  for (int r = 0; r < STATE_PER_LANE; ++r) {
    int q = threadIdx.x + r * blockDim.x;
    if (q < STATE_WORDS) ((unsigned*)L)[q] = w[r];
  }
  __syncthreads();
}
__device__ inline void state_store(DevState* st, const DevState* L) {
  __syncthreads();
  for (int q = threadIdx.x; q < STATE_WORDS; q += blockDim.x) ((unsigned*)st)[q] = ((const unsigned*)L)[q];
}

// compute_new_timestep tail (dependent_functions.F90:1143-1277); part holds
// per-(component, plane) fluid/solid/total sums and max|u| from c2r_ubcp
__global__ void __launch_bounds__(SCAL_T) k_scalar_cnt(DevState* stg, const double* part, int nx, int rks,
                                                       const double* velbdy, int nb, double* sconst) {
  __shared__ DevState L;
  __shared__ double sh[12 * 16];
  const int t = threadIdx.x;
  // every global load of the kernel is in flight before the first use
  unsigned sw[STATE_PER_LANE];
  state_get(sw, stg);
  double v[12];
  for (int q = 0; q < 12; ++q) v[q] = 0.0;
  // mx[3]: the fastest particle |velbdy| (dependent_functions.F90:1156-1160)
  double mx[4] = {PIB_SMALL_NUMBER, PIB_SMALL_NUMBER, PIB_SMALL_NUMBER, PIB_SMALL_NUMBER};
  for (int i = t; i < nx; i += SCAL_T)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double* o = part + ((size_t)c * nx + i) * 4;
      v[c * 3 + 0] += o[0];
      v[c * 3 + 1] += o[1];
      v[c * 3 + 2] += o[2];
      mx[c] = fmax(mx[c], o[3]);
    }
  for (int m = t; m < nb; m += SCAL_T) {
    double vb[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      vb[n] = velbdy[(size_t)n * nb + m];
      v[9 + n] += vb[n];
    }
    mx[3] = fmax(mx[3], sqrt(vb[0] * vb[0] + vb[1] * vb[1] + vb[2] * vb[2]));
  }
  state_put(&L, sw);
  DevState* st = &L;
  block_sums_w<12>(v, sh);
  double mxc[4] = {mx[0], mx[1], mx[2], mx[3]};
  block_maxes_w<4>(mxc, sh);
  if (t == 0) {
    pib_scalars& s = st->s;
    double umax_tmp = fmax(fmax(mxc[0], mxc[1]), mxc[2]);
    double dtnew;
    if (umax_tmp * s.dt / st->dx > st->cfl_max)
      dtnew = st->cfl_max * st->dx / umax_tmp;
    else
      dtnew = s.dt;
    // moving particles (:1161-1164): the new dt still comes from the fluid's umax_tmp
    if (st->move_particles && mxc[3] * s.dt / st->dx > st->cfl_max) dtnew = st->cfl_max * st->dx / umax_tmp;
    s.dtnew = dtnew;
    s.umax_tmp = umax_tmp;
    double ntot = (double)((long long)st->nx * st->my * st->mz);
    for (int n = 0; n < 3; ++n) {
      s.ufmean_old[n] = s.ufmean[n];
      s.ufmean[n] = v[n * 3 + 0] / s.count_fluid;
      s.usmean_temp[n] = v[n * 3 + 1] / s.count_solid;
      s.umean_temp[n] = v[n * 3 + 2] / ntot;
    }
    double meanslip[3], mix[3];
    for (int n = 0; n < 3; ++n) {
      double msv = nb > 0 ? v[9 + n] / (double)nb : 0.0;
      s.usmean_act[n] = msv;
      s.dufmeandt[n] = -s.cf * (s.ufmean[n] - s.ufmean_old[n]);
      meanslip[n] = (s.ufmean[n] - s.usmean_des[n]) * (1.0 - s.maxvolfrac);
      double us = 0.0 + s.mean_volfrac * msv;
      s.usmean[n] = us / s.mean_volfrac;
      mix[n] = (1.0 - s.maxvolfrac) * (s.usmean[n] - s.ufmean[n]);
    }
    s.meanslipmod = sqrt(meanslip[0] * meanslip[0] + meanslip[1] * meanslip[1] + meanslip[2] * meanslip[2]);
    s.mixmeanslipmod = sqrt(mix[0] * mix[0] + mix[1] * mix[1] + mix[2] * mix[2]);
    if (rks == st->itrmax && st->adaptive) s.dt = dtnew;
    s.t = s.t + s.dt * (st->coef[rks - 1][0] + st->coef[rks - 1][1]);
    double tempt = s.ramp_frac_time * (double)(float)(st->nx) * st->dx / st->ucharmod;
    if (st->soft_start && s.t < tempt) {
      s.cf = -1.0 / s.dt;
      s.cforig = s.cf;
      s.cf = s.cf * (s.t / tempt);
    } else {
      s.cf = -1.0 / s.dt;
      s.cforig = s.cf;
    }
    // the stage's constants for kernels that run beside the drag's state
    // update: cf, umean, the stage coefficients
    sconst[0] = s.cf;
    for (int n = 0; n < 3; ++n) sconst[1 + n] = s.umean[n];
    for (int q = 0; q < 4; ++q) sconst[4 + q] = st->coef[rks - 1][q];
  }
  state_store(stg, &L);
}

// calc_pres_visc_drag per-body totals (boundary_condition.F90:1631-1645) from
// the (body, chunk) partials of k_drag_points, the compute_mpg_at_n tail
// (boundary_condition.F90:373-435), the per-body forces, then the bcset
// diagnostics of the phase (boundary_condition.F90:723-884).  One workgroup
// of DRAG_T lanes; lane items are (body, quantity) pairs.
__global__ void __launch_bounds__(DRAG_T) k_scalar_drag(DevState* stg, const double* __restrict__ dpart, int G,
                                                        int nb, int rks, const double* __restrict__ rad,
                                                        double* __restrict__ pres, double* __restrict__ visc,
                                                        double* __restrict__ torq, double* __restrict__ force,
                                                        double* __restrict__ force_chem, double* __restrict__ ferr_ring) {
  __shared__ DevState L;
  __shared__ double sh[12 * 16];
  __shared__ double mpg_s[3];
  const int t = threadIdx.x;
  unsigned sw[STATE_PER_LANE];
  state_get(sw, stg);
  double v[12];
  for (int q = 0; q < 12; ++q) v[q] = 0.0;
  // items (body m, quantity q); two items per lane per round, their partial
  // loads (the first 16 chunks) and previous totals in flight together
  const int nitem = nb * 9;
  for (int j0 = t; j0 < nitem; j0 += 2 * DRAG_T) {
    double x[2][16], old[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int j = j0 + u * DRAG_T;
      const bool ok = j < nitem;
      const int m = ok ? j / 9 : 0, q = ok ? j - m * 9 : 0;
#pragma unroll
      for (int r = 0; r < 16; ++r) x[u][r] = ok && r < G ? dpart[((size_t)m * G + r) * 9 + q] : 0.0;
      const double* acc = q < 3 ? pres : (q < 6 ? visc : torq);
      old[u] = ok && rks != 1 ? acc[(size_t)(q % 3) * nb + m] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int j = j0 + u * DRAG_T;
      if (j >= nitem) continue;
      const int m = j / 9, q = j - m * 9, d = q % 3;
      const size_t k = (size_t)d * nb + m;
      // the chunk sums in chunk order
      double tot = 0.0;
#pragma unroll
      for (int r = 0; r < 16; ++r)
        if (r < G) tot += x[u][r];
      for (int g = 16; g < G; ++g) tot += dpart[((size_t)m * G + g) * 9 + q];
      if (q < 3) {
        double pn = old[u] - tot;
        pres[k] = pn;
        v[d] += tot;     // pres_total
        v[6 + d] += pn;  // sum of pres(m)
      } else if (q < 6) {
        double vn = old[u] + tot;
        visc[k] = vn;
        v[3 + d] += tot;  // visc_total
        v[9 + d] += vn;   // sum of visc(m)
      } else {
        torq[k] = old[u] + tot;
      }
    }
  }
  state_put(&L, sw);
  DevState* st = &L;
  block_sums_w<12>(v, sh);
  pib_scalars& s = st->s;
  const double PI = 3.14159265358979323846;
  if (t == 0) {
    double norm_factor = (3.0 * PI * st->vis * (s.mixmeanslipmod + PIB_SMALL_NUMBER) * st->char_length) * (double)nb;
    st->norm_factor = norm_factor;
    for (int n = 0; n < 3; ++n) {
      s.visc_total_old[n] = s.visc_total[n];
      s.pres_total[n] = v[n];
      s.visc_total[n] = v[3 + n];
    }
    const double* pa = v + 6;
    const double* va = v + 9;
    s.pres_drag = sqrt(pa[0] * pa[0] + pa[1] * pa[1] + pa[2] * pa[2]) / norm_factor / (1.0 - s.maxvolfrac);
    s.visc_drag = sqrt(va[0] * va[0] + va[1] * va[1] + va[2] * va[2]) / norm_factor / (1.0 - s.maxvolfrac);
    s.total_drag = s.pres_drag + s.visc_drag;
    if (!st->set_mpg) {
      for (int n = 0; n < 3; ++n) {
        if (st->move_particles) {  // boundary_condition.F90:390-397
          const double rs = st->rhos, rf = st->rhof, mv = s.mean_volfrac;
          s.mpg[n] = s.cf * (s.ufmean_des[n] - s.ufmean[n]) + s.cf * (s.usmean[n] - s.usmean_des[n]);
          s.mpg[n] = s.mpg[n] + (s.pres_total[n] - (s.visc_total[n])) * (1.0 / (rf * (1.0 - mv)) + 1.0 / (mv * rs)) /
                                    st->voldom;
          s.mpg[n] = s.mpg[n] / (1.0 / rf - 1.0 / rs);
          s.frame_accln[n] = -s.mpg[n] / rs - (s.pres_total[n] - s.visc_total[n]) / (mv * rs * st->voldom) +
                             s.cf * (s.usmean_des[n] - s.usmean[n]);
        } else {
          s.mpg[n] = s.cf * (s.ufmean_des[n] - s.ufmean[n]) +
                     (s.pres_total[n] / st->voldom - (s.visc_total[n]) / st->voldom) / (1.0 - s.maxvolfrac);
        }
        if (st->include_frmeanfluid) s.mpg[n] = s.mpg[n] + s.frmeanfluid[n];
      }
    } else if (st->move_particles) {  // :414-417
      for (int n = 0; n < 3; ++n)
        s.frame_accln[n] = -s.mpg[n] / st->rhos -
                           (s.pres_total[n] - s.visc_total[n]) / (s.mean_volfrac * st->rhos * st->voldom) +
                           s.cf * (s.usmean_des[n] - s.usmean[n]);
    }
    for (int n = 0; n < 3; ++n) mpg_s[n] = s.mpg[n];
  }
  __syncthreads();  // pres/visc of every body and mpg_s are in place
  const double c12 = st->coef[rks - 1][0] + st->coef[rks - 1][1];
  double f6[6] = {0, 0, 0, 0, 0, 0};
  for (int j = t; j < nb * 3; j += DRAG_T) {
    const int m = j / 3, n = j - m * 3;
    const size_t k = (size_t)n * nb + m;
    double d = 2.0 * rad[m] * st->dx;
    double vol = d * d * d;
    double fc = c12 * (pres[k] + visc[k]);
    double fo = fc - c12 * mpg_s[n] * PI * vol / 6.0;
    force[k] = fo;
    force_chem[k] = fc;
    f6[n] += fo;
    f6[3 + n] += fc;
  }
  block_sums_w<6>(f6, sh);
  if (t == 0) {
    double af[3], ac[3];
    for (int n = 0; n < 3; ++n) {
      af[n] = (0.0 + s.mean_volfrac * (f6[n] / nb)) / s.mean_volfrac;
      ac[n] = (0.0 + s.mean_volfrac * (f6[3 + n] / nb)) / s.mean_volfrac;
    }
    double nd = sqrt(af[0] * af[0] + af[1] * af[1] + af[2] * af[2]);
    double ndc = sqrt(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]);
    double mpg_avg = sqrt(s.mpg[0] * s.mpg[0] + s.mpg[1] * s.mpg[1] + s.mpg[2] * s.mpg[2]);
    s.total_drag_mpg = mpg_avg * st->voldom / st->norm_factor;
    double nf;
    if (st->set_mpg)
      nf = (3.0 * PI * st->vis * (s.mixmeanslipmod + PIB_SMALL_NUMBER) * st->dia);
    else
      nf = (3.0 * PI * st->vis * st->ucharmod * st->dia);
    s.norm_drag = nd / nf;
    s.norm_drag_chem = ndc / nf;
    if (s.norm_drag > 0.0) s.ferror = fabs(s.norm_drag - s.fold) / s.norm_drag;
    else s.ferror = 1.0;
    s.fold = s.norm_drag;
    // the step's last stage enters its ferror in the history (:828-833): a
    // ring of nerr slots in place of the reference's shifted array
    if (rks == st->itrmax) {
      const int q = st->ferr_pos;
      ferr_ring[q] = s.ferror;
      st->ferr_pos = q + 1 == st->nerr ? 0 : q + 1;
    }
    if (!(s.norm_drag <= (double)1.0E+06f) && st->status == 0) st->status = PIB_ERR_BLOWUP;
  }
  state_store(stg, &L);
}

// frmean / frmeanfluid over all reversal pairs (boundary_condition.F90:632-633)
// from the (body, chunk) partials of k_forcing_points
__global__ void __launch_bounds__(SCAL_T) k_reduce_frmean(DevState* st, const double* fpart, int nrec) {
  __shared__ double sh[6 * 16];
  reduce_frmean_body(st, fpart, nrec, sh);
}

// velstep tail: check_divergence max, umean update (steptheflow.F90:93-99)
__global__ void __launch_bounds__(SCAL_T) k_scalar_vel(DevState* stg, const double* part, int nblk, int* status) {
  __shared__ DevState L;
  __shared__ double sh[16];
  unsigned sw[STATE_PER_LANE];
  state_get(sw, stg);
  double mx[1] = {0.0};
  for (int b = threadIdx.x; b < nblk; b += 4 * SCAL_T) {
    double a[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) a[r] = b + r * SCAL_T < nblk ? part[b + r * SCAL_T] : 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) mx[0] = fmax(mx[0], a[r]);
  }
  state_put(&L, sw);
  block_maxes_w<1>(mx, sh);
  if (threadIdx.x == 0) {
    DevState* st = &L;
    st->s.divmax = mx[0];
    for (int n = 0; n < 3; ++n)
      st->s.umean[n] = st->s.umean[n] + (-st->s.mpg[n] + st->s.frmean[n] - st->s.frame_accln[n]) * st->s.dt;
    if (st->move_particles)  // steptheflow.F90:101-104
      for (int n = 0; n < 3; ++n) {
        st->s.frame_vel[n] = st->s.frame_vel[n] + st->s.frame_accln[n] * st->s.dt;
        st->s.frame_pos[n] = st->s.frame_pos[n] + st->s.frame_vel[n] * st->s.dt;
      }
    int e = *status;
    if (e && st->status == 0) st->status = e;
    *status = 0;  // the velstep kernels only ever raise it
  }
  state_store(stg, &L);
}

// ===========================================================================
// transform-layer drop-in: planes with arbitrary strides
// ===========================================================================
template <typename T>
struct OpRawR2C {
  const T* a;
  C2<T>* b;
  long long lda, ldb;
  int my, my2, q;
  __device__ void load2(int y, int z, T& x0, T& x1) {
    const T* p = a + (size_t)q * lda + (size_t)z * my + y;
    x0 = p[0];
    x1 = p[1];
  }
  __device__ void store(int, int ky, int kz, C2<T> v) { b[(size_t)q * ldb + (size_t)kz * my2 + ky] = v; }
};
template <typename T>
struct OpRawC2R {
  const C2<T>* b;
  T* a;
  long long lda, ldb;
  int my, my2, q;
  __device__ C2<T> load(int ky, int kz) { return b[(size_t)q * ldb + (size_t)kz * my2 + ky]; }
  __device__ void store2(int, int y, int z, T x0, T x1) {
    T* p = a + (size_t)q * lda + (size_t)z * my + y;
    p[0] = x0;
    p[1] = x1;
  }
};
template <typename T>
__global__ void k_raw_r2c(PlaneGeom G, PlaneTabs<T> tb, const T* a, long long lda, C2<T>* b, long long ldb,
                          C2<T>* scr, size_t pslot) {
  extern __shared__ __align__(16) unsigned char smem[];
  OpRawR2C<T> op{a, b, lda, ldb, G.my, G.my2, (int)blockIdx.x};
  plane_r2c_sh<0, 0>(scr ? scr + blockIdx.x * pslot : nullptr, G, tb, op, smem);
}
template <typename T>
__global__ void k_raw_c2r(PlaneGeom G, PlaneTabs<T> tb, const C2<T>* b, long long ldb, T* a, long long lda,
                          C2<T>* scr, size_t pslot) {
  extern __shared__ __align__(16) unsigned char smem[];
  OpRawC2R<T> op{b, a, lda, ldb, G.my, G.my2, (int)blockIdx.x};
  plane_c2r_sh<0, 0>(scr ? scr + blockIdx.x * pslot : nullptr, G, tb, op, smem);
}

// batched cyclic / plain Thomas, rows [i*nsys + s]
__global__ void k_tridiag(int n, int nsys, const double* a, const double* b, const double* c, const double* r,
                          double* x, int cyclic, double* scr) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  size_t ms = nsys;
  double* g = scr;
  double* y1 = scr + (size_t)n * ms;
  double* y2 = scr + (size_t)2 * n * ms;
  double* y3 = scr + (size_t)3 * n * ms;
  auto A = [&](int i) { return a[(size_t)i * ms + s]; };
  auto B = [&](int i) { return b[(size_t)i * ms + s]; };
  auto Cc = [&](int i) { return c[(size_t)i * ms + s]; };
  double auxv0 = 0, auxvn = 0;
  double b0 = B(0), bn = B(n - 1);
  if (cyclic) {
    auxv0 = Cc(n - 1) / 1.0;
    auxvn = A(0) / 1.0;
    b0 = b0 - 1.0 * auxv0;
    bn = bn - 1.0 * auxvn;
  }
  double bet = b0;
  double v1 = r[s] / bet, v2 = 1.0 / bet, v3 = 0.0;
  y1[s] = v1;
  y2[s] = v2;
  y3[s] = v3;
  g[s] = 0;
  for (int j = 1; j < n; ++j) {
    double gam = Cc(j - 1) / bet;
    double bj = (j == n - 1) ? bn : B(j);
    bet = bj - A(j) * gam;
    double r2 = (j == n - 1) ? 1.0 : 0.0;
    v1 = (r[(size_t)j * ms + s] - A(j) * v1) / bet;
    v2 = (r2 - A(j) * v2) / bet;
    g[(size_t)j * ms + s] = gam;
    y1[(size_t)j * ms + s] = v1;
    y2[(size_t)j * ms + s] = v2;
  }
  for (int j = n - 2; j >= 0; --j) {
    double gg = g[(size_t)(j + 1) * ms + s];
    v1 = y1[(size_t)j * ms + s] - gg * v1;
    v2 = y2[(size_t)j * ms + s] - gg * v2;
    y1[(size_t)j * ms + s] = v1;
    y2[(size_t)j * ms + s] = v2;
  }
  double vz = 0, vy = 0;
  if (cyclic) {
    vz = auxv0 * y2[s] + auxvn * y2[(size_t)(n - 1) * ms + s];
    vy = auxv0 * y1[s] + auxvn * y1[(size_t)(n - 1) * ms + s];
  }
  for (int i = 0; i < n; ++i) {
    double yy = y1[(size_t)i * ms + s];
    x[(size_t)i * ms + s] = cyclic ? yy - vy * y2[(size_t)i * ms + s] / (1.0 + vz) : yy;
  }
}

// Workspace of pib_tridiag_batched_noalloc: the four scratch vectors of
// k_tridiag for at most PIB_TRIDIAG_NOALLOC_MAX rows in all.  Each device the
// library runs on holds its own copy until the process ends.  Host code never
// passes this name to a kernel: it uses the device address that
// hipGetSymbolAddress returns, kept per device in scratch_dev.
static __device__ __align__(256) double scratch[4 * (size_t)PIB_TRIDIAG_NOALLOC_MAX];

// ===========================================================================
// host side
// ===========================================================================
#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      throw PibFail(PIB_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    }                                                                              \
  } while (0)

struct PibFail {
  int code;
  std::string msg;
  PibFail(int c, std::string m) : code(c), msg(std::move(m)) {}
};

static bool make_plan(int n, Fft1D& p) {
  p.n = n;
  p.np = 0;
  int m = n;
  while (m % 8 == 0 && p.np < 8) { p.rad[p.np++] = 8; m /= 8; }
  while (m % 4 == 0 && p.np < 8) { p.rad[p.np++] = 4; m /= 4; }
  while (m % 2 == 0 && p.np < 8) { p.rad[p.np++] = 2; m /= 2; }
  while (m % 3 == 0 && p.np < 8) { p.rad[p.np++] = 3; m /= 3; }
  while (m % 5 == 0 && p.np < 8) { p.rad[p.np++] = 5; m /= 5; }
  return m == 1;
}
static std::vector<int> rev_table(const Fft1D& p) {
  std::vector<int> r(p.n);
  for (int k = 0; k < p.n; ++k) {
    int kk = k, span = p.n, pos = 0;
    for (int s = 0; s < p.np; ++s) {
      span /= p.rad[s];
      pos += (kk % p.rad[s]) * span;
      kk /= p.rad[s];
    }
    r[k] = pos;
  }
  return r;
}
template <typename T>
static std::vector<C2<T>> tw_table(int n, int count) {
  std::vector<C2<T>> t(count);
  const long double TWO_PI = 6.283185307179586476925286766559005768L;
  for (int k = 0; k < count; ++k) {
    long double a = -TWO_PI * (long double)k / (long double)n;
    t[k].x = (T)cosl(a);
    t[k].y = (T)sinl(a);
  }
  return t;
}

// What a batch of contexts (pib_batch_*) needs of each context, whatever its
// precision.
struct Base {
  virtual ~Base() {}
  virtual hipStream_t main_stream() const = 0;
  virtual bool graph_ready() const = 0;  // eager steps have set every kernel attribute
  virtual int graph_parity() = 0;        // the nl/onl buffer assignment of the next step
  virtual hipGraph_t capture_step() = 0; // the next step as a graph; the host state is kept
  // the next step enqueued on the context's stream while that stream takes
  // part in a capture begun elsewhere; the host state is kept
  virtual void step_into_capture() = 0;
  virtual void advance_step() = 0;       // the host side of a step replayed elsewhere
  virtual void batch_eager_step() = 0;   // one eager step on the context's own streams
  virtual unsigned long long graph_generation() const = 0;  // bumped when captured graphs go stale
};

template <typename T>
struct Impl : Base {
  pib_config cfg;
  int nx, my, mz, my2, hy, itrmax;
  double coef[3][4];
  hipStream_t stream = nullptr;
  // second stream for the form_nl branch of a stage, forked from and joined
  // back into `stream` with the two events
  hipStream_t stream2 = nullptr, stream_main = nullptr;
  hipEvent_t evfork = nullptr, evjoin = nullptr;
  // forked by default where the transforms hold one workgroup per CU (FP64
  // planes fill the LDS), so that the two branches share the CUs; per-launch
  // timing always runs one stream so that each kernel's time is its own
  bool use_fork = sizeof(T) == 8;
  bool forked = false;
  PlaneGeom G;
  PlaneTabs<T> tb;
  int nt_plane = 256;   // threads per plane workgroup
  size_t lds_plane = 0;
  // device memory
  std::vector<void*> allocs;
  T *wy, *wz, *w2;
  C2<T> *u, *p, *nl, *onl, *ff;
  T *ubcp, *pbcp, *nlbc, *onlbc, *diffn, *omega, *fr, *ppr, *qin;
  double* sconst;  // cf, umean and the coefficients of the current stage
  C2<T>* nlp = nullptr;  // the eight spectra of form_nl by separate transforms
  double* keep_dev = nullptr;  // per-particle force sums parked across a particle update
  size_t keep_n = 0;
  uint8_t* mask;
  DevState* dst;
  double* part;
  size_t part_n = 0;
  T* vscr;
  double* vscr64;
  int* dstatus;
  // scalar transport, one species (allocated by the first scalar_set)
  C2<double>* phi[4] = {nullptr, nullptr, nullptr, nullptr};  // phif, nlphif, onlphif, ffphi
  double* sscr = nullptr;
  // heat_ratio, heat_ratio_new, heat_ratio_old, gamma, phi_fluid_mean,
  // phi_solid_mean on the device (the host values are those last set)
  double* sratio = nullptr;
  double sratio_h[8] = {};
  // the whole scalar step: physical arrays (planes 0..nx+1) and per-body data
  T* sphys = nullptr;
  T *s_nlp = nullptr, *s_onp = nullptr;  // the physical nl and onl of phi (they swap roles a step)
  bool snlp_valid = false;                // s_nlp = ff2cr(nlphif) of the fields as they stand
  double* spart = nullptr;
  double *sflux = nullptr, *sphisurf = nullptr;
  int ssurf_n = -1;
  double* sspt = nullptr;  // per spreading-record point: the scalar forcing's two evaluations
  C2<double>* snlz = nullptr;  // the w phi term of nlphif between k_scal_phi and k_scal_xflux
  T* sph2 = nullptr;           // the second copy of phi that k_scal_phi's w phi workgroups transform
  size_t sspt_n = 0;
  std::vector<double> host_rad;
  double sc_gamma = 0, heat_ratio = 1, heat_ratio_new = 1, heat_ratio_old = 1;
  // particles
  int nbody = 0, nbnd = 0;
  double *xc = nullptr, *rad = nullptr, *radi = nullptr, *rado = nullptr, *vel = nullptr, *angv = nullptr;
  double *bnd = nullptr, *cdn = nullptr;
  uint8_t* if_rev = nullptr;
  int* nrpr = nullptr;
  double *pres = nullptr, *visc = nullptr, *torq = nullptr, *force = nullptr, *force_chem = nullptr;
  double *drag_part = nullptr, *frc_part = nullptr, *frc_pts = nullptr;
  // ferror history ring, PIB_NERR_MAX slots (nerr of them in use)
  double* ferr_ring = nullptr;
  int ntarget = 0, nspread_blk = 0;
  int* spread_bt = nullptr;  // target range of each spreading workgroup
  int* spread_rs = nullptr;  // its first record (toff of its first target)
  int *tnode = nullptr, *toff = nullptr, *rpt = nullptr;
  double* rw = nullptr;
  uint8_t* rdist = nullptr;
  bool have_points = false, have_particles = false, nlbc_valid = true;
  int nl_pending = 0;  // rks of a pib_nonlinear_begin not yet finished
  std::vector<std::pair<void*, size_t>> pallocs;  // particle-dependent allocations and their bytes
  std::multimap<size_t, void*> ppool;  // those of the previous particle set, reused by size
  hipEvent_t ev[5];
  bool timed = false;
  // optional per-launch timing: events recorded around every launch on `stream`
  bool ktime = false;
  int vel_variant = 1;  // 0 thread-per-mode, 1 tiled (L=8/4), 2 tiled (L=4/2)
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  std::vector<std::pair<const char*, std::pair<hipEvent_t, hipEvent_t>>> krec;
  const char* kname = nullptr;
  hipEvent_t kbeg = nullptr;
  hipEvent_t get_ev() {
    if (evused == evpool.size()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      evpool.push_back(e);
    }
    return evpool[evused++];
  }
  void kb(const char* name) {
    if (!ktime) return;
    kname = name;
    kbeg = get_ev();
    HIPCHK(hipEventRecord(kbeg, stream));
  }
  void ke() {
    if (!ktime) return;
    hipEvent_t e = get_ev();
    HIPCHK(hipEventRecord(e, stream));
    krec.push_back({kname, {kbeg, e}});
  }

  template <typename V>
  // zero = false for a buffer that its producer writes whole before any read
  V* alloc(size_t n, bool part_dep = false, bool zero = true) {
    const size_t bytes = n * sizeof(V) + 64;
    void* ptr = nullptr;
    if (part_dep) {
      auto it = ppool.find(bytes);
      if (it != ppool.end()) {
        ptr = it->second;
        ppool.erase(it);
      }
    }
    if (!ptr) HIPCHK(hipMalloc(&ptr, bytes));
    if (zero) HIPCHK(hipMemsetAsync(ptr, 0, bytes, stream));
    if (part_dep)
      pallocs.push_back({ptr, bytes});
    else
      allocs.push_back(ptr);
    return (V*)ptr;
  }
  void drain_pool() {
    for (auto& e : ppool) (void)hipFree(e.second);
    ppool.clear();
  }
  template <typename V>
  V* upload(const std::vector<V>& h) {
    V* d = alloc<V>(h.size());
    HIPCHK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice, stream));
    return d;
  }

  explicit Impl(const pib_config& c) : cfg(c) {
    nx = c.nx;
    my = c.my;
    mz = c.mz;
    my2 = my / 2 + 1;
    hy = my / 2;
    if (c.rk3) {
      double cc[3][4] = {{4.0 / 15.0, 4.0 / 15.0, 8.0 / 15.0, 0.0},
                         {1.0 / 15.0, 1.0 / 15.0, 5.0 / 12.0, -17.0 / 60.0},
                         {1.0 / 6.0, 1.0 / 6.0, 3.0 / 4.0, -5.0 / 12.0}};
      memcpy(coef, cc, sizeof cc);
      itrmax = 3;
    } else {
      memset(coef, 0, sizeof coef);
      coef[0][0] = 0.5;
      coef[0][1] = 0.5;
      coef[0][2] = 3.0 / 2.0;
      coef[0][3] = -0.5;
      itrmax = 1;
    }
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    stream_main = stream;
    HIPCHK(hipEventCreateWithFlags(&evfork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&evjoin, hipEventDisableTiming));
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    // plane FFT geometry
    G.my = my;
    G.mz = mz;
    G.my2 = my2;
    G.hy = hy;
    G.ld = my2;
    if (!make_plan(hy, G.py) || !make_plan(mz, G.pz))
      throw PibFail(PIB_ERR_UNSUPPORTED, "my/2 and mz must factor into 2, 3, 5");
    size_t E = (size_t)my2 * mz;
    nt_plane = E <= 256 ? 64 : (E <= 2500 ? 256 : (E <= 5000 ? 512 : 1024));
    sh = shape_of(hy, mz);
    // form_nl by separate transforms, except on the FP32 128^2 plane, whose
    // in-place form runs its three transforms at four waves per SIMD and
    // measured faster (the autotuner times both)
    nl_sep = !(sizeof(T) == 4 && sh == 1);
    // rows of the generic transform, or of the compiled shape's (padded) layout
    lds_plane = (size_t)mz * std::max(G.ld, shape_ld(sh)) * sizeof(C2<T>);
    size_t red = (size_t)std::max(nt_plane, shape_nt(sh)) * sizeof(double);
    if (lds_plane < red) lds_plane = red;
    lds_plane = (lds_plane + 15) / 16 * 16;
    G.tab_off = (int)lds_plane;
    lds_plane += tab_lds_bytes<T>(hy, mz);
    lds_plane += (size_t)mz * sizeof(C2<T>);  // form_nl: ky = my/2 column accumulators
    // a plane that does not fit the 160 KB LDS of a workgroup is transformed
    // in a slot of HBM scratch (the generic transform; LDS holds the tables)
    hbm_needed = lds_plane > 160 * 1024;
    pslot = (size_t)mz * G.ld;
    lds_hbm = (tab_lds_bytes<T>(hy, mz) + 15) / 16 * 16 + 4096;
    if (hbm_needed) set_plane_hbm(true);
    tb.twy = upload(tw_table<T>(hy, hy));
    tb.twz = upload(tw_table<T>(mz, mz));
    tb.wr = upload(tw_table<T>(my, hy + 1));
    tb.revy = upload(rev_table(G.py));
    tb.revz = upload(rev_table(G.pz));
    // wavenumbers (initialize_flo.F90:735-763), dy = dz = dx
    const long double TWO_PI = 6.283185307179586476925286766559005768L;
    std::vector<T> hwy(my2), hwz(mz), hw2((size_t)my2 * mz);
    std::vector<double> dwy(my2), dwz(mz);
    const double twopi = 2.0 * (4.0 * atan(1.0));
    for (int i = 1; i <= my2; ++i) dwy[i - 1] = twopi * (double)(i - 1) / ((double)my * c.dx);
    for (int i = 1; i <= mz / 2; ++i) {
      dwz[i - 1] = twopi * (double)(i - 1) / ((double)mz * c.dx);
      dwz[mz - i] = -twopi * (double)i / ((double)mz * c.dx);
    }
    (void)TWO_PI;
    for (int j = 0; j < my2; ++j) hwy[j] = (T)dwy[j];
    for (int k = 0; k < mz; ++k) hwz[k] = (T)dwz[k];
    for (int k = 0; k < mz; ++k)
      for (int j = 0; j < my2; ++j) {
        double wyy = -(dwy[j] * dwy[j]), wzz = -(dwz[k] * dwz[k]);
        hw2[(size_t)k * my2 + j] = (T)(-(wyy + wzz));
      }
    wy = upload(hwy);
    wz = upload(hwz);
    w2 = upload(hw2);
    size_t nsc = (size_t)3 * nx * E, nsr = (size_t)3 * nx * my * mz;
    u = alloc<C2<T>>(nsc);
    nl = alloc<C2<T>>(nsc);
    onl = alloc<C2<T>>(nsc);
    ff = alloc<C2<T>>(nsc);
    p = alloc<C2<T>>((size_t)nx * E);
    // fields the IBM stencils gather from carry GATHER_PAD elements past
    // their end (read, never used, by a y-pair load at the last node)
    ubcp = alloc<T>(nsr + GATHER_PAD);
    nlbc = alloc<T>(nsr + GATHER_PAD);
    onlbc = alloc<T>(nsr + GATHER_PAD);
    diffn = alloc<T>(nsr + GATHER_PAD);
    qin = alloc<T>(nsr + GATHER_PAD);
    nlp = alloc<C2<T>>((size_t)8 * nx * E);
    sconst = alloc<double>(8);
    omega = alloc<T>(nsr + GATHER_PAD);
    fr = alloc<T>(nsr);
    ppr = alloc<T>(nsr + GATHER_PAD);
    pbcp = alloc<T>((size_t)nx * my * mz + GATHER_PAD);
    mask = alloc<uint8_t>((size_t)nx * my * mz + GATHER_PAD);
    fill_mask();
    dst = alloc<DevState>(1);
    ferr_ring = alloc<double>(PIB_NERR_MAX, false, false);
    part_n = std::max<size_t>((size_t)3 * nx * 4, (size_t)(E / 8 + 2));
    part = alloc<double>(part_n);
    vscr = alloc<T>((size_t)6 * nx * E);
    vscr64 = alloc<double>((size_t)6 * nx);
    dstatus = alloc<int>(1);
    DevState h;
    memset(&h, 0, sizeof h);
    memcpy(h.coef, coef, sizeof coef);
    h.dx = c.dx;
    h.vis = c.vis;
    h.cfl_max = c.cfl_max;
    h.ucharmod = c.ucharmod;
    h.char_length = c.char_length;
    h.dia = c.dia;
    h.voldom = c.voldom;
    h.dx3 = pow(c.dx, 3.0);
    h.dr = c.dr;
    h.nx = nx;
    h.my = my;
    h.mz = mz;
    h.adaptive = c.adaptive;
    h.soft_start = c.soft_start;
    h.set_mpg = c.set_mpg;
    h.include_frmeanfluid = c.include_frmeanfluid;
    h.itrmax = itrmax;
    h.revernorm = c.revernorm;
    h.move_particles = c.move_particles;
    h.rhos = c.rhos;
    h.rhof = c.rhof;
    h.s.count_fluid = (double)((long long)nx * my * mz);
    // initialize_flo.F90:695: ferror_array = 1 over the default NERR_STEPS
    h.nerr = PIB_NERR_DEFAULT;
    h.ferr_pos = 0;
    std::vector<double> ones(PIB_NERR_MAX, 1.0);
    HIPCHK(hipMemcpyAsync(dst, &h, sizeof h, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(ferr_ring, ones.data(), sizeof(double) * PIB_NERR_MAX, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  ~Impl() override {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    for (auto& c : ck_stack) (void)hipFree(c.first);
    for (void* q : ck_pool) (void)hipFree(q);
    for (void* q : rec_blocks) (void)hipFree(q);
    if (keep_dev) (void)hipFree(keep_dev);
    free_adj_records();
    for (auto& q : pallocs) (void)hipFree(q.first);
    drain_pool();
    for (void* q : allocs) (void)hipFree(q);
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (auto& e : evpool) (void)hipEventDestroy(e);
    if (stream2) (void)hipStreamSynchronize(stream2);
    if (evfork) (void)hipEventDestroy(evfork);
    if (evjoin) (void)hipEventDestroy(evjoin);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void fill_mask() {
    size_t n = (size_t)nx * my * mz;
    k_fill_u8<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(mask, n, 1);
    HIPCHK(hipGetLastError());
  }
  // planes in HBM: one slot of pslot entries per workgroup of a plane launch
  // (at most 6 transforms x nx planes), and 3 x nx more for the form_nl
  // branch, which may run beside the others on stream2
  bool hbm_needed = false, plane_hbm = false;
  C2<T>* pscr = nullptr;
  size_t pslot = 0, lds_hbm = 0;
  void set_plane_hbm(bool on) {
    if (on && !pscr) pscr = alloc<C2<T>>((size_t)14 * nx * pslot);
    plane_hbm = on;
  }
  size_t plane_lds() const { return plane_hbm ? lds_hbm : lds_plane; }
  KArgs<T> args(int rks) const {
    KArgs<T> A;
    A.nx = nx;
    A.my = my;
    A.mz = mz;
    A.my2 = my2;
    A.hy = hy;
    A.dx = (T)cfg.dx;
    A.dx2 = (T)(cfg.dx * cfg.dx);
    A.vis = (T)cfg.vis;
    A.wy = wy;
    A.wz = wz;
    A.w2 = w2;
    A.G = G;
    // six transforms a launch on the main stream, eight (the separate
    // transforms of form_nl) on the second
    A.pscr = plane_hbm ? pscr + (stream == stream_main ? 0 : (size_t)6 * nx * pslot) : nullptr;
    A.pslot = pslot;
    if (plane_hbm) A.G.tab_off = 0;
    A.tb = tb;
    A.u = u;
    A.p = p;
    A.nl = nl;
    A.onl = onl;
    A.ff = ff;
    A.nlp = nlp;
    A.ubcp = ubcp;
    A.pbcp = pbcp;
    A.nlbc = nlbc;
    A.onlbc = onlbc;
    A.diffn = diffn;
    A.omega = omega;
    A.fr = fr;
    A.ppr = ppr;
    A.qin = qin;
    A.sconst = sconst;
    A.mask = mask;
    A.st = dst;
    A.part = part;
    A.rks = rks;
    A.xcd_map = 1;
    return A;
  }
  Particles parts() const {
    Particles P;
    P.xc = xc;
    P.rad = rad;
    P.radi = radi;
    P.rado = rado;
    P.vel = vel;
    P.angv = angv;
    P.bnd = bnd;
    P.cd = cdn;
    P.if_rev = if_rev;
    P.nrpr_active = nrpr;
    P.nbody = nbody;
    P.nbnd = nbnd;
    return P;
  }
  bool capturing = false;
  template <typename K>
  void set_lds(K kern, size_t bytes) {
    if (capturing) return;  // set by the eager steps that precede any capture
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  }
  // plane transforms: the compiled shape sh (pib_fft2.h) when the plane has
  // one and fft_compiled is set, else the generic mixed-radix transform
  int sh = 0;
  bool fft_compiled = true;
  int cur_sh() const { return fft_compiled && !plane_hbm ? sh : 0; }
  int plane_nt(int s) const { return s ? shape_nt(s) : nt_plane; }
  template <int KIND, int SH>
  void launch_c2r(int rks, int ncomp) {
    static const char* names[] = {"c2r_ubcp",  "c2r_pbcp",  "c2r_nlbc",  "c2r_onlbc", "c2r_visc",
                                  "c2r_omega", "c2r_pgrad", "c2r_inner", "c2r_inner_nl"};
    auto kern = k_c2r<T, KIND, SH>;
    set_lds(kern, plane_lds());
    kb(names[KIND]);
    kern<<<dim3(nx, ncomp), plane_nt(SH), plane_lds(), stream>>>(args(rks));
    HIPCHK(hipGetLastError());
    ke();
  }
  template <int KA, int NA, int KB, int NB, int SH>
  void launch_c2r_pair(int rks, const char* name) {
    auto kern = k_c2r_pair<T, KA, NA, KB, SH>;
    set_lds(kern, plane_lds());
    kb(name);
    kern<<<dim3(nx, NA + NB), plane_nt(SH), plane_lds(), stream>>>(args(rks));
    HIPCHK(hipGetLastError());
    ke();
  }
  template <int KA, int NA, int KB, int NB>
  void c2r_pair(int rks, const char* name) {
    switch (cur_sh()) {
      case 1: launch_c2r_pair<KA, NA, KB, NB, 1>(rks, name); break;
      case 2: launch_c2r_pair<KA, NA, KB, NB, 2>(rks, name); break;
      case 3: launch_c2r_pair<KA, NA, KB, NB, 3>(rks, name); break;
      case 4: launch_c2r_pair<KA, NA, KB, NB, 4>(rks, name); break;
      case 5: launch_c2r_pair<KA, NA, KB, NB, 5>(rks, name); break;
      default: launch_c2r_pair<KA, NA, KB, NB, 0>(rks, name);
    }
  }
  template <int KIND>
  void c2r(int rks, int ncomp) {
    switch (cur_sh()) {
      case 1: launch_c2r<KIND, 1>(rks, ncomp); break;
      case 2: launch_c2r<KIND, 2>(rks, ncomp); break;
      case 3: launch_c2r<KIND, 3>(rks, ncomp); break;
      case 4: launch_c2r<KIND, 4>(rks, ncomp); break;
      case 5: launch_c2r<KIND, 5>(rks, ncomp); break;
      default: launch_c2r<KIND, 0>(rks, ncomp);
    }
  }

  // ---------------- set-up ----------------
  void set_points(int n, const double* b, const double* c) {
    if (have_particles) throw PibFail(PIB_ERR_STATE, "set_points must precede set_particles");
    nbnd = n;
    std::vector<double> hb(b, b + 3 * (size_t)n), hc(c, c + 3 * (size_t)n);
    bnd = upload(hb);
    cdn = upload(hc);
    HIPCHK(hipStreamSynchronize(stream));
    have_points = true;
  }
  void set_particles(int nb, const double* hxc, const double* hrad, const double* hvel, const double* hangv,
                     long long* count_solid, int* nrpr_active) {
    if (nb > 0 && !have_points) throw PibFail(PIB_ERR_STATE, "set_points must precede set_particles");
    const int nmin = std::min(nx, std::min(my, mz));
    std::vector<double> hr(nb), hri(nb), hro(nb);
    for (int m = 0; m < nb; ++m) {
      if (!(hrad[m] > cfg.dr && hrad[m] + cfg.dr + 2 < 0.5 * nmin))
        throw PibFail(PIB_ERR_ARG, "radbdy must exceed dr and stay below half the box");
      const int ext[3] = {nx, my, mz};
      for (int d = 0; d < 3; ++d) {
        double v = hxc[(size_t)d * nb + m];
        if (!(v >= -ext[d] && v <= 2.0 * ext[d]))
          throw PibFail(PIB_ERR_ARG, "particle centre outside [-n, 2n] grid units");
      }
      hr[m] = hrad[m];
      hri[m] = hrad[m] - cfg.dr;
      hro[m] = hrad[m] + cfg.dr;
    }
    HIPCHK(hipStreamSynchronize(stream));
    drop_graphs();
    ++part_gen;
    // the per-particle force sums of the step live on when the same bed
    // moves (the reference's grid_nodes_insphere leaves them alone): parked
    // on the device while the buffers are handed out again
    const bool keep = have_particles && nb == nbody && nb > 0;
    if (keep) {
      if (keep_n < (size_t)15 * nb) {
        if (keep_dev) HIPCHK(hipFree(keep_dev));
        keep_dev = nullptr;
        HIPCHK(hipMalloc(&keep_dev, (size_t)15 * nb * sizeof(double)));
        keep_n = (size_t)15 * nb;
      }
      const double* src[5] = {pres, visc, torq, force, force_chem};
      for (int a = 0; a < 5; ++a)
        HIPCHK(hipMemcpyAsync(keep_dev + (size_t)a * 3 * nb, src[a], 3 * nb * sizeof(double),
                              hipMemcpyDeviceToDevice, stream));
    }
    // a moved bed of the same size reuses the previous set's buffers
    for (auto& q : pallocs) ppool.insert({q.second, q.first});
    pallocs.clear();
    nbody = nb;
    size_t n3 = (size_t)3 * (nb > 0 ? nb : 1);
    xc = alloc<double>(n3, true, nb == 0);
    rad = alloc<double>(n3, true, nb == 0);
    radi = alloc<double>(n3, true, nb == 0);
    rado = alloc<double>(n3, true, nb == 0);
    vel = alloc<double>(n3, true, !hvel);
    angv = alloc<double>(n3, true, !hangv);
    pres = alloc<double>(n3, true, !keep);
    visc = alloc<double>(n3, true, !keep);
    torq = alloc<double>(n3, true, !keep);
    force = alloc<double>(n3, true, !keep);
    force_chem = alloc<double>(n3, true, !keep);
    nrpr = alloc<int>(nb > 0 ? nb : 1, true);
    long long* cnt = alloc<long long>(nb > 0 ? nb : 1, true);
    host_rad = hr;
    if (nb > 0) {
      HIPCHK(hipMemcpyAsync(xc, hxc, n3 * sizeof(double), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(rad, hr.data(), nb * sizeof(double), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(radi, hri.data(), nb * sizeof(double), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(rado, hro.data(), nb * sizeof(double), hipMemcpyHostToDevice, stream));
      if (hvel) HIPCHK(hipMemcpyAsync(vel, hvel, n3 * sizeof(double), hipMemcpyHostToDevice, stream));
      if (hangv) HIPCHK(hipMemcpyAsync(angv, hangv, n3 * sizeof(double), hipMemcpyHostToDevice, stream));
      if (keep) {
        double* dstp[5] = {pres, visc, torq, force, force_chem};
        for (int a = 0; a < 5; ++a)
          HIPCHK(hipMemcpyAsync(dstp[a], keep_dev + (size_t)a * 3 * nb, 3 * nb * sizeof(double),
                                hipMemcpyDeviceToDevice, stream));
      }
    }
    // solid mask and counts
    fill_mask();
    std::vector<long long> hc(nb > 0 ? nb : 1, 0);
    std::vector<int> hn(nb > 0 ? nb : 1, 0), hcnt(nb > 0 ? nb : 1, 0);
    if (nb > 0) {
      k_mask_body<<<nb, 256, 0, stream>>>(mask, xc, rad, nb, nx, my, mz, cnt);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(hc.data(), cnt, nb * sizeof(long long), hipMemcpyDeviceToHost, stream));
      // neighbour sets (k_if_rev reads at most MAXNB members; a larger count
      // is refused below), active reversal pairs
      int* nbl = alloc<int>((size_t)nb * MAXNB, true, false);
      int* nbc = alloc<int>(nb, true, false);
      k_neighbours<<<nb, 256, 0, stream>>>(xc, rad, nb, cfg.dx, nx * cfg.dx, my * cfg.dx, mz * cfg.dx, nbl, nbc);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(hcnt.data(), nbc, nb * sizeof(int), hipMemcpyDeviceToHost, stream));
      size_t npts = (size_t)nb * nbnd;
      if_rev = alloc<uint8_t>(npts, true, false);
      Particles P = parts();
      k_if_rev<<<(unsigned)((npts + 255) / 256), 256, 0, stream>>>(P, nx, my, mz, nbl, nbc, if_rev);
      HIPCHK(hipGetLastError());
      k_count_active<<<nb, 256, 0, stream>>>(if_rev, nbnd, nrpr);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(hn.data(), nrpr, nb * sizeof(int), hipMemcpyDeviceToHost, stream));
      const size_t G = (nbnd + 255) / 256;
      // written whole by their kernels every step before any read
      drag_part = alloc<double>((size_t)nb * G * 9, true, false);
      const size_t kinds = cfg.dobnd ? 2 : 1;  // reversal pairs, and the surface points of calc_bnd_forcing
      frc_part = alloc<double>(kinds * nb * G * 6, true, false);
      frc_pts = alloc<double>(kinds * npts * 6, true, false);
      build_spreading();  // waits for the stream: the copies above are in
      for (int m = 0; m < nb; ++m)
        if (hcnt[m] > MAXNB) throw PibFail(PIB_ERR_ARG, "more than 96 neighbours of one sphere (overlapping bed?)");
    }
    long long cs = 0;
    for (int m = 0; m < nb; ++m) cs += hc[m];
    // scalars that depend on the mask, set on the device
    k_state_counts<<<1, 64, 0, stream>>>(dst, nb, nbnd, (double)cs, (double)((long long)nx * my * mz));
    HIPCHK(hipGetLastError());
    if (count_solid) *count_solid = cs;
    if (nrpr_active)
      for (int m = 0; m < nb; ++m) nrpr_active[m] = hn[m];
    // zero the forcing field: only target nodes are written each step
    HIPCHK(hipMemsetAsync(fr, 0, (size_t)3 * nx * my * mz * sizeof(T), stream));
    HIPCHK(hipStreamSynchronize(stream));
    drain_pool();
    have_particles = true;
  }
  void build_spreading() {
    const int dobnd = cfg.dobnd ? 1 : 0;
    size_t nrec = (size_t)nbody * nbnd * 8 * (dobnd ? 2 : 1);
    size_t N = (size_t)nx * my * mz;
    unsigned int sentinel = (unsigned int)N;
    int bits = 1;
    while ((1ull << bits) <= (unsigned long long)N) ++bits;
    // every buffer but the run counts is written whole by its producer
    unsigned int* key = alloc<unsigned int>(nrec, true, false);
    unsigned int* key_s = alloc<unsigned int>(nrec, true, false);
    int* rec = alloc<int>(nrec, true, false);
    int* rec_s = alloc<int>(nrec, true, false);
    double* w = alloc<double>(nrec, true, false);
    uint8_t* dist = alloc<uint8_t>(nrec, true, false);
    Particles P = parts();
    k_spread_records<<<(unsigned)((nrec + 255) / 256), 256, 0, stream>>>(P, mask, nx, my, mz,
                                                                          cfg.include_frmeanfluid, cfg.dorpr, dobnd,
                                                                          sentinel, key, rec, w, dist);
    HIPCHK(hipGetLastError());
    size_t tmp_bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, key_s, rec, rec_s, (int)nrec, 0, bits,
                                              stream));
    void* tmp = alloc<unsigned char>(tmp_bytes + 16, true, false);
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, key_s, rec, rec_s, (int)nrec, 0, bits, stream));
    // run-length encode the sorted keys
    unsigned int* ukeys = alloc<unsigned int>(nrec + 1, true, false);
    int* counts = alloc<int>(nrec + 1, true);  // zero past the runs: the scan and maximum read it whole
    int* nruns = alloc<int>(1, true, false);
    size_t tb2 = 0;
    HIPCHK(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb2, key_s, ukeys, counts, nruns, (int)nrec, stream));
    void* tmp2 = alloc<unsigned char>(tb2 + 16, true, false);
    HIPCHK(hipcub::DeviceRunLengthEncode::Encode(tmp2, tb2, key_s, ukeys, counts, nruns, (int)nrec, stream));
    // target count, scan and maximum on the device, one wait for the sizes
    int* meta = alloc<int>(4, true, false);
    k_runs_fix<<<1, 64, 0, stream>>>(ukeys, counts, nruns, sentinel, meta);
    HIPCHK(hipGetLastError());
    // target nodes are the run keys; record offsets their exclusive prefix sum
    tnode = (int*)ukeys;
    toff = alloc<int>(nrec + 2, true, false);
    int* dmax = alloc<int>(1, true, false);
    size_t tb3 = 0, tb4 = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, counts, toff, (int)nrec + 1, stream));
    HIPCHK(hipcub::DeviceReduce::Max(nullptr, tb4, counts, dmax, (int)nrec, stream));
    void* tmp3 = alloc<unsigned char>(std::max(tb3, tb4) + 16, true, false);
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp3, tb3, counts, toff, (int)nrec + 1, stream));
    HIPCHK(hipcub::DeviceReduce::Max(tmp3, tb4, counts, dmax, (int)nrec, stream));
    k_runs_total<<<1, 64, 0, stream>>>(toff, meta);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(meta + 2, dmax, sizeof(int), hipMemcpyDeviceToDevice, stream));
    int hm[3] = {0, 0, 0};
    HIPCHK(hipMemcpyAsync(hm, meta, sizeof hm, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    ntarget = hm[0];
    const int acc = hm[1], maxcnt = ntarget > 0 ? hm[2] : 0;
    if (maxcnt > SPREAD_CAP)
      throw PibFail(PIB_ERR_UNSUPPORTED, "more reversal records on one node than a workgroup holds");
    // workgroup ranges: the targets whose first record falls in one bucket of
    // B = SPREAD_CAP - maxcnt + 1 records, so that a range holds at most
    // SPREAD_CAP records (buffers sized by the record bound, so that a moved
    // bed of the same size finds them in the reuse pool)
    const int B = SPREAD_CAP - maxcnt + 1;
    nspread_blk = ntarget > 0 ? (acc + B - 1) / B : 0;
    spread_bt = alloc<int>(nrec + 2, true, false);
    k_bucket_starts<<<(unsigned)((ntarget + 1 + 255) / 256), 256, 0, stream>>>(toff, ntarget, B, nspread_blk,
                                                                                 spread_bt);
    HIPCHK(hipGetLastError());
    spread_rs = alloc<int>(nrec + 2, true, false);
    k_range_starts<<<(unsigned)((nspread_blk + 1 + 255) / 256), 256, 0, stream>>>(toff, spread_bt, nspread_blk,
                                                                                    spread_rs);
    HIPCHK(hipGetLastError());
    rpt = alloc<int>(nrec + 1, true, false);
    rw = alloc<double>(nrec + 1, true, false);
    rdist = alloc<uint8_t>(nrec + 1, true, false);
    if (acc > 0) {
      k_gather_records<<<(acc + 255) / 256, 256, 0, stream>>>(acc, rec_s, w, dist, nbody, nbnd, dobnd, rpt, rw,
                                                                rdist);
      HIPCHK(hipGetLastError());
    }
  }

  // ---------------- per-step routines ----------------
  void compute_new_timestep(int rks) {
    c2r_pair<C2R_UBCP, 3, C2R_PBCP, 1>(rks, "c2r_ubcp_pbcp");
    kb("scalar_cnt");
    k_scalar_cnt<<<1, SCAL_T, 0, stream>>>(dst, part, nx, rks, vel ? vel : part, vel ? nbody : 0, sconst);
    HIPCHK(hipGetLastError());
    ke();
  }
  // form_nl keeps NIT spectral accumulators per lane across its three
  // transforms; at most 512 lanes per workgroup leave each lane 256 VGPRs
  int nt_nl() const { return nt_plane < 512 ? nt_plane : 512; }
  // form_nl (and, with particles, the qin transform) of a stage: eight
  // separate transforms a plane, nl written by the qin transform or
  // k_nl_combine (nl_sep); else the in-place three-stage form
  bool nl_sep = true;  // form_nl by separate transforms (PIB_TUNE_NL)
  void nl_terms(int rks, bool inner) {
    if (nl_sep) {
      switch (cur_sh()) {
        case 1: nl_products<1>(rks); break;
        case 2: nl_products<2>(rks); break;
        case 3: nl_products<3>(rks); break;
        case 4: nl_products<4>(rks); break;
        case 5: nl_products<5>(rks); break;
        default: nl_products<0>(rks);
      }
      if (inner) {
        c2r<C2R_INNER_NL>(rks, 3);
      } else {
        kb("nl_combine");
        const size_t n = (size_t)3 * nx * my2 * mz;
        k_nl_combine<T><<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(args(rks));
        HIPCHK(hipGetLastError());
        ke();
      }
      return;
    }
    form_nl(rks);
    if (inner) c2r<C2R_INNER>(rks, 3);
  }
  template <int SH>
  void nl_products(int rks) {
    const bool up = cfg.discr_scheme == PIB_SCHEME_UPWIND, mm = cfg.limiter == PIB_LIMITER_MINMOD;
    auto kern = !up ? k_nl_products<T, false, false, SH> : (!mm ? k_nl_products<T, true, false, SH>
                                                                : k_nl_products<T, true, true, SH>);
    set_lds(kern, plane_lds());
    kb("nl_products");
    kern<<<dim3(nx, 8), plane_nt(SH), plane_lds(), stream>>>(args(rks));
    HIPCHK(hipGetLastError());
    ke();
  }
  void form_nl(int rks) {
    bool up0 = cfg.discr_scheme == PIB_SCHEME_UPWIND, mm0 = cfg.limiter == PIB_LIMITER_MINMOD;
    switch (cur_sh()) {
      case 1: return launch_nl_sh<1>(up0, mm0, rks);
      case 2: return launch_nl_sh<2>(up0, mm0, rks);
      case 3: return launch_nl_sh<3>(up0, mm0, rks);
      case 4: return launch_nl_sh<4>(up0, mm0, rks);
      case 5: return launch_nl_sh<5>(up0, mm0, rks);
      default: break;
    }
    size_t E = (size_t)my2 * mz;
    int nit = (int)((E + nt_nl() - 1) / nt_nl());
    bool up = cfg.discr_scheme == PIB_SCHEME_UPWIND, mm = cfg.limiter == PIB_LIMITER_MINMOD;
    if (plane_hbm || nit > 26) {
      // the in-place form: no accumulators, the plane wherever it lives
      if (!up) run_nl(k_form_nl<T, 1, false, false, -1>, rks, nt_plane);
      else if (!mm) run_nl(k_form_nl<T, 1, true, false, -1>, rks, nt_plane);
      else run_nl(k_form_nl<T, 1, true, true, -1>, rks, nt_plane);
      return;
    }
    if (!launch_nl<1>(nit, up, mm, rks) && !launch_nl<2>(nit, up, mm, rks) && !launch_nl<3>(nit, up, mm, rks) &&
        !launch_nl<5>(nit, up, mm, rks) && !launch_nl<9>(nit, up, mm, rks) && !launch_nl<13>(nit, up, mm, rks) &&
        !launch_nl<17>(nit, up, mm, rks) && !launch_nl<26>(nit, up, mm, rks))
      throw PibFail(PIB_ERR_UNSUPPORTED, "plane too large for the register-accumulated form_nl");
  }
  template <int NIT>
  bool launch_nl(int nit, bool up, bool mm, int rks) {
    if (nit > NIT) return false;
    if (!up) run_nl(k_form_nl<T, NIT, false, false, 0>, rks, nt_nl());
    else if (!mm) run_nl(k_form_nl<T, NIT, true, false, 0>, rks, nt_nl());
    else run_nl(k_form_nl<T, NIT, true, true, 0>, rks, nt_nl());
    return true;
  }
  template <int SH>
  void launch_nl_sh(bool up, bool mm, int rks) {
    constexpr int NT = NlOcc<T, SH>::NT;
    if (!up) run_nl(k_form_nl<T, 1, false, false, SH>, rks, NT);
    else if (!mm) run_nl(k_form_nl<T, 1, true, false, SH>, rks, NT);
    else run_nl(k_form_nl<T, 1, true, true, SH>, rks, NT);
  }
  template <typename K>
  void run_nl(K kern, int rks, int nt, const char* name = "form_nl") {
    set_lds(kern, plane_lds());
    kb(name);
    kern<<<dim3(nx, 3), nt, plane_lds(), stream>>>(args(rks));
    HIPCHK(hipGetLastError());
    ke();
  }
  void nonlinear(int rks) { nonlinear_timed(rks, false); }
  void bcset(int rks) {
    bcset_drag(rks);
    bcset_forcing(rks);
  }
  // bcset up to where the reference moves particles (boundary_condition.F90:
  // 517-538): compute_omega, compute_mpg_at_n and the diagnostics (calc_visc's
  // diffn goes into qin, on the form_nl branch)
  void bcset_drag(int rks) {
    if (skip_drag) return;  // a recorded step
    Particles P = parts();
    c2r<C2R_OMEGA>(rks, 3);
    const int G = (nbnd + 255) / 256;
    kb("drag_points");
    const unsigned nbg = (unsigned)((nbody * G + 7) / 8 * 8);
    k_drag_points<T><<<nbg, 256, 0, stream>>>(args(rks), P, drag_part, G);
    HIPCHK(hipGetLastError());
    ke();
    kb("scalar_drag");
    k_scalar_drag<<<1, DRAG_T, 0, stream>>>(dst, drag_part, G, nbody, rks, rad, pres, visc, torq, force, force_chem,
                                            ferr_ring);
    HIPCHK(hipGetLastError());
    ke();
  }
  // the rest of bcset (:557-900): calc_pgrad and the reversal forcing
  void bcset_forcing(int rks) {
    Particles P = parts();
    const int G = (nbnd + 255) / 256;
    const unsigned nbg = (unsigned)((nbody * G + 7) / 8 * 8);
    c2r<C2R_PGRAD>(rks, 3);
    join_branches();
    if (cfg.dorpr || cfg.dobnd) {
      if (cfg.dorpr) {
        kb("forcing_points");
        k_forcing_points<T><<<nbg, 256, 0, stream>>>(args(rks), P, frc_pts, frc_part, G);
        HIPCHK(hipGetLastError());
        ke();
      } else {
        // the reversal rows of the partials and forces stay zero
        k_zero_f64<<<(unsigned)((nbody * G * 6 + 255) / 256), 256, 0, stream>>>(frc_part, (size_t)nbody * G * 6);
        HIPCHK(hipGetLastError());
      }
      if (cfg.dobnd) {
        kb("bnd_points");
        k_bnd_points<T><<<nbg, 256, 0, stream>>>(args(rks), P, frc_pts, frc_part, G);
        HIPCHK(hipGetLastError());
        ke();
      }
      const int nfrec = G * nbody * (cfg.dobnd ? 2 : 1);
      if (ntarget > 0) {
        // the spreading ranges, then one workgroup for the frmean reduction
        kb("spread");
        k_spread<T><<<(nspread_blk + 1 + 7) / 8 * 8, 256, 0, stream>>>(args(rks), spread_bt, spread_rs,
                                                                      nspread_blk, tnode, toff, rpt, rw, rdist,
                                                                      frc_pts, dst, frc_part, nfrec);
        HIPCHK(hipGetLastError());
        ke();
      } else {
        kb("reduce_frmean");
        k_reduce_frmean<<<1, SCAL_T, 0, stream>>>(dst, frc_part, nfrec);
        HIPCHK(hipGetLastError());
        ke();
      }
    }
    switch (cur_sh()) {
      case 1: r2c_force<1>(rks); break;
      case 2: r2c_force<2>(rks); break;
      case 3: r2c_force<3>(rks); break;
      case 4: r2c_force<4>(rks); break;
      case 5: r2c_force<5>(rks); break;
      default: r2c_force<0>(rks);
    }
  }
  // tiled velstep: M modes per workgroup, L x rows per lane; false when the
  // shape does not fit (nx % L, workgroup size, the (0,0)-mode workgroup)
  template <int L, int M>
  bool velstep_tiled(int rks) {
    if (nx % L != 0 || nx % 2 != 0) return false;
    int bd = M * (nx / L);
    if (bd > 512 || nx / 2 > bd) return false;
    int nm = my2 * mz;
    int ntiles = (nm + M - 1) / M;
    if ((size_t)ntiles + 1 > part_n) return false;
    size_t lds = std::max(vel_lds_bytes(bd), (size_t)bd * sizeof(double));
    auto kern = bd <= 256 ? k_velstep_tiled<T, L, M, 256> : k_velstep_tiled<T, L, M, 512>;
    set_lds(kern, lds);
    kb("velstep");
    kern<<<ntiles + 1, bd, lds, stream>>>(args(rks), ntiles, dstatus);
    HIPCHK(hipGetLastError());
    ke();
    vel_tail(ntiles + 1);
    return true;
  }
  void vel_tail(int nblk) {
    kb("scalar_vel");
    k_scalar_vel<<<1, SCAL_T, 0, stream>>>(dst, part, nblk, dstatus);
    HIPCHK(hipGetLastError());
    ke();
  }
  template <int SH>
  void r2c_force(int rks) {
    auto kern = k_r2c_force<T, SH>;
    set_lds(kern, plane_lds());
    kb("r2c_force");
    kern<<<dim3(nx, 3), plane_nt(SH), plane_lds(), stream>>>(args(rks));
    HIPCHK(hipGetLastError());
    ke();
  }
  void velstep(int rks) {
    if (vel_variant > 0) {
      bool done;
      if constexpr (sizeof(T) == 4)
        done = (vel_variant == 1 && velstep_tiled<8, 16>(rks)) || velstep_tiled<4, 16>(rks);
      else
        done = (vel_variant == 1 && velstep_tiled<4, 8>(rks)) || velstep_tiled<2, 8>(rks);
      if (done) return;
    }
    // thread-per-mode path for shapes the tiled kernel does not take
    int nm = my2 * mz;
    int nblk = (nm + 255) / 256;
    if ((size_t)nblk > part_n) throw PibFail(PIB_ERR_ARG, "partial buffer too small");
    kb("velstep");
    k_velstep<T><<<nblk, 256, 0, stream>>>(args(rks), vscr, vscr64, dstatus);
    HIPCHK(hipGetLastError());
    ke();
    vel_tail(nblk);
  }
  // ---- whole steps replayed from HIP graphs ----
  // A step swaps the nl/onl buffers, so it is captured twice, once per buffer
  // assignment; a step that is timed per launch or per phase runs eagerly.
  // Graphs are captured only after eager steps have run, so every kernel
  // attribute is already set, and are dropped when the particle set or a
  // kernel form changes.
  bool use_graph = true;
  int eager_steps = 0;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  C2<T>* nl_a = nullptr;  // the nl buffer of assignment 0
  unsigned long long graph_gen = 0;
  void drop_graphs() {
    ++graph_gen;
    for (auto& g : gexec)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
    for (auto& g : agexec)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
    for (auto& kv : agexec_rec) (void)hipGraphExecDestroy(kv.second);
    agexec_rec.clear();
    for (auto& kv : rgexec) (void)hipGraphExecDestroy(kv.second);
    rgexec.clear();
    adj_eager = 0;
    lite_eager = 0;
  }
  void eager_step(bool rec) {
    ++state_gen;
    for (int rks = 1; rks <= itrmax; ++rks) {
      bool r = rec && rks == itrmax;
      if (r) HIPCHK(hipEventRecord(ev[0], stream));
      compute_new_timestep(rks);
      if (r) HIPCHK(hipEventRecord(ev[1], stream));
      nonlinear_timed(rks, r);
      if (r) HIPCHK(hipEventRecord(ev[3], stream));
      velstep(rks);
      if (r) HIPCHK(hipEventRecord(ev[4], stream));
    }
  }
  hipStream_t main_stream() const override { return stream; }
  bool graph_ready() const override { return use_graph && !ktime && eager_steps >= 2 && nlbc_valid; }
  int graph_parity() override {
    if (!nl_a) nl_a = nl;
    return nl == nl_a ? 0 : 1;
  }
  unsigned long long graph_generation() const override { return graph_gen; }
  void step_into_capture() override {
    C2<T>*nl0 = nl, *onl0 = onl;
    bool v0 = nlbc_valid;
    auto gen0 = state_gen;
    auto restore = [&] {
      capturing = false;
      nl = nl0;
      onl = onl0;
      nlbc_valid = v0;
      state_gen = gen0;
    };
    capturing = true;
    try {
      eager_step(false);
    } catch (...) {
      restore();
      throw;
    }
    restore();
  }
  hipGraph_t capture_step() override {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      step_into_capture();
    } catch (...) {
      (void)hipStreamEndCapture(stream, &g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(stream, &g));
    return g;
  }
  void advance_step() override {
    ++state_gen;
    std::swap(nl, onl);
    nlbc_valid = true;
    timed = false;
  }
  void batch_eager_step() override {
    if (!have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    eager_step(false);
    ++eager_steps;
    timed = false;
  }
  void graph_step() {
    int par = graph_parity();
    if (!gexec[par]) {
      hipGraph_t g = capture_step();
      hipError_t e = hipGraphInstantiate(&gexec[par], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHK(e);
    }
    HIPCHK(hipGraphLaunch(gexec[par], stream));
    advance_step();
  }
  void step_async(int nsteps) {
    if (!have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    for (int n = 0; n < nsteps; ++n) {
      bool last = n == nsteps - 1;
      if (use_graph && !ktime && !last && eager_steps >= 2 && nlbc_valid) {
        graph_step();
        continue;
      }
      eager_step(last);
      ++eager_steps;
    }
    timed = nsteps > 0;
  }
  // nonlinear + bcset of one stage.  onl <- nl (nl.F90:99-108): the buffers
  // swap roles.  With particles, the form_nl branch (form_nl and the qin
  // transform, which takes nl, onl and u as spectra) runs on stream2 beside
  // the bcset branch up to the pressure gradient (omega transform, drag,
  // drag scalars, pgrad transform): the two write disjoint buffers, and the
  // form_nl branch reads the stage's constants from sconst, not from the
  // device state the drag scalars rewrite.  The branches join before the
  // forcing.  nlbcp, onlbcp and diffn are not formed by the step; the
  // field getters transform them on request.
  // part 0: the whole of nonlinear; 1: up to the drag (nonlinear_begin);
  // 2: the forcing (nonlinear_finish)
  void nonlinear_timed(int rks, bool rec, int part = 0) {
    if (part == 2) {
      if (nbody > 0) bcset_forcing(rks);
      return;
    }
    join_branches();
    std::swap(nl, onl);
    // part 1 leaves the branch running for the caller's DEM and particle
    // update; nonlinear_finish joins it before the forcing
    forked = use_fork && !ktime && nbody > 0;
    if (forked) {
      HIPCHK(hipEventRecord(evfork, stream));
      HIPCHK(hipStreamWaitEvent(stream2, evfork, 0));
      std::swap(stream, stream2);
    }
    nl_terms(rks, nbody > 0);
    if (forked) {
      std::swap(stream, stream2);
      HIPCHK(hipEventRecord(evjoin, stream2));
    }
    nlbc_valid = true;
    if (rec) HIPCHK(hipEventRecord(ev[2], stream));
    if (nbody > 0) {
      if (part == 1)
        bcset_drag(rks);
      else
        bcset(rks);
    }
  }
  void join_branches() {
    if (!forked) return;
    HIPCHK(hipStreamWaitEvent(stream, evjoin, 0));
    forked = false;
  }
  void set_timing(int on) {
    HIPCHK(hipStreamSynchronize(stream));
    ktime = on != 0;
    krec.clear();
    evused = 0;
  }
  int kernel_times(int maxn, char* names, double* total_ms, int* launches) {
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<std::string> nm;
    std::vector<double> tot;
    std::vector<int> cnt;
    for (auto& r : krec) {
      float f = 0;
      HIPCHK(hipEventElapsedTime(&f, r.second.first, r.second.second));
      size_t q = 0;
      while (q < nm.size() && nm[q] != r.first) ++q;
      if (q == nm.size()) {
        nm.push_back(r.first);
        tot.push_back(0);
        cnt.push_back(0);
      }
      tot[q] += f;
      cnt[q] += 1;
    }
    int n = (int)nm.size();
    for (int q = 0; q < n && q < maxn; ++q) {
      snprintf(names + 32 * q, 32, "%s", nm[q].c_str());
      total_ms[q] = tot[q];
      launches[q] = cnt[q];
    }
    krec.clear();
    evused = 0;
    return n;
  }
  int check_status() {
    int st = 0;
    HIPCHK(hipMemcpyAsync(&st, &dst->status, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return st;
  }
  void phase_times(double* ms) {
    HIPCHK(hipStreamSynchronize(stream));
    if (!timed) {
      for (int q = 0; q < 4; ++q) ms[q] = 0;
      return;
    }
    for (int q = 0; q < 4; ++q) {
      float f = 0;
      HIPCHK(hipEventElapsedTime(&f, ev[q], ev[q + 1]));
      ms[q] = f;
    }
  }

  // ---------------- state transfer ----------------
  // spectral (nx+1, my2, mz, ncomp) reference layout <-> device [c][i][kz][ky]
  void put_spec(C2<T>* d, int ncomp, const double* h, bool copy_plane) {
    int ni = copy_plane ? nx + 1 : nx;
    size_t E = (size_t)my2 * mz;
    std::vector<C2<T>> t((size_t)ncomp * nx * E);
    for (int c = 0; c < ncomp; ++c)
      for (int k = 0; k < mz; ++k)
        for (int j = 0; j < my2; ++j)
          for (int i = 0; i < nx; ++i) {
            size_t hi = ((((size_t)c * mz + k) * my2 + j) * ni + i) * 2;
            C2<T>& v = t[(((size_t)c * nx + i) * mz + k) * my2 + j];
            v.x = (T)h[hi];
            v.y = (T)h[hi + 1];
          }
    HIPCHK(hipMemcpyAsync(d, t.data(), t.size() * sizeof(C2<T>), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  void get_spec(const C2<T>* d, int ncomp, double* h, bool copy_plane) {
    int ni = copy_plane ? nx + 1 : nx;
    size_t E = (size_t)my2 * mz;
    std::vector<C2<T>> t((size_t)ncomp * nx * E);
    HIPCHK(hipMemcpyAsync(t.data(), d, t.size() * sizeof(C2<T>), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int c = 0; c < ncomp; ++c)
      for (int k = 0; k < mz; ++k)
        for (int j = 0; j < my2; ++j)
          for (int i = 0; i < ni; ++i) {
            size_t hi = ((((size_t)c * mz + k) * my2 + j) * ni + i) * 2;
            const C2<T>& v = t[(((size_t)c * nx + (i == nx ? 0 : i)) * mz + k) * my2 + j];
            h[hi] = (double)v.x;
            h[hi + 1] = (double)v.y;
          }
  }
  void get_real(const T* d, int ncomp, double* h, bool copy_plane, double sub[3] = nullptr) {
    int ni = copy_plane ? nx + 1 : nx;
    size_t P = (size_t)my * mz;
    std::vector<T> t((size_t)ncomp * nx * P);
    HIPCHK(hipMemcpyAsync(t.data(), d, t.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int c = 0; c < ncomp; ++c)
      for (int k = 0; k < mz; ++k)
        for (int j = 0; j < my; ++j)
          for (int i = 0; i < ni; ++i) {
            size_t hi = (((size_t)c * mz + k) * my + j) * ni + i;
            double v = (double)t[(((size_t)c * nx + (i == nx ? 0 : i)) * mz + k) * my + j];
            if (sub) v = (i == nx) ? 0.0 : v - sub[c];
            h[hi] = v;
          }
  }
  void put_real(T* d, int ncomp, const double* h, bool copy_plane) {
    int ni = copy_plane ? nx + 1 : nx;
    size_t P = (size_t)my * mz;
    std::vector<T> t((size_t)ncomp * nx * P);
    for (int c = 0; c < ncomp; ++c)
      for (int k = 0; k < mz; ++k)
        for (int j = 0; j < my; ++j)
          for (int i = 0; i < nx; ++i)
            t[(((size_t)c * nx + i) * mz + k) * my + j] = (T)h[(((size_t)c * mz + k) * my + j) * ni + i];
    HIPCHK(hipMemcpyAsync(d, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  // scalar fields: (nx+1, my2, mz) reference layout <-> device [i][kz][ky], all nx+1 planes
  C2<double>* scalar_field(int f) {
    if (!phi[0]) throw PibFail(PIB_ERR_STATE, "pib_scalar_set first");
    return phi[f - PIB_F_PHI];
  }
  void put_phi(C2<double>* d, const double* h) {
    size_t E = (size_t)my2 * mz;
    std::vector<C2<double>> t((size_t)(nx + 1) * E);
    for (int k = 0; k < mz; ++k)
      for (int j = 0; j < my2; ++j)
        for (int i = 0; i <= nx; ++i) {
          size_t hi = (((size_t)k * my2 + j) * (nx + 1) + i) * 2;
          t[(size_t)i * E + (size_t)k * my2 + j] = {h[hi], h[hi + 1]};
        }
    HIPCHK(hipMemcpyAsync(d, t.data(), t.size() * sizeof(C2<double>), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  void get_phi(const C2<double>* d, double* h) {
    size_t E = (size_t)my2 * mz;
    std::vector<C2<double>> t((size_t)(nx + 1) * E);
    HIPCHK(hipMemcpyAsync(t.data(), d, t.size() * sizeof(C2<double>), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int k = 0; k < mz; ++k)
      for (int j = 0; j < my2; ++j)
        for (int i = 0; i <= nx; ++i) {
          size_t hi = (((size_t)k * my2 + j) * (nx + 1) + i) * 2;
          const C2<double>& v = t[(size_t)i * E + (size_t)k * my2 + j];
          h[hi] = v.x;
          h[hi + 1] = v.y;
        }
  }
  void scalar_set(double gamma, double hr, double hrn) {
    if (!(gamma > 0) || !(hr > 0) || !(hrn > 0)) throw PibFail(PIB_ERR_ARG, "gamma and heat ratios must be > 0");
    if (!phi[0]) {
      // phi[0] marks the fields as present, so it is set only once every
      // buffer exists (a failed alloc leaves the context without them; the
      // buffers already made are freed with the context)
      size_t n = (size_t)(nx + 1) * my2 * mz;
      C2<double>* f[4];
      for (auto& p : f) p = alloc<C2<double>>(n);
      double* scr = alloc<double>((size_t)4 * nx * my2 * mz);
      sscr = scr;
      sratio = alloc<double>(8);
      for (int q = 3; q >= 0; --q) phi[q] = f[q];
      heat_ratio_old = hr;
    } else {
      scalar_get(sratio_h);  // heat_ratio_old as the device holds it
      heat_ratio_old = sratio_h[2];
    }
    sc_gamma = gamma;
    heat_ratio = hr;
    heat_ratio_new = hrn;
    sratio_h[0] = hr;
    sratio_h[1] = hrn;
    sratio_h[2] = heat_ratio_old;
    sratio_h[3] = gamma;
    HIPCHK(hipMemcpyAsync(sratio, sratio_h, 4 * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  void scalar_ratios(const double* r) {
    if (!phi[0]) throw PibFail(PIB_ERR_STATE, "pib_scalar_set first");
    if (!r || !(r[0] > 0) || !(r[1] > 0) || !(r[2] > 0)) throw PibFail(PIB_ERR_ARG, "heat ratios must be > 0");
    double cur[4];
    scalar_get(cur);
    sratio_h[0] = r[0];
    sratio_h[1] = r[1];
    sratio_h[2] = r[2];
    sratio_h[3] = cur[3];
    HIPCHK(hipMemcpyAsync(sratio, sratio_h, 4 * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  void scalar_get(double* out) {
    if (!phi[0]) throw PibFail(PIB_ERR_STATE, "pib_scalar_set first");
    double r[8];
    HIPCHK(hipMemcpyAsync(r, sratio, sizeof r, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int q = 0; q < 4; ++q) out[q] = r[q];
  }
  void scalstep_modes(int rks) {
    if (rks < 1 || rks > itrmax) throw PibFail(PIB_ERR_ARG, "rks out of range");
    if (!phi[0]) throw PibFail(PIB_ERR_STATE, "pib_scalar_set first");
    const int E = my2 * mz;
    ScalArgs a{nx, my, mz, my2, rks, sc_gamma, sratio, dst, phi[0], phi[1], phi[2], phi[3], sscr};
    // a zero pivot is raised in the device status word, like the fluid
    // solves', and reported by the next call that waits on the stream
    kb("phistep");
    k_phistep<<<(E + PHI_T - 1) / PHI_T, PHI_T, 0, stream>>>(a, &dst->status);
    HIPCHK(hipGetLastError());
    ke();
    k_ratio_roll<<<1, 64, 0, stream>>>(sratio);
    HIPCHK(hipGetLastError());
  }
  // ---- the whole scalar step (stepthescalar.F90:51-95) ----
  void scalar_surface(const double* ps) {
    if (!have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    if (nbody > 0 && !ps) throw PibFail(PIB_ERR_ARG, "phisurfall is null");
    const int n = nbody > 0 ? nbody : 1;
    sphisurf = alloc<double>(n);
    sflux = alloc<double>(n);
    if (nbody > 0) HIPCHK(hipMemcpyAsync(sphisurf, ps, nbody * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    ssurf_n = nbody;
  }
  template <int SH>
  void scal_launch(int rks, const ScalStep<T>& S) {
    const int nt = plane_nt(SH);
    const size_t lds = plane_lds();
    const KArgs<T> A = args(rks);
    {
      auto kern = k_scal_phi<T, SH>;
      set_lds(kern, lds);
      kb("scal_phi");
      kern<<<dim3(nx + 1, 2), nt, lds, stream>>>(A, S);
      HIPCHK(hipGetLastError());
      ke();
    }
    k_scal_ratio<T><<<1, 256, 0, stream>>>(A, S);
    HIPCHK(hipGetLastError());
    if (S.upwind) {
      const size_t n = (size_t)(nx + 1) * my * mz;
      kb("scal_faces");
      k_scal_faces<T><<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(A, S);
      HIPCHK(hipGetLastError());
      ke();
    }
    {
      auto kern = cfg.discr_scheme == PIB_SCHEME_UPWIND ? k_scal_xflux<T, SH, true> : k_scal_xflux<T, SH, false>;
      set_lds(kern, lds);
      kb("scal_xflux");
      kern<<<nx, nt, lds, stream>>>(A, S);
      HIPCHK(hipGetLastError());
      ke();
    }
    {
      auto kern = k_scal_phys<T, SH>;
      set_lds(kern, lds);
      kb("scal_phys");
      kern<<<dim3(nx + 1, S.skip_onl ? 5 : 6), nt, lds, stream>>>(A, S);
      HIPCHK(hipGetLastError());
      ke();
    }
    k_fill_t<T><<<(unsigned)((S.np + 255) / 256), 256, 0, stream>>>(S.fr, S.np);
    HIPCHK(hipGetLastError());
    if (nbody > 0) {
      kb("scal_bnd");
      k_scal_bnd<T><<<nbody, 256, 0, stream>>>(A, parts(), S);
      HIPCHK(hipGetLastError());
      ke();
      if (cfg.dorpr) {
        const int n = nbody * nbnd;
        kb("scal_rev");
        k_scal_rev<T><<<(n + 255) / 256, 256, 0, stream>>>(A, parts(), S, rks);
        HIPCHK(hipGetLastError());
        ke();
      }
      if (nspread_blk > 0) {
        kb("scal_spread");
        k_scal_spread<T><<<(nspread_blk + 7) / 8 * 8, 256, 0, stream>>>(A, S, spread_bt, spread_rs, nspread_blk,
                                                                         tnode, toff,
                                                                         rpt, rw, rdist);
        HIPCHK(hipGetLastError());
        ke();
      }
    }
    {
      auto kern = k_scal_ff<T, SH>;
      set_lds(kern, lds);
      kb("scal_ff");
      kern<<<nx, nt, lds, stream>>>(A, S);
      HIPCHK(hipGetLastError());
      ke();
    }
  }
  void scalstep(int rks) {
    if (rks < 1 || rks > itrmax) throw PibFail(PIB_ERR_ARG, "rks out of range");
    if (!phi[0]) throw PibFail(PIB_ERR_STATE, "pib_scalar_set first");
    if (!have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    if (ssurf_n != nbody) throw PibFail(PIB_ERR_STATE, "pib_scalar_surface first");
    if (nx < 3) throw PibFail(PIB_ERR_ARG, "the scalar step needs nx >= 3");
    if (nl_pending) throw PibFail(PIB_ERR_STATE, "pib_nonlinear_finish first");
    const size_t np = (size_t)(nx + 2) * my * mz;
    if (!sphys) {
      sphys = alloc<T>(9 * np);
      spart = alloc<double>((size_t)4 * (nx + 1));
      s_nlp = sphys + np;
      s_onp = sphys + 2 * np;
      snlp_valid = false;
    }
    // onlphif <- nlphif (nlphi.F90:73-86): the buffers swap roles; nlphif
    // is rewritten whole.  The physical onl is then the last step's nl.
    std::swap(phi[1], phi[2]);
    std::swap(s_nlp, s_onp);
    ScalStep<T> S;
    S.phif = phi[0];
    S.nlphif = phi[1];
    S.onlphif = phi[2];
    S.ffphi = phi[3];
    S.ph = sphys;
    S.nlp = s_nlp;
    S.onp = s_onp;
    S.skip_onl = snlp_valid ? 1 : 0;
    S.df = sphys + 3 * np;
    S.g = sphys + 4 * np;
    S.fr = sphys + 7 * np;
    S.fdd = sphys + 8 * np;
    S.np = np;
    S.part = spart;
    S.ratio = sratio;
    S.flux_body = sflux;
    S.phisurf = sphisurf;
    const size_t nspt = (size_t)(cfg.dobnd ? 2 : 1) * nbody * nbnd * 4;
    if (nspt > sspt_n) {
      sspt = alloc<double>(nspt);
      sspt_n = nspt;
    }
    S.spt = sspt;
    if (!snlz) {
      snlz = alloc<C2<double>>((size_t)nx * my2 * mz);
      sph2 = alloc<T>(np);
    }
    S.nlz = snlz;
    S.ph2 = sph2;
    S.dx = cfg.dx;
    S.upwind = cfg.discr_scheme == PIB_SCHEME_UPWIND;
    S.dobnd = cfg.dobnd;
    S.dorpr = cfg.dorpr;
    S.frmeanfluid = cfg.include_frmeanfluid;
    switch (cur_sh()) {
      case 1: scal_launch<1>(rks, S); break;
      case 2: scal_launch<2>(rks, S); break;
      case 3: scal_launch<3>(rks, S); break;
      case 4: scal_launch<4>(rks, S); break;
      case 5: scal_launch<5>(rks, S); break;
      default: scal_launch<0>(rks, S);
    }
    snlp_valid = true;
    scalstep_modes(rks);
  }
  // flux_body (nbody), and (flux_global, phi_fluid_mean, phi_solid_mean, 0)
  void scalar_flux(double* fb, double* out4) {
    if (!sphys || ssurf_n != nbody) throw PibFail(PIB_ERR_STATE, "no scalar step has run");
    std::vector<double> f(nbody > 0 ? nbody : 1);
    double r[8];
    if (nbody > 0) HIPCHK(hipMemcpyAsync(f.data(), sflux, nbody * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(r, sratio, sizeof r, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    // flux_global = sum of flux_body dA1 in body order (bc_scalar.F90:185)
    double fg = 0.0;
    for (int m = 0; m < nbody; ++m) {
      const double rb = host_rad[m];
      const double da1 = 4.0 * 3.14159265358979323846264338327950288 * (rb * cfg.dx) * (rb * cfg.dx) / (double)(float)nbnd;
      fg = fg + f[m] * da1;
      if (fb) fb[m] = f[m];
    }
    if (out4) {
      out4[0] = fg;
      out4[1] = r[4];
      out4[2] = r[5];
      out4[3] = 0.0;
    }
  }
  void set_field(int f, const double* h) {
    ++state_gen;
    if (f >= PIB_F_PHI && f <= PIB_F_FFPHI) {
      snlp_valid = false;
      return put_phi(scalar_field(f), h);
    }
    switch (f) {
      case PIB_F_U: put_spec(u, 3, h, true); break;
      case PIB_F_P: put_spec(p, 1, h, false); break;
      case PIB_F_NL: put_spec(nl, 3, h, true); nlbc_valid = false; break;
      case PIB_F_ONL: put_spec(onl, 3, h, true); break;
      case PIB_F_FF: put_spec(ff, 3, h, true); break;
      case PIB_F_UBCP: put_real(ubcp, 3, h, true); break;
      case PIB_F_PBCP: put_real(pbcp, 1, h, false); break;
      case PIB_F_AU: need_adj(); put_spec(J.au, 3, h, true); break;
      case PIB_F_AP: need_adj(); put_spec(J.ap, 1, h, false); break;
      case PIB_F_ANL: need_adj(); put_spec(J.anl, 3, h, true); break;
      default: throw PibFail(PIB_ERR_ARG, "field cannot be set");
    }
  }
  void get_field(int f, double* h) {
    if (f >= PIB_F_PHI && f <= PIB_F_FFPHI) return get_phi(scalar_field(f), h);
    switch (f) {
      case PIB_F_U: get_spec(u, 3, h, true); break;
      case PIB_F_P: get_spec(p, 1, h, false); break;
      case PIB_F_NL: get_spec(nl, 3, h, true); break;
      case PIB_F_ONL: get_spec(onl, 3, h, true); break;
      case PIB_F_FF: get_spec(ff, 3, h, true); break;
      case PIB_F_UBCP: get_real(ubcp, 3, h, true); break;
      case PIB_F_PBCP: get_real(pbcp, 1, h, false); break;
      // not formed by the step (the forcing takes qin): transformed here
      // from nl, onl and u as they stand
      case PIB_F_NLBCP:
        c2r<C2R_NLBC>(1, 3);
        get_real(nlbc, 3, h, true);
        break;
      case PIB_F_ONLBCP:
        c2r<C2R_ONLBC>(1, 3);
        get_real(onlbc, 3, h, true);
        break;
      case PIB_F_DIFFN:
        c2r<C2R_VISC>(1, 3);
        get_real(diffn, 3, h, true);
        break;
      case PIB_F_PPR: get_real(ppr, 3, h, false); break;
      case PIB_F_OMEGA: get_real(omega, 3, h, true); break;
      case PIB_F_AU: need_adj(); get_spec(J.au, 3, h, true); break;
      case PIB_F_AP: need_adj(); get_spec(J.ap, 1, h, false); break;
      case PIB_F_ANL: need_adj(); get_spec(J.anl, 3, h, true); break;
      case PIB_F_FR: {
        DevState s;
        HIPCHK(hipMemcpyAsync(&s, dst, sizeof s, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        get_real(fr, 3, h, true, s.s.frmean);
        break;
      }
      case PIB_F_MASK: {
        std::vector<uint8_t> t((size_t)nx * my * mz);
        HIPCHK(hipMemcpyAsync(t.data(), mask, t.size(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int k = 0; k < mz; ++k)
          for (int j = 0; j < my; ++j)
            for (int i = 0; i < nx; ++i)
              h[((size_t)k * my + j) * nx + i] = t[((size_t)i * mz + k) * my + j] ? 1.0 : 0.0;
        break;
      }
      default: throw PibFail(PIB_ERR_ARG, "unknown field");
    }
  }
  void set_scalars(const pib_scalars* s) {
    ++state_gen;
    // ordered after the work already queued on the main stream
    DevState h;
    HIPCHK(hipMemcpyAsync(&h, dst, sizeof h, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    double cs = h.s.count_solid, cf = h.s.count_fluid, mv = h.s.maxvolfrac;
    h.s = *s;
    if (have_particles) {
      h.s.count_solid = cs;
      h.s.count_fluid = cf;
      h.s.maxvolfrac = mv;
    }
    HIPCHK(hipMemcpyAsync(dst, &h, sizeof h, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  // ferror_array(1:n) newest first, as the reference shifts it
  // (boundary_condition.F90:828-833), and ferror_hist = SUM(...)/n summed in
  // that order
  void get_ferror_history(int n, double* arr, double* hist) {
    DevState h;
    HIPCHK(hipMemcpyAsync(&h, dst, sizeof h, hipMemcpyDeviceToHost, stream));
    std::vector<double> ring(PIB_NERR_MAX);
    HIPCHK(hipMemcpyAsync(ring.data(), ferr_ring, sizeof(double) * PIB_NERR_MAX, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (n != h.nerr) throw PibFail(PIB_ERR_ARG, "the history holds NERR_STEPS entries: n must equal it");
    double sum = 0.0;
    for (int k = 0; k < n; ++k) {
      const double v = ring[(size_t)((h.ferr_pos - 1 - k + 2 * n) % n)];
      if (arr) arr[k] = v;
      sum += v;
    }
    if (hist) *hist = sum / n;
  }
  // a new history of n entries, newest first (ones when arr is null, as
  // initialize_flo.F90:695 sets it; a restart passes the saved array)
  void set_ferror_history(int n, const double* arr) {
    if (n < 1 || n > PIB_NERR_MAX) throw PibFail(PIB_ERR_ARG, "NERR_STEPS out of range (1..PIB_NERR_MAX)");
    std::vector<double> ring(PIB_NERR_MAX, 1.0);
    // slot n-1-k holds entry k, so the next ferror (slot 0) drops the oldest
    for (int k = 0; k < n; ++k) ring[(size_t)(n - 1 - k)] = arr ? arr[k] : 1.0;
    ++state_gen;
    // ordered after the work already queued on the main stream
    DevState h;
    HIPCHK(hipMemcpyAsync(&h, dst, sizeof h, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    h.nerr = n;
    h.ferr_pos = 0;
    HIPCHK(hipMemcpyAsync(dst, &h, sizeof h, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(ferr_ring, ring.data(), sizeof(double) * PIB_NERR_MAX, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  // copies ordered on the main stream only (a form_nl branch in flight on
  // the second stream writes neither)
  void get_scalars(pib_scalars* s) {
    DevState h;
    HIPCHK(hipMemcpyAsync(&h, dst, sizeof h, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    *s = h.s;
  }
  void get_forces(double* f, double* pr, double* vi, double* tq) {
    size_t n = (size_t)3 * nbody * sizeof(double);
    if (nbody == 0) return;
    if (f) HIPCHK(hipMemcpyAsync(f, force, n, hipMemcpyDeviceToHost, stream));
    if (pr) HIPCHK(hipMemcpyAsync(pr, pres, n, hipMemcpyDeviceToHost, stream));
    if (vi) HIPCHK(hipMemcpyAsync(vi, visc, n, hipMemcpyDeviceToHost, stream));
    if (tq) HIPCHK(hipMemcpyAsync(tq, torq, n, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }

  // ---------------- drop-in transforms ----------------
  void ff2rc_planes(const double* a, long long lda, int np, double* b, long long ldb) {
    size_t na = (size_t)np * my * mz, nb2 = (size_t)np * my2 * mz;
    std::vector<T> ha(na);
    for (int q = 0; q < np; ++q)
      for (size_t e = 0; e < (size_t)my * mz; ++e) ha[(size_t)q * my * mz + e] = (T)a[(size_t)q * lda + e];
    T* da = alloc<T>(na, true);
    C2<T>* db = alloc<C2<T>>(nb2, true);
    HIPCHK(hipMemcpyAsync(da, ha.data(), na * sizeof(T), hipMemcpyHostToDevice, stream));
    auto kern = k_raw_r2c<T>;
    PlaneGeom g = G;
    C2<T>* scr = nullptr;
    if (plane_hbm) {
      g.tab_off = 0;
      scr = alloc<C2<T>>((size_t)np * pslot, true);
    }
    set_lds(kern, plane_lds());
    kern<<<np, nt_plane, plane_lds(), stream>>>(g, tb, da, (long long)my * mz, db, (long long)my2 * mz, scr, pslot);
    HIPCHK(hipGetLastError());
    std::vector<C2<T>> hb(nb2);
    HIPCHK(hipMemcpyAsync(hb.data(), db, nb2 * sizeof(C2<T>), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int q = 0; q < np; ++q)
      for (size_t e = 0; e < (size_t)my2 * mz; ++e) {
        b[2 * ((size_t)q * ldb + e)] = (double)hb[(size_t)q * my2 * mz + e].x;
        b[2 * ((size_t)q * ldb + e) + 1] = (double)hb[(size_t)q * my2 * mz + e].y;
      }
    free_tail(scr ? 3 : 2);
  }
  void ff2cr_planes(const double* b, long long ldb, int np, double* a, long long lda) {
    size_t na = (size_t)np * my * mz, nb2 = (size_t)np * my2 * mz;
    std::vector<C2<T>> hb(nb2);
    for (int q = 0; q < np; ++q)
      for (size_t e = 0; e < (size_t)my2 * mz; ++e) {
        hb[(size_t)q * my2 * mz + e].x = (T)b[2 * ((size_t)q * ldb + e)];
        hb[(size_t)q * my2 * mz + e].y = (T)b[2 * ((size_t)q * ldb + e) + 1];
      }
    T* da = alloc<T>(na, true);
    C2<T>* db = alloc<C2<T>>(nb2, true);
    HIPCHK(hipMemcpyAsync(db, hb.data(), nb2 * sizeof(C2<T>), hipMemcpyHostToDevice, stream));
    auto kern = k_raw_c2r<T>;
    PlaneGeom g = G;
    C2<T>* scr = nullptr;
    if (plane_hbm) {
      g.tab_off = 0;
      scr = alloc<C2<T>>((size_t)np * pslot, true);
    }
    set_lds(kern, plane_lds());
    kern<<<np, nt_plane, plane_lds(), stream>>>(g, tb, db, (long long)my2 * mz, da, (long long)my * mz, scr, pslot);
    HIPCHK(hipGetLastError());
    std::vector<T> ha(na);
    HIPCHK(hipMemcpyAsync(ha.data(), da, na * sizeof(T), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int q = 0; q < np; ++q)
      for (size_t e = 0; e < (size_t)my * mz; ++e) a[(size_t)q * lda + e] = (double)ha[(size_t)q * my * mz + e];
    free_tail(scr ? 3 : 2);
  }
  void free_tail(int n) {
    for (int q = 0; q < n && !pallocs.empty(); ++q) {
      (void)hipFree(pallocs.back().first);
      pallocs.pop_back();
    }
  }

  // ---------------- adjoint of the fixed-bed step ----------------
  AdjArgs<T> J{};
  hipGraphExec_t agexec[4] = {};  // (nl buffer, adjoint buffer) assignments
  std::map<int, hipGraphExec_t> agexec_rec;  // steps on recorded fields: (record depth, adjoint buffer)
  C2<T>* adj_au0 = nullptr;  // the au buffer of adjoint assignment 0
  int adj_eager = 0;
  bool have_adj = false;
  int adj_nbody = -1;
  double* wdev = nullptr;
  int part_gen = 0, adj_gen = -1;  // particle sets seen, and the one the adjoint's records are of
  std::vector<void*> adj_rec_mem;
  void free_adj_records() {
    for (void* q : adj_rec_mem) (void)hipFree(q);
    adj_rec_mem.clear();
  }
  void adjoint_begin() {
    if (!have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    const size_t E = (size_t)my2 * mz, nsc = (size_t)3 * nx * E, np = (size_t)nx * E;
    const size_t npr = (size_t)nx * my * mz;
    if (!have_adj) {
      J.au = alloc<C2<T>>(nsc);
      J.aunew = alloc<C2<T>>(nsc);
      J.anl = alloc<C2<T>>(nsc);
      J.anlp = alloc<C2<T>>(nsc);
      J.aff = alloc<C2<T>>(nsc);
      J.Z = alloc<C2<T>>(nsc);
      J.ap = alloc<C2<T>>(np);
      J.apnew = alloc<C2<T>>(np);
      for (T** f : {&J.aubcp, &J.aom, &J.afr, &J.appr, &J.S}) *f = alloc<T>(3 * npr);
      J.apbcp = alloc<T>(npr);
      J.Q = alloc<T>(8 * npr);
      J.sc = alloc<double>(12);
      J.part = alloc<double>((size_t)3 * nx);
      J.vs = alloc<double>((size_t)12 * nx * E);
      have_adj = true;
    }
    if (adj_nbody != nbody) {
      J.axc = alloc<double>((size_t)3 * std::max(nbody, 1));
      wdev = alloc<double>((size_t)3 * std::max(nbody, 1));
      J.coef = alloc<double>((size_t)6 * std::max(nbody * nbnd, 1));
      adj_nbody = nbody;
    }
    if (adj_gen != part_gen) {
      HIPCHK(hipStreamSynchronize(stream));
      free_adj_records();
      adjoint_build_records(parts(), nx, my, mz, stream, J.rec, adj_rec_mem);
      adj_gen = part_gen;
    }
    rec_depth = 0;
    HIPCHK(hipMemsetAsync(J.au, 0, nsc * sizeof(C2<T>), stream));
    HIPCHK(hipMemsetAsync(J.anl, 0, nsc * sizeof(C2<T>), stream));
    HIPCHK(hipMemsetAsync(J.ap, 0, np * sizeof(C2<T>), stream));
    HIPCHK(hipMemsetAsync(J.sc, 0, 12 * sizeof(double), stream));
    HIPCHK(hipMemsetAsync(J.axc, 0, (size_t)3 * nbody * sizeof(double), stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  void need_adj() const {
    if (!have_adj) throw PibFail(PIB_ERR_STATE, "pib_adjoint_begin first");
  }
  void adjoint_step(const double* wforce) {
    need_adj();
    if (adj_nbody != nbody || adj_gen != part_gen)
      throw PibFail(PIB_ERR_STATE, "particle set changed since pib_adjoint_begin");
    const pib_config& c = cfg;
    if (c.rk3 || c.adaptive || !c.set_mpg || c.discr_scheme != PIB_SCHEME_CENTER || !c.dorpr || c.dobnd ||
        c.include_frmeanfluid || c.move_particles)
      throw PibFail(PIB_ERR_UNSUPPORTED, "adjoint: rk3=0, adaptive=0, set_mpg, center, dorpr, fixed bed only");
    const bool drag = wforce != nullptr && nbody > 0;
    J.wforce = nullptr;
    if (drag) {
      HIPCHK(hipMemcpyAsync(wdev, wforce, (size_t)3 * nbody * sizeof(double), hipMemcpyHostToDevice, stream));
      J.wforce = wdev;
    }
    // without a force weight, the recorded fields of the step when there are
    // any: the kernels read them where they were recorded
    const void* rec = nullptr;
    if (!drag && rec_depth > 0) rec = rec_blocks[--rec_depth];
    // steps without a force weight replay from HIP graphs, one per (nl buffer,
    // adjoint buffer) assignment, or per (record, adjoint buffer) on recorded
    // fields, once an eager adjoint step has set every kernel attribute
    const bool graph = use_graph && !drag && !ktime && adj_eager >= 1 && (rec || nlbc_valid);
    if (!graph) {
      adjoint_body(drag, rec);
      ++adj_eager;
      return;
    }
    if (!nl_a) nl_a = nl;
    if (!adj_au0) adj_au0 = J.au;
    const int apar = J.au == adj_au0 ? 0 : 1;
    hipGraphExec_t* slot = rec ? &agexec_rec[2 * rec_depth + apar] : &agexec[(nl == nl_a ? 0 : 1) + 2 * apar];
    if (!*slot) {
      C2<T>*nl0 = nl, *onl0 = onl;
      const bool v0 = nlbc_valid;
      const AdjArgs<T> J0 = J;
      auto restore = [&] {
        nl = nl0;
        onl = onl0;
        nlbc_valid = v0;
        J = J0;
      };
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      capturing = true;
      try {
        adjoint_body(false, rec);
      } catch (...) {
        capturing = false;
        (void)hipStreamEndCapture(stream, &g);
        restore();
        throw;
      }
      capturing = false;
      HIPCHK(hipStreamEndCapture(stream, &g));
      HIPCHK(hipGraphInstantiate(slot, g, nullptr, nullptr, 0));
      HIPCHK(hipGraphDestroy(g));
      restore();
    }
    HIPCHK(hipGraphLaunch(*slot, stream));
    if (!rec) {
      std::swap(nl, onl);
      nlbc_valid = true;
    }
    std::swap(J.au, J.aunew);
    std::swap(J.ap, J.apnew);
    std::swap(J.anl, J.anlp);
  }
  // the work of one adjoint step on the stream: the step's forward fields
  // that the adjoint reads -- compute_new_timestep (ubcp, pbcp, t, cf),
  // nonlinear's onl <- nl, form_nl and qin, bcset's ppr (and omega when the
  // step's drag is weighed); the forward drag, forcing and ff are not needed
  // -- then the transposed kernels.  rec: a record block whose fields and
  // device state the kernels read instead.
  void adjoint_body(bool drag, const void* rec) {
    if (!rec) {
      compute_new_timestep(1);
      std::swap(nl, onl);
      nl_terms(1, nbody > 0);
      nlbc_valid = true;
      if (nbody > 0) {
        if (drag) c2r<C2R_OMEGA>(1, 3);
        c2r<C2R_PGRAD>(1, 3);
      }
    }
    KArgs<T> A = args(1);
    if (rec) {
      const size_t cs3 = (size_t)3 * nx * my * mz;
      T* b = (T*)rec;
      A.ubcp = b;
      A.qin = b + cs3;
      A.ppr = b + 2 * cs3;
      A.st = (DevState*)(b + 3 * cs3);
    }
    AdjGeom g{cur_sh(), nt_plane, plane_lds(), nbody, nbnd, stream, coef[0][0] + coef[0][1], capturing};
    adjoint_step_kernels<T>(A, parts(), J, g);
    std::swap(J.au, J.aunew);
    std::swap(J.ap, J.apnew);
    std::swap(J.anl, J.anlp);
  }
  void adjoint_get(double* dxc, double* dmpg, double* dum) {
    need_adj();
    double h[12];
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(h, J.sc, sizeof h, hipMemcpyDeviceToHost));
    if (dxc && nbody > 0) HIPCHK(hipMemcpy(dxc, J.axc, (size_t)3 * nbody * sizeof(double), hipMemcpyDeviceToHost));
    for (int n = 0; n < 3; ++n) {
      if (dmpg) dmpg[n] = h[3 + n];
      if (dum) dum[n] = h[n];
    }
  }
  void adjoint_set_umean(const double* dum) {
    need_adj();
    HIPCHK(hipMemcpyAsync(J.sc, dum, 3 * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }

  // ---------------- checkpoint stack of the forward state ----------------
  std::vector<std::pair<void*, bool>> ck_stack;  // device copy, nlbc_valid
  std::vector<void*> ck_pool;
  size_t part_bytes(int q) const {
    const size_t E = (size_t)my2 * mz;
    const size_t sz[4] = {(size_t)3 * nx * E * sizeof(C2<T>), (size_t)3 * nx * E * sizeof(C2<T>),
                          (size_t)nx * E * sizeof(C2<T>), sizeof(DevState)};
    return sz[q];
  }
  // offsets of u, nl, p and the device state in one checkpoint block
  size_t ck_parts(size_t* off) const {
    size_t o = 0;
    for (int q = 0; q < 4; ++q) {
      off[q] = o;
      o += (part_bytes(q) + 255) / 256 * 256;
    }
    off[4] = o;
    return o;
  }
  void checkpoint_push() {
    size_t off[5];
    const size_t bytes = ck_parts(off);
    void* m = nullptr;
    if (!ck_pool.empty()) {
      m = ck_pool.back();
      ck_pool.pop_back();
    } else {
      HIPCHK(hipMalloc(&m, bytes));
    }
    char* b = (char*)m;
    const void* src[4] = {u, nl, p, dst};
    for (int q = 0; q < 4; ++q)
      HIPCHK(hipMemcpyAsync(b + off[q], src[q], part_bytes(q), hipMemcpyDeviceToDevice, stream));
    ck_stack.push_back({m, nlbc_valid});
  }
  // one whole step, replayed from the step graphs once they exist
  void one_step() {
    if (use_graph && !ktime && eager_steps >= 2 && nlbc_valid) {
      graph_step();
      return;
    }
    eager_step(false);
    ++eager_steps;
  }
  void checkpoint_steps(int nsteps) {
    if (!have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    for (int n = 0; n < nsteps; ++n) {
      checkpoint_push();
      one_step();
    }
  }
  // ---- the recorded forward fields of steps to be reversed ----
  // After a whole step, ubcp, qin and ppr still hold that step's values
  // (velstep writes only u and p), together with the step's t, dt and cf in
  // the device state: what the adjoint of the step reads besides the mask
  // and the bed.  adjoint_record runs steps whose ubcp, qin and ppr are
  // written straight into the record block (the context's buffers are
  // pointed there for the step) and copies the device state after it;
  // adjoint_step then takes them instead of recomputing them.  Record blocks
  // by stack depth (a segment's records are pushed and taken in stack order,
  // so a depth's block always holds the same step position and the graphs
  // of steps on recorded fields stay valid); each block is the three
  // 3-component fields, then the device state, then GATHER_PAD elements for
  // the stencil loads past the last field.
  std::vector<void*> rec_blocks;
  int rec_depth = 0;
  // state_gen counts whatever changes the forward state (steps, restores,
  // sets)
  unsigned long long state_gen = 0;
  // recorded steps skip the drag and its diagnostics (the vorticity, the
  // per-particle forces, norm_drag): with the mean pressure gradient set and
  // the bed fixed nothing of the state depends on them
  bool skip_drag = false;
  int lite_eager = 0;  // eager recorded steps since the graphs were dropped
  // graphs of recorded steps, by the buffers they read and write
  std::map<std::array<const void*, 5>, hipGraphExec_t> rgexec;
  size_t rec_bytes() const {
    return (size_t)9 * nx * my * mz * sizeof(T) + sizeof(DevState) + GATHER_PAD * sizeof(T);
  }
  void rec_step() {
    if (use_graph && !ktime && eager_steps >= 2 && nlbc_valid && (!skip_drag || lite_eager >= 1)) {
      const std::array<const void*, 5> key{nl, ubcp, qin, ppr, skip_drag ? this : nullptr};
      hipGraphExec_t& slot = rgexec[key];
      if (!slot) {
        hipGraph_t g = capture_step();
        hipError_t e = hipGraphInstantiate(&slot, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) rgexec.erase(key);
        HIPCHK(e);
      }
      HIPCHK(hipGraphLaunch(slot, stream));
      advance_step();
      return;
    }
    eager_step(false);
    if (skip_drag) ++lite_eager;
    else ++eager_steps;
  }
  void adjoint_record(int nsteps) {
    need_adj();
    if (nbody == 0) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    if (nsteps <= 0) return;
    const size_t cs3 = (size_t)3 * nx * my * mz, fb = cs3 * sizeof(T);
    T *const ub0 = ubcp, *const qi0 = qin, *const pp0 = ppr;
    // back to the context's own buffers, holding the last step's fields
    auto back = [&](int done) {
      skip_drag = false;
      if (done > 0) {
        HIPCHK(hipMemcpyAsync(ub0, ubcp, fb, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(qi0, qin, fb, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(pp0, ppr, fb, hipMemcpyDeviceToDevice, stream));
      }
      ubcp = ub0;
      qin = qi0;
      ppr = pp0;
    };
    skip_drag = cfg.set_mpg && !cfg.move_particles;
    int done = 0;
    try {
      for (; done < nsteps; ++done) {
        if (rec_depth == (int)rec_blocks.size()) {
          void* m = nullptr;
          HIPCHK(hipMalloc(&m, rec_bytes()));
          rec_blocks.push_back(m);
        }
        T* b = (T*)rec_blocks[rec_depth];
        ubcp = b;
        qin = b + cs3;
        ppr = b + 2 * cs3;
        rec_step();
        HIPCHK(hipMemcpyAsync(b + 3 * cs3, dst, sizeof(DevState), hipMemcpyDeviceToDevice, stream));
        ++rec_depth;
      }
    } catch (...) {
      back(done);
      throw;
    }
    back(done);
  }
  // ---------------- autotuner ----------------
  // a step form: (PIB_TUNE_FFT, PIB_TUNE_VELSTEP, PIB_TUNE_STREAMS, PIB_TUNE_NL)
  std::array<int, 4> form() const {
    return {plane_hbm && !hbm_needed ? 2 : (fft_compiled ? 1 : 0), vel_variant, use_fork ? 1 : 0, nl_sep ? 1 : 0};
  }
  void set_form(const std::array<int, 4>& f) {
    HIPCHK(hipStreamSynchronize(stream));
    drop_graphs();
    fft_compiled = f[0] == 1;
    set_plane_hbm(f[0] == 2 || hbm_needed);
    vel_variant = f[1];
    use_fork = f[2] == 1;
    nl_sep = f[3] == 1;
  }
  // GPU ms per graph-replayed step of the current form from the current
  // state, which is restored afterwards
  double time_form(int nsteps) {
    checkpoint_push();
    for (int n = 0; n < 2; ++n) {  // kernel attributes, then both graphs captured
      eager_step(false);
      ++eager_steps;
    }
    for (int n = 0; n < 2; ++n) graph_step();
    HIPCHK(hipEventRecord(ev[0], stream));
    for (int n = 0; n < nsteps; ++n) graph_step();
    HIPCHK(hipEventRecord(ev[1], stream));
    HIPCHK(hipEventSynchronize(ev[1]));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    checkpoint_pop();
    timed = false;
    return ms / nsteps;
  }
  void autotune(int nsteps, bool force, int* picks, double* ms_out) {
    if (!have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    if (nsteps < 1) throw PibFail(PIB_ERR_ARG, "nsteps < 1");
    if (!use_graph) throw PibFail(PIB_ERR_STATE, "the autotuner times graph-replayed steps (PIB_TUNE_GRAPH = 1)");
    // the x-flux scheme too: the forms' costs differ between central and upwind
    const std::array<int, 6> key{nx, my, mz, (int)sizeof(T), (int)cfg.discr_scheme, (int)cfg.limiter};
    std::array<int, 4> pick = form();
    double best = 0;
    bool cached = false;
    {
      std::lock_guard<std::mutex> lk(tune_mutex());
      auto it = tune_cache().find(key);
      if (it != tune_cache().end() && !force) {
        pick = it->second;
        cached = true;
      }
    }
    if (!cached) {
      if (pick[0] == 2) pick[0] = 1;
      set_form(pick);
      best = time_form(nsteps);
      auto axis = [&](int a, std::initializer_list<int> vals) {
        for (int v : vals) {
          if (v == pick[a]) continue;
          auto f = pick;
          f[a] = v;
          set_form(f);
          double t = time_form(nsteps);
          if (t < best) {
            best = t;
            pick = f;
          }
        }
      };
      if (sh != 0 && !hbm_needed) axis(0, {0, 1});
      axis(1, {0, 1, 2});
      if (nbody > 0) axis(2, {0, 1});
      axis(3, {0, 1});
      std::lock_guard<std::mutex> lk(tune_mutex());
      tune_cache()[key] = pick;
    }
    set_form(pick);
    if (picks)
      for (int q = 0; q < 3; ++q) picks[q] = pick[q];
    if (ms_out) *ms_out = best;
  }
  static std::mutex& tune_mutex() {
    static std::mutex m;
    return m;
  }
  static std::map<std::array<int, 6>, std::array<int, 4>>& tune_cache() {
    static std::map<std::array<int, 6>, std::array<int, 4>> c;
    return c;
  }
  void checkpoint_pop() {
    if (ck_stack.empty()) throw PibFail(PIB_ERR_STATE, "checkpoint stack is empty");
    size_t off[5];
    ck_parts(off);
    auto top = ck_stack.back();
    ck_stack.pop_back();
    const char* b = (const char*)top.first;
    void* dstp[4] = {u, nl, p, dst};
    for (int q = 0; q < 4; ++q)
      HIPCHK(hipMemcpyAsync(dstp[q], b + off[q], part_bytes(q), hipMemcpyDeviceToDevice, stream));
    nlbc_valid = top.second;
    ++state_gen;
    ck_pool.push_back(top.first);  // the copies out of it precede any reuse on the stream
  }
};

struct Ctx {
  int prec;
  Base* impl;
  std::string err;
};

template <typename F>
int guard(Ctx* c, F f) {
  if (!c) return PIB_ERR_ARG;
  try {
    f();
    return PIB_OK;
  } catch (const PibFail& e) {
    c->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    c->err = e.what();
    return PIB_ERR_HIP;
  }
}

// Entry points that may run while the form_nl branch of pib_nonlinear_begin
// is still in flight on the second stream (it writes nl and qin only): the
// drag results, the device state, the particle update and the finish, which
// joins the branch itself before the forcing.
template <typename F>
int dispatch_keep_fork(void* ctx, F f) {
  Ctx* c = (Ctx*)ctx;
  return guard(c, [&] {
    if (c->prec == PIB_FP64) f(*static_cast<Impl<double>*>(c->impl));
    else f(*static_cast<Impl<float>*>(c->impl));
  });
}
// Every other entry point first makes the main stream wait for that branch.
template <typename F>
int dispatch(void* ctx, F f) {
  return dispatch_keep_fork(ctx, [&](auto& I) {
    I.join_branches();
    f(I);
  });
}

int finish_status(void* ctx, int st) {
  if (st != PIB_OK) return st;
  int dev = 0;
  int r = dispatch_keep_fork(ctx, [&](auto& I) { dev = I.check_status(); });
  if (r != PIB_OK) return r;
  if (dev) {
    Ctx* c = (Ctx*)ctx;
    c->err = dev == PIB_ERR_BLOWUP ? "norm_drag > 1e6 (blow-up)" : "zero pivot in tridiagonal solve";
    return dev;
  }
  return PIB_OK;
}

// Device address of `scratch` on each device, resolved once per device with
// that device current (at pib_create, or at the first no-alloc call).
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_dev[MAX_DEVICES] = {};
int scratch_setup(int dev, double** out, std::string& err) {
  if (dev < 0 || dev >= MAX_DEVICES) {
    err = "device ordinal out of range";
    return PIB_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_dev[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return PIB_ERR_HIP;
    }
    scratch_dev[dev] = (double*)ptr;
  }
  *out = scratch_dev[dev];
  return PIB_OK;
}

}  // namespace

extern "C" {

int pib_create(const pib_config* cfg, void** ctx) {
  if (!cfg || !ctx) return PIB_ERR_ARG;
  *ctx = nullptr;
  if (cfg->nx < 4 || cfg->my < 4 || cfg->mz < 4 || (cfg->my & 1) || (cfg->mz & 1)) return PIB_ERR_ARG;
  if (cfg->precision != PIB_FP64 && cfg->precision != PIB_FP32_BULK) return PIB_ERR_ARG;
  Ctx* c = new Ctx{cfg->precision, nullptr, ""};
  int st = guard(c, [&] {
    if (cfg->precision == PIB_FP64) c->impl = new Impl<double>(*cfg);
    else c->impl = new Impl<float>(*cfg);
  });
  if (st == PIB_OK) {  // the no-alloc entry point's workspace on this device
    int dev = 0;
    double* ptr = nullptr;
    st = hipGetDevice(&dev) == hipSuccess ? scratch_setup(dev, &ptr, c->err) : PIB_ERR_HIP;
    if (st != PIB_OK) delete c->impl;
  }
  if (st != PIB_OK) {
    fprintf(stderr, "pib_create: %s\n", c->err.c_str());
    delete c;
    return st;
  }
  *ctx = c;
  return PIB_OK;
}
int pib_destroy(void* ctx) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return PIB_OK;
  delete c->impl;
  delete c;
  return PIB_OK;
}
const char* pib_last_error(void* ctx) { return ctx ? ((Ctx*)ctx)->err.c_str() : "null context"; }
int pib_device_info(char* name, int len, int* cu_count) {
  hipDeviceProp_t prop;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return PIB_ERR_HIP;
  if (name && len > 0) {
    snprintf(name, len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  return PIB_OK;
}
int pib_set_points(void* ctx, int nbnd, const double* bndpts, const double* cd) {
  if (nbnd <= 0 || !bndpts || !cd) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) { I.set_points(nbnd, bndpts, cd); });
}
int pib_set_particles(void* ctx, int nbody, const double* xc, const double* radbdy, const double* velbdy,
                      const double* angv, long long* count_solid, int* nrpr_active) {
  if (nbody < 0 || (nbody > 0 && (!xc || !radbdy))) return PIB_ERR_ARG;
  return dispatch_keep_fork(ctx,
                            [&](auto& I) { I.set_particles(nbody, xc, radbdy, velbdy, angv, count_solid, nrpr_active); });
}
int pib_set_field(void* ctx, int field, const double* host) {
  if (!host) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) { I.set_field(field, host); });
}
int pib_get_field(void* ctx, int field, double* host) {
  if (!host) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) { I.get_field(field, host); });
}
int pib_scalar_set(void* ctx, double gamma, double heat_ratio, double heat_ratio_new) {
  return dispatch(ctx, [&](auto& I) { I.scalar_set(gamma, heat_ratio, heat_ratio_new); });
}
int pib_scalar_get(void* ctx, double* out) {
  if (!out) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) { I.scalar_get(out); });
}
int pib_scalar_ratios(void* ctx, const double* hr3) {
  return dispatch(ctx, [&](auto& I) { I.scalar_ratios(hr3); });
}
int pib_scalstep(void* ctx, int rks) {
  return dispatch(ctx, [&](auto& I) { I.scalstep(rks); });
}
int pib_scalar_surface(void* ctx, const double* phisurfall) {
  return dispatch(ctx, [&](auto& I) { I.scalar_surface(phisurfall); });
}
int pib_scalar_flux(void* ctx, double* flux_body, double* out4) {
  return dispatch(ctx, [&](auto& I) { I.scalar_flux(flux_body, out4); });
}
int pib_scalstep_modes(void* ctx, int rks) {
  return dispatch(ctx, [&](auto& I) { I.scalstep_modes(rks); });
}
int pib_set_scalars(void* ctx, const pib_scalars* s) {
  if (!s) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) { I.set_scalars(s); });
}
int pib_get_scalars(void* ctx, pib_scalars* s) {
  if (!s) return PIB_ERR_ARG;
  return dispatch_keep_fork(ctx, [&](auto& I) { I.get_scalars(s); });
}
int pib_get_ferror_history(void* ctx, int n, double* ferror_array, double* ferror_hist) {
  return dispatch_keep_fork(ctx, [&](auto& I) { I.get_ferror_history(n, ferror_array, ferror_hist); });
}
int pib_set_ferror_history(void* ctx, int n, const double* ferror_array) {
  return dispatch(ctx, [&](auto& I) { I.set_ferror_history(n, ferror_array); });
}
int pib_get_forces(void* ctx, double* force, double* pres, double* visc, double* torq) {
  return dispatch_keep_fork(ctx, [&](auto& I) { I.get_forces(force, pres, visc, torq); });
}
int pib_compute_new_timestep(void* ctx, int rks) {
  int st = dispatch(ctx, [&](auto& I) {
    if (rks < 1 || rks > I.itrmax) throw PibFail(PIB_ERR_ARG, "rks out of range");
    if (!I.have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    I.compute_new_timestep(rks);
  });
  return finish_status(ctx, st);
}
int pib_nonlinear(void* ctx, int rks) {
  int st = dispatch(ctx, [&](auto& I) {
    if (rks < 1 || rks > I.itrmax) throw PibFail(PIB_ERR_ARG, "rks out of range");
    if (!I.have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    I.nonlinear(rks);
  });
  return finish_status(ctx, st);
}
int pib_nonlinear_begin(void* ctx, int rks) {
  int st = dispatch(ctx, [&](auto& I) {
    if (rks < 1 || rks > I.itrmax) throw PibFail(PIB_ERR_ARG, "rks out of range");
    if (!I.have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    I.nonlinear_timed(rks, false, 1);
    I.nl_pending = rks;
  });
  return finish_status(ctx, st);
}
int pib_nonlinear_finish(void* ctx, int rks) {
  int st = dispatch_keep_fork(ctx, [&](auto& I) {
    if (I.nl_pending != rks) throw PibFail(PIB_ERR_STATE, "pib_nonlinear_begin with this rks first");
    if (!I.have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    I.nonlinear_timed(rks, false, 2);
    I.nl_pending = 0;
  });
  return finish_status(ctx, st);
}
int pib_velstep(void* ctx, int rks) {
  int st = dispatch(ctx, [&](auto& I) {
    if (rks < 1 || rks > I.itrmax) throw PibFail(PIB_ERR_ARG, "rks out of range");
    if (!I.have_particles) throw PibFail(PIB_ERR_STATE, "pib_set_particles first");
    I.velstep(rks);
  });
  return finish_status(ctx, st);
}
int pib_adjoint_begin(void* ctx) {
  return dispatch(ctx, [&](auto& I) { I.adjoint_begin(); });
}
int pib_adjoint_step(void* ctx, const double* wforce) {
  return dispatch(ctx, [&](auto& I) { I.adjoint_step(wforce); });
}
int pib_adjoint_get(void* ctx, double* dxc, double* dmpg, double* dumean) {
  int st = dispatch(ctx, [&](auto& I) { I.adjoint_get(dxc, dmpg, dumean); });
  return finish_status(ctx, st);
}
int pib_adjoint_set_umean(void* ctx, const double* dumean) {
  if (!dumean) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) { I.adjoint_set_umean(dumean); });
}
int pib_checkpoint_push(void* ctx) {
  return dispatch(ctx, [&](auto& I) { I.checkpoint_push(); });
}
int pib_checkpoint_pop(void* ctx) {
  return dispatch(ctx, [&](auto& I) { I.checkpoint_pop(); });
}
int pib_adjoint_record(void* ctx, int nsteps) {
  if (nsteps < 0) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) { I.adjoint_record(nsteps); });
}
int pib_checkpoint_steps(void* ctx, int nsteps) {
  if (nsteps < 0) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) { I.checkpoint_steps(nsteps); });
}
int pib_step_async(void* ctx, int nsteps) {
  if (nsteps < 0) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) { I.step_async(nsteps); });
}
int pib_sync(void* ctx) {
  // a form_nl branch left in flight by pib_nonlinear_begin is drained too
  return finish_status(ctx, dispatch(ctx, [&](auto& I) { (void)I; }));
}
int pib_step(void* ctx, int nsteps) { return finish_status(ctx, pib_step_async(ctx, nsteps)); }
int pib_set_tuning(void* ctx, int key, int value) {
  if (key == PIB_TUNE_VELSTEP) {
    if (value < 0 || value > 2) return PIB_ERR_ARG;
    return dispatch(ctx, [&](auto& I) {
      HIPCHK(hipStreamSynchronize(I.stream));
      I.drop_graphs();
      I.vel_variant = value;
    });
  }
  if (key == PIB_TUNE_FFT) {
    if (value < 0 || value > 2) return PIB_ERR_ARG;
    return dispatch(ctx, [&](auto& I) {
      HIPCHK(hipStreamSynchronize(I.stream));
      I.drop_graphs();
      I.fft_compiled = value == 1;
      I.set_plane_hbm(value == 2 || I.hbm_needed);
    });
  }
  if (key == PIB_TUNE_STREAMS) {
    if (value < 0 || value > 1) return PIB_ERR_ARG;
    return dispatch(ctx, [&](auto& I) {
      HIPCHK(hipStreamSynchronize(I.stream));
      I.drop_graphs();
      I.use_fork = value == 1;
    });
  }
  if (key == PIB_TUNE_NL) {
    if (value < 0 || value > 1) return PIB_ERR_ARG;
    return dispatch(ctx, [&](auto& I) {
      HIPCHK(hipStreamSynchronize(I.stream));
      I.drop_graphs();
      I.nl_sep = value == 1;
    });
  }
  if (key == PIB_TUNE_GRAPH) {
    if (value < 0 || value > 1) return PIB_ERR_ARG;
    return dispatch(ctx, [&](auto& I) {
      HIPCHK(hipStreamSynchronize(I.stream));
      I.drop_graphs();
      I.use_graph = value == 1;
    });
  }
  return PIB_ERR_ARG;
}
int pib_get_tuning(void* ctx, int key, int* value) {
  if (!value) return PIB_ERR_ARG;
  if (key != PIB_TUNE_VELSTEP && key != PIB_TUNE_GRAPH && key != PIB_TUNE_FFT && key != PIB_TUNE_STREAMS &&
      key != PIB_TUNE_NL)
    return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) {
    const auto f = I.form();
    *value = key == PIB_TUNE_FFT ? f[0]
             : key == PIB_TUNE_VELSTEP ? f[1]
             : key == PIB_TUNE_STREAMS ? f[2]
             : key == PIB_TUNE_NL ? f[3]
                                  : (I.use_graph ? 1 : 0);
  });
}
int pib_autotune(void* ctx, int nsteps, int force, int* picks, double* ms) {
  return dispatch(ctx, [&](auto& I) { I.autotune(nsteps, force != 0, picks, ms); });
}

// ---------------- batched realizations ----------------
// Each batch step is one launch of a graph captured across the batch's and
// the contexts' streams: a fork from the batch stream into every context's
// stream, each context's whole step on its own stream, and a join back, so
// the contexts' steps are independent branches of the one graph.
struct Batch {
  std::vector<Ctx*> ctx;
  hipStream_t stream = nullptr;
  std::vector<hipEvent_t> ev;  // a context's stream -> the batch stream
  hipEvent_t ev_done = nullptr;  // the batch stream -> every context's stream
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<int> key[2];  // nl/onl assignment of every context
  std::vector<unsigned long long> gen[2];
  bool joined = false;
  std::string err;
  ~Batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& e : exec)
      if (e) (void)hipGraphExecDestroy(e);
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    if (ev_done) (void)hipEventDestroy(ev_done);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void join() {  // the batch stream waits for the contexts' earlier work
    for (size_t i = 0; i < ctx.size(); ++i) {
      HIPCHK(hipEventRecord(ev[i], ctx[i]->impl->main_stream()));
      HIPCHK(hipStreamWaitEvent(stream, ev[i], 0));
    }
    joined = true;
  }
  void release() {  // the contexts' later work waits for the batch's
    HIPCHK(hipEventRecord(ev_done, stream));
    for (auto c : ctx) HIPCHK(hipStreamWaitEvent(c->impl->main_stream(), ev_done, 0));
    joined = false;
  }
  void build(int slot, const std::vector<int>& k, const std::vector<unsigned long long>& g) {
    if (exec[slot]) {
      HIPCHK(hipStreamSynchronize(stream));
      (void)hipGraphExecDestroy(exec[slot]);
      exec[slot] = nullptr;
    }
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      HIPCHK(hipEventRecord(ev_done, stream));
      for (size_t i = 0; i < ctx.size(); ++i) {
        hipStream_t cs = ctx[i]->impl->main_stream();
        HIPCHK(hipStreamWaitEvent(cs, ev_done, 0));
        ctx[i]->impl->step_into_capture();
        HIPCHK(hipEventRecord(ev[i], cs));
        HIPCHK(hipStreamWaitEvent(stream, ev[i], 0));
      }
    } catch (...) {
      (void)hipStreamEndCapture(stream, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    HIPCHK(hipStreamEndCapture(stream, &graph));
    hipError_t e = hipGraphInstantiate(&exec[slot], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    key[slot] = k;
    gen[slot] = g;
  }
  void step(int nsteps) {
    const size_t n = ctx.size();
    std::vector<int> k(n);
    std::vector<unsigned long long> g(n);
    for (int s = 0; s < nsteps; ++s) {
      bool ready = true;
      for (auto c : ctx) ready = ready && c->impl->graph_ready();
      if (!ready) {  // the eager steps that precede any capture, per context
        if (joined) release();
        for (auto c : ctx) c->impl->batch_eager_step();
        continue;
      }
      if (!joined) join();
      for (size_t i = 0; i < n; ++i) {
        k[i] = ctx[i]->impl->graph_parity();
        g[i] = ctx[i]->impl->graph_generation();
      }
      int slot = -1;
      for (int q = 0; q < 2; ++q)
        if (exec[q] && key[q] == k && gen[q] == g) slot = q;
      if (slot < 0) {  // keep the other assignment's graph while it is current
        slot = (exec[0] && gen[0] == g) ? 1 : 0;
        build(slot, k, g);
      }
      HIPCHK(hipGraphLaunch(exec[slot], stream));
      for (auto c : ctx) c->impl->advance_step();
    }
    if (joined) release();
  }
};

int pib_batch_create(void* const* ctxs, int n, void** batch) {
  if (!ctxs || n < 1 || !batch) return PIB_ERR_ARG;
  *batch = nullptr;
  Batch* b = new Batch;
  for (int i = 0; i < n; ++i) {
    Ctx* c = (Ctx*)ctxs[i];
    if (!c || !c->impl) {
      delete b;
      return PIB_ERR_ARG;
    }
    for (auto o : b->ctx)
      if (o == c) {
        delete b;
        return PIB_ERR_ARG;
      }
    b->ctx.push_back(c);
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming) != hipSuccess) {
    delete b;
    return PIB_ERR_HIP;
  }
  b->ev.assign(n, nullptr);
  for (auto& e : b->ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      delete b;
      return PIB_ERR_HIP;
    }
  *batch = b;
  return PIB_OK;
}
int pib_batch_step(void* batch, int nsteps) {
  Batch* b = (Batch*)batch;
  if (!b || nsteps < 0) return PIB_ERR_ARG;
  try {
    b->step(nsteps);
    return PIB_OK;
  } catch (const PibFail& e) {
    b->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    b->err = e.what();
    return PIB_ERR_HIP;
  }
}
int pib_batch_sync(void* batch) {
  Batch* b = (Batch*)batch;
  if (!b) return PIB_ERR_ARG;
  if (hipStreamSynchronize(b->stream) != hipSuccess) {
    b->err = "hipStreamSynchronize failed";
    return PIB_ERR_HIP;
  }
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    int st = finish_status(b->ctx[i], PIB_OK);
    if (st != PIB_OK) {
      b->err = "context " + std::to_string(i) + ": " + b->ctx[i]->err;
      return st;
    }
  }
  return PIB_OK;
}
int pib_batch_destroy(void* batch) {
  delete (Batch*)batch;
  return PIB_OK;
}
const char* pib_batch_last_error(void* batch) {
  return batch ? ((Batch*)batch)->err.c_str() : "null batch";
}

int pib_set_timing(void* ctx, int on) {
  return dispatch(ctx, [&](auto& I) { I.set_timing(on); });
}
int pib_kernel_times(void* ctx, int maxn, char* names, double* total_ms, int* launches, int* count) {
  if (!names || !total_ms || !launches || !count || maxn < 0) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) { *count = I.kernel_times(maxn, names, total_ms, launches); });
}
int pib_phase_times(void* ctx, double* ms4) {
  if (!ms4) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) { I.phase_times(ms4); });
}
int pib_ff2rc_planes(void* ctx, const double* a, long long lda, int nplanes, double* b, long long ldb) {
  if (!a || !b || nplanes <= 0) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) {
    if (lda < (long long)I.my * I.mz || ldb < (long long)I.my2 * I.mz) throw PibFail(PIB_ERR_ARG, "leading dim");
    I.ff2rc_planes(a, lda, nplanes, b, ldb);
  });
}
int pib_ff2cr_planes(void* ctx, const double* b, long long ldb, int nplanes, double* a, long long lda) {
  if (!a || !b || nplanes <= 0) return PIB_ERR_ARG;
  return dispatch(ctx, [&](auto& I) {
    if (lda < (long long)I.my * I.mz || ldb < (long long)I.my2 * I.mz) throw PibFail(PIB_ERR_ARG, "leading dim");
    I.ff2cr_planes(b, ldb, nplanes, a, lda);
  });
}
int pib_tridiag_batched(void* ctx, int n, int nsys, const double* a, const double* b, const double* c,
                        const double* r, double* x, int cyclic) {
  if (n < 3 || nsys < 1 || !a || !b || !c || !r || !x) return PIB_ERR_ARG;
  Ctx* cc = (Ctx*)ctx;
  return guard(cc, [&] {
    size_t N = (size_t)n * nsys;
    double *d[6], *scr;
    for (auto& q : d) HIPCHK(hipMalloc(&q, N * sizeof(double)));
    HIPCHK(hipMalloc(&scr, 4 * N * sizeof(double)));
    HIPCHK(hipMemcpy(d[0], a, N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d[1], b, N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d[2], c, N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d[3], r, N * sizeof(double), hipMemcpyHostToDevice));
    k_tridiag<<<(nsys + 255) / 256, 256>>>(n, nsys, d[0], d[1], d[2], d[3], d[4], cyclic, scr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(x, d[4], N * sizeof(double), hipMemcpyDeviceToHost));
    for (auto& q : d) (void)hipFree(q);
    (void)hipFree(scr);
  });
}
int pib_tridiag_batched_noalloc(void* ctx, int n, int nsys, const double* d_a, const double* d_b,
                                const double* d_c, const double* d_r, double* d_x, int cyclic, void* stream) {
  Ctx* cc = (Ctx*)ctx;
  std::string err;
  int st = PIB_OK;
  if (n < 3 || nsys < 1 || !d_a || !d_b || !d_c || !d_r || !d_x) {
    err = "n >= 3, nsys >= 1 and five device pointers";
    st = PIB_ERR_ARG;
  } else if ((long long)n * nsys > (long long)PIB_TRIDIAG_NOALLOC_MAX) {
    err = "n*nsys exceeds PIB_TRIDIAG_NOALLOC_MAX (" + std::to_string(PIB_TRIDIAG_NOALLOC_MAX) + " rows)";
    st = PIB_ERR_ARG;
  } else {
    int dev = 0;
    double* scr = nullptr;
    if (hipGetDevice(&dev) != hipSuccess) {
      err = "hipGetDevice failed";
      st = PIB_ERR_HIP;
    } else {
      st = scratch_setup(dev, &scr, err);
    }
    if (st == PIB_OK) {
      // every scratch entry the kernel reads it has written first in the same
      // launch, so nothing is cleared
      k_tridiag<<<(nsys + 255) / 256, 256, 0, (hipStream_t)stream>>>(n, nsys, d_a, d_b, d_c, d_r, d_x, cyclic, scr);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        err = std::string("k_tridiag launch: ") + hipGetErrorString(e);
        st = PIB_ERR_HIP;
      }
    }
  }
  if (st != PIB_OK && cc) cc->err = err;
  return st;
}

}  // extern "C"
