// pib_adjoint.hip -- kernels of the adjoint of the fixed-bed step (no
// counterpart in the reference, which has no adjoint).  Each kernel is the
// transpose of one forward piece of pib_kernels.h / pib_velstep.h, applied in
// the reverse order of the step.  The transposed line solves run one lane
// per (ky,kz) mode in double; the transposed plane transforms reuse the plane
// FFT engine with their own load/store functors; the transposed IBM
// interpolations scatter with atomic adds, so their sums (and the gradient
// sums) are not bit-reproducible from run to run, unlike the forward step.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "pib_adjoint.h"
#include "pib_velstep.h"

namespace {

#define ADJCHK(x)                                                                           \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

__device__ __forceinline__ C2<double> cd(double x, double y) { return mk(x, y); }
template <typename T>
__device__ __forceinline__ C2<double> d2(C2<T> v) { return mk((double)v.x, (double)v.y); }
template <typename T>
__device__ __forceinline__ C2<T> t2(C2<double> v) { return mk((T)v.x, (T)v.y); }
// (i k) z
__device__ __forceinline__ C2<double> ik(double k, C2<double> z) { return mk(-k * z.y, k * z.x); }
__device__ __forceinline__ C2<double> sc(double s, C2<double> z) { return mk(s * z.x, s * z.y); }
__device__ __forceinline__ int wrp(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
// weight of column ky in the c2r sum: 1 at ky = 0 and my/2, else 2
template <typename T>
__device__ __forceinline__ double ky_weight(const KArgs<T>& A, int ky) { return (ky == 0 || ky == A.my2 - 1) ? 1.0 : 2.0; }

// up to six buffers zeroed in one launch (grid.y = buffer); a fill kernel,
// so that it may be captured in a graph
template <typename T>
struct ZeroSet {
  T* p[6];
  size_t n[6];
  int k;
};
template <typename T>
__global__ void k_adj_zero_set(ZeroSet<T> z) {
  const int b = blockIdx.y;
  T* p = z.p[b];
  const size_t n = z.n[b];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = (T)0;
}

// ===========================================================================
// velstep (steptheflow.F90:52-160), transposed per (ky,kz) mode.  Forward:
//   pressure  p~ = M^-1 [c3 D(nl) + D(ff) + c4 D(onl)]/c12
//   ustep     u*_d = H^-1 [(1 - dt nu c1 w2) u_d + dt nu c1 L(u_d)/dx2
//                         + dt (c3 nl_d + ff_d + c4 onl_d) - dt c12 G_d(p~)]
//   divcorr   phi = M^-1 D(u*)/(c12 dt);  u_d = u*_d - c12 dt G_d(phi);
//             p = p~ + phi - c2 dt nu (L(phi)/dx2 - w2 phi)
// with the staggered divergence D(v)_i = (v1_{i+1} - v1_i)/dx
// + (i wy)(v2_{i+1} + v2_i)/2 + (i wz)(v3_{i+1} + v3_i)/2, the gradient
// G_1(p)_i = (p_i - p_{i-1})/dx, G_{2,3}(p)_i = (i wy|wz)(p_{i-1} + p_i)/2,
// L the periodic second difference; M = tridiag(1/dx2 - w2/4,
// -2/dx2 - w2/2) cyclic ((0,0): rows 1..nx-1, re only, p(nx) = 0),
// H = tridiag(-dt nu c2/dx2, 1 + dt nu c2 (2/dx2 + w2)) cyclic.
// Scratch lines (mode-major, row r of line k at vs[(k*nx + r)*nm + q]):
// 0..5 re/im of u*_bar per component, 6,7 p~_bar, 8,9 work, 10 the Thomas
// multipliers, 11 the Sherman-Morrison vector.
// ===========================================================================
struct Lines {
  double* vs;
  int nx, nm, q;
  __device__ double& operator()(int k, int r) const { return vs[((size_t)k * nx + r) * nm + q]; }
};
// in place: A^-1 x for the symmetric cyclic Toeplitz matrix with a on the
// off-diagonals and corners and b on the diagonal (Sherman-Morrison with
// u = (gamma, 0, .., 0, a), v = (1, 0, .., 0, a/gamma), gamma = -b)
__device__ void cyc_solve(const Lines& V, int kr, int ki, double a, double b) {
  const int n = V.nx;
  const double gam = -b;
  double bet = b - gam;
  double yr = V(kr, 0) / bet, yi = V(ki, 0) / bet, z = gam / bet;
  V(kr, 0) = yr;
  V(ki, 0) = yi;
  V(11, 0) = z;
  for (int r = 1; r < n; ++r) {
    const double g = a / bet;
    const double d = (r == n - 1) ? b - a * a / gam : b;
    bet = d - a * g;
    V(10, r) = g;
    yr = (V(kr, r) - a * yr) / bet;
    yi = (V(ki, r) - a * yi) / bet;
    z = ((r == n - 1 ? a : 0.0) - a * z) / bet;
    V(kr, r) = yr;
    V(ki, r) = yi;
    V(11, r) = z;
  }
  for (int r = n - 2; r >= 0; --r) {
    const double g = V(10, r + 1);
    yr = V(kr, r) - g * yr;
    yi = V(ki, r) - g * yi;
    z = V(11, r) - g * z;
    V(kr, r) = yr;
    V(ki, r) = yi;
    V(11, r) = z;
  }
  const double f = a / gam;
  const double vz = V(11, 0) + f * V(11, n - 1);
  const double fr = (V(kr, 0) + f * V(kr, n - 1)) / (1.0 + vz);
  const double fi = (V(ki, 0) + f * V(ki, n - 1)) / (1.0 + vz);
  for (int r = 0; r < n; ++r) {
    const double zr = V(11, r);
    V(kr, r) -= fr * zr;
    V(ki, r) -= fi * zr;
  }
}
// the (0,0) mode of the pressure matrix: rows 0..n-2 of re, row n-1 and im zero
__device__ void tri0_solve(const Lines& V, int kr, int ki, double a, double b) {
  const int n = V.nx - 1;
  double bet = b;
  double y = V(kr, 0) / bet;
  V(kr, 0) = y;
  for (int r = 1; r < n; ++r) {
    const double g = a / bet;
    bet = b - a * g;
    V(10, r) = g;
    y = (V(kr, r) - a * y) / bet;
    V(kr, r) = y;
  }
  for (int r = n - 2; r >= 0; --r) {
    y = V(kr, r) - V(10, r + 1) * y;
    V(kr, r) = y;
  }
  V(kr, n) = 0.0;
  for (int r = 0; r <= n; ++r) V(ki, r) = 0.0;
}

template <typename T>
__global__ void __launch_bounds__(64) k_adj_velstep(KArgs<T> A, AdjArgs<T> J, int nm) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nm) return;
  const int ky = q % A.my2, kz = q / A.my2, nx = A.nx;
  const size_t E = (size_t)nm;
  const DevState* st = A.st;
  const double dx = st->dx, dx2 = dx * dx, vis = st->vis, dt = st->s.dt;
  const double w2 = (double)A.w2[q], wy = (double)A.wy[ky], wz = (double)A.wz[kz];
  const double c1 = st->coef[0][0], c2 = st->coef[0][1], c3 = st->coef[0][2], c4 = st->coef[0][3];
  const double c12 = c1 + c2, alp = c12 * dt, cv = c2 * dt * vis;
  const double ma = 1.0 / dx2 - 0.25 * w2, mb = -2.0 / dx2 - 0.5 * w2;
  const double ha = -dt * vis * c2 / dx2, hb = 1.0 + dt * vis * c2 * (2.0 / dx2 + w2);
  const bool zero = q == 0;
  const Lines V{J.vs, nx, nm, q};
  auto at = [&](int c, int i) { return ((size_t)c * nx + i) * E + q; };
  auto msolve = [&](int kr, int ki) {
    if (zero) tri0_solve(V, kr, ki, ma, mb);
    else cyc_solve(V, kr, ki, ma, mb);
  };
  // u*_bar = u_bar; p~_bar = p_bar
  for (int i = 0; i < nx; ++i) {
    for (int d = 0; d < 3; ++d) {
      C2<double> v = d2(J.au[at(d, i)]);
      V(2 * d, i) = v.x;
      V(2 * d + 1, i) = v.y;
    }
    C2<double> p = d2(J.ap[at(0, i)]);
    V(6, i) = p.x;
    V(7, i) = p.y;
  }
  // divcorr^T: phi_bar = p_bar - cv (L(p_bar)/dx2 - w2 p_bar) - alp G^T(u_bar)
  for (int i = 0; i < nx; ++i) {
    const int im = wrp(i - 1, nx), ip = wrp(i + 1, nx);
    C2<double> pm = cd(V(6, im), V(7, im)), pc = cd(V(6, i), V(7, i)), pp = cd(V(6, ip), V(7, ip));
    C2<double> lap = (pm - sc(2.0, pc)) + pp;
    C2<double> ph = pc - sc(cv, sc(1.0 / dx2, lap) - sc(w2, pc));
    C2<double> g = sc(1.0 / dx, cd(V(0, i) - V(0, ip), V(1, i) - V(1, ip)));
    g = g + sc(0.5, ik(-wy, cd(V(2, i) + V(2, ip), V(3, i) + V(3, ip))));
    g = g + sc(0.5, ik(-wz, cd(V(4, i) + V(4, ip), V(5, i) + V(5, ip))));
    ph = ph - sc(alp, g);
    V(8, i) = ph.x;
    V(9, i) = ph.y;
  }
  msolve(8, 9);
  // u*_bar += D^T(dv_bar/alp)
  double sr_prev = V(8, nx - 1) / alp, si_prev = V(9, nx - 1) / alp;
  for (int i = 0; i < nx; ++i) {
    const double sr = V(8, i) / alp, si = V(9, i) / alp;
    C2<double> sm = cd(sr_prev, si_prev), s = cd(sr, si);
    C2<double> g1 = ik(-wy, sc(0.5, sm + s)), g2 = ik(-wz, sc(0.5, sm + s));
    V(0, i) += (sm.x - s.x) / dx;
    V(1, i) += (sm.y - s.y) / dx;
    V(2, i) += g1.x;
    V(3, i) += g1.y;
    V(4, i) += g2.x;
    V(5, i) += g2.y;
    sr_prev = sr;
    si_prev = si;
  }
  // ustep^T, one component at a time
  const double e1 = 1.0 - dt * vis * c1 * w2, e2 = dt * vis * c1 / dx2;
  for (int d = 0; d < 3; ++d) {
    cyc_solve(V, 2 * d, 2 * d + 1, ha, hb);
    for (int i = 0; i < nx; ++i) {
      const int im = wrp(i - 1, nx), ip = wrp(i + 1, nx);
      C2<double> qm = cd(V(2 * d, im), V(2 * d + 1, im)), qc = cd(V(2 * d, i), V(2 * d + 1, i)),
                 qp = cd(V(2 * d, ip), V(2 * d + 1, ip));
      C2<double> un = sc(e1, qc) + sc(e2, (qm - sc(2.0, qc)) + qp);
      J.aunew[at(d, i)] = t2<T>(un);
      const size_t ix = at(d, i);
      J.anl[ix] = t2<T>(d2(J.anl[ix]) + sc(dt * c3, qc));
      J.aff[ix] = t2<T>(sc(dt, qc));
      J.anlp[ix] = t2<T>(sc(dt * c4, qc));
      C2<double> g;
      if (d == 0) g = sc(1.0 / dx, qc - qp);
      else g = sc(0.5, ik(d == 1 ? -wy : -wz, qc + qp));
      V(6, i) -= dt * c12 * g.x;
      V(7, i) -= dt * c12 * g.y;
    }
  }
  // pressure^T
  msolve(6, 7);
  double sr_prev2 = V(6, nx - 1) / c12, si_prev2 = V(7, nx - 1) / c12;
  for (int i = 0; i < nx; ++i) {
    const double sr = V(6, i) / c12, si = V(7, i) / c12;
    C2<double> sm = cd(sr_prev2, si_prev2), s = cd(sr, si);
    C2<double> g[3] = {sc(1.0 / dx, sm - s), sc(0.5, ik(-wy, sm + s)), sc(0.5, ik(-wz, sm + s))};
    for (int d = 0; d < 3; ++d) {
      const size_t ix = at(d, i);
      J.anl[ix] = t2<T>(d2(J.anl[ix]) + sc(c3, g[d]));
      J.aff[ix] = t2<T>(d2(J.aff[ix]) + g[d]);
      J.anlp[ix] = t2<T>(d2(J.anlp[ix]) + sc(c4, g[d]));
    }
    sr_prev2 = sr;
    si_prev2 = si;
  }
}

// The same transposes on the tiles of the forward velstep (pib_velstep.h):
// M modes per workgroup, L x rows per lane in registers, every line solved
// by the partition method with the forward's symmetric matrices (their own
// transposes).  The (0,0) mode runs in double in the last workgroup (L = 2,
// one mode); its pressure/divcorr solves keep rows 1..nx-1 of the real part,
// which is the transpose of the forward's pinned solve.
// the last row of the previous segment for NL lines, through LDS
template <int NL, typename T, typename R, int L, int M>
__device__ __forceinline__ void prev_last(const VTile<T, R, L, M>& V, C2<R> (*x)[L], C2<R>* out) {
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NL; ++c) V.S.el[c * V.bd + V.tid] = x[c][L - 1];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NL; ++c) out[c] = V.S.el[c * V.bd + V.nbl(V.s - 1)];
}
template <typename T, typename R, int L, int M>
__device__ void vel_tile_adj(const KArgs<T>& A, const AdjArgs<T>& J, int q0, bool mean, unsigned char* raw) {
  VTile<T, R, L, M> V(A, q0, mean, raw);
  const DevState* st = A.st;
  const R c1 = (R)st->coef[0][0], c2 = (R)st->coef[0][1], c3 = (R)st->coef[0][2], c4 = (R)st->coef[0][3];
  const R dx = (R)st->dx, dx2 = dx * dx, dt = (R)st->s.dt, vis = (R)st->vis;
  const R idx = (R)1 / dx, idx2 = (R)1 / dx2;
  const R c12 = c1 + c2, alp = c12 * dt, cv = (c2 * dt) * vis;
  const int i0 = V.s * L;
  const R wy = V.wy, wz = V.wz, w2 = V.w2;
  const R ap = idx2 - (R)0.25 * w2, bp = (R)-2 * idx2 - (R)0.5 * w2;
  C2<R> u[3][L], pb[1][L];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int k = 0; k < L; ++k) u[c][k] = V.ld(J.au, c, i0 + k);
#pragma unroll
    for (int k = 0; k < L; ++k) pin(u[c][k]);
  }
#pragma unroll
  for (int k = 0; k < L; ++k) pb[0][k] = V.ld(J.ap, 0, i0 + k);
  // ---- divcorr^T: phi_bar = p_bar - cv (L(p_bar)/dx2 - w2 p_bar) - alp G^T(u_bar) ----
  C2<R> phi[1][L];
  {
    C2<R> un[3], pn[1], pp[1];
    V.template next_first<3>(u, un);
    V.template next_first<1>(pb, pn);
    prev_last<1>(V, pb, pp);
#pragma unroll
    for (int k = 0; k < L; ++k) {
      C2<R> pm = k > 0 ? pb[0][k - 1] : pp[0];
      C2<R> pq = k < L - 1 ? pb[0][k + 1] : pn[0];
      C2<R> n0 = k < L - 1 ? u[0][k + 1] : un[0];
      C2<R> n1 = k < L - 1 ? u[1][k + 1] : un[1];
      C2<R> n2 = k < L - 1 ? u[2][k + 1] : un[2];
      C2<R> lap = (pm - cscal((R)2, pb[0][k])) + pq;
      C2<R> ph = pb[0][k] - cscal(cv, cscal(idx2, lap) - cscal(w2, pb[0][k]));
      C2<R> g = (cscal(idx, u[0][k] - n0) + cik((R)-0.5 * wy, u[1][k] + n1)) + cik((R)-0.5 * wz, u[2][k] + n2);
      phi[0][k] = ph - cscal(alp, g);
      if (mean) phi[0][k].y = (R)0;
    }
  }
  {
    Toep<R, L> tp;
    toep_init(tp, ap, bp, V.st);
    V.template stage<1>(tp);
    V.template solve<1>(tp, phi, mean);
  }
  if (mean)
#pragma unroll
    for (int k = 0; k < L; ++k) phi[0][k].y = (R)0;
  // u*_bar = u_bar + D^T(dv_bar/alp)
  {
    C2<R> fp[1];
    prev_last<1>(V, phi, fp);
    const R ia = (R)1 / alp;
#pragma unroll
    for (int k = 0; k < L; ++k) {
      C2<R> sm = cscal(ia, k > 0 ? phi[0][k - 1] : fp[0]), sk = cscal(ia, phi[0][k]);
      u[0][k] = u[0][k] + cscal(idx, sm - sk);
      u[1][k] = u[1][k] + cik((R)-0.5 * wy, sm + sk);
      u[2][k] = u[2][k] + cik((R)-0.5 * wz, sm + sk);
    }
  }
  // ---- ustep^T ----
  {
    Toep<R, L> th;
    const R ah = -dt * vis * c2 * idx2;
    const R bh = (R)1 + dt * vis * c2 * ((R)2 * idx2 + w2);
    toep_init(th, ah, bh, V.st);
    V.template stage<3>(th);
    V.template solve<3>(th, u, false);
  }
  {
    C2<R> qn[3], qp[3];
    V.template next_first<3>(u, qn);
    prev_last<3>(V, u, qp);
    const R e1 = (R)1 - dt * vis * c1 * w2, e2 = dt * vis * c1 * idx2;
#pragma unroll
    for (int k = 0; k < L; ++k) {
      C2<R> g = mk((R)0, (R)0);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        C2<R> qm = k > 0 ? u[c][k - 1] : qp[c];
        C2<R> qq = k < L - 1 ? u[c][k + 1] : qn[c];
        C2<R> qc = u[c][k];
        V.stv(J.aunew, c, i0 + k, cscal(e1, qc) + cscal(e2, (qm - cscal((R)2, qc)) + qq));
        V.stv(J.anl, c, i0 + k, V.ld(J.anl, c, i0 + k) + cscal(dt * c3, qc));
        V.stv(J.aff, c, i0 + k, cscal(dt, qc));
        V.stv(J.anlp, c, i0 + k, cscal(dt * c4, qc));
        if (c == 0) g = g + cscal(idx, qc - qq);
        else g = g + cik((R)-0.5 * (c == 1 ? wy : wz), qc + qq);
      }
      pb[0][k] = pb[0][k] - cscal(dt * c12, g);
      if (mean) pb[0][k].y = (R)0;
    }
  }
  // ---- pressure^T ----
  {
    Toep<R, L> tp;
    toep_init(tp, ap, bp, V.st);
    V.template stage<1>(tp);
    V.template solve<1>(tp, pb, mean);
  }
  if (mean)
#pragma unroll
    for (int k = 0; k < L; ++k) pb[0][k].y = (R)0;
  {
    C2<R> fp[1];
    prev_last<1>(V, pb, fp);
    const R ic = (R)1 / c12;
#pragma unroll
    for (int k = 0; k < L; ++k) {
      C2<R> sm = cscal(ic, k > 0 ? pb[0][k - 1] : fp[0]), sk = cscal(ic, pb[0][k]);
      C2<R> g[3] = {cscal(idx, sm - sk), cik((R)-0.5 * wy, sm + sk), cik((R)-0.5 * wz, sm + sk)};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        V.stv(J.anl, c, i0 + k, V.ld(J.anl, c, i0 + k) + cscal(c3, g[c]));
        V.stv(J.aff, c, i0 + k, V.ld(J.aff, c, i0 + k) + g[c]);
        V.stv(J.anlp, c, i0 + k, V.ld(J.anlp, c, i0 + k) + cscal(c4, g[c]));
      }
    }
  }
}
template <typename T, int L, int M, int NT>
__global__ void __launch_bounds__(NT, NT <= 256 ? 3 : 1) k_adj_velstep_tiled(KArgs<T> A, AdjArgs<T> J, int ntiles) {
  extern __shared__ __align__(16) unsigned char vlds[];
  if ((int)blockIdx.x == ntiles) vel_tile_adj<T, double, 2, 1>(A, J, 0, true, vlds);
  else vel_tile_adj<T, T, L, M>(A, J, (int)blockIdx.x * M, false, vlds);
}

// umean += (-mpg + frmean) dt and ff = r2c(fr - frmean), transposed for the
// mean terms; with the force weights, the mpg term of
// force = c12 (pres + visc) - c12 mpg pi vol/6.  One workgroup.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_scal(KArgs<T> A, AdjArgs<T> J, int nbody, const double* rad) {
  __shared__ double sh[3 * 16];
  const DevState* st = A.st;
  const size_t E = (size_t)A.my2 * A.mz;
  double v[3] = {0, 0, 0};
  for (int i = threadIdx.x; i < A.nx; i += blockDim.x)
    for (int c = 0; c < 3; ++c) v[c] += (double)J.aff[((size_t)c * A.nx + i) * E].x;
  block_sums_w<3>(v, sh);
  double w[3] = {0, 0, 0};
  if (J.wforce)
    for (int m = threadIdx.x; m < nbody; m += blockDim.x) {
      const double vol = pow(2.0 * rad[m] * st->dx, 3.0);
      for (int c = 0; c < 3; ++c) w[c] += J.wforce[c * nbody + m] * vol;
    }
  block_sums_w<3>(w, sh);
  if (threadIdx.x == 0) {
    const double dt = st->s.dt, c12 = st->coef[0][0] + st->coef[0][1];
    const double N = (double)A.nx * A.my * A.mz;
    for (int c = 0; c < 3; ++c) {
      const double um = J.sc[c];
      J.sc[6 + c] = (um * dt - v[c]) / N;
      J.sc[3 + c] += -um * dt - c12 * w[c] * 3.14159265358979323846 / 6.0;
    }
  }
}

// ===========================================================================
// transposed plane transforms
// ===========================================================================
// r2c^T: a physical field from spectral data Y: c2r(Y/c_ky)/(my*mz)
enum { ADJ_AFR = 0, ADJ_NLQ = 1 };
template <typename T, int KIND>
struct OpR2CT {
  const KArgs<T>& A;
  const AdjArgs<T>& J;
  int c, i;
  __device__ C2<T> load(int ky, int kz) {
    const double s = 1.0 / (ky_weight(A, ky) * (double)A.my * (double)A.mz);
    const size_t E = (size_t)A.my2 * A.mz;
    auto g = [&](int cc) { return d2(J.anl[((size_t)cc * A.nx + i) * E + (size_t)kz * A.my2 + ky]); };
    C2<double> v;
    if constexpr (KIND == ADJ_AFR) {
      v = d2(J.aff[((size_t)c * A.nx + i) * E + (size_t)kz * A.my2 + ky]);
    } else {
      // the products of form_nl, transposed: nl = -[(i wy) F(P_c2) + (i wz) F(P_c3) + F(E_c)]
      const double wy = (double)A.wy[ky], wz = (double)A.wz[kz];
      switch (c) {
        case 0: v = ik(wy, g(0)); break;               // u1 u2
        case 1: v = ik(wz, g(0)); break;               // u1 u3
        case 2: v = ik(wy, g(1)); break;               // u2 u2
        case 3: v = ik(wz, g(1)) + ik(wy, g(2)); break; // u2 u3
        case 4: v = ik(wz, g(2)); break;               // u3 u3
        default: v = sc(-1.0, g(c - 5));               // x-flux derivative of component c-5
      }
    }
    return t2<T>(sc(s, v));
  }
  __device__ void store2(int, int y, int z, T a, T b) {
    T* dst = KIND == ADJ_AFR ? J.afr : J.Q;
    const size_t o = (((size_t)c * A.nx + i) * A.mz + z) * A.my + y;
    dst[o] = a;
    dst[o + 1] = b;
  }
};
template <typename T, int KIND, int SH>
__global__ void __launch_bounds__(1024, (PlaneOcc<T, SH>::WPE)) k_adj_r2c_t(KArgs<T> A, AdjArgs<T> J) {
  extern __shared__ __align__(16) unsigned char smem[];
  C2<T>* buf = SH == 0 ? plane_slot(A) : nullptr;
  OpR2CT<T, KIND> op{A, J, (int)blockIdx.y, plane_of(A)};
  plane_c2r_sh<SH, 0>(buf, A.G, A.tb, op, smem);
}

// c2r^T: spectral data from a physical field x: c_ky*my*mz*r2c(x), written
// or added to dst; the (0,0) entry, which is the sum of x over the plane, also
// goes to part[c*nx + i] when part is given.
template <typename T>
struct OpC2RT {
  const KArgs<T>& A;
  const T* src;
  C2<T>* dst;
  double* part;
  int c, i;
  bool add;
  __device__ void load2(int y, int z, T& a, T& b) {
    const size_t o = (((size_t)c * A.nx + i) * A.mz + z) * A.my + y;
    a = src[o];
    b = src[o + 1];
  }
  __device__ void store(int, int ky, int kz, C2<T> v) {
    const double s = ky_weight(A, ky) * (double)A.my * (double)A.mz;
    C2<double> w = sc(s, d2(v));
    const size_t o = (((size_t)c * A.nx + i) * A.mz + kz) * A.my2 + ky;
    dst[o] = add ? t2<T>(d2(dst[o]) + w) : t2<T>(w);
    if (part && ky == 0 && kz == 0) part[(size_t)c * A.nx + i] = w.x;
  }
};
template <typename T, int SH>
__global__ void __launch_bounds__(1024, (PlaneOcc<T, SH>::WPE)) k_adj_c2r_t(KArgs<T> A, const T* src, C2<T>* dst,
                                                                           double* part, int add) {
  extern __shared__ __align__(16) unsigned char smem[];
  C2<T>* buf = SH == 0 ? plane_slot(A) : nullptr;
  OpC2RT<T> op{A, src, dst, part, (int)blockIdx.y, plane_of(A), add != 0};
  plane_r2c_sh<SH, 0>(buf, A.G, A.tb, op, smem);
}

// sc[off + c] += sum over planes of part[c*nx + i] (one workgroup)
__global__ void __launch_bounds__(256) k_adj_sum_part(const double* part, int nx, double* sc, int off) {
  __shared__ double sh[3 * 16];
  double v[3] = {0, 0, 0};
  for (int i = threadIdx.x; i < nx; i += blockDim.x)
    for (int c = 0; c < 3; ++c) v[c] += part[(size_t)c * nx + i];
  block_sums_w<3>(v, sh);
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) sc[off + c] += v[c];
}

// ===========================================================================
// element-wise transposes on the spectral grid, (i, mode) per lane
// ===========================================================================
enum { ADJ_PGRAD = 0, ADJ_OMEGA = 1, ADJ_SFIELD = 2 };
template <typename T, int KIND>
__global__ void k_adj_modes(KArgs<T> A, AdjArgs<T> J) {
  const size_t E = (size_t)A.my2 * A.mz;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)A.nx * E) return;
  const int i = (int)(e / E);
  const size_t q = e - (size_t)i * E;
  const int ky = (int)(q % A.my2), kz = (int)(q / A.my2);
  const double wy = (double)A.wy[ky], wz = (double)A.wz[kz];
  auto Z = [&](int c, int ii) { return d2(J.Z[((size_t)c * A.nx + ii) * E + q]); };
  auto at = [&](int c) { return ((size_t)c * A.nx + i) * E + q; };
  const int im = wrp(i - 1, A.nx), ip = wrp(i + 1, A.nx);
  if constexpr (KIND == ADJ_PGRAD) {
    // calc_pgrad: ppr = c2r((p_{i+1} - p_{i-1})/2dx, (i wy) p, (i wz) p) + mpg
    const double dx = (double)A.dx;
    C2<double> v = sc(1.0 / (2.0 * dx), Z(0, im) - Z(0, ip)) + ik(-wy, Z(1, i)) + ik(-wz, Z(2, i));
    J.apnew[at(0)] = t2<T>(v);  // the first term of the step's adjoint of p
  } else if constexpr (KIND == ADJ_OMEGA) {
    // compute_omega: w1 = c2r((i wy) u3 - (i wz) u2), w2 = c2r((i wz) u1) - .., w3 = .. - c2r((i wy) u1)
    J.aunew[at(0)] = t2<T>(d2(J.aunew[at(0)]) + ik(wy, Z(2, i)) + ik(-wz, Z(1, i)));
    J.aunew[at(1)] = t2<T>(d2(J.aunew[at(1)]) + ik(wz, Z(0, i)));
    J.aunew[at(2)] = t2<T>(d2(J.aunew[at(2)]) + ik(-wy, Z(0, i)));
  } else {
    // Z = c2r^T(S): nlbc = c2r(nl) and onlbc = c2r(onl) with the adjoints
    // -c3 S, -c4 S; calc_visc, diffn = c2r(vis (L(u)/dx2 - w2 u)), with -c12 S
    const DevState* st = A.st;
    const double c3 = st->coef[0][2], c4 = st->coef[0][3], c12 = st->coef[0][0] + st->coef[0][1];
    const double dx2 = (double)A.dx * (double)A.dx, w2 = (double)A.w2[q], vis = (double)A.vis;
    for (int c = 0; c < 3; ++c) {
      const C2<double> zc = Z(c, i);
      C2<double> lap = (Z(c, im) - sc(2.0, zc)) + Z(c, ip);
      J.aunew[at(c)] = t2<T>(d2(J.aunew[at(c)]) + sc(-c12 * vis, sc(1.0 / dx2, lap) - sc(w2, zc)));
      J.anl[at(c)] = t2<T>(d2(J.anl[at(c)]) + sc(-c3, zc));
      J.anlp[at(c)] = t2<T>(d2(J.anlp[at(c)]) + sc(-c4, zc));
    }
  }
}

// ===========================================================================
// element-wise transposes on the physical grid, one node per lane
// ===========================================================================
// form_nl's products (central), the inner-point share cf S of the reversal
// forcing and, when the step's drag is weighed, compute_omega's
// x-differences of ubcp, transposed into aubcp
template <typename T>
__global__ void k_adj_phys(KArgs<T> A, AdjArgs<T> J, int drag) {
  const size_t P = (size_t)A.my * A.mz, cs = (size_t)A.nx * P;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= cs) return;
  const int i = (int)(e / P);
  const size_t n = e - (size_t)i * P;
  const int im = wrp(i - 1, A.nx), ip = wrp(i + 1, A.nx);
  const size_t om = (size_t)im * P + n, op = (size_t)ip * P + n;
  const double tdx = 2.0 * (double)A.dx;
  const double u1 = (double)A.ubcp[e], u2 = (double)A.ubcp[cs + e], u3 = (double)A.ubcp[2 * cs + e];
  auto Q = [&](int t, size_t o) { return (double)J.Q[(size_t)t * cs + o]; };
  const double F1 = (Q(5, om) - Q(5, op)) / tdx, F2 = (Q(6, om) - Q(6, op)) / tdx, F3 = (Q(7, om) - Q(7, op)) / tdx;
  const double q12 = Q(0, e), q13 = Q(1, e), q22 = Q(2, e), q23 = Q(3, e), q33 = Q(4, e);
  double a1 = q12 * u2 + q13 * u3 + 2.0 * F1 * u1 + F2 * u2 + F3 * u3;
  double a2 = q12 * u1 + 2.0 * q22 * u2 + q23 * u3 + F2 * u1;
  double a3 = q13 * u1 + q23 * u2 + 2.0 * q33 * u3 + F3 * u1;
  if (drag) {
    a2 += ((double)J.aom[2 * cs + om] - (double)J.aom[2 * cs + op]) / tdx;
    a3 -= ((double)J.aom[cs + om] - (double)J.aom[cs + op]) / tdx;
  }
  const double cf = A.st->s.cf;
  a1 += cf * (double)J.S[e];
  a2 += cf * (double)J.S[cs + e];
  a3 += cf * (double)J.S[2 * cs + e];
  J.aubcp[e] = (T)a1;  // the first terms of the step's adjoint of ubcp
  J.aubcp[cs + e] = (T)a2;
  J.aubcp[2 * cs + e] = (T)a3;
}

// ===========================================================================
// IBM, transposed
// ===========================================================================
// value of the trilinear interpolant of f, with coef * its position gradient
// added to g
template <typename T>
__device__ double interp_g(const Sten& s, const T* f, double coef, double g[3]) {
  double acc = 0.0;
  const double x1 = s.z[0], y1 = s.z[1], z1 = s.z[2];
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        const double v = (double)f[s.node(i, j, k)];
        const double X = i ? x1 : 1.0 - x1, Y = j ? y1 : 1.0 - y1, Z = k ? z1 : 1.0 - z1;
        acc += v * X * Y * Z;
        g[0] += coef * v * (i ? 1.0 : -1.0) * Y * Z;
        g[1] += coef * v * X * (j ? 1.0 : -1.0) * Z;
        g[2] += coef * v * X * Y * (k ? 1.0 : -1.0);
      }
  return acc;
}
template <typename T>
__device__ void scatter8(const Sten& s, T* f, double v) {
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) unsafeAtomicAdd(f + s.node(i, j, k), (T)(s.w(i, j, k) * v));
}
__device__ void add_grad(double* axc, int nbody, int m, double* g, double* sh) {
  block_sums_w<3>(g, sh);
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; ++d) unsafeAtomicAdd(axc + (size_t)d * nbody + m, g[d]);
}

// value of the trilinear interpolant of f; its position gradient in g
template <typename T>
__device__ __forceinline__ double interp_vg(const Sten& s, const T* f, double g[3]) {
  double acc = 0.0;
  g[0] = g[1] = g[2] = 0.0;
  const double x1 = s.z[0], y1 = s.z[1], z1 = s.z[2];
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const double v = (double)f[s.node(i, j, k)];
        const double X = i ? x1 : 1.0 - x1, Y = j ? y1 : 1.0 - y1, Z = k ? z1 : 1.0 - z1;
        acc += v * X * Y * Z;
        g[0] += v * (i ? 1.0 : -1.0) * Y * Z;
        g[1] += v * X * (j ? 1.0 : -1.0) * Z;
        g[2] += v * X * Y * (k ? 1.0 : -1.0);
      }
  return acc;
}

// calc_inner_reversal_forcing (boundary_condition.F90:1104-1479) transposed:
// per active reversal pair, one pass of gathers gives force_tmp and the
// position gradients of its interpolations; the adjoint of force_tmp comes
// from the spread nodes (solid share, the force_dist redistribution,
// frmean); the position gradient of the spreading weights; and the scatters:
// h at the inner point into S, cf kappa P h at the outer point into aubcp,
// c12 h at the pressure point into appr.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_forcing_points(KArgs<T> A, Particles P, AdjArgs<T> J, int G) {
  __shared__ double sh[3 * 16];
  const int w = xcd_linear(blockIdx.x, gridDim.x);
  const int m = w / G, gi = w - m * G;
  if (m >= P.nbody) return;
  const int l = gi * blockDim.x + threadIdx.x;
  double gx[3] = {0, 0, 0};
  if (l < P.nbnd && P.if_rev[(size_t)m * P.nbnd + l]) {
    const DevState* st = A.st;
    const double c12 = st->coef[0][0] + st->coef[0][1];
    const double cf = st->s.cf, dx3 = st->dx3;
    const double rb = P.rad[m], ri = P.radi[m], ro = P.rado[m];
    double xli[3], xlo[3], rad = 0.0;
    for (int n = 0; n < 3; ++n) {
      const double b = P.bnd[l * 3 + n], xcn = P.xc[n * P.nbody + m];
      xli[n] = xcn + b * ri;
      xlo[n] = xcn + b * ro;
      rad = rad + (b * ri) * (b * ri);
    }
    rad = sqrt(rad);
    double nv[3];
    for (int n = 0; n < 3; ++n) nv[n] = (P.bnd[l * 3 + n] * ri) / rad;
    const double xpi[3] = {xli[0] - 0.5, xli[1], xli[2]};
    const size_t cs = (size_t)A.nx * A.mz * A.my;
    Sten so, sp, si;
    make_sten(so, xlo, A.nx, A.my, A.mz);
    make_sten(sp, xpi, A.nx, A.my, A.mz);
    make_sten(si, xli, A.nx, A.my, A.mz);
    // force_tmp of this pair with the gradients of its interpolations:
    // Gi[d] = d/dpos of [cf uli - c3 nll - c4 onll - c12 dfll + c12 ppll]_d
    // (the first four terms are the one field qin at the inner point),
    // Go[d] = d/dpos of ulo_d
    double ulo[3], Go[3][3], Gi[3][3], g[3];
    for (int n = 0; n < 3; ++n) {
      ulo[n] = interp_vg(so, A.ubcp + n * cs, g);
      for (int e = 0; e < 3; ++e) Go[n][e] = g[e];
    }
    const double w1 = P.angv[m], w2v = P.angv[P.nbody + m], w3 = P.angv[2 * P.nbody + m];
    const double crossp[3] = {w2v * nv[2] - w3 * nv[1], w3 * nv[0] - w1 * nv[2], w1 * nv[1] - w2v * nv[0]};
    double linvel[3], unmag = 0.0, ubnmag = 0.0;
    for (int n = 0; n < 3; ++n) linvel[n] = P.vel[n * P.nbody + m] + crossp[n] * rb * st->dx;
    for (int d = 0; d < 3; ++d) {
      unmag += nv[d] * ulo[d];
      ubnmag += nv[d] * linvel[d];
    }
    const double rdi = ri * st->dx;
    const double K = 4.0 * 3.14159265358979323846 * (rdi * rdi) / (double)P.nrpr_active[m] * st->dr * st->dx;
    const double rn = st->revernorm == 1 ? 1.0 : 0.0, kap = (rb - ri) / (ro - rb);
    double ft[3];
    for (int d = 0; d < 3; ++d) {
      const double unorm = nv[d] * unmag, utang = ulo[d] - unorm;
      const double ubnorm = nv[d] * ubnmag, ubtang = linvel[d] - ubnorm;
      const double utangi = (ubtang * (ro - ri) - utang * (rb - ri)) / (ro - rb);
      const double unormi = rn * (ubnorm * (ro - ri) - unorm * (rb - ri)) / (ro - rb);
      double f = cf * -(utangi + unormi) + interp_vg(si, A.qin + d * cs, g);
      for (int e = 0; e < 3; ++e) Gi[d][e] = g[e];
      const double pp = interp_vg(sp, A.ppr + d * cs, g);
      for (int e = 0; e < 3; ++e) Gi[d][e] += c12 * g[e];
      f = f + c12 * pp;
      ft[d] = f * K;
    }
    // spreading, transposed
    bool solid[8];
    double fb[3][8];
    int count_so = 0;
    double wfl = 0.0, wall = 0.0;
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int q = (k * 2 + j) * 2 + i;
          const unsigned nd = si.node(i, j, k);
          solid[q] = A.mask[nd] == 0;
          for (int d = 0; d < 3; ++d) fb[d][q] = (double)J.afr[d * cs + nd];
          const double wq = si.w(i, j, k) / dx3;
          wall += wq;
          if (solid[q]) ++count_so;
          else wfl += wq;
        }
    double ftb[3], Ad[3];
    for (int d = 0; d < 3; ++d) {
      double ssum = 0.0, direct = 0.0;
#pragma unroll
      for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            const int q = (k * 2 + j) * 2 + i;
            if (solid[q]) {
              ssum += fb[d][q];
              direct += si.w(i, j, k) / dx3 * fb[d][q];
            }
          }
      Ad[d] = count_so > 0 ? ssum / (double)count_so : 0.0;
      ftb[d] = direct + (count_so > 0 ? wfl * Ad[d] : 0.0) + J.sc[6 + d] * wall;
    }
    {
      const double x1 = si.z[0], y1 = si.z[1], z1 = si.z[2];
#pragma unroll
      for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            const int q = (k * 2 + j) * 2 + i;
            double gq = 0.0;
            for (int d = 0; d < 3; ++d) gq += ft[d] / dx3 * ((solid[q] ? fb[d][q] : Ad[d]) + J.sc[6 + d]);
            const double X = i ? x1 : 1.0 - x1, Y = j ? y1 : 1.0 - y1, Z = k ? z1 : 1.0 - z1;
            gx[0] += gq * (i ? 1.0 : -1.0) * Y * Z;
            gx[1] += gq * X * (j ? 1.0 : -1.0) * Z;
            gx[2] += gq * X * Y * (k ? 1.0 : -1.0);
          }
    }
    // force_tmp, transposed onto the interpolated fields
    double h[3], Ph[3];
    for (int d = 0; d < 3; ++d) h[d] = K * ftb[d];
    const double nh = nv[0] * h[0] + nv[1] * h[1] + nv[2] * h[2];
    for (int d = 0; d < 3; ++d) Ph[d] = (h[d] - nv[d] * nh) + rn * nv[d] * nh;
    for (int d = 0; d < 3; ++d)
      for (int e = 0; e < 3; ++e) gx[e] += h[d] * Gi[d][e] + cf * kap * Ph[d] * Go[d][e];
    double* co = J.coef + ((size_t)m * P.nbnd + l) * 6;
    for (int d = 0; d < 3; ++d) {
      co[d] = h[d];
      co[3 + d] = cf * kap * Ph[d];
    }
  }
  add_grad(J.axc, P.nbody, m, gx, sh);
}

// The spread of the per-point adjoints through node-sorted records, one
// lane per target node (no atomics; each node's records in a fixed order).
// two record sets in one launch: set 0 spreads h into S (inner points), set 1
// cf kappa P h into aubcp (outer points), set 2 c12 h into appr (pressure
// points)
template <typename T>
__global__ void k_adj_rec_gather3(AdjArgs<T> J, int k0, int k1, int b1, double c12, size_t cs) {
  // workgroups [0, b1) take record set k0, the rest set k1
  const int b = blockIdx.x;
  const int q = b < b1 ? k0 : k1;
  const int t = (b < b1 ? b : b - b1) * blockDim.x + threadIdx.x;
  const AdjRecords& R = J.rec[q];
  if (t >= R.ntarget) return;
  const int off = q == 1 ? 3 : 0;
  const double scale = q == 2 ? c12 : 1.0;
  T* dst = q == 0 ? J.S : (q == 1 ? J.aubcp : J.appr);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int r = R.toff[t]; r < R.toff[t + 1]; ++r) {
    const double* c = J.coef + (size_t)R.rpt[r] * 6 + off;
    const double w = R.rw[r];
    a0 += w * c[0];
    a1 += w * c[1];
    a2 += w * c[2];
  }
  const size_t n = (size_t)R.tnode[t];
  dst[n] = (T)((double)dst[n] + scale * a0);
  dst[cs + n] = (T)((double)dst[cs + n] + scale * a1);
  dst[2 * cs + n] = (T)((double)dst[2 * cs + n] + scale * a2);
}

// (point, corner) records of one stencil kind for every active reversal pair
__global__ void k_adj_records(Particles P, int kind, int nx, int my, int mz, unsigned int sentinel, unsigned int* key,
                              int* rec, double* w) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P.nbody * P.nbnd) return;
  const int m = idx / P.nbnd, l = idx - m * P.nbnd;
  const bool act = P.if_rev[idx] != 0;
  const double r = kind == ADJ_REC_OUTER ? P.rado[m] : P.radi[m];
  double x[3];
  for (int n = 0; n < 3; ++n) x[n] = P.xc[n * P.nbody + m] + P.bnd[l * 3 + n] * r;
  if (kind == ADJ_REC_PRES) x[0] -= 0.5;
  Sten s;
  make_sten(s, x, nx, my, mz);
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        const size_t q = (size_t)idx * 8 + (k * 2 + j) * 2 + i;
        key[q] = act ? s.node(i, j, k) : sentinel;
        rec[q] = (int)q;
        w[q] = s.w(i, j, k);
      }
}
__global__ void k_adj_rec_gather_fields(int n, const int* rec_sorted, const double* w, int* rpt, double* rw) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int r = rec_sorted[t];
  rpt[t] = r / 8;
  rw[t] = w[r];
}

// calc_pres_visc_drag (boundary_condition.F90:1481-1645) transposed for the
// weighted force of the step: per surface point, the adjoints of pbcp (at
// the pressure point) and of omega, and the position gradient.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_drag_points(KArgs<T> A, Particles P, AdjArgs<T> J, int G) {
  __shared__ double sh[3 * 16];
  const int w = xcd_linear(blockIdx.x, gridDim.x);
  const int m = w / G, gi = w - m * G;
  if (m >= P.nbody) return;
  const int l = gi * blockDim.x + threadIdx.x;
  double gx[3] = {0, 0, 0};
  if (l < P.nbnd) {
    const DevState* st = A.st;
    const double c12 = st->coef[0][0] + st->coef[0][1];
    const double rb = P.rad[m], rdx = rb * st->dx;
    const double da1 = 4.0 * 3.14159265358979323846 * (rdx * rdx) / (double)P.nbnd;
    double xl[3];
    for (int n = 0; n < 3; ++n) xl[n] = P.xc[n * P.nbody + m] + P.bnd[l * 3 + n] * rb;
    const double xpb[3] = {xl[0] - 0.5, xl[1], xl[2]};
    Sten sp, sv;
    make_sten(sp, xpb, A.nx, A.my, A.mz);
    make_sten(sv, xl, A.nx, A.my, A.mz);
    const double cd1 = P.cd[l * 3], cd2 = P.cd[l * 3 + 1], cd3 = P.cd[l * 3 + 2];
    const double w0 = J.wforce[m], w1 = J.wforce[P.nbody + m], w2 = J.wforce[2 * P.nbody + m];
    const double plb = -c12 * da1 * (w0 * cd1 + w1 * cd2 + w2 * cd3);
    const double cw[3] = {cd2 * w2 - cd3 * w1, cd3 * w0 - cd1 * w2, cd1 * w1 - cd2 * w0};
    const size_t cs = (size_t)A.nx * A.mz * A.my;
    interp_g(sp, A.pbcp, plb, gx);
    scatter8(sp, J.apbcp, plb);
    for (int c = 0; c < 3; ++c) {
      const double vb = c12 * st->vis * da1 * cw[c];
      interp_g(sv, A.omega + c * cs, vb, gx);
      scatter8(sv, J.aom + c * cs, vb);
    }
  }
  add_grad(J.axc, P.nbody, m, gx, sh);
}

// ===========================================================================
// host side: the launches of one adjoint step
// ===========================================================================
template <typename K>
void set_lds(K kern, size_t bytes, bool capturing = false) {
  if (capturing) return;  // set by the eager step that precedes any capture
  ADJCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}
template <typename T, int KIND, int SH>
void r2c_t(const KArgs<T>& A, const AdjArgs<T>& J, const AdjGeom& g, int ncomp) {
  auto kern = k_adj_r2c_t<T, KIND, SH>;
  set_lds(kern, g.lds, g.capturing);
  kern<<<dim3(A.nx, ncomp), SH ? shape_nt(SH) : g.nt_plane, g.lds, g.stream>>>(A, J);
  ADJCHK(hipGetLastError());
}
template <typename T, int KIND>
void r2c_t_any(const KArgs<T>& A, const AdjArgs<T>& J, const AdjGeom& g, int ncomp) {
  switch (g.sh) {
    case 1: return r2c_t<T, KIND, 1>(A, J, g, ncomp);
    case 2: return r2c_t<T, KIND, 2>(A, J, g, ncomp);
    case 3: return r2c_t<T, KIND, 3>(A, J, g, ncomp);
    default: return r2c_t<T, KIND, 0>(A, J, g, ncomp);
  }
}
template <typename T, int SH>
void c2r_t(const KArgs<T>& A, const AdjGeom& g, int ncomp, const T* src, C2<T>* dst, double* part, bool add) {
  auto kern = k_adj_c2r_t<T, SH>;
  set_lds(kern, g.lds, g.capturing);
  kern<<<dim3(A.nx, ncomp), SH ? shape_nt(SH) : g.nt_plane, g.lds, g.stream>>>(A, src, dst, part, add ? 1 : 0);
  ADJCHK(hipGetLastError());
}
template <typename T>
void c2r_t_any(const KArgs<T>& A, const AdjGeom& g, int ncomp, const T* src, C2<T>* dst, double* part, bool add) {
  switch (g.sh) {
    case 1: return c2r_t<T, 1>(A, g, ncomp, src, dst, part, add);
    case 2: return c2r_t<T, 2>(A, g, ncomp, src, dst, part, add);
    case 3: return c2r_t<T, 3>(A, g, ncomp, src, dst, part, add);
    default: return c2r_t<T, 0>(A, g, ncomp, src, dst, part, add);
  }
}

// the tiled transposed velstep when the shape fits it (as the forward's)
template <typename T, int L, int M>
bool vel_adj_tiled(const KArgs<T>& A, const AdjArgs<T>& J, hipStream_t s, bool capturing) {
  const int nx = A.nx;
  if (nx % L != 0 || nx % 2 != 0) return false;
  const int bd = M * (nx / L);
  if (bd > 512 || nx / 2 > bd) return false;
  const int ntiles = (A.my2 * A.mz + M - 1) / M;
  const size_t lds = std::max(vel_lds_bytes(bd), (size_t)bd * sizeof(double));
  auto kern = bd <= 256 ? k_adj_velstep_tiled<T, L, M, 256> : k_adj_velstep_tiled<T, L, M, 512>;
  set_lds(kern, lds, capturing);
  kern<<<ntiles + 1, bd, lds, s>>>(A, J, ntiles);
  ADJCHK(hipGetLastError());
  return true;
}

}  // namespace

void adjoint_build_records(const Particles& P, int nx, int my, int mz, hipStream_t s, AdjRecords rec[3],
                           std::vector<void*>& mem) {
  auto dev = [&](size_t bytes) {
    void* p = nullptr;
    ADJCHK(hipMalloc(&p, bytes + 64));
    mem.push_back(p);
    return p;
  };
  const size_t nrec = (size_t)P.nbody * P.nbnd * 8, N = (size_t)nx * my * mz;
  const unsigned int sentinel = (unsigned int)N;
  int bits = 1;
  while ((1ull << bits) <= (unsigned long long)N) ++bits;
  unsigned int* key = (unsigned int*)dev(nrec * 4);
  unsigned int* key_s = (unsigned int*)dev(nrec * 4);
  int* recn = (int*)dev(nrec * 4);
  int* rec_s = (int*)dev(nrec * 4);
  double* w = (double*)dev(nrec * 8);
  int* counts = (int*)dev((nrec + 1) * 4);
  int* nruns = (int*)dev(4);
  size_t tb1 = 0, tb2 = 0, tb3 = 0;
  ADJCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, key, key_s, recn, rec_s, (int)nrec, 0, bits, s));
  ADJCHK(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb2, key_s, key, counts, nruns, (int)nrec, s));
  ADJCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, counts, recn, (int)nrec + 1, s));
  void* tmp = dev(std::max(tb1, std::max(tb2, tb3)));
  const size_t npts = (size_t)P.nbody * P.nbnd;
  for (int kind = 0; kind < 3; ++kind) {
    AdjRecords& R = rec[kind];
    R.ntarget = 0;
    if (nrec == 0) continue;
    unsigned int* ukeys = (unsigned int*)dev((nrec + 1) * 4);
    int* toff = (int*)dev((nrec + 2) * 4);
    k_adj_records<<<(unsigned)((npts + 255) / 256), 256, 0, s>>>(P, kind, nx, my, mz, sentinel, key, recn, w);
    ADJCHK(hipGetLastError());
    ADJCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb1, key, key_s, recn, rec_s, (int)nrec, 0, bits, s));
    ADJCHK(hipcub::DeviceRunLengthEncode::Encode(tmp, tb2, key_s, ukeys, counts, nruns, (int)nrec, s));
    int hruns = 0;
    ADJCHK(hipMemcpyAsync(&hruns, nruns, sizeof(int), hipMemcpyDeviceToHost, s));
    ADJCHK(hipStreamSynchronize(s));
    unsigned int lastkey = 0;
    if (hruns > 0) ADJCHK(hipMemcpy(&lastkey, ukeys + hruns - 1, sizeof(unsigned int), hipMemcpyDeviceToHost));
    const int nt = hruns - (hruns > 0 && lastkey >= sentinel ? 1 : 0);
    ADJCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tb3, counts, toff, nt + 1, s));
    int acc = 0;
    ADJCHK(hipMemcpyAsync(&acc, toff + nt, sizeof(int), hipMemcpyDeviceToHost, s));
    ADJCHK(hipStreamSynchronize(s));
    int* rpt = (int*)dev((size_t)(acc + 1) * 4);
    double* rw = (double*)dev((size_t)(acc + 1) * 8);
    if (acc > 0) {
      k_adj_rec_gather_fields<<<(acc + 255) / 256, 256, 0, s>>>(acc, rec_s, w, rpt, rw);
      ADJCHK(hipGetLastError());
    }
    R.ntarget = nt;
    R.tnode = (const int*)ukeys;
    R.toff = toff;
    R.rpt = rpt;
    R.rw = rw;
  }
  ADJCHK(hipStreamSynchronize(s));
}

template <typename T>
void adjoint_step_kernels(const KArgs<T>& A, const Particles& P, const AdjArgs<T>& J, const AdjGeom& g) {
  const hipStream_t s = g.stream;
  const size_t E = (size_t)A.my2 * A.mz, N = (size_t)A.nx * A.my * A.mz;
  const size_t nspec = (size_t)A.nx * E;
  const unsigned ebl = (unsigned)((nspec + 255) / 256), pbl = (unsigned)((N + 255) / 256);
  const bool drag = J.wforce != nullptr && P.nbody > 0;
  {
    // aubcp and apnew are first written by k_adj_phys and k_adj_modes<PGRAD>
    ZeroSet<T> z{};
    z.p[0] = J.S;
    z.p[1] = J.appr;
    z.n[0] = z.n[1] = 3 * N;
    z.k = 2;
    if (drag) {
      z.p[2] = J.aom;
      z.p[3] = J.apbcp;
      z.n[2] = 3 * N;
      z.n[3] = N;
      z.k = 4;
    }
    k_adj_zero_set<T><<<dim3((unsigned)((3 * N + 255) / 256), z.k), 256, 0, s>>>(z);
    ADJCHK(hipGetLastError());
  }
  // velstep
  const int nm = (int)E;
  bool tiled;
  const bool cap = g.capturing;
  if constexpr (sizeof(T) == 4)
    tiled = vel_adj_tiled<T, 8, 16>(A, J, s, cap) || vel_adj_tiled<T, 4, 16>(A, J, s, cap);
  else
    tiled = vel_adj_tiled<T, 4, 8>(A, J, s, cap) || vel_adj_tiled<T, 2, 8>(A, J, s, cap);
  if (!tiled) {
    k_adj_velstep<T><<<(nm + 63) / 64, 64, 0, s>>>(A, J, nm);
    ADJCHK(hipGetLastError());
  }
  // umean update and the mean of fr; the drag's mpg term
  k_adj_scal<T><<<1, 256, 0, s>>>(A, J, P.nbody, P.rad);
  ADJCHK(hipGetLastError());
  // ff = r2c(fr - frmean)
  r2c_t_any<T, ADJ_AFR>(A, J, g, 3);
  // the reversal forcing and the drag
  const int G = (P.nbnd + 255) / 256;
  const unsigned nbg = (unsigned)((P.nbody * G + 7) / 8 * 8);
  if (P.nbody > 0) {
    k_adj_forcing_points<T><<<nbg, 256, 0, s>>>(A, P, J, G);
    ADJCHK(hipGetLastError());
    const int bi = (J.rec[ADJ_REC_INNER].ntarget + 255) / 256, bp = (J.rec[ADJ_REC_PRES].ntarget + 255) / 256;
    if (bi + bp > 0) {
      k_adj_rec_gather3<T><<<bi + bp, 256, 0, s>>>(J, ADJ_REC_INNER, ADJ_REC_PRES, bi, g.c12, N);
      ADJCHK(hipGetLastError());
    }
    if (drag) {
      k_adj_drag_points<T><<<nbg, 256, 0, s>>>(A, P, J, G);
      ADJCHK(hipGetLastError());
    }
  }
  // calc_pgrad -> p and mpg
  c2r_t_any<T>(A, g, 3, J.appr, J.Z, J.part, false);
  k_adj_sum_part<<<1, 256, 0, s>>>(J.part, A.nx, J.sc, 3);
  ADJCHK(hipGetLastError());
  k_adj_modes<T, ADJ_PGRAD><<<ebl, 256, 0, s>>>(A, J);
  ADJCHK(hipGetLastError());
  // compute_omega -> u (its x-differences of ubcp go with the products below)
  if (drag) {
    c2r_t_any<T>(A, g, 3, J.aom, J.Z, nullptr, false);
    k_adj_modes<T, ADJ_OMEGA><<<ebl, 256, 0, s>>>(A, J);
    ADJCHK(hipGetLastError());
  }
  // nlbc, onlbc, diffn at the inner points: -> nl, onl, u
  c2r_t_any<T>(A, g, 3, J.S, J.Z, nullptr, false);
  k_adj_modes<T, ADJ_SFIELD><<<ebl, 256, 0, s>>>(A, J);
  ADJCHK(hipGetLastError());
  // form_nl and the sign of nonlinear
  r2c_t_any<T, ADJ_NLQ>(A, J, g, 8);
  k_adj_phys<T><<<pbl, 256, 0, s>>>(A, J, drag ? 1 : 0);
  ADJCHK(hipGetLastError());
  {
    const int bo = (J.rec[ADJ_REC_OUTER].ntarget + 255) / 256;
    if (P.nbody > 0 && bo > 0) {
      k_adj_rec_gather3<T><<<bo, 256, 0, s>>>(J, ADJ_REC_OUTER, ADJ_REC_OUTER, bo, g.c12, N);
      ADJCHK(hipGetLastError());
    }
  }
  // ubcp = c2r(u) + umean, pbcp = c2r(p)
  c2r_t_any<T>(A, g, 3, J.aubcp, J.aunew, J.part, true);
  k_adj_sum_part<<<1, 256, 0, s>>>(J.part, A.nx, J.sc, 0);
  ADJCHK(hipGetLastError());
  if (drag) c2r_t_any<T>(A, g, 1, J.apbcp, J.apnew, nullptr, true);
}

template void adjoint_step_kernels<float>(const KArgs<float>&, const Particles&, const AdjArgs<float>&, const AdjGeom&);
template void adjoint_step_kernels<double>(const KArgs<double>&, const Particles&, const AdjArgs<double>&,
                                           const AdjGeom&);
