// pib_velstep.h -- tiled velstep (steptheflow.F90:52-160) for gfx950.
//
// One workgroup owns a tile of M consecutive (ky,kz) modes; in the spectral
// layout S[c][i][kz][ky] those modes are contiguous for every (c,i), so each
// row of a tile is one M*sizeof(complex) burst (128 B for 16 float or 8
// double modes).  P = nx/L lanes per mode each hold L consecutive x rows of
// every line in registers; lane t = s*M + m (segment s, mode m).
//
// The three solves of the step -- pressure (poisson.F90:55-239), ustep
// (usteptime.F90:45-258) and divcorr (poisson.F90:241-508) -- are symmetric
// Toeplitz tridiagonal systems along x, cyclic for every mode except the
// pressure/divcorr solves of the (0,0) mode, which the reference solves
// non-cyclic on rows 1..nx-1 with p(nx) = 0.  The reference runs Thomas with
// a Sherman-Morrison correction on each x-line; here each line is solved by
// a partition method: the L-1 interior rows of every segment are eliminated
// against the two separators around it (X_{s-1} = row sL-1, X_s = row
// sL+L-1), which leaves a symmetric cyclic tridiagonal system in the P
// separators (pinned X_{P-1} = 0 for the (0,0) mode); one lane per (mode,
// right-hand side) solves that with Thomas + Sherman-Morrison, and every
// lane then back-substitutes its interior rows.  The arithmetic is that of
// an exact direct solve, rounded differently from the reference's.
//
// The pressure solution replaces p (the reference's p = pold + (x - pold));
// it is stored right after the pressure solve and updated in place by
// divcorr.  The (0,0) mode is processed in double precision by the last
// workgroup, with L = 2 and one mode (its pressure is rounded to the bulk
// type when stored between the two phases).
#pragma once
#include "pib_kernels.h"

// Interior block (size L-1) of a symmetric Toeplitz (a, b, a) system and the
// reduced separator system it induces.
template <typename R, int L>
struct Toep {
  R a, b;
  R gam[L];  // Thomas multipliers (index 1..L-2)
  R ib[L];   // inverse pivots (0..L-2)
  R v[L];    // T'^{-1}(a e_0): interior response to a unit left separator
  R al, be;  // reduced system off-diagonal and diagonal
};

template <typename R, int L>
__device__ __forceinline__ void toep_init(Toep<R, L>& t, R a, R b, int& st) {
  t.a = a;
  t.b = b;
  R bet = b;
  if (bet == (R)0) {
    bet = (R)1.0e-30;
    st = PIB_ERR_PIVOT;
  }
  t.ib[0] = (R)1 / bet;
  t.gam[0] = (R)0;
#pragma unroll
  for (int k = 1; k < L - 1; ++k) {
    t.gam[k] = a * t.ib[k - 1];
    bet = b - a * t.gam[k];
    if (bet == (R)0) {
      bet = (R)1.0e-30;
      st = PIB_ERR_PIVOT;
    }
    t.ib[k] = (R)1 / bet;
  }
  R y[L];
  y[0] = a * t.ib[0];
#pragma unroll
  for (int k = 1; k < L - 1; ++k) y[k] = (-a * y[k - 1]) * t.ib[k];
#pragma unroll
  for (int k = L - 3; k >= 0; --k) y[k] = y[k] - t.gam[k + 1] * y[k + 1];
#pragma unroll
  for (int k = 0; k < L - 1; ++k) t.v[k] = y[k];
  // T' is symmetric and persymmetric, so T'^{-1}(a e_{L-2}) is v reversed
  t.al = -a * t.v[L - 2];
  t.be = b - (R)2 * a * t.v[0];
}

// interior rows 0..L-2 of d solved in place (row L-1 untouched)
template <typename R, int L>
__device__ __forceinline__ void toep_interior(const Toep<R, L>& t, C2<R>* d) {
  d[0] = cscal(t.ib[0], d[0]);
#pragma unroll
  for (int k = 1; k < L - 1; ++k) d[k] = cscal(t.ib[k], d[k] - cscal(t.a, d[k - 1]));
#pragma unroll
  for (int k = L - 3; k >= 0; --k) d[k] = d[k] - cscal(t.gam[k + 1], d[k + 1]);
}

// Cyclic symmetric tridiagonal (al, be, al) of size P in place on r; z and g
// are P-long work vectors.  Thomas on the Sherman-Morrison-modified matrix
// with gamma = -be, then x = y - (v.y)/(1 + v.z) z.  Entry k of each vector
// is at k*ld: the systems of a workgroup are interleaved, so that the lanes
// solving neighbouring systems touch neighbouring LDS words at every step.
template <typename R>
__device__ void reduced_cyclic(C2<R>* r, R* z, R* g, int P, int ld, R al, R be, int& st) {
  if (P == 1) {
    R d = be + (R)2 * al;
    if (d == (R)0) {
      d = (R)1.0e-30;
      st = PIB_ERR_PIVOT;
    }
    r[0] = cscal((R)1 / d, r[0]);
    return;
  }
  if (P == 2) {
    R det = be * be - (R)4 * al * al;
    if (det == (R)0) {
      det = (R)1.0e-30;
      st = PIB_ERR_PIVOT;
    }
    C2<R> r0 = r[0], r1 = r[ld];
    R id = (R)1 / det;
    r[0] = cscal(id, cscal(be, r0) - cscal((R)2 * al, r1));
    r[ld] = cscal(id, cscal(be, r1) - cscal((R)2 * al, r0));
    return;
  }
  const R gm = -be;
  R bet = be - gm;
  if (bet == (R)0) {
    bet = (R)1.0e-30;
    st = PIB_ERR_PIVOT;
  }
  R ib = (R)1 / bet;
  C2<R> y = cscal(ib, r[0]);
  r[0] = y;
  R zz = gm * ib;
  z[0] = zz;
  g[0] = (R)0;
  for (int k = 1; k < P; ++k) {
    R gk = al * ib;
    R dk = (k == P - 1) ? be - al * al / gm : be;
    bet = dk - al * gk;
    if (bet == (R)0) {
      bet = (R)1.0e-30;
      st = PIB_ERR_PIVOT;
    }
    ib = (R)1 / bet;
    g[k * ld] = gk;
    y = cscal(ib, r[k * ld] - cscal(al, y));
    r[k * ld] = y;
    R uk = (k == P - 1) ? al : (R)0;
    zz = (uk - al * zz) * ib;
    z[k * ld] = zz;
  }
  for (int k = P - 2; k >= 0; --k) {
    R gk1 = g[(k + 1) * ld];
    y = r[k * ld] - cscal(gk1, y);
    r[k * ld] = y;
    zz = z[k * ld] - gk1 * zz;
    z[k * ld] = zz;
  }
  R den = (R)1 + z[0] + al * z[(P - 1) * ld] / gm;
  C2<R> f = cscal((R)1 / den, r[0] + cscal(al / gm, r[(P - 1) * ld]));
  for (int k = 0; k < P; ++k) r[k * ld] = r[k * ld] - cscal(z[k * ld], f);
}

// Non-cyclic (al, be, al) on separators 0..P-2 with X_{P-1} = 0 (entry k at
// k*ld as above).
template <typename R>
__device__ void reduced_pinned(C2<R>* r, R* g, int P, int ld, R al, R be, int& st) {
  const int n = P - 1;
  if (n > 0) {
    R bet = be;
    if (bet == (R)0) {
      bet = (R)1.0e-30;
      st = PIB_ERR_PIVOT;
    }
    C2<R> y = cscal((R)1 / bet, r[0]);
    r[0] = y;
    for (int k = 1; k < n; ++k) {
      R gk = al / bet;
      g[k * ld] = gk;
      bet = be - al * gk;
      if (bet == (R)0) {
        bet = (R)1.0e-30;
        st = PIB_ERR_PIVOT;
      }
      y = cscal((R)1 / bet, r[k * ld] - cscal(al, y));
      r[k * ld] = y;
    }
    for (int k = n - 2; k >= 0; --k) {
      y = r[k * ld] - cscal(g[(k + 1) * ld], y);
      r[k * ld] = y;
    }
  }
  r[(P - 1) * ld] = mk((R)0, (R)0);
}

// LDS carve-up of one tile workgroup (bd = blockDim.x lanes):
//   ef, el  [3][bd] complex      first / last row exchange between segments
//   rr      [P][NR*M] complex    reduced right-hand sides -> separators
//                                (separator k of system j = r*M + m)
//   rz, rg  [P][NR*M] real       reduced-solve work vectors
template <typename R>
struct VLds {
  C2<R>* ef;
  C2<R>* el;
  C2<R>* rr;
  R* rz;
  R* rg;
  __device__ VLds(unsigned char* raw, int bd) {
    ef = (C2<R>*)raw;
    el = ef + 3 * bd;
    rr = el + 3 * bd;
    rz = (R*)(rr + 3 * bd);
    rg = rz + 3 * bd;
  }
};
// bytes of dynamic LDS a tile workgroup of bd lanes needs (double covers float)
static inline size_t vel_lds_bytes(int bd) {
  return (size_t)bd * (6 * sizeof(C2<double>) + 3 * sizeof(C2<double>) + 6 * sizeof(double));
}

// The value is complete at this point of the program: its loads and
// arithmetic are not deferred past here.  Bounds how many loads are in
// flight (and so the registers) of the per-component load phases.
template <typename R>
__device__ __forceinline__ void pin(C2<R>& v) { __asm__ __volatile__("" : "+v"(v.x), "+v"(v.y)); }

template <typename T, typename R, int L, int M>
struct VTile {
  const KArgs<T>& A;
  const int nx, P, nthr, nm, bd, tid, q0;
  const bool mean;
  bool lane, act;
  int s, m, q;
  R wy, wz, w2;
  VLds<R> S;
  int st;   // pivot failures of this lane's own mode (counted when act)
  int sst;  // pivot failures of reduced systems this lane solved
  __device__ VTile(const KArgs<T>& a, int q0_, bool mean_, unsigned char* raw)
      : A(a), nx(a.nx), P(a.nx / L), nthr(M * (a.nx / L)), nm(a.my2 * a.mz), bd(blockDim.x), tid(threadIdx.x),
        q0(q0_), mean(mean_), S(raw, blockDim.x), st(0), sst(0) {
    lane = tid < nthr;
    s = lane ? tid / M : 0;
    m = lane ? tid % M : 0;
    q = q0 + m;
    act = lane && q < nm && (mean || q != 0);
    int qq = act ? q : 0;
    int ky = qq % A.my2, kz = qq / A.my2;
    wy = act ? (R)A.wy[ky] : (R)0;
    wz = act ? (R)A.wz[kz] : (R)0;
    w2 = act ? (R)A.w2[(size_t)kz * A.my2 + ky] : (R)1;
  }
  __device__ __forceinline__ int nbl(int s2) const {
    s2 = s2 < 0 ? s2 + P : (s2 >= P ? s2 - P : s2);
    return s2 * M + m;
  }
  __device__ __forceinline__ size_t at(int c, int row) const {
    row = row < 0 ? row + nx : (row >= nx ? row - nx : row);
    return ((size_t)c * nx + row) * nm + q;
  }
  __device__ __forceinline__ C2<R> ld(const C2<T>* f, int c, int row) const {
    if (!act) return mk((R)0, (R)0);
    C2<T> v = f[at(c, row)];
    return mk((R)v.x, (R)v.y);
  }
  __device__ __forceinline__ void stv(C2<T>* f, int c, int row, C2<R> v) const {
    if (act) f[at(c, row)] = mk((T)v.x, (T)v.y);
  }
  // row 0 of the next segment for NL lines
  template <int NL>
  __device__ void next_first(C2<R> (*x)[L], C2<R>* out) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NL; ++c) S.ef[c * bd + tid] = x[c][0];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NL; ++c) out[c] = S.ef[c * bd + nbl(s + 1)];
  }
  // solve NR lines x[r][0..L-1] with matrix t; pinned: (0,0)-mode pressure form
  template <int NR>
  __device__ void solve(const Toep<R, L>& t, C2<R> (*x)[L], bool pinned) {
#pragma unroll
    for (int r = 0; r < NR; ++r) toep_interior(t, x[r]);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NR; ++r) S.ef[r * bd + tid] = x[r][0];
    __syncthreads();
    if (lane) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        C2<R> yn = S.ef[r * bd + nbl(s + 1)];
        C2<R> rhs = (x[r][L - 1] - cscal(t.a, x[r][L - 2])) - cscal(t.a, yn);
        S.rr[s * (NR * M) + r * M + m] = rhs;
      }
    }
    __syncthreads();
    for (int j = tid; j < NR * M; j += bd) {
      // lane j = r*M + mm solves right-hand side r of mode mm with that
      // mode's reduced coefficients, staged by its s = 0 lane (a workgroup
      // of fewer than NR*M lanes takes several systems per lane)
      int r = j / M, mm = j % M;
      int e = 0;
      R al = S.rg[j], be = S.rz[j];
      if (pinned) reduced_pinned(S.rr + j, S.rg + j, P, NR * M, al, be, e);
      else reduced_cyclic(S.rr + j, S.rz + j, S.rg + j, P, NR * M, al, be, e);
      int qm = q0 + mm;
      if (e && qm < nm && (mean || qm != 0)) sst = e;
    }
    __syncthreads();
    if (lane) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int o = r * M + m;
        C2<R> Xl = S.rr[(s == 0 ? P - 1 : s - 1) * (NR * M) + o], Xs = S.rr[s * (NR * M) + o];
#pragma unroll
        for (int k = 0; k < L - 1; ++k) x[r][k] = (x[r][k] - cscal(t.v[k], Xl)) - cscal(t.v[L - 2 - k], Xs);
        x[r][L - 1] = Xs;
      }
    }
  }
  // stage the reduced coefficients of this lane's mode where the solving
  // lanes read them (must precede solve<NR>; rz/rg slots 0 of each system)
  template <int NR>
  __device__ void stage(const Toep<R, L>& t) {
    __syncthreads();
    if (lane && s == 0) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        S.rg[r * M + m] = t.al;
        S.rz[r * M + m] = t.be;
      }
    }
  }
};

template <typename T, typename R, int L, int M>
__device__ void vel_tile(const KArgs<T>& A, int q0, bool mean, unsigned char* raw, double& dmax, int& status) {
  VTile<T, R, L, M> V(A, q0, mean, raw);
  const DevState* st = A.st;
  const int rks = A.rks;
  const R c1 = (R)st->coef[rks - 1][0], c2 = (R)st->coef[rks - 1][1], c3 = (R)st->coef[rks - 1][2],
          c4 = (R)st->coef[rks - 1][3];
  const R dx = (R)st->dx, dx2 = dx * dx, dt = (R)st->s.dt, vis = (R)st->vis;
  const R idx = (R)1 / dx, idx2 = (R)1 / dx2;
  const R c12 = c1 + c2;
  const int i0 = V.s * L;
  const R wy = V.wy, wz = V.wz, w2 = V.w2;
  // ---- h_c = c3 nl_c + ff_c + c4 onl_c (rows i0..i0+L-1) ----
  C2<R> h[3][L];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int k = 0; k < L; ++k)
      h[c][k] = (cscal(c3, V.ld(A.nl, c, i0 + k)) + V.ld(A.ff, c, i0 + k)) + cscal(c4, V.ld(A.onl, c, i0 + k));
    // one component's 3L loads in flight at a time
#pragma unroll
    for (int k = 0; k < L; ++k) pin(h[c][k]);
  }
  // ---- pressure (poisson.F90:55-239) ----
  // the pressure/divcorr matrix is set up again for divcorr rather than kept
  // live through ustep (registers bound the workgroups per CU)
  const R ap = idx2 - (R)0.25 * w2, bp = (R)-2 * idx2 - (R)0.5 * w2;
  C2<R> hn[3];
  V.template next_first<3>(h, hn);
  C2<R> p[1][L];
  {
    const R ic = (R)1 / c12;
#pragma unroll
    for (int k = 0; k < L; ++k) {
      C2<R> n0 = k < L - 1 ? h[0][k + 1] : hn[0];
      C2<R> n1 = k < L - 1 ? h[1][k + 1] : hn[1];
      C2<R> n2 = k < L - 1 ? h[2][k + 1] : hn[2];
      C2<R> t = (cscal(idx, n0 - h[0][k]) + cik((R)0.5 * wy, n1 + h[1][k])) + cik((R)0.5 * wz, n2 + h[2][k]);
      p[0][k] = cscal(ic, t);
    }
  }
  {
    Toep<R, L> tp;
    toep_init(tp, ap, bp, V.st);
    V.template stage<1>(tp);
    V.template solve<1>(tp, p, mean);
  }
  if (mean)  // the (0,0) mode keeps only the real part (poisson.F90:139)
#pragma unroll
    for (int k = 0; k < L; ++k) p[0][k].y = (R)0;
  // the pressure goes to p now and is read back for the divcorr update, so
  // it holds no registers through the ustep and divcorr solves
#pragma unroll
  for (int k = 0; k < L; ++k) V.stv(A.p, 0, i0 + k, p[0][k]);
  // ---- ustep (usteptime.F90:45-258) ----
  C2<R> pl;
  {
    __syncthreads();
    V.S.el[V.tid] = p[0][L - 1];
    __syncthreads();
    pl = V.S.el[V.nbl(V.s - 1)];
  }
  {
    const R f = (R)1 - dt * vis * c1 * w2;
    const R sg = dt * c12 * idx, sh = dt * c12 * (R)0.5, sl = dt * vis * c1 * idx2;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      C2<R> uu[L + 2];
#pragma unroll
      for (int k = 0; k < L + 2; ++k) uu[k] = V.ld(A.u, c, i0 + k - 1);
      const R kk = c == 1 ? wy : wz;
#pragma unroll
      for (int k = 0; k < L; ++k) {
        C2<R> pm = k > 0 ? p[0][k - 1] : pl;
        C2<R> r = cscal(f, uu[k + 1]) + cscal(dt, h[c][k]);
        if (c == 0) r = r - cscal(sg, p[0][k] - pm);
        else r = r - cik(kk * sh, pm + p[0][k]);
        C2<R> lap = (uu[k] + uu[k + 2]) - cscal((R)2, uu[k + 1]);
        h[c][k] = r + cscal(sl, lap);
      }
#pragma unroll
      for (int k = 0; k < L; ++k) pin(h[c][k]);
    }
  }
  {
    Toep<R, L> th;
    const R ah = -dt * vis * c2 * idx2;
    const R bh = (R)1 + dt * vis * c2 * ((R)2 * idx2 + w2);
    toep_init(th, ah, bh, V.st);
    V.template stage<3>(th);
    V.template solve<3>(th, h, false);
  }
  // ---- divcorr (poisson.F90:241-508) ----
  const R alp = c12 * dt;
  C2<R> phi[1][L];
  {
    V.template next_first<3>(h, hn);
    const R ia = (R)1 / alp;
#pragma unroll
    for (int k = 0; k < L; ++k) {
      C2<R> n0 = k < L - 1 ? h[0][k + 1] : hn[0];
      C2<R> n1 = k < L - 1 ? h[1][k + 1] : hn[1];
      C2<R> n2 = k < L - 1 ? h[2][k + 1] : hn[2];
      C2<R> t = (cscal(idx, n0 - h[0][k]) + cik((R)0.5 * wy, h[1][k] + n1)) + cik((R)0.5 * wz, h[2][k] + n2);
      phi[0][k] = cscal(ia, t);
    }
  }
  {
    Toep<R, L> tp;
    toep_init(tp, ap, bp, V.st);
    V.template stage<1>(tp);
    V.template solve<1>(tp, phi, mean);
  }
  if (mean)  // poisson.F90:320
#pragma unroll
    for (int k = 0; k < L; ++k) phi[0][k].y = (R)0;
  C2<R> fm, fp;
  {
    __syncthreads();
    V.S.ef[V.tid] = phi[0][0];
    V.S.el[V.tid] = phi[0][L - 1];
    __syncthreads();
    fm = V.S.el[V.nbl(V.s - 1)];
    fp = V.S.ef[V.nbl(V.s + 1)];
  }
  {
    const R cv = (c2 * dt) * vis;
    const R ga = alp * idx, gh = alp * (R)0.5;
#pragma unroll
    for (int k = 0; k < L; ++k) {
      C2<R> dm = k > 0 ? phi[0][k - 1] : fm;
      C2<R> dp = k < L - 1 ? phi[0][k + 1] : fp;
      C2<R> di = phi[0][k];
      C2<R> sm = dm + di;
      h[0][k] = h[0][k] - cscal(ga, di - dm);
      h[1][k] = h[1][k] - cik(gh * wy, sm);
      h[2][k] = h[2][k] - cik(gh * wz, sm);
      C2<R> lap = (dm - cscal((R)2, di)) + dp;
      C2<R> corr = cscal(idx2, lap) - cscal(w2, di);
      V.stv(A.p, 0, i0 + k, (V.ld(A.p, 0, i0 + k) + di) - cscal(cv, corr));
    }
#pragma unroll
    for (int k = 0; k < L; ++k) {
      V.stv(A.u, 0, i0 + k, h[0][k]);
      V.stv(A.u, 1, i0 + k, h[1][k]);
      V.stv(A.u, 2, i0 + k, h[2][k]);
    }
  }
  // ---- check_divergence of the new velocity ----
  V.template next_first<3>(h, hn);
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < L; ++k) {
    C2<R> n0 = k < L - 1 ? h[0][k + 1] : hn[0];
    C2<R> n1 = k < L - 1 ? h[1][k + 1] : hn[1];
    C2<R> n2 = k < L - 1 ? h[2][k + 1] : hn[2];
    C2<R> t = (cscal(idx, n0 - h[0][k]) + cik((R)0.5 * wy, h[1][k] + n1)) + cik((R)0.5 * wz, h[2][k] + n2);
    double mg = sqrt((double)t.x * (double)t.x + (double)t.y * (double)t.y);
    mx = fmax(mx, mg);
  }
  if (V.act) dmax = fmax(dmax, mx);
  if (V.act && V.st) status = V.st;
  if (V.sst) status = V.sst;
}

// grid: ntiles tile workgroups + 1 workgroup for the (0,0) mode;
// block = M * nx / L lanes, at most NT.  Workgroups of up to 256 lanes are
// held to three per CU (168 registers), so that the 128^3 tile counts fill
// the chip in one (FP32) or two (FP64) rounds of workgroups.
template <typename T, int L, int M, int NT>
__global__ void __launch_bounds__(NT, NT <= 256 ? 3 : 1) k_velstep_tiled(KArgs<T> A, int ntiles, int* status) {
  extern __shared__ __align__(16) unsigned char vlds[];
  double dmax = 0.0;
  int st = 0;
  if ((int)blockIdx.x == ntiles) vel_tile<T, double, 2, 1>(A, 0, true, vlds, dmax, st);
  else vel_tile<T, T, L, M>(A, (int)blockIdx.x * M, false, vlds, dmax, st);
  if (st) *status = st;
  __syncthreads();
  double mx[1] = {dmax};
  if (blockDim.x % 64 == 0) block_maxes_w<1>(mx, (double*)vlds);
  else mx[0] = block_max_any(dmax, (double*)vlds);
  if (threadIdx.x == 0) A.part[blockIdx.x] = mx[0];
}
