// pib_scalar_nl.h -- the explicit half of scalstep on the device: the
// nonlinear term (nlphi_graduphi, nlphi.F90:56-155, with form_nlphi and
// compute_phim_heat_ratio, nl_allphi.F90:59-536) and the immersed-boundary
// forcing of the scalar (bcsetscal with calc_diffn, calc_bndrev_data and
// calc_inner_rev_data, bc_scalar.F90:65-1033), serial branch, x periodic with
// the heat-ratio twist phi(x + L) = phi(x) / heat_ratio, one species.
//
// Physical scalar arrays hold the reference's planes 0..nx+1, [i][z][y]; the
// spectral fields (FP64 in both precisions) planes 1..nx+1 as 0..nx,
// [p][kz][ky].  The transforms run in the context's precision T on the plane
// FFT engine.  The heat ratios live on the device (heat_ratio,
// heat_ratio_new, heat_ratio_old, gamma), so no step waits on the host.
//
// Launch order (one stream): k_scal_phi (phi, products, y/z terms, plane
// sums) -> k_scal_ratio (heat_ratio_new, phase means) -> [k_scal_faces, the
// upwind face fluxes] -> k_scal_xflux (the
// x-flux derivative, the sign, plane nx+1) -> k_scal_phys (nl, onl, diffn and
// grad phi in physical space) -> fill fr -> k_scal_bnd (surface flux and
// DOBND forcing, one workgroup per body) -> k_scal_rev (reversal pairs, one
// lane per pair) -> k_scal_spread (the per-point values gathered onto the
// grid through the fluid's node-sorted records) -> k_scal_ff (ff2rc of fr)
// -> k_phistep -> k_ratio_roll.  No atomics: every sum has a fixed order.
#pragma once

template <typename T>
struct ScalStep {
  C2<double>* phif;
  C2<double>* nlphif;
  C2<double>* nlz;  // the wz ff2rc(w phi) term of nlphif, planes 1..nx, until k_scal_xflux adds it
  T* ph2;           // phi of planes 1..nx as the second k_scal_phi workgroup of a plane transforms it
  const C2<double>* onlphif;
  C2<double>* ffphi;
  T *ph, *nlp, *onp, *df, *g, *fr, *fdd;  // physical, planes 0..nx+1; g holds 3 arrays
  size_t np;                              // entries of one physical array
  double* part;       // per plane p < nx: fluid and solid sums of phi; plane nx: sum u phi, sum u, count
  double* ratio;      // heat_ratio, heat_ratio_new, heat_ratio_old, gamma, phi_fluid_mean, phi_solid_mean
  double* flux_body;  // (nbody)
  const double* phisurf;  // (nbody)
  // per point of the spreading records: (tempor, force_dist) of the first
  // evaluation and of the one a period down; rows as the fluid's records
  // number them (reversal pairs m*nbnd + l, surface points after them)
  double* spt;
  double dx;
  int upwind, dobnd, dorpr, frmeanfluid;
  int skip_onl;  // onp already holds ff2cr(onlphif): the previous step's nlp
};

template <typename T>
__device__ __forceinline__ size_t sps(const KArgs<T>& A, int i, int y, int z) {
  return ((size_t)i * A.mz + z) * A.my + y;
}
template <typename T>
__device__ __forceinline__ size_t sss(const KArgs<T>& A, int p, int ky, int kz) {
  return ((size_t)p * A.mz + kz) * A.my2 + ky;
}
__device__ __forceinline__ C2<double> cikd(double k, C2<double> z) { return mk(-k * z.y, k * z.x); }
__device__ __forceinline__ C2<double> cmulr(C2<double> a, double s) { return mk(a.x * s, a.y * s); }
__device__ __forceinline__ C2<double> cdivr(C2<double> a, double s) { return mk(a.x / s, a.y / s); }
template <typename T>
__device__ __forceinline__ C2<double> dbl(C2<T> v) { return mk((double)v.x, (double)v.y); }
// the fluid flag of reference plane i in 0..nx+1 (planes 0 and nx+1 are the
// periodic copies of nx and 1, dependent_functions.F90:913-915)
template <typename T>
__device__ __forceinline__ bool fluid_at(const KArgs<T>& A, int i, int y, int z) {
  int pl = i == 0 ? A.nx - 1 : (i == A.nx + 1 ? 0 : i - 1);
  return A.mask[((size_t)pl * A.mz + z) * A.my + y] != 0;
}

template <typename T>
__global__ void k_fill_t(T* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (T)0;
}

// ---------------------------------------------------------------------------
// k_scal_phi, per reference plane i = p+1 in 1..nx+1 (nl_allphi.F90:97-160,
// 380-419, 448-483): phi = ff2cr(phif), the central x-flux u phi, the sums of
// the phase means (planes 1..nx) and of compute_phim_heat_ratio (plane
// nx+1), then nlphif = wy ff2rc(v phi) + wz ff2rc(w phi) (planes 1..nx; the
// plane nx+1 value is set by k_scal_xflux).
// LAST: plane nx+1 (its sums weigh phi with u for heat_ratio_new); a
// functor of its own, so that the other planes' functor carries no branch
// for it (with the branch, the compiler kept the kernel arguments in
// scratch memory)
template <typename T, bool LAST>
struct OpPhiC2R {
  const KArgs<T>& A;
  const ScalStep<T>& S;
  int p;
  double s0, s1, s2;
  bool own;  // the workgroup that writes fdd and the plane sums
  __device__ C2<T> load(int ky, int kz) {
    C2<double> v = S.phif[sss(A, p, ky, kz)];
    return mk((T)v.x, (T)v.y);
  }
  __device__ void one(int y, int z, T v) {
    const size_t o = sps(A, p + 1, y, z);
    if (!own) {  // the w phi workgroup: its own copy
      S.ph2[o] = v;
      return;
    }
    S.ph[o] = v;
    const int pu = LAST ? 0 : p;
    const T u1 = A.ubcp[ridx(A, 0, pu, y, z)];
    S.fdd[o] = S.upwind ? (T)0 : u1 * v;
    const bool fl = A.mask[((size_t)pu * A.mz + z) * A.my + y] != 0;
    if (!LAST) {
      if (fl) s0 += (double)v;
      else s1 += (double)v;
    } else if (fl) {
      s0 += (double)u1 * (double)v;
      s1 += (double)u1;
      s2 += 1.0;
    }
  }
  __device__ void store2(int, int y, int z, T a, T b) {
    one(y, z, a);
    one(y + 1, z, b);
  }
};
template <typename T, int ST>
struct OpPhiProd {  // ST 0: ff2rc(v phi) times wy into nlphif, 1: ff2rc(w phi) times wz into nlz
  const KArgs<T>& A;
  const ScalStep<T>& S;
  int p;
  __device__ T val(int y, int z) const {
    return A.ubcp[ridx(A, ST + 1, p, y, z)] * (ST == 0 ? S.ph : S.ph2)[sps(A, p + 1, y, z)];
  }
  __device__ void load2(int y, int z, T& a, T& b) {
    a = val(y, z);
    b = val(y + 1, z);
  }
  __device__ void store(int, int ky, int kz, C2<T> v) {
    if (ST == 0) S.nlphif[sss(A, p, ky, kz)] = cikd((double)A.wy[ky], dbl(v));
    else S.nlz[sss(A, p, ky, kz)] = cikd((double)A.wz[kz], dbl(v));
  }
};
// ff2cr of plane p of phif into ph, with fdd and the plane sums when own
template <typename T, int SH, bool LAST>
__device__ void phi_plane(const KArgs<T>& A, const ScalStep<T>& S, int p, bool own, C2<T>* buf,
                          unsigned char* smem) {
  OpPhiC2R<T, LAST> op{A, S, p, 0.0, 0.0, 0.0, own};
  plane_c2r_sh<SH, 0>(buf, A.G, A.tb, op, smem);
  if (own) {
    double sm[3] = {op.s0, op.s1, op.s2};
    block_sums_w<3>(sm, (double*)smem);
    if (threadIdx.x == 0)
      for (int q = 0; q < 3; ++q) S.part[(size_t)p * 4 + q] = sm[q];
  }
}
template <typename T, int SH>
__global__ void __launch_bounds__(1024, (PlaneOcc<T, SH>::WPE)) k_scal_phi(KArgs<T> A, ScalStep<T> S) {
  extern __shared__ __align__(16) unsigned char smem[];
  // grid (nx + 1, 2): workgroup row 0 transforms phi into ph, keeps fdd and
  // the plane sums and forms the v phi term; row 1 transforms phi again into
  // its own copy ph2 and forms the w phi term, so a plane's three transforms
  // run as two in each of two workgroups
  const int p = blockIdx.x, half = blockIdx.y;
  if (half == 1 && p == A.nx) return;
  C2<T>* buf = plane_slot(A);
  if (p == A.nx) {
    phi_plane<T, SH, true>(A, S, p, true, buf, smem);
    return;
  }
  phi_plane<T, SH, false>(A, S, p, half == 0, buf, smem);
  __syncthreads();  // this workgroup's phi stores are visible to it
  if (half == 0) {
    OpPhiProd<T, 0> o0{A, S, p};
    plane_r2c_sh<SH, 0>(buf, A.G, A.tb, o0, smem);
  } else {
    OpPhiProd<T, 1> o1{A, S, p};
    plane_r2c_sh<SH, 0>(buf, A.G, A.tb, o1, smem);
  }
}

// heat_ratio_new = 1/phim(nx+1), phim = sum u phi / (um count), um = sum u /
// count over the fluid of plane nx+1 (nl_allphi.F90:448-483); the phase
// means of phi (:380-419).  One workgroup of 256 lanes; the plane sums in a
// fixed order.
template <typename T>
__global__ void __launch_bounds__(256) k_scal_ratio(KArgs<T> A, ScalStep<T> S) {
  __shared__ double sh[2 * 16];
  double v[2] = {0.0, 0.0};
  for (int p = threadIdx.x; p < A.nx; p += blockDim.x) {
    v[0] += S.part[(size_t)p * 4];
    v[1] += S.part[(size_t)p * 4 + 1];
  }
  block_sums_w<2>(v, sh);
  if (threadIdx.x != 0) return;
  const double* q = S.part + (size_t)A.nx * 4;
  const double um = q[1] / q[2];
  const double phim = q[0] / (um * q[2]);
  S.ratio[1] = 1.0 / phim;
  S.ratio[4] = v[0] / A.st->s.count_fluid;
  S.ratio[5] = A.st->s.count_solid > 0 ? v[1] / A.st->s.count_solid : 0.0;
}

// ---------------------------------------------------------------------------
// k_scal_xflux, per reference plane i = p+1 in 1..nx (nl_allphi.F90:176-375,
// 393-410): the x-derivative of the x-flux, central (u phi differenced, with
// fdderiv(0) = fdderiv(nx) heat_ratio) or upwind (MUSCL with central slopes
// and the Rusanov flux at faces 0..nx, the end faces from the twisted wrap),
// added to nlphif and negated; plane nx+1 = plane 1 / heat_ratio.
template <typename T, bool UP>
struct OpPhiFlux {
  const KArgs<T>& A;
  const ScalStep<T>& S;
  int p;
  double h;
  __device__ double PH(int i, int y, int z) const { return (double)S.ph[sps(A, i, y, z)]; }
  __device__ double U1(int i, int y, int z) const {  // reference plane i in 1..nx+1
    return (double)A.ubcp[ridx(A, 0, i > A.nx ? i - 1 - A.nx : i - 1, y, z)];
  }
  // slopes of cell c in 1..nx
  __device__ double sphi(int c, int y, int z) const {
    int im1 = c - 1;
    double pfm = 1.0;
    if (im1 < 1) {
      im1 = A.nx;
      pfm = h;
    }
    return (PH(c + 1, y, z) - pfm * PH(im1, y, z)) / (2.0 * S.dx);
  }
  __device__ double su(int c, int y, int z) const {
    int im1 = c - 1 < 1 ? A.nx : c - 1;
    return (U1(c + 1, y, z) - U1(im1, y, z)) / (2.0 * S.dx);
  }
  // Rusanov flux at face i + 1/2, i in 0..nx
  __device__ double face(int i, int y, int z) const {
    const int nx = A.nx;
    const double hd = S.dx / 2.0;
    double um, Um, up, Up;
    if (i == 0) {
      um = U1(nx, y, z) + ((U1(1, y, z) - U1(nx - 1, y, z)) / (2.0 * S.dx)) * hd;
      Um = PH(nx, y, z) * h + ((PH(1, y, z) - PH(nx - 1, y, z) * h) / (2.0 * S.dx)) * hd;
    } else {
      um = U1(i, y, z) + su(i, y, z) * hd;
      Um = PH(i, y, z) + sphi(i, y, z) * hd;
    }
    if (i == nx) {
      up = U1(1, y, z) - ((U1(2, y, z) - U1(nx, y, z)) / (2.0 * S.dx)) * hd;
      Up = PH(1, y, z) / h - ((PH(2, y, z) / h - PH(nx, y, z)) / (2.0 * S.dx)) * hd;
    } else {
      up = U1(i + 1, y, z) - su(i + 1, y, z) * hd;
      Up = PH(i + 1, y, z) - sphi(i + 1, y, z) * hd;
    }
    const double ms = fmax(fabs(um), fabs(up));
    const double f = 0.5 * (up * Up) + 0.5 * (um * Um);
    return f - 0.5 * ms * (Up - Um);
  }
  __device__ T val(int y, int z) const {
    const int i = p + 1;
    if constexpr (UP) {  // the faces are in fdd (k_scal_faces)
      return (T)(((double)S.fdd[sps(A, i, y, z)] - (double)S.fdd[sps(A, i - 1, y, z)]) / S.dx);
    } else {
      const double fm = i == 1 ? (double)S.fdd[sps(A, A.nx, y, z)] * h : (double)S.fdd[sps(A, i - 1, y, z)];
      return (T)(((double)S.fdd[sps(A, i + 1, y, z)] - fm) / (2.0 * S.dx));
    }
  }
  __device__ void load2(int y, int z, T& a, T& b) {
    a = val(y, z);
    b = val(y + 1, z);
  }
  __device__ void store(int, int ky, int kz, C2<T> v) {
    C2<double>* o = S.nlphif + sss(A, p, ky, kz);
    C2<double> s = (*o + S.nlz[sss(A, p, ky, kz)]) + dbl(v);
    C2<double> r = mk(-s.x, -s.y);
    *o = r;
    if (p == 0) S.nlphif[sss(A, A.nx, ky, kz)] = mk(r.x / h, r.y / h);
  }
};
// the upwind flux of every face i + 1/2, i = 0..nx, into fdd (the
// reference's fdderiv(0:nx), nl_allphi.F90:325-340), one lane a point
template <typename T>
__global__ void __launch_bounds__(256) k_scal_faces(KArgs<T> A, ScalStep<T> S) {
  const size_t plane = (size_t)A.my * A.mz;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)(A.nx + 1) * plane) return;
  const int i = (int)(q / plane);
  const int r = (int)(q - (size_t)i * plane), z = r / A.my, y = r - z * A.my;
  OpPhiFlux<T, true> op{A, S, 0, S.ratio[0]};
  S.fdd[q] = (T)op.face(i, y, z);
}
template <typename T, int SH, bool UP>
__global__ void __launch_bounds__(1024) k_scal_xflux(KArgs<T> A, ScalStep<T> S) {
  extern __shared__ __align__(16) unsigned char smem[];
  OpPhiFlux<T, UP> op{A, S, (int)blockIdx.x, S.ratio[0]};
  plane_r2c_sh<SH, 0>(plane_slot(A), A.G, A.tb, op, smem);
}

// ---------------------------------------------------------------------------
// k_scal_phys, grid (planes 1..nx+1, 6): the physical nl and onl of phi
// (nlphi.F90:119-132) and calc_diffn's diffn and grad phi (bc_scalar.F90:
// 935-1033), whose x differences wrap with the heat ratio.  onlphif is the
// previous step's nlphif, so its physical field is the previous step's nl
// when nothing changed the fields in between (grid (planes, 5), skip_onl).
template <typename T>
struct OpPhiPhys {
  const KArgs<T>& A;
  const ScalStep<T>& S;
  int p, kind;
  double h, g;
  __device__ C2<double> P(int i, int ky, int kz) const { return S.phif[sss(A, i - 1, ky, kz)]; }
  __device__ C2<T> load(int ky, int kz) {
    C2<double> v;
    if (kind == 0) {
      v = S.nlphif[sss(A, p, ky, kz)];
    } else if (kind == 1) {
      v = S.onlphif[sss(A, p, ky, kz)];
    } else {
      const int nx = A.nx, i = p + 1;
      const C2<double> p0 = P(i, ky, kz);
      C2<double> pm, pp;
      if (i == 1) {
        pm = cmulr(P(nx, ky, kz), h);
        pp = P(2, ky, kz);
      } else if (i == nx) {
        pm = P(i - 1, ky, kz);
        pp = cdivr(P(1, ky, kz), h);
      } else if (i == nx + 1) {
        pm = P(i - 1, ky, kz);
        pp = cdivr(P(2, ky, kz), h);
      } else {
        pm = P(i - 1, ky, kz);
        pp = P(i + 1, ky, kz);
      }
      if (kind == 2) {
        C2<double> f = cmulr((pm - cmulr(p0, 2.0)) + pp, 1.0 / (S.dx * S.dx));
        f = f - cmulr(p0, (double)A.w2[(size_t)kz * A.my2 + ky]);
        v = cmulr(f, g);
      } else if (kind == 3) {
        v = cdivr(pp - pm, 2.0 * S.dx);
      } else if (kind == 4) {
        v = cikd((double)A.wy[ky], p0);
      } else {
        v = cikd((double)A.wz[kz], p0);
      }
    }
    return mk((T)v.x, (T)v.y);
  }
  __device__ void store2(int, int y, int z, T a, T b) {
    T* dst = kind == 0 ? S.nlp : (kind == 1 ? S.onp : (kind == 2 ? S.df : S.g + (size_t)(kind - 3) * S.np));
    const size_t o = sps(A, p + 1, y, z);
    dst[o] = a;
    dst[o + 1] = b;
  }
};
template <typename T, int SH>
__global__ void __launch_bounds__(1024) k_scal_phys(KArgs<T> A, ScalStep<T> S) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int kind = (int)blockIdx.y + (S.skip_onl && blockIdx.y >= 1 ? 1 : 0);
  OpPhiPhys<T> op{A, S, (int)blockIdx.x, kind, S.ratio[0], S.ratio[3]};
  plane_c2r_sh<SH, 0>(plane_slot(A), A.G, A.tb, op, smem);
}

// ff2rc of fr, planes 1..nx (bc_scalar.F90:328-346)
template <typename T>
struct OpPhiFr {
  const KArgs<T>& A;
  const ScalStep<T>& S;
  int p;
  __device__ void load2(int y, int z, T& a, T& b) {
    const size_t o = sps(A, p + 1, y, z);
    a = S.fr[o];
    b = S.fr[o + 1];
  }
  __device__ void store(int, int ky, int kz, C2<T> v) { S.ffphi[sss(A, p, ky, kz)] = dbl(v); }
};
template <typename T, int SH>
__global__ void __launch_bounds__(1024) k_scal_ff(KArgs<T> A, ScalStep<T> S) {
  extern __shared__ __align__(16) unsigned char smem[];
  OpPhiFr<T> op{A, S, (int)blockIdx.x};
  plane_r2c_sh<SH, 0>(plane_slot(A), A.G, A.tb, op, smem);
}

// ---------------------------------------------------------------------------
// interpolate_phidata (bc_scalar.F90:1036-1153), order-2 stencil: x nodes
// below 1 wrap up a period with the factor heat_ratio (heat_ratio_old for
// onl), above nx+1 down a period with 1/heat_ratio; phi of plane nx+1 is
// phi(1)/heat_ratio (nl_allphi.F90:411).
struct PhiSten {
  int ib, jb, kb;
  double w[8];  // (k*2+j)*2+i
};
__device__ inline void phi_sten(PhiSten& s, const int pc[3], const double pos[3]) {
  s.ib = pc[0];
  s.jb = pc[1];
  s.kb = pc[2];
  const double zx = pos[0] - s.ib, zy = pos[1] - s.jb, zz = pos[2] - s.kb;
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i)
        s.w[(k * 2 + j) * 2 + i] = (i ? zx : 1.0 - zx) * (j ? zy : 1.0 - zy) * (k ? zz : 1.0 - zz);
}
// a y or z node index into 1..n (a full period wrap, in bounds for any input)
__device__ inline int wrap1y(int j, int n) {
  int r = (j - 1) % n;
  return (r < 0 ? r + n : r) + 1;
}
// an x node index the one-period rules left outside 0..nx+1 (a centre more
// than a period outside the box): folded into 1..nx so that no access leaves
// the arrays; the reference has no such case
__device__ inline int xfold(int ii, int nx) {
  if (ii >= 0 && ii <= nx + 1) return ii;
  return wrap1y(ii, nx);
}
template <typename T>
__device__ inline void interp_phi(const KArgs<T>& A, const ScalStep<T>& S, const PhiSten& st, bool inner, double h,
                                  double ho, double* ul, double* nll, double* onll, double* dfll) {
  double vs[8], ns[8], os[8], ds[8];
  const int nx = A.nx;
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        int ii = st.ib + i;
        const int y = wrap1y(st.jb + j, A.my) - 1, z = wrap1y(st.kb + k, A.mz) - 1;
        double pf = 1.0, pfo = 1.0;
        if (ii < 1) {
          ii = ii + nx;
          pf = h;
          pfo = ho;
        } else if (ii > nx + 1) {
          ii = ii - nx;
          pf = 1.0 / h;
        }
        ii = xfold(ii, nx);
        const int q = (k * 2 + j) * 2 + i;
        const double ph = ii == nx + 1 ? (double)S.ph[sps(A, 1, y, z)] / h : (double)S.ph[sps(A, ii, y, z)];
        vs[q] = ph * pf;
        if (inner) {
          ns[q] = (double)S.nlp[sps(A, ii, y, z)] * pf;
          os[q] = (double)S.onp[sps(A, ii, y, z)] * pfo;
          ds[q] = (double)S.df[sps(A, ii, y, z)] * pf;
        }
      }
  double s = 0.0;
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) s = s + vs[(k * 2 + j) * 2 + i] * st.w[(k * 2 + j) * 2 + i];
  *ul = s;
  if (inner) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        for (int k = 0; k < 2; ++k) {
          const int q = (k * 2 + j) * 2 + i;
          a = a + ns[q] * st.w[q];
          b = b + os[q] * st.w[q];
          c = c + ds[q] * st.w[q];
        }
    *nll = a;
    *onll = b;
    *dfll = c;
  }
}
__device__ inline int cellof_d(double x) { return x < 0.0 ? (int)(x - 1.0) : (int)x; }

// calc_bndrev_data (bc_scalar.F90:385-556), one workgroup per body: the
// surface flux sum of n . grad phi, and with DOBND the surface forcing
// cf (phi - phisurf) dA1 dx of each point for k_scal_spread (its x nodes
// wrap into 1..nx+1; plane nx+1 is not transformed).
template <typename T>
__global__ void __launch_bounds__(256) k_scal_bnd(KArgs<T> A, Particles P, ScalStep<T> S) {
  __shared__ double sh[64];
  const int m = blockIdx.x;
  const int nx = A.nx, nb = P.nbnd;
  const DevState* st = A.st;
  const double h = S.ratio[0], dx = st->dx, dx3 = st->dx3, cf = st->s.cf;
  const double rb = P.rad[m];
  const double da1 = 4.0 * 3.14159265358979323846264338327950288 * (rb * dx) * (rb * dx) / (double)(float)nb;
  double fl = 0.0;
  for (int l = threadIdx.x; l < nb; l += blockDim.x) {
    double xl[3];
    int is[3];
    for (int n = 0; n < 3; ++n) {
      xl[n] = P.xc[(size_t)n * P.nbody + m] + P.bnd[(size_t)l * 3 + n] * rb;
      is[n] = cellof_d(xl[n]);
    }
    if (!(1 <= is[0] && is[0] <= nx)) {
      const int c0 = is[0];
      const double x0 = xl[0];
      if (c0 < 1) {
        is[0] = nx + c0;
        xl[0] = nx + x0;
      }
      if (c0 >= nx + 1) {
        is[0] = c0 - nx;
        xl[0] = x0 - nx;
      }
    }
    PhiSten stn;
    phi_sten(stn, is, xl);
    double phil, dum;
    interp_phi(A, S, stn, false, h, 1.0, &phil, &dum, &dum, &dum);
    double tempor = 0.0;
    if (S.dobnd) {
      tempor = ((cf * (phil - S.phisurf[m])) * da1 * 1.0) * dx;
      // the share of x node nx+1 (a point in cell nx) stays on a plane that
      // is not transformed: the second set carries it as zero
      double* o = S.spt + ((size_t)P.nbody * nb + (size_t)m * nb + l) * 4;
      o[0] = tempor;
      o[1] = 0.0;
      o[2] = is[0] == nx ? 0.0 : tempor;
      o[3] = 0.0;
    }
    double gb[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        for (int k = 0; k < 2; ++k) {
          int ii = stn.ib + i;
          const int y = wrap1y(stn.jb + j, A.my) - 1, z = wrap1y(stn.kb + k, A.mz) - 1;
          double pf = 1.0;
          if (ii < 1) {
            ii = ii + nx;
            pf = h;
          }
          if (ii > nx + 1) {
            ii = ii - nx;
            pf = 1.0 / h;
          }
          ii = xfold(ii, nx);
          const double w = stn.w[(k * 2 + j) * 2 + i];
          for (int n = 0; n < 3; ++n) gb[n] = gb[n] + (double)S.g[(size_t)n * S.np + sps(A, ii, y, z)] * pf * w;
        }
    fl = fl + (gb[0] * P.bnd[(size_t)l * 3] + gb[1] * P.bnd[(size_t)l * 3 + 1] + gb[2] * P.bnd[(size_t)l * 3 + 2]);
  }
  double v[1] = {fl};
  block_sums_w<1>(v, sh);
  if (threadIdx.x == 0) S.flux_body[m] = v[0];
}

// calc_inner_rev_data (bc_scalar.F90:558-932), one lane per (body, pair):
// a pair whose inner point lies in cell nx is evaluated twice, the second
// time one period down (cell 0); its forcing cf (phi_i - phi_ref) - c3 nl -
// c4 onl - (c1 + c2) diffn, times dA2 dx, goes to the stencil's solid nodes
// with the fluid nodes' share redistributed evenly over them (to every node
// with include_frmeanfluid).  Of the two evaluations, only the first's x node
// nx and the second's x node 1 land on transformed planes: the pair's row
// holds (tempor, force_dist) of each, and k_scal_spread takes the second for
// the stencil's upper x nodes.
template <typename T>
__global__ void __launch_bounds__(256) k_scal_rev(KArgs<T> A, Particles P, ScalStep<T> S, int rks) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P.nbody * P.nbnd) return;
  const int m = q / P.nbnd, l = q - m * P.nbnd;
  if (!P.if_rev[(size_t)m * P.nbnd + l]) return;
  const int nx = A.nx;
  const DevState* st = A.st;
  const double h = S.ratio[0], ho = S.ratio[2], dx = st->dx, dx3 = st->dx3, cf = st->s.cf;
  const double c1 = st->coef[rks - 1][0], c2 = st->coef[rks - 1][1], c3 = st->coef[rks - 1][2],
               c4 = st->coef[rks - 1][3];
  const double rb = P.rad[m], ri = P.radi[m], ro = P.rado[m];
  const double da2 =
      4.0 * 3.14159265358979323846264338327950288 * (ri * dx) * (ri * dx) / (double)(float)P.nrpr_active[m];
  double xli[3], bn[3];
  int iii[3];
  for (int n = 0; n < 3; ++n) {
    bn[n] = P.bnd[(size_t)l * 3 + n];
    xli[n] = P.xc[(size_t)n * P.nbody + m] + bn[n] * ri;
    iii[n] = cellof_d(xli[n]);
  }
  int ct = iii[0];
  double xlt = xli[0];
  if (!(1 <= ct && ct <= nx)) {
    if (iii[0] < 1) {
      ct = nx + iii[0];
      xlt = nx + xli[0];
    }
    if (iii[0] >= nx + 1) {
      ct = iii[0] - nx;
      xlt = xli[0] - nx;
    }
  }
  const int nev = ct == nx ? 2 : 1;
  for (int ev = 1; ev <= nev; ++ev) {
    if (ev == 1) {
      iii[0] = ct;
      xli[0] = xlt;
    } else if (iii[0] >= nx) {
      iii[0] -= nx;
      xli[0] -= nx;
    }
    double xlo[3];
    int io[3];
    for (int n = 0; n < 3; ++n) {
      xlo[n] = xli[n] + bn[n] * (ro - ri);
      io[n] = cellof_d(xlo[n]);
    }
    PhiSten so, si;
    phi_sten(so, io, xlo);
    phi_sten(si, iii, xli);
    double philo, phili, nlphil, onlphil, dfll, dum;
    interp_phi(A, S, so, false, h, ho, &philo, &dum, &dum, &dum);
    interp_phi(A, S, si, true, h, ho, &phili, &nlphil, &onlphil, &dfll);
    philo = S.phisurf[m] * (ro - ri) - philo * (rb - ri);
    philo = philo / (ro - rb);
    const double phil = cf * (-philo + phili);
    double tempor = 0.0 + phil;
    tempor = tempor - c3 * nlphil - c4 * onlphil;
    tempor = tempor - (c2 + c1) * dfll;
    tempor = tempor * da2 * 1.0 * dx;
    int count_fl = 0;
    double force_fl = 0.0;
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
          const int ii = xfold(si.ib + i, nx);
          const int y = wrap1y(si.jb + j, A.my) - 1, z = wrap1y(si.kb + k, A.mz) - 1;
          if (fluid_at(A, ii, y, z)) {
            count_fl += 1;
            force_fl = force_fl + si.w[(k * 2 + j) * 2 + i] * tempor / dx3;
          }
        }
    const double force_dist = count_fl < 8 ? force_fl / (double)(8 - count_fl) : 0.0;
    double* o = S.spt + ((size_t)m * P.nbnd + l) * 4;
    if (ev == 1) {
      o[0] = tempor;
      o[1] = force_dist;
    }
    o[2] = tempor;  // a single evaluation: its values for the upper nodes too
    o[3] = force_dist;
  }
}

// The scalar forcing on the grid from the fluid's node-sorted spreading
// records (the same stencils, modulo the x period): one lane per target node
// adds its records in their sorted order, each (w tempor)/dx^3 plus the
// redistributed share when the record carries it, with the upper-x-node
// records taking the pair's second evaluation.  No atomics: the sums are
// deterministic.  rdist bit 0: redistributed share; bit 1: upper x node.
template <typename T>
__global__ void __launch_bounds__(256) k_scal_spread(KArgs<T> A, ScalStep<T> S, const int* bt, const int* rs,
                                                     int nrange, const int* tnode, const int* toff, const int* rpt,
                                                     const double* rw, const uint8_t* rdist) {
  __shared__ double cv[SPREAD_CAP];
  const int w = xcd_linear(blockIdx.x, gridDim.x);
  if (w >= nrange) return;
  const int t0 = bt[w], t1 = bt[w + 1];
  const int r0 = rs[w], r1 = rs[w + 1];  // toff[t0], toff[t1]
  const double dx3 = A.st->dx3;
  // each record's value staged in LDS (a range holds at most SPREAD_CAP
  // records), then one lane per node sums its records in record order
  for (int r = r0 + (int)threadIdx.x; r < r1; r += 256) {
    const uint8_t f = rdist[r];
    const double* o = S.spt + (size_t)rpt[r] * 4 + ((f & 2) ? 2 : 0);
    cv[r - r0] = (rw[r] * o[0]) / dx3 + ((f & 1) ? o[1] : 0.0);
  }
  __syncthreads();
  const size_t plane = (size_t)A.my * A.mz;
  for (int t = t0 + (int)threadIdx.x; t < t1; t += 256) {
    double acc = 0.0;
    for (int r = toff[t]; r < toff[t + 1]; ++r) acc = acc + cv[r - r0];
    S.fr[(size_t)tnode[t] + plane] = (T)acc;
  }
}
