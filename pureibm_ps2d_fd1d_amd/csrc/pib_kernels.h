// pib_kernels.h -- device kernels of the per-time-step fluid solve.
//
// Device layout (x planes slowest, so every (y,z) plane is one contiguous
// block for the plane FFT, and the per-(ky,kz)-mode x-lines of the
// tridiagonal solves are read mode-major, coalesced across threads):
//   spectral  S[c][i][kz][ky]  complex, i = 0..nx-1 (the reference's plane
//             nx+1 is the periodic copy of plane 1 and is not stored)
//   physical  R[c][i][z][y]    real
//   mask      M[i][z][y]       1 = fluid (fluid_atijk)
// Grid positions follow the reference: velocity node i (1-based) sits at
// x = i, pressure node i at x = i + 1/2; particle centres xc are in grid
// units.
#pragma once
#include "pib_fft2.h"
#include "../../include/pureibm_gpu.h"

#define PIB_SMALL_NUMBER ((double)1.0E-35f)

// Device-resident scalar state: the reference's global_data scalars plus
// the step-internal values the kernels hand to each other.
struct DevState {
  pib_scalars s;
  double coef[3][4];
  double norm_factor;
  double dx, vis, cfl_max, ucharmod, char_length, dia, voldom, dx3, dr;
  int nx, my, mz, nbody, nbnd, adaptive, soft_start, set_mpg, include_frmeanfluid, itrmax;
  int revernorm;
  int status;  // 0 ok, else a PIB_ERR_* code raised on the device
  int move_particles;
  // the drag-error history (boundary_condition.F90:828-833): its length
  // nerr_steps and the ring slot the next step's ferror goes to
  int nerr, ferr_pos, pad_;
  double rhos, rhof;
};

template <typename T>
struct KArgs {
  int nx, my, mz, my2, hy;
  T dx, dx2, vis;
  const T* wy;
  const T* wz;
  const T* w2;
  PlaneGeom G;
  PlaneTabs<T> tb;
  C2<T>* u;
  C2<T>* p;
  C2<T>* nl;
  C2<T>* onl;
  C2<T>* ff;
  // the six product spectra of the central form_nl, [q][i][kz][ky] with q:
  // 0 uu, 1 uv, 2 uw, 3 vv, 4 vw, 5 ww
  C2<T>* nlp;
  T* ubcp;
  T* pbcp;
  T* nlbc;
  T* onlbc;
  T* diffn;
  T* omega;
  T* fr;
  T* ppr;
  // the inner-point combination of the reversal forcing, one field for the
  // four that calc_inner_reversal_forcing interpolates at the same point:
  // qin = cf ubcp - c3 nlbc - c4 onlbc - (c1 + c2) diffn
  T* qin;
  // the stage's cf, umean[3] and coefficients c1..c4, written by the step's
  // first scalar kernel; read where the device state may be rewritten beside
  const double* sconst;
  const uint8_t* mask;
  DevState* st;
  double* part;  // per-workgroup partial results
  // planes that do not fit the LDS of a workgroup live in HBM: one slot of
  // pslot entries per workgroup of a launch (null when planes fit the LDS)
  C2<T>* pscr;
  size_t pslot;
  int rks;
  int xcd_map;   // map workgroups to planes so neighbouring planes share an XCD
};

// This workgroup's plane slot in HBM, or null when planes live in LDS.
template <typename T>
__device__ __forceinline__ C2<T>* plane_slot(const KArgs<T>& A) {
  if (!A.pscr) return nullptr;
  return A.pscr + (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * A.pslot;
}
template <typename T>
__device__ __forceinline__ size_t sidx(const KArgs<T>& A, int c, int i, int ky, int kz) {
  return (((size_t)c * A.nx + i) * A.mz + kz) * A.my2 + ky;
}
template <typename T>
__device__ __forceinline__ size_t ridx(const KArgs<T>& A, int c, int i, int y, int z) {
  return (((size_t)c * A.nx + i) * A.mz + z) * A.my + y;
}
__device__ __forceinline__ int wrapn(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Plane index of this workgroup.  With xcd_map, the round-robin dispatch of
// consecutive workgroups over the 8 XCDs is undone so that one XCD works on a
// contiguous run of planes (their x-neighbours then hit the same L2).
template <typename T>
__device__ __forceinline__ int plane_of(const KArgs<T>& A) {
  int b = blockIdx.x;
  if (A.xcd_map && (A.nx % 8) == 0) {
    int per = A.nx / 8;
    return (b % 8) * per + b / 8;
  }
  return b;
}

// Work index of a workgroup in a 1-D grid of nblk workgroups (nblk a multiple
// of 8).  The hardware deals consecutive workgroups round-robin to the 8
// XCDs; this mapping gives each XCD a contiguous run of work indices, so the
// workgroups that gather around one sphere share one L2.
__device__ __forceinline__ int xcd_linear(int b, int nblk) {
  const int per = nblk >> 3;
  return (b & 7) * per + (b >> 3);
}

// ---------------------------------------------------------------------------
// deterministic block reductions (power-of-two block sizes)
// ---------------------------------------------------------------------------
__device__ inline double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}
__device__ inline double block_max(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// Block-wide sums of K values per lane in one deterministic tree; any block
// size; sh holds K * blockDim.x doubles.
template <int K>
__device__ inline void block_sums(double* v, double* sh) {
  const int bd = blockDim.x, t = threadIdx.x;
  int n2 = 1;
  while (n2 < bd) n2 <<= 1;
  for (int q = 0; q < K; ++q) sh[q * bd + t] = v[q];
  __syncthreads();
  for (int s = n2 / 2; s > 0; s >>= 1) {
    if (t < s && t + s < bd)
      for (int q = 0; q < K; ++q) sh[q * bd + t] += sh[q * bd + t + s];
    __syncthreads();
  }
  for (int q = 0; q < K; ++q) v[q] = sh[q * bd];
  __syncthreads();
}
// Block sums for blocks whose size is a multiple of 64 (at most 1024).  Each
// wavefront reduces its 64 values with DPP lane moves (VALU, no LDS
// traffic): within each quad (xor 1, xor 2), then across the quads of each
// 16-lane row (rotate 4, rotate 8); the four row totals are read from lanes
// 0, 16, 32, 48 and added as (r0 + r1) + (r2 + r3).  The (at most 16) wave
// totals of value q then sit in one 16-lane row and are added the same way.
// Every association is fixed, so results repeat bit for bit.  sh holds
// K * 16 doubles.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
// DPP controls: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_ror:4, row_ror:8
#define PIB_DPP_QXOR1 0xB1
#define PIB_DPP_QXOR2 0x4E
#define PIB_DPP_ROR4 0x124
#define PIB_DPP_ROR8 0x128
// sum / max over each 16-lane row; every lane of the row gets a total, lane 0
// of the row the one with the documented association
__device__ __forceinline__ double row_sum(double v) {
  v = v + dpp_move<PIB_DPP_QXOR1>(v);
  v = v + dpp_move<PIB_DPP_QXOR2>(v);
  v = v + dpp_move<PIB_DPP_ROR4>(v);
  return v + dpp_move<PIB_DPP_ROR8>(v);
}
__device__ __forceinline__ double row_max(double v) {
  v = fmax(v, dpp_move<PIB_DPP_QXOR1>(v));
  v = fmax(v, dpp_move<PIB_DPP_QXOR2>(v));
  v = fmax(v, dpp_move<PIB_DPP_ROR4>(v));
  return fmax(v, dpp_move<PIB_DPP_ROR8>(v));
}
__device__ inline double wave_sum(double v) {
  v = row_sum(v);
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}
__device__ inline double wave_max(double v) {
  v = row_max(v);
  return fmax(fmax(lane_value(v, 0), lane_value(v, 16)), fmax(lane_value(v, 32), lane_value(v, 48)));
}
template <int K, bool MAX>
__device__ inline void block_reduce_w(double* v, double* sh) {
  const int t = threadIdx.x, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < K; ++q) v[q] = MAX ? wave_max(v[q]) : wave_sum(v[q]);
  if (nw == 1) return;
  if ((t & 63) == 0)
#pragma unroll
    for (int q = 0; q < K; ++q) sh[q * 16 + (t >> 6)] = v[q];
  __syncthreads();
  // slot idx = q*16 + w: row q of 16 lanes holds the wave totals of value q
  constexpr int NCH = (K * 16 + 63) / 64;
  double r[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    const int idx = ch * (int)blockDim.x + t;
    double x = MAX ? -__builtin_huge_val() : 0.0;
    if (idx < K * 16 && (idx & 15) < nw) x = sh[idx];
    r[ch] = MAX ? row_max(x) : row_sum(x);
  }
  __syncthreads();
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    const int idx = ch * (int)blockDim.x + t;
    if (idx < K * 16 && (idx & 15) == 0) sh[idx >> 4] = r[ch];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) v[q] = sh[q];
  __syncthreads();
}
template <int K>
__device__ inline void block_sums_w(double* v, double* sh) {
  block_reduce_w<K, false>(v, sh);
}
template <int K>
__device__ inline void block_maxes_w(double* v, double* sh) {
  block_reduce_w<K, true>(v, sh);
}
// deterministic max over any block size
__device__ inline double block_max_any(double v, double* sh) {
  const int bd = blockDim.x, t = threadIdx.x;
  int n2 = 1;
  while (n2 < bd) n2 <<= 1;
  sh[t] = v;
  __syncthreads();
  for (int s = n2 / 2; s > 0; s >>= 1) {
    if (t < s && t + s < bd) sh[t] = fmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// ===========================================================================
// c2r ops (physical fields from spectral ones)
// ===========================================================================
// compute_new_timestep (dependent_functions.F90:1088-1122): ubcp = c2r(u) +
// umean, with the fluid/solid/total sums and max|u| of the same loop.
template <typename T>
struct OpUbcp {
  const KArgs<T>& A;
  int c, i;
  double sf, ss, sa, mx, um;
  __device__ OpUbcp(const KArgs<T>& a, int c_, int i_) : A(a), c(c_), i(i_), sf(0), ss(0), sa(0), mx(PIB_SMALL_NUMBER) {
    um = A.st->s.umean[c];
  }
  __device__ C2<T> load(int ky, int kz) { return A.u[sidx(A, c, i, ky, kz)]; }
  __device__ void one(int y, int z, T v) {
    T w = v + (T)um;
    A.ubcp[ridx(A, c, i, y, z)] = w;
    double d = (double)w;
    if (A.mask[((size_t)i * A.mz + z) * A.my + y]) sf += d;
    else ss += d;
    sa += d;
    mx = fmax(mx, fabs(d));
  }
  __device__ void store2(int, int y, int z, T a, T b) {
    one(y, z, a);
    one(y + 1, z, b);
  }
};
// plain copy-out with an optional additive device scalar
template <typename T>
struct OpPlain {
  const C2<T>* src;
  T* dst;
  const KArgs<T>& A;
  size_t sbase, dbase;
  T add;
  __device__ C2<T> load(int ky, int kz) { return src[sbase + (size_t)kz * A.my2 + ky]; }
  __device__ void store2(int, int y, int z, T a, T b) {
    size_t o = dbase + (size_t)z * A.my + y;
    dst[o] = a + add;
    dst[o + 1] = b + add;
  }
};
// calc_visc (boundary_condition.F90:1767-1815):
// diffn = c2r( vis*((1/dx2)*(u(i-1) - 2u(i) + u(i+1)) - w2 u(i)) )
template <typename T>
struct OpVisc {
  const KArgs<T>& A;
  int c, i;
  __device__ C2<T> load(int ky, int kz) {
    int im = wrapn(i - 1, A.nx), ip = wrapn(i + 1, A.nx);
    C2<T> um = A.u[sidx(A, c, im, ky, kz)], uc = A.u[sidx(A, c, i, ky, kz)], up = A.u[sidx(A, c, ip, ky, kz)];
    C2<T> d2 = (um - cscal((T)2, uc)) + up;
    C2<T> f = cscal((T)1 / A.dx2, d2);
    f = f - cscal(A.w2[(size_t)kz * A.my2 + ky], uc);
    return cscal(A.vis, f);
  }
  __device__ void store2(int, int y, int z, T a, T b) {
    size_t o = ridx(A, c, i, y, z);
    A.diffn[o] = a;
    A.diffn[o + 1] = b;
  }
};
// compute_omega (boundary_condition.F90:1648-1765); omega_1 is transformed as
// one combined spectrum wy*u3 - wz*u2 (dwdy - dvdz by linearity).
template <typename T>
struct OpOmega {
  const KArgs<T>& A;
  int c, i;
  __device__ C2<T> load(int ky, int kz) {
    T wy = A.wy[ky], wz = A.wz[kz];
    if (c == 0) return cik(wy, A.u[sidx(A, 2, i, ky, kz)]) - cik(wz, A.u[sidx(A, 1, i, ky, kz)]);
    if (c == 1) return cik(wz, A.u[sidx(A, 0, i, ky, kz)]);  // dudz
    return cik(wy, A.u[sidx(A, 0, i, ky, kz)]);              // dudy
  }
  __device__ T val(int y, int z, T v) {
    if (c == 0) return v;
    int im = wrapn(i - 1, A.nx), ip = wrapn(i + 1, A.nx);
    T twodx = (T)2 * A.dx;
    if (c == 1) return v - (A.ubcp[ridx(A, 2, ip, y, z)] - A.ubcp[ridx(A, 2, im, y, z)]) / twodx;
    return (A.ubcp[ridx(A, 1, ip, y, z)] - A.ubcp[ridx(A, 1, im, y, z)]) / twodx - v;
  }
  __device__ void store2(int, int y, int z, T a, T b) {
    size_t o = ridx(A, c, i, y, z);
    A.omega[o] = val(y, z, a);
    A.omega[o + 1] = val(y + 1, z, b);
  }
};
// calc_pgrad (boundary_condition.F90:1817-1896): grad p at pressure points + mpg
template <typename T>
struct OpPgrad {
  const KArgs<T>& A;
  int c, i;
  T add;
  __device__ C2<T> load(int ky, int kz) {
    if (c == 0) {
      int im = wrapn(i - 1, A.nx), ip = wrapn(i + 1, A.nx);
      C2<T> d = A.p[sidx(A, 0, ip, ky, kz)] - A.p[sidx(A, 0, im, ky, kz)];
      T twodx = (T)2 * A.dx;
      return mk(d.x / twodx, d.y / twodx);
    }
    C2<T> pv = A.p[sidx(A, 0, i, ky, kz)];
    return c == 1 ? cik(A.wy[ky], pv) : cik(A.wz[kz], pv);
  }
  __device__ void store2(int, int y, int z, T a, T b) {
    size_t o = ridx(A, c, i, y, z);
    A.ppr[o] = a + add;
    A.ppr[o + 1] = b + add;
  }
};

// The inner-point combination of calc_inner_reversal_forcing
// (boundary_condition.F90:1104-1479) and calc_bnd_forcing (:902-1102): the
// forcing interpolates cf ubcp, nlbcp, onlbcp and diffn at one point and
// combines them linearly with coefficients fixed for the stage, so the
// combination is formed on the spectra and transformed once:
// qin = c2r(cf u - c3 nl - c4 onl - c12 vis (delta_x^2 u / dx2 - w2 u))
//       + cf umean
template <typename T>
__device__ C2<T> nl_from_products(const KArgs<T>& A, int c, int i, int ky, int kz);
template <typename T, bool NLP>
struct OpInner {
  const KArgs<T>& A;
  int c, i;
  T cf, c3, c4, c12v, add;
  __device__ OpInner(const KArgs<T>& a, int c_, int i_) : A(a), c(c_), i(i_) {
    const double* k = A.sconst;
    cf = (T)k[0];
    c3 = (T)k[6];
    c4 = (T)k[7];
    c12v = (T)((k[4] + k[5]) * (double)A.vis);
    add = (T)(k[0] * k[1 + c]);
  }
  __device__ C2<T> load(int ky, int kz) {
    int im = wrapn(i - 1, A.nx), ip = wrapn(i + 1, A.nx);
    const size_t o = sidx(A, c, i, ky, kz);
    C2<T> um = A.u[sidx(A, c, im, ky, kz)], uc = A.u[o], up = A.u[sidx(A, c, ip, ky, kz)];
    C2<T> d2 = (um - cscal((T)2, uc)) + up;
    C2<T> v = cscal((T)1 / A.dx2, d2) - cscal(A.w2[(size_t)kz * A.my2 + ky], uc);
    C2<T> nlv;
    if constexpr (NLP) {  // nl formed here from the product spectra, and stored
      nlv = nl_from_products(A, c, i, ky, kz);
      A.nl[o] = nlv;
    } else {
      nlv = A.nl[o];
    }
    return ((cscal(cf, uc) - cscal(c3, nlv)) - cscal(c4, A.onl[o])) - cscal(c12v, v);
  }
  __device__ void store2(int, int y, int z, T a, T b) {
    size_t o = ridx(A, c, i, y, z);
    A.qin[o] = a + add;
    A.qin[o + 1] = b + add;
  }
};

// generic c2r launcher body: grid (nplanes, ncomp)
enum { C2R_UBCP = 0, C2R_PBCP, C2R_NLBC, C2R_ONLBC, C2R_VISC, C2R_OMEGA, C2R_PGRAD, C2R_INNER, C2R_INNER_NL };

template <typename T, int KIND, int SH>
__device__ __forceinline__ void c2r_plane(const KArgs<T>& A, unsigned char* smem, int c, int i) {
  C2<T>* buf = SH == 0 ? plane_slot(A) : nullptr;
  if constexpr (KIND == C2R_UBCP) {
    OpUbcp<T> op(A, c, i);
    plane_c2r_sh<SH, 0>(buf, A.G, A.tb, op, smem);
    double* sh = (double*)smem;
    double sm[3] = {op.sf, op.ss, op.sa}, mx[1] = {op.mx};
    block_sums_w<3>(sm, sh);
    block_maxes_w<1>(mx, sh);
    if (threadIdx.x == 0) {
      double* o = A.part + ((size_t)c * A.nx + i) * 4;
      o[0] = sm[0]; o[1] = sm[1]; o[2] = sm[2]; o[3] = mx[0];
    }
  } else if constexpr (KIND == C2R_PBCP) {
    OpPlain<T> op{A.p, A.pbcp, A, (size_t)i * A.mz * A.my2, (size_t)i * A.mz * A.my, (T)0};
    plane_c2r_sh<SH, 0>(buf, A.G, A.tb, op, smem);
  } else if constexpr (KIND == C2R_NLBC || KIND == C2R_ONLBC) {
    const C2<T>* src = KIND == C2R_NLBC ? A.nl : A.onl;
    T* dst = KIND == C2R_NLBC ? A.nlbc : A.onlbc;
    OpPlain<T> op{src, dst, A, ((size_t)c * A.nx + i) * A.mz * A.my2, ((size_t)c * A.nx + i) * A.mz * A.my, (T)0};
    plane_c2r_sh<SH, 0>(buf, A.G, A.tb, op, smem);
  } else if constexpr (KIND == C2R_VISC) {
    OpVisc<T> op{A, c, i};
    plane_c2r_sh<SH, 0>(buf, A.G, A.tb, op, smem);
  } else if constexpr (KIND == C2R_OMEGA) {
    OpOmega<T> op{A, c, i};
    plane_c2r_sh<SH, 0>(buf, A.G, A.tb, op, smem);
  } else if constexpr (KIND == C2R_INNER || KIND == C2R_INNER_NL) {
    OpInner<T, KIND == C2R_INNER_NL> op(A, c, i);
    plane_c2r_sh<SH, 0>(buf, A.G, A.tb, op, smem);
  } else {
    OpPgrad<T> op{A, c, i, (T)A.st->s.mpg[c]};
    plane_c2r_sh<SH, 0>(buf, A.G, A.tb, op, smem);
  }
}
// Waves per SIMD the FP32 transforms of the 128^2 plane are compiled for:
// 8 (64 VGPRs), so that two 1024-lane workgroups share a CU -- the padded
// plane (74 KB) leaves room for two in the 160 KB LDS.  Others: no bound.
template <typename T, int SH>
struct PlaneOcc {
  static constexpr int WPE = (SH == 1 && sizeof(T) == 4) ? 8 : 1;
};
// grid (planes, components of KIND)
template <typename T, int KIND, int SH>
__global__ void __launch_bounds__(1024, (PlaneOcc<T, SH>::WPE)) k_c2r(KArgs<T> A) {
  extern __shared__ __align__(16) unsigned char smem[];
  c2r_plane<T, KIND, SH>(A, smem, blockIdx.y, plane_of(A));
}
// two independent transforms in one launch: grid (planes, NA + NB), the
// first NA rows of workgroups do KA, the rest KB
template <typename T, int KA, int NA, int KB, int SH>
__global__ void __launch_bounds__(1024, (PlaneOcc<T, SH>::WPE)) k_c2r_pair(KArgs<T> A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int y = blockIdx.y;
  if (y < NA) c2r_plane<T, KA, SH>(A, smem, y, plane_of(A));
  else c2r_plane<T, KB, SH>(A, smem, y - NA, plane_of(A));
}

// ===========================================================================
// r2c: fr - frmean -> ff (boundary_condition.F90:678-701)
// ===========================================================================
template <typename T>
struct OpForce {
  const KArgs<T>& A;
  int c, i;
  T fm;
  __device__ void load2(int y, int z, T& a, T& b) {
    size_t o = ridx(A, c, i, y, z);
    a = A.fr[o] - fm;
    b = A.fr[o + 1] - fm;
  }
  __device__ void store(int, int ky, int kz, C2<T> v) { A.ff[sidx(A, c, i, ky, kz)] = v; }
};
template <typename T, int SH>
__global__ void __launch_bounds__(1024, (PlaneOcc<T, SH>::WPE)) k_r2c_force(KArgs<T> A) {
  extern __shared__ __align__(16) unsigned char smem[];
  C2<T>* buf = SH == 0 ? plane_slot(A) : nullptr;
  const int i = plane_of(A);
  const int c = blockIdx.y;
  OpForce<T> op{A, c, i, (T)A.st->s.frmean[c]};
  plane_r2c_sh<SH, 0>(buf, A.G, A.tb, op, smem);
}


// ===========================================================================
// form_nl (nl_allflow.F90:71-327) + the sign flip of nonlinear (nl.F90:141-153)
// One workgroup per (plane i, component c): three forward transforms
// accumulated in registers, nl_c = -( wy F(u_c v) + wz F(u_c w) + F(D_c) ),
// D_c = x-derivative of the x-flux of u_c.
// ===========================================================================
template <typename T, bool UPWIND, bool MINMOD>
struct NlFlux {
  const KArgs<T>& A;
  __device__ T u(int c, int i, int y, int z) const { return A.ubcp[ridx(A, c, wrapn(i, A.nx), y, z)]; }
  // MUSCL slope of component c in cell i (nl_allflow.F90:171-207)
  __device__ T slope(int c, int i, int y, int z) const {
    T um = u(c, i - 1, y, z), uc = u(c, i, y, z), up = u(c, i + 1, y, z);
    T s = (up - um) / ((T)2 * A.dx);
    T phi = (T)1;
    if (MINMOD) {
      phi = (uc - um) / (up - uc + (T)PIB_SMALL_NUMBER);
      if (phi <= (T)0) phi = (T)0;
      else phi = fmin((T)1, phi);
    }
    return s * phi * (T)1;
  }
  // Rusanov flux of component c at interface i+1/2 (nl_allflow.F90:212-294)
  __device__ T flux(int c, int i, int y, int z) const {
    const T h = A.dx / (T)2;
    T s1i = slope(0, i, y, z), s1p = slope(0, i + 1, y, z);
    T uminus = u(0, i, y, z) + s1i * h;
    T uplus = u(0, i + 1, y, z) - s1p * h;
    T Um, Up;
    if (c == 0) {
      Um = uminus;
      Up = uplus;
    } else {
      Um = u(c, i, y, z) + slope(c, i, y, z) * h;
      Up = u(c, i + 1, y, z) - slope(c, i + 1, y, z) * h;
    }
    T a = (T)0.5 * uminus * Um;
    T b = (T)0.5 * uplus * Up;
    T ms = fmax(fabs(uminus), fabs(uplus));
    return (a + b) - (T)0.5 * ms * (Up - Um);
  }
  __device__ T dx_flux(int c, int i, int y, int z) const {
    if (UPWIND) return (flux(c, i, y, z) - flux(c, i - 1, y, z)) / A.dx;
    T fp = u(0, i + 1, y, z) * u(c, i + 1, y, z);
    T fm = u(0, i - 1, y, z) * u(c, i - 1, y, z);
    return (fp - fm) / ((T)2 * A.dx);
  }
};

template <typename T, int NIT, bool UPWIND, bool MINMOD>
struct OpNl {
  const KArgs<T>& A;
  int c, i, stage;  // stage 0: y-product, 1: z-product, 2: x-flux derivative
  C2<T> acc[NIT];
  NlFlux<T, UPWIND, MINMOD> fx;
  C2<T>* nyq;  // LDS accumulators of the ky = my/2 column (outputs with it = -1)
  __device__ void load2(int y, int z, T& a, T& b) {
    a = val(y, z);
    b = val(y + 1, z);
  }
  __device__ T val(int y, int z) const {
    if (stage == 2) return fx.dx_flux(c, i, y, z);
    // stage 0: u_c * u_2, stage 1: u_c * u_3 (products as the reference forms
    // them: uv, vv, vw / uw, vw, ww)
    size_t o = ridx(A, 0, i, y, z);
    size_t cs = (size_t)A.nx * A.mz * A.my;
    T uc = A.ubcp[o + (size_t)c * cs], us = A.ubcp[o + (size_t)(stage + 1) * cs];
    return uc * us;
  }
  bool rmw = false;     // stage 2 adds to the partial sum kept in nl and writes -(sum)
  bool direct = false;  // every stage updates nl in place (no accumulators)
  __device__ void store(int it, int ky, int kz, C2<T> v) {
    if (rmw || direct) {
      C2<T>* p = A.nl + sidx(A, c, i, ky, kz);
      if (stage == 0) {
        *p = cik(A.wy[ky], v);
      } else if (stage == 1) {
        *p = *p + cik(A.wz[kz], v);
      } else {
        C2<T> a = *p + v;
        *p = mk(-a.x, -a.y);
      }
      return;
    }
    C2<T>& a = it < 0 ? nyq[kz] : acc[it];
    if (stage == 0) a = cik(A.wy[ky], v);
    else if (stage == 1) a = a + cik(A.wz[kz], v);
    else a = a + v;
  }
  // the same with the stage fixed at compile time (in-place form)
  template <int ST>
  __device__ T val_at(int y, int z) const {
    if constexpr (ST == 2) {
      return fx.dx_flux(c, i, y, z);
    } else {
      size_t o = ridx(A, 0, i, y, z);
      size_t cs = (size_t)A.nx * A.mz * A.my;
      return A.ubcp[o + (size_t)c * cs] * A.ubcp[o + (size_t)(ST + 1) * cs];
    }
  }
  template <int ST>
  __device__ void store_at(int ky, int kz, C2<T> v) {
    C2<T>* p = A.nl + sidx(A, c, i, ky, kz);
    if constexpr (ST == 0) {
      *p = cik(A.wy[ky], v);
    } else if constexpr (ST == 1) {
      *p = *p + cik(A.wz[kz], v);
    } else {
      C2<T> a = *p + v;
      *p = mk(-a.x, -a.y);
    }
  }
};
// one stage of the in-place form as an Op of its own
template <int ST, class O, typename T>
struct OpNlStage {
  O& o;
  __device__ void load2(int y, int z, T& a, T& b) {
    a = o.template val_at<ST>(y, z);
    b = o.template val_at<ST>(y + 1, z);
  }
  __device__ void store(int, int ky, int kz, C2<T> v) { o.template store_at<ST>(ky, kz, v); }
};

// SH = 0 (generic plane): the three transforms in one workgroup with
// register accumulators, stored negated at the end.  Compiled shape, and
// SH = -1 (generic plane, for planes whose accumulators do not fit the
// registers or whose buffer is in HBM): the three transforms in one
// workgroup, each updating nl in place (the lane that writes an output
// element is the lane that updates it).  All add in the order
// ((wy F0 + wz F1) + F2).
// Workgroup size and waves per SIMD of form_nl: the FP32 128^2 plane runs
// 512 lanes at 4 waves per SIMD (128 VGPRs), so that two workgroups share a
// CU; the generic in-place form (SH = -1) takes up to 1024 lanes; others use
// the shape's lanes with no bound.
template <typename T, int SH>
struct NlOcc {
  static constexpr int NT = (SH == 1 && sizeof(T) == 4) ? 512 : (SH == 0 ? 512 : (SH < 0 ? 1024 : Shape<SH>::NT));
  static constexpr int WPE = (SH == 1 && sizeof(T) == 4) ? 4 : 1;
};
template <typename T, int NIT, bool UPWIND, bool MINMOD, int SH>
__global__ void __launch_bounds__((NlOcc<T, SH>::NT), (NlOcc<T, SH>::WPE)) k_form_nl(KArgs<T> A) {
  extern __shared__ __align__(16) unsigned char smem[];
  C2<T>* buf = (C2<T>*)smem;
  const int i = plane_of(A);
  const int c = blockIdx.y;
  OpNl<T, NIT, UPWIND, MINMOD> op{A, c, i, 0, {}, {A}, nullptr};
  if constexpr (SH == 0) {
    for (int s = 0; s < 3; ++s) {
      op.stage = s;
      plane_r2c<NIT>(buf, A.G, A.tb, op);
    }
    const int nt = blockDim.x;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      int e = threadIdx.x + it * nt;
      if (e < A.my2 * A.mz) {
        int kz = e / A.my2, ky = e - kz * A.my2;
        C2<T> v = op.acc[it];
        A.nl[sidx(A, c, i, ky, kz)] = mk(-v.x, -v.y);
      }
    }
  } else if constexpr (SH < 0) {
    C2<T>* pb = plane_slot(A);
    using O = OpNl<T, NIT, UPWIND, MINMOD>;
    OpNlStage<0, O, T> o0{op};
    plane_r2c_sh<0, 0>(pb, A.G, A.tb, o0, smem);
    OpNlStage<1, O, T> o1{op};
    plane_r2c_sh<0, 0>(pb, A.G, A.tb, o1, smem);
    OpNlStage<2, O, T> o2{op};
    plane_r2c_sh<0, 0>(pb, A.G, A.tb, o2, smem);
  } else {
    constexpr int HY = Shape<SH>::HY, MZ = Shape<SH>::MZ, NT = NlOcc<T, SH>::NT;
    op.nyq = (C2<T>*)(smem + A.G.tab_off + tab_lds_bytes<T>(HY, MZ));
    using O = OpNl<T, NIT, UPWIND, MINMOD>;
    OpNlStage<0, O, T> o0{op};
    plane_r2c2<HY, MZ, NT>(buf, A.G, A.tb, o0);
    OpNlStage<1, O, T> o1{op};
    plane_r2c2<HY, MZ, NT>(buf, A.G, A.tb, o1);
    OpNlStage<2, O, T> o2{op};
    plane_r2c2<HY, MZ, NT>(buf, A.G, A.tb, o2);
  }
}

// ===========================================================================
// form_nl (nl_allflow.F90:71-327) by separate transforms: per plane the five
// distinct products of the y- and z-fluxes, uv, uw, vv, vw, ww (as the
// reference forms them, u_c u_s), and the x-flux derivative D_c of each
// component (central or MUSCL/Rusanov, as the in-place form has it) are
// transformed once each -- eight transforms a plane where the in-place form
// runs nine -- and
//   nl_c(i) = -((wy F(u_c v) + wz F(u_c w)) + F(D_c))
// is formed where it is consumed (the qin transform, or k_nl_combine), in
// the in-place form's order of terms, so nl is written once.
// ===========================================================================
template <typename T, bool UPWIND, bool MINMOD>
struct OpProd {
  const KArgs<T>& A;
  int a, b, i, q;
  NlFlux<T, UPWIND, MINMOD> fx;
  __device__ T val(int y, int z) const {
    if (q >= 5) return fx.dx_flux(q - 5, i, y, z);
    return A.ubcp[ridx(A, a, i, y, z)] * A.ubcp[ridx(A, b, i, y, z)];
  }
  __device__ void load2(int y, int z, T& va, T& vb) {
    va = val(y, z);
    vb = val(y + 1, z);
  }
  __device__ void store(int, int ky, int kz, C2<T> v) { A.nlp[sidx(A, q, i, ky, kz)] = v; }
};
// grid (planes, 8 spectra): q = 0 uv, 1 uw, 2 vv, 3 vw, 4 ww, 5 + c: D_c
template <typename T, bool UPWIND, bool MINMOD, int SH>
__global__ void __launch_bounds__(1024, (PlaneOcc<T, SH>::WPE)) k_nl_products(KArgs<T> A) {
  extern __shared__ __align__(16) unsigned char smem[];
  C2<T>* buf = SH == 0 ? plane_slot(A) : nullptr;
  const int i = plane_of(A);
  const int q = blockIdx.y;
  const int a = q < 2 ? 0 : (q < 4 ? 1 : 2);
  const int b = q < 2 ? q + 1 : (q < 4 ? q - 1 : 2);
  OpProd<T, UPWIND, MINMOD> op{A, a, b, i, q, {A}};
  plane_r2c_sh<SH, 0>(buf, A.G, A.tb, op, smem);
}
template <typename T>
__device__ C2<T> nl_from_products(const KArgs<T>& A, int c, int i, int ky, int kz) {
  // F(u_c v) and F(u_c w) of component c: uv, uw / vv, vw / vw, ww
  const int qy = c == 0 ? 0 : (c == 1 ? 2 : 3), qz = c == 0 ? 1 : (c == 1 ? 3 : 4);
  const C2<T> py = A.nlp[sidx(A, qy, i, ky, kz)];
  const C2<T> pz = A.nlp[sidx(A, qz, i, ky, kz)];
  const C2<T> px = A.nlp[sidx(A, 5 + c, i, ky, kz)];
  const C2<T> a = (cik(A.wy[ky], py) + cik(A.wz[kz], pz)) + px;
  return mk(-a.x, -a.y);
}
// nl from the products where no qin transform follows (a bed without
// particles): one lane per spectral entry (c, i, ky, kz)
template <typename T>
__global__ void __launch_bounds__(256) k_nl_combine(KArgs<T> A) {
  const size_t E = (size_t)A.my2 * A.mz;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * (size_t)A.nx * E) return;
  const size_t ci = e / E, q = e - ci * E;
  const int c = (int)(ci / A.nx), i = (int)(ci - (size_t)c * A.nx);
  const int kz = (int)(q / A.my2), ky = (int)(q - (size_t)kz * A.my2);
  A.nl[e] = nl_from_products(A, c, i, ky, kz);
}

// ===========================================================================
// IBM: trilinear ('lpi' order 2) stencils, dependent_functions.F90:45-275,
// 612-781; interpolation.F90:202-433, 740-751.  For order 2 with periodic
// x, y, z the stencil base is the cell pc itself; nodes pc, pc+1 wrapped by
// one period into 1..n.
// ===========================================================================
struct Sten {
  unsigned oi[2], ok[2], oj[2];  // per-axis parts of the 0-based wrapped node index
  double z[3];                   // position within the cell
  // node (i, j, k) of the stencil: (ii*mz + kk)*my + jj
  __device__ unsigned node(int i, int j, int k) const { return oi[i] + ok[k] + oj[j]; }
  // xval(i)*yval(j)*zval(k), formed where it is used
  __device__ double w(int i, int j, int k) const {
    return (i ? z[0] : 1.0 - z[0]) * (j ? z[1] : 1.0 - z[1]) * (k ? z[2] : 1.0 - z[2]);
  }
};
__device__ inline int cellof(double x) { return x < 0.0 ? (int)(x - 1.0) : (int)x; }
// one-period wrap of the reference, written as a full modulo so that any
// input maps into 1..n
__device__ inline int wrap1(int v, int n) {
  int r = (v - 1) % n;
  if (r < 0) r += n;
  return r + 1;
}
__device__ inline void make_sten(Sten& s, const double pos[3], int nx, int my, int mz) {
  int pc[3] = {cellof(pos[0]), cellof(pos[1]), cellof(pos[2])};
  for (int d = 0; d < 3; ++d) s.z[d] = pos[d] - (double)pc[d];
  for (int q = 0; q < 2; ++q) {
    s.oi[q] = (unsigned)(wrap1(pc[0] + q, nx) - 1) * (unsigned)mz * (unsigned)my;
    s.oj[q] = (unsigned)(wrap1(pc[1] + q, my) - 1);
    s.ok[q] = (unsigned)(wrap1(pc[2] + q, mz) - 1) * (unsigned)my;
  }
}
// The two y corners (j = 0, 1) of a stencil row are adjacent in memory
// unless the stencil wraps in y, so one 2-element load (4-byte aligned for
// float, which the hardware's unaligned mode takes) at corner j = 0 fetches
// both; a wrapping lane then loads its j = 1 corner on its own.  The pair
// load at the last node of a field reads one element past it: gathered
// fields are allocated GATHER_PAD elements longer.  v[k][i][j].
#define GATHER_PAD 64
template <typename T>
__device__ inline void gather8(const Sten& s, const T* f, T (&v)[2][2][2]) {
  typedef T Two __attribute__((ext_vector_type(2)));
  const bool wrap = s.oj[1] != s.oj[0] + 1u;
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const unsigned n = s.oi[i] + s.ok[k];
      Two t;
      __builtin_memcpy(&t, f + n + s.oj[0], sizeof t);
      v[k][i][0] = t.x;
      v[k][i][1] = t.y;
      if (wrap) v[k][i][1] = f[n + s.oj[1]];
    }
}
// interp_threed_scalar order (k, j, i)
template <typename T>
__device__ inline double interp_kji(const Sten& s, const T* f) {
  T v[2][2][2];
  gather8(s, f, v);
  double acc = 0.0;
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) acc = acc + (double)v[k][i][j] * s.w(i, j, k);
  return acc;
}
// array_dot_product_3d order (i, j, k), general_funcs.F90:126-143
template <typename T>
__device__ inline double interp_ijk(const Sten& s, const T* f) {
  T v[2][2][2];
  gather8(s, f, v);
  double acc = 0.0;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) acc = acc + (double)v[k][i][j] * s.w(i, j, k);
  return acc;
}
// The value is complete at this point of the program: the loads and
// arithmetic it depends on are not deferred past here.  Ends a group of
// gathers, so that the compiler keeps one group's loads in flight at a time
// instead of hoisting every gather of a point (registers, occupancy).
template <typename X>
__device__ __forceinline__ void pin_reg(X& x) {
  __asm__ __volatile__("" : "+v"(x));
}

struct Particles {
  const double* xc;   // (nbody,3) column-major
  const double* rad;  // radbdy
  const double* radi; // radibdy
  const double* rado; // radobdy
  const double* vel;  // (nbody,3)
  const double* angv; // (nbody,3)
  const double* bnd;  // (3,nbnd)
  const double* cd;   // (3,nbnd)
  const uint8_t* if_rev;     // (nbnd, nbody)
  const int* nrpr_active;
  int nbody, nbnd;
};

// calc_pres_visc_drag (boundary_condition.F90:1481-1645).  Work item (m, g)
// (xcd_linear over nbody*G workgroups) takes surface points g*blockDim.x + t
// of body m and writes
// the workgroup sums of their contributions, dpart[(m*G + g)*9 + q]:
// q = 0..2 pressure part (subtracted from pres(m)), 3..5 viscous part,
// 6..8 torque.
template <typename T>
__global__ void __launch_bounds__(256) k_drag_points(KArgs<T> A, Particles P, double* dpart, int G) {
  __shared__ double sh[9 * 16];
  const int w = xcd_linear(blockIdx.x, gridDim.x);
  const int m = w / G, g = w - m * G;
  if (m >= P.nbody) return;  // padding of the grid to a multiple of 8
  const int l = g * blockDim.x + threadIdx.x;
  double o[9];
  for (int q = 0; q < 9; ++q) o[q] = 0.0;
  if (l < P.nbnd) {
    const DevState* st = A.st;
    double rb = P.rad[m];
    double xl[3], xpb[3];
    for (int n = 0; n < 3; ++n) xl[n] = P.xc[n * P.nbody + m] + P.bnd[l * 3 + n] * rb;
    xpb[0] = xl[0] - 0.5;
    xpb[1] = xl[1];
    xpb[2] = xl[2];
    Sten sp, sv;
    make_sten(sp, xpb, A.nx, A.my, A.mz);
    make_sten(sv, xl, A.nx, A.my, A.mz);
    double pl = interp_ijk(sp, A.pbcp);
    size_t cs = (size_t)A.nx * A.mz * A.my;
    double vort[3];
    for (int n = 0; n < 3; ++n) vort[n] = interp_kji(sv, A.omega + n * cs);
    double cd1 = P.cd[l * 3 + 0], cd2 = P.cd[l * 3 + 1], cd3 = P.cd[l * 3 + 2];
    double df[3];
    df[0] = -(vort[2] * cd2 - vort[1] * cd3);
    df[1] = -(vort[0] * cd3 - vort[2] * cd1);
    df[2] = -(vort[1] * cd1 - vort[0] * cd2);
    double rdx = rb * st->dx;
    double da1 = 4.0 * 3.14159265358979323846 * (rdx * rdx) / (double)P.nbnd;
    double cdv[3] = {cd1, cd2, cd3};
    double tf[3];
    for (int d = 0; d < 3; ++d) {
      o[d] = pl * cdv[d] * da1;
      tf[d] = df[d] * st->vis * da1;
      o[3 + d] = tf[d];
    }
    double b1 = P.bnd[l * 3], b2 = P.bnd[l * 3 + 1], b3 = P.bnd[l * 3 + 2];
    o[6] = (b2 * tf[2] - b3 * tf[1]) * rb * st->dx;
    o[7] = (b3 * tf[0] - b1 * tf[2]) * rb * st->dx;
    o[8] = (b1 * tf[1] - b2 * tf[0]) * rb * st->dx;
  }
  block_sums_w<9>(o, sh);
  if (threadIdx.x == 0)
    for (int q = 0; q < 9; ++q) dpart[((size_t)m * G + g) * 9 + q] = o[q];
}

// calc_inner_reversal_forcing (boundary_condition.F90:1104-1479), serial path:
// per active reversal pair, force_tmp and the solid-node redistribution share.
// Work item (m, g) (xcd_linear over nbody*G workgroups) takes points
// l = g*blockDim.x + t of body m; out[(m*nbnd + l)*6 + 0..2] force_tmp,
// [3..5] force_dist; the workgroup
// sums of the frmean / frmeanfluid parts go to fpart[(m*G + g)*6 + 0..5].
template <typename T>
__device__ void forcing_point(const KArgs<T>& A, const Particles& P, int m, int l, double* o, double* fsum);
template <typename T>
__global__ void __launch_bounds__(256) k_forcing_points(KArgs<T> A, Particles P, double* out, double* fpart,
                                                        int G) {
  __shared__ double sh[6 * 16];
  const int w = xcd_linear(blockIdx.x, gridDim.x);
  const int m = w / G, g = w - m * G;
  if (m >= P.nbody) return;  // padding of the grid to a multiple of 8
  const int l = g * blockDim.x + threadIdx.x;
  double fs[6] = {0, 0, 0, 0, 0, 0};
  if (l < P.nbnd) {
    double* o = out + ((size_t)m * P.nbnd + l) * 6;
    if (!P.if_rev[(size_t)m * P.nbnd + l]) {
      for (int q = 0; q < 6; ++q) o[q] = 0.0;
    } else {
      forcing_point(A, P, m, l, o, fs);
    }
  }
  block_sums_w<6>(fs, sh);
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; ++q) fpart[((size_t)m * G + g) * 6 + q] = fs[q];
}
template <typename T>
__device__ void forcing_point(const KArgs<T>& A, const Particles& P, int m, int l, double* o, double* fsum) {
  const DevState* st = A.st;
  const int rks = A.rks;
  const double c1 = st->coef[rks - 1][0], c2 = st->coef[rks - 1][1];
  double rb = P.rad[m], ri = P.radi[m], ro = P.rado[m];
  double xli[3], xlo[3], rad = 0.0;
  for (int n = 0; n < 3; ++n) {
    double b = P.bnd[l * 3 + n];
    double xcn = P.xc[n * P.nbody + m];
    xli[n] = xcn + b * ri;
    xlo[n] = xcn + b * ro;
    rad = rad + (b * ri) * (b * ri);
  }
  rad = sqrt(rad);
  double snorm[3];
  for (int n = 0; n < 3; ++n) snorm[n] = (P.bnd[l * 3 + n] * ri) / rad;
  double xpi[3] = {xli[0] - 0.5, xli[1], xli[2]};
  size_t cs = (size_t)A.nx * A.mz * A.my;
  Sten so, sp, si;
  make_sten(so, xlo, A.nx, A.my, A.mz);
  make_sten(sp, xpi, A.nx, A.my, A.mz);
  make_sten(si, xli, A.nx, A.my, A.mz);
  // gathers in groups of three interpolations; ft accumulates in the
  // reference's order of terms
  double ulo[3];
  for (int n = 0; n < 3; ++n) ulo[n] = interp_kji(so, A.ubcp + n * cs);
  for (int n = 0; n < 3; ++n) pin_reg(ulo[n]);
  double w1 = P.angv[m], w2 = P.angv[P.nbody + m], w3 = P.angv[2 * P.nbody + m];
  double crossp[3] = {w2 * snorm[2] - w3 * snorm[1], w3 * snorm[0] - w1 * snorm[2], w1 * snorm[1] - w2 * snorm[0]};
  double linvel[3], unmag = 0.0, ubnmag = 0.0;
  for (int n = 0; n < 3; ++n) linvel[n] = P.vel[n * P.nbody + m] + crossp[n] * rb * st->dx;
  for (int d = 0; d < 3; ++d) {
    unmag = unmag + snorm[d] * ulo[d];
    ubnmag = ubnmag + snorm[d] * linvel[d];
  }
  double ft[3];
  for (int d = 0; d < 3; ++d) {
    double unorm = snorm[d] * unmag;
    double utang = ulo[d] - unorm;
    double ubnorm = snorm[d] * ubnmag;
    double ubtang = linvel[d] - ubnorm;
    double utangi = (ubtang * (ro - ri) - utang * (rb - ri)) / (ro - rb);
    double unormi = 0.0;
    if (st->revernorm == 1) unormi = (ubnorm * (ro - ri) - unorm * (rb - ri)) / (ro - rb);
    ft[d] = -(utangi + unormi);  // -ucar
  }
  // cf (uli - ucar) - c3 nll - c4 onll + c12 (ppll - dfll), the four
  // inner-point terms as one interpolation of qin
  for (int n = 0; n < 3; ++n) ft[n] = st->s.cf * ft[n] + interp_kji(si, A.qin + n * cs);
  for (int n = 0; n < 3; ++n) pin_reg(ft[n]);
  for (int n = 0; n < 3; ++n) ft[n] = ft[n] + (c1 + c2) * interp_kji(sp, A.ppr + n * cs);
  for (int n = 0; n < 3; ++n) pin_reg(ft[n]);
  double rdi = ri * st->dx;
  double da2 = 4.0 * 3.14159265358979323846 * (rdi * rdi) / (double)P.nrpr_active[m];
  for (int n = 0; n < 3; ++n) ft[n] = ft[n] * da2 * st->dr * st->dx;
  int count_fl = 0;
  double force_fl[3] = {0, 0, 0}, fm[3] = {0, 0, 0}, fmf[3] = {0, 0, 0};
  uint8_t mk8[2][2][2];
  gather8(si, A.mask, mk8);
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        bool fl = mk8[k][i][j] != 0;
        double w = si.w(i, j, k);
        if (fl) count_fl++;
        for (int n = 0; n < 3; ++n) {
          double tp = w * ft[n];
          if (fl) force_fl[n] = force_fl[n] + tp / st->dx3;
          if (fl) fmf[n] = fmf[n] + tp / st->dx3;
          fm[n] = fm[n] + tp / st->dx3;
        }
      }
  int count_so = 8 - count_fl;
  for (int n = 0; n < 3; ++n) {
    o[n] = ft[n];
    o[3 + n] = count_so > 0 ? force_fl[n] / (double)count_so : 0.0;
    fsum[n] = fm[n];
    fsum[3 + n] = fmf[n];
  }
}

// calc_bnd_forcing (boundary_condition.F90:902-1102): the direct forcing at
// every surface point, interpolated as at an inner reversal point and
// spread to all eight nodes of its stencil (fluid or solid, no
// redistribution).  Work item (m, g) as above; out[(nbody*nbnd + m*nbnd + l)*6
// + 0..2] force_tmp, [3..5] zero; the workgroup sums of the frmean /
// frmeanfluid parts go to fpart[(nbody*G + m*G + g)*6 + 0..5].  The
// rotation term uses the point's own normal cd (the reference's stale SNORM
// gives the same for angv = 0, the case of every reference run).
template <typename T>
__global__ void __launch_bounds__(256) k_bnd_points(KArgs<T> A, Particles P, double* out, double* fpart, int G) {
  __shared__ double sh[6 * 16];
  const int w = xcd_linear(blockIdx.x, gridDim.x);
  const int m = w / G, g = w - m * G;
  if (m >= P.nbody) return;  // padding of the grid to a multiple of 8
  const int l = g * blockDim.x + threadIdx.x;
  double fs[6] = {0, 0, 0, 0, 0, 0};
  if (l < P.nbnd) {
    const DevState* st = A.st;
    const int rks = A.rks;
    const double c1 = st->coef[rks - 1][0], c2 = st->coef[rks - 1][1];
    const double rb = P.rad[m];
    double xl[3];
    for (int n = 0; n < 3; ++n) xl[n] = P.xc[n * P.nbody + m] + P.bnd[l * 3 + n] * rb;
    const double xpb[3] = {xl[0] - 0.5, xl[1], xl[2]};
    const size_t cs = (size_t)A.nx * A.mz * A.my;
    Sten sv, sp;
    make_sten(sv, xl, A.nx, A.my, A.mz);
    make_sten(sp, xpb, A.nx, A.my, A.mz);
    const double w1 = P.angv[m], w2 = P.angv[P.nbody + m], w3 = P.angv[2 * P.nbody + m];
    const double sn[3] = {P.cd[l * 3], P.cd[l * 3 + 1], P.cd[l * 3 + 2]};
    const double crossp[3] = {w2 * sn[2] - w3 * sn[1], w3 * sn[0] - w1 * sn[2], w1 * sn[1] - w2 * sn[0]};
    double ft[3];
    for (int n = 0; n < 3; ++n) {
      const double linvel = P.vel[n * P.nbody + m] + crossp[n] * rb * st->dx;
      ft[n] = st->s.cf * (-linvel) + interp_kji(sv, A.qin + n * cs);
    }
    for (int n = 0; n < 3; ++n) pin_reg(ft[n]);
    for (int n = 0; n < 3; ++n) ft[n] = ft[n] + (c1 + c2) * interp_kji(sp, A.ppr + n * cs);
    const double rdx = rb * st->dx;
    const double da1 = 4.0 * 3.14159265358979323846 * (rdx * rdx) / (double)P.nbnd;
    for (int n = 0; n < 3; ++n) ft[n] = ft[n] * da1 * st->dr * st->dx;
    uint8_t mk8[2][2][2];
    gather8(sv, A.mask, mk8);
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
          const double wq = sv.w(i, j, k);
          for (int n = 0; n < 3; ++n) {
            const double tp = wq * ft[n] / st->dx3;
            fs[n] += tp;
            if (mk8[k][i][j] && st->include_frmeanfluid) fs[3 + n] += tp;
          }
        }
    double* o = out + ((size_t)P.nbody * P.nbnd + (size_t)m * P.nbnd + l) * 6;
    for (int n = 0; n < 3; ++n) {
      o[n] = ft[n];
      o[3 + n] = 0.0;
    }
  }
  block_sums_w<6>(fs, sh);
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; ++q) fpart[((size_t)P.nbody * G + (size_t)m * G + g) * 6 + q] = fs[q];
}

// Atomics-free spreading.  The records (point, corner) that deposit on a
// grid node were binned to that node at set-up (stable sort: the reference's
// (body, point, corner) order within a node) and the target nodes were cut
// into workgroup ranges of at most SPREAD_CAP records (bt).  A workgroup
// first computes the contribution of each of its records with coalesced,
// batched loads -- the weighted share (w F)/dx^3 plus the redistributed
// fluid share -- into LDS (component-major, 24 KB), then one lane per target
// node adds its records in the reference's order.  The reference adds the
// two shares to fr one after the other; here they are added together first
// (a rounding-level difference).
#define SPREAD_CAP 1024
// frmean and frmeanfluid (boundary_condition.F90:678-701) from the
// (body, chunk) partials of the forcing: one workgroup, fixed order
__device__ inline void reduce_frmean_body(DevState* st, const double* fpart, int nrec, double* sh) {
  double a[6] = {0, 0, 0, 0, 0, 0};
  for (int p = threadIdx.x; p < nrec; p += blockDim.x)
    for (int q = 0; q < 6; ++q) a[q] += fpart[(size_t)p * 6 + q];
  block_sums_w<6>(a, sh);
  if (threadIdx.x == 0) {
    double ntot = (double)((long long)st->nx * st->my * st->mz);
    double cfl = st->s.count_fluid;
    double r[6];
    for (int n = 0; n < 3; ++n) {
      r[n] = a[n] / ntot;
      r[3 + n] = a[3 + n] / cfl;
    }
    for (int n = 0; n < 3; ++n) {
      st->s.frmean[n] = r[n];
      st->s.frmeanfluid[n] = r[3 + n];
    }
  }
}
// The spreading, and (work index nrange) the frmean reduction, which reads
// only the forcing's partials and writes only frmean/frmeanfluid, so it
// runs beside the spreading workgroups instead of as a launch of its own.
template <typename T>
__global__ void __launch_bounds__(256) k_spread(KArgs<T> A, const int* bt, const int* rs, int nrange,
                                                const int* tnode, const int* toff, const int* rpt, const double* rw,
                                                const uint8_t* rdist, const double* pts, DevState* stw,
                                                const double* fpart, int nfrec) {
  __shared__ double cv[3 * SPREAD_CAP];
  const int w = xcd_linear(blockIdx.x, gridDim.x);  // neighbouring ranges share an L2
  if (w == nrange) {
    reduce_frmean_body(stw, fpart, nfrec, cv);
    return;
  }
  if (w > nrange) return;
  const int t0 = bt[w], t1 = bt[w + 1];
  const int r0 = rs[w], r1 = rs[w + 1];  // toff[t0], toff[t1]
  const double dx3 = A.st->dx3;
  constexpr int CH = SPREAD_CAP / 256;
  int rp[CH];
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    int r = r0 + (int)threadIdx.x + k * 256;
    rp[k] = r < r1 ? rpt[r] : -1;
  }
  // this lane's nodes (a range holds at most SPREAD_CAP records, so at most
  // SPREAD_CAP nodes): their record bounds, loaded before the barrier
  int tb[CH], te[CH], tn[CH];
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    const int t = t0 + (int)threadIdx.x + k * 256;
    const bool ok = t < t1;
    tb[k] = ok ? toff[t] : 0;
    te[k] = ok ? toff[t + 1] : 0;
    tn[k] = ok ? tnode[t] : 0;
  }
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    int r = r0 + (int)threadIdx.x + k * 256;
    if (rp[k] >= 0) {
      const double* o = pts + (size_t)rp[k] * 6;
      double wt = rw[r];
      bool d = (rdist[r] & 1) != 0;
      int q = r - r0;
      for (int n = 0; n < 3; ++n) cv[n * SPREAD_CAP + q] = (wt * o[n]) / dx3 + (d ? o[3 + n] : 0.0);
    }
  }
  __syncthreads();
  const size_t cs = (size_t)A.nx * A.mz * A.my;
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    if (t0 + (int)threadIdx.x + k * 256 >= t1) break;
    double acc[3] = {0, 0, 0};
    for (int r = tb[k]; r < te[k]; ++r) {
      int q = r - r0;
      for (int n = 0; n < 3; ++n) acc[n] = acc[n] + cv[n * SPREAD_CAP + q];
    }
    for (int n = 0; n < 3; ++n) A.fr[n * cs + tn[k]] = (T)acc[n];
  }
}

// ===========================================================================
// velstep (steptheflow.F90:52-160): per (ky,kz) mode, one thread runs
// pressure (poisson.F90:55-239), ustep (usteptime.F90:45-258) and divcorr
// (poisson.F90:241-508) along x.  Thomas sweeps spill their intermediates to
// mode-major scratch (coalesced across threads).  The (0,0) mode always runs
// in double precision.
// ===========================================================================
template <typename T, typename R>
struct ModeIO {
  const KArgs<T>& A;
  int ky, kz;
  size_t base, step;  // element (c,i) at base + (c*nx + i)*step
  __device__ C2<R> get(const C2<T>* f, int c, int i) const {
    C2<T> v = f[base + ((size_t)c * A.nx + i) * step];
    return mk((R)v.x, (R)v.y);
  }
  __device__ void put(C2<T>* f, int c, int i, C2<R> v) const {
    f[base + ((size_t)c * A.nx + i) * step] = mk((T)v.x, (T)v.y);
  }
};

template <typename R>
__device__ __forceinline__ C2<R> ciR(R k, C2<R> z) { return mk(-k * z.y, k * z.x); }

// Forward Thomas sweep with the Sherman-Morrison second rhs (auxu = e_1 + e_n)
// for rows r(j) supplied by `rhs(j)`; writes y1 (re), y3 (im), y2 (auxz), gam
// to scratch.  b(0), b(n-1) already corrected.  Returns the pivot status.
template <typename R, class RHS>
__device__ int thomas3(int n, R a, R b, R b0, R bn, R c, RHS rhs, R* g_s, R* y1_s, R* y2_s, R* y3_s,
                       size_t ms, bool store_aux) {
  int st = 0;
  R bet = b0;
  if (bet == (R)0) st = PIB_ERR_PIVOT;
  C2<R> r0 = rhs(0);
  R y1 = r0.x / bet, y2 = (R)1 / bet, y3 = r0.y / bet;
  y1_s[0] = y1;
  y3_s[0] = y3;
  if (store_aux) { y2_s[0] = y2; g_s[0] = (R)0; }
  for (int j = 1; j < n; ++j) {
    R gam = c / bet;
    R bj = (j == n - 1) ? bn : b;
    bet = bj - a * gam;
    if (bet == (R)0) { bet = (R)1.0e-30; st = PIB_ERR_PIVOT; }
    C2<R> rj = rhs(j);
    R r2 = (j == n - 1) ? (R)1 : (R)0;
    y1 = (rj.x - a * y1) / bet;
    y3 = (rj.y - a * y3) / bet;
    y1_s[(size_t)j * ms] = y1;
    y3_s[(size_t)j * ms] = y3;
    if (store_aux) {
      y2 = (r2 - a * y2) / bet;
      y2_s[(size_t)j * ms] = y2;
      g_s[(size_t)j * ms] = gam;
    }
  }
  return st;
}
// backward substitution of the (re, im[, aux]) solutions in scratch
template <typename R>
__device__ void thomas_back(int n, const R* g_s, R* y1_s, R* y2_s, R* y3_s, size_t ms, bool aux) {
  R n1 = y1_s[(size_t)(n - 1) * ms], n3 = y3_s[(size_t)(n - 1) * ms];
  R n2 = aux ? y2_s[(size_t)(n - 1) * ms] : (R)0;
  for (int j = n - 2; j >= 0; --j) {
    R g = g_s[(size_t)(j + 1) * ms];
    R v1 = y1_s[(size_t)j * ms] - g * n1;
    R v3 = y3_s[(size_t)j * ms] - g * n3;
    y1_s[(size_t)j * ms] = v1;
    y3_s[(size_t)j * ms] = v3;
    n1 = v1;
    n3 = v3;
    if (aux) {
      R v2 = y2_s[(size_t)j * ms] - g * n2;
      y2_s[(size_t)j * ms] = v2;
      n2 = v2;
    }
  }
}

// scr: six scratch vectors [s][i][mode] with mode stride ms; qo is this
// mode's column in them.
template <typename T, typename R>
__device__ int run_mode(const KArgs<T>& A, int q, R* scr, size_t ms, size_t qo, double& divmax) {
  const int nx = A.nx;
  const int ky = q % A.my2, kz = q / A.my2;
  const DevState* st = A.st;
  const int rks = A.rks;
  const R c1 = (R)st->coef[rks - 1][0], c2 = (R)st->coef[rks - 1][1], c3 = (R)st->coef[rks - 1][2],
          c4 = (R)st->coef[rks - 1][3];
  const R dx = (R)st->dx, dx2 = dx * dx, dt = (R)st->s.dt, vis = (R)st->vis;
  const R wy = (R)A.wy[ky], wz = (R)A.wz[kz], w2 = (R)A.w2[(size_t)kz * A.my2 + ky];
  ModeIO<T, R> io{A, ky, kz, (size_t)kz * A.my2 + ky, (size_t)A.mz * A.my2};
  // scratch: 0 gam_p, 1 auxz_p, 2 gam_h, 3 auxz_h, 4 y_re, 5 y_im
  R* gp = scr + 0 * nx * ms + qo;
  R* zp = scr + 1 * nx * ms + qo;
  R* gh = scr + 2 * nx * ms + qo;
  R* zh = scr + 3 * nx * ms + qo;
  R* yr = scr + 4 * nx * ms + qo;
  R* yi = scr + 5 * nx * ms + qo;
  const bool mean = (ky == 0 && kz == 0);
  int status = 0;
  // ---------------- pressure (poisson.F90:55-239) ----------------
  const R ap = (R)1 / dx2 - (R)0.25 * w2;
  const R bp = (R)-2 / dx2 - (R)0.5 * w2;
  const R cp = ap;
  auto prhs = [&](int i) -> C2<R> {
    int i1 = wrapn(i + 1, nx);
    C2<R> t = mk((R)0, (R)0);
    C2<R> nl1 = io.get(A.nl, 0, i1) - io.get(A.nl, 0, i);
    C2<R> ff1 = io.get(A.ff, 0, i1) - io.get(A.ff, 0, i);
    t = (t + mk(c3 * nl1.x / dx, c3 * nl1.y / dx)) + mk(ff1.x / dx, ff1.y / dx);
    C2<R> nl2 = io.get(A.nl, 1, i1) + io.get(A.nl, 1, i);
    C2<R> ff2 = io.get(A.ff, 1, i1) + io.get(A.ff, 1, i);
    C2<R> a2 = ciR(c3 * wy, nl2), b2 = ciR(wy, ff2);
    t = (t + mk(a2.x / (R)2, a2.y / (R)2)) + mk(b2.x / (R)2, b2.y / (R)2);
    C2<R> nl3 = io.get(A.nl, 2, i1) + io.get(A.nl, 2, i);
    C2<R> ff3 = io.get(A.ff, 2, i1) + io.get(A.ff, 2, i);
    C2<R> a3 = ciR(c3 * wz, nl3), b3 = ciR(wz, ff3);
    t = (t + mk(a3.x / (R)2, a3.y / (R)2)) + mk(b3.x / (R)2, b3.y / (R)2);
    C2<R> on1 = io.get(A.onl, 0, i1) - io.get(A.onl, 0, i);
    t = t + mk(c4 * on1.x / dx, c4 * on1.y / dx);
    C2<R> on2 = ciR(c4 * wy, io.get(A.onl, 1, i1) + io.get(A.onl, 1, i));
    t = t + mk(on2.x / (R)2, on2.y / (R)2);
    C2<R> on3 = ciR(c4 * wz, io.get(A.onl, 2, i1) + io.get(A.onl, 2, i));
    t = t + mk(on3.x / (R)2, on3.y / (R)2);
    t = mk((R)1 * t.x, (R)1 * t.y);  // half*(g(i)+g(i+1)) with g = 1
    return mk(t.x / (c1 + c2), t.y / (c1 + c2));
  };
  R vdotz = 0, vdotyr = 0, vdotyi = 0;
  if (mean) {
    // non-cyclic, size nx-1, p(nx) = 0 (poisson.F90:127-142)
    int s2 = thomas3(nx - 1, ap, bp, bp, bp, cp, prhs, gp, yr, zp, yi, ms, true);
    if (s2) status = s2;
    thomas_back(nx - 1, gp, yr, zp, yi, ms, true);
    yr[(size_t)(nx - 1) * ms] = 0;
    for (int i = 0; i < nx; ++i) {
      yi[(size_t)i * ms] = 0;
      zp[(size_t)i * ms] = 0;
    }
  } else {
    const R b0 = bp - (R)1 * (cp / (R)1);
    const R bn = bp - (R)1 * (ap / (R)1);
    int s2 = thomas3(nx, ap, bp, b0, bn, cp, prhs, gp, yr, zp, yi, ms, true);
    if (s2) status = s2;
    thomas_back(nx, gp, yr, zp, yi, ms, true);
    const R auxv0 = cp / (R)1, auxvn = ap / (R)1;
    size_t ln = (size_t)(nx - 1) * ms;
    vdotz = auxv0 * zp[0] + auxvn * zp[ln];
    vdotyr = auxv0 * yr[0] + auxvn * yr[ln];
    vdotyi = auxv0 * yi[0] + auxvn * yi[ln];
  }
  for (int i = 0; i < nx; ++i) {
    size_t o = (size_t)i * ms;
    R tr = yr[o] - vdotyr * zp[o] / ((R)1 + vdotz);
    R ti = yi[o] - vdotyi * zp[o] / ((R)1 + vdotz);
    C2<R> pold = io.get(A.p, 0, i);
    C2<R> d = mk(tr - pold.x, ti - pold.y);
    io.put(A.p, 0, i, pold + mk((R)1 * d.x, (R)1 * d.y));
  }
  // ---------------- ustep (usteptime.F90:45-258) ----------------
  const R g = (R)1;
  const R ah = (-g * dt * vis * c2 / dx2);
  const R bh = ((R)1 + dt * vis * c2 * (g * (R)2 / dx2 + w2));
  const R ch = ah;
  const R b0h = bh - (R)1 * (ch / (R)1);
  const R bnh = bh - (R)1 * (ah / (R)1);
  const R auxv0h = ch / (R)1, auxvnh = ah / (R)1;
  const R c12 = c1 + c2;
  R vdotzh = 0;
  for (int c = 0; c < 3; ++c) {
    const R kk = c == 1 ? wy : wz;
    auto urhs = [&](int i) -> C2<R> {
      C2<R> uu = io.get(A.u, c, i);
      R f = (R)1 - dt * vis * c1 * w2;
      C2<R> r = mk(f * uu.x, f * uu.y);
      C2<R> n = io.get(A.nl, c, i), fv = io.get(A.ff, c, i), on = io.get(A.onl, c, i);
      r = (r + mk(dt * c3 * n.x, dt * c3 * n.y)) + mk(dt * fv.x, dt * fv.y);
      r = r + mk(dt * c4 * on.x, dt * c4 * on.y);
      int im = wrapn(i - 1, nx), ip = wrapn(i + 1, nx);
      C2<R> pi_ = io.get(A.p, 0, i), pm = io.get(A.p, 0, im);
      if (c == 0) {
        C2<R> dpv = pi_ - pm;
        R s = dt * c12;
        r = r - mk(s * dpv.x / dx, s * dpv.y / dx);
      } else {
        R s = dt * c12 * (R)0.5;
        C2<R> sp = i == 0 ? (pm + pi_) : (pm + pi_);
        r = r - ciR(kk, mk(s * sp.x, s * sp.y));
      }
      C2<R> um = io.get(A.u, c, im), up = io.get(A.u, c, ip), uc = io.get(A.u, c, i);
      C2<R> lap;
      if (i == 0) lap = (up - mk((R)2 * uc.x, (R)2 * uc.y)) + um;
      else lap = (um + up) - mk((R)2 * uc.x, (R)2 * uc.y);
      R s = g * dt * vis * c1;
      return r + mk(s * lap.x / dx2, s * lap.y / dx2);
    };
    int s2 = thomas3(nx, ah, bh, b0h, bnh, ch, urhs, gh, yr, zh, yi, ms, c == 0);
    if (s2) status = s2;
    thomas_back(nx, gh, yr, zh, yi, ms, c == 0);
    size_t ln = (size_t)(nx - 1) * ms;
    if (c == 0) vdotzh = auxv0h * zh[0] + auxvnh * zh[ln];
    R vr = auxv0h * yr[0] + auxvnh * yr[ln];
    R vi = auxv0h * yi[0] + auxvnh * yi[ln];
    for (int i = 0; i < nx; ++i) {
      size_t o = (size_t)i * ms;
      io.put(A.u, c, i, mk(yr[o] - vr * zh[o] / ((R)1 + vdotzh), yi[o] - vi * zh[o] / ((R)1 + vdotzh)));
    }
  }
  // ---------------- divcorr (poisson.F90:241-508) ----------------
  const R alp = (c1 + c2) * dt;
  auto drhs = [&](int i) -> C2<R> {
    int i1 = wrapn(i + 1, nx);
    C2<R> d = io.get(A.u, 0, i1) - io.get(A.u, 0, i);
    C2<R> t = mk(d.x / dx, d.y / dx);
    C2<R> s2v = ciR(wy, io.get(A.u, 1, i) + io.get(A.u, 1, i1));
    t = t + mk((R)0.5 * s2v.x, (R)0.5 * s2v.y);
    C2<R> s3v = ciR(wz, io.get(A.u, 2, i) + io.get(A.u, 2, i1));
    t = t + mk((R)0.5 * s3v.x, (R)0.5 * s3v.y);
    return mk(((R)0.5 * ((R)1 + (R)1) * t.x) / alp, ((R)0.5 * ((R)1 + (R)1) * t.y) / alp);
  };
  if (mean) {
    thomas3(nx - 1, ap, bp, bp, bp, cp, drhs, gp, yr, zp, yi, ms, false);
    thomas_back(nx - 1, gp, yr, zp, yi, ms, false);
    yr[(size_t)(nx - 1) * ms] = 0;
    for (int i = 0; i < nx; ++i) yi[(size_t)i * ms] = 0;
    vdotyr = 0;
    vdotyi = 0;
  } else {
    const R b0 = bp - (R)1 * (cp / (R)1);
    const R bn = bp - (R)1 * (ap / (R)1);
    thomas3(nx, ap, bp, b0, bn, cp, drhs, gp, yr, zp, yi, ms, false);
    thomas_back(nx, gp, yr, zp, yi, ms, false);
    const R auxv0 = cp / (R)1, auxvn = ap / (R)1;
    size_t ln = (size_t)(nx - 1) * ms;
    vdotyr = auxv0 * yr[0] + auxvn * yr[ln];
    vdotyi = auxv0 * yi[0] + auxvn * yi[ln];
  }
  // div(i) into (yr, yi) in place
  for (int i = 0; i < nx; ++i) {
    size_t o = (size_t)i * ms;
    R tr = yr[o] - vdotyr * zp[o] / ((R)1 + vdotz);
    R ti = yi[o] - vdotyi * zp[o] / ((R)1 + vdotz);
    yr[o] = tr;
    yi[o] = ti;
  }
  auto divv = [&](int i) -> C2<R> { size_t o = (size_t)i * ms; return mk(yr[o], yi[o]); };
  // u(t+dt) = u* - grad(phi)
  for (int i = 0; i < nx; ++i) {
    C2<R> di = divv(i), dm = divv(wrapn(i - 1, nx));
    C2<R> u1 = io.get(A.u, 0, i), u2 = io.get(A.u, 1, i), u3 = io.get(A.u, 2, i);
    C2<R> dd = di - dm;
    C2<R> sm = i == 0 ? (di + dm) : (dm + di);
    u1 = u1 - mk(alp * dd.x / dx, alp * dd.y / dx);
    u2 = u2 - ciR(alp * (R)0.5 * wy, sm);
    u3 = u3 - ciR(alp * (R)0.5 * wz, sm);
    io.put(A.u, 0, i, u1);
    io.put(A.u, 1, i, u2);
    io.put(A.u, 2, i, u3);
  }
  // pressure update and the new divergence (check_divergence)
  const R cv = (c2 * dt) * vis;
  double dmax = 0.0;
  for (int i = 0; i < nx; ++i) {
    C2<R> dm = divv(wrapn(i - 1, nx)), di = divv(i), dp = divv(wrapn(i + 1, nx));
    C2<R> lap = (dm - mk((R)2 * di.x, (R)2 * di.y)) + dp;
    C2<R> corr = mk(lap.x / dx2, lap.y / dx2) - mk(w2 * di.x, w2 * di.y);
    C2<R> pv = io.get(A.p, 0, i);
    io.put(A.p, 0, i, (pv + di) - mk(cv * corr.x, cv * corr.y));
    int i1 = wrapn(i + 1, nx);
    C2<R> d = io.get(A.u, 0, i1) - io.get(A.u, 0, i);
    C2<R> t = mk(d.x / dx, d.y / dx);
    C2<R> s2v = ciR(wy, io.get(A.u, 1, i) + io.get(A.u, 1, i1));
    t = t + mk((R)0.5 * s2v.x, (R)0.5 * s2v.y);
    C2<R> s3v = ciR(wz, io.get(A.u, 2, i) + io.get(A.u, 2, i1));
    t = t + mk((R)0.5 * s3v.x, (R)0.5 * s3v.y);
    double mg = sqrt((double)t.x * (double)t.x + (double)t.y * (double)t.y);
    if (mg > dmax) dmax = mg;
  }
  divmax = dmax;
  return status;
}

template <typename T>
__global__ void __launch_bounds__(256) k_velstep(KArgs<T> A, T* scr, double* scr64, int* status) {
  __shared__ double sh[256];
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  int nm = A.my2 * A.mz;
  double dmax = 0.0;
  if (q < nm) {
    int st;
    if (q == 0) st = run_mode<T, double>(A, 0, scr64, 1, 0, dmax);
    else st = run_mode<T, T>(A, q, scr, (size_t)nm, (size_t)q, dmax);
    if (st) *status = st;
  }
  double m = block_max(dmax, sh);
  if (threadIdx.x == 0) A.part[blockIdx.x] = m;
}
