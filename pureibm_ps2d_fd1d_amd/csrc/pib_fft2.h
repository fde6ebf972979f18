// pib_fft2.h -- plane FFT for power-of-two plane sizes, two stages per axis.
//
// Same transforms and the same Op interface as pib_fft.h (ff2rc / ff2cr of
// fftw_interface.F90:256-338), for my/2 = HY and mz = MZ known at compile
// time.  Each 1-D transform of length N = R1*R2 is a four-step FFT whose two
// stages run in registers (radix-4/8/16), with the stage that stays in LDS
// arranged to work in place:
//   decimation in time  n = n1 + R1 n2, k = k2 + R2 k1:
//     A: per n1, DFT_R2 over n2, times W_N^(n1 k2), to slot n1 R2 + k2
//     B: per k2, DFT_R1 over n1 of slots n1 R2 + k2, to slots k2 + R2 k1
//        (the same slots: in place)
//   decimation in frequency  n = R2 a + b, k = c + R1 d:
//     A: per b, DFT_R1 over a of slots R2 a + b, times W_N^(b c), to slots
//        R2 c + b (the same slots: in place)
//     B: per c, DFT_R2 over b of slots R2 c + b, giving X[c + R1 d]
// so that the first stage of a transform can read straight from HBM (DIT)
// and the last can write straight to HBM (DIF):
//   c2r: z DIT-A (HBM -> LDS) | z DIT-B | merge | y DIF-A | y DIF-B (-> HBM)
//   r2c: y DIT-A (HBM -> LDS) | y DIT-B | split | z DIF-A | z DIF-B (-> HBM)
// four barriers and eight plane-sized LDS sweeps per transform.  Results are
// in natural order, so no slot tables are needed.
//
// Every phase walks its items as e = threadIdx.x + j*NT, j < ceil(items/NT).
// `it` of op.store / op.store2 is j*R + r (walk round j, output r of the
// item): below r2c_nout / c2r_nout.
#pragma once
#include "pib_fft.h"

template <int N>
struct Split2;
template <>
struct Split2<8> {
  static constexpr int A = 2, B = 4;
};
template <>
struct Split2<16> {
  static constexpr int A = 4, B = 4;
};
template <>
struct Split2<32> {
  static constexpr int A = 4, B = 8;
};
template <>
struct Split2<64> {
  static constexpr int A = 8, B = 8;
};
template <>
struct Split2<128> {
  static constexpr int A = 8, B = 16;
};
template <>
struct Split2<256> {
  static constexpr int A = 16, B = 16;
};
template <>
struct Split2<48> {
  static constexpr int A = 6, B = 8;
};
template <>
struct Split2<96> {
  static constexpr int A = 8, B = 12;
};
template <>
struct Split2<80> {
  static constexpr int A = 10, B = 8;
};
template <>
struct Split2<160> {
  static constexpr int A = 10, B = 16;
};

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// e^{-2 pi i j/16} for the products j = n1*k2 of a 4 x 4 split
template <typename T>
__device__ __forceinline__ C2<T> w16(int j) {
  const T c1 = (T)0.92387953251128675612818318939678828682;  // cos(pi/8)
  const T s1 = (T)0.38268343236508977172845998403039886676;  // sin(pi/8)
  const T h = (T)0.70710678118654752440084436210484903928;   // cos(pi/4)
  switch (j) {
    case 1: return mk(c1, -s1);
    case 2: return mk(h, -h);
    case 3: return mk(s1, -c1);
    case 4: return mk((T)0, (T)-1);
    case 6: return mk(-h, -h);
    case 9: return mk(-c1, s1);
    default: return mk((T)1, (T)0);
  }
}

template <bool INV, typename T>
__device__ __forceinline__ void dft16(C2<T>* v) {
  C2<T> a[4][4];
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) a[n1][n2] = v[n1 + 4 * n2];
    dft4<INV>(a[n1]);
#pragma unroll
    for (int k2 = 1; k2 < 4; ++k2)
      if (n1 > 0) {
        C2<T> w = w16<T>(n1 * k2);
        if (INV) w.y = -w.y;
        a[n1][k2] = cmul(a[n1][k2], w);
      }
  }
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    C2<T> b[4] = {a[0][k2], a[1][k2], a[2][k2], a[3][k2]};
    dft4<INV>(b);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) v[k2 + 4 * k1] = b[k1];
  }
}
// e^{-2 pi i j/N} for the products j = n1*k2 of the composite splits below
// (6 = 2 x 3, 10 = 2 x 5, 12 = 4 x 3)
template <typename T, int N>
__device__ __forceinline__ C2<T> wN(int j) {
  if constexpr (N == 6) {
    switch (j) {
      case 1: return mk((T)0.5, (T)-0.86602540378443864676372317075293618);
      case 2: return mk((T)-0.5, (T)-0.86602540378443864676372317075293618);
      default: return mk((T)1, (T)0);
    }
  } else if constexpr (N == 10) {
    switch (j) {
      case 1: return mk((T)0.80901699437494742410229341718281906, (T)-0.58778525229247312916870595463907277);
      case 2: return mk((T)0.30901699437494742410229341718281906, (T)-0.95105651629515357211643933337938214);
      case 3: return mk((T)-0.30901699437494742410229341718281906, (T)-0.95105651629515357211643933337938214);
      case 4: return mk((T)-0.80901699437494742410229341718281906, (T)-0.58778525229247312916870595463907277);
      default: return mk((T)1, (T)0);
    }
  } else {
    switch (j) {
      case 1: return mk((T)0.86602540378443864676372317075293618, (T)-0.5);
      case 2: return mk((T)0.5, (T)-0.86602540378443864676372317075293618);
      case 3: return mk((T)0, (T)-1);
      case 4: return mk((T)-0.5, (T)-0.86602540378443864676372317075293618);
      case 6: return mk((T)-1, (T)0);
      default: return mk((T)1, (T)0);
    }
  }
}
template <bool INV, int R, typename T>
__device__ __forceinline__ void dft2n(C2<T>* v);
// N = R1 R2 as a four-step DFT in registers: n = n1 + R1 n2, k = k2 + R2 k1
template <bool INV, int R1, int R2, typename T>
__device__ __forceinline__ void dft_ct(C2<T>* v) {
  C2<T> a[R1][R2];
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) a[n1][n2] = v[n1 + R1 * n2];
    dft2n<INV, R2>(a[n1]);
#pragma unroll
    for (int k2 = 1; k2 < R2; ++k2)
      if (n1 > 0) {
        C2<T> w = wN<T, R1 * R2>(n1 * k2);
        if (INV) w.y = -w.y;
        a[n1][k2] = cmul(a[n1][k2], w);
      }
  }
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) {
    C2<T> b[R1];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) b[k1] = a[k1][k2];
    dft2n<INV, R1>(b);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) v[k2 + R2 * k1] = b[k1];
  }
}
template <bool INV, int R, typename T>
__device__ __forceinline__ void dft2n(C2<T>* v) {
  if constexpr (R == 2) dft2<INV>(v);
  else if constexpr (R == 3) dft3<INV>(v);
  else if constexpr (R == 4) dft4<INV>(v);
  else if constexpr (R == 5) dft5<INV>(v);
  else if constexpr (R == 6) dft_ct<INV, 2, 3>(v);
  else if constexpr (R == 8) dft8<INV>(v);
  else if constexpr (R == 10) dft_ct<INV, 2, 5>(v);
  else if constexpr (R == 12) dft_ct<INV, 4, 3>(v);
  else dft16<INV>(v);
}
// v * W^j (forward) or v * conj(W^j) (inverse), W^j from a table in LDS
template <bool INV, typename T>
__device__ __forceinline__ C2<T> twmul(C2<T> v, const C2<T>* tw, int j) {
  if (j == 0) return v;
  C2<T> w = tw[j];
  if (INV) w.y = -w.y;
  return cmul(v, w);
}

// LDS layout of the plane: row z holds the HY + 1 entries of ky (or y pair)
// at slots ys(y).  For HY = 64 the rows are padded: one empty slot after
// every 8 entries and the ky = HY entry last (slot 71), with rows 72 slots
// apart.  Then the strided y-stage accesses (stride 8 within a row, two rows
// per 16 lanes) fall in distinct banks.  The z stages walk the rows by slot
// (W slots per row, empty slots idle) so that their 16-lane groups stay on
// consecutive slots too.
template <int HY, int MZ, int NT>
struct Plane2 {
  static constexpr bool PAD = HY == 64;
  static constexpr int MY2 = HY + 1, LD = PAD ? HY + HY / 8 : HY + 1, W = LD;
  __device__ static __forceinline__ int ys(int y) {
    if constexpr (PAD) return y < HY ? y + (y >> 3) : HY + HY / 8 - 1;
    else return y;
  }
  // the ky of slot s, or -1 for an empty slot
  __device__ static __forceinline__ int yof(int s) {
    if constexpr (PAD) {
      if (s == HY + HY / 8 - 1) return HY;
      return s % 9 == 8 || s >= HY + HY / 8 - 1 ? -1 : s - s / 9;
    } else {
      return s;
    }
  }
  static constexpr int ZA = Split2<MZ>::A, ZB = Split2<MZ>::B;
  static constexpr int YA = Split2<HY>::A, YB = Split2<HY>::B;
  // outputs per lane of the last phase (the Op's `it` bound)
  static constexpr int c2r_nout = cdiv(MZ * YA, NT) * YB;  // y DIF-B items (z, c), YB outputs each
  // z DIF-B items (slot of ky < HY, c), ZB outputs each; the ky = HY
  // column's items go to lanes c < ZA with it = -1
  static constexpr int r2c_nout = cdiv(W * ZA, NT) * ZB;
};

// ---------------------------------------------------------------------------
// Inverse: (HY+1, MZ) half spectrum -> real plane, unscaled; Im of the ky = 0
// and ky = HY entries is ignored after the z transform.
//   z: DIT with R1 = ZA, R2 = ZB;  y: DIF with R1 = YA, R2 = YB.
template <int HY, int MZ, int NT, typename T, class Op>
__device__ void plane_c2r2(C2<T>* buf, const PlaneGeom& G, const PlaneTabs<T>& gtb, Op& op) {
  using P = Plane2<HY, MZ, NT>;
  constexpr int LD = P::LD, W = P::W, ZA = P::ZA, ZB = P::ZB, YA = P::YA, YB = P::YB;
  const int t = threadIdx.x;
  const PlaneTabs<T> tb = stage_tabs(G, gtb, (unsigned char*)buf + G.tab_off);
  __syncthreads();  // the first stage already multiplies by twiddles
  // z DIT-A: item (slot s, n1), s fastest; reads HBM
#pragma unroll
  for (int j = 0; j < cdiv(W * ZA, NT); ++j) {
    const int e = t + j * NT;
    const int s = e % W, n1 = e / W, ky = P::yof(s);
    if (e < W * ZA && ky >= 0) {
      C2<T> v[ZB];
#pragma unroll
      for (int n2 = 0; n2 < ZB; ++n2) v[n2] = op.load(ky, n1 + ZA * n2);
      dft2n<true, ZB>(v);
#pragma unroll
      for (int k2 = 0; k2 < ZB; ++k2) buf[(n1 * ZB + k2) * LD + s] = twmul<true>(v[k2], tb.twz, n1 * k2);
    }
  }
  __syncthreads();
  // z DIT-B: item (slot s, k2), in place
#pragma unroll
  for (int j = 0; j < cdiv(W * ZB, NT); ++j) {
    const int e = t + j * NT;
    const int s = e % W, k2 = e / W;
    if (e < W * ZB && P::yof(s) >= 0) {
      C2<T> v[ZA];
#pragma unroll
      for (int n1 = 0; n1 < ZA; ++n1) v[n1] = buf[(n1 * ZB + k2) * LD + s];
      dft2n<true, ZA>(v);
#pragma unroll
      for (int k1 = 0; k1 < ZA; ++k1) buf[(k2 + ZB * k1) * LD + s] = v[k1];
    }
  }
  __syncthreads();
  // merge: Z_k = A_k + i B_k, A_k = X_k + conj X_{HY-k},
  // B_k = (X_k - conj X_{HY-k}) e^{+2 pi i k/my}; pairs (k, HY-k) in place
  constexpr int NP = HY / 2 + 1;
#pragma unroll
  for (int j = 0; j < cdiv(MZ * NP, NT); ++j) {
    const int e = t + j * NT;
    if (e < MZ * NP) {
      const int k = e % NP, z = e / NP;
      C2<T>* row = buf + z * LD;
      if (k == 0) {
        T x0 = row[0].x, xh = row[P::ys(HY)].x;
        row[0] = mk(x0 + xh, x0 - xh);
      } else if (2 * k == HY) {
        C2<T> a = row[P::ys(k)];
        row[P::ys(k)] = mk(a.x + a.x, -(a.y + a.y));
      } else {
        C2<T> a = row[P::ys(k)], b = row[P::ys(HY - k)];
        C2<T> A = mk(a.x + b.x, a.y - b.y);
        C2<T> D = mk(a.x - b.x, a.y + b.y);
        C2<T> B = cmul(D, cconj(tb.wr[k]));
        row[P::ys(k)] = mk(A.x - B.y, A.y + B.x);
        row[P::ys(HY - k)] = mk(A.x + B.y, B.x - A.y);
      }
    }
  }
  __syncthreads();
  // y DIF-A: item (z, b), b fastest; slots R2 a + b, in place
#pragma unroll
  for (int j = 0; j < cdiv(MZ * YB, NT); ++j) {
    const int e = t + j * NT;
    if (e < MZ * YB) {
      const int b = e % YB, z = e / YB;
      C2<T>* row = buf + z * LD;
      C2<T> v[YA];
#pragma unroll
      for (int a = 0; a < YA; ++a) v[a] = row[P::ys(YB * a + b)];
      dft2n<true, YA>(v);
#pragma unroll
      for (int c = 0; c < YA; ++c) row[P::ys(YB * c + b)] = twmul<true>(v[c], tb.twy, b * c);
    }
  }
  __syncthreads();
  // y DIF-B: item (z, c), writes y = 2m, 2m+1 for m = c + YA d
#pragma unroll
  for (int j = 0; j < cdiv(MZ * YA, NT); ++j) {
    const int e = t + j * NT;
    if (e < MZ * YA) {
      const int c = e % YA, z = e / YA;
      const C2<T>* row = buf + z * LD;
      C2<T> v[YB];
#pragma unroll
      for (int b = 0; b < YB; ++b) v[b] = row[P::ys(YB * c + b)];
      dft2n<true, YB>(v);
#pragma unroll
      for (int d = 0; d < YB; ++d) op.store2(j * YB + d, 2 * (c + YA * d), z, v[d].x, v[d].y);
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Forward: real plane -> (HY+1, MZ) half spectrum scaled by 1/(my*mz).
//   y: DIT with R1 = YA, R2 = YB;  z: DIF with R1 = ZA, R2 = ZB.
template <int HY, int MZ, int NT, typename T, class Op>
__device__ void plane_r2c2(C2<T>* buf, const PlaneGeom& G, const PlaneTabs<T>& gtb, Op& op) {
  using P = Plane2<HY, MZ, NT>;
  constexpr int LD = P::LD, W = P::W, ZA = P::ZA, ZB = P::ZB, YA = P::YA, YB = P::YB;
  const int t = threadIdx.x;
  const PlaneTabs<T> tb = stage_tabs(G, gtb, (unsigned char*)buf + G.tab_off);
  __syncthreads();  // the first stage already multiplies by twiddles
  // y DIT-A: item (z, n1), n1 fastest; pairs (y, y+1) from HBM as one complex
#pragma unroll
  for (int j = 0; j < cdiv(MZ * YA, NT); ++j) {
    const int e = t + j * NT;
    if (e < MZ * YA) {
      const int n1 = e % YA, z = e / YA;
      C2<T> v[YB];
#pragma unroll
      for (int n2 = 0; n2 < YB; ++n2) op.load2(2 * (n1 + YA * n2), z, v[n2].x, v[n2].y);
      dft2n<false, YB>(v);
      C2<T>* row = buf + z * LD;
#pragma unroll
      for (int k2 = 0; k2 < YB; ++k2) row[P::ys(n1 * YB + k2)] = twmul<false>(v[k2], tb.twy, n1 * k2);
    }
  }
  __syncthreads();
  // y DIT-B: item (z, k2), in place
#pragma unroll
  for (int j = 0; j < cdiv(MZ * YB, NT); ++j) {
    const int e = t + j * NT;
    if (e < MZ * YB) {
      const int k2 = e % YB, z = e / YB;
      C2<T>* row = buf + z * LD;
      C2<T> v[YA];
#pragma unroll
      for (int n1 = 0; n1 < YA; ++n1) v[n1] = row[P::ys(n1 * YB + k2)];
      dft2n<false, YA>(v);
#pragma unroll
      for (int k1 = 0; k1 < YA; ++k1) row[P::ys(k2 + YB * k1)] = v[k1];
    }
  }
  __syncthreads();
  // split: X_k = E_k + w^k O_k, X_{hy-k} = conj(E_k - w^k O_k); pairs in place
  constexpr int NP = HY / 2 + 1;
#pragma unroll
  for (int j = 0; j < cdiv(MZ * NP, NT); ++j) {
    const int e = t + j * NT;
    if (e < MZ * NP) {
      const int k = e % NP, z = e / NP;
      C2<T>* row = buf + z * LD;
      if (k == 0) {
        C2<T> z0 = row[0];
        row[0] = mk(z0.x + z0.y, (T)0);
        row[P::ys(HY)] = mk(z0.x - z0.y, (T)0);
      } else if (2 * k == HY) {
        row[P::ys(k)] = cconj(row[P::ys(k)]);
      } else {
        C2<T> a = row[P::ys(k)], b = row[P::ys(HY - k)];
        C2<T> E = mk((a.x + b.x) * (T)0.5, (a.y - b.y) * (T)0.5);
        C2<T> O = mk((a.y + b.y) * (T)0.5, (b.x - a.x) * (T)0.5);
        C2<T> wO = cmul(tb.wr[k], O);
        row[P::ys(k)] = E + wO;
        row[P::ys(HY - k)] = cconj(E - wO);
      }
    }
  }
  __syncthreads();
  // z DIF-A: item (slot s, b), s fastest; rows R2 a + b, in place
#pragma unroll
  for (int j = 0; j < cdiv(W * ZB, NT); ++j) {
    const int e = t + j * NT;
    const int s = e % W, b = e / W;
    if (e < W * ZB && P::yof(s) >= 0) {
      C2<T> v[ZA];
#pragma unroll
      for (int a = 0; a < ZA; ++a) v[a] = buf[(ZB * a + b) * LD + s];
      dft2n<false, ZA>(v);
#pragma unroll
      for (int c = 0; c < ZA; ++c) buf[(ZB * c + b) * LD + s] = twmul<false>(v[c], tb.twz, b * c);
    }
  }
  __syncthreads();
  // z DIF-B: item (slot s, c), outputs kz = c + ZA d; columns ky < HY
  // first, then the ky = HY column on lanes c < ZA (it = -1)
  const T scale = (T)((double)(2 * HY) * (double)MZ);
#pragma unroll
  for (int j = 0; j < cdiv(W * ZA, NT); ++j) {
    const int e = t + j * NT;
    const int s = e % W, c = e / W, ky = P::yof(s);
    if (e < W * ZA && ky >= 0 && ky < HY) {
      C2<T> v[ZB];
#pragma unroll
      for (int b = 0; b < ZB; ++b) v[b] = buf[(ZB * c + b) * LD + s];
      dft2n<false, ZB>(v);
#pragma unroll
      for (int d = 0; d < ZB; ++d) op.store(j * ZB + d, ky, c + ZA * d, mk(v[d].x / scale, v[d].y / scale));
    }
  }
  if (t < ZA) {
    const int c = t;
    C2<T> v[ZB];
#pragma unroll
    for (int b = 0; b < ZB; ++b) v[b] = buf[(ZB * c + b) * LD + P::ys(HY)];
    dft2n<false, ZB>(v);
#pragma unroll
    for (int d = 0; d < ZB; ++d) op.store(-1, HY, c + ZA * d, mk(v[d].x / scale, v[d].y / scale));
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Compiled plane shapes: SH selects (HY, MZ) and the workgroup size; SH = 0
// is the generic mixed-radix transform of pib_fft.h (any 2,3,5-smooth size).
template <int SH>
struct Shape {
  static constexpr int HY = 0, MZ = 0, NT = 0;
};
template <>
struct Shape<1> {
  static constexpr int HY = 64, MZ = 128, NT = 1024;
};
template <>
struct Shape<2> {
  static constexpr int HY = 32, MZ = 64, NT = 256;
};
template <>
struct Shape<3> {
  static constexpr int HY = 16, MZ = 32, NT = 128;
};
// the 96^2 and 160^2 planes (radix 3 and 5 in the split lengths)
template <>
struct Shape<4> {
  static constexpr int HY = 48, MZ = 96, NT = 512;
};
template <>
struct Shape<5> {
  static constexpr int HY = 80, MZ = 160, NT = 1024;
};
static inline int shape_of(int hy, int mz) {
  if (hy == 64 && mz == 128) return 1;
  if (hy == 32 && mz == 64) return 2;
  if (hy == 16 && mz == 32) return 3;
  if (hy == 48 && mz == 96) return 4;
  if (hy == 80 && mz == 160) return 5;
  return 0;
}
template <int SH>
constexpr int shape_ld_c() {
  return Plane2<Shape<SH>::HY, Shape<SH>::MZ, Shape<SH>::NT>::LD;
}
// LDS row stride (in complex entries) of a compiled shape's plane buffer
static inline int shape_ld(int sh) {
  switch (sh) {
    case 1: return shape_ld_c<1>();
    case 2: return shape_ld_c<2>();
    case 3: return shape_ld_c<3>();
    case 4: return shape_ld_c<4>();
    case 5: return shape_ld_c<5>();
    default: return 0;
  }
}
static inline int shape_nt(int sh) {
  switch (sh) {
    case 1: return Shape<1>::NT;
    case 2: return Shape<2>::NT;
    case 3: return Shape<3>::NT;
    case 4: return Shape<4>::NT;
    case 5: return Shape<5>::NT;
    default: return 0;
  }
}
// lds: the workgroup's LDS, which holds the plane (and the tables); hbm: the
// workgroup's plane slot in HBM, or null.  The generic transform (SH = 0)
// works in the slot when there is one; the two calls are separate so that
// each is compiled for its own address space.  A compiled shape always
// works in LDS.
template <int SH, int NIT, typename T, class Op>
__device__ __forceinline__ void plane_c2r_sh(C2<T>* hbm, const PlaneGeom& G, const PlaneTabs<T>& tb, Op& op,
                                             unsigned char* lds) {
  if constexpr (SH == 0) {
    if (hbm) plane_c2r<NIT>(hbm, G, tb, op, lds);
    else plane_c2r<NIT>((C2<T>*)lds, G, tb, op, lds);
  } else {
    plane_c2r2<Shape<SH>::HY, Shape<SH>::MZ, Shape<SH>::NT>((C2<T>*)lds, G, tb, op);
  }
}
template <int SH, int NIT, typename T, class Op>
__device__ __forceinline__ void plane_r2c_sh(C2<T>* hbm, const PlaneGeom& G, const PlaneTabs<T>& tb, Op& op,
                                             unsigned char* lds) {
  if constexpr (SH == 0) {
    if (hbm) plane_r2c<NIT>(hbm, G, tb, op, lds);
    else plane_r2c<NIT>((C2<T>*)lds, G, tb, op, lds);
  } else {
    plane_r2c2<Shape<SH>::HY, Shape<SH>::MZ, Shape<SH>::NT>((C2<T>*)lds, G, tb, op);
  }
}
