// pib_fft.h -- LDS-resident 2-D real<->complex FFT of one (y, z) plane per
// workgroup, for gfx950.
//
// Semantics are those of the reference's transform layer
// (fftw_interface.F90:256-338):
//   ff2rc: b(ky,kz) = sum_{y,z} a(y,z) e^{-2 pi i (ky y/my + kz z/mz)} / (my*mz),
//          ky = 0..my/2 (the contiguous y axis is the halved one);
//   ff2cr: a(y,z) = sum_{ky,kz} b e^{+2 pi i (...)}, unscaled, computed
//          row-column: inverse c2c along z for every ky column, then c2r along
//          y ignoring Im of the ky = 0 and ky = my/2 entries.
//
// Layout in LDS: the plane is an array of mz rows of `ld` complex slots
// (ld >= my/2+1).  The real y-transform of length my is done as a complex
// transform of length hy = my/2 on (x[2m], x[2m+1]) pairs with the standard
// split/merge step; the z-transform runs down the my/2+1 columns.  Every 1-D
// transform is an in-place mixed-radix (2,3,4,5,8) decimation-in-frequency
// FFT: natural-order input, digit-reversed output (slot tables revy/revz),
// so no reordering pass is needed between the y and z passes.
//
// Loads and stores go through an `Op` functor so that the producer or
// consumer arithmetic of each use (derivatives, products, means, reductions)
// is fused into the transform:
//   r2c: op.load2(y, z, &a, &b)  values at (y, z), (y+1, z), y even
//        op.store(it, ky, kz, v) spectral value (already scaled)
//   c2r: op.load(ky, kz) -> v
//        op.store2(it, y, z, a, b)  physical values at (y, z), (y+1, z)
// `it` is the iteration index of the calling thread's fixed element walk
// (element e = threadIdx.x + it*blockDim.x), so an op may keep per-thread
// register accumulators across several transforms.  NIT > 0 unrolls that
// walk for a known bound (NIT*blockDim.x covers the plane); NIT = 0 is the
// plain strided loop for ops without accumulators.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

template <typename T>
struct C2 {
  T x, y;
};
template <typename T>
__device__ __forceinline__ C2<T> mk(T a, T b) {
  C2<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ C2<T> operator+(C2<T> a, C2<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ C2<T> operator-(C2<T> a, C2<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ C2<T> cconj(C2<T> a) { return mk(a.x, -a.y); }
template <typename T>
__device__ __forceinline__ C2<T> cscal(T s, C2<T> a) { return mk(s * a.x, s * a.y); }
// (i k) * z for a pure-imaginary wavenumber i k
template <typename T>
__device__ __forceinline__ C2<T> cik(T k, C2<T> z) { return mk(-k * z.y, k * z.x); }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename T>
__device__ __forceinline__ C2<T> rot90(C2<T> a) {
  return INV ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

// In-place DIF plan of one 1-D length: radices applied in order.
struct Fft1D {
  int n, np;
  int rad[8];
};

struct PlaneGeom {
  int my, mz, my2, hy, ld;
  int tab_off;  // byte offset from the plane buffer to the LDS copy of the tables
  Fft1D py, pz;
};

template <typename T>
struct PlaneTabs {
  const C2<T>* twy;  // e^{-2 pi i t/hy}, t < hy
  const C2<T>* twz;  // e^{-2 pi i t/mz}, t < mz
  const C2<T>* wr;   // e^{-2 pi i k/my}, k <= hy
  const int* revy;   // slot of natural index k in a length-hy DIF output
  const int* revz;   // slot of natural index k in a length-mz DIF output
};

// bytes of LDS that stage_tabs needs after tab_off
template <typename T>
__host__ __device__ inline size_t tab_lds_bytes(int hy, int mz) {
  return (size_t)(2 * hy + mz + 1) * sizeof(C2<T>) + (size_t)(hy + mz) * sizeof(int);
}
// Copies the twiddle and slot tables into LDS (they are read in every pass);
// visible after the next barrier, which every transform's load phase ends
// with.
template <typename T>
__device__ inline PlaneTabs<T> stage_tabs(const PlaneGeom& G, const PlaneTabs<T>& g, unsigned char* lds) {
  C2<T>* twy = (C2<T>*)lds;
  C2<T>* twz = twy + G.hy;
  C2<T>* wr = twz + G.mz;
  int* revy = (int*)(wr + G.hy + 1);
  int* revz = revy + G.hy;
  for (int e = threadIdx.x; e < G.hy; e += blockDim.x) {
    twy[e] = g.twy[e];
    revy[e] = g.revy[e];
  }
  for (int e = threadIdx.x; e < G.mz; e += blockDim.x) {
    twz[e] = g.twz[e];
    revz[e] = g.revz[e];
  }
  for (int e = threadIdx.x; e <= G.hy; e += blockDim.x) wr[e] = g.wr[e];
  PlaneTabs<T> t;
  t.twy = twy;
  t.twz = twz;
  t.wr = wr;
  t.revy = revy;
  t.revz = revz;
  return t;
}

// loads of a plane are issued in chunks of PIB_LD_CHUNK per lane so that
// that many global reads are in flight before the LDS writes
#define PIB_LD_CHUNK 8

// ---- radix-R DFT kernels, y_q = sum_r v_r w^{r q}, w = e^{-/+ 2 pi i/R} ----
template <bool INV, typename T>
__device__ __forceinline__ void dft2(C2<T>* v) {
  C2<T> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}
template <bool INV, typename T>
__device__ __forceinline__ void dft4(C2<T>* v) {
  C2<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
  C2<T> t2 = v[1] + v[3], t3 = rot90<INV>(v[1] - v[3]);
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  v[1] = t1 + t3;
  v[3] = t1 - t3;
}
template <bool INV, typename T>
__device__ __forceinline__ void dft3(C2<T>* v) {
  const T s3 = (T)0.86602540378443864676372317075294;  // sin(2 pi/3)
  C2<T> t = v[1] + v[2];
  C2<T> d = v[1] - v[2];
  C2<T> m = v[0] - cscal((T)0.5, t);
  // forward: y1 = m - i s3 d, y2 = m + i s3 d
  C2<T> s = INV ? mk(-s3 * d.y, s3 * d.x) : mk(s3 * d.y, -s3 * d.x);
  v[0] = v[0] + t;
  v[1] = m + s;
  v[2] = m - s;
}
template <bool INV, typename T>
__device__ __forceinline__ void dft5(C2<T>* v) {
  const T c1 = (T)0.30901699437494742410229341718282;   // cos(2 pi/5)
  const T c2 = (T)-0.80901699437494742410229341718282;  // cos(4 pi/5)
  const T s1 = (T)0.95105651629515357211643933337938;   // sin(2 pi/5)
  const T s2 = (T)0.58778525229247312916870595463907;   // sin(4 pi/5)
  C2<T> a1 = v[1] + v[4], b1 = v[1] - v[4];
  C2<T> a2 = v[2] + v[3], b2 = v[2] - v[3];
  C2<T> m1 = v[0] + cscal(c1, a1) + cscal(c2, a2);
  C2<T> m2 = v[0] + cscal(c2, a1) + cscal(c1, a2);
  // n1 = s1 b1 + s2 b2, n2 = s2 b1 - s1 b2; forward multiplies by -i
  C2<T> n1 = cscal(s1, b1) + cscal(s2, b2);
  C2<T> n2 = cscal(s2, b1) - cscal(s1, b2);
  C2<T> r1 = rot90<INV>(n1), r2 = rot90<INV>(n2);
  v[0] = v[0] + a1 + a2;
  v[1] = m1 + r1;
  v[4] = m1 - r1;
  v[2] = m2 + r2;
  v[3] = m2 - r2;
}
template <bool INV, typename T>
__device__ __forceinline__ void dft8(C2<T>* v) {
  const T h = (T)0.70710678118654752440084436210485;
  // radix-2 x radix-4: even/odd split
  C2<T> e[4] = {v[0], v[2], v[4], v[6]};
  C2<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(e);
  dft4<INV>(o);
  // twiddle o[q] by w8^q
  C2<T> o1 = INV ? mk(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                 : mk(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
  C2<T> o2 = rot90<INV>(o[2]);
  C2<T> o3 = INV ? mk(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                 : mk(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
  v[0] = e[0] + o[0];
  v[4] = e[0] - o[0];
  v[1] = e[1] + o1;
  v[5] = e[1] - o1;
  v[2] = e[2] + o2;
  v[6] = e[2] - o2;
  v[3] = e[3] + o3;
  v[7] = e[3] - o3;
}
template <bool INV, int R, typename T>
__device__ __forceinline__ void dftR(C2<T>* v) {
  if constexpr (R == 2) dft2<INV>(v);
  else if constexpr (R == 3) dft3<INV>(v);
  else if constexpr (R == 4) dft4<INV>(v);
  else if constexpr (R == 5) dft5<INV>(v);
  else dft8<INV>(v);
}

// One DIF pass of radix R over L lines of length n (sub-transform length
// `span`).  line_fast selects which index varies fastest across threads.
template <bool INV, int R, typename T>
__device__ __forceinline__ void dif_pass(C2<T>* buf, int L, int ls, int es, int n, int span,
                                         const C2<T>* __restrict__ tw, bool line_fast) {
  const int stride = span / R;
  const int nb = n / R;
  const int total = L * nb;
  const int twmul = n / span;
  for (int b = threadIdx.x; b < total; b += blockDim.x) {
    int line, bf;
    if (line_fast) {
      line = b % L;
      bf = b / L;
    } else {
      bf = b % nb;
      line = b / nb;
    }
    const int g = bf / stride;
    const int j = bf - g * stride;
    C2<T>* p = buf + line * ls + (g * span + j) * es;
    C2<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = p[r * stride * es];
    dftR<INV, R>(v);
    p[0] = v[0];
#pragma unroll
    for (int q = 1; q < R; ++q) {
      C2<T> w = tw[j * q * twmul];
      if (INV) w.y = -w.y;
      p[q * stride * es] = (j == 0) ? v[q] : cmul(v[q], w);
    }
  }
}

template <bool INV, typename T>
__device__ void fft_lines(C2<T>* buf, int L, int ls, int es, const Fft1D& pl,
                          const C2<T>* __restrict__ tw, bool line_fast) {
  int span = pl.n;
  for (int s = 0; s < pl.np; ++s) {
    const int R = pl.rad[s];
    switch (R) {
      case 2: dif_pass<INV, 2>(buf, L, ls, es, pl.n, span, tw, line_fast); break;
      case 3: dif_pass<INV, 3>(buf, L, ls, es, pl.n, span, tw, line_fast); break;
      case 4: dif_pass<INV, 4>(buf, L, ls, es, pl.n, span, tw, line_fast); break;
      case 5: dif_pass<INV, 5>(buf, L, ls, es, pl.n, span, tw, line_fast); break;
      default: dif_pass<INV, 8>(buf, L, ls, es, pl.n, span, tw, line_fast); break;
    }
    span /= R;
    __syncthreads();
  }
}

// ---- plane transforms ----------------------------------------------------
// Forward: real plane -> (my/2+1, mz) half spectrum scaled by 1/(my*mz).
// buf is the plane buffer: the workgroup's LDS, or a slot in HBM for planes
// that do not fit it; the tables go to lds + tab_off (lds = buf when null).
template <int NIT, typename T, class Op>
__device__ void plane_r2c(C2<T>* buf, const PlaneGeom& G, const PlaneTabs<T>& gtb, Op& op,
                          unsigned char* lds = nullptr) {
  const int hy = G.hy, ld = G.ld, mz = G.mz, my2 = G.my2;
  const int nt = blockDim.x;
  const PlaneTabs<T> tb = stage_tabs(G, gtb, (lds ? lds : (unsigned char*)buf) + G.tab_off);
  for (int e0 = threadIdx.x; e0 < mz * hy; e0 += PIB_LD_CHUNK * nt) {
    C2<T> v[PIB_LD_CHUNK];
#pragma unroll
    for (int u = 0; u < PIB_LD_CHUNK; ++u) {
      int e = e0 + u * nt;
      if (e < mz * hy) {
        int z = e / hy, m = e - z * hy;
        op.load2(2 * m, z, v[u].x, v[u].y);
      }
    }
#pragma unroll
    for (int u = 0; u < PIB_LD_CHUNK; ++u) {
      int e = e0 + u * nt;
      if (e < mz * hy) {
        int z = e / hy, m = e - z * hy;
        buf[z * ld + m] = v[u];
      }
    }
  }
  __syncthreads();
  fft_lines<false>(buf, mz, ld, 1, G.py, tb.twy, false);
  // split step: X_k = E_k + w^k O_k, X_{hy-k} = conj(E_k - w^k O_k)
  const int npair = hy / 2 + 1;
  for (int e = threadIdx.x; e < mz * npair; e += nt) {
    int z = e / npair, k = e - z * npair;
    C2<T>* row = buf + z * ld;
    if (k == 0) {
      C2<T> z0 = row[0];
      row[0] = mk(z0.x + z0.y, (T)0);
      row[hy] = mk(z0.x - z0.y, (T)0);
    } else if (2 * k == hy) {
      int s = tb.revy[k];
      row[s] = cconj(row[s]);
    } else {
      int sa = tb.revy[k], sb = tb.revy[hy - k];
      C2<T> a = row[sa], b = row[sb];
      C2<T> E = mk((a.x + b.x) * (T)0.5, (a.y - b.y) * (T)0.5);
      C2<T> O = mk((a.y + b.y) * (T)0.5, (b.x - a.x) * (T)0.5);
      C2<T> wO = cmul(tb.wr[k], O);
      row[sa] = E + wO;
      row[sb] = cconj(E - wO);
    }
  }
  __syncthreads();
  fft_lines<false>(buf, my2, 1, ld, G.pz, tb.twz, true);
  const T scale = (T)((double)G.my * (double)G.mz);
  auto put = [&](int it, int e) {
    int kz = e / my2, ky = e - kz * my2;
    int col = (ky == hy) ? hy : tb.revy[ky];
    C2<T> v = buf[tb.revz[kz] * ld + col];
    op.store(it, ky, kz, mk(v.x / scale, v.y / scale));
  };
  if constexpr (NIT == 0) {
    // no per-thread accumulators: plain strided walk
    for (int e = threadIdx.x, it = 0; e < my2 * mz; e += nt, ++it) put(it, e);
  } else {
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      int e = threadIdx.x + it * nt;
      if (e < my2 * mz) put(it, e);
    }
  }
  __syncthreads();
}

// Inverse: (my/2+1, mz) half spectrum -> real plane, unscaled (buf and lds as
// for plane_r2c).
template <int NIT, typename T, class Op>
__device__ void plane_c2r(C2<T>* buf, const PlaneGeom& G, const PlaneTabs<T>& gtb, Op& op,
                          unsigned char* lds = nullptr) {
  const int hy = G.hy, ld = G.ld, mz = G.mz, my2 = G.my2;
  const int nt = blockDim.x;
  const PlaneTabs<T> tb = stage_tabs(G, gtb, (lds ? lds : (unsigned char*)buf) + G.tab_off);
  for (int e0 = threadIdx.x; e0 < my2 * mz; e0 += PIB_LD_CHUNK * nt) {
    C2<T> v[PIB_LD_CHUNK];
#pragma unroll
    for (int u = 0; u < PIB_LD_CHUNK; ++u) {
      int e = e0 + u * nt;
      if (e < my2 * mz) {
        int kz = e / my2, ky = e - kz * my2;
        v[u] = op.load(ky, kz);
      }
    }
#pragma unroll
    for (int u = 0; u < PIB_LD_CHUNK; ++u) {
      int e = e0 + u * nt;
      if (e < my2 * mz) {
        int kz = e / my2, ky = e - kz * my2;
        buf[kz * ld + ky] = v[u];
      }
    }
  }
  __syncthreads();
  fft_lines<true>(buf, my2, 1, ld, G.pz, tb.twz, true);
  // merge step: Z_k = A_k + i B_k, A_k = X_k + conj X_{hy-k},
  // B_k = (X_k - conj X_{hy-k}) e^{+2 pi i k/my}; Im X_0 = Im X_hy = 0
  const int npair = hy / 2 + 1;
  for (int e = threadIdx.x; e < mz * npair; e += nt) {
    int r = e / npair, k = e - r * npair;
    C2<T>* row = buf + r * ld;
    if (k == 0) {
      T x0 = row[0].x, xh = row[hy].x;
      row[0] = mk(x0 + xh, x0 - xh);
    } else if (2 * k == hy) {
      C2<T> a = row[k];
      row[k] = mk(a.x + a.x, -(a.y + a.y));
    } else {
      C2<T> a = row[k], b = row[hy - k];
      C2<T> A = mk(a.x + b.x, a.y - b.y);
      C2<T> D = mk(a.x - b.x, a.y + b.y);
      C2<T> B = cmul(D, cconj(tb.wr[k]));
      row[k] = mk(A.x - B.y, A.y + B.x);
      row[hy - k] = mk(A.x + B.y, B.x - A.y);
    }
  }
  __syncthreads();
  fft_lines<true>(buf, mz, ld, 1, G.py, tb.twy, false);
  auto put = [&](int it, int e) {
    int z = e / hy, m = e - z * hy;
    C2<T> v = buf[tb.revz[z] * ld + tb.revy[m]];
    op.store2(it, 2 * m, z, v.x, v.y);
  };
  if constexpr (NIT == 0) {
    for (int e = threadIdx.x, it = 0; e < mz * hy; e += nt, ++it) put(it, e);
  } else {
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      int e = threadIdx.x + it * nt;
      if (e < mz * hy) put(it, e);
    }
  }
  __syncthreads();
}
